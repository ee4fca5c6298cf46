/* idgx — IDG gridding/degridding hot path for AMD Instinct MI355X (gfx950), C ABI.
 *
 * Every entry point takes plain pointers and sizes.  Arrays prefixed d_ live in device memory (HBM) and must
 * stay valid until the work queued on `stream` has finished; h_ arrays are host memory, borrowed for the call.
 * `stream` is a hipStream_t (NULL = the default stream).  Nothing here calls exit(): every function returns an
 * idgx_status, and idgx_strerror() names it.
 *
 * Array layouts are the reference benchmark's (gridding/inc/types.hpp, gridding/inc/common_gridding.cuh:7-26):
 *   uvw            float[nr_rows][3]                 metres
 *   wavenumbers    float[nr_channels]                2*pi*f/c, rad/m
 *   visibilities   complex64[nr_rows][nr_channels][4]  {XX, XY, YX, YY}; row = time_offset + t
 *   spheroidal     float[S][S]                       taper
 *   aterms         complex64[nr_aterm_slots][nr_stations][S][S][4]   2x2 Jones {xx, xy, yx, yy}
 *   subgrids       complex64[nr_subgrids][4][S][S]
 *   metadata       idgx_metadata[nr_subgrids]        the reference's 36-byte Metadata record
 *   channels       idgx_channels[nr_subgrids]        optional per-subgrid channel window (NULL = all)
 *
 * Subgrid-local pixel coordinates and phase offsets follow the reference: l = (x + 0.5 - S/2) * image_size / S,
 * u_offset = (x0 + S/2 - G/2) * 2*pi / image_size (gridding/inc/math.hpp:20-37,
 * gridding/src/gridding_benchmark.cpp:74-81).
 */
#ifndef IDGX_H
#define IDGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Plan record, layout-identical to the reference's Metadata (gridding/inc/types.hpp:7-25). */
typedef struct
{
    int32_t baseline_offset;
    int32_t time_offset;
    int32_t nr_timesteps;
    int32_t aterm_index;
    uint32_t station1;
    uint32_t station2;
    int32_t x, y, z; /* subgrid corner on the master grid; z = w-layer */
} idgx_metadata;

/* Channel window of one subgrid: channels [channel_begin, channel_begin + nr_channels). */
typedef struct
{
    int32_t channel_begin;
    int32_t nr_channels;
} idgx_channels;

typedef enum
{
    IDGX_OK = 0,
    IDGX_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad size, inconsistent plan */
    IDGX_ERR_INVALID_METADATA = 2, /* a subgrid reads rows/channels/stations/A-terms outside the arrays */
    IDGX_ERR_UNSUPPORTED = 3,      /* valid request this build does not implement (e.g. subgrid size) */
    IDGX_ERR_DEVICE = 4,           /* HIP runtime error (launch, copy, event) */
    IDGX_ERR_OUT_OF_MEMORY = 5
} idgx_status;

typedef enum
{
    IDGX_PLAIN_FP32 = 0,   /* FP32 throughout, phases rounded exactly as the reference CPU code */
    IDGX_MIXED_BF16X3 = 1, /* phasor x data contraction on BF16 MFMA, hi/lo split, 3 products, FP32 accumulate */
    IDGX_MIXED_FP16X3 = 2, /* same on FP16 MFMA */
    IDGX_MIXED_BF16 = 3,   /* single BF16 product (accuracy reference point) */
    IDGX_MIXED_FP16 = 4    /* single FP16 product */
} idgx_precision;

typedef enum
{
    IDGX_ATERM_ADJOINT = 0,  /* degridder applies A1 * P * A2^H: the exact adjoint of the gridder */
    IDGX_ATERM_REFERENCE = 1 /* degridder applies A1^H * P * A2, as the reference CPU degridder does
                                (gridding/src/degridding_benchmark.cpp:113) */
} idgx_aterm_convention;

typedef enum
{
    /* Plain FP32 path: evaluate sin/cos with the ROCm device library's sincosf (full-range reduction, ~1 ulp)
     * instead of the default FP32 reduction to a fraction of a turn (two FMAs with a two-part 1/(2 pi)) followed by
     * v_sin_f32/v_cos_f32. */
    IDGX_FLAG_ACCURATE_SINCOS = 1,
    /* Mixed forms: evaluate the per-pair phase qu x' + qv y' + dw n as an FP32 MFMA (v_mfma_f32_16x16x4_f32)
     * instead of three VALU fused multiply-adds.  Same accuracy; a tuning choice. */
    IDGX_FLAG_PHASE_MFMA = 2
} idgx_flags;

typedef struct
{
    int32_t nr_subgrids;
    int32_t subgrid_size;      /* S; 8 <= S <= 128, even */
    int32_t grid_size;         /* G */
    int32_t nr_channels;       /* channels per visibility row; length of wavenumbers */
    int32_t nr_stations;
    int32_t nr_aterm_slots;
    int64_t nr_rows;           /* rows of uvw / visibilities */
    int32_t max_nr_timesteps;  /* upper bound of metadata[].nr_timesteps (sizes the launch) */
    float image_size;          /* radians */
    float w_step_in_lambda;    /* w-layer spacing; 0 = no w-layers (the reference's value) */
    int32_t precision;         /* idgx_precision */
    int32_t aterm_convention;  /* idgx_aterm_convention, degridder only */
    int32_t flags;             /* bitwise OR of idgx_flags */
} idgx_plan;

/* Checks a plan and its host-side metadata: every subgrid must read rows, channels, stations and A-term
 * slots inside the arrays the plan describes.  The GPU entry points do not read device metadata on the host,
 * so callers validate with this first; the kernels additionally skip (and zero) any subgrid that would read
 * out of bounds. */
int idgx_check_plan(const idgx_plan *plan, const idgx_metadata *h_metadata, const idgx_channels *h_channels);

/* Gridder: subgrids[s] = taper * A1^H (sum_{t,c} V[t,c] e^{i phase}) A2, phase = phase_offset - k_c (u l + v m + w n).
 * Replaces gridding_benchmark_launcher<gpu>::launch's kernel (gridding/src/gridding_benchmark.cu:19-246,326)
 * and the CPU loop it mirrors (gridding/src/gridding_benchmark.cpp:56-165).  Subgrids are overwritten. */
int idgx_grid(const idgx_plan *plan, const idgx_metadata *d_metadata, const idgx_channels *d_channels,
              const float *d_uvw, const float *d_wavenumbers, const void *d_visibilities, const float *d_spheroidal,
              const void *d_aterms, void *d_subgrids, void *stream);

/* Degridder: V[t,c] = sum_pixels e^{-i phase} A1 (taper * P) A2^H  (the adjoint of idgx_grid; with
 * IDGX_ATERM_REFERENCE the A-term is A1^H P A2 as in the reference CPU path).  Replaces
 * degridding_benchmark_launcher<gpu>::launch's kernel (gridding/src/degridding_benchmark.cu:106-203,283) and the
 * CPU loop (gridding/src/degridding_benchmark.cpp:58-181).  Only the (row, channel) entries the plan covers are
 * written; the rest of d_visibilities is left untouched. */
int idgx_degrid(const idgx_plan *plan, const idgx_metadata *d_metadata, const idgx_channels *d_channels,
                const float *d_uvw, const float *d_wavenumbers, const void *d_subgrids, const float *d_spheroidal,
                const void *d_aterms, void *d_visibilities, void *stream);

/* ---- Host-buffer calls: the reference launcher's contract in one call --------------------------------------
 *
 * For callers that hold only host arrays (a cgo / ctypes / JNI binding; the C++ drop-in launcher is built on
 * them).  Each call checks the plan against h_metadata, allocates the device buffers, runs `iterations` uploads,
 * `iterations` kernel runs and `iterations` downloads of the result on a private stream, frees everything and
 * fills `timing` (may be NULL).  Plain FP32 runs the one-shot kernels (idgx_grid / idgx_degrid); the mixed forms run
 * as a one-class batch (idgx_batch_*) created outside the timed runs.  Host layouts are the reference's:
 * uvw [nr_rows][3] float (metres), wavenumbers [nr_channels], visibilities [nr_rows][nr_channels][4] complex64,
 * spheroidal [S][S], aterms [slots][stations][S][S][4] complex64, subgrids [nr_subgrids][4][S][S] complex64;
 * h_channels (per-subgrid channel windows) may be NULL: every subgrid then covers all channels, as in the
 * reference's plans.
 * Replaces gridding_benchmark_launcher<gpu>::launch (gridding/src/gridding_benchmark.cu:248-361) and
 * degridding_benchmark_launcher<gpu>::launch (gridding/src/degridding_benchmark.cu:205-318), whose interface is
 * gridding/inc/gridding_benchmark.h:10-21 and gridding/inc/degridding_benchmark.h:18-28. */

/* Per-iteration timings, as the reference's benchmark_result (gridding/inc/common_gridding.h:5-12). */
typedef struct
{
    double in_transfer_time;  /* microseconds per upload of the inputs */
    double out_transfer_time; /* microseconds per download of the result */
    double in_bandwidth;      /* MB/s (bytes per microsecond) */
    double out_bandwidth;     /* MB/s */
    double compute_time;      /* microseconds per kernel run */
} idgx_timing;

/* Gridder from host buffers: h_subgrids receives the subgrids. */
int idgx_host_grid(const idgx_plan *plan, const idgx_metadata *h_metadata, const idgx_channels *h_channels,
                   const float *h_uvw, const float *h_wavenumbers, const void *h_visibilities,
                   const float *h_spheroidal, const void *h_aterms, void *h_subgrids, int iterations,
                   idgx_timing *timing);

/* Degridder from host buffers: h_visibilities receives every row and channel (entries no subgrid covers are 0). */
int idgx_host_degrid(const idgx_plan *plan, const idgx_metadata *h_metadata, const idgx_channels *h_channels,
                     const float *h_uvw, const float *h_wavenumbers, const void *h_subgrids,
                     const float *h_spheroidal, const void *h_aterms, void *h_visibilities, int iterations,
                     idgx_timing *timing);

/* ---- Stages after the gridder / before the degridder (no reference counterpart: the reference benchmark stops
 * at image-domain subgrids, gridding/src/gridding_benchmark.cpp:155-163) ------------------------------------- */

typedef enum
{
    IDGX_FFT_FORWARD = 0, /* image-domain subgrid -> uv subgrid (after the gridder) */
    IDGX_FFT_BACKWARD = 1 /* uv subgrid -> image-domain subgrid (before the degridder); adjoint of FORWARD */
} idgx_fft_direction;

/* In-place batched 2-D transform of d_subgrids (complex64 [nr_subgrids][4][S][S]) that maps the gridder's pixel
 * x, at l = (x + 1/2 - S/2) image_size / S, onto uv cell k at u = u_centre + (k - S/2) / image_size:
 *   FORWARD  G[k] = sum_x I[x] exp(+2 pi i (k - S/2)(x - S/2 + 1/2) / S)   (in y and x)
 *   BACKWARD I[x] = sum_k G[k] exp(-2 pi i (k - S/2)(x - S/2 + 1/2) / S)
 * unnormalised, so BACKWARD(FORWARD(I)) = S^2 I.  S must be 2^a 3^b <= 64 (8, 12, 16, 24, 32, 48, 64). */
int idgx_subgrid_fft(const idgx_plan *plan, void *d_subgrids, int direction, void *stream);

/* Conflict-free order of the adder: tiles of tile_size x tile_size grid cells (tile_size = S), coloured by
 * tile-coordinate parity; tiles [color_begin[c], color_begin[c+1]) have colour c, and tile t adds subgrids
 * order[tile_start[t] .. tile_start[t+1]) in that order. */
typedef struct
{
    int32_t tile_size;
    int32_t nr_tiles;
    int32_t color_begin[5];
    int32_t z_min;     /* w-stacking: subgrid s goes to layer metadata[s].z - z_min ... */
    int32_t nr_layers; /* ... of a grid complex64 [nr_layers][4][G][G]; 1 = a single grid, z ignored */
} idgx_adder_plan;

/* Builds the adder order on the host from the plan's metadata (single grid, z ignored).  h_order has nr_subgrids
 * entries, h_tile_start nr_subgrids + 1 (only nr_tiles + 1 are used). */
int idgx_adder_plan_build(const idgx_plan *plan, const idgx_metadata *h_metadata, idgx_adder_plan *adder_plan,
                          int32_t *h_order, int32_t *h_tile_start);

/* The same for w-stacking onto nr_layers grids: every metadata[s].z must lie in [z_min, z_min + nr_layers)
 * (IDGX_ERR_INVALID_METADATA otherwise).  A subgrid's phase reference in the gridder is
 * w_offset = (z + 0.5) w_step_in_lambda, as in gridding/src/gridding_benchmark.cpp:80. */
int idgx_adder_plan_build_layered(const idgx_plan *plan, const idgx_metadata *h_metadata, int z_min, int nr_layers,
                                  idgx_adder_plan *adder_plan, int32_t *h_order, int32_t *h_tile_start);

/* Adder: d_grid[layer][pol][y0 + y][x0 + x] += d_subgrids[s][pol][y][x] for every subgrid (cells outside the
 * grid are skipped); layer = z - z_min for a layered adder plan, else 0.  d_grid is complex64
 * [nr_layers][4][G][G]; d_order / d_tile_start are the device copies of the arrays built by
 * idgx_adder_plan_build(_layered).  The summation order per cell is fixed, so results are bit-reproducible. */
int idgx_add_subgrids(const idgx_plan *plan, const idgx_adder_plan *adder_plan, const idgx_metadata *d_metadata,
                      const int32_t *d_order, const int32_t *d_tile_start, const void *d_subgrids, void *d_grid,
                      void *stream);

/* Splitter: d_subgrids[s][pol][y][x] = d_grid[pol][y0 + y][x0 + x] (zero outside the grid); the adjoint of the
 * adder. */
int idgx_split_subgrids(const idgx_plan *plan, const idgx_metadata *d_metadata, const void *d_grid, void *d_subgrids,
                        void *stream);

/* Splitter from w-layers: subgrid s reads layer metadata[s].z - z_min of d_grid (complex64 [nr_layers][4][G][G]);
 * a subgrid of no layer reads zeros.  The adjoint of the layered adder. */
int idgx_split_subgrids_layered(const idgx_plan *plan, const idgx_metadata *d_metadata, int z_min, int nr_layers,
                                const void *d_grid, void *d_subgrids, void *stream);

/* ---- Master grid and w-stacking (the step after the adder; the reference's fft/ benchmark times a plain 2-D FFT
 * of this size, fft/inc/fft_benchmark.h:65-99) -------------------------------------------------------------------
 *
 * Master-grid cell K is at u = (K - G/2) / image_size and image pixel X at l = (X - G/2) image_size / G.  In place
 * on d_grid (complex64 [nr_layers][4][G][G]), in each dimension:
 *   FORWARD  (image -> uv):  G[K] = sum_X I[X] exp(+2 pi i (K - G/2)(X - G/2) / G)
 *   BACKWARD (uv -> image):  I[X] = sum_K G[K] exp(-2 pi i (K - G/2)(X - G/2) / G)
 * unnormalised (BACKWARD(FORWARD(I)) = G^2 I), the signs of idgx_subgrid_fft.  G must be a multiple of 4.  For
 * G = 1024 .. 16384 (powers of two) the transform is this library's three-pass kernel set; the first call for a G on
 * a device allocates a scratch of up to 64 MB, or one plane (G^2 x 8 B) when a plane is larger, kept for the life
 * of the process.  Other sizes run idgx_fft2d's transform (in place, batched over layers and correlations) between
 * two sign-flip passes: this library's kernels for G = 16 .. 512 and 3 x 2^k from 3072 (a scratch of one plane),
 * rocFFT otherwise (plans cached per device, G and batch; the first call for a size allocates them).  Calls that
 * overlap on one device share the scratch, so issue them on one stream. */
int idgx_grid_fft(const idgx_plan *plan, void *d_grid, int nr_layers, int direction, void *stream);

/* idgx_grid_fft that allocates nothing (no hipMalloc / hipFree, no library plan), for callers that cannot allocate:
 * code under hipGraph capture, tight serving loops, harnesses that forbid device allocation.  No counterpart in the
 * reference.  Same kernels and results as idgx_grid_fft, bit for bit; its scratch is a fixed 128 MiB array in the
 * library's device code (one copy per GPU, held from the library's load until the process ends), so G must be
 * 1024, 2048 or 4096 (IDGX_ERR_UNSUPPORTED otherwise: larger planes do not fit the array, and other sizes would
 * need rocFFT, which allocates).  Every call on a GPU shares that GPU's array: calls that overlap on one GPU (two
 * streams, two host threads) race, so issue them on one stream. */
int idgx_grid_fft_noalloc(const idgx_plan *plan, void *d_grid, int nr_layers, int direction, void *stream);

/* The bounding box {x0, x1, y0, y1} of a plan's subgrids on the master grid, cells [x0, x1) x [y0, y1) clamped to
 * [0, G): the only cells the adder writes and the splitter reads.  {0, 0, 0, 0} for an empty plan. */
int idgx_grid_box(const idgx_plan *plan, const idgx_metadata *h_metadata, int32_t box[4]);

/* idgx_grid_fft restricted to a box of cells {x0, x1, y0, y1} of every plane (0 <= x0 <= x1 <= G, same for y), for
 * grids whose uv cells outside the box carry nothing (a plan's subgrids cover the box of idgx_grid_box):
 *   BACKWARD (uv -> image): the grid must be zero outside the box.  The column transforms run only over the
 *     64-column tiles that meet [x0, x1), their first pass reads only rows [y0, y1), and the row transforms read
 *     only those columns.  The image is complete
 *     and equals idgx_grid_fft's bit for bit.
 *   FORWARD (image -> uv): only the box's cells of the uv grid are produced (with the rest of the 64-column tiles
 *     around it), equal to idgx_grid_fft's to FP32 rounding: the row transforms run first, dense and in place,
 *     then the column transforms only over the box's column tiles.  The other cells are left undefined.  A box
 *     whose column tiles and rows cover most of the grid, where that order would move more data than the full
 *     transform, runs the full transform (every cell produced, bit for bit idgx_grid_fft's).
 * Grid sizes outside the three-pass kernels' range (see idgx_grid_fft) run the full transform.  No counterpart in
 * the reference (its benchmark stops at the subgrids); it is the sparse master-grid transform of an imaging cycle
 * whose uv coverage is smaller than the grid. */
int idgx_grid_fft_box(const idgx_plan *plan, void *d_grid, int nr_layers, int direction, const int32_t box[4],
                      void *stream);

/* Zeroes every cell that idgx_grid_fft_box(BACKWARD) reads for this box (rows [y0, y1) of the 64-column tiles that
 * meet [x0, x1); the whole grid for sizes outside the three-pass kernels' range), with a fill kernel: enough to start
 * an imaging cycle's grid when the adder writes only inside the box, even if a FORWARD transform left the rest of
 * the grid undefined. */
int idgx_grid_clear_box(const idgx_plan *plan, void *d_grid, int nr_layers, const int32_t box[4], void *stream);

/* Plain batched 2-D DFT, the transform of the reference's FFT benchmark (heFFTe fft2d, fft/inc/fft_benchmark.h:65-99,
 * run by fft/src/fft_benchmark.cpp:240-345): for b < batch,
 *   d_out[b][ky][kx] = sum_{y < ny, x < nx} d_in[b][y][x] exp(sign 2 pi i (ky y / ny + kx x / nx)),
 * unnormalised, sign = -1 (heFFTe forward) or +1 (backward); complex64 planes, or complex128 with
 * precision = IDGX_FFT_FP64.  d_in == d_out transforms in place.  This library's kernels run:
 *   - FP32 64 x 64 and 128 x 128: one pass with the plane in registers;
 *   - other nx, ny powers of two in 16 .. 512, FP32 or FP64: LDS tiles (one pass when both sides are <= 128 and the
 *     plane fits 132 KB of LDS, i.e. up to 128 x 64 in FP64; else a row pass and a column pass, 64 MB of planes at
 *     a time);
 *   - other power-of-two shapes with nx up to 16384 (FP32) / 8192 (FP64) and ny up to 16384 (FP64 every such
 *     shape; FP32 those with a side below 1024): rows in one LDS pass, columns longer than 512 in two four-step
 *     passes through a scratch (their column tiles, 32 .. 128 columns, must divide nx);
 *   - FP32, nx and ny powers of two in 1024 .. 16384: the three-pass kernels;
 *   - FP32, nx = ny = 3 x (a power of two in 1024 .. 8192), e.g. 24576: three passes as well, the transforms of the
 *     sub-rows x = 3 x' + b, then two column passes each opening with a radix-3 step (along y, then along x).
 * Planes hold at most 2^31 - 1 points.
 * The first call for a plane size on a device allocates the scratch those need (the three-pass kernels and the
 * four-step columns: 64 MB of planes, or one plane when larger; the 3 x 2^k sides: one plane), kept until the
 * process ends.  Other shapes run rocFFT
 * (the first call for a shape creates its plan).  Calls that overlap on one device share the scratch, so issue
 * them on one stream. */
enum
{
    IDGX_FFT_FP32 = 0,
    IDGX_FFT_FP64 = 1
};
int idgx_fft2d(const void *d_in, void *d_out, int nx, int ny, int64_t batch, int sign, int precision, void *stream);

/* 1 when idgx_fft2d runs this library's own kernels for the shape and precision, 0 when it runs rocFFT, negative
 * (-IDGX_ERR_INVALID_ARGUMENT) for a bad argument.  No GPU needed. */
int idgx_fft2d_own_kernels(int nx, int ny, int precision);

/* W-stacking in the image domain, with w_z = (z + 0.5) w_step_in_lambda for layer z - z_min and
 * n = 1 - sqrt(1 - l^2 - m^2) at image pixel (X, Y):
 *   to_image:   d_image[pol][Y][X] = sum_z d_layers[z - z_min][pol][Y][X] exp(-2 pi i w_z n)
 *   from_image: d_layers[z - z_min][pol][Y][X] = d_image[pol][Y][X] exp(+2 pi i w_z n)   (the adjoint)
 * d_image is complex64 [4][G][G] and must not alias d_layers. */
int idgx_wstack_to_image(const idgx_plan *plan, const void *d_layers, int z_min, int nr_layers, void *d_image,
                         void *stream);
int idgx_wstack_from_image(const idgx_plan *plan, const void *d_image, int z_min, int nr_layers, void *d_layers,
                           void *stream);

/* The image side of an imaging cycle in one call: the boxed master-grid transform and the w-screens together.
 *   idgx_grid_to_image: d_image = idgx_wstack_to_image(idgx_grid_fft_box(BACKWARD, d_grid)), the grid's layers
 *     zero outside the box as idgx_grid_fft_box requires; d_grid is left undefined (it serves as scratch);
 *   idgx_image_to_grid: d_grid = idgx_grid_fft_box(FORWARD, idgx_wstack_from_image(d_image)) in the box's cells
 *     (the rest of the grid undefined, as for idgx_grid_fft_box FORWARD).
 * For the three-pass sizes (G = 1024 ... 16384) the screens and the sum over layers run inside the row pass: each
 * image row is transformed for every layer while n(l, m) stays in registers, so neither the layers' image-domain
 * planes nor a separate w-screen pass touch HBM.  The screens there take v_sin / v_cos of the FP64-reduced fraction
 * of a turn (about 1e-7 absolute), not FP64 sincospi.  idgx_grid_to_image keeps a scratch of
 * nr_layers x G x (box width rounded out to 64-column tiles) cells per device, allocated on first use and grown
 * (after a device synchronisation, which also makes a growing call unsuitable for hipGraph capture) when a call
 * needs more; calls that overlap on one device share it, so issue them from one host thread on one stream.  Other
 * sizes run the two calls one after the other.
 * No counterpart in the reference (its benchmark stops at the subgrids).  d_image must not alias d_grid. */
int idgx_grid_to_image(const idgx_plan *plan, void *d_grid, int z_min, int nr_layers, const int32_t box[4],
                       void *d_image, void *stream);
int idgx_image_to_grid(const idgx_plan *plan, const void *d_image, int z_min, int nr_layers, const int32_t box[4],
                       void *d_grid, void *stream);

/* ---- Device-side planner (no reference counterpart: the reference draws random metadata,
 * gridding/src/gridding_initialization.cpp:158-182) -------------------------------------------------------------
 *
 * Turns time segments of baselines into subgrid metadata and channel windows on the GPU.  Segment i covers rows
 * seg_baseline[i] * nr_times + [seg_t0[i], seg_t1[i]) of uvw (float [rows][3], metres); a segment must not cross an
 * A-term slot boundary and must not be empty.  Its channels are grouped greedily: a group closes when adding the
 * next channel would make the uv footprint of its samples (metres x frequency / c x image_size grid cells) wider
 * than subgrid_size - support cells in u or v.  Each group becomes a subgrid centred on its footprint, dropped if it
 * would leave the grid; with w_step_in_lambda > 0, its w-layer is z = floor(middle of the segment's w range x the
 * group's centre frequency / c / w_step).  Arithmetic is FP64 without contraction, so the plan is bit-identical
 * to the package's host planner.  Subgrids are written in segment order, each segment's groups in channel order
 * (sorted by first row and first channel when the segments are sorted by baseline and t0). */
typedef struct
{
    int32_t subgrid_size;
    int32_t grid_size;
    int32_t support;        /* cells of the subgrid kept free of samples (taper support) */
    int32_t nr_times;       /* rows per baseline */
    int32_t nr_channels;
    int32_t aterm_interval; /* timesteps per A-term slot; 0 = one slot */
    double image_size;
    double w_step_in_lambda; /* 0 = no w-layers */
} idgx_planner;

/* Bytes of device workspace idgx_plan_segments needs for nr_segments segments. */
size_t idgx_plan_workspace_bytes(int64_t nr_segments);

/* Writes at most `capacity` subgrids to d_metadata / d_channels and returns their total number in *h_nr_subgrids
 * (call again with a larger capacity when it exceeds `capacity`).  d_frequencies are the channel frequencies in Hz
 * (double [nr_channels]); d_station1 / d_station2 the stations of each baseline.  Synchronises `stream`. */
int idgx_plan_segments(const idgx_planner *planner, const float *d_uvw, const double *d_frequencies,
                       const uint32_t *d_station1, const uint32_t *d_station2, const int32_t *d_seg_baseline,
                       const int32_t *d_seg_t0, const int32_t *d_seg_t1, int64_t nr_segments, void *d_workspace,
                       idgx_metadata *d_metadata, idgx_channels *d_channels, int64_t capacity,
                       int64_t *h_nr_subgrids, void *stream);

/* ---- Batched execution: persistent launches over a whole plan ------------------------------------------------
 *
 * A batch holds up to 4 subgrid classes (one subgrid size each).  Its work queues of (class, subgrid, block)
 * items are built on the host at creation, one per class and direction, sorted heaviest first, and consumed by
 * persistent workgroups through device atomic counters.  A batch of several classes starts in single-launch
 * mode: the whole plan is one launch per direction over a merged queue of every class's items, heaviest first
 * across classes, the gridder's items each with their class's tiles per wave.  idgx_batch_set_single_launch(b,
 * -1, -1) switches to one launch per class, each class with its own tile shape (a one-class batch runs that way
 * from the start; for one class the two modes coincide).  Optional per-class visibility selections (flagging,
 * ragged timesteps) restrict each subgrid to the listed entries: subgrid s uses vis_index[vis_offset[s] ..
 * vis_offset[s+1]), each entry (t << 16) | c relative to the subgrid's first row and channel window.  Creation
 * allocates the queue memory; the execution entry points allocate nothing, and idgx_batch_capture records them into
 * a hipGraph (no memset nodes) for idgx_batch_launch.
 * Every precision batches: the mixed forms run their persistent MFMA kernels (shapes idgx_batch_shape_info), plain
 * FP32 the persistent forms of the plain kernels (shapes idgx_batch_plain_shape_info), which skip unselected
 * (flagged) samples instead of multiplying them by zero and otherwise round like idgx_grid / idgx_degrid.  They
 * replace the reference's one-block-per-subgrid launches over every sample (gridding/src/gridding_benchmark.cu:
 * 323-332, gridding/src/degridding_benchmark.cu:283).  All classes of a batch share precision and flags.  A batch
 * is used from one host thread at a time. */

typedef struct idgx_batch idgx_batch;

enum
{
    IDGX_STAGE_GRID = 1,
    IDGX_STAGE_DEGRID = 2
};

typedef struct
{
    idgx_plan plan;                  /* all classes share precision, channels, rows, stations, grid and image size */
    const idgx_metadata *h_metadata; /* host copy of the class's metadata */
    const idgx_channels *h_channels; /* host copy of the channel windows, or NULL */
    const int64_t *h_vis_offset;     /* host copy of the selection offsets (nr_subgrids + 1), or NULL (dense) */
    int64_t nr_vis_index;            /* length of the class's selection list */
} idgx_batch_class;

typedef struct
{
    const idgx_metadata *d_metadata;
    const idgx_channels *d_channels;
    const uint32_t *d_vis_index; /* NULL together with d_vis_offset = dense */
    const int64_t *d_vis_offset;
    const float *d_spheroidal; /* [S][S] of this class */
    const void *d_aterms;      /* [slots][stations][S][S][4] of this class */
    void *d_subgrids;          /* [n][4][S][S]: gridder output, degridder input */
} idgx_batch_class_data;

typedef struct
{
    const float *d_uvw;
    const float *d_wavenumbers;
    const void *d_visibilities; /* gridder input */
    void *d_visibilities_out;   /* degridder output (may be the same buffer) */
    idgx_batch_class_data cls[4];
} idgx_batch_data;

int idgx_batch_create(const idgx_batch_class *classes, int nr_classes, idgx_batch **batch);
void idgx_batch_destroy(idgx_batch *batch);

/* Tile shapes of the mixed forms: shape i = (tiles of 16 pixels / visibilities per wave, LDS staging depth, phases
 * on the matrix core). */
int idgx_batch_shape_count(void);
int idgx_batch_shape_info(int shape, int *tiles_per_wave, int *lds_batch, int *phase_mfma);
/* Shapes of a plain FP32 batch: shape i = (gridder pixels per thread, degridder visibilities per thread).  The
 * shape arguments of the calls below index this table when the batch's precision is IDGX_PLAIN_FP32, and with
 * IDGX_SHAPE_TILES_PER_CLASS each gridder item takes its class's pixels per thread. */
int idgx_batch_plain_shape_count(void);
int idgx_batch_plain_shape_info(int shape, int *gridder_pixels_per_thread, int *degridder_visibilities_per_thread);
/* set_shape sets every class's shapes; get_shape reports class 0's. */
int idgx_batch_set_shape(idgx_batch *batch, int gridder_shape, int degridder_shape);
int idgx_batch_get_shape(const idgx_batch *batch, int *gridder_shape, int *degridder_shape);
int idgx_batch_set_class_shape(idgx_batch *batch, int cls, int gridder_shape, int degridder_shape);
int idgx_batch_get_class_shape(const idgx_batch *batch, int cls, int *gridder_shape, int *degridder_shape);

/* Gridder shape of a single launch in which every item takes the tiles per wave of its class's gridder shape
 * (LDS batch 256, phases on the VALU). */
enum
{
    IDGX_SHAPE_TILES_PER_CLASS = -2
};

/* Single-launch mode per direction: one persistent launch over the merged queue of all classes with the given
 * shape (for the gridder also IDGX_SHAPE_TILES_PER_CLASS); -1 switches that direction to per-class launches.
 * Defaults: several classes (IDGX_SHAPE_TILES_PER_CLASS, the default degridder shape); one class (-1, -1).  The
 * per-class shapes are kept.  For a given shape the results equal the per-class launches' bit for bit (every item
 * writes its own outputs).  get reports -1 for a direction in per-class mode. */
int idgx_batch_set_single_launch(idgx_batch *batch, int gridder_shape, int degridder_shape);
int idgx_batch_get_single_launch(const idgx_batch *batch, int *gridder_shape, int *degridder_shape);

/* Times every shape `repeats` times per class and direction on the given data (overwriting its outputs), keeps
 * each class's fastest and reports the sum over classes of their average times (ms).  A direction in
 * single-launch mode then times, on the whole direction, its tuned per-class launches run together and every shape
 * of its merged launch (gridder: also IDGX_SHAPE_TILES_PER_CLASS), keeps the fastest (per-class launches switch
 * the direction to per-class mode, as idgx_batch_set_single_launch with -1 would) and reports that time.  A
 * direction in per-class mode stays in it. */
int idgx_batch_autotune(idgx_batch *batch, const idgx_batch_data *data, int repeats, void *stream,
                        float *best_gridder_ms, float *best_degridder_ms);

/* Persistent gridder / degridder launches (one per class, or one in single-launch mode) over every subgrid of every
 * class. */
int idgx_batch_grid(idgx_batch *batch, const idgx_batch_data *data, void *stream);
int idgx_batch_degrid(idgx_batch *batch, const idgx_batch_data *data, void *stream);

/* Records the selected stages (IDGX_STAGE_*) for `data` into a hipGraph (stream must not be NULL);
 * idgx_batch_launch replays it.  Changing the shape drops the graph. */
int idgx_batch_capture(idgx_batch *batch, const idgx_batch_data *data, int stages, void *stream);
int idgx_batch_launch(idgx_batch *batch, void *stream);

/* Persistent gridder / degridder kernel launches that the latest idgx_batch_grid / idgx_batch_degrid (or the
 * capture that recorded them) issued: 1 in single-launch mode, one per class with work otherwise. */
int idgx_batch_last_launches(const idgx_batch *batch, int *gridder_launches, int *degridder_launches);

/* Queue sizes (all classes) and resident workgroups (the largest launch) of the current shapes. */
int idgx_batch_items(const idgx_batch *batch, int *gridder_items, int *degridder_items, int *gridder_workgroups,
                     int *degridder_workgroups);

/* Work per thread of the plain FP32 one-shot kernels for a plan (idgx_grid / idgx_degrid with IDGX_PLAIN_FP32):
 * gridder pixels per thread (4, or 3 where that leaves fewer lanes idle: S = 24, 48) and degridder channels per
 * thread (4 without channel windows when the channel count is a multiple of 4, else 2).  They name the kernel
 * instantiation a profile reports (gridder_plain<pixels, 1>, degridder_plain<channels, 1>).  No GPU needed. */
int idgx_plain_kernel_shape(const idgx_plan *plan, int has_channel_windows, int *gridder_pixels_per_thread,
                            int *degridder_channels_per_thread);

/* Name of a status code. */
const char *idgx_strerror(int status);

/* Library version string. */
const char *idgx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* IDGX_H */
