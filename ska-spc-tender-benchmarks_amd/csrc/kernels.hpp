// Internal kernel-launch interface of libidgx (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "idgx.h"
#include "idgx_device.hpp"

namespace idgx
{
// Arguments shared by the gridder and degridder kernels, passed by value.
struct KernelArgs
{
    PlanDev p;
    const idgx_metadata *metadata;
    const idgx_channels *channels; // may be null: every subgrid covers all channels
    const float *uvw;
    const float *wavenumbers;
    const void *visibilities; // gridder input
    void *visibilities_out;   // degridder output
    const float *spheroidal;
    const void *aterms;
    const void *subgrids; // degridder input
    void *subgrids_out;   // gridder output
    // Optional visibility selection (flagging, ragged timesteps): subgrid s grids / degrids the entries
    // vis_index[vis_offset[s] .. vis_offset[s+1]), each (t << 16) | c relative to the subgrid's first row and
    // channel window.  Null = every (t, c) of the subgrid's window.
    const uint32_t *vis_index;
    const long long *vis_offset;
    // Mixed degridder pixels prepared once per subgrid (null: every work item computes them).  Per pixel slot
    // 8 words: for each correlation the packed high and low (re, im) parts of taper x A-term-applied pixel,
    // FP16 forms scaled by 2^-prepared_scale[s]; slot_coord holds (x', y', n, 0) of every slot (shared by the
    // class; n = 0 in the padding slots of a row).
    const uint4 *prepared;
    const int *prepared_scale;
    const float4 *slot_coord;
};

// Persistent-kernel work over up to kMaxClasses subgrid classes (one subgrid size each): the items are
// (class << 28 | subgrid, block), in nr_queues queues (items[queue_begin[k] .. queue_begin[k + 1])), each heaviest
// first; workgroups take items with an atomic counter per queue.  Queue k is the home of workgroups
// k, k + nr_queues, ...: with nr_queues = the XCD count, the workgroups of one XCD (the hardware dispatches them
// round-robin) share a queue, and every item of a subgrid is in one queue, so a subgrid's data is fetched into one
// XCD's L2 only.  A workgroup whose home queue is empty takes items from the others, so the queues drain together.
constexpr int kMaxClasses = 4;
constexpr int kQueues = 8; // XCDs of the MI355X
struct QueueArgs
{
    KernelArgs cls[kMaxClasses];
    const uint2 *items;
    int queue_begin[kQueues + 1];
    int nr_queues;
    int *counter;           // [nr_queues], zeroed by a reset kernel before every launch
    int tiles[kMaxClasses]; // gridder tiles per wave of each class (shape kShapeTilesPerClass only)
};

// Takes the next queue item (all threads of the workgroup see the same one); returns false when every queue is
// exhausted.  Counters only grow, and a workgroup tries each queue at most once per call, so every workgroup leaves
// after a bounded number of calls.
__device__ __forceinline__ bool next_queue_item(const QueueArgs &q, int &cls, int &s, int &blk)
{
    __shared__ int s_item;
    if (threadIdx.x == 0)
    {
        int item = -1;
        const int home = static_cast<int>(blockIdx.x % static_cast<unsigned>(q.nr_queues));
        for (int k = 0; k < q.nr_queues && item < 0; ++k)
        {
            const int qi = home + k < q.nr_queues ? home + k : home + k - q.nr_queues;
            const int size = q.queue_begin[qi + 1] - q.queue_begin[qi];
            if (size == 0)
                continue;
            const int i = atomicAdd(q.counter + qi, 1);
            if (i < size)
                item = q.queue_begin[qi] + i;
        }
        s_item = item;
    }
    __syncthreads();
    // The same value in every lane: read it into a scalar register, so that the item and the class's arguments
    // (QueueArgs::cls[cls]) are scalar loads rather than per-lane copies held in vector registers.
    const int i = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (i < 0)
        return false;
    const uint2 it = q.items[i];
    cls = __builtin_amdgcn_readfirstlane(static_cast<int>(it.x >> 28));
    s = __builtin_amdgcn_readfirstlane(static_cast<int>(it.x & 0x0fffffffu));
    blk = __builtin_amdgcn_readfirstlane(static_cast<int>(it.y));
    return true;
}

// Gridder queue "shape" of a single launch over several classes in which each item takes its class's tiles per
// wave (QueueArgs::tiles: 2, 3, 4, 6 or 8), with the default LDS batch and phases on the VALU.
constexpr int kShapeTilesPerClass = IDGX_SHAPE_TILES_PER_CLASS;

// Tile shape of the mixed kernels: gridder = tiles of 16 pixels per wave (PT) and visibilities staged per LDS
// batch (VB); degridder = tiles of 16 visibilities per wave (VT) and pixels staged per LDS batch (PB).
struct MixedShape
{
    int tiles;
    int batch;
    int phase_mfma; // phases from v_mfma_f32_16x16x4_f32 instead of VALU fused multiply-adds
};

// Pixel units of the mixed gridder.  The subgrid is cut into strips wd columns wide (wd the largest of 16, 8, 4, 2
// that divides S) and row blocks of tiles * rt rows (rt = 16 / wd); one unit (strip, row block) holds one wave's
// pixels, four consecutive units one workgroup item.  Row r of tile t of unit u is pixel
//   x = (u % ns) * wd + r % wd,   y = (u / ns) * rows + t * rt + r / wd,
// so a lane's column x is the same in every tile of its unit.
struct GridUnits
{
    int wd, rt, ns, rows, units;
};

__host__ __device__ inline GridUnits grid_units(int S, int tiles)
{
    GridUnits g;
    g.wd = S % 16 == 0 ? 16 : S % 8 == 0 ? 8 : S % 4 == 0 ? 4 : 2;
    g.rt = 16 / g.wd;
    g.ns = S / g.wd;
    g.rows = tiles * g.rt;
    g.units = g.ns * ((S + g.rows - 1) / g.rows);
    return g;
}

// Pixel slots of the mixed degridder: rows padded to a multiple of 4 columns, so that the 4 pixels a lane group
// takes from one 16-slot block share a row.  Slots with x >= S carry zero B words.
__host__ __device__ inline int degrid_row_stride(int S)
{
    return (S + 3) & ~3;
}

hipError_t launch_grid_plain(const KernelArgs &a, bool accurate_sincos, hipStream_t stream);
hipError_t launch_degrid_plain(const KernelArgs &a, bool accurate_sincos, hipStream_t stream);
int plain_grid_pixels_for(int S);                                  // pixels per thread of launch_grid_plain
int plain_degrid_channels_for(bool channel_windows, int nr_channels); // channels per thread of launch_degrid_plain

// Persistent work-queue forms of the plain FP32 kernels (batches).  Shape i: gridder pixels per thread and
// degridder visibilities per thread; kShapeTilesPerClass (gridder) gives every item its class's pixels per thread
// (QueueArgs::tiles).
struct PlainShape
{
    int pixels;  // gridder pixels per thread (a workgroup item covers 256 x pixels pixels of one subgrid)
    int entries; // degridder visibilities per thread (an item covers 256 x entries selected visibilities)
};
int plain_shape_count();
PlainShape plain_shape(int i);
int plain_default_shape(bool gridder, int S);
int plain_grid_items_per_subgrid(int shape, int S);
long long plain_grid_item_pixels(int shape, int S, int item);
int plain_degrid_visibilities_per_item(int shape);
hipError_t launch_grid_plain_queue(const QueueArgs &q, bool accurate_sincos, int shape, int workgroups,
                                   hipStream_t st);
hipError_t launch_degrid_plain_queue(const QueueArgs &q, bool accurate_sincos, int shape, int workgroups,
                                     hipStream_t st);
int plain_queue_workgroups(bool accurate_sincos, bool gridder, int shape);
hipError_t launch_grid_mixed(const KernelArgs &a, int precision, bool phase_mfma, hipStream_t stream);
hipError_t launch_degrid_mixed(const KernelArgs &a, int precision, bool phase_mfma, hipStream_t stream);

// Persistent work-queue forms (the queue items are built on the host, see batch.cpp).
int mixed_shape_count();
MixedShape mixed_shape(int i);
int mixed_default_shape(bool gridder, int S); // shape the one-shot kernels use for subgrid size S
int grid_items_per_subgrid(int shape, int S);            // gridder items of one subgrid
long long grid_item_pixels(int shape, int S, int item);  // pixels gridder item `item` of a subgrid writes
int degrid_visibilities_per_item(int shape);             // visibilities one degridder item covers
hipError_t launch_grid_mixed_queue(const QueueArgs &q, int precision, int shape, int workgroups, hipStream_t st);
hipError_t launch_degrid_mixed_queue(const QueueArgs &q, int precision, int shape, int workgroups, hipStream_t st);
hipError_t launch_reset_counters(int *counters, int n, hipStream_t st); // counters[0 .. n) = 0
// Degridder pixel preparation for a.prepared / a.prepared_scale / a.slot_coord (writes through those pointers).
size_t degrid_prepared_bytes(int S, long long nr_subgrids); // bytes of a.prepared
hipError_t launch_degrid_prepare(const KernelArgs &a, int precision, hipStream_t st);
int queue_workgroups(int precision, bool gridder, int shape); // resident workgroups for a persistent launch

bool subgrid_fft_supported(int S);
hipError_t launch_subgrid_fft(void *subgrids, long long nr_subgrids, int S, bool forward, hipStream_t stream);
hipError_t launch_adder(const idgx_metadata *md, const int32_t *order, const int32_t *tile_start,
                        const int32_t *color_begin, int S, int T, int G, int z_min, int nr_layers,
                        const void *subgrids, void *grid, hipStream_t st);
hipError_t launch_splitter(const idgx_metadata *md, long long nr_subgrids, int S, int G, int z_min, int nr_layers,
                           const void *grid, void *subgrids, hipStream_t st);

size_t plan_workspace_bytes(long long nr_segments);
hipError_t launch_plan(const idgx_planner &pl, const float *uvw, const double *frequencies, const uint32_t *station1,
                       const uint32_t *station2, const int32_t *seg_bl, const int32_t *seg_t0, const int32_t *seg_t1,
                       long long nr_segments, void *workspace, idgx_metadata *metadata, idgx_channels *channels,
                       long long capacity, long long *h_count, hipStream_t st);

bool grid_fft_supported(int G);
// box: nullptr, or {x0, x1, y0, y1} (0 <= x0 <= x1 <= G, same for y): BACKWARD takes the grid as zero outside the
// box, FORWARD produces only the box's cells.  Sizes outside the three-pass kernels' range ignore it.
hipError_t launch_grid_fft(void *grid, int G, long long nr_layers, bool forward, hipStream_t st,
                           const int *box = nullptr);
// Zeroes the cells a boxed BACKWARD transform reads (rows [y0, y1) of the 64-column tiles around [x0, x1); the whole
// grid for sizes outside the three-pass kernels' range).
hipError_t launch_grid_clear_box(void *grid, int G, long long nr_layers, const int *box, hipStream_t st);
// The same transform (no box) with the scratch in a file-scope array instead of a plane allocated on first use:
// G in 1024 .. 4096 (one 4096 x 4096 plane fills the array).
bool grid_fft_noalloc_supported(int G);
hipError_t launch_grid_fft_noalloc(void *grid, int G, long long nr_layers, bool forward, hipStream_t st);
bool fft2d_supported(int nx, int ny);
// Whether launch_fft2d runs this library's kernels for the shape (else rocFFT).
bool fft2d_own_kernels(int nx, int ny, bool fp64);
// Plain 2-D DFT through LDS line transforms, FP32 or FP64: small planes (both sides powers of two in 16 .. 512:
// one or two passes), larger power-of-two sides with rows of up to 16384 (FP32) / 8192 (FP64) points in one pass and
// columns longer than 512 in two four-step passes (the column tile must divide nx).
bool small_fft_supported(int nx, int ny);
bool lds_fft_supported(int nx, int ny, bool fp64);
hipError_t launch_lds_fft2d(const void *in, void *out, int nx, int ny, long long batch, int sign, bool fp64,
                            hipStream_t st);
// Scratch of the plain transforms, per (device, user, size), allocated on first use and kept for the process.
hipError_t fft_plane_scratch(long long elements, float2 *&out, int user);
// Plain 2-D DFT of `batch` planes of ny rows of nx points, out = sum in exp(sign 2 pi i (...)), unnormalised;
// in == out is in place.
hipError_t launch_fft2d(const void *in, void *out, int nx, int ny, long long batch, int sign, bool fp64,
                        hipStream_t st);
hipError_t launch_wstack(const void *src, void *dst, int G, int z_min, int nr_layers, float image_size,
                         float w_step, bool to_image, hipStream_t st);
// The boxed master-grid transform and the w-screens of launch_wstack in one chain (box as for launch_grid_fft):
//   to_image:  the image from BACKWARD-transformed layers (launch_grid_fft, then launch_wstack to the image); the
//              grid is left undefined;
//   !to_image: the layers from the image (launch_wstack from the image, then launch_grid_fft FORWARD), in the box's
//              cells.
// Three-pass sizes fold the screens and the layer sum into the row pass, so the layers' image-domain planes are
// never stored; other sizes run the two calls one after the other.
hipError_t launch_grid_wstack(void *grid, void *image, int G, int z_min, int nr_layers, float image_size,
                              float w_step, bool to_image, const int *box, hipStream_t st);
} // namespace idgx
