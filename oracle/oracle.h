/* Oracle for the IDG gridding/degridding hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and only as the
 * checker (or, for cpu_baseline, as the timed CPU restatement).  The product path never calls it.
 *
 * Two families live here:
 *  - *_fp32: a faithful CPU restatement of the reference's OpenMP CPU gridder/degridder (same loop nest, same
 *    FP32 operation order, same per-routine rounding of the subgrid offsets, glibc sinf/cosf).  Compiled with
 *    -ffp-contract=off it reproduces the reference's output bit for bit; tests/test_oracle.py pins that against
 *    golden vectors produced from the compiled reference sources (tests/golden/).
 *  - *_fp64: the same mathematics evaluated in double precision (phases, sin/cos, accumulation), used as the
 *    accuracy yardstick.  Pinned by closed forms (single visibility / single pixel), the adjoint dot-product
 *    identity, and agreement with the FP32 restatement on short baselines.
 *  - subgrid FFT and adder/splitter restatements for the stages after the gridder (fft2d_*, adder_*, splitter_*),
 *    pinned by brute-force DFT, round trips and brute-force placement.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Plan record: same layout as the reference's Metadata (gridding/inc/types.hpp:7-25), 36 bytes. */
typedef struct
{
    int32_t baseline_offset;
    int32_t time_offset;
    int32_t nr_timesteps;
    int32_t aterm_index;
    uint32_t station1;
    uint32_t station2;
    int32_t x, y, z;
} oracle_metadata;

typedef struct
{
    int32_t nr_subgrids;
    int32_t subgrid_size;
    int32_t grid_size;
    int32_t nr_channels; /* channels per visibility row (row stride) */
    int32_t nr_stations;
    float image_size;
    float w_step_in_lambda;
} oracle_params;

/* Per-subgrid channel window [channel_begin, channel_begin + nr_channels); NULL = every channel. */
typedef struct
{
    int32_t channel_begin;
    int32_t nr_channels;
} oracle_channels;

/* A-term convention of the degridder:
 *   ORACLE_ATERM_ADJOINT   P <- A1 * P * A2^H  (the adjoint of the gridder's A1^H * P * A2)
 *   ORACLE_ATERM_REFERENCE P <- A1^H * P * A2  (what the reference CPU degridder applies,
 *                                               gridding/src/degridding_benchmark.cpp:113) */
enum
{
    ORACLE_ATERM_ADJOINT = 0,
    ORACLE_ATERM_REFERENCE = 1
};

/* uvw: float[rows][3]; wavenumbers: float[nr_channels]; visibilities: float[rows][nr_channels][4][2];
 * spheroidal: float[S][S]; aterms: float[slots][stations][S][S][4][2]; subgrids: float[n][4][S][S][2]. */
int oracle_grid_fp32(const oracle_params *p, const oracle_metadata *metadata, const oracle_channels *channels,
                     const float *uvw, const float *wavenumbers, const float *visibilities, const float *spheroidal,
                     const float *aterms, float *subgrids, int nthreads);

int oracle_degrid_fp32(const oracle_params *p, const oracle_metadata *metadata, const oracle_channels *channels,
                       const float *uvw, const float *wavenumbers, const float *subgrids, const float *spheroidal,
                       const float *aterms, float *visibilities, int aterm_convention, int nthreads);

/* FP64 oracle: same argument meaning, outputs in double (subgrids: double[n][4][S][S][2],
 * visibilities: double[rows][nr_channels][4][2]); the degridder always applies the adjoint A-term. */
int oracle_grid_fp64(const oracle_params *p, const oracle_metadata *metadata, const oracle_channels *channels,
                     const float *uvw, const float *wavenumbers, const float *visibilities, const float *spheroidal,
                     const float *aterms, double *subgrids, int nthreads);

int oracle_degrid_fp64(const oracle_params *p, const oracle_metadata *metadata, const oracle_channels *channels,
                       const float *uvw, const float *wavenumbers, const float *subgrids, const float *spheroidal,
                       const float *aterms, double *visibilities, int nthreads);

/* Batched 2-D DFT of complex subgrids in double, data double[batch][S][S][2], in place.
 * sign = -1: forward  X[k] = sum_n x[n] e^{-2 pi i k.n / S};  sign = +1: inverse without 1/S^2 scaling. */
int oracle_dft2d_fp64(double *data, int batch, int S, int sign, int nthreads);

/* Adder: grid[pol][gy][gx] += subgrid[s][pol][y][x] at gy = meta.y + y, gx = meta.x + x (pixels that fall
 * outside the grid are skipped); grid double[4][G][G][2], subgrids float[n][4][S][S][2].  Splitter: the
 * transpose, subgrid[s][pol][y][x] = grid[pol][gy][gx] (0 outside the grid).  fftshift conventions of the
 * product are applied by the caller on the test side, this is the plain placement. */
int oracle_adder_fp64(const oracle_metadata *metadata, int nr_subgrids, int S, int G, const float *subgrids,
                      double *grid);
int oracle_splitter_fp64(const oracle_metadata *metadata, int nr_subgrids, int S, int G, const double *grid,
                         double *subgrids);

#ifdef __cplusplus
}
#endif
