// Master-grid transform and w-stacking for gfx950.
//
// Master grid: complex64 [nr_layers][4][G][G].  Cell K of a row is at u = (K - G/2) / image_size (the adder puts
// subgrid cell k of a subgrid with corner x0 at K = x0 + k, and the subgrid FFT centres cell k at
// u_centre + (k - S/2) / image_size with u_centre = (x0 + S/2 - G/2) / image_size).  Image pixel X is at
// l = (X - G/2) image_size / G.  The transforms, in each dimension, are
//     FORWARD  (image -> uv):  G[K] = sum_X I[X] exp(+2 pi i (K - G/2)(X - G/2) / G)
//     BACKWARD (uv -> image):  I[X] = sum_K G[K] exp(-2 pi i (K - G/2)(X - G/2) / G)
// (the signs of the subgrid FFT, unnormalised).  Since (K - G/2)(X - G/2) = K X - (G/2)(K + X) + G^2/4, they are a
// plain DFT between two checkerboard sign flips (-1)^(x + y) when G is a multiple of 4.
//
// G = 1024 .. 16384 (powers of two): three passes over each plane, the sign flips folded into the first loads and
// the last stores, each pass a batched Stockham FFT (radix-16 stages in registers, exchanges through LDS, then one
// radix-2/4/8 stage); twiddles from a two-level LDS table of FP64-rounded values (rows) or from the hardware sine
// and cosine of exact fractions of a turn (columns).  Columns are split as G = N1 x N2 (n = N2 n1 + n2,
// k = k1 + N1 k2):
//   pass 1 (in place): for each n2, the N1-point transforms over n1 of rows N2 n1 + n2, times W_G^(n2 k1), back to
//           rows N2 k1 + n2;
//   pass 2 (to a plane-sized scratch): for each k1, the N2-point transforms over n2 of rows N2 k1 + n2, to rows
//           k1 + N1 k2;
//   pass 3 (scratch -> grid): the G-point transform of each row, the whole row in LDS.
// The column passes take tiles of 64 columns (one wave reads a 512-byte row segment), so every pass reads and
// writes whole cache lines: 3 x (read + write) of the grid, against 2 checkerboard passes plus the library
// transform otherwise.  All three kernels are persistent (as many workgroups as are resident); the column kernels
// load their next tile into registers while they transform the current one (their LDS barriers do not wait for
// those loads), the row kernel holds only its row, which keeps two workgroups on a CU.  Other sizes run the plain
// batched 2-D transform (launch_fft2d, over correlations and layers: LDS or register kernels for sides 16 .. 512,
// four-step columns for the other powers of two, the radix-3 three-pass set for 3 x 2^k from 3072, rocFFT for the
// rest) between two checkerboard passes.
//
// W-stacking.  A subgrid with w-layer z is gridded relative to w_z = (z + 0.5) w_step (in wavelengths), so layer
// z of the dirty image carries exp(+2 pi i w_z n(l, m)) relative to the true one, n = 1 - sqrt(1 - l^2 - m^2)
// written as (l^2 + m^2) / (1 + sqrt(1 - l^2 - m^2)) (gridding/inc/math.hpp:29-37).  to_image sums the layers
// after multiplying each by exp(-2 pi i w_z n); from_image is its adjoint.  The phases are reduced in FP64.  Either
// as a pass of its own over the image-domain layers, or (three-pass sizes) inside the row pass, which then
// transforms an image row of every layer and applies the screens there (wstack_row_kernel).
#include <hipfft/hipfft.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "idgx_device.hpp"
#include "kernels.hpp"

namespace idgx
{
namespace
{
constexpr int kThreads = 256;

// Element (y, x) of every plane times (-1)^(x + y); two elements (x even, x + 1) per thread.
__global__ __launch_bounds__(kThreads) void checkerboard_kernel(float4 *grid, long long nr_pairs, int G)
{
    const long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= nr_pairs)
        return;
    const long long e = 2 * i; // first element of the pair
    const int x = static_cast<int>(e % G);
    const int y = static_cast<int>((e / G) % G);
    const float sg = ((x + y) & 1) ? -1.F : 1.F;
    float4 v = grid[i];
    grid[i] = make_float4(sg * v.x, sg * v.y, -sg * v.z, -sg * v.w);
}

__device__ __forceinline__ double image_n(int X, int Y, int G, double image_size)
{
    const double l = (X - 0.5 * G) * image_size / G, m = (Y - 0.5 * G) * image_size / G;
    const double r2 = l * l + m * m;
    return r2 >= 1.0 ? 1.0 : r2 / (1.0 + sqrt(1.0 - r2));
}

// exp(sign 2 pi i w n) with w = (z + 0.5) w_step, reduced to a fraction of a turn in FP64.
__device__ __forceinline__ cfloat wscreen(double n, int z, double w_step, double sign)
{
    double turns = (z + 0.5) * w_step * n;
    turns -= rint(turns);
    double sn, cs;
    sincospi(2.0 * sign * turns, &sn, &cs);
    return {static_cast<float>(cs), static_cast<float>(sn)};
}

// d_image[pol][Y][X] = sum_layers d_layers[layer][pol][Y][X] exp(-2 pi i w_z n(X, Y)).
__global__ __launch_bounds__(kThreads) void wstack_to_image_kernel(const float2 *layers, float2 *image, int G,
                                                                   int z_min, int nr_layers, double image_size,
                                                                   double w_step)
{
    const long long cell = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    const long long GG = static_cast<long long>(G) * G;
    if (cell >= GG)
        return;
    const int Y = static_cast<int>(cell / G), X = static_cast<int>(cell - static_cast<long long>(Y) * G);
    const double n = image_n(X, Y, G, image_size);
    cfloat acc[kCorr] = {{0.F, 0.F}, {0.F, 0.F}, {0.F, 0.F}, {0.F, 0.F}};
    for (int layer = 0; layer < nr_layers; ++layer)
    {
        const cfloat w = wscreen(n, z_min + layer, w_step, -1.0);
#pragma unroll
        for (int pol = 0; pol < kCorr; ++pol)
        {
            const float2 v = layers[(static_cast<long long>(layer) * kCorr + pol) * GG + cell];
            acc[pol].re = __builtin_fmaf(v.x, w.re, __builtin_fmaf(-v.y, w.im, acc[pol].re));
            acc[pol].im = __builtin_fmaf(v.x, w.im, __builtin_fmaf(v.y, w.re, acc[pol].im));
        }
    }
#pragma unroll
    for (int pol = 0; pol < kCorr; ++pol)
        image[pol * GG + cell] = make_float2(acc[pol].re, acc[pol].im);
}

// d_layers[layer][pol][Y][X] = d_image[pol][Y][X] exp(+2 pi i w_z n(X, Y)).
__global__ __launch_bounds__(kThreads) void wstack_from_image_kernel(const float2 *image, float2 *layers, int G,
                                                                     int z_min, int nr_layers, double image_size,
                                                                     double w_step)
{
    const long long cell = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    const long long GG = static_cast<long long>(G) * G;
    if (cell >= GG)
        return;
    const int Y = static_cast<int>(cell / G), X = static_cast<int>(cell - static_cast<long long>(Y) * G);
    const double n = image_n(X, Y, G, image_size);
    float2 v[kCorr];
#pragma unroll
    for (int pol = 0; pol < kCorr; ++pol)
        v[pol] = image[pol * GG + cell];
    for (int layer = 0; layer < nr_layers; ++layer)
    {
        const cfloat w = wscreen(n, z_min + layer, w_step, 1.0);
#pragma unroll
        for (int pol = 0; pol < kCorr; ++pol)
            layers[(static_cast<long long>(layer) * kCorr + pol) * GG + cell] =
                make_float2(v[pol].x * w.re - v[pol].y * w.im, v[pol].x * w.im + v[pol].y * w.re);
    }
}

// ---- Stockham passes ------------------------------------------------------------------------------------------
constexpr int kColTile = 64; // columns per column-pass tile

__device__ __forceinline__ float2 f2add(float2 a, float2 b)
{
    return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 f2sub(float2 a, float2 b)
{
    return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 f2mul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// cos and sin of 2 pi m / 16.
__device__ constexpr float cos16(int m)
{
    constexpr float t[16] = {1.F,
                             0.92387953251128674F,
                             0.70710678118654752F,
                             0.38268343236508977F,
                             0.F,
                             -0.38268343236508977F,
                             -0.70710678118654752F,
                             -0.92387953251128674F,
                             -1.F,
                             -0.92387953251128674F,
                             -0.70710678118654752F,
                             -0.38268343236508977F,
                             0.F,
                             0.38268343236508977F,
                             0.70710678118654752F,
                             0.92387953251128674F};
    return t[m & 15];
}
__device__ constexpr float sin16(int m)
{
    return cos16(m + 12);
}

// In-register DFT of length R (a power of two <= 16), natural order in and out:
// v[k] <- sum_n v[n] exp(SG 2 pi i n k / R), by radix-2 decimation in time.
template <int R, int SG> __device__ __forceinline__ void dft(float2 *v)
{
    if constexpr (R > 1)
    {
        float2 e[R / 2], o[R / 2];
#pragma unroll
        for (int i = 0; i < R / 2; ++i)
        {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        dft<R / 2, SG>(e);
        dft<R / 2, SG>(o);
#pragma unroll
        for (int k = 0; k < R / 2; ++k)
        {
            float2 t = o[k];
            if (4 * k == R)
                t = make_float2(-SG * t.y, SG * t.x);
            else if (k != 0)
                t = f2mul(make_float2(cos16(16 * k / R), SG * sin16(16 * k / R)), t);
            v[k] = f2add(e[k], t);
            v[k + R / 2] = f2sub(e[k], t);
        }
    }
}

enum FftPass
{
    kPassCol1 = 0,
    kPassCol2 = 1
};

struct FftArgs
{
    const float2 *src; // nplanes planes of gy rows of gx: column pass 1 input (the same planes as grid when in place)
    float2 *grid;      // column pass 1 output, column pass 2 input, row pass output
    float2 *scratch;   // nplanes planes
    int gx, gy;        // row length, column length
    int N1, N2;        // gy = N1 x N2
    int log_gy, log_tiles; // log2(gy), log2(gx / kColTile)
    int log_n1, log_n2;
    int nplanes;       // planes of this launch
    int centred;       // 1: the (-1)^(x + y) sign flips of the centred transform, 0: a plain DFT
    int ct0, nct;      // column passes: column tiles [ct0, ct0 + nct) of every row set (all gx / kColTile: dense)
    int cx0, cx1;      // row pass: scratch columns outside [cx0, cx1) are read as zero (0, gx: dense)
    int ry0, nry;      // row pass: rows [ry0, ry0 + nry) of each plane (0, gy: dense)
    int rows_first;    // 1: the row pass runs first, in place (scratch == grid), and takes the input signs
    int iy0, iy1;      // column pass 1: input rows outside [iy0, iy1) are zero and not read (0, gy: dense)
    // Column pass 2 stores row y, column x of plane pl at scratch[pl splane + y spitch + x - sx0]: (gx gy, gx, 0)
    // for whole planes, (gy nc, nc, first column) for a box of nc columns stored compactly.
    int spitch, sx0;
    long long splane;
};

constexpr int ilog2(int n)
{
    return n <= 1 ? 0 : 1 + ilog2(n / 2);
}

// W_G^q = exp(SG 2 pi i q / G) for 0 <= q < G (G a power of two) from v_cos / v_sin on the exact fraction q / G of a
// turn (absolute error <= 1.1e-7, measured on the device): no table and no memory access inside the stages.
template <int SG> __device__ __forceinline__ float2 twiddle(int q, int log_g)
{
    const float turns = __builtin_ldexpf(static_cast<float>(q), -log_g);
    return make_float2(__builtin_amdgcn_cosf(turns), SG * __builtin_amdgcn_sinf(turns));
}

// LDS exchanges.  Slots hold one float: the exchange sends the real parts, then the imaginary parts, through the
// same array, which halves the LDS a workgroup holds.  Column tiles interleave their kColTile transforms (a wave
// touches one row, so no two lanes share a bank); rows pad one slot per 16 so the radix-16 scatter is
// conflict-free.  The exchanges use lds_barrier(), so a prefetch of the next tile stays in flight across them.

// Row pass, persistent: each workgroup transforms rows blockIdx.x, blockIdx.x + gridDim.x, ... of one plane
// (scratch -> grid), L = G points per row, one radix-16 butterfly per thread and stage.  Twiddles
// W_L^q = exp(SG 2 pi i q / L) come from a two-level table in LDS, formed where they are used, so no stage waits on
// global memory and none holds twiddles in registers across rows.  A thread holds only its row's 16 points (no
// prefetch of the next row), which keeps the kernel within the 64 VGPRs that let two workgroups share a CU (at
// L = 16384: 2 x 1024 threads, 2 x 72 KB of LDS); each hides the other's loads and barriers.  Measured against the
// form that loaded the next row during the transform (97-109 VGPRs): 3-4 % faster at 1024 .. 4096, 0.5 % at 8192,
// 3-6 % at 16384, and 10 % on the boxed transform to the image, whose row reads are mostly skipped.
// The second __launch_bounds__ argument is HIP's minimum number of waves per SIMD (amdgpu-waves-per-eu): two
// workgroups of L / 16 threads per CU put max(2, L / 2048) waves on each SIMD, which caps the kernel at 512 / that
// many VGPRs (64 at L = 16384), so the two-workgroup occupancy described above is enforced, not incidental.  The
// plain rows of 8192 points take four workgroups per CU (8 waves per SIMD, 64 VGPRs), which they fit without
// spilling; the R3 form there would spill and keeps two.
template <int L, bool R3> constexpr int fft_row_waves_per_simd()
{
    return L == 8192 && !R3 ? 8 : L / 2048 > 2 ? L / 2048 : 2;
}

// Twiddles W_L^q = exp(SG 2 pi i q / L) of the row transforms as t_hi[q / 64] t_lo[q % 64]: two LDS reads and a
// complex product instead of a sine and a cosine at quarter rate.  The L / 64 + 64 entries are rounded from FP64
// values once per workgroup of L / 16 threads (the caller synchronises before use).
template <int L, int SG> __device__ __forceinline__ void row_twiddle_tables(float2 (&t_hi)[L / 64], float2 (&t_lo)[64])
{
    for (int i = threadIdx.x; i < L / 64 + 64; i += L / 16)
    {
        const int qq = i < L / 64 ? 64 * i : i - L / 64;
        double sn, cs;
        sincospi(2.0 * qq / L, &sn, &cs);
        const float2 w = make_float2(static_cast<float>(cs), static_cast<float>(SG * sn));
        if (i < L / 64)
            t_hi[i] = w;
        else
            t_lo[i - L / 64] = w;
    }
}

// The L-point transform of one row held by the workgroup's L / 16 threads, 16 values each: thread j0 holds elements
// j0 + r L / 16 in v[r].  Stages: radix 16 (L = 16^a x {1, 2, 4, 8}, a >= 2) with an LDS exchange through buf
// (L + L / 16 floats) after each, then a radix-kRl stage when L is not a power of 16.  Output element k, the idx-th
// (0 .. 15, a compile-time index after unrolling) of this thread's, goes to emit(idx, k, value).  The first exchange
// waits for every thread's reads of buf from the previous call.
template <int L, int SG, class Emit>
__device__ __forceinline__ void row_transform(float2 (&v)[16], int j0, float (&buf)[L + L / 16],
                                              const float2 (&t_hi)[L / 64], const float2 (&t_lo)[64], Emit &&emit)
{
    constexpr int kLog = ilog2(L);
    constexpr int kStages16 = kLog / 4;
    constexpr int kRl = 1 << (kLog % 4);
    constexpr int kQ = L / 16;
    static_assert(kStages16 >= 2, "L = 16^a x {1, 2, 4, 8}, a >= 2");
    // LDS slot of element i is i + i / 16.  Every access below is a per-thread base plus a compile-time offset:
    // for the elements j + t kQ (kQ a multiple of 16) the slot is slot(j) + t (kQ + kQ / 16).
    constexpr int kQs = kQ + kQ / 16;
    constexpr int kRs = L / kRl + L / (16 * kRl);
    auto tw = [&](int qq) {
        qq &= L - 1;
        return f2mul(t_hi[qq >> 6], t_lo[qq & 63]);
    };
    int Ns = 1;
#pragma unroll
    for (int st = 0; st < kStages16; ++st)
    {
        if (st > 0)
        {
            const int m = (j0 & (Ns - 1)) * (L / (16 * Ns));
#pragma unroll
            for (int r = 1; r < 16; ++r)
                v[r] = f2mul(v[r], tw(m * r));
        }
        dft<16, SG>(v);
        if (st == kStages16 - 1 && kRl == 1)
        {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                emit(r, j0 + r * Ns, v[r]);
        }
        else
        {
            // Output r of butterfly j0 goes to element d + r Ns: slot wb + r (Ns + Ns / 16), or 17 j0 + r when
            // Ns = 1.  The next stage reads elements j0 + t kQ (slot rb + t kQs), or for the radix-kRl stage
            // j0 + b kQ + r (L / kRl) (slot rb + b kQs + r kRs).
            const int d = (j0 / Ns) * Ns * 16 + (j0 & (Ns - 1));
            const int wb = d + (d >> 4), ws = Ns == 1 ? 1 : Ns + Ns / 16;
            const int rb = j0 + (j0 >> 4);
            const bool next16 = st + 1 < kStages16;
            auto rslot = [&](int t) { return next16 ? rb + t * kQs : rb + (t / kRl) * kQs + (t % kRl) * kRs; };
            float re[16];
            lds_barrier(); // the previous exchange's reads are done
#pragma unroll
            for (int r = 0; r < 16; ++r)
                buf[wb + r * ws] = v[r].x;
            lds_barrier();
#pragma unroll
            for (int t = 0; t < 16; ++t)
                re[t] = buf[rslot(t)];
            lds_barrier();
#pragma unroll
            for (int r = 0; r < 16; ++r)
                buf[wb + r * ws] = v[r].y;
            lds_barrier();
#pragma unroll
            for (int t = 0; t < 16; ++t)
                v[t] = make_float2(re[t], buf[rslot(t)]);
        }
        Ns *= 16;
    }
    if constexpr (kRl > 1)
    {
#pragma unroll
        for (int b = 0; b < 16 / kRl; ++b)
        {
            const int j = j0 + b * kQ;
            float2 u[kRl];
#pragma unroll
            for (int r = 0; r < kRl; ++r)
                u[r] = v[b * kRl + r];
#pragma unroll
            for (int r = 1; r < kRl; ++r)
                u[r] = f2mul(u[r], tw(j * r));
            dft<kRl, SG>(u);
#pragma unroll
            for (int r = 0; r < kRl; ++r)
                emit(b * kRl + r, j + r * Ns, u[r]);
        }
    }
}

// R3: the first pass of the N x N planes of fft_col3_kernel (N = 3 L = A.gx, not centred, dense): the L-point
// transforms of the sub-rows x = 3 x' + b (b = 0, 1, 2) of each input row (A.scratch), stored as segment b
// (columns [b L, b L + L)) of that row in A.grid.
template <int L, int SG, bool R3 = false>
__global__ __launch_bounds__(L / 16, (fft_row_waves_per_simd<L, R3>())) void fft_row_kernel(FftArgs A)
{
    constexpr int kQ = L / 16;
    __shared__ float buf[L + L / 16];
    __shared__ float2 t_hi[L / 64], t_lo[64];
    row_twiddle_tables<L, SG>(t_hi, t_lo);
    __syncthreads();

    const bool flip = A.centred != 0;
    // Work item i is row ry0 + i % nry of plane i / nry; planes are contiguous, so its global row is
    // plane * gy + row.  Scratch columns outside [cx0, cx1) are taken as zero without being read.  R3: item i is
    // sub-row b of row g (below).
    const int rows = R3 ? 3 * A.nplanes * A.gx : A.nplanes * A.nry;
    constexpr int kIn = R3 ? 3 : 1; // input element stride
    auto global_row = [&](int i) { return (i / A.nry) * A.gy + A.ry0 + i % A.nry; };
    const bool sign_in = flip && A.rows_first; // (-1)^(x + y) of the centred transform's input
    // R3 addresses a row as a workgroup-uniform row pointer plus the thread's 32-bit element offset (fewer VGPRs
    // than 64-bit element indices there); the plain rows keep 64-bit indices, whose constant parts fold into the
    // instructions' offsets.
    auto elem = [](int k) {
        if constexpr (R3)
            return static_cast<unsigned>(k);
        else
            return static_cast<size_t>(k);
    };
    auto load = [&](int gr, const float2 *in_row, int col) {
        float2 v = col >= A.cx0 && col < A.cx1 ? in_row[elem(kIn * col)] : make_float2(0.F, 0.F);
        if (sign_in && ((col + gr) & 1)) // gy is even: the row's parity within its plane is that of gr
            v = make_float2(-v.x, -v.y);
        return v;
    };
    // The L-point transform of the row at in_row (global row gr, for the input signs).
    auto row_fft = [&](int gr, const float2 *in_row, auto &&emit) {
        // An opaque copy of the thread index per row: what derives from it (twiddles, slots) is formed where it is
        // used instead of being hoisted out of the row loop and held in registers for every row.
        int j0 = threadIdx.x;
        asm volatile("" : "+v"(j0));
        float2 v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r)
            v[r] = load(gr, in_row, j0 + r * kQ);
        row_transform<L, SG>(v, j0, buf, t_hi, t_lo, emit);
    };
    for (int i = blockIdx.x; i < rows; i += gridDim.x)
    {
        if constexpr (R3)
        {
            // Item i is sub-row b = (i % 24) / 8 of row g = 8 (i / 24) + i % 8: with a grid that is a multiple of
            // 24, the three sub-rows of a row go to workgroups w, w + 8, w + 16, which share an XCD (workgroups are
            // dispatched to the 8 XCDs in turn) and run in step, so each cache line of the input row, a third of
            // which each sub-row reads, is fetched into one L2 once.
            const int N = A.gx;
            const int grp = i / 24, rem = i - 24 * grp;
            const int g = 8 * grp + (rem & 7), b = rem >> 3;
            float2 *out_row = A.grid + static_cast<size_t>(g) * N + static_cast<size_t>(b) * L;
            row_fft(g, A.scratch + static_cast<size_t>(g) * N + b,
                    [&](int, int k, float2 val) { out_row[elem(k)] = val; });
        }
        else
        {
            const int y = global_row(i);
            float2 *out_row = A.grid + static_cast<size_t>(y) * L;
            row_fft(y, A.scratch + static_cast<size_t>(y) * L, [&](int, int k, float2 val) {
                out_row[elem(k)] = (flip && (k & 1)) ? make_float2(-val.x, -val.y) : val; // k odd <=> j0 odd
            });
        }
    }
}

// Output idx of row_transform's emit is element j0 + row_out_t<L>(idx) L / 16 of the row: the 16 elements a thread
// loads, in the order its last stage produces them.
template <int L> constexpr int row_out_t(int idx)
{
    constexpr int kRl = 1 << (ilog2(L) % 4);
    return kRl == 1 ? idx : idx / kRl + (16 / kRl) * (idx % kRl);
}

// W-stacked row pass: the image side of the w-stacked master-grid transform, one image row (row y of one
// correlation plane) per item with every w-layer in it, so that neither the layers' image-domain planes nor a
// separate w-screen pass go through HBM.  Layer z (w_z = (z_min + z + 0.5) w_step) has the screen
// exp(SG 2 pi i w_z n) at pixel (X, Y), n = image_n(X, Y): n in FP64 once per item and point, w_z n reduced to a
// fraction of a turn in FP64, then v_cos / v_sin of that fraction (absolute error about 1e-7).  The transforms carry
// the row pass's share of the centred signs.
//   SG = -1, to the image (after the column passes): image[y][k] = sum_z screen_z(k) (-1)^k F_z[k], F_z the
//     transform of layer z's row y, read at rows_in + z layer_stride + y pitch as columns [cx0, cx1) (stored from
//     cx0 on; the other columns are zero);
//   SG = +1, from the image (before the column passes): for each layer z, the transform of
//     (-1)^(x + y) screen_z(x) image[y][x], with the output signs (-1)^k, goes to columns [cx0, cx1) of row
//     rows_out + z layer_stride + y pitch.
struct WsArgs
{
    const float2 *image_in;
    float2 *image_out;
    const float2 *rows_in;
    float2 *rows_out;
    long long layer_stride;
    int pitch, nr_layers, z_min, cx0, cx1;
    double image_size, w_step;
};

template <int L, int SG> __global__ __launch_bounds__(L / 16, (L >= 2048 ? 4 : 2)) void wstack_row_kernel(WsArgs W)
{
    constexpr int kQ = L / 16;
    __shared__ float buf[L + L / 16];
    __shared__ float2 t_hi[L / 64], t_lo[64];
    row_twiddle_tables<L, SG>(t_hi, t_lo);
    __syncthreads();
    for (int y = blockIdx.x; y < L; y += gridDim.x)
    {
        int j0 = threadIdx.x; // opaque per row, as in the row pass
        asm volatile("" : "+v"(j0));
        // The thread's points j0 + t kQ (t < 16), in and out: n once per row; (-1)^k = (-1)^j0 (kQ is even).
        double nn[16];
#pragma unroll
        for (int t = 0; t < 16; ++t)
            nn[t] = image_n(j0 + t * kQ, y, L, W.image_size);
        const bool odd = (j0 & 1) != 0;
        auto screen = [&](double wz, int t) {
            // An opaque copy: the screen is formed where it is used, not ahead of the transform (which would hold 16
            // screens in registers across it).
            double n = nn[t];
            asm volatile("" : "+v"(n));
            double turns = wz * n;
            turns -= __builtin_rint(turns);
            const float f = static_cast<float>(turns);
            return make_float2(__builtin_amdgcn_cosf(f), SG * __builtin_amdgcn_sinf(f));
        };
        if constexpr (SG < 0)
        {
            float2 acc[16];
#pragma unroll
            for (int t = 0; t < 16; ++t)
                acc[t] = make_float2(0.F, 0.F);
            for (int z = 0; z < W.nr_layers; ++z)
            {
                const double wz = (W.z_min + z + 0.5) * W.w_step;
                const float2 *in_row = W.rows_in + z * W.layer_stride + static_cast<long long>(y) * W.pitch;
                int jz = j0; // opaque per layer: what derives from it is formed per layer, not held across layers
                asm volatile("" : "+v"(jz));
                float2 v[16];
#pragma unroll
                for (int r = 0; r < 16; ++r)
                {
                    const int col = jz + r * kQ;
                    v[r] = col >= W.cx0 && col < W.cx1 ? in_row[col - W.cx0] : make_float2(0.F, 0.F);
                }
                row_transform<L, SG>(v, jz, buf, t_hi, t_lo, [&](int idx, int, float2 val) {
                    const int t = row_out_t<L>(idx);
                    const float2 w = screen(wz, t);
                    acc[t].x = __builtin_fmaf(val.x, w.x, __builtin_fmaf(-val.y, w.y, acc[t].x));
                    acc[t].y = __builtin_fmaf(val.x, w.y, __builtin_fmaf(val.y, w.x, acc[t].y));
                });
            }
            float2 *out = W.image_out + static_cast<size_t>(y) * L + j0;
#pragma unroll
            for (int t = 0; t < 16; ++t)
                out[t * kQ] = odd ? make_float2(-acc[t].x, -acc[t].y) : acc[t];
        }
        else
        {
            const float sg_in = ((j0 + y) & 1) ? -1.F : 1.F; // (-1)^(x + y), x = j0 + t kQ
            const float2 *in = W.image_in + static_cast<size_t>(y) * L + j0;
            float2 img[16];
#pragma unroll
            for (int t = 0; t < 16; ++t)
            {
                const float2 g = in[t * kQ];
                img[t] = make_float2(sg_in * g.x, sg_in * g.y);
            }
            for (int z = 0; z < W.nr_layers; ++z)
            {
                const double wz = (W.z_min + z + 0.5) * W.w_step;
                float2 v[16];
#pragma unroll
                for (int t = 0; t < 16; ++t)
                    v[t] = f2mul(img[t], screen(wz, t));
                float2 *out_row = W.rows_out + z * W.layer_stride + static_cast<long long>(y) * W.pitch;
                int jz = j0;
                asm volatile("" : "+v"(jz));
                row_transform<L, SG>(v, jz, buf, t_hi, t_lo, [&](int, int k, float2 val) {
                    if (k >= W.cx0 && k < W.cx1)
                        out_row[k] = odd ? make_float2(-val.x, -val.y) : val;
                });
            }
        }
    }
}

// Column passes, persistent: tiles of kColTile columns x one transform set (pass 1: rows N2 n1 + n2 for one n2; pass 2:
// rows N2 k1 + n2 for one k1), tile index = x tile + (G / kColTile) x (n2 or k1), taken blockIdx.x, blockIdx.x +
// gridDim.x, ...  The next tile is loaded into registers while the current one is transformed.  L = 16 kRl:
// a radix-16 stage from global memory, one LDS exchange, a radix-kRl stage to global memory.
// ROWS (column pass 1 of a box): input rows outside [iy0, iy1) are zero and not read.
template <int PASS, int L, int SG, bool ROWS = false>
__global__ __launch_bounds__(kColTile * L / 16) void fft_col_kernel(FftArgs A, int nr_tiles)
{
    constexpr int kLog = ilog2(L);
    constexpr int kRl = 1 << (kLog % 4);
    constexpr int kQ = L / 16;
    static_assert(kLog / 4 == 1 && kRl > 1, "L = 32, 64 or 128");
    __shared__ float buf[L * kColTile];
    const int c = threadIdx.x & (kColTile - 1);
    const int G = A.gx; // row stride
    // Tile index = (plane x Q + q) x nct + (column tile - ct0), q = n2 (pass 1) or k1 (pass 2), Q = N2 or N1.
    const int log_q = PASS == kPassCol1 ? A.log_n2 : A.log_n1;
    const size_t plane_elems = static_cast<size_t>(G) * A.gy;
    const float2 *in = PASS == kPassCol1 ? A.src : A.grid;
    auto decode = [&](int tile, int &pl, int &x, int &q) {
        const int rest = tile / A.nct;
        x = (A.ct0 + tile - rest * A.nct) * kColTile + c;
        q = rest & ((1 << log_q) - 1);
        pl = rest >> log_q;
    };
    auto element = [&](int tile, int i) {
        int pl, x, q;
        decode(tile, pl, x, q);
        const int y = PASS == kPassCol1 ? A.N2 * i + q : A.N2 * q + i;
        return pl * plane_elems + static_cast<size_t>(y) * G + x;
    };
    auto load = [&](int tile, int i) {
        const int q = (tile / A.nct) & ((1 << log_q) - 1);
        return static_cast<unsigned>(A.N2 * i + q - A.iy0) < static_cast<unsigned>(A.iy1 - A.iy0)
                   ? in[element(tile, i)]
                   : make_float2(0.F, 0.F);
    };

    float2 nxt[16];
    int tile = blockIdx.x;
    if (tile < nr_tiles)
    {
        const int j = threadIdx.x / kColTile;
#pragma unroll
        for (int r = 0; r < 16; ++r)
            nxt[r] = ROWS ? load(tile, j + r * kQ) : in[element(tile, j + r * kQ)];
    }
    for (; tile < nr_tiles; tile += gridDim.x)
    {
        int j0 = threadIdx.x / kColTile; // opaque per tile: keeps twiddles and slots out of registers across tiles
        asm volatile("" : "+v"(j0));
        int pl, x, q;
        decode(tile, pl, x, q);
        const size_t base = pl * plane_elems;
        const size_t sbase = pl * static_cast<size_t>(A.splane) - A.sx0;
        float2 v[16];
        // Pass 1 applies the input sign (-1)^(x + y), y = N2 n1 + n2 with N2 even: (-1)^(x + n2) for the whole thread.
        const float sg_in = PASS == kPassCol1 && A.centred && !A.rows_first && ((x + q) & 1) ? -1.F : 1.F;
#pragma unroll
        for (int r = 0; r < 16; ++r)
            v[r] = make_float2(sg_in * nxt[r].x, sg_in * nxt[r].y);
        const int tn = tile + gridDim.x;
        if (tn < nr_tiles)
        {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                nxt[r] = ROWS ? load(tn, j0 + r * kQ) : in[element(tn, j0 + r * kQ)];
        }
        dft<16, SG>(v);
        // Output r of butterfly j0 is element 16 j0 + r; the radix-kRl stage takes elements j0 + b kQ + r (L / kRl).
        auto rslot = [&](int t) { return (j0 + (t / kRl) * kQ + (t % kRl) * (L / kRl)) * kColTile + c; };
        float re[16];
        lds_barrier(); // the previous tile's reads are done
#pragma unroll
        for (int r = 0; r < 16; ++r)
            buf[(16 * j0 + r) * kColTile + c] = v[r].x;
        lds_barrier();
#pragma unroll
        for (int t = 0; t < 16; ++t)
            re[t] = buf[rslot(t)];
        lds_barrier();
#pragma unroll
        for (int r = 0; r < 16; ++r)
            buf[(16 * j0 + r) * kColTile + c] = v[r].y;
        lds_barrier();
#pragma unroll
        for (int t = 0; t < 16; ++t)
            v[t] = make_float2(re[t], buf[rslot(t)]);
#pragma unroll
        for (int b = 0; b < 16 / kRl; ++b)
        {
            const int j = j0 + b * kQ;
            float2 u[kRl];
#pragma unroll
            for (int r = 0; r < kRl; ++r)
                u[r] = v[b * kRl + r];
#pragma unroll
            for (int r = 1; r < kRl; ++r)
                u[r] = f2mul(u[r], twiddle<SG>(j * r, kLog));
            dft<kRl, SG>(u);
#pragma unroll
            for (int r = 0; r < kRl; ++r)
            {
                const int k = j + r * (L / kRl);
                if constexpr (PASS == kPassCol1)
                    A.grid[base + static_cast<size_t>(A.N2 * k + q) * G + x] =
                        f2mul(u[r], twiddle<SG>(q * k, A.log_gy));
                else
                {
                    // Only the rows the row pass reads ([ry0, ry0 + nry), all rows when dense).
                    const int row = q + A.N1 * k;
                    if (static_cast<unsigned>(row - A.ry0) < static_cast<unsigned>(A.nry))
                        A.scratch[sbase + static_cast<size_t>(row) * A.spitch + x] =
                            (A.centred && (row & 1)) ? make_float2(-u[r].x, -u[r].y) : u[r];
                }
            }
        }
    }
}

// y[r] = sum_q W_3^(r q) x_q, W_3 = exp(SG 2 pi i / 3) = -1/2 + SG i sqrt(3) / 2.
template <int SG> __device__ __forceinline__ void dft3(float2 a, float2 b, float2 c, float2 *y)
{
    constexpr float kS3 = 0.866025403784438647F * SG;
    const float2 s = f2add(b, c), d = f2sub(b, c);
    const float2 m = make_float2(a.x - 0.5F * s.x, a.y - 0.5F * s.y);
    y[0] = f2add(a, s);
    y[1] = make_float2(m.x - kS3 * d.y, m.y + kS3 * d.x);
    y[2] = make_float2(m.x + kS3 * d.y, m.y - kS3 * d.x);
}

// Column passes of N x N planes with N = 3 M (M = N1 N2 a power of two), after the R3 row pass, which leaves in
// segment b (columns [b M, b M + M)) of row y the M-point transforms R_b of the sub-rows x = 3 x' + b.  Along x
// (decimation in time),
//   X_x[y][k + M q] = sum_{b < 3} W_3^(b q) W_N^(b k) R_b[y][k]        (W_n = exp(SG 2 pi i / n)),
// and along y (decimation in frequency),
//   X[3 k + r][.] = sum_{y < M} W_M^(k y) t_r[y][.],   t_r[y] = W_N^(r y) sum_{q < 3} W_3^(r q) X_x[y + M q][.].
// The x-axis combination acts on whole columns and the y-axis steps on each column alone, so they commute: column
// pass 1 (in place) forms t_0, t_1, t_2 from rows y, y + M, y + 2 M of the row pass's output and runs the first
// four-step stage of their M-point transforms (as fft_col_kernel's pass 1), storing t_r's rows in row block r (rows
// [r M, r M + M)): a tile writes the rows it read.  Column pass 2 reads columns k, k + M, k + 2 M of a row block's
// rows, forms the x-axis combination, runs the second stage (as pass 2) on each of its three results q and stores
// them at columns [q M, q M + M) of output row 3 (k1 + N1 k2) + r.  Every pass reads and writes whole cache lines.
// No prefetch of the next tile: a thread already holds three sequences.
template <int PASS, int L, int SG>
__global__ __launch_bounds__(kColTile * L / 16) void fft_col3_kernel(FftArgs A, int nr_tiles)
{
    constexpr int kLog = ilog2(L);
    constexpr int kRl = 1 << (kLog % 4);
    constexpr int kQ = L / 16;
    static_assert(kLog / 4 == 1 && kRl > 1, "L = 32, 64 or 128");
    __shared__ float buf[L * kColTile];
    const int c = threadIdx.x & (kColTile - 1);
    const int N = A.gx, M = A.gy;
    const size_t block_elems = static_cast<size_t>(M) * N; // one row block: M rows of N
    const int log_q = PASS == kPassCol1 ? A.log_n2 : A.log_n1;
    const float inv_n = 1.F / static_cast<float>(N);
    auto rot = [&](float2 v, int n) { // v W_N^n for 0 <= n < N, from sin / cos of the fraction n / N of a turn
        const float t = SG * static_cast<float>(n) * inv_n;
        return f2mul(v, make_float2(__builtin_amdgcn_cosf(t), __builtin_amdgcn_sinf(t)));
    };
    for (int tile = blockIdx.x; tile < nr_tiles; tile += gridDim.x)
    {
        int j0 = threadIdx.x / kColTile; // opaque per tile: keeps twiddles and slots out of registers across tiles
        asm volatile("" : "+v"(j0));
        // Tile index = (blk x Q + q) x nct + column tile: pass 1 blk = plane, q = n2, nct = N / kColTile; pass 2 blk =
        // 3 plane + row block, q = k1, nct = M / kColTile.
        const int rest = tile / A.nct;
        const int x0 = (tile - rest * A.nct) * kColTile; // the tile's first column
        const int x = x0 + c;
        const int q = rest & ((1 << log_q) - 1);
        const int blk = rest >> log_q;
        // Element r of sequence qq: pass 1 row N2 (j0 + r kQ) + q + M qq, column x; pass 2 row N2 q + j0 + r kQ, column
        // x + M qq.  Addresses: a pointer that is uniform over the workgroup (tile, r, qq) plus the thread's 32-bit
        // offset (j0 rows of the pass's row step, c columns).
        const float2 *in = PASS == kPassCol1 ? A.src + 3 * static_cast<size_t>(blk) * block_elems
                                                   + static_cast<size_t>(q) * N + x0
                                             : A.grid + static_cast<size_t>(blk) * block_elems
                                                   + static_cast<size_t>(A.N2) * q * N + x0;
        const int row_step = PASS == kPassCol1 ? A.N2 : 1; // rows between consecutive elements of a sequence
        const unsigned toff = static_cast<unsigned>(j0 * row_step * N + c);
        float2 t[3][16];
#pragma unroll
        for (int qq = 0; qq < 3; ++qq)
#pragma unroll
            for (int r = 0; r < 16; ++r)
            {
                const size_t u = PASS == kPassCol1 ? static_cast<size_t>(row_step * r * kQ + M * qq) * N
                                                   : static_cast<size_t>(r * kQ) * N + static_cast<size_t>(M) * qq;
                t[qq][r] = (in + u)[toff];
            }
        // The radix-3 steps: pass 1 twiddles the butterfly's outputs by W_N^(r y), pass 2 its inputs by W_N^(b k)
        // (y, k < M, so 2 y, 2 k < N).
#pragma unroll
        for (int r = 0; r < 16; ++r)
        {
            float2 y3[3];
            if constexpr (PASS == kPassCol1)
            {
                const int n = A.N2 * (j0 + r * kQ) + q;
                dft3<SG>(t[0][r], t[1][r], t[2][r], y3);
                t[0][r] = y3[0];
                t[1][r] = rot(y3[1], n);
                t[2][r] = rot(y3[2], 2 * n);
            }
            else
            {
                dft3<SG>(t[0][r], rot(t[1][r], x), rot(t[2][r], 2 * x), y3);
                t[0][r] = y3[0];
                t[1][r] = y3[1];
                t[2][r] = y3[2];
            }
        }
#pragma unroll
        for (int s = 0; s < 3; ++s)
        {
            float2 *v = t[s];
            dft<16, SG>(v);
            // As fft_col_kernel: output r of butterfly j0 is element 16 j0 + r; the radix-kRl stage takes elements
            // j0 + b kQ + r (L / kRl).
            auto rslot = [&](int e) { return (j0 + (e / kRl) * kQ + (e % kRl) * (L / kRl)) * kColTile + c; };
            float re[16];
            lds_barrier(); // the previous sequence's reads are done
#pragma unroll
            for (int r = 0; r < 16; ++r)
                buf[(16 * j0 + r) * kColTile + c] = v[r].x;
            lds_barrier();
#pragma unroll
            for (int e = 0; e < 16; ++e)
                re[e] = buf[rslot(e)];
            lds_barrier();
#pragma unroll
            for (int r = 0; r < 16; ++r)
                buf[(16 * j0 + r) * kColTile + c] = v[r].y;
            lds_barrier();
#pragma unroll
            for (int e = 0; e < 16; ++e)
                v[e] = make_float2(re[e], buf[rslot(e)]);
#pragma unroll
            for (int b = 0; b < 16 / kRl; ++b)
            {
                const int j = j0 + b * kQ;
                float2 u[kRl];
#pragma unroll
                for (int r = 0; r < kRl; ++r)
                    u[r] = v[b * kRl + r];
#pragma unroll
                for (int r = 1; r < kRl; ++r)
                    u[r] = f2mul(u[r], twiddle<SG>(j * r, kLog));
                dft<kRl, SG>(u);
#pragma unroll
                for (int r = 0; r < kRl; ++r)
                {
                    // k = j0 + (b kQ + r L / kRl): the thread's j0 rows of the pass's output row step plus a
                    // uniform part.
                    const int k = j + r * (L / kRl);
                    const int ku = b * kQ + r * (L / kRl);
                    if constexpr (PASS == kPassCol1)
                    {
                        float2 *o = A.grid + (3 * static_cast<size_t>(blk) * M + s * M + A.N2 * ku + q) * N + x0;
                        o[static_cast<unsigned>(j0 * A.N2 * N + c)] = f2mul(u[r], twiddle<SG>(q * k, A.log_gy));
                    }
                    else
                    {
                        // Row block blk % 3 of plane blk / 3, k1 = q: output row 3 (q + N1 k) + blk % 3.
                        const int plane = blk / 3, rb = blk - 3 * plane;
                        float2 *o = A.scratch + (static_cast<size_t>(plane) * N + rb + 3 * (q + A.N1 * ku)) * N +
                                    s * M + x0;
                        o[static_cast<unsigned>(3 * j0 * A.N1 * N + c)] = u[r];
                    }
                }
            }
        }
    }
}

// Persistent grid: as many workgroups as are resident at once (at most one per tile or row).
hipError_t resident_blocks(const void *fn, int threads, int work, int &blocks)
{
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0);
    blocks = std::max(1, std::min(work, cus * std::max(per_cu, 1)));
    return e;
}

template <int PASS, int L, int SG, bool ROWS>
hipError_t launch_cols_kernel(const FftArgs &A, int nr_tiles, hipStream_t st)
{
    constexpr int threads = kColTile * L / 16;
    int blocks = 0;
    const hipError_t e = resident_blocks(reinterpret_cast<const void *>(&fft_col_kernel<PASS, L, SG, ROWS>), threads,
                                         nr_tiles, blocks);
    if (e != hipSuccess)
        return e;
    fft_col_kernel<PASS, L, SG, ROWS><<<blocks, threads, 0, st>>>(A, nr_tiles);
    return hipGetLastError();
}

template <int PASS, int L, int SG> hipError_t launch_cols(const FftArgs &A, int nr_tiles, hipStream_t st)
{
    if constexpr (PASS == kPassCol1)
    {
        if (A.iy0 > 0 || A.iy1 < A.gy)
            return launch_cols_kernel<PASS, L, SG, true>(A, nr_tiles, st);
    }
    return launch_cols_kernel<PASS, L, SG, false>(A, nr_tiles, st);
}

template <int PASS, int SG> hipError_t launch_col_pass(const FftArgs &A, int L, int nr_tiles, hipStream_t st)
{
    switch (L)
    {
    case 32:
        return launch_cols<PASS, 32, SG>(A, nr_tiles, st);
    case 64:
        return launch_cols<PASS, 64, SG>(A, nr_tiles, st);
    case 128:
        return launch_cols<PASS, 128, SG>(A, nr_tiles, st);
    default:
        return hipErrorInvalidValue;
    }
}

template <int L, int SG> hipError_t launch_rows(const FftArgs &A, hipStream_t st)
{
    int blocks = 0;
    const hipError_t e =
        resident_blocks(reinterpret_cast<const void *>(&fft_row_kernel<L, SG>), L / 16, A.nplanes * A.nry, blocks);
    if (e != hipSuccess)
        return e;
    fft_row_kernel<L, SG><<<blocks, L / 16, 0, st>>>(A);
    return hipGetLastError();
}

template <int SG> hipError_t launch_row_pass(const FftArgs &A, hipStream_t st)
{
    switch (A.gx)
    {
    case 1024:
        return launch_rows<1024, SG>(A, st);
    case 2048:
        return launch_rows<2048, SG>(A, st);
    case 4096:
        return launch_rows<4096, SG>(A, st);
    case 8192:
        return launch_rows<8192, SG>(A, st);
    case 16384:
        return launch_rows<16384, SG>(A, st);
    default:
        return hipErrorInvalidValue;
    }
}

template <int L, int SG> hipError_t launch_wstack_rows(const WsArgs &W, hipStream_t st)
{
    int blocks = 0;
    const hipError_t e =
        resident_blocks(reinterpret_cast<const void *>(&wstack_row_kernel<L, SG>), L / 16, L, blocks);
    if (e != hipSuccess)
        return e;
    wstack_row_kernel<L, SG><<<blocks, L / 16, 0, st>>>(W);
    return hipGetLastError();
}

template <int SG> hipError_t launch_wstack_row_pass(const WsArgs &W, int G, hipStream_t st)
{
    switch (G)
    {
    case 1024:
        return launch_wstack_rows<1024, SG>(W, st);
    case 2048:
        return launch_wstack_rows<2048, SG>(W, st);
    case 4096:
        return launch_wstack_rows<4096, SG>(W, st);
    case 8192:
        return launch_wstack_rows<8192, SG>(W, st);
    case 16384:
        return launch_wstack_rows<16384, SG>(W, st);
    default:
        return hipErrorInvalidValue;
    }
}

bool stockham_size(int G)
{
    return G >= 1024 && G <= 16384 && (G & (G - 1)) == 0;
}

// Scratch per (device, user, size), created on first use and kept for the life of the process.  The user tells
// apart buffers that are live at the same time: the three-pass transform's (user 0) and the radix-3 decimation's
// sub-planes (user 1), which the three-pass transform of those sub-planes must not share.
std::mutex g_scratch_mutex;
std::map<std::tuple<int, int, long long>, float2 *> g_scratch;

hipError_t fft_scratch(long long elements, float2 *&out, int user = 0)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    const auto key = std::make_tuple(dev, user, elements);
    auto it = g_scratch.find(key);
    if (it != g_scratch.end())
    {
        out = it->second;
        return hipSuccess;
    }
    float2 *p = nullptr;
    e = hipMalloc(&p, static_cast<size_t>(elements) * sizeof(float2));
    if (e != hipSuccess)
        return e;
    g_scratch[key] = p;
    out = p;
    return hipSuccess;
}

// Scratch of the allocation-free transform (idgx_grid_fft_noalloc): a file-scope array of fixed size, one copy per
// GPU for the life of the process, big enough for one group of planes as stockham_fft2d groups them (64 MB, or one
// plane when a plane is larger), i.e. planes of up to 4096 x 4096 points.  Its device address comes from
// hipGetSymbolAddress, taken once per device with that device current; host code never uses the array's name as a
// device address.  Every call on a GPU shares that GPU's array, so calls that overlap on one GPU race; no call needs
// its contents to start at zero (column pass 2 writes every element that the row pass reads).
constexpr long long kStaticScratchBytes = 128LL << 20;
static __device__ __align__(256) float2 scratch[kStaticScratchBytes / sizeof(float2)];

constexpr int kMaxDevices = 64;
std::mutex g_static_scratch_mutex;
float2 *g_static_scratch[kMaxDevices] = {};

hipError_t static_scratch(float2 *&out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev < 0 || dev >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_static_scratch_mutex);
    if (!g_static_scratch[dev])
    {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_static_scratch[dev] = static_cast<float2 *>(p);
    }
    out = g_static_scratch[dev];
    return hipSuccess;
}

// Planes of gy rows of gx points (both powers of two in 1024 .. 16384), src -> dst (in place when equal).  Each
// launch takes a group of planes, so that small planes still fill the chip; the scratch holds one group, at most
// kScratchBytes (or one plane when a plane is larger).  Kept small enough that the scratch written by column pass 2
// is still in the 256 MB Infinity Cache when the row pass reads it (measured at G = 4096: one 128 MB plane per
// group 0.57 ms per direction, two planes 0.64 ms).
constexpr long long kScratchBytes = 64LL << 20;

// Copies rows [y0, y0 + ny) x columns [c0, c0 + nc) of every plane from src to dst (planes of G x G).
__global__ __launch_bounds__(kThreads) void copy_box_kernel(const float2 *src, float2 *dst, int G, int y0, int ny,
                                                            int c0, int nc, long long total)
{
    const long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= total)
        return;
    const long long r = i / nc; // plane * ny + row
    const int c = c0 + static_cast<int>(i - r * nc);
    const long long pl = r / ny;
    const size_t e = (static_cast<size_t>(pl) * G + y0 + (r - pl * ny)) * G + c;
    dst[e] = src[e];
}

// Restriction of a transform to a box of cells [x0, x1) x [y0, y1) of every plane.  kSparseInput: the input is
// zero outside the box, so the column passes take only the column tiles that meet [x0, x1) and the row pass reads
// only those columns (zeros elsewhere); the whole output is produced.  kRowsOut: only the box's cells (within the
// column tiles around it) of the output are produced: the row pass runs first, in place and dense, then the column
// passes take only the box's column tiles and column pass 2 stores only its rows into the scratch, from which they
// are copied back (one dense pass instead of two; the rounding differs from the column-first order).
enum BoxMode
{
    kFullTransform = 0,
    kSparseInput = 1,
    kRowsOut = 2
};
struct FftBox
{
    int mode = kFullTransform;
    int x0 = 0, x1 = 0, y0 = 0, y1 = 0;
    bool rows_done = false; // kRowsOut: the planes already hold their row transforms (the w-stacked row pass)
};

// Dense three-pass arguments for planes of gy rows of gx points (pointers and plane count left to the caller).
FftArgs fft_args(int gx, int gy, bool centred)
{
    FftArgs A{};
    A.gx = gx;
    A.gy = gy;
    A.log_gy = ilog2(gy);
    A.log_n1 = ilog2(gy) / 2;
    A.log_n2 = ilog2(gy) - A.log_n1;
    A.N1 = 1 << A.log_n1;
    A.N2 = 1 << A.log_n2;
    A.log_tiles = ilog2(gx / kColTile);
    A.centred = centred ? 1 : 0;
    A.ct0 = 0;
    A.nct = gx / kColTile;
    A.cx0 = 0;
    A.cx1 = gx;
    A.ry0 = 0;
    A.nry = gy;
    A.rows_first = 0;
    A.splane = static_cast<long long>(gx) * gy;
    A.spitch = gx;
    A.sx0 = 0;
    A.iy0 = 0;
    A.iy1 = gy;
    return A;
}

// no_alloc: the scratch is the file-scope array (the caller has checked that a group fits it) instead of a plane
// allocated on first use.
template <int SG>
hipError_t stockham_fft2d(const float2 *src, float2 *dst, int gx, int gy, long long planes, bool centred,
                          hipStream_t st, const FftBox &box = FftBox(), bool no_alloc = false)
{
    FftArgs A = fft_args(gx, gy, centred);
    const long long plane = static_cast<long long>(gx) * gy;
    const long long plane_bytes = plane * static_cast<long long>(sizeof(float2));
    const long long group = std::max(1LL, std::min(planes, kScratchBytes / plane_bytes));
    hipError_t e = hipSuccess;
    if (no_alloc)
    {
        if (group * plane * static_cast<long long>(sizeof(float2)) > kStaticScratchBytes)
            return hipErrorInvalidValue;
        e = static_scratch(A.scratch);
    }
    else
        e = fft_scratch(group * plane, A.scratch);
    if (e != hipSuccess)
        return e;
    if (box.mode == kSparseInput)
    {
        const int t0 = box.x0 / kColTile, t1 = (box.x1 + kColTile - 1) / kColTile;
        A.ct0 = t0;
        A.nct = std::max(0, t1 - t0);
        A.cx0 = t0 * kColTile;
        A.cx1 = std::max(A.cx0, t1 * kColTile);
        A.iy0 = box.y0;
        A.iy1 = std::max(box.y0, box.y1);
    }
    A.rows_first = 0;
    if (box.mode == kRowsOut)
    {
        const int t0 = box.x0 / kColTile, t1 = (box.x1 + kColTile - 1) / kColTile;
        const int ny = std::max(0, box.y1 - box.y0);
        if (src != dst)
            return hipErrorInvalidValue; // the grid transforms run in place
        if (t1 <= t0 || ny == 0)
            return hipSuccess; // nothing to produce
        for (long long p = 0; p < planes && e == hipSuccess; p += group)
        {
            A.nplanes = static_cast<int>(std::min(group, planes - p));
            float2 *g = dst + static_cast<size_t>(p) * plane;
            // 1. Rows, in place and dense, with the input signs.
            if (!box.rows_done)
            {
                FftArgs R = A;
                R.scratch = g;
                R.grid = g;
                R.rows_first = 1;
                e = launch_row_pass<SG>(R, st);
            }
            // 2. Columns of the box's column tiles; column pass 2 stores the box's rows into the scratch.
            FftArgs C = A;
            C.src = g;
            C.grid = g;
            C.rows_first = 1;
            C.ct0 = t0;
            C.nct = t1 - t0;
            C.ry0 = box.y0;
            C.nry = ny;
            const long long col_sets = static_cast<long long>(C.nplanes) * C.nct;
            if (e == hipSuccess)
                e = launch_col_pass<kPassCol1, SG>(C, C.N1, static_cast<int>(col_sets * C.N2), st);
            if (e == hipSuccess)
                e = launch_col_pass<kPassCol2, SG>(C, C.N2, static_cast<int>(col_sets * C.N1), st);
            // 3. The box back into the grid.
            const int nc = (t1 - t0) * kColTile;
            const long long total = static_cast<long long>(C.nplanes) * ny * nc;
            if (e == hipSuccess)
            {
                copy_box_kernel<<<static_cast<unsigned>((total + kThreads - 1) / kThreads), kThreads, 0, st>>>(
                    C.scratch, g, gx, box.y0, ny, t0 * kColTile, nc, total);
                e = hipGetLastError();
            }
        }
        return e;
    }
    for (long long p = 0; p < planes && e == hipSuccess; p += group)
    {
        A.nplanes = static_cast<int>(std::min(group, planes - p));
        A.src = src + static_cast<size_t>(p) * plane;
        A.grid = dst + static_cast<size_t>(p) * plane;
        const long long col_sets = static_cast<long long>(A.nplanes) * A.nct;
        if (col_sets > 0)
        {
            e = launch_col_pass<kPassCol1, SG>(A, A.N1, static_cast<int>(col_sets * A.N2), st);
            if (e == hipSuccess)
                e = launch_col_pass<kPassCol2, SG>(A, A.N2, static_cast<int>(col_sets * A.N1), st);
        }
        if (e == hipSuccess && A.nry > 0)
            e = launch_row_pass<SG>(A, st);
    }
    return e;
}

// hipFFT plans per (device, nx, ny, batch, type), created on first use and kept for the life of the process.
std::mutex g_plan_mutex;
std::map<std::tuple<int, int, int, long long, int>, hipfftHandle> g_plans;

hipError_t fft_plan(int nx, int ny, long long batch, hipfftType type, hipfftHandle &out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    std::lock_guard<std::mutex> lock(g_plan_mutex);
    const auto key = std::make_tuple(dev, nx, ny, batch, static_cast<int>(type));
    auto it = g_plans.find(key);
    if (it != g_plans.end())
    {
        out = it->second;
        return hipSuccess;
    }
    hipfftHandle h;
    int n[2] = {ny, nx}; // rows, row length
    if (hipfftPlanMany(&h, 2, n, nullptr, 1, nx * ny, nullptr, 1, nx * ny, type, static_cast<int>(batch)) !=
        HIPFFT_SUCCESS)
        return hipErrorUnknown;
    g_plans[key] = h;
    out = h;
    return hipSuccess;
}

hipError_t checkerboard(void *grid, long long planes, int G, hipStream_t st)
{
    const long long pairs = planes * G * (G / 2);
    const long long blocks = (pairs + kThreads - 1) / kThreads;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    checkerboard_kernel<<<static_cast<unsigned>(blocks), kThreads, 0, st>>>(static_cast<float4 *>(grid), pairs, G);
    return hipGetLastError();
}
} // namespace

bool grid_fft_supported(int G)
{
    return G >= 4 && G % 4 == 0;
}

// Zeroes columns [c0, c0 + width) of rows [y0, y0 + ny) of every plane; rows = planes x ny (width even: two
// cells per 16-byte store).
__global__ __launch_bounds__(kThreads) void clear_band_kernel(float4 *grid, long long rows, int G, int y0, int ny,
                                                              int c0, int width)
{
    const long long half = width / 2;
    const long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= rows * half)
        return;
    const long long r = i / half, c = c0 / 2 + (i - r * half);
    const long long plane = r / ny, y = y0 + (r - plane * ny);
    grid[(plane * G + y) * (G / 2) + c] = make_float4(0.F, 0.F, 0.F, 0.F);
}

hipError_t launch_grid_clear_box(void *grid, int G, long long nr_layers, const int *box, hipStream_t st)
{
    const long long planes = nr_layers * kCorr;
    int c0 = 0, c1 = G, y0 = 0, y1 = G;
    if (stockham_size(G))
    {
        // The cells the boxed BACKWARD transform reads: the box's rows of the column tiles around [x0, x1).
        c0 = box[0] / kColTile * kColTile;
        c1 = std::max(c0, (box[1] + kColTile - 1) / kColTile * kColTile);
        y0 = box[2];
        y1 = std::max(y0, box[3]);
    }
    const long long ny = y1 - y0;
    const long long items = planes * ny * ((c1 - c0) / 2);
    if (items == 0)
        return hipSuccess;
    const long long blocks = (items + kThreads - 1) / kThreads;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    clear_band_kernel<<<static_cast<unsigned>(blocks), kThreads, 0, st>>>(
        static_cast<float4 *>(grid), planes * ny, G, y0, static_cast<int>(ny), c0, c1 - c0);
    return hipGetLastError();
}

hipError_t launch_grid_fft(void *grid, int G, long long nr_layers, bool forward, hipStream_t st, const int *box)
{
    const long long planes = nr_layers * kCorr;
    if (planes > 0x7fffffffLL || static_cast<long long>(G) * G > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    float2 *g = static_cast<float2 *>(grid);
    if (stockham_size(G))
    {
        // With a box: the image direction needs only the box's rows; the uv grid is zero outside the box.
        FftBox b;
        if (box)
        {
            b.mode = forward ? kRowsOut : kSparseInput;
            b.x0 = box[0];
            b.x1 = box[1];
            b.y0 = box[2];
            b.y1 = box[3];
            if (forward)
            {
                // In plane-sized reads + writes, the boxed transform to uv costs 1 (dense rows) + fx (column pass 1)
                // + fx (1 + fy) / 2 (column pass 2) + fx fy (copy back), fx and fy being the shares of the grid's
                // columns (whole tiles) and rows that the box takes; the full transform costs 3 and produces every
                // cell.  A box that covers most of the grid takes the full transform (measured at G = 4096 with a
                // box of 4095 x 4005 cells: 0.69 ms boxed, 0.55 ms full).
                const int t0 = b.x0 / kColTile, t1 = (b.x1 + kColTile - 1) / kColTile;
                const double fx = static_cast<double>(t1 - t0) * kColTile / G;
                const double fy = static_cast<double>(b.y1 - b.y0) / G;
                if (1.5 * fx * (1.0 + fy) >= 2.0)
                    b.mode = kFullTransform;
            }
        }
        return forward ? stockham_fft2d<1>(g, g, G, G, planes, true, st, b)
                       : stockham_fft2d<-1>(g, g, G, G, planes, true, st, b);
    }
    // Other sizes: the plain 2-D transform in place between the two sign flips (this library's kernels for sides
    // 16 .. 512, the other powers of two and 3 x 2^k from 3072; rocFFT for the rest).  FORWARD has the + sign.
    hipError_t e = checkerboard(grid, planes, G, st);
    if (e != hipSuccess)
        return e;
    e = launch_fft2d(grid, grid, G, G, planes, forward ? 1 : -1, false, st);
    if (e != hipSuccess)
        return e;
    return checkerboard(grid, planes, G, st);
}

// ---- sides 3 x 2^k (3072 .. 24576): one radix-3 step per axis folded into the three passes ---------------------
//
// N = 3 M: the R3 row pass (M-point transforms of the sub-rows x = 3 x' + b, input -> scratch), column pass 1 of
// fft_col3_kernel (a radix-3 step along y, then the first four-step stage; in place) and column pass 2 (the radix-3
// step along x, then the second stage; scratch -> output).  Three passes over HBM, as for the power-of-two sides;
// the scratch is one plane.
template <int PASS, int L, int SG> hipError_t launch_cols3(const FftArgs &A, int nr_tiles, hipStream_t st)
{
    constexpr int threads = kColTile * L / 16;
    int blocks = 0;
    const hipError_t e = resident_blocks(reinterpret_cast<const void *>(&fft_col3_kernel<PASS, L, SG>), threads,
                                         nr_tiles, blocks);
    if (e != hipSuccess)
        return e;
    fft_col3_kernel<PASS, L, SG><<<blocks, threads, 0, st>>>(A, nr_tiles);
    return hipGetLastError();
}

template <int PASS, int SG> hipError_t launch_col3_pass(const FftArgs &A, int L, int nr_tiles, hipStream_t st)
{
    switch (L)
    {
    case 32:
        return launch_cols3<PASS, 32, SG>(A, nr_tiles, st);
    case 64:
        return launch_cols3<PASS, 64, SG>(A, nr_tiles, st);
    case 128:
        return launch_cols3<PASS, 128, SG>(A, nr_tiles, st);
    default:
        return hipErrorInvalidValue;
    }
}

template <int L, int SG> hipError_t launch_rows3(const FftArgs &A, hipStream_t st)
{
    const int items = 3 * A.nplanes * A.gx; // a multiple of 24: M >= 1024
    int blocks = 0;
    const hipError_t e =
        resident_blocks(reinterpret_cast<const void *>(&fft_row_kernel<L, SG, true>), L / 16, items, blocks);
    if (e != hipSuccess)
        return e;
    blocks = std::max(24, blocks / 24 * 24); // the three sub-rows of a row on one XCD, in step
    fft_row_kernel<L, SG, true><<<blocks, L / 16, 0, st>>>(A);
    return hipGetLastError();
}

template <int SG> hipError_t launch_row3_pass(const FftArgs &A, hipStream_t st)
{
    switch (A.gy)
    {
    case 1024:
        return launch_rows3<1024, SG>(A, st);
    case 2048:
        return launch_rows3<2048, SG>(A, st);
    case 4096:
        return launch_rows3<4096, SG>(A, st);
    case 8192:
        return launch_rows3<8192, SG>(A, st);
    default:
        return hipErrorInvalidValue;
    }
}

// 3 M with M a three-pass size whose N x N plane stays within 2^31 - 1 points (M <= 8192).
bool three_times_stockham(int n)
{
    return n % 3 == 0 && stockham_size(n / 3) && n / 3 <= 8192;
}

template <int SG>
hipError_t radix3_fft2d(const float2 *src, float2 *dst, int N, long long planes, hipStream_t st)
{
    const int M = N / 3;
    const long long plane = static_cast<long long>(N) * N;
    FftArgs A;
    hipError_t e = fft_scratch(plane, A.scratch, 1);
    if (e != hipSuccess)
        return e;
    A.gx = N;
    A.gy = M;
    A.log_gy = ilog2(M);
    A.log_n1 = A.log_gy / 2;
    A.log_n2 = A.log_gy - A.log_n1;
    A.N1 = 1 << A.log_n1;
    A.N2 = 1 << A.log_n2;
    A.log_tiles = 0;
    A.nplanes = 1;
    A.centred = 0;
    A.ct0 = 0;
    A.cx0 = 0;
    A.cx1 = M; // the row pass reads whole segments
    A.ry0 = 0;
    A.nry = M;
    A.rows_first = 0;
    A.splane = plane; // (fft_col3_kernel stores with its own layout)
    A.spitch = N;
    A.sx0 = 0;
    A.iy0 = 0;
    A.iy1 = M;
    float2 *sc = A.scratch;
    for (long long pl = 0; pl < planes && e == hipSuccess; ++pl)
    {
        FftArgs R = A; // rows: input -> scratch
        R.scratch = const_cast<float2 *>(src + pl * plane);
        R.grid = sc;
        e = launch_row3_pass<SG>(R, st);
        FftArgs C = A; // column pass 1: scratch, in place
        C.src = sc;
        C.grid = sc;
        C.nct = N / kColTile;
        if (e == hipSuccess)
            e = launch_col3_pass<kPassCol1, SG>(C, A.N1, A.N2 * C.nct, st);
        C.scratch = dst + pl * plane; // column pass 2: scratch -> output
        C.nct = M / kColTile;
        if (e == hipSuccess)
            e = launch_col3_pass<kPassCol2, SG>(C, A.N2, 3 * A.N1 * C.nct, st);
    }
    return e;
}

bool grid_fft_noalloc_supported(int G)
{
    return stockham_size(G) && static_cast<long long>(G) * G * static_cast<long long>(sizeof(float2)) <=
                                   kStaticScratchBytes;
}

hipError_t launch_grid_fft_noalloc(void *grid, int G, long long nr_layers, bool forward, hipStream_t st)
{
    const long long planes = nr_layers * kCorr;
    if (!grid_fft_noalloc_supported(G) || planes > 0x7fffffffLL)
        return hipErrorInvalidValue;
    float2 *g = static_cast<float2 *>(grid);
    return forward ? stockham_fft2d<1>(g, g, G, G, planes, true, st, FftBox(), true)
                   : stockham_fft2d<-1>(g, g, G, G, planes, true, st, FftBox(), true);
}

hipError_t fft_plane_scratch(long long elements, float2 *&out, int user)
{
    return fft_scratch(elements, out, user);
}

bool fft2d_own_kernels(int nx, int ny, bool fp64)
{
    return lds_fft_supported(nx, ny, fp64) ||
           (!fp64 && ((stockham_size(nx) && stockham_size(ny)) || (nx == ny && three_times_stockham(nx))));
}

bool fft2d_supported(int nx, int ny)
{
    return nx >= 1 && ny >= 1 && static_cast<long long>(nx) * ny <= 0x7fffffffLL;
}

hipError_t launch_fft2d(const void *in, void *out, int nx, int ny, long long batch, int sign, bool fp64,
                        hipStream_t st)
{
    if (batch < 1 || batch > 0x7fffffffLL || !fft2d_supported(nx, ny) || (sign != 1 && sign != -1))
        return hipErrorInvalidValue;
    if (small_fft_supported(nx, ny))
        return launch_lds_fft2d(in, out, nx, ny, batch, sign, fp64, st);
    if (!fp64 && stockham_size(nx) && stockham_size(ny))
    {
        const float2 *src = static_cast<const float2 *>(in);
        float2 *dst = static_cast<float2 *>(out);
        return sign > 0 ? stockham_fft2d<1>(src, dst, nx, ny, batch, false, st)
                        : stockham_fft2d<-1>(src, dst, nx, ny, batch, false, st);
    }
    if (!fp64 && nx == ny && three_times_stockham(nx))
    {
        const float2 *src = static_cast<const float2 *>(in);
        float2 *dst = static_cast<float2 *>(out);
        return sign > 0 ? radix3_fft2d<1>(src, dst, nx, batch, st) : radix3_fft2d<-1>(src, dst, nx, batch, st);
    }
    if (lds_fft_supported(nx, ny, fp64))
        return launch_lds_fft2d(in, out, nx, ny, batch, sign, fp64, st);
    hipfftHandle plan;
    hipError_t e = fft_plan(nx, ny, batch, fp64 ? HIPFFT_Z2Z : HIPFFT_C2C, plan);
    if (e != hipSuccess)
        return e;
    if (hipfftSetStream(plan, st) != HIPFFT_SUCCESS)
        return hipErrorUnknown;
    const int dir = sign < 0 ? HIPFFT_FORWARD : HIPFFT_BACKWARD;
    hipfftResult r;
    if (fp64)
        r = hipfftExecZ2Z(plan, static_cast<hipfftDoubleComplex *>(const_cast<void *>(in)),
                          static_cast<hipfftDoubleComplex *>(out), dir);
    else
        r = hipfftExecC2C(plan, static_cast<hipfftComplex *>(const_cast<void *>(in)), static_cast<hipfftComplex *>(out),
                          dir);
    return r == HIPFFT_SUCCESS ? hipSuccess : hipErrorUnknown;
}

hipError_t launch_wstack(const void *src, void *dst, int G, int z_min, int nr_layers, float image_size,
                         float w_step, bool to_image, hipStream_t st)
{
    const long long cells = static_cast<long long>(G) * G;
    const long long blocks = (cells + kThreads - 1) / kThreads;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    if (to_image)
        wstack_to_image_kernel<<<static_cast<unsigned>(blocks), kThreads, 0, st>>>(
            static_cast<const float2 *>(src), static_cast<float2 *>(dst), G, z_min, nr_layers, image_size, w_step);
    else
        wstack_from_image_kernel<<<static_cast<unsigned>(blocks), kThreads, 0, st>>>(
            static_cast<const float2 *>(src), static_cast<float2 *>(dst), G, z_min, nr_layers, image_size, w_step);
    return hipGetLastError();
}

namespace
{
// Scratch of the w-stacked transform to the image (column pass 2's output for one correlation's layers, the box's
// columns only), per device: grown when a call needs more (the device is synchronised and the old one freed first),
// kept for the life of the process.
std::mutex g_ws_mutex;
std::map<int, std::pair<float2 *, long long>> g_ws_scratch;

hipError_t wstack_scratch(long long elements, float2 *&out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    auto &s = g_ws_scratch[dev];
    if (s.second < elements)
    {
        if (s.first)
        {
            e = hipDeviceSynchronize();
            if (e == hipSuccess)
                e = hipFree(s.first);
            if (e != hipSuccess)
                return e;
            s = {nullptr, 0};
        }
        float2 *p = nullptr;
        e = hipMalloc(&p, static_cast<size_t>(elements) * sizeof(float2));
        if (e != hipSuccess)
            return e;
        s = {p, elements};
    }
    out = s.first;
    return hipSuccess;
}
} // namespace

hipError_t launch_grid_wstack(void *grid, void *image, int G, int z_min, int nr_layers, float image_size,
                              float w_step, bool to_image, const int *box, hipStream_t st)
{
    const long long planes = static_cast<long long>(nr_layers) * kCorr;
    const long long plane = static_cast<long long>(G) * G;
    if (planes > 0x7fffffffLL || plane > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    hipError_t e = hipSuccess;
    if (!stockham_size(G))
    {
        if (to_image)
        {
            e = launch_grid_fft(grid, G, nr_layers, false, st, box);
            return e == hipSuccess ? launch_wstack(grid, image, G, z_min, nr_layers, image_size, w_step, true, st) : e;
        }
        e = launch_wstack(image, grid, G, z_min, nr_layers, image_size, w_step, false, st);
        return e == hipSuccess ? launch_grid_fft(grid, G, nr_layers, true, st, box) : e;
    }
    // The box's whole column tiles, as the boxed transform takes them.
    const int t0 = box[0] / kColTile, t1 = std::max(t0, (box[1] + kColTile - 1) / kColTile);
    WsArgs W{};
    W.nr_layers = nr_layers;
    W.z_min = z_min;
    W.cx0 = t0 * kColTile;
    W.cx1 = t1 * kColTile;
    W.image_size = image_size;
    W.w_step = w_step;
    float2 *g = static_cast<float2 *>(grid), *img = static_cast<float2 *>(image);
    if (to_image)
    {
        // Per correlation: the column passes of its layers (the grid is zero outside the box's columns) into the
        // scratch, then the row pass of all its layers into its image plane.
        const int nc = W.cx1 - W.cx0;
        float2 *sc = nullptr;
        if (nc > 0)
            e = wstack_scratch(static_cast<long long>(nr_layers) * G * nc, sc);
        W.rows_in = sc;
        W.layer_stride = static_cast<long long>(G) * nc;
        W.pitch = nc;
        for (int p = 0; p < kCorr && e == hipSuccess; ++p)
        {
            for (int z = 0; z < nr_layers && nc > 0 && e == hipSuccess; ++z)
            {
                FftArgs C = fft_args(G, G, true);
                C.src = C.grid = g + (static_cast<long long>(z) * kCorr + p) * plane;
                C.scratch = sc + static_cast<long long>(z) * G * nc;
                C.splane = static_cast<long long>(G) * nc;
                C.spitch = nc;
                C.sx0 = W.cx0;
                C.nplanes = 1;
                C.ct0 = t0;
                C.nct = t1 - t0;
                C.iy0 = box[2];
                C.iy1 = std::max(box[2], box[3]);
                e = launch_col_pass<kPassCol1, -1>(C, C.N1, C.nct * C.N2, st);
                if (e == hipSuccess)
                    e = launch_col_pass<kPassCol2, -1>(C, C.N2, C.nct * C.N1, st);
            }
            W.image_out = img + p * plane;
            if (e == hipSuccess)
                e = launch_wstack_row_pass<-1>(W, G, st);
        }
        return e;
    }
    // From the image: the row pass of every layer of each correlation (the box's columns of each layer row), then
    // the boxed transform's column passes over all planes.
    W.layer_stride = kCorr * plane;
    W.pitch = G;
    for (int p = 0; p < kCorr && e == hipSuccess; ++p)
    {
        W.image_in = img + p * plane;
        W.rows_out = g + p * plane;
        e = launch_wstack_row_pass<1>(W, G, st);
    }
    FftBox b;
    b.mode = kRowsOut;
    b.x0 = box[0];
    b.x1 = box[1];
    b.y0 = box[2];
    b.y1 = box[3];
    b.rows_done = true;
    return e == hipSuccess ? stockham_fft2d<1>(g, g, G, G, planes, true, st, b) : e;
}
} // namespace idgx
