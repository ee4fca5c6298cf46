// Batched 2-D DFTs through LDS line transforms for the plain FFT entry point (idgx_fft2d), FP32 or FP64:
// out[b][ky][kx] = sum in[b][y][x] exp(sign 2 pi i (ky y / ny + kx x / nx)), unnormalised, nx and ny powers of two.
//
// Everything happens in LDS tiles; every thread of a workgroup holds 16 complex points of a tile at a time.
// - FP32 64 x 64 and 128 x 128 planes: one pass with the plane in registers (plane_regs_kernel below).
// - Other planes with both sides <= 128 whose padded copy fits 132 KB of LDS (128 x 64 in FP64, and every smaller
//   plane) take one pass over HBM: a workgroup loads whole planes (several when a plane has fewer than 4096
//   points), transforms their rows, then their columns, and writes them once.
// - Otherwise, with ny <= 512: rows (max(1, 4096 / nx) whole rows per tile, contiguous loads; nx up to 16384 in
//   FP32, 8192 in FP64, whose padded row fills 128 KB; FP32 rows of 256 and 512 points in registers,
//   rows_regs_kernel below), then columns in place (a tile of ny rows x C columns,
//   C = max(4096 / ny, 128 bytes / point), so that every row segment is a whole 128-byte line); 64 MB of planes at
//   a time, so that the column pass reads the row pass's output from the memory-side cache.
// - With ny > 512: columns of ny = N1 N2 points in two four-step passes over tiles of C columns (A, in place: the
//   N1-point DFTs over rows n1 N2 + n2, times W_ny^(n2 k1); B: the N2-point DFTs over rows k1 N2 + n2, into a
//   scratch at rows k1 + N1 k2), then the rows from the scratch back into the output; 64 MB of planes (or one plane)
//   at a time.
// A line of n points is transformed by decimation-in-frequency stages of radix 4 to 16 (n = 16: 16; 32: 8 x 4;
// 64: 8 x 8; 128: 16 x 8; 256: 16 x 16; 512: 8 x 8 x 8; 1024: 16 x 8 x 8; ... 16384: 16 x 16 x 8 x 8).  A stage's
// small DFTs run in registers, their twiddles W_L^(j k) come from a per-workgroup table of W_n^t (sincospi in FP64,
// rounded once for FP32; above 512 points two tables, W_n^(64 a) and W_n^b, multiplied), and the last stage writes
// its outputs to their natural positions, so that no digit reversal is left for the stores.
// Bank conflicts: consecutive lanes take consecutive lines; rows are padded to nx + 1 points, so that lanes on
// consecutive rows fall in different banks, and lanes on consecutive columns read consecutive addresses.
// Workgroups loop over tiles (grid-stride), so that the twiddle table is computed once per workgroup; barriers are
// LDS-only, so that a tile's global stores need not complete before the next tile's loads start.  (Loading the next
// tile into registers during the current one's transform measured no faster: 64 x 64 planes run at 4.15 TB/s of
// HBM traffic either way.)
#include "kernels.hpp"

#include <algorithm>
#include <utility>

namespace idgx
{
namespace
{
constexpr int kPointsPerThread = 16;
constexpr int kRowTile = 4096;         // points per tile of the row pass, and the smallest one-pass tile
constexpr long long kMaxGrid = 8192;   // workgroups of a launch (each loops over tiles)
constexpr long long kOnePassLds = 132LL << 10;

template <class T> struct Vec;
template <> struct Vec<float>
{
    using type = float2;
};
template <> struct Vec<double>
{
    using type = double2;
};

// cos(2 pi q / 16) and sin(2 pi q / 16) for any integer q, from cos(pi / 8), cos(pi / 4) and cos(3 pi / 8).
constexpr double kCosEighthPi[5] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173, 0.0};
constexpr double cos16(int q)
{
    q &= 15;
    if (q > 8)
        q = 16 - q;
    return q <= 4 ? kCosEighthPi[q] : -kCosEighthPi[8 - q];
}
constexpr double sin16(int q)
{
    return cos16(q - 4);
}

template <class V> __device__ __forceinline__ V cadd(V a, V b)
{
    return {a.x + b.x, a.y + b.y};
}
template <class V> __device__ __forceinline__ V csub(V a, V b)
{
    return {a.x - b.x, a.y - b.y};
}
template <class V> __device__ __forceinline__ V cmul(V a, V w)
{
    return {a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x};
}

// v exp(SG 2 pi i K / R), R dividing 16; the factors 1 and +-i cost no multiplication.
template <int R, int K, int SG, class V> __device__ __forceinline__ V rotate(V v)
{
    using T = decltype(v.x);
    constexpr int q = K * (16 / R);
    if constexpr (q % 16 == 0)
        return v;
    else if constexpr (q % 16 == 4)
        return SG > 0 ? V{-v.y, v.x} : V{v.y, -v.x};
    else
    {
        constexpr T c = static_cast<T>(cos16(q)), s = static_cast<T>(SG * sin16(q));
        return {v.x * c - v.y * s, v.x * s + v.y * c};
    }
}

template <int R, int SG, class V> __device__ __forceinline__ void dft(V *a);

template <int R, int K, int SG, class V> __device__ __forceinline__ void butterfly(V *a, const V *e, const V *o)
{
    const V t = rotate<R, K, SG>(o[K]);
    a[K] = cadd(e[K], t);
    a[K + R / 2] = csub(e[K], t);
}

template <int R, int SG, class V, int... K>
__device__ __forceinline__ void combine_halves(V *a, const V *e, const V *o, std::integer_sequence<int, K...>)
{
    (butterfly<R, K, SG>(a, e, o), ...);
}

// R-point DFT in registers, natural order in and out (radix-2 decimation in time, unrolled at compile time).
template <int R, int SG, class V> __device__ __forceinline__ void dft(V *a)
{
    if constexpr (R == 2)
    {
        const V t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    }
    else
    {
        constexpr int H = R / 2;
        V e[H], o[H];
#pragma unroll
        for (int i = 0; i < H; ++i)
        {
            e[i] = a[2 * i];
            o[i] = a[2 * i + 1];
        }
        dft<H, SG>(e);
        dft<H, SG>(o);
        combine_halves<R, SG>(a, e, o, std::make_integer_sequence<int, H>());
    }
}

// Radix sequence of an N-point line (count stages; r[count..3] = 1).
template <int C, int R0, int R1 = 1, int R2 = 1, int R3 = 1> struct RadixList
{
    static constexpr int count = C;
    static constexpr int r[4] = {R0, R1, R2, R3};
};
template <int N> struct Radices;
template <> struct Radices<16> : RadixList<1, 16> {};
template <> struct Radices<32> : RadixList<2, 8, 4> {};
template <> struct Radices<64> : RadixList<2, 8, 8> {};
template <> struct Radices<128> : RadixList<2, 16, 8> {};
template <> struct Radices<256> : RadixList<2, 16, 16> {};
template <> struct Radices<512> : RadixList<3, 8, 8, 8> {};
template <> struct Radices<1024> : RadixList<3, 16, 8, 8> {};
template <> struct Radices<2048> : RadixList<3, 16, 16, 8> {};
template <> struct Radices<4096> : RadixList<3, 16, 16, 16> {};
template <> struct Radices<8192> : RadixList<4, 16, 16, 8, 4> {};
template <> struct Radices<16384> : RadixList<4, 16, 16, 8, 8> {};

// Natural frequency of output 0 of the last stage's block blk.  After stages of radix r0, r1, ... the position
// d0 (N / r0) + d1 (N / r0 r1) + ... + ds holds frequency d0 + r0 d1 + r0 r1 d2 + ..., and blk = position / rs, whose
// digits are d0 .. d(s-1), d0 the most significant.
template <int N> __device__ __forceinline__ int natural_base(int blk)
{
    using Rs = Radices<N>;
    int k = 0, mult = 1;
#pragma unroll
    for (int i = 0; i + 1 < Rs::count; ++i)
    {
        int span = 1;
#pragma unroll
        for (int j = i + 1; j + 1 < Rs::count; ++j)
            span *= Rs::r[j];
        const int d = blk / span;
        blk -= d * span;
        k += d * mult;
        mult *= Rs::r[i];
    }
    return k;
}

// Twiddles W_N^t (t < N) of a line: one table of N values for N <= 512; above, W_N^(64 a) W_N^b from two tables of
// N / 64 and 64 values (one complex multiply per twiddle, in T).
template <int N> constexpr int twiddle_words()
{
    return N <= 512 ? N : N / 64 + 64;
}
template <int N, class V> struct Twiddles
{
    const V *tab;
    __device__ __forceinline__ V operator()(int t) const
    {
        if constexpr (N <= 512)
            return tab[t];
        else
            return cmul(tab[t >> 6], tab[N / 64 + (t & 63)]);
    }
};

// Fills the twiddle words of an N-point line: exp(SG 2 pi i t / N), computed in FP64 and rounded once.
template <int N, int SG, int THREADS, class V> __device__ __forceinline__ void twiddle_table(V *tw)
{
    using T = decltype(tw->x);
    for (int i = threadIdx.x; i < twiddle_words<N>(); i += THREADS)
    {
        const int t = N <= 512 ? i : i < N / 64 ? 64 * i : i - N / 64;
        double s, c;
        sincospi(2.0 * t / N, &s, &c);
        tw[i] = V{static_cast<T>(c), static_cast<T>(SG * s)};
    }
}

// One non-final DIF stage of radix R on blocks of L points, in place: for each line, block and j < Q = L / R, the R
// points at j + m Q become their DFT y_k times W_L^(j k) = W_N^(j k N / L).  Consecutive threads take consecutive
// lines.
template <int N, int R, int L, int NLINES, int THREADS, int SG, class V, class Addr>
__device__ __forceinline__ void dif_stage(V *lds, Twiddles<N, V> tw, Addr addr)
{
    constexpr int Q = L / R, BUTTERFLIES = NLINES * (N / R);
    static_assert(BUTTERFLIES % THREADS == 0, "whole butterflies per thread");
#pragma unroll
    for (int it = 0; it < BUTTERFLIES / THREADS; ++it)
    {
        const int b = threadIdx.x + it * THREADS;
        const int line = b % NLINES, bf = b / NLINES, blk = bf / Q, j = bf % Q;
        const int base = blk * L + j;
        V a[R];
#pragma unroll
        for (int m = 0; m < R; ++m)
            a[m] = lds[addr(line, base + m * Q)];
        dft<R, SG>(a);
#pragma unroll
        for (int k = 1; k < R; ++k)
            a[k] = cmul(a[k], tw(j * k * (N / L)));
#pragma unroll
        for (int k = 0; k < R; ++k)
            lds[addr(line, base + k * Q)] = a[k];
    }
    lds_barrier();
}

// The last stage (blocks of R points): every thread reads its blocks into registers, the workgroup waits, and output
// k of block blk goes to its natural position natural_base(blk) + (N / R) k.
template <int N, int R, int NLINES, int THREADS, int SG, class V, class Addr>
__device__ __forceinline__ void last_stage(V *lds, Addr addr)
{
    constexpr int BUTTERFLIES = NLINES * (N / R), ITER = BUTTERFLIES / THREADS;
    static_assert(BUTTERFLIES % THREADS == 0, "whole butterflies per thread");
    V a[ITER][R];
#pragma unroll
    for (int it = 0; it < ITER; ++it)
    {
        const int b = threadIdx.x + it * THREADS;
        const int line = b % NLINES, blk = b / NLINES;
#pragma unroll
        for (int m = 0; m < R; ++m)
            a[it][m] = lds[addr(line, blk * R + m)];
    }
    lds_barrier();
#pragma unroll
    for (int it = 0; it < ITER; ++it)
    {
        const int b = threadIdx.x + it * THREADS;
        const int line = b % NLINES, blk = b / NLINES;
        dft<R, SG>(a[it]);
        const int k0 = natural_base<N>(blk);
#pragma unroll
        for (int k = 0; k < R; ++k)
            lds[addr(line, k0 + (N / R) * k)] = a[it][k];
    }
    lds_barrier();
}

// Stages STAGE .. count-1 on blocks of L points.
template <int N, int STAGE, int L, int NLINES, int THREADS, int SG, class V, class Addr>
__device__ __forceinline__ void stages_from(V *lds, Twiddles<N, V> tw, Addr addr)
{
    using Rs = Radices<N>;
    if constexpr (STAGE + 1 == Rs::count)
        last_stage<N, Rs::r[STAGE], NLINES, THREADS, SG>(lds, addr);
    else
    {
        dif_stage<N, Rs::r[STAGE], L, NLINES, THREADS, SG>(lds, tw, addr);
        stages_from<N, STAGE + 1, L / Rs::r[STAGE], NLINES, THREADS, SG>(lds, tw, addr);
    }
}

// Every stage on NLINES lines of N points; element e of line l is lds[addr(l, e)].  Natural order in and out; the
// caller has synchronised after filling the lines (and the twiddle words), and every stage ends with a barrier.
template <int N, int NLINES, int THREADS, int SG, class V, class Addr>
__device__ __forceinline__ void fft_lines(V *lds, const V *tw, Addr addr)
{
    stages_from<N, 0, N, NLINES, THREADS, SG>(lds, Twiddles<N, V>{tw}, addr);
}

template <int NX, int NY> constexpr int planes_per_tile()
{
    return NX * NY >= kRowTile ? 1 : kRowTile / (NX * NY);
}
template <int NX, int NY> constexpr int plane_threads()
{
    return NX * NY * planes_per_tile<NX, NY>() / kPointsPerThread;
}

// One pass: P whole planes per tile; LDS rows padded to NX + 1 points.
template <int NX, int NY, int SG, class T>
__global__ __launch_bounds__((plane_threads<NX, NY>())) void small_fft_planes_kernel(const typename Vec<T>::type *in,
                                                                                   typename Vec<T>::type *out,
                                                                                   long long planes)
{
    using V = typename Vec<T>::type;
    constexpr int P = planes_per_tile<NX, NY>(), THREADS = plane_threads<NX, NY>();
    constexpr int PITCH = NX + 1, PLANE = NX * NY;
    __shared__ V lds[P * NY * PITCH];
    __shared__ V tw[NX + NY];
    twiddle_table<NX, SG, THREADS>(tw);
    twiddle_table<NY, SG, THREADS>(tw + NX);
    const long long tiles = (planes + P - 1) / P;
    for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x)
    {
        const long long p0 = tile * P;
        const int np = static_cast<int>(planes - p0 < P ? planes - p0 : P);
        const V *src = in + p0 * PLANE;
        for (int i = threadIdx.x; i < np * PLANE; i += THREADS)
            lds[(i / NX) * PITCH + i % NX] = src[i];
        lds_barrier();
        fft_lines<NX, P * NY, THREADS, SG>(lds, tw, [](int line, int e) { return line * PITCH + e; });
        fft_lines<NY, P * NX, THREADS, SG>(
            lds, tw + NX, [](int line, int e) { return ((line / NX) * NY + e) * PITCH + line % NX; });
        V *dst = out + p0 * PLANE;
        for (int i = threadIdx.x; i < np * PLANE; i += THREADS)
            dst[i] = lds[(i / NX) * PITCH + i % NX];
        lds_barrier();
    }
}

// One pass for FP32 N x N planes, N = 64 or 128, with the plane in registers (16 points per thread) instead of in
// LDS.  Thread t is lane j = t % R of line a = t / R (R = N / 16): a row in the row phase, a column in the column
// phase.  Each phase is X[k + 16 q] = sum_j W_R^(j q) W_N^(j k) sum_m W_16^(m k) x[j + R m]: a radix-16 DFT over
// the thread's points j + R m, the twiddles W_N^(j k), then for k = 16 j / R + kk (kk < 16 / R) a radix-R DFT over
// the R lanes of the line, whose values an LDS exchange brings together.  Between the phases a transpose through
// LDS gives thread (a, j) the points j + R m of column a.  LDS holds one float per point (the real parts, then the
// imaginary parts, through the same array): 69 KB at N = 128, so two 1024-thread workgroups share a CU, where the
// complex plane in LDS (small_fft_planes_kernel) takes 132 KB and leaves one.  Line exchanges pad one slot per 16
// points (lanes j of 32-lane halves of a wave land in distinct banks both ways); the transpose pitch N + 32 / R
// makes its column reads conflict-free.  At N = 128 the 64-VGPR limit leaves 24 bytes per lane in scratch; a form
// that recomputed the line indices in each phase spilled less and measured slower (71.0 against 64.7 ns per plane).
template <int N> constexpr int plane_regs_waves_per_simd()
{
    return N * N / 16 / 64 / 4 * 2; // two workgroups of N^2 / 16 threads per CU
}

template <int N, int SG>
__global__ __launch_bounds__(N * N / 16, plane_regs_waves_per_simd<N>()) void plane_regs_kernel(const float2 *in,
                                                                                               float2 *out,
                                                                                               long long planes)
{
    constexpr int THREADS = N * N / 16, R = N / 16, KPT = 16 / R;
    constexpr int LP = N + N / 16;      // line pitch of the line exchanges (one pad slot per 16 points)
    constexpr int TP = N + 32 / R;      // row pitch of the transpose
    constexpr int WORDS = N * (LP > TP ? LP : TP);
    static_assert(N == 64 || N == 128, "N = 64 or 128");
    __shared__ float buf[WORDS];
    // tw[16 j + k] = W_N^(j k) (j < R, k < 16): a lane's 15 twiddles are one base address plus constant offsets.
    __shared__ float2 tw[N];
    for (int i = threadIdx.x; i < N; i += THREADS)
    {
        double sn, cs;
        sincospi(2.0 * ((i / 16) * (i % 16)) / N, &sn, &cs);
        tw[i] = make_float2(static_cast<float>(cs), static_cast<float>(SG * sn));
    }
    lds_barrier();
    auto phys = [](int p) { return p + p / 16; };
    // The new real parts go straight into v[].x (the old ones are in LDS by then) while v[].y still holds the
    // imaginary parts to send: no temporary array.
    auto exchange = [&](float2 *v, auto widx, auto ridx) {
        lds_barrier(); // the previous exchange's reads are done
#pragma unroll
        for (int r = 0; r < 16; ++r)
            buf[widx(r)] = v[r].x;
        lds_barrier();
#pragma unroll
        for (int r = 0; r < 16; ++r)
            v[r].x = buf[ridx(r)];
        lds_barrier();
#pragma unroll
        for (int r = 0; r < 16; ++r)
            buf[widx(r)] = v[r].y;
        lds_barrier();
#pragma unroll
        for (int r = 0; r < 16; ++r)
            v[r].y = buf[ridx(r)];
    };
    // One phase on the thread's line: points j + R m in v[m] -> X[KPT j + kk + 16 q] in v[kk R + q].
    auto line_fft = [&](float2 *v, int a, int j) {
        dft<16, SG>(v);
#pragma unroll
        for (int k = 1; k < 16; ++k)
            v[k] = cmul(v[k], tw[16 * j + k]);
        // y[k] of lane j at line position k R + j; lane j' takes positions 16 j' .. 16 j' + 15 = (k = KPT j' + kk,
        // lane jj) for kk < KPT, jj < R.
        exchange(
            v, [&](int k) { return a * LP + phys(k * R + j); }, [&](int e) { return a * LP + phys(16 * j + e); });
#pragma unroll
        for (int kk = 0; kk < KPT; ++kk)
            dft<R, SG>(v + kk * R);
    };
    const int nplanes = static_cast<int>(planes); // the entry point admits batches below 2^31
    for (int p = blockIdx.x; p < nplanes; p += gridDim.x)
    {
        int t = threadIdx.x; // opaque per plane: indices are formed where they are used
        asm volatile("" : "+v"(t));
        const int a = t / R, j = t % R;
        const float2 *src = in + static_cast<size_t>(p) * N * N;
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m)
            v[m] = src[a * N + j + R * m];
        line_fft(v, a, j); // rows
        // Transpose: X of row a at column K = KPT j + kk + 16 q (in v[kk R + q]) -> thread (c, j) takes rows j + R m
        // of column c = its line a.
        exchange(
            v, [&](int e) { return a * TP + KPT * j + e / R + 16 * (e % R); },
            [&](int m) { return (j + R * m) * TP + a; });
        line_fft(v, a, j); // columns
        float2 *dst = out + static_cast<size_t>(p) * N * N;
#pragma unroll
        for (int e = 0; e < 16; ++e)
            dst[(KPT * j + e / R + 16 * (e % R)) * N + a] = v[e];
    }
}

// Rows of N = 256 or 512 points (FP32) with the row in registers: 256 threads, 4096 / N rows per tile; thread
// (r, j) of the R = N / 16 lanes of row r takes its points j + R m.  A radix-16 DFT over m and the twiddles
// W_N^(j k) give y_j[k], and an exchange through LDS within the row gives
//   N = 256: lane k the values y_j[k] (j < 16), whose radix-16 DFT over j is X[k + 16 q];
//   N = 512: lane 2 kk + j1 the values y_{j1 + 2 j2}[kk] (j2 < 16); their radix-16 DFT over j2 is Z_j1[q], and
//            X[kk + 16 q + 256 j1] = Z_0[q] + (-1)^j1 W_32^q Z_1[q], a butterfly between neighbouring lanes (a DPP
//            lane swap).
// LDS holds one float per point (real parts, then imaginary parts) with one pad slot per 16 points.
template <int N, int SG>
__global__ __launch_bounds__(256) void rows_regs_kernel(const float2 *in, float2 *out, long long rows)
{
    constexpr int THREADS = 256, R = N / 16, ROWS = THREADS / R, LP = N + N / 16;
    static_assert(N == 256 || N == 512, "N = 256 or 512");
    __shared__ float buf[ROWS * LP];
    // tw[16 j + k] = W_N^(j k) for j < R, k < 16 (at N = 512, j = 16 gives W_32^k).
    __shared__ float2 tw[N];
    for (int i = threadIdx.x; i < N; i += THREADS)
    {
        double sn, cs;
        sincospi(2.0 * ((i / 16) * (i % 16)) / N, &sn, &cs);
        tw[i] = make_float2(static_cast<float>(cs), static_cast<float>(SG * sn));
    }
    lds_barrier();
    auto phys = [](int p) { return p + p / 16; };
    const long long tiles = (rows + ROWS - 1) / ROWS;
    for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x)
    {
        int t = threadIdx.x; // opaque per tile: indices are formed where they are used
        asm volatile("" : "+v"(t));
        const int r = t / R, j = t % R;
        const long long row = tile * ROWS + r;
        const bool live = row < rows; // lanes of a missing last-tile row still join the barriers
        const float2 *src = in + (live ? row : 0) * N;
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m)
            v[m] = live ? src[j + R * m] : make_float2(0.F, 0.F);
        dft<16, SG>(v);
#pragma unroll
        for (int k = 1; k < 16; ++k)
            v[k] = cmul(v[k], tw[16 * j + k]);
        // y_j[k] at row position k R + j; the reads gather 16 values per lane (see above).
        auto widx = [&](int k) { return r * LP + phys(k * R + j); };
        auto ridx = [&](int e) {
            if constexpr (N == 256)
                return r * LP + phys(16 * j + e);
            else
                return r * LP + phys(32 * (j >> 1) + (j & 1) + 2 * e);
        };
        lds_barrier(); // the previous tile's reads are done
#pragma unroll
        for (int k = 0; k < 16; ++k)
            buf[widx(k)] = v[k].x;
        lds_barrier();
#pragma unroll
        for (int e = 0; e < 16; ++e)
            v[e].x = buf[ridx(e)];
        lds_barrier();
#pragma unroll
        for (int k = 0; k < 16; ++k)
            buf[widx(k)] = v[k].y;
        lds_barrier();
#pragma unroll
        for (int e = 0; e < 16; ++e)
            v[e].y = buf[ridx(e)];
        dft<16, SG>(v);
        float2 *dst = out + (live ? row : 0) * N;
        if constexpr (N == 256)
        {
            if (live)
            {
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    dst[j + 16 * q] = v[q];
            }
        }
        else
        {
            const int kk = j >> 1, j1 = j & 1;
            // Neighbouring lanes 2 kk and 2 kk + 1 swap values: DPP quad_perm [1, 0, 3, 2].
            auto swap = [](float x) {
                return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
            };
#pragma unroll
            for (int q = 0; q < 16; ++q)
            {
                if (j1)
                    v[q] = cmul(v[q], tw[256 + q]);
                const float2 o = make_float2(swap(v[q].x), swap(v[q].y));
                v[q] = j1 ? make_float2(o.x - v[q].x, o.y - v[q].y) : make_float2(v[q].x + o.x, v[q].y + o.y);
            }
            if (live)
            {
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    dst[kk + 16 * q + 256 * j1] = v[q];
            }
        }
    }
}

// Row pass: max(1, kRowTile / NX) whole rows (of any plane) per tile, in -> out; NX up to 16384 (FP32) or 8192
// (FP64), whose padded row fills 128 KB of LDS.
template <int NX> constexpr int rows_per_tile()
{
    return NX >= kRowTile ? 1 : kRowTile / NX;
}
template <int NX> constexpr int row_threads()
{
    return rows_per_tile<NX>() * NX / kPointsPerThread;
}

template <int NX, int SG, class T>
__global__ __launch_bounds__((row_threads<NX>())) void small_fft_rows_kernel(const typename Vec<T>::type *in,
                                                                           typename Vec<T>::type *out, long long rows)
{
    using V = typename Vec<T>::type;
    constexpr int ROWS = rows_per_tile<NX>(), THREADS = row_threads<NX>(), PITCH = NX + 1;
    __shared__ V lds[ROWS * PITCH];
    __shared__ V tw[twiddle_words<NX>()];
    twiddle_table<NX, SG, THREADS>(tw);
    const long long tiles = (rows + ROWS - 1) / ROWS;
    for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x)
    {
        const long long r0 = tile * ROWS;
        const int nr = static_cast<int>(rows - r0 < ROWS ? rows - r0 : ROWS);
        const V *src = in + r0 * NX;
        for (int i = threadIdx.x; i < nr * NX; i += THREADS)
            lds[(i / NX) * PITCH + i % NX] = src[i];
        lds_barrier();
        fft_lines<NX, ROWS, THREADS, SG>(lds, tw, [](int line, int e) { return line * PITCH + e; });
        V *dst = out + r0 * NX;
        for (int i = threadIdx.x; i < nr * NX; i += THREADS)
            dst[i] = lds[(i / NX) * PITCH + i % NX];
        lds_barrier();
    }
}

template <int NY, class T> constexpr int column_tile()
{
    constexpr int by_points = kRowTile / NY, by_line = 128 / static_cast<int>(sizeof(typename Vec<T>::type));
    return by_points > by_line ? by_points : by_line;
}

// Column pass, in place: tiles of NY rows x C columns of one plane (C consecutive points of every row).
template <int NY, int SG, class T>
__global__ __launch_bounds__((NY * column_tile<NY, T>() / kPointsPerThread)) void small_fft_columns_kernel(
    typename Vec<T>::type *data, int nx, long long tiles)
{
    using V = typename Vec<T>::type;
    constexpr int C = column_tile<NY, T>(), THREADS = NY * C / kPointsPerThread;
    __shared__ V lds[NY * C];
    __shared__ V tw[NY];
    twiddle_table<NY, SG, THREADS>(tw);
    const int per_plane = nx / C;
    for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x)
    {
        const long long plane = tile / per_plane;
        const int c0 = static_cast<int>(tile - plane * per_plane) * C;
        V *p = data + plane * static_cast<long long>(nx) * NY + c0;
        for (int i = threadIdx.x; i < NY * C; i += THREADS)
            lds[i] = p[static_cast<long long>(i / C) * nx + i % C];
        lds_barrier();
        fft_lines<NY, C, THREADS, SG>(lds, tw, [](int line, int e) { return e * C + line; });
        for (int i = threadIdx.x; i < NY * C; i += THREADS)
            p[static_cast<long long>(i / C) * nx + i % C] = lds[i];
        lds_barrier();
    }
}

// Columns of NY = N1 N2 > 512 points in two passes over HBM (four-step), on tiles of C consecutive columns; row
// index n = n1 N2 + n2 in, k = k1 + N1 k2 out.
// Pass A, src -> dst at the same positions (in place when equal): for each n2, the N1-point DFT over n1, times
// W_NY^(n2 k1), stored at row k1 N2 + n2.
template <int N1, int N2, int SG, class T>
__global__ __launch_bounds__((N1 * column_tile<N1, T>() / kPointsPerThread)) void fft_columns_a_kernel(
    const typename Vec<T>::type *src, typename Vec<T>::type *dst, int nx, long long tiles)
{
    using V = typename Vec<T>::type;
    constexpr int C = column_tile<N1, T>(), THREADS = N1 * C / kPointsPerThread, NY = N1 * N2;
    __shared__ V lds[N1 * C];
    __shared__ V tw[twiddle_words<N1>()];
    __shared__ V step[N1];
    twiddle_table<N1, SG, THREADS>(tw);
    const int per_row = nx / C;
    for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x)
    {
        const long long plane = tile / (static_cast<long long>(N2) * per_row);
        const int rest = static_cast<int>(tile - plane * N2 * per_row), n2 = rest / per_row;
        const long long base = plane * NY * static_cast<long long>(nx) + (rest - n2 * per_row) * C;
        for (int i = threadIdx.x; i < N1 * C; i += THREADS)
            lds[i] = src[base + static_cast<long long>((i / C) * N2 + n2) * nx + i % C];
        for (int k1 = threadIdx.x; k1 < N1; k1 += THREADS)
        {
            double sn, cs;
            sincospi(2.0 * (n2 * k1) / NY, &sn, &cs);
            step[k1] = V{static_cast<T>(cs), static_cast<T>(SG * sn)};
        }
        lds_barrier();
        fft_lines<N1, C, THREADS, SG>(lds, tw, [](int line, int e) { return e * C + line; });
        for (int i = threadIdx.x; i < N1 * C; i += THREADS)
            dst[base + static_cast<long long>((i / C) * N2 + n2) * nx + i % C] = cmul(lds[i], step[i / C]);
        lds_barrier();
    }
}

// Pass B, src -> dst (distinct): for each k1, the N2-point DFT over the rows k1 N2 + n2, stored at rows k1 + N1 k2.
template <int N1, int N2, int SG, class T>
__global__ __launch_bounds__((N2 * column_tile<N2, T>() / kPointsPerThread)) void fft_columns_b_kernel(
    const typename Vec<T>::type *src, typename Vec<T>::type *dst, int nx, long long tiles)
{
    using V = typename Vec<T>::type;
    constexpr int C = column_tile<N2, T>(), THREADS = N2 * C / kPointsPerThread, NY = N1 * N2;
    __shared__ V lds[N2 * C];
    __shared__ V tw[twiddle_words<N2>()];
    twiddle_table<N2, SG, THREADS>(tw);
    const int per_row = nx / C;
    for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x)
    {
        const long long plane = tile / (static_cast<long long>(N1) * per_row);
        const int rest = static_cast<int>(tile - plane * N1 * per_row), k1 = rest / per_row;
        const long long base = plane * NY * static_cast<long long>(nx) + (rest - k1 * per_row) * C;
        for (int i = threadIdx.x; i < N2 * C; i += THREADS)
            lds[i] = src[base + static_cast<long long>(k1 * N2 + i / C) * nx + i % C];
        lds_barrier();
        fft_lines<N2, C, THREADS, SG>(lds, tw, [](int line, int e) { return e * C + line; });
        for (int i = threadIdx.x; i < N2 * C; i += THREADS)
            dst[base + static_cast<long long>(k1 + N1 * (i / C)) * nx + i % C] = lds[i];
        lds_barrier();
    }
}

unsigned grid_for(long long tiles)
{
    return static_cast<unsigned>(tiles < kMaxGrid ? tiles : kMaxGrid);
}

template <int N, int SG> hipError_t rows_regs_pass(const void *in, void *out, long long rows, hipStream_t st)
{
    constexpr int ROWS = 256 / (N / 16);
    rows_regs_kernel<N, SG><<<grid_for((rows + ROWS - 1) / ROWS), 256, 0, st>>>(static_cast<const float2 *>(in),
                                                                                static_cast<float2 *>(out), rows);
    return hipGetLastError();
}

template <int N, int SG> hipError_t plane_regs_pass(const void *in, void *out, long long batch, hipStream_t st)
{
    plane_regs_kernel<N, SG><<<grid_for(batch), N * N / 16, 0, st>>>(static_cast<const float2 *>(in),
                                                                      static_cast<float2 *>(out), batch);
    return hipGetLastError();
}

template <int NX, int NY, int SG, class T>
hipError_t planes_pass(const void *in, void *out, long long batch, hipStream_t st)
{
    using V = typename Vec<T>::type;
    constexpr int P = planes_per_tile<NX, NY>();
    small_fft_planes_kernel<NX, NY, SG, T><<<grid_for((batch + P - 1) / P), plane_threads<NX, NY>(), 0, st>>>(
        static_cast<const V *>(in), static_cast<V *>(out), batch);
    return hipGetLastError();
}

template <int NX, int SG, class T> hipError_t rows_pass(const void *in, void *out, long long rows, hipStream_t st)
{
    using V = typename Vec<T>::type;
    constexpr int ROWS = rows_per_tile<NX>();
    small_fft_rows_kernel<NX, SG, T><<<grid_for((rows + ROWS - 1) / ROWS), row_threads<NX>(), 0, st>>>(
        static_cast<const V *>(in), static_cast<V *>(out), rows);
    return hipGetLastError();
}

template <int NY, int SG, class T> hipError_t columns_pass(void *data, int nx, long long batch, hipStream_t st)
{
    using V = typename Vec<T>::type;
    constexpr int C = column_tile<NY, T>();
    if (nx % C != 0)
        return hipErrorInvalidValue;
    const long long tiles = batch * (nx / C);
    small_fft_columns_kernel<NY, SG, T><<<grid_for(tiles), NY * C / kPointsPerThread, 0, st>>>(static_cast<V *>(data),
                                                                                             nx, tiles);
    return hipGetLastError();
}

template <int SG, class T> hipError_t rows_any(int nx, const void *in, void *out, long long rows, hipStream_t st)
{
    switch (nx)
    {
    case 16:
        return rows_pass<16, SG, T>(in, out, rows, st);
    case 32:
        return rows_pass<32, SG, T>(in, out, rows, st);
    case 64:
        return rows_pass<64, SG, T>(in, out, rows, st);
    case 128:
        return rows_pass<128, SG, T>(in, out, rows, st);
    case 256:
        if constexpr (sizeof(T) == sizeof(float))
            return rows_regs_pass<256, SG>(in, out, rows, st);
        else
            return rows_pass<256, SG, T>(in, out, rows, st);
    case 512:
        if constexpr (sizeof(T) == sizeof(float))
            return rows_regs_pass<512, SG>(in, out, rows, st);
        else
            return rows_pass<512, SG, T>(in, out, rows, st);
    case 1024:
        return rows_pass<1024, SG, T>(in, out, rows, st);
    case 2048:
        return rows_pass<2048, SG, T>(in, out, rows, st);
    case 4096:
        return rows_pass<4096, SG, T>(in, out, rows, st);
    case 8192:
        return rows_pass<8192, SG, T>(in, out, rows, st);
    case 16384:
        if constexpr (sizeof(T) == 4) // a padded FP64 row of 16384 points does not fit LDS
            return rows_pass<16384, SG, T>(in, out, rows, st);
        else
            return hipErrorInvalidValue;
    default:
        return hipErrorInvalidValue;
    }
}

template <int SG, class T> hipError_t columns_any(int ny, void *data, int nx, long long batch, hipStream_t st)
{
    switch (ny)
    {
    case 16:
        return columns_pass<16, SG, T>(data, nx, batch, st);
    case 32:
        return columns_pass<32, SG, T>(data, nx, batch, st);
    case 64:
        return columns_pass<64, SG, T>(data, nx, batch, st);
    case 128:
        return columns_pass<128, SG, T>(data, nx, batch, st);
    case 256:
        return columns_pass<256, SG, T>(data, nx, batch, st);
    case 512:
        return columns_pass<512, SG, T>(data, nx, batch, st);
    default:
        return hipErrorInvalidValue;
    }
}

template <int N1, int N2, int SG, class T>
hipError_t four_step_columns(const void *src, void *dst, void *scratch, int nx, long long planes, hipStream_t st)
{
    using V = typename Vec<T>::type;
    constexpr int CT1 = column_tile<N1, T>(), CT2 = column_tile<N2, T>();
    if (nx % CT1 != 0 || nx % CT2 != 0)
        return hipErrorInvalidValue;
    const long long ta = planes * N2 * (nx / CT1), tb = planes * N1 * (nx / CT2);
    fft_columns_a_kernel<N1, N2, SG, T><<<grid_for(ta), N1 * CT1 / kPointsPerThread, 0, st>>>(
        static_cast<const V *>(src), static_cast<V *>(dst), nx, ta);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    fft_columns_b_kernel<N1, N2, SG, T><<<grid_for(tb), N2 * CT2 / kPointsPerThread, 0, st>>>(
        static_cast<const V *>(dst), static_cast<V *>(scratch), nx, tb);
    return hipGetLastError();
}

// Splits NY = N1 x N2 of the four-step column passes.
template <int SG, class T>
hipError_t long_columns(int ny, const void *src, void *dst, void *scratch, int nx, long long planes, hipStream_t st)
{
    switch (ny)
    {
    case 1024:
        return four_step_columns<32, 32, SG, T>(src, dst, scratch, nx, planes, st);
    case 2048:
        return four_step_columns<64, 32, SG, T>(src, dst, scratch, nx, planes, st);
    case 4096:
        return four_step_columns<64, 64, SG, T>(src, dst, scratch, nx, planes, st);
    case 8192:
        return four_step_columns<128, 64, SG, T>(src, dst, scratch, nx, planes, st);
    case 16384:
        return four_step_columns<128, 128, SG, T>(src, dst, scratch, nx, planes, st);
    default:
        return hipErrorInvalidValue;
    }
}

int column_tile_of(int n, bool fp64)
{
    const int by_points = kRowTile / n, by_line = fp64 ? 8 : 16;
    return by_points > by_line ? by_points : by_line;
}

void column_split(int ny, int &n1, int &n2)
{
    n1 = ny >= 8192 ? 128 : ny >= 2048 ? 64 : 32;
    n2 = ny / n1;
}

// One-pass plane shapes: both sides <= 128 and the padded tile within kOnePassLds.
template <int NX, int NY, class T> constexpr bool one_pass()
{
    return NX <= 128 && NY <= 128 &&
           static_cast<long long>(NX + 1) * NY * planes_per_tile<NX, NY>() *
                   static_cast<long long>(sizeof(typename Vec<T>::type)) <=
               kOnePassLds;
}

template <int SG, class T, int NX, int NY>
bool try_planes(int nx, int ny, const void *in, void *out, long long b, hipStream_t st, hipError_t &e)
{
    if constexpr (one_pass<NX, NY, T>())
    {
        if (nx == NX && ny == NY)
        {
            e = planes_pass<NX, NY, SG, T>(in, out, b, st);
            return true;
        }
    }
    return false;
}

template <int SG, class T, int NX, int... LY>
bool try_row(int nx, int ny, const void *in, void *out, long long b, hipStream_t st, hipError_t &e,
             std::integer_sequence<int, LY...>)
{
    return (try_planes<SG, T, NX, (16 << LY)>(nx, ny, in, out, b, st, e) || ...);
}

// Sides 16, 32, 64, 128 (16 << 0 .. 16 << 3) in both directions.
template <int SG, class T, int... LX>
bool planes_any(int nx, int ny, const void *in, void *out, long long b, hipStream_t st, hipError_t &e,
                std::integer_sequence<int, LX...>)
{
    return (try_row<SG, T, (16 << LX)>(nx, ny, in, out, b, st, e, std::make_integer_sequence<int, 4>()) || ...);
}

bool pow2_in(int n, int lo, int hi)
{
    return n >= lo && n <= hi && (n & (n - 1)) == 0;
}

// Planes per launch group of the four-step path: the scratch holds one group (64 MB, or one plane when larger).
constexpr long long kGroupBytes = 64LL << 20;

template <int SG, class T>
hipError_t lds_fft(const void *in, void *out, int nx, int ny, long long batch, hipStream_t st)
{
    using V = typename Vec<T>::type;
    hipError_t e = hipSuccess;
    if constexpr (sizeof(T) == sizeof(float))
    {
        if (nx == ny && nx == 64)
            return plane_regs_pass<64, SG>(in, out, batch, st);
        if (nx == ny && nx == 128)
            return plane_regs_pass<128, SG>(in, out, batch, st);
    }
    if (planes_any<SG, T>(nx, ny, in, out, batch, st, e, std::make_integer_sequence<int, 4>()))
        return e;
    const long long plane = static_cast<long long>(nx) * ny;
    const long long plane_bytes = plane * static_cast<long long>(sizeof(V));
    const long long group = std::max(1LL, std::min(batch, kGroupBytes / plane_bytes));
    if (ny <= 512)
    {
        // Rows in -> out, then the columns in place, a group of planes at a time: the column pass finds the row
        // pass's output of its group in the memory-side cache instead of HBM.
        const V *src = static_cast<const V *>(in);
        V *dst = static_cast<V *>(out);
        for (long long p = 0; p < batch && e == hipSuccess; p += group)
        {
            const long long n = std::min(group, batch - p);
            e = rows_any<SG, T>(nx, src + p * plane, dst + p * plane, n * ny, st);
            if (e == hipSuccess)
                e = columns_any<SG, T>(ny, dst + p * plane, nx, n, st);
        }
        return e;
    }
    // Four-step columns (in -> out, out -> scratch), then the rows scratch -> out, a group of planes at a time.
    float2 *scratch = nullptr;
    e = fft_plane_scratch(group * plane * static_cast<long long>(sizeof(V) / sizeof(float2)), scratch, 2);
    const V *src = static_cast<const V *>(in);
    V *dst = static_cast<V *>(out);
    for (long long p = 0; p < batch && e == hipSuccess; p += group)
    {
        const long long n = std::min(group, batch - p);
        e = long_columns<SG, T>(ny, src + p * plane, dst + p * plane, scratch, nx, n, st);
        if (e == hipSuccess)
            e = rows_any<SG, T>(nx, scratch, dst + p * plane, n * ny, st);
    }
    return e;
}
} // namespace

bool lds_fft_supported(int nx, int ny, bool fp64)
{
    if (!pow2_in(nx, 16, fp64 ? 8192 : 16384) || !pow2_in(ny, 16, 16384))
        return false;
    if (ny <= 512)
        return true; // one pass, or a column tile of at most 256 columns, which divides nx
    int n1, n2;
    column_split(ny, n1, n2);
    return nx % column_tile_of(n1, fp64) == 0 && nx % column_tile_of(n2, fp64) == 0;
}

bool small_fft_supported(int nx, int ny)
{
    return pow2_in(nx, 16, 512) && pow2_in(ny, 16, 512);
}

hipError_t launch_lds_fft2d(const void *in, void *out, int nx, int ny, long long batch, int sign, bool fp64,
                            hipStream_t st)
{
    if (!lds_fft_supported(nx, ny, fp64) || batch < 1 || (sign != 1 && sign != -1))
        return hipErrorInvalidValue;
    if (fp64)
        return sign > 0 ? lds_fft<1, double>(in, out, nx, ny, batch, st)
                        : lds_fft<-1, double>(in, out, nx, ny, batch, st);
    return sign > 0 ? lds_fft<1, float>(in, out, nx, ny, batch, st) : lds_fft<-1, float>(in, out, nx, ny, batch, st);
}
} // namespace idgx
