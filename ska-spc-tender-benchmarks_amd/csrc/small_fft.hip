// Batched 2-D DFTs of small planes for the plain FFT entry point (idgx_fft2d): sides nx, ny powers of two in
// 16 .. 512, FP32 or FP64, out[b][ky][kx] = sum in[b][y][x] exp(sign 2 pi i (ky y / ny + kx x / nx)), unnormalised.
//
// Everything happens in LDS tiles; every thread of a workgroup holds 16 complex points of a tile at a time.
// - A plane with both sides <= 128 whose padded copy fits 132 KB of LDS (128 x 128 in FP32, 128 x 64 in FP64, and
//   every smaller plane) takes one pass over HBM: a workgroup loads whole planes (several when a plane has fewer
//   than 4096 points), transforms their rows, then their columns, and writes them once.
// - Larger planes take two passes: rows (4096 / nx whole rows per tile, contiguous loads), then columns in place (a
//   tile of ny rows x C columns, C = max(4096 / ny, 128 bytes / point), so that every row segment is a whole 128-byte
//   line).
// A line of n points is transformed by decimation-in-frequency stages of radix 8 or 16 (n = 16: 16; 32: 8 x 4;
// 64: 8 x 8; 128: 16 x 8; 256: 16 x 16; 512: 8 x 8 x 8).  A stage's small DFTs run in registers, their twiddles
// W_L^(j k) come from a per-workgroup table of W_n^t (t < n, sincospi in FP64, rounded once for FP32), and the last
// stage writes its outputs to their natural positions, so that no digit reversal is left for the stores.
// Bank conflicts: consecutive lanes take consecutive lines; rows are padded to nx + 1 points, so that lanes on
// consecutive rows fall in different banks, and lanes on consecutive columns read consecutive addresses.
// Workgroups loop over tiles (grid-stride), so that the twiddle table is computed once per workgroup; barriers are
// LDS-only, so that a tile's global stores need not complete before the next tile's loads start.  (Loading the next
// tile into registers during the current one's transform measured no faster: 64 x 64 planes run at 4.15 TB/s of
// HBM traffic either way.)
#include "kernels.hpp"

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

// Radix sequence of an N-point line.
template <int N> struct Radices;
template <> struct Radices<16>
{
    static constexpr int count = 1, r0 = 16, r1 = 1, r2 = 1;
};
template <> struct Radices<32>
{
    static constexpr int count = 2, r0 = 8, r1 = 4, r2 = 1;
};
template <> struct Radices<64>
{
    static constexpr int count = 2, r0 = 8, r1 = 8, r2 = 1;
};
template <> struct Radices<128>
{
    static constexpr int count = 2, r0 = 16, r1 = 8, r2 = 1;
};
template <> struct Radices<256>
{
    static constexpr int count = 2, r0 = 16, r1 = 16, r2 = 1;
};
template <> struct Radices<512>
{
    static constexpr int count = 3, r0 = 8, r1 = 8, r2 = 8;
};

// Natural frequency of output 0 of the last stage's block blk.  After stages of radix r0, r1, ... the position
// d0 (N / r0) + d1 (N / r0 r1) + ... + ds holds frequency d0 + r0 d1 + r0 r1 d2 + ..., and blk = position / rs.
template <int N> __device__ __forceinline__ int natural_base(int blk)
{
    using Rs = Radices<N>;
    if constexpr (Rs::count == 1)
        return 0;
    else if constexpr (Rs::count == 2)
        return blk;
    else
        return blk / Rs::r1 + Rs::r0 * (blk % Rs::r1);
}

// One non-final DIF stage of radix R on blocks of L points, in place: for each line, block and j < Q = L / R, the R
// points at j + m Q become their DFT y_k times W_L^(j k) (tw[t] = W_N^t, so W_L^(j k) = tw[j k N / L]).
// Consecutive threads take consecutive lines.
template <int N, int R, int L, int NLINES, int THREADS, int SG, class V, class Addr>
__device__ __forceinline__ void dif_stage(V *lds, const V *tw, Addr addr)
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
            a[k] = cmul(a[k], tw[j * k * (N / L)]);
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

// Every stage on NLINES lines of N points; element e of line l is lds[addr(l, e)].  Natural order in and out; the
// caller has synchronised after filling the lines, and every stage ends with a barrier.
template <int N, int NLINES, int THREADS, int SG, class V, class Addr>
__device__ __forceinline__ void fft_lines(V *lds, const V *tw, Addr addr)
{
    using Rs = Radices<N>;
    if constexpr (Rs::count == 1)
        last_stage<N, Rs::r0, NLINES, THREADS, SG>(lds, addr);
    else if constexpr (Rs::count == 2)
    {
        dif_stage<N, Rs::r0, N, NLINES, THREADS, SG>(lds, tw, addr);
        last_stage<N, Rs::r1, NLINES, THREADS, SG>(lds, addr);
    }
    else
    {
        dif_stage<N, Rs::r0, N, NLINES, THREADS, SG>(lds, tw, addr);
        dif_stage<N, Rs::r1, N / Rs::r0, NLINES, THREADS, SG>(lds, tw, addr);
        last_stage<N, Rs::r2, NLINES, THREADS, SG>(lds, addr);
    }
}

// tw[t] = exp(SG 2 pi i t / n) for t < n, computed in FP64 and rounded once.
template <int SG, int THREADS, class V> __device__ __forceinline__ void twiddle_table(V *tw, int n)
{
    using T = decltype(tw->x);
    for (int t = threadIdx.x; t < n; t += THREADS)
    {
        double s, c;
        sincospi(2.0 * t / n, &s, &c);
        tw[t] = V{static_cast<T>(c), static_cast<T>(SG * s)};
    }
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
    twiddle_table<SG, THREADS>(tw, NX);
    twiddle_table<SG, THREADS>(tw + NX, NY);
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

// Row pass: kRowTile / NX whole rows (of any plane) per tile, in -> out.
template <int NX, int SG, class T>
__global__ __launch_bounds__(kRowTile / kPointsPerThread) void small_fft_rows_kernel(const typename Vec<T>::type *in,
                                                                                    typename Vec<T>::type *out,
                                                                                    long long rows)
{
    using V = typename Vec<T>::type;
    constexpr int ROWS = kRowTile / NX, THREADS = kRowTile / kPointsPerThread, PITCH = NX + 1;
    __shared__ V lds[ROWS * PITCH];
    __shared__ V tw[NX];
    twiddle_table<SG, THREADS>(tw, NX);
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
    twiddle_table<SG, THREADS>(tw, NY);
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

unsigned grid_for(long long tiles)
{
    return static_cast<unsigned>(tiles < kMaxGrid ? tiles : kMaxGrid);
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
    constexpr int ROWS = kRowTile / NX;
    small_fft_rows_kernel<NX, SG, T><<<grid_for((rows + ROWS - 1) / ROWS), kRowTile / kPointsPerThread, 0, st>>>(
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
        return rows_pass<256, SG, T>(in, out, rows, st);
    case 512:
        return rows_pass<512, SG, T>(in, out, rows, st);
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

bool small_side(int n)
{
    return n >= 16 && n <= 512 && (n & (n - 1)) == 0;
}

template <int SG, class T> hipError_t small_fft(const void *in, void *out, int nx, int ny, long long batch, hipStream_t st)
{
    hipError_t e = hipSuccess;
    if (planes_any<SG, T>(nx, ny, in, out, batch, st, e, std::make_integer_sequence<int, 4>()))
        return e;
    e = rows_any<SG, T>(nx, in, out, batch * ny, st);
    if (e == hipSuccess)
        e = columns_any<SG, T>(ny, out, nx, batch, st);
    return e;
}
} // namespace

bool small_fft_supported(int nx, int ny)
{
    return small_side(nx) && small_side(ny);
}

hipError_t launch_small_fft2d(const void *in, void *out, int nx, int ny, long long batch, int sign, bool fp64,
                              hipStream_t st)
{
    if (!small_fft_supported(nx, ny) || batch < 1 || (sign != 1 && sign != -1))
        return hipErrorInvalidValue;
    if (fp64)
        return sign > 0 ? small_fft<1, double>(in, out, nx, ny, batch, st)
                        : small_fft<-1, double>(in, out, nx, ny, batch, st);
    return sign > 0 ? small_fft<1, float>(in, out, nx, ny, batch, st) : small_fft<-1, float>(in, out, nx, ny, batch, st);
}
} // namespace idgx
