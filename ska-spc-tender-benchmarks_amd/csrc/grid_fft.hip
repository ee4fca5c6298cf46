// Master-grid transform and w-stacking for gfx950.
//
// Master grid: complex64 [nr_layers][4][G][G].  Cell K of a row is at u = (K - G/2) / image_size (the adder puts
// subgrid cell k of a subgrid with corner x0 at K = x0 + k, and the subgrid FFT centres cell k at
// u_centre + (k - S/2) / image_size with u_centre = (x0 + S/2 - G/2) / image_size).  Image pixel X is at
// l = (X - G/2) image_size / G.  The transforms, in each dimension, are
//     FORWARD  (image -> uv):  G[K] = sum_X I[X] exp(+2 pi i (K - G/2)(X - G/2) / G)
//     BACKWARD (uv -> image):  I[X] = sum_K G[K] exp(-2 pi i (K - G/2)(X - G/2) / G)
// (the signs of the subgrid FFT, unnormalised).  Since (K - G/2)(X - G/2) = K X - (G/2)(K + X) + G^2/4, they are a
// plain DFT between two checkerboard sign flips (-1)^(x + y) when G is a multiple of 4; the DFT itself is rocFFT
// (hipFFT C2C, batched over correlations and layers).
//
// W-stacking.  A subgrid with w-layer z is gridded relative to w_z = (z + 0.5) w_step (in wavelengths), so layer
// z of the dirty image carries exp(+2 pi i w_z n(l, m)) relative to the true one, n = 1 - sqrt(1 - l^2 - m^2)
// written as (l^2 + m^2) / (1 + sqrt(1 - l^2 - m^2)) (gridding/inc/math.hpp:29-37).  to_image sums the layers
// after multiplying each by exp(-2 pi i w_z n); from_image is its adjoint.  The phases are reduced in FP64.
#include <hipfft/hipfft.h>

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

// hipFFT plans per (device, G, batch), created on first use and kept for the life of the process.
std::mutex g_plan_mutex;
std::map<std::tuple<int, int, long long>, hipfftHandle> g_plans;

hipError_t fft_plan(int G, long long batch, hipfftHandle &out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    std::lock_guard<std::mutex> lock(g_plan_mutex);
    const auto key = std::make_tuple(dev, G, batch);
    auto it = g_plans.find(key);
    if (it != g_plans.end())
    {
        out = it->second;
        return hipSuccess;
    }
    hipfftHandle h;
    int n[2] = {G, G};
    if (hipfftPlanMany(&h, 2, n, nullptr, 1, G * G, nullptr, 1, G * G, HIPFFT_C2C, static_cast<int>(batch)) !=
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

hipError_t launch_grid_fft(void *grid, int G, long long nr_layers, bool forward, hipStream_t st)
{
    const long long planes = nr_layers * kCorr;
    if (planes > 0x7fffffffLL || static_cast<long long>(G) * G > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    hipfftHandle plan;
    hipError_t e = fft_plan(G, planes, plan);
    if (e != hipSuccess)
        return e;
    e = checkerboard(grid, planes, G, st);
    if (e != hipSuccess)
        return e;
    if (hipfftSetStream(plan, st) != HIPFFT_SUCCESS)
        return hipErrorUnknown;
    // FORWARD has the + sign: hipFFT's BACKWARD (exp(+2 pi i ...)); BACKWARD the - sign: hipFFT's FORWARD.
    hipfftComplex *d = static_cast<hipfftComplex *>(grid);
    if (hipfftExecC2C(plan, d, d, forward ? HIPFFT_BACKWARD : HIPFFT_FORWARD) != HIPFFT_SUCCESS)
        return hipErrorUnknown;
    return checkerboard(grid, planes, G, st);
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
} // namespace idgx
