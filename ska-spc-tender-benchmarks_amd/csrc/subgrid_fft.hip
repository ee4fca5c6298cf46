// Batched subgrid 2-D FFT in LDS for gfx950.
//
// The gridder's image-domain pixel x sits at l = (x + 1/2 - S/2) image_size / S, and uv cell k of a subgrid is
// at u = u_centre + (k - S/2) / image_size.  The transform that takes one to the other is
//     forward  (image -> uv):  G[k] = sum_x I[x] exp(+2 pi i (k - S/2)(x - S/2 + 1/2) / S)
//     backward (uv -> image):  I[x] = sum_k G[k] exp(-2 pi i (k - S/2)(x - S/2 + 1/2) / S)
// in each dimension (backward is the adjoint of forward; backward(forward(I)) = S^2 I).  With S even it is a
// plain DFT between two diagonal factors: forward = diag(post) DFT+ diag((-1)^x), post[k] = e^{i pi (k - S/2)
// (1 + 1/S)}; backward = diag((-1)^x) DFT- diag(conj(post)).
//
// Layout: one workgroup per subgrid (4 correlation planes) for S <= 32, one per plane above.  The planes are
// loaded with 16-byte loads, all in flight before the first LDS store, into LDS (rows padded by one element);
// every thread transforms one row held entirely in registers (a compile-time mixed radix-2/3 recursion,
// twiddles from an LDS table), then one column, and the result is stored the same way.  The pass is HBM-bound:
// 2 x 8 B per complex sample per direction.
#include "idgx_device.hpp"
#include "kernels.hpp"

namespace idgx
{
namespace
{
__device__ __forceinline__ cfloat cmulf(cfloat a, cfloat b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// out[k] = sum_n in[n * STRIDE] w^{SIGN k n}, w = e^{2 pi i / N}; tw[j] = e^{2 pi i j / S} with S = N * TWS.
template <int N, int STRIDE, int TWS, int SIGN>
__device__ __forceinline__ void fft_rec(const cfloat *in, cfloat *out, const float2 *tw)
{
    if constexpr (N == 1)
    {
        out[0] = in[0];
    }
    else if constexpr (N % 2 == 0)
    {
        cfloat e[N / 2], o[N / 2];
        fft_rec<N / 2, STRIDE * 2, TWS * 2, SIGN>(in, e, tw);
        fft_rec<N / 2, STRIDE * 2, TWS * 2, SIGN>(in + STRIDE, o, tw);
#pragma unroll
        for (int k = 0; k < N / 2; ++k)
        {
            const float2 w = tw[k * TWS];
            const cfloat t = cmulf({w.x, SIGN * w.y}, o[k]);
            out[k] = {e[k].re + t.re, e[k].im + t.im};
            out[k + N / 2] = {e[k].re - t.re, e[k].im - t.im};
        }
    }
    else
    {
        static_assert(N % 3 == 0, "subgrid sizes are 2^a 3^b");
        cfloat a[N / 3], b[N / 3], c[N / 3];
        fft_rec<N / 3, STRIDE * 3, TWS * 3, SIGN>(in, a, tw);
        fft_rec<N / 3, STRIDE * 3, TWS * 3, SIGN>(in + STRIDE, b, tw);
        fft_rec<N / 3, STRIDE * 3, TWS * 3, SIGN>(in + 2 * STRIDE, c, tw);
        constexpr float kHalf = -0.5F;
        constexpr float kS = SIGN * 0.866025403784438647F; // sin(2 pi / 3) with the transform's sign
#pragma unroll
        for (int k = 0; k < N / 3; ++k)
        {
            const float2 w1 = tw[k * TWS], w2 = tw[2 * k * TWS];
            const cfloat B = cmulf({w1.x, SIGN * w1.y}, b[k]);
            const cfloat Cc = cmulf({w2.x, SIGN * w2.y}, c[k]);
            const cfloat sum = {B.re + Cc.re, B.im + Cc.im};
            const cfloat dif = {B.re - Cc.re, B.im - Cc.im};
            // w3 = -1/2 + i kS, w3^2 = -1/2 - i kS
            const cfloat m = {a[k].re + kHalf * sum.re, a[k].im + kHalf * sum.im};
            const cfloat r = {-kS * dif.im, kS * dif.re}; // i kS (B - C)
            out[k] = {a[k].re + sum.re, a[k].im + sum.im};
            out[k + N / 3] = {m.re + r.re, m.im + r.im};
            out[k + 2 * N / 3] = {m.re - r.re, m.im - r.im};
        }
    }
}

// One plane of the workgroup: rows then columns, with the shift factors of the forward / backward transform.
template <int S, bool FORWARD>
__device__ __forceinline__ void transform_line(cfloat *line_lds, int step, const float2 *tw, const float2 *post)
{
    cfloat x[S], y[S];
#pragma unroll
    for (int j = 0; j < S; ++j)
        x[j] = line_lds[j * step];
    if constexpr (FORWARD)
    {
#pragma unroll
        for (int j = 1; j < S; j += 2)
            x[j] = {-x[j].re, -x[j].im};
        fft_rec<S, 1, 1, 1>(x, y, tw);
#pragma unroll
        for (int k = 0; k < S; ++k)
            y[k] = cmulf(y[k], {post[k].x, post[k].y});
    }
    else
    {
#pragma unroll
        for (int k = 0; k < S; ++k)
            x[k] = cmulf(x[k], {post[k].x, -post[k].y});
        fft_rec<S, 1, 1, -1>(x, y, tw);
#pragma unroll
        for (int j = 1; j < S; j += 2)
            y[j] = {-y[j].re, -y[j].im};
    }
#pragma unroll
    for (int j = 0; j < S; ++j)
        line_lds[j * step] = y[j];
}

template <int S, int PLANES, bool FORWARD>
__global__ __launch_bounds__(S *PLANES) void subgrid_fft_kernel(float2 *subgrids, long long nr_planes)
{
    constexpr int LD = S + 1; // padded row length
    constexpr int kThreadsFft = S * PLANES;
    constexpr int kPairs = PLANES * S * S / 2;      // float4 (two complex samples) per workgroup
    constexpr int kPerThread = kPairs / kThreadsFft; // = S / 2
    __shared__ cfloat s_data[PLANES * S * LD];
    __shared__ float2 s_tw[S];
    __shared__ float2 s_post[S];

    const int tid = threadIdx.x;
    const long long plane0 = static_cast<long long>(blockIdx.x) * PLANES;
    // Twiddles e^{2 pi i k / S} and the output factors e^{i pi (k - S/2)(1 + 1/S)} = (-1)^(k - S/2)
    // e^{i pi (k - S/2) / S}, from exactly reduced arguments.
    for (int k = tid; k < S; k += kThreadsFft)
    {
        float sn, cs;
        sincospif(2.F * static_cast<float>(k) / static_cast<float>(S), &sn, &cs);
        s_tw[k] = make_float2(cs, sn);
        sincospif(static_cast<float>(k - S / 2) / static_cast<float>(S), &sn, &cs);
        const float sg = ((k - S / 2) & 1) ? -1.F : 1.F;
        s_post[k] = make_float2(sg * cs, sg * sn);
    }
    // Coalesced 16-byte loads of PLANES consecutive planes, all issued before the first LDS store (only planes
    // < nr_planes exist; the others read as zero).
    const float4 *src = reinterpret_cast<const float4 *>(subgrids) + plane0 * (S * S / 2);
    float4 v[kPerThread];
#pragma unroll
    for (int j = 0; j < kPerThread; ++j)
    {
        const int i = tid + j * kThreadsFft;
        const int p = (2 * i) / (S * S);
        v[j] = plane0 + p < nr_planes ? src[i] : make_float4(0.F, 0.F, 0.F, 0.F);
    }
#pragma unroll
    for (int j = 0; j < kPerThread; ++j)
    {
        const int e = 2 * (tid + j * kThreadsFft);
        const int p = e / (S * S), rem = e - p * S * S;
        const int r = rem / S, c = rem - r * S;
        s_data[(p * S + r) * LD + c] = {v[j].x, v[j].y};
        s_data[(p * S + r) * LD + c + 1] = {v[j].z, v[j].w};
    }
    __syncthreads();
    const int p = tid / S, line = tid - p * S;
    transform_line<S, FORWARD>(&s_data[(p * S + line) * LD], 1, s_tw, s_post); // row `line`
    __syncthreads();
    transform_line<S, FORWARD>(&s_data[p * S * LD + line], LD, s_tw, s_post); // column `line`
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPerThread; ++j)
    {
        const int e = 2 * (tid + j * kThreadsFft);
        const int q = e / (S * S), rem = e - q * S * S;
        const int r = rem / S, c = rem - r * S;
        const cfloat a = s_data[(q * S + r) * LD + c], b = s_data[(q * S + r) * LD + c + 1];
        v[j] = make_float4(a.re, a.im, b.re, b.im);
    }
    float4 *dst = reinterpret_cast<float4 *>(subgrids) + plane0 * (S * S / 2);
#pragma unroll
    for (int j = 0; j < kPerThread; ++j)
    {
        const int i = tid + j * kThreadsFft;
        if (plane0 + (2 * i) / (S * S) < nr_planes)
            dst[i] = v[j];
    }
}

template <int S> hipError_t fft_launch(float2 *subgrids, long long planes, bool forward, hipStream_t st)
{
    constexpr int PLANES = S <= 32 ? 4 : 1;
    const long long blocks = (planes + PLANES - 1) / PLANES;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    if (forward)
        subgrid_fft_kernel<S, PLANES, true><<<static_cast<unsigned>(blocks), S * PLANES, 0, st>>>(subgrids, planes);
    else
        subgrid_fft_kernel<S, PLANES, false><<<static_cast<unsigned>(blocks), S * PLANES, 0, st>>>(subgrids, planes);
    return hipGetLastError();
}
} // namespace

bool subgrid_fft_supported(int S)
{
    return S == 8 || S == 12 || S == 16 || S == 24 || S == 32 || S == 48 || S == 64;
}

hipError_t launch_subgrid_fft(void *subgrids, long long nr_subgrids, int S, bool forward, hipStream_t st)
{
    float2 *d = static_cast<float2 *>(subgrids);
    const long long planes = nr_subgrids * kCorr;
    switch (S)
    {
    case 8:
        return fft_launch<8>(d, planes, forward, st);
    case 12:
        return fft_launch<12>(d, planes, forward, st);
    case 16:
        return fft_launch<16>(d, planes, forward, st);
    case 24:
        return fft_launch<24>(d, planes, forward, st);
    case 32:
        return fft_launch<32>(d, planes, forward, st);
    case 48:
        return fft_launch<48>(d, planes, forward, st);
    case 64:
        return fft_launch<64>(d, planes, forward, st);
    default:
        return hipErrorInvalidValue;
    }
}
} // namespace idgx
