// Device helpers shared by the idgx kernels (gfx950).
//
// Geometry (pixel l/m/n and subgrid uvw offsets) is evaluated with the rounding points of the reference CPU
// code: these helpers are only ever included in translation units compiled under
// `#pragma clang fp contract(off)` for the parts that must round like the reference, and fused multiply-adds
// are written explicitly (__builtin_fmaf) where fusion is wanted.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "idgx.h"

namespace idgx
{
constexpr int kCorr = 4;

struct cfloat
{
    float re, im;
};

// ---- geometry --------------------------------------------------------------------------------------------------

// float(3.141592653589793), the reference's pi() (gridding/inc/math.hpp:13-16).
constexpr float kPiF = 3.14159274101257324f;

// l (or m) of pixel x: ((x + 0.5f) - S/2.f) * image_size / S, rounding after every step
// (gridding/inc/math.hpp:20-24).  Division and sqrt are IEEE correctly rounded in HIP device code
// (clang's -fhip-fp32-correctly-rounded-divide-sqrt is on by default), as on the host.
__device__ __forceinline__ float pixel_lm(int x, int S, float image_size)
{
#pragma clang fp contract(off)
    float a = static_cast<float>(x) + 0.5F;
    a = a - static_cast<float>(S) / 2.F;
    a = a * image_size;
    return a / static_cast<float>(S);
}

// n = r2 / (1 + sqrt(1 - r2)), clamped to 1 (gridding/inc/math.hpp:31-37).
__device__ __forceinline__ float pixel_n(float l, float m)
{
#pragma clang fp contract(off)
    const float r2 = (l * l) + (m * m);
    return r2 > 1.F ? 1.F : r2 / (1.F + __builtin_sqrtf(1.F - r2));
}

struct Offsets
{
    float u, v, w;
};

// Offsets as the reference CPU gridder rounds them: all in float (gridding/src/gridding_benchmark.cpp:68,74-81).
__device__ __forceinline__ Offsets gridder_offsets(const idgx_metadata &m, int S, int G, float image_size,
                                                   float w_step)
{
#pragma clang fp contract(off)
    const float w_lambda = static_cast<float>(static_cast<double>(w_step) * (m.z + 0.5));
    const float half_s = static_cast<float>(S) * 0.5F;
    const float half_g = static_cast<float>(G) * 0.5F;
    const float scale = 2.F * kPiF / image_size;
    Offsets o;
    o.u = ((static_cast<float>(m.x) + half_s) - half_g) * scale;
    o.v = ((static_cast<float>(m.y) + half_s) - half_g) * scale;
    o.w = (2.F * kPiF) * w_lambda;
    return o;
}

// Offsets as the reference CPU degridder rounds them: float base, double scale, back to float
// (gridding/src/degridding_benchmark.cpp:72,123-129).
__device__ __forceinline__ Offsets degridder_offsets(const idgx_metadata &m, int S, int G, float image_size,
                                                     float w_step)
{
#pragma clang fp contract(off)
    const float w_lambda = static_cast<float>(static_cast<double>(w_step) * (m.z + 0.5));
    const double scale = 2 * M_PI / static_cast<double>(image_size);
    const float bu = (static_cast<float>(m.x) + static_cast<float>(S) / 2) - static_cast<float>(G) / 2;
    const float bv = (static_cast<float>(m.y) + static_cast<float>(S) / 2) - static_cast<float>(G) / 2;
    Offsets o;
    o.u = static_cast<float>(static_cast<double>(bu) * scale);
    o.v = static_cast<float>(static_cast<double>(bv) * scale);
    o.w = static_cast<float>(2 * M_PI * static_cast<double>(w_lambda));
    return o;
}

// ---- sin/cos ---------------------------------------------------------------------------------------------------

enum SincosMode
{
    kSincosOcml = 0, // ROCm device library sincosf: full range reduction, ~1 ulp
    kSincosHw = 1    // reduction to a fraction of a turn in FP32 FMAs (two-constant 1/(2 pi)), then v_sin/v_cos_f32
};

// 1/(2 pi) split for the reduction: kInvTwoPi = float(1/(2 pi)), kInvTwoPiLo = 1/(2 pi) - kInvTwoPi.
constexpr float kInvTwoPi = 0.159154943091895336f;
constexpr float kInvTwoPiLo = 6.42063816e-09f;

template <int MODE> __device__ __forceinline__ void sincos_rad(float x, float &s, float &c)
{
    if constexpr (MODE == kSincosOcml)
    {
        sincosf(x, &s, &c);
    }
    else
    {
        // x / (2 pi) - k in revolutions: x kInvTwoPi - k is formed exactly inside the first FMA and rounded once
        // (|result| <= ~0.5), the low part of 1/(2 pi) added by the second; absolute error below 6e-8 revolutions
        // for |x| < 5e5, against 7e-8 for a Cody-Waite reduction in radians followed by a scaling to revolutions.
        const float k = __builtin_rintf(x * kInvTwoPi);
        const float rev = __builtin_fmaf(x, kInvTwoPiLo, __builtin_fmaf(x, kInvTwoPi, -k));
        s = __builtin_amdgcn_sinf(rev);
        c = __builtin_amdgcn_cosf(rev);
    }
}

// Workgroup b of n -> item, XCD-contiguous: workgroups reach the 8 XCDs round-robin (b % 8), so XCD x takes the
// items [x per + min(x, rem), ...) in order.  A bijection for any n (and any XCD count: only locality changes).
__device__ __forceinline__ int xcd_contiguous(int b, int n)
{
    constexpr int kXcds = 8;
    const int per = n / kXcds, rem = n % kXcds, x = b % kXcds, k = b / kXcds;
    return x * per + min(x, rem) + k;
}

// LDS-only workgroup barrier: this wave's LDS accesses are complete, then the workgroup meets.  Unlike
// __syncthreads() it does not wait for outstanding global loads, so a persistent kernel's prefetch of its next
// item stays in flight across the barriers of the current one (global memory is not exchanged through it).
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// ---- 2x2 Jones algebra, row-major {xx, xy, yx, yy} -------------------------------------------------------------

__device__ __forceinline__ cfloat cmul(cfloat a, cfloat b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cfloat cadd(cfloat a, cfloat b)
{
    return {a.re + b.re, a.im + b.im};
}
__device__ __forceinline__ cfloat cconj(cfloat a)
{
    return {a.re, -a.im};
}

// c = a * b
__device__ __forceinline__ void mat2_mul(const cfloat *a, const cfloat *b, cfloat *c)
{
    c[0] = cadd(cmul(a[0], b[0]), cmul(a[1], b[2]));
    c[1] = cadd(cmul(a[0], b[1]), cmul(a[1], b[3]));
    c[2] = cadd(cmul(a[2], b[0]), cmul(a[3], b[2]));
    c[3] = cadd(cmul(a[2], b[1]), cmul(a[3], b[3]));
}

__device__ __forceinline__ void mat2_herm(const cfloat *a, cfloat *h)
{
    h[0] = cconj(a[0]);
    h[1] = cconj(a[2]);
    h[2] = cconj(a[1]);
    h[3] = cconj(a[3]);
}

// P <- A1^H P A2 (gridder; gridding/inc/math.hpp:87-100).
__device__ __forceinline__ void apply_a1h_p_a2(cfloat *p, const cfloat *a1, const cfloat *a2)
{
    cfloat h[4], t[4];
    mat2_herm(a1, h);
    mat2_mul(h, p, t);
    mat2_mul(t, a2, p);
}

// P <- A1 P A2^H (degridder adjoint; gridding/inc/common_gridding.cuh:82-94).
__device__ __forceinline__ void apply_a1_p_a2h(cfloat *p, const cfloat *a1, const cfloat *a2)
{
    cfloat h[4], t[4];
    mat2_mul(a1, p, t);
    mat2_herm(a2, h);
    mat2_mul(t, h, p);
}

__device__ __forceinline__ void load_jones(const float4 *aterms, size_t pixel_index, cfloat *a)
{
    const float4 lo = aterms[2 * pixel_index];
    const float4 hi = aterms[2 * pixel_index + 1];
    a[0] = {lo.x, lo.y};
    a[1] = {lo.z, lo.w};
    a[2] = {hi.x, hi.y};
    a[3] = {hi.z, hi.w};
}

// ---- plan checks done on the device ----------------------------------------------------------------------------

struct PlanDev
{
    int nr_subgrids, S, G, nr_channels, nr_stations, nr_aterm_slots, max_nr_timesteps;
    long long nr_rows;
    float image_size, w_step;
    int convention;
};

// First row of a subgrid: (baseline_offset - metadata[0].baseline_offset) + time_offset
// (gridding/src/gridding_benchmark.cpp:36-37,61).
__device__ __forceinline__ long long first_row(const idgx_metadata *md, const idgx_metadata &m)
{
    return static_cast<long long>(m.baseline_offset) - md[0].baseline_offset + m.time_offset;
}

// True when the subgrid reads only inside the arrays described by the plan.
__device__ __forceinline__ bool subgrid_in_bounds(const PlanDev &p, const idgx_metadata &m, long long row0, int c0,
                                                  int nc)
{
    return m.nr_timesteps >= 0 && m.nr_timesteps <= p.max_nr_timesteps && row0 >= 0 &&
           row0 + m.nr_timesteps <= p.nr_rows && c0 >= 0 && nc >= 0 && c0 + nc <= p.nr_channels &&
           m.station1 < static_cast<uint32_t>(p.nr_stations) && m.station2 < static_cast<uint32_t>(p.nr_stations) &&
           m.aterm_index >= 0 && m.aterm_index < p.nr_aterm_slots;
}
} // namespace idgx
