// Mixed-precision gridder and degridder for gfx950: FP64 phase-argument reduction, FP32 phasors, and the
// phasor x data contraction on v_mfma_f32_16x16x32_{bf16,f16} with FP32 accumulation.
//
// Phase reduction.  With l = (x' + 1/2) delta, x' = x - S/2 (an integer), delta = image_size / S, and
//     du = k_c u_t / 2pi - u_off,  dv = ...,  dw = k_c w_t / 2pi - w_off      (wavelengths, formed in FP64,
// u_off = (x0 + S/2 - G/2) / image_size the subgrid's phase centre), the phase in revolutions is
//     du l + dv m + dw n = qu x' + qv y' + dw n + (qu + qv) / 2,     qu = du delta, qv = dv delta.
// Because x', y' are integers, qu and qv can be reduced mod 1 in FP64 without changing the phasor, and the
// constant (qu + qv)/2 mod 1 becomes one rotation per visibility (folded into the gridder's B operand, or
// applied to the degridder's output).  The per-pair phase is then bounded by ~S/2 + |dw n| revolutions for any
// input, so three FP32 operations evaluate it to ~1e-6 revolutions and v_sin_f32/v_cos_f32 (domain +-256
// revolutions) take it directly.  This replaces the reference's FP32 phase_offset - k (u l + v m + w n)
// (gridding/src/gridding_benchmark.cpp:112-121), whose two large terms cancel.
//
// Contraction as a real GEMM.  pixel[p][corr] = sum_v phasor[p][v] * vis[v][corr] (gridder) is
//     D[p][n] = sum_k A[p][k] B[k][n],   A[p][2v] = cos, A[p][2v+1] = sin,
//     B[2v][re] = a, B[2v+1][re] = -b, B[2v][im] = b, B[2v+1][im] = a   (vis = a + ib, 4 correlations x re/im)
// which fills 8 of the 16 MFMA columns; the other 8 carry the low half of B, so one MFMA with A_hi gives
// A_hi*B_hi + A_hi*B_lo and one with A_lo gives A_lo*B_hi (+ A_lo*B_lo): the three-product split in two
// instructions.  Columns n and n+8 are summed at the end.  The degridder is the same with pixels and
// visibilities exchanging roles (vis[v][corr] = sum_p phasor[v][p] * pixel[p][corr]).
//
// Operand layout of v_mfma_f32_16x16x32 (one wave): lane l holds A[l%16][8*(l/16) .. +7], B[8*(l/16) .. +7][l%16]
// and D[4*(l/16) .. +3][l%16].  So a lane evaluates phasors of one row (pixel, or visibility) against four
// k-pairs, packs (cos, sin) pairs straight into A registers, and reads its 16-byte B fragment from LDS.
//
// Phases.  By default the VALU forms phi = qu x' + qv y' + dw n with two FMAs per pair: the gridder's pixel
// mapping keeps a lane's column x' fixed across its tiles (strips of 16 / 8 / 4 / 2 columns), and the
// degridder's pixel slots run in rows padded to a multiple of 4 so the four pixels of a lane group share y', so
// one of the three terms is formed once per visibility.  The PM shapes take the phases from the matrix core
// instead (phi is a rank-3 product: one v_mfma_f32_16x16x4_f32 per 16 x 16 tile, whose D layout, lane l holding
// D[4*(l/16)+j][l%16], is exactly the A layout above); that frees two VALU ops per pair but measured slower, so
// the autotuner rarely picks it.
//
// hi/lo split (FP16): hi = v_cvt_pk_f16_f32 of (cos, sin); lo = the exact residual from v_fma_mix_f32 (reading
// the half straight from the packed register), packed again: 4 VALU ops per pair.  FP16 B operands are scaled
// by a power of two so that their largest value is ~2^14 (per LDS batch, or per subgrid for prepared degridder
// pixels), which keeps any data range in range; the accumulators are rescaled exactly when the scale grows.
//
// Workgroup numbering spreads the workgroups of one subgrid over the XCDs (block b = sub_block * n + s): subgrid
// costs are uneven, and grouping a subgrid's workgroups on one XCD (for L2 reuse) measured 1.2-2x slower because
// the round-robin dispatch then waits on the most loaded XCD.
#include "idgx_device.hpp"
#include "kernels.hpp"

namespace idgx
{
namespace
{
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

enum DType
{
    kBF16 = 0,
    kFP16 = 1
};

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
// Default tile shape: gridder 8 tiles of 16 pixels per wave with 256 visibilities per LDS batch, degridder 4 tiles
// of 16 visibilities per wave with 256 pixels per LDS batch.
constexpr int kDefaultPT = 8, kDefaultVB = 256, kDefaultVT = 4, kDefaultPB = 256;

// Round a pair of floats to two 16-bit values (RNE), x0 in the low half.
template <int DT> __device__ __forceinline__ uint32_t pack2(float x0, float x1)
{
    if constexpr (DT == kBF16)
        return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){x0, x1}, bf16x2));
    else
        return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){x0, x1}, f16x2));
}

template <int DT> __device__ __forceinline__ void unpack2(uint32_t p, float &y0, float &y1)
{
    if constexpr (DT == kBF16)
    {
        y0 = __builtin_bit_cast(float, p << 16);
        y1 = __builtin_bit_cast(float, p & 0xffff0000u);
    }
    else
    {
        const f16x2 h = __builtin_bit_cast(f16x2, p);
        y0 = static_cast<float>(h[0]);
        y1 = static_cast<float>(h[1]);
    }
}

// x - float(half of h) in one v_fma_mix_f32 (the half is read straight from the packed register).  The result
// is exact: x and its RNE fp16 rounding are close enough for the difference to be a float.  (Writing the rounded
// residual straight into a register half with v_fma_mixlo/mixhi_f16 saves the pack but measured slower.)
__device__ __forceinline__ float sub_f16_lo(float x, uint32_t h)
{
    float r;
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(x));
    return r;
}
__device__ __forceinline__ float sub_f16_hi(float x, uint32_t h)
{
    float r;
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(x));
    return r;
}

// hi = round(x), lo = round(x - hi), both packed as pairs.
template <int DT> __device__ __forceinline__ void split2(float x0, float x1, uint32_t &hi, uint32_t &lo)
{
    hi = pack2<DT>(x0, x1);
    if constexpr (DT == kFP16)
    {
        lo = pack2<DT>(sub_f16_lo(x0, hi), sub_f16_hi(x1, hi));
    }
    else
    {
        float h0, h1;
        unpack2<DT>(hi, h0, h1);
        lo = pack2<DT>(x0 - h0, x1 - h1);
    }
}

// FP16 B operands are scaled by a power of two per LDS batch so that the largest staged value is about 2^14:
// no overflow (65504) for any data, and the hi/lo split keeps its relative precision.  The accumulators are
// rescaled exactly when a later batch needs a larger scale, and the final result is multiplied back.
constexpr int kF16TargetExp = 14;

__device__ __forceinline__ int scale_exponent(float maxabs)
{
    if (!(maxabs > 0.F))
        return -126;
    int e;
    (void)frexpf(maxabs, &e); // maxabs in [2^(e-1), 2^e)
    return e - kF16TargetExp;
}

// Phases of a 16 x 16 tile: D[i][j] = sum_k A[i][k] B[k][j], K = 4 (fourth component zero).
__device__ __forceinline__ f32x4 phase_mfma(float a, float b)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, (f32x4){0.F, 0.F, 0.F, 0.F}, 0, 0, 0);
}


__device__ __forceinline__ f32x4 mfma(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, u32x4 b, f32x4 c,
                                      std::integral_constant<int, kBF16>)
{
    const u32x4 a = {a0, a1, a2, a3};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

__device__ __forceinline__ f32x4 mfma(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, u32x4 b, f32x4 c,
                                      std::integral_constant<int, kFP16>)
{
    const u32x4 a = {a0, a1, a2, a3};
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0,
                                                  0, 0);
}

// Writes the B-operand words of one k-pair item (a visibility for the gridder, a pixel for the degridder) with
// complex values v[0..3] into a 1 KiB LDS block laid out as [lane][4 words]: item q of the block occupies word
// (q % 4) of lanes (q / 4) * 16 + n, n = column.  Columns 0-7 carry the high parts, 8-15 the low parts.
// Packed (re, im) high and low parts of the four correlations (low parts zero without the split).
template <int DT, bool SPLIT>
__device__ __forceinline__ void split_item(const cfloat *v, uint32_t *hi, uint32_t *lo)
{
#pragma unroll
    for (int pol = 0; pol < 4; ++pol)
    {
        split2<DT>(v[pol].re, v[pol].im, hi[pol], lo[pol]);
        if (!SPLIT)
            lo[pol] = 0;
    }
}

__device__ __forceinline__ void write_b_words(uint32_t *blk, int q, const uint32_t *hip, const uint32_t *lop)
{
    const int kg = q >> 2, slot = q & 3;
    uint32_t *base = blk + kg * 64 + slot;
#pragma unroll
    for (int pol = 0; pol < 4; ++pol)
    {
        const uint32_t hi = hip[pol], lo = lop[pol];
        // hi = (re | im << 16).  Column re: (cos row re, sin row -im); column im: (cos row im, sin row re).
        const uint32_t re_col_hi = (hi & 0xffffu) | ((hi ^ 0x80000000u) & 0xffff0000u);
        const uint32_t im_col_hi = (hi >> 16) | (hi << 16);
        const uint32_t re_col_lo = (lo & 0xffffu) | ((lo ^ 0x80000000u) & 0xffff0000u);
        const uint32_t im_col_lo = (lo >> 16) | (lo << 16);
        base[(2 * pol) * 4] = re_col_hi;
        base[(2 * pol + 1) * 4] = im_col_hi;
        base[(8 + 2 * pol) * 4] = re_col_lo;
        base[(8 + 2 * pol + 1) * 4] = im_col_lo;
    }
}

template <int DT, bool SPLIT> __device__ __forceinline__ void write_b_item(uint32_t *blk, int q, const cfloat *v)
{
    uint32_t hi[4], lo[4];
    split_item<DT, SPLIT>(v, hi, lo);
    write_b_words(blk, q, hi, lo);
}

__device__ __forceinline__ void channel_window(const KernelArgs &a, int s, int &c0, int &nc)
{
    if (a.channels)
    {
        const idgx_channels w = a.channels[s];
        c0 = w.channel_begin;
        nc = w.nr_channels;
    }
    else
    {
        c0 = 0;
        nc = a.p.nr_channels;
    }
}

struct CentreOffsets
{
    double u, v, w; // wavelengths
};

__device__ __forceinline__ CentreOffsets centre_offsets(const PlanDev &p, const idgx_metadata &m)
{
    const double inv_image = 1.0 / static_cast<double>(p.image_size);
    CentreOffsets o;
    o.u = (m.x + p.S * 0.5 - p.G * 0.5) * inv_image;
    o.v = (m.y + p.S * 0.5 - p.G * 0.5) * inv_image;
    o.w = static_cast<double>(p.w_step) * (m.z + 0.5);
    return o;
}

constexpr double kInv2Pi = 0.15915494309189533577;

// Reduced phase coordinates of one visibility (revolutions), signed by sgn: per-pixel-step slopes qu, qv in
// [-1/2, 1/2], the w slope dw per unit n, and the constant rot = (qu + qv)/2 mod 1.
struct Reduced
{
    float qu, qv, dw, rot;
};

__device__ __forceinline__ Reduced reduce_visibility(const float *uvw, float k, const CentreOffsets &o, double delta,
                                                     float sgn)
{
    const double kk = static_cast<double>(k) * kInv2Pi;
    const double qu = __builtin_fma(kk, static_cast<double>(uvw[0]), -o.u) * delta;
    const double qv = __builtin_fma(kk, static_cast<double>(uvw[1]), -o.v) * delta;
    const double dw = __builtin_fma(kk, static_cast<double>(uvw[2]), -o.w);
    double rot = 0.5 * (qu + qv);
    rot -= __builtin_rint(rot);
    Reduced r;
    r.qu = sgn * static_cast<float>(qu - __builtin_rint(qu));
    r.qv = sgn * static_cast<float>(qv - __builtin_rint(qv));
    r.dw = sgn * static_cast<float>(dw);
    r.rot = sgn * static_cast<float>(rot);
    return r;
}

// c * e^{2 pi i rot} for the four correlations.
__device__ __forceinline__ void rotate4(cfloat *v, float rot)
{
    const float sn = __builtin_amdgcn_sinf(rot), cs = __builtin_amdgcn_cosf(rot);
#pragma unroll
    for (int pol = 0; pol < 4; ++pol)
        v[pol] = {v[pol].re * cs - v[pol].im * sn, v[pol].re * sn + v[pol].im * cs};
}

__device__ __forceinline__ float pixel_lm_f(int x, int S, float image_size)
{
    return (static_cast<float>(x) + 0.5F - 0.5F * static_cast<float>(S)) * (image_size / static_cast<float>(S));
}

// x' = x - S/2, exact in float.
__device__ __forceinline__ float pixel_offset(int x, int S)
{
    return static_cast<float>(x - S / 2);
}

__device__ __forceinline__ float pixel_n_f(float l, float m)
{
    const float r2 = l * l + m * m;
    return r2 > 1.F ? 1.F : r2 / (1.F + __builtin_sqrtf(1.F - r2));
}

// The visibility v of subgrid s: (t, c) relative to its first row / channel window, and the subgrid's count.
struct Selection
{
    const uint32_t *index; // null: dense (t, c) = (v / nc, v % nc)
    int nc;
    __device__ __forceinline__ void at(int v, int &t, int &c) const
    {
        if (index)
        {
            const uint32_t e = index[v];
            t = static_cast<int>(e >> 16);
            c = static_cast<int>(e & 0xffffu);
        }
        else
        {
            t = v / nc;
            c = v - t * nc;
        }
    }
};

__device__ __forceinline__ int selection_of(const KernelArgs &a, int s, int nr_timesteps, int nc, Selection &sel)
{
    sel.nc = nc;
    if (a.vis_index)
    {
        const long long o0 = a.vis_offset[s], o1 = a.vis_offset[s + 1];
        sel.index = a.vis_index + o0;
        return static_cast<int>(o1 - o0);
    }
    sel.index = nullptr;
    return nr_timesteps * nc;
}

// Selected entries must lie inside the subgrid's window (entries outside are treated as empty visibilities).
__device__ __forceinline__ bool entry_ok(int t, int c, int nr_timesteps, int nc)
{
    return t >= 0 && t < nr_timesteps && c >= 0 && c < nc;
}

// Largest value over the workgroup (non-negative floats order like their bit patterns).
__device__ __forceinline__ float block_max(float v, unsigned &slot)
{
    if (threadIdx.x == 0)
        slot = 0u;
    __syncthreads();
    atomicMax(&slot, __float_as_uint(v));
    __syncthreads();
    const float m = __uint_as_float(slot);
    __syncthreads();
    return m;
}

// Degridder input pixel p: taper x (A1 P A2^H, or A1^H P A2 in the reference convention) of the subgrid values.
__device__ __forceinline__ void degrid_pixel(const KernelArgs &a, const idgx_metadata &m, const float2 *sub, int p,
                                             cfloat *px)
{
    const int P = a.p.S * a.p.S;
    const float taper = a.spheroidal[p];
#pragma unroll
    for (int pol = 0; pol < 4; ++pol)
    {
        const float2 g = sub[pol * P + p];
        px[pol] = {taper * g.x, taper * g.y};
    }
    const float4 *aterms = reinterpret_cast<const float4 *>(a.aterms);
    const size_t abase = static_cast<size_t>(m.aterm_index) * a.p.nr_stations;
    cfloat a1[4], a2[4];
    load_jones(aterms, (abase + m.station1) * P + p, a1);
    load_jones(aterms, (abase + m.station2) * P + p, a2);
    if (a.p.convention == IDGX_ATERM_REFERENCE)
        apply_a1h_p_a2(px, a1, a2);
    else
        apply_a1_p_a2h(px, a1, a2);
}

template <int PT, int VB> struct GridderShared
{
    unsigned maxbits;
    u32x4 b[(VB / 16) * 64];
    float coef[(VB / 16) * 64];
    float pix[kWaves * PT * 16 * 8];
};

template <int VT, int PB> struct DegridderShared
{
    unsigned maxbits;
    u32x4 b[(PB / 16) * 64];
    float coord[(PB / 16) * 64];
    float vis[kWaves * VT * 16 * 8];
};

// ---------------------------------------------------------------------------------------------------------------
// Gridder item: 4 waves x PT tiles x 16 pixels (block pblk) of subgrid s; visibilities staged VB at a time.
// ---------------------------------------------------------------------------------------------------------------
template <int DT, bool SPLIT, int PT, int VB, bool PM, class Shared>
__device__ __forceinline__ void gridder_item(const KernelArgs &a, int s, int pblk, Shared &sm)
{
    constexpr int kBlocks = VB / 16;
    constexpr int kPix = kWaves * PT * 16;
    static_assert(sizeof(sm.b) == sizeof(u32x4) * kBlocks * 64 && sizeof(sm.pix) >= sizeof(float) * kPix * 8,
                  "LDS staging of the item's shape");
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int S = a.p.S, P = S * S;
    const idgx_metadata m = a.metadata[s];
    int c0, nc;
    channel_window(a, s, c0, nc);
    const long long row0 = first_row(a.metadata, m);
    const bool ok = subgrid_in_bounds(a.p, m, row0, c0, nc);
    const long long C = a.p.nr_channels;

    // This wave's unit: a strip of gu.wd columns by gu.rows rows; lane row r of tile t is pixel (px, y0 + t*rt).
    const GridUnits gu = grid_units(S, PT);
    const int unit = kWaves * pblk + wave;
    const bool active = unit < gu.units; // wave-uniform
    const int px = (unit % gu.ns) * gu.wd + (lane & 15) % gu.wd;
    const int py0 = (unit / gu.ns) * gu.rows + (lane & 15) / gu.wd;

    // Pixel coordinates of row lane%16 of each tile: x' (the same in every tile), y', n; for PM only component
    // lane/16 of them (the B operand of the phase MFMA).  Rows past the subgrid repeat the last row.
    const float xp = pixel_offset(px, S);
    const float lx = pixel_lm_f(px, S, a.p.image_size);
    float coord[PT], yp[PT], nn[PT];
#pragma unroll
    for (int t = 0; t < PT; ++t)
    {
        const int y = min(py0 + t * gu.rt, S - 1);
        const int kk = lane >> 4;
        yp[t] = pixel_offset(y, S);
        nn[t] = pixel_n_f(lx, pixel_lm_f(y, S, a.p.image_size));
        coord[t] = kk == 0 ? xp : kk == 1 ? yp[t] : kk == 2 ? nn[t] : 0.F;
    }
    f32x4 acc[PT];
#pragma unroll
    for (int t = 0; t < PT; ++t)
        acc[t] = (f32x4){0.F, 0.F, 0.F, 0.F};
    int scale_exp = -126; // FP16: current power-of-two scale of the staged B operand

    Selection sel;
    const int V = ok ? selection_of(a, s, m.nr_timesteps, nc, sel) : 0;
    const CentreOffsets off = centre_offsets(a.p, m);
    const double delta = static_cast<double>(a.p.image_size) / S;
    const float4 *vis = reinterpret_cast<const float4 *>(a.visibilities);
    uint32_t *s_bw = reinterpret_cast<uint32_t *>(sm.b);

    for (int v0 = 0; v0 < V; v0 += VB)
    {
        __syncthreads();
        // Stage visibilities: reduced phase coefficients and B-operand words (scaled for FP16).
        constexpr int kPer = (VB + kThreads - 1) / kThreads;
        cfloat val[kPer][4];
        Reduced rd[kPer];
        float vmax = 0.F;
#pragma unroll
        for (int i = 0; i < kPer; ++i)
        {
            const int q = tid + i * kThreads;
            const int v = v0 + q;
#pragma unroll
            for (int pol = 0; pol < 4; ++pol)
                val[i][pol] = {0.F, 0.F};
            rd[i] = {0.F, 0.F, 0.F, 0.F};
            int t = 0, c = 0;
            if (q < VB && v < V)
                sel.at(v, t, c);
            if (q < VB && v < V && entry_ok(t, c, m.nr_timesteps, nc))
            {
                const long long r = row0 + t;
                rd[i] = reduce_visibility(a.uvw + 3 * r, a.wavenumbers[c0 + c], off, delta, -1.F);
                const float4 x0 = vis[(r * C + c0 + c) * 2], x1 = vis[(r * C + c0 + c) * 2 + 1];
                val[i][0] = {x0.x, x0.y};
                val[i][1] = {x0.z, x0.w};
                val[i][2] = {x1.x, x1.y};
                val[i][3] = {x1.z, x1.w};
                rotate4(val[i], rd[i].rot);
#pragma unroll
                for (int pol = 0; pol < 4; ++pol)
                    vmax = fmaxf(vmax, fmaxf(fabsf(val[i][pol].re), fabsf(val[i][pol].im)));
            }
        }
        if constexpr (DT == kFP16)
        {
            const int e_new = max(scale_exp, scale_exponent(block_max(vmax, sm.maxbits)));
            if (e_new > scale_exp)
            {
                const float f = ldexpf(1.F, scale_exp - e_new);
#pragma unroll
                for (int t = 0; t < PT; ++t)
                    acc[t] *= f;
                scale_exp = e_new;
            }
            const float inv = ldexpf(1.F, -scale_exp);
#pragma unroll
            for (int i = 0; i < kPer; ++i)
#pragma unroll
                for (int pol = 0; pol < 4; ++pol)
                    val[i][pol] = {val[i][pol].re * inv, val[i][pol].im * inv};
        }
#pragma unroll
        for (int i = 0; i < kPer; ++i)
        {
            const int q = tid + i * kThreads;
            if (q >= VB)
                break;
            const int blk = q >> 4, qq = q & 15;
            write_b_item<DT, SPLIT>(s_bw + blk * 256, qq, val[i]);
            if constexpr (PM)
            {
                float *cf = sm.coef + blk * 64 + qq;
                cf[0] = rd[i].qu;
                cf[16] = rd[i].qv;
                cf[32] = rd[i].dw;
                cf[48] = 0.F;
            }
            else
            {
                float *cf = sm.coef + blk * 64 + (qq >> 2) * 16 + (qq & 3);
                cf[0] = rd[i].qu;
                cf[4] = rd[i].qv;
                cf[8] = rd[i].dw;
            }
        }
        __syncthreads();
        const int nblk = active ? min(kBlocks, (V - v0 + 15) >> 4) : 0;
        for (int b = 0; b < nblk; ++b)
        {
            const u32x4 bf = sm.b[b * 64 + lane];
            f32x4 phs[PT];
            if constexpr (PM)
            {
                const float coef = sm.coef[b * 64 + lane];
#pragma unroll
                for (int t = 0; t < PT; ++t)
                    phs[t] = phase_mfma(coef, coord[t]); // phi(visibility 4*(lane/16)+j, pixel lane%16)
            }
            else
            {
                // phi = qu x' + qv y' + dw n with the x' term formed once per visibility: 2 FMAs per phase.
                const f32x4 *cf = reinterpret_cast<const f32x4 *>(sm.coef + b * 64 + (lane >> 4) * 16);
                const f32x4 du = cf[0], dv = cf[1], dw = cf[2];
                f32x4 ux;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    ux[j] = du[j] * xp;
#pragma unroll
                for (int t = 0; t < PT; ++t)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        phs[t][j] = __builtin_fmaf(dv[j], yp[t], __builtin_fmaf(dw[j], nn[t], ux[j]));
            }
#pragma unroll
            for (int t = 0; t < PT; ++t)
            {
                uint32_t hi[4], lo[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                {
                    const float sn = __builtin_amdgcn_sinf(phs[t][j]);
                    const float cs = __builtin_amdgcn_cosf(phs[t][j]);
                    if constexpr (SPLIT)
                        split2<DT>(cs, sn, hi[j], lo[j]);
                    else
                        hi[j] = pack2<DT>(cs, sn);
                }
                acc[t] = mfma(hi[0], hi[1], hi[2], hi[3], bf, acc[t], std::integral_constant<int, DT>());
                if constexpr (SPLIT)
                    acc[t] = mfma(lo[0], lo[1], lo[2], lo[3], bf, acc[t], std::integral_constant<int, DT>());
            }
        }
    }

    // Sum the high and low column halves, gather each pixel's 8 values in LDS, apply A-term and taper.
    __syncthreads();
    const float unscale = DT == kFP16 ? ldexpf(1.F, scale_exp) : 1.F;
#pragma unroll
    for (int t = 0; t < PT; ++t)
    {
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const float v = (acc[t][j] + __shfl_xor(acc[t][j], 8)) * unscale;
            const int col = lane & 15;
            if (col < 8)
                sm.pix[((wave * PT + t) * 16 + (lane >> 4) * 4 + j) * 8 + col] = v;
        }
    }
    __syncthreads();
    float2 *out = reinterpret_cast<float2 *>(a.subgrids_out) + static_cast<size_t>(s) * kCorr * P;
    const float4 *aterms = reinterpret_cast<const float4 *>(a.aterms);
    for (int q = tid; q < kPix; q += kThreads)
    {
        // Slot q = (wave w, tile t, row r) of the accumulator gather above.
        const int w = q / (PT * 16), t = (q >> 4) % PT, r = q & 15;
        const int u = kWaves * pblk + w;
        const int y = (u / gu.ns) * gu.rows + t * gu.rt + r / gu.wd;
        if (u >= gu.units || y >= S)
            continue;
        const int p = y * S + (u % gu.ns) * gu.wd + r % gu.wd;
        const float *v = sm.pix + q * 8;
        cfloat px[4] = {{v[0], v[1]}, {v[2], v[3]}, {v[4], v[5]}, {v[6], v[7]}};
        if (ok)
        {
            const size_t base = static_cast<size_t>(m.aterm_index) * a.p.nr_stations;
            cfloat a1[4], a2[4];
            load_jones(aterms, (base + m.station1) * P + p, a1);
            load_jones(aterms, (base + m.station2) * P + p, a2);
            apply_a1h_p_a2(px, a1, a2);
            const float taper = a.spheroidal[p];
#pragma unroll
            for (int pol = 0; pol < 4; ++pol)
                out[pol * P + p] = make_float2(px[pol].re * taper, px[pol].im * taper);
        }
        else
        {
#pragma unroll
            for (int pol = 0; pol < 4; ++pol)
                out[pol * P + p] = make_float2(0.F, 0.F);
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Degridder item: chunk vblk of 4 waves x VT tiles x 16 visibilities of subgrid s.  The subgrid's pixels stream
// through LDS PB at a time (taper and A-term applied, split into B fragments); the accumulators stay in
// registers across batches, so each visibility is written once.
// ---------------------------------------------------------------------------------------------------------------
// PREP: the pixels come prepared (a.prepared, a.prepared_scale, a.slot_coord: the persistent kernels, whose batch
// prepares them before every pass); otherwise the item applies taper and A-terms as it stages them (the one-shot
// kernel).  A compile-time choice, so the persistent kernel carries no registers for the staging it never runs.
template <int DT, bool SPLIT, int VT, int PB, bool PM, bool PREP>
__device__ __forceinline__ void degridder_item(const KernelArgs &a, int s, int vblk, DegridderShared<VT, PB> &sm)
{
    constexpr int kChunk = kWaves * VT * 16;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int S = a.p.S, P = S * S;
    const idgx_metadata m = a.metadata[s];
    int c0, nc;
    channel_window(a, s, c0, nc);
    const long long row0 = first_row(a.metadata, m);
    if (!subgrid_in_bounds(a.p, m, row0, c0, nc) || nc == 0)
        return; // uniform: nothing of this subgrid is written
    Selection sel;
    const int V = selection_of(a, s, m.nr_timesteps, nc, sel);
    const int v0 = vblk * kChunk;
    if (v0 >= V)
        return; // uniform
    const long long C = a.p.nr_channels;
    const CentreOffsets off = centre_offsets(a.p, m);
    const double delta = static_cast<double>(a.p.image_size) / S;
    const float2 *sub = reinterpret_cast<const float2 *>(a.subgrids) + static_cast<size_t>(s) * kCorr * P;
    uint32_t *s_bw = reinterpret_cast<uint32_t *>(sm.b);

    // The FP64 reduction of each of the chunk's visibilities, once: thread tid takes visibilities tid, tid + 256, ...
    // (the ones it writes in the epilogue, so their output rotations stay in registers) and shares qu, qv, dw
    // through sm.vis, which the epilogue alone writes.  Each lane then reads visibility lane%16 of each of its tiles;
    // for PM only component lane/16 of them (the B operand of the phase MFMA).
    constexpr int kPerThread = (kChunk + kThreads - 1) / kThreads;
    float rot[kPerThread];
#pragma unroll
    for (int i = 0; i < kPerThread; ++i)
    {
        const int q = tid + i * kThreads;
        rot[i] = 0.F;
        if (q >= kChunk)
            break;
        const int v = v0 + q;
        float ru = 0.F, rv = 0.F, rw = 0.F;
        int tt = 0, c = 0;
        if (v < V)
            sel.at(v, tt, c);
        if (v < V && entry_ok(tt, c, m.nr_timesteps, nc))
        {
            const Reduced rd = reduce_visibility(a.uvw + 3 * (row0 + tt), a.wavenumbers[c0 + c], off, delta, 1.F);
            ru = rd.qu;
            rv = rd.qv;
            rw = rd.dw;
            rot[i] = rd.rot;
        }
        sm.vis[3 * q] = ru;
        sm.vis[3 * q + 1] = rv;
        sm.vis[3 * q + 2] = rw;
    }
    __syncthreads();
    float coef[VT], qu[VT], qv[VT], qw[VT];
#pragma unroll
    for (int t = 0; t < VT; ++t)
    {
        const int q = (wave * VT + t) * 16 + (lane & 15);
        qu[t] = sm.vis[3 * q];
        qv[t] = sm.vis[3 * q + 1];
        qw[t] = sm.vis[3 * q + 2];
        const int kk = lane >> 4;
        coef[t] = kk == 0 ? qu[t] : kk == 1 ? qv[t] : kk == 2 ? qw[t] : 0.F;
    }
    f32x4 acc[VT];
#pragma unroll
    for (int t = 0; t < VT; ++t)
        acc[t] = (f32x4){0.F, 0.F, 0.F, 0.F};
    const bool wave_live = v0 + wave * VT * 16 < V; // wave-uniform
    int scale_exp = -126; // FP16: current power-of-two scale of the staged B operand
    if constexpr (DT == kFP16 && PREP)
        scale_exp = a.prepared_scale[s];

    // Pixel slots: rows of Sp = S rounded up to a multiple of 4, so each lane group's 4 slots share a row.
    const int Sp = degrid_row_stride(S), nslots = S * Sp;
    for (int pb = 0; pb < nslots; pb += PB)
    {
        const int np = min(PB, nslots - pb);
        __syncthreads();
        constexpr int kPer = (PB + kThreads - 1) / kThreads;
        if constexpr (PREP)
        {
            // Prepared pixels: copy the split words and the slot coordinates into LDS.
            const uint4 *src = a.prepared + (static_cast<size_t>(s) * nslots + pb) * 2;
#pragma unroll
            for (int i = 0; i < kPer; ++i)
            {
                const int q = tid + i * kThreads;
                if (q >= PB)
                    break;
                uint32_t hi[4] = {0u, 0u, 0u, 0u}, lo[4] = {0u, 0u, 0u, 0u};
                float cx = 0.F, cy = 0.F, cn = 0.F;
                if (q < np)
                {
                    const uint4 w0 = src[2 * q], w1 = src[2 * q + 1];
                    hi[0] = w0.x, lo[0] = w0.y, hi[1] = w0.z, lo[1] = w0.w;
                    hi[2] = w1.x, lo[2] = w1.y, hi[3] = w1.z, lo[3] = w1.w;
                    const float4 c = a.slot_coord[pb + q];
                    cx = c.x;
                    cy = c.y;
                    cn = c.z;
                }
                const int blk = q >> 4, qq = q & 15;
                write_b_words(s_bw + blk * 256, qq, hi, lo);
                if constexpr (PM)
                {
                    float *r = sm.coord + blk * 64 + qq;
                    r[0] = cx;
                    r[16] = cy;
                    r[32] = cn;
                    r[48] = 0.F;
                }
                else
                {
                    float *r = sm.coord + blk * 64 + (qq >> 2) * 16 + (qq & 3);
                    r[0] = cx;
                    r[4] = cy;
                    r[8] = cn;
                }
            }
        }
        else
        {
            cfloat px[kPer][4];
            float cx[kPer], cy[kPer], cn[kPer];
            float vmax = 0.F;
#pragma unroll
            for (int i = 0; i < kPer; ++i)
            {
                const int q = tid + i * kThreads;
#pragma unroll
                for (int pol = 0; pol < 4; ++pol)
                    px[i][pol] = {0.F, 0.F};
                cx[i] = cy[i] = cn[i] = 0.F;
                const int sy = (pb + q) / Sp, sx = pb + q - sy * Sp;
                if (q < np)
                {
                    cx[i] = pixel_offset(sx, S);
                    cy[i] = pixel_offset(sy, S);
                }
                if (q < np && sx < S)
                {
                    degrid_pixel(a, m, sub, sy * S + sx, px[i]);
                    cn[i] = pixel_n_f(pixel_lm_f(sx, S, a.p.image_size), pixel_lm_f(sy, S, a.p.image_size));
#pragma unroll
                    for (int pol = 0; pol < 4; ++pol)
                        vmax = fmaxf(vmax, fmaxf(fabsf(px[i][pol].re), fabsf(px[i][pol].im)));
                }
            }
            if constexpr (DT == kFP16)
            {
                const int e_new = max(scale_exp, scale_exponent(block_max(vmax, sm.maxbits)));
                if (e_new > scale_exp)
                {
                    const float f = ldexpf(1.F, scale_exp - e_new);
#pragma unroll
                    for (int t = 0; t < VT; ++t)
                        acc[t] *= f;
                    scale_exp = e_new;
                }
                const float inv = ldexpf(1.F, -scale_exp);
#pragma unroll
                for (int i = 0; i < kPer; ++i)
#pragma unroll
                    for (int pol = 0; pol < 4; ++pol)
                        px[i][pol] = {px[i][pol].re * inv, px[i][pol].im * inv};
            }
#pragma unroll
            for (int i = 0; i < kPer; ++i)
            {
                const int q = tid + i * kThreads;
                if (q >= PB)
                    break;
                const int blk = q >> 4, qq = q & 15;
                write_b_item<DT, SPLIT>(s_bw + blk * 256, qq, px[i]);
                if constexpr (PM)
                {
                    float *r = sm.coord + blk * 64 + qq;
                    r[0] = cx[i];
                    r[16] = cy[i];
                    r[32] = cn[i];
                    r[48] = 0.F;
                }
                else
                {
                    float *r = sm.coord + blk * 64 + (qq >> 2) * 16 + (qq & 3);
                    r[0] = cx[i];
                    r[4] = cy[i];
                    r[8] = cn[i];
                }
            }
        }
        __syncthreads();
        // A wave whose visibilities all lie past V (the subgrid's last chunk) only stages and waits.
        const int nblk = wave_live ? (np + 15) >> 4 : 0;
        for (int b = 0; b < nblk; ++b)
        {
            const u32x4 bf = sm.b[b * 64 + lane];
            f32x4 phs[VT];
            if constexpr (PM)
            {
                const float cd = sm.coord[b * 64 + lane];
#pragma unroll
                for (int t = 0; t < VT; ++t)
                    phs[t] = phase_mfma(cd, coef[t]); // phi(pixel 4*(lane/16)+j, visibility lane%16)
            }
            else
            {
                // The lane group's 4 pixels share y': its term is formed once per visibility, 2 FMAs per phase.
                const float *g = sm.coord + b * 64 + (lane >> 4) * 16;
                const f32x4 x4 = *reinterpret_cast<const f32x4 *>(g), n4 = *reinterpret_cast<const f32x4 *>(g + 8);
                const float y = g[4];
#pragma unroll
                for (int t = 0; t < VT; ++t)
                {
                    const float vy = qv[t] * y;
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        phs[t][j] = __builtin_fmaf(qu[t], x4[j], __builtin_fmaf(qw[t], n4[j], vy));
                }
            }
#pragma unroll
            for (int t = 0; t < VT; ++t)
            {
                uint32_t hi[4], lo[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                {
                    const float sn = __builtin_amdgcn_sinf(phs[t][j]);
                    const float cs = __builtin_amdgcn_cosf(phs[t][j]);
                    if constexpr (SPLIT)
                        split2<DT>(cs, sn, hi[j], lo[j]);
                    else
                        hi[j] = pack2<DT>(cs, sn);
                }
                acc[t] = mfma(hi[0], hi[1], hi[2], hi[3], bf, acc[t], std::integral_constant<int, DT>());
                if constexpr (SPLIT)
                    acc[t] = mfma(lo[0], lo[1], lo[2], lo[3], bf, acc[t], std::integral_constant<int, DT>());
            }
        }
    }

    // Combine column halves, gather 8 values per visibility in LDS, write 32 B per visibility.
    const float unscale = DT == kFP16 ? ldexpf(1.F, scale_exp) : 1.F;
#pragma unroll
    for (int t = 0; t < VT; ++t)
    {
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const float v = (acc[t][j] + __shfl_xor(acc[t][j], 8)) * unscale;
            const int col = lane & 15;
            if (col < 8)
                sm.vis[((wave * VT + t) * 16 + (lane >> 4) * 4 + j) * 8 + col] = v;
        }
    }
    __syncthreads();
    float4 *vout = reinterpret_cast<float4 *>(a.visibilities_out);
#pragma unroll
    for (int i = 0; i < kPerThread; ++i)
    {
        const int q = tid + i * kThreads, v = v0 + q;
        if (q >= kChunk || v >= V)
            break;
        int tt, c;
        sel.at(v, tt, c);
        if (!entry_ok(tt, c, m.nr_timesteps, nc))
            continue;
        const size_t o = ((row0 + tt) * C + c0 + c) * 2;
        const float *x = sm.vis + q * 8;
        cfloat val[4] = {{x[0], x[1]}, {x[2], x[3]}, {x[4], x[5]}, {x[6], x[7]}};
        rotate4(val, rot[i]);
        vout[o] = make_float4(val[0].re, val[0].im, val[1].re, val[1].im);
        vout[o + 1] = make_float4(val[2].re, val[2].im, val[3].re, val[3].im);
    }
    __syncthreads(); // LDS is reused by the next item of a persistent workgroup
}

// ---------------------------------------------------------------------------------------------------------------
// Kernels: one-shot grids (workgroup b = item (b % n, b / n), so a subgrid's workgroups spread over the XCDs)
// and persistent workgroups fed from the device work queue.
// ---------------------------------------------------------------------------------------------------------------
// Resident workgroups per CU the kernels ask for: 4 (gridder) and 5 (degridder), which leaves 128 / 96 VGPRs (a
// few bytes spill outside the inner loops; measured 2 % faster than 3 / 4), fewer where the shape's LDS or its
// accumulators do not allow it.
constexpr int kLdsBytes = 160 * 1024;
template <int PT, int VB> constexpr int gridder_min_blocks()
{
    constexpr int fit = kLdsBytes / static_cast<int>(sizeof(GridderShared<PT, VB>));
    return fit < 4 ? fit : 4;
}
template <int VT, int PB> constexpr int degridder_min_blocks()
{
    constexpr int fit = kLdsBytes / static_cast<int>(sizeof(DegridderShared<VT, PB>));
    constexpr int regs = VT >= 8 ? 3 : VT >= 6 ? 4 : 5;
    return fit < regs ? fit : regs;
}

template <int DT, bool SPLIT, int PT, int VB, bool PM>
__global__ __launch_bounds__(kThreads, (gridder_min_blocks<PT, VB>())) void gridder_mfma(KernelArgs a)
{
    __shared__ GridderShared<PT, VB> sm;
    const int pblk = blockIdx.x / a.p.nr_subgrids;
    gridder_item<DT, SPLIT, PT, VB, PM>(a, blockIdx.x - pblk * a.p.nr_subgrids, pblk, sm);
}

template <int DT, bool SPLIT, int VT, int PB, bool PM>
__global__ __launch_bounds__(kThreads, (degridder_min_blocks<VT, PB>())) void degridder_mfma(KernelArgs a)
{
    __shared__ DegridderShared<VT, PB> sm;
    const int vblk = blockIdx.x / a.p.nr_subgrids;
    degridder_item<DT, SPLIT, VT, PB, PM, false>(a, blockIdx.x - vblk * a.p.nr_subgrids, vblk, sm);
}

template <int DT, bool SPLIT, int PT, int VB, bool PM>
__global__ __launch_bounds__(kThreads, (gridder_min_blocks<PT, VB>())) void gridder_mfma_queue(QueueArgs q)
{
    __shared__ GridderShared<PT, VB> sm;
    int cls, s, blk;
    while (next_queue_item(q, cls, s, blk))
    {
        gridder_item<DT, SPLIT, PT, VB, PM>(q.cls[cls], s, blk, sm);
        __syncthreads();
    }
}

// One launch over items of several subgrid classes, each item with its class's tiles per wave (q.tiles[class]),
// LDS batch kDefaultVB, phases on the VALU.  The LDS holds the largest tile count's staging; every variant's
// registers fit the same occupancy (4 waves per SIMD), so a class loses nothing to sharing the launch.
template <int DT, bool SPLIT>
__global__ __launch_bounds__(kThreads, (gridder_min_blocks<8, kDefaultVB>())) void gridder_mfma_queue_tiles(QueueArgs q)
{
    __shared__ GridderShared<8, kDefaultVB> sm;
    int cls, s, blk;
    while (next_queue_item(q, cls, s, blk))
    {
        switch (q.tiles[cls]) // workgroup-uniform
        {
        case 2:
            gridder_item<DT, SPLIT, 2, kDefaultVB, false>(q.cls[cls], s, blk, sm);
            break;
        case 3:
            gridder_item<DT, SPLIT, 3, kDefaultVB, false>(q.cls[cls], s, blk, sm);
            break;
        case 4:
            gridder_item<DT, SPLIT, 4, kDefaultVB, false>(q.cls[cls], s, blk, sm);
            break;
        case 6:
            gridder_item<DT, SPLIT, 6, kDefaultVB, false>(q.cls[cls], s, blk, sm);
            break;
        default:
            gridder_item<DT, SPLIT, 8, kDefaultVB, false>(q.cls[cls], s, blk, sm);
            break;
        }
        __syncthreads();
    }
}

template <int DT, bool SPLIT, int VT, int PB, bool PM>
__global__ __launch_bounds__(kThreads, (degridder_min_blocks<VT, PB>())) void degridder_mfma_queue(QueueArgs q)
{
    __shared__ DegridderShared<VT, PB> sm;
    int cls, s, blk;
    while (next_queue_item(q, cls, s, blk))
        degridder_item<DT, SPLIT, VT, PB, PM, true>(q.cls[cls], s, blk, sm);
}

__global__ void reset_counters_kernel(int *counters, int n)
{
    if (static_cast<int>(threadIdx.x) < n)
        counters[threadIdx.x] = 0;
}

// One workgroup per subgrid: the taper x A-term-applied pixels of every slot, split into the packed words the
// degridder stages (FP16: scaled by 2^-scale so that the largest value is about 2^14).  The workgroup of subgrid 0
// also writes the slot table (x', y', n), so that the degridder's staging reads a slot's coordinates instead of
// dividing by the row stride.  Subgrids outside the arrays are skipped (the degridder skips them too).  Each XCD
// takes a contiguous run of subgrids (plan order: baseline by baseline), so the stations' A-terms it reads stay in
// its own L2 instead of every XCD streaming the whole table.
template <int DT, bool SPLIT, int CACHED>
__global__ __launch_bounds__(kThreads) void degrid_prepare_kernel(KernelArgs a)
{
    __shared__ unsigned maxbits;
    const int s = xcd_contiguous(blockIdx.x, gridDim.x), tid = threadIdx.x;
    const int S = a.p.S, P = S * S, Sp = degrid_row_stride(S), nslots = S * Sp;
    float4 *slot_coord = const_cast<float4 *>(a.slot_coord);
    if (s == 0)
    {
        for (int q = tid; q < nslots; q += kThreads)
        {
            const int sy = q / Sp, sx = q - sy * Sp;
            const float n = sx < S ? pixel_n_f(pixel_lm_f(sx, S, a.p.image_size), pixel_lm_f(sy, S, a.p.image_size))
                                   : 0.F;
            slot_coord[q] = make_float4(pixel_offset(sx, S), pixel_offset(sy, S), n, 0.F);
        }
    }
    const idgx_metadata m = a.metadata[s];
    int c0, nc;
    channel_window(a, s, c0, nc);
    const long long row0 = first_row(a.metadata, m);
    if (!subgrid_in_bounds(a.p, m, row0, c0, nc) || nc == 0)
        return; // uniform
    const float2 *sub = reinterpret_cast<const float2 *>(a.subgrids) + static_cast<size_t>(s) * kCorr * P;
    uint4 *dst = const_cast<uint4 *>(a.prepared) + static_cast<size_t>(s) * nslots * 2;
    // Slots per thread held in registers between the FP16 scale and the split, so that the subgrid is read once:
    // 4 covers S <= 32, 9 covers S <= 48, 16 covers S <= 64; larger subgrids read it twice.
    constexpr int kCached = CACHED;
    if (nslots <= kCached * kThreads)
    {
        // One pass over the subgrid: pixels stay in registers while the workgroup finds the FP16 scale.
        cfloat px[kCached][4];
        float vmax = 0.F;
#pragma unroll
        for (int k = 0; k < kCached; ++k)
        {
            const int q = tid + k * kThreads;
            const int sy = q / Sp, sx = q - sy * Sp;
#pragma unroll
            for (int pol = 0; pol < 4; ++pol)
                px[k][pol] = {0.F, 0.F};
            if (q < nslots && sx < S)
            {
                degrid_pixel(a, m, sub, sy * S + sx, px[k]);
#pragma unroll
                for (int pol = 0; pol < 4; ++pol)
                    vmax = fmaxf(vmax, fmaxf(fabsf(px[k][pol].re), fabsf(px[k][pol].im)));
            }
        }
        float inv1 = 1.F;
        if constexpr (DT == kFP16)
        {
            const int e = scale_exponent(block_max(vmax, maxbits));
            if (tid == 0)
                const_cast<int *>(a.prepared_scale)[s] = e;
            inv1 = ldexpf(1.F, -e);
        }
#pragma unroll
        for (int k = 0; k < kCached; ++k)
        {
            const int q = tid + k * kThreads;
            if (q >= nslots)
                break;
            uint32_t hi[4], lo[4];
#pragma unroll
            for (int pol = 0; pol < 4; ++pol)
                px[k][pol] = {px[k][pol].re * inv1, px[k][pol].im * inv1};
            split_item<DT, SPLIT>(px[k], hi, lo);
            dst[2 * q] = make_uint4(hi[0], lo[0], hi[1], lo[1]);
            dst[2 * q + 1] = make_uint4(hi[2], lo[2], hi[3], lo[3]);
        }
        return;
    }
    float inv = 1.F;
    if constexpr (DT == kFP16)
    {
        float vmax = 0.F;
        for (int q = tid; q < nslots; q += kThreads)
        {
            const int sy = q / Sp, sx = q - sy * Sp;
            if (sx < S)
            {
                cfloat px[4];
                degrid_pixel(a, m, sub, sy * S + sx, px);
#pragma unroll
                for (int pol = 0; pol < 4; ++pol)
                    vmax = fmaxf(vmax, fmaxf(fabsf(px[pol].re), fabsf(px[pol].im)));
            }
        }
        const int e = scale_exponent(block_max(vmax, maxbits));
        if (tid == 0)
            const_cast<int *>(a.prepared_scale)[s] = e;
        inv = ldexpf(1.F, -e);
    }
    for (int q = tid; q < nslots; q += kThreads)
    {
        const int sy = q / Sp, sx = q - sy * Sp;
        uint32_t hi[4] = {0u, 0u, 0u, 0u}, lo[4] = {0u, 0u, 0u, 0u};
        if (sx < S)
        {
            cfloat px[4];
            degrid_pixel(a, m, sub, sy * S + sx, px);
#pragma unroll
            for (int pol = 0; pol < 4; ++pol)
                px[pol] = {px[pol].re * inv, px[pol].im * inv};
            split_item<DT, SPLIT>(px, hi, lo);
        }
        dst[2 * q] = make_uint4(hi[0], lo[0], hi[1], lo[1]);
        dst[2 * q + 1] = make_uint4(hi[2], lo[2], hi[3], lo[3]);
    }
}

// ---- launchers ------------------------------------------------------------------------------------------------


// Gridder tiles per wave for subgrid size S.  A wave with more tiles spreads each visibility's staged work over
// more pixels (issue efficiency relative to 8 tiles, measured on sizes the tiles cover exactly: 6 -> 0.965,
// 4 -> 0.89, 3 -> 0.845, 2 -> 0.78), but the last row block and the last item of a subgrid can leave tiles and
// waves idle (coverage = S^2 / pixels of the items).  The choice maximises coverage x efficiency: S = 32 and 64
// take 8, S = 48 takes 6, S = 24 takes 3, S = 16 takes 4.
constexpr int kGridTileChoices[] = {8, 6, 4, 3, 2};
constexpr double kGridTileEfficiency[] = {1.0, 0.965, 0.89, 0.845, 0.78};

double grid_coverage(int S, int tiles)
{
    const GridUnits gu = grid_units(S, tiles);
    const long long items = (gu.units + kWaves - 1) / kWaves;
    return static_cast<double>(S) * S / static_cast<double>(items * kWaves * tiles * 16);
}

int grid_tiles_for(int S)
{
    int best = kGridTileChoices[0];
    double best_score = -1.0;
    for (int i = 0; i < static_cast<int>(sizeof(kGridTileChoices) / sizeof(int)); ++i)
    {
        const double score = grid_coverage(S, kGridTileChoices[i]) * kGridTileEfficiency[i];
        if (score > best_score + 1e-9) // ties keep the larger tile count
        {
            best_score = score;
            best = kGridTileChoices[i];
        }
    }
    return best;
}

template <int DT, bool SPLIT, int PT, bool PM> hipError_t grid_mfma_tiles(const KernelArgs &a, hipStream_t st)
{
    const GridUnits gu = grid_units(a.p.S, PT);
    const long long blocks = static_cast<long long>(a.p.nr_subgrids) * ((gu.units + kWaves - 1) / kWaves);
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    gridder_mfma<DT, SPLIT, PT, kDefaultVB, PM><<<static_cast<unsigned>(blocks), kThreads, 0, st>>>(a);
    return hipGetLastError();
}

template <int DT, bool SPLIT> hipError_t grid_mfma(const KernelArgs &a, bool pm, hipStream_t st)
{
    if (pm)
        return grid_mfma_tiles<DT, SPLIT, kDefaultPT, true>(a, st);
    switch (grid_tiles_for(a.p.S))
    {
    case 8:
        return grid_mfma_tiles<DT, SPLIT, 8, false>(a, st);
    case 6:
        return grid_mfma_tiles<DT, SPLIT, 6, false>(a, st);
    case 4:
        return grid_mfma_tiles<DT, SPLIT, 4, false>(a, st);
    case 3:
        return grid_mfma_tiles<DT, SPLIT, 3, false>(a, st);
    default:
        return grid_mfma_tiles<DT, SPLIT, 2, false>(a, st);
    }
}

template <int DT, bool SPLIT> hipError_t degrid_mfma(const KernelArgs &a, bool pm, hipStream_t st)
{
    constexpr int chunk = kWaves * kDefaultVT * 16;
    const long long vmax = static_cast<long long>(a.p.max_nr_timesteps) * a.p.nr_channels;
    const long long per = vmax > 0 ? (vmax + chunk - 1) / chunk : 1;
    const long long blocks = static_cast<long long>(a.p.nr_subgrids) * per;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    if (pm)
        degridder_mfma<DT, SPLIT, kDefaultVT, kDefaultPB, true><<<static_cast<unsigned>(blocks), kThreads, 0, st>>>(a);
    else
        degridder_mfma<DT, SPLIT, kDefaultVT, kDefaultPB, false><<<static_cast<unsigned>(blocks), kThreads, 0, st>>>(a);
    return hipGetLastError();
}

// Tile shapes the persistent kernels are built for (the autotuner picks among them): {tiles per wave, LDS batch,
// phases on the matrix core}.
constexpr MixedShape kShapes[] = {{4, 256, 0}, {2, 256, 0}, {8, 256, 0}, {4, 128, 0}, {4, 512, 0},
                                  {4, 256, 1}, {8, 256, 1}, {2, 256, 1}, {3, 256, 0}, {6, 256, 0}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

template <int DT, bool SPLIT, int I> void *queue_fn(bool gridder)
{
    constexpr MixedShape sh = kShapes[I];
    if (gridder)
        return reinterpret_cast<void *>(&gridder_mfma_queue<DT, SPLIT, sh.tiles, sh.batch, sh.phase_mfma != 0>);
    return reinterpret_cast<void *>(&degridder_mfma_queue<DT, SPLIT, sh.tiles, sh.batch, sh.phase_mfma != 0>);
}

template <int DT, bool SPLIT> void *queue_kernel(bool gridder, int shape)
{
    switch (shape)
    {
    case 0:
        return queue_fn<DT, SPLIT, 0>(gridder);
    case 1:
        return queue_fn<DT, SPLIT, 1>(gridder);
    case 2:
        return queue_fn<DT, SPLIT, 2>(gridder);
    case 3:
        return queue_fn<DT, SPLIT, 3>(gridder);
    case 4:
        return queue_fn<DT, SPLIT, 4>(gridder);
    case 5:
        return queue_fn<DT, SPLIT, 5>(gridder);
    case 6:
        return queue_fn<DT, SPLIT, 6>(gridder);
    case 7:
        return queue_fn<DT, SPLIT, 7>(gridder);
    case 8:
        return queue_fn<DT, SPLIT, 8>(gridder);
    case 9:
        return queue_fn<DT, SPLIT, 9>(gridder);
    default:
        return nullptr;
    }
}

void *queue_kernel_for(int precision, bool gridder, int shape)
{
    if (shape == kShapeTilesPerClass)
    {
        if (!gridder)
            return nullptr;
        switch (precision)
        {
        case IDGX_MIXED_BF16X3:
            return reinterpret_cast<void *>(&gridder_mfma_queue_tiles<kBF16, true>);
        case IDGX_MIXED_FP16X3:
            return reinterpret_cast<void *>(&gridder_mfma_queue_tiles<kFP16, true>);
        case IDGX_MIXED_BF16:
            return reinterpret_cast<void *>(&gridder_mfma_queue_tiles<kBF16, false>);
        case IDGX_MIXED_FP16:
            return reinterpret_cast<void *>(&gridder_mfma_queue_tiles<kFP16, false>);
        default:
            return nullptr;
        }
    }
    switch (precision)
    {
    case IDGX_MIXED_BF16X3:
        return queue_kernel<kBF16, true>(gridder, shape);
    case IDGX_MIXED_FP16X3:
        return queue_kernel<kFP16, true>(gridder, shape);
    case IDGX_MIXED_BF16:
        return queue_kernel<kBF16, false>(gridder, shape);
    case IDGX_MIXED_FP16:
        return queue_kernel<kFP16, false>(gridder, shape);
    default:
        return nullptr;
    }
}

hipError_t launch_queue(const QueueArgs &q, int precision, bool gridder, int shape, int workgroups, hipStream_t st)
{
    void *fn = queue_kernel_for(precision, gridder, shape);
    if (!fn || workgroups <= 0)
        return hipErrorInvalidValue;
    // The persistent degridder reads prepared pixels only (launch_degrid_prepare fills them first).
    for (int k = 0; k < kMaxClasses && !gridder; ++k)
        if (q.cls[k].p.nr_subgrids > 0 && (!q.cls[k].prepared || !q.cls[k].prepared_scale || !q.cls[k].slot_coord))
            return hipErrorInvalidValue;
    QueueArgs arg = q;
    void *args[] = {&arg};
    return hipLaunchKernel(fn, dim3(workgroups), dim3(kThreads), args, 0, st);
}
} // namespace

hipError_t launch_grid_mixed(const KernelArgs &a, int precision, bool pm, hipStream_t st)
{
    switch (precision)
    {
    case IDGX_MIXED_BF16X3:
        return grid_mfma<kBF16, true>(a, pm, st);
    case IDGX_MIXED_FP16X3:
        return grid_mfma<kFP16, true>(a, pm, st);
    case IDGX_MIXED_BF16:
        return grid_mfma<kBF16, false>(a, pm, st);
    case IDGX_MIXED_FP16:
        return grid_mfma<kFP16, false>(a, pm, st);
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_degrid_mixed(const KernelArgs &a, int precision, bool pm, hipStream_t st)
{
    switch (precision)
    {
    case IDGX_MIXED_BF16X3:
        return degrid_mfma<kBF16, true>(a, pm, st);
    case IDGX_MIXED_FP16X3:
        return degrid_mfma<kFP16, true>(a, pm, st);
    case IDGX_MIXED_BF16:
        return degrid_mfma<kBF16, false>(a, pm, st);
    case IDGX_MIXED_FP16:
        return degrid_mfma<kFP16, false>(a, pm, st);
    default:
        return hipErrorInvalidValue;
    }
}

int mixed_shape_count()
{
    return kNumShapes;
}

// The one-shot kernels' tile shapes for S = 32, as persistent-kernel shapes.
constexpr int kDefaultGridShape = 2, kDefaultDegridShape = 0;
static_assert(kShapes[kDefaultGridShape].tiles == kDefaultPT && kShapes[kDefaultGridShape].batch == kDefaultVB &&
                  kShapes[kDefaultGridShape].phase_mfma == 0,
              "default gridder shape");
static_assert(kShapes[kDefaultDegridShape].tiles == kDefaultVT && kShapes[kDefaultDegridShape].batch == kDefaultPB &&
                  kShapes[kDefaultDegridShape].phase_mfma == 0,
              "default degridder shape");

int mixed_default_shape(bool gridder, int S)
{
    if (!gridder)
        return kDefaultDegridShape;
    const int tiles = grid_tiles_for(S);
    for (int i = 0; i < kNumShapes; ++i)
        if (kShapes[i].tiles == tiles && kShapes[i].batch == kDefaultVB && !kShapes[i].phase_mfma)
            return i;
    return kDefaultGridShape;
}

MixedShape mixed_shape(int i)
{
    return kShapes[i];
}

int grid_items_per_subgrid(int shape, int S)
{
    return (grid_units(S, kShapes[shape].tiles).units + kWaves - 1) / kWaves;
}

long long grid_item_pixels(int shape, int S, int item)
{
    const GridUnits gu = grid_units(S, kShapes[shape].tiles);
    long long n = 0;
    for (int u = kWaves * item; u < min(kWaves * (item + 1), gu.units); ++u)
        n += static_cast<long long>(gu.wd) * min(gu.rows, S - (u / gu.ns) * gu.rows);
    return n;
}

int degrid_visibilities_per_item(int shape)
{
    return kWaves * kShapes[shape].tiles * 16;
}

hipError_t launch_grid_mixed_queue(const QueueArgs &q, int precision, int shape, int workgroups, hipStream_t st)
{
    return launch_queue(q, precision, true, shape, workgroups, st);
}

hipError_t launch_degrid_mixed_queue(const QueueArgs &q, int precision, int shape, int workgroups, hipStream_t st)
{
    return launch_queue(q, precision, false, shape, workgroups, st);
}

size_t degrid_prepared_bytes(int S, long long nr_subgrids)
{
    return static_cast<size_t>(nr_subgrids) * S * degrid_row_stride(S) * 2 * sizeof(uint4);
}

hipError_t launch_degrid_prepare(const KernelArgs &a, int precision, hipStream_t st)
{
    if (a.p.nr_subgrids <= 0)
        return hipSuccess;
    if (!a.prepared || !a.prepared_scale || !a.slot_coord)
        return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(a.p.nr_subgrids));
    // FP16 needs the subgrid's maximum before the split: hold up to 4, 9 or 16 slots per thread, so that subgrids
    // up to S = 32, 48 or 64 and their A-terms are read once (BF16 has no scale and streams the slots in one pass
    // either way).
    const int nslots = a.p.S * degrid_row_stride(a.p.S);
    const int cached = nslots <= 4 * kThreads ? 4 : nslots <= 9 * kThreads ? 9 : 16;
    switch (precision)
    {
    case IDGX_MIXED_BF16X3:
        degrid_prepare_kernel<kBF16, true, 4><<<grid, kThreads, 0, st>>>(a);
        break;
    case IDGX_MIXED_FP16X3:
        if (cached == 16)
            degrid_prepare_kernel<kFP16, true, 16><<<grid, kThreads, 0, st>>>(a);
        else if (cached == 9)
            degrid_prepare_kernel<kFP16, true, 9><<<grid, kThreads, 0, st>>>(a);
        else
            degrid_prepare_kernel<kFP16, true, 4><<<grid, kThreads, 0, st>>>(a);
        break;
    case IDGX_MIXED_BF16:
        degrid_prepare_kernel<kBF16, false, 4><<<grid, kThreads, 0, st>>>(a);
        break;
    case IDGX_MIXED_FP16:
        if (cached == 16)
            degrid_prepare_kernel<kFP16, false, 16><<<grid, kThreads, 0, st>>>(a);
        else if (cached == 9)
            degrid_prepare_kernel<kFP16, false, 9><<<grid, kThreads, 0, st>>>(a);
        else
            degrid_prepare_kernel<kFP16, false, 4><<<grid, kThreads, 0, st>>>(a);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_reset_counters(int *counters, int n, hipStream_t st)
{
    if (n < 1 || n > 64)
        return hipErrorInvalidValue;
    reset_counters_kernel<<<1, 64, 0, st>>>(counters, n);
    return hipGetLastError();
}

int queue_workgroups(int precision, bool gridder, int shape)
{
    void *fn = queue_kernel_for(precision, gridder, shape);
    if (!fn)
        return 0;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    return cus * per_cu;
}
} // namespace idgx
