// Oracle for the IDG gridding/degridding hot path.  TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Build: oracle/Makefile, `-O3 -ffp-contract=off -fopenmp`.  Contraction must stay off: the reference is built
// without FMA contraction (as shipped it is -O0, gridding CMake `-g`, CMakeLists.txt:37) and one fused
// multiply-add in the phase moves results by ~5e-5 relative.
//
// FP32 restatement: every float expression below is written so that it rounds at the same points, in the same
// order, as the reference CPU code it cites.  The loop nest and the OpenMP structure (serial subgrid loop, a
// parallel region per subgrid) are kept so that the same code serves as the timed CPU baseline.

#include "oracle.h"

#include <cmath>
#include <complex>
#include <cstddef>
#include <cstring>
#include <vector>

#include <omp.h>

namespace
{
using cf = std::complex<float>;
using cd = std::complex<double>;

// The reference's pi() is a float constant built from a double literal (gridding/inc/math.hpp:13-16).
constexpr float kPiF = static_cast<float>(3.141592653589793);
constexpr int kCorr = 4;

// compute_l / compute_m (gridding/inc/math.hpp:20-29): ((x + 0.5f) - S/2.f) * image_size / S, float steps.
inline float pixel_lm(int x, int S, float image_size)
{
    float a = static_cast<float>(x) + 0.5F;
    a = a - static_cast<float>(S) / 2.F;
    a = a * image_size;
    return a / static_cast<float>(S);
}

// compute_n (gridding/inc/math.hpp:31-37): n = r2 / (1 + sqrt(1 - r2)), clamped to 1 when r2 > 1.
inline float pixel_n(float l, float m)
{
    const float r2 = (l * l) + (m * m);
    if (r2 > 1.F)
        return 1.F;
    return r2 / (1.F + std::sqrt(1.F - r2));
}

struct Offsets
{
    float u, v, w;
};

// Gridder's subgrid offsets in float (gridding/src/gridding_benchmark.cpp:68,74-81).
inline Offsets gridder_offsets(const oracle_params &p, const oracle_metadata &m)
{
    const float w_lambda = static_cast<float>(p.w_step_in_lambda * (m.z + 0.5));
    const float half_s = static_cast<float>(p.subgrid_size) * 0.5F;
    const float half_g = static_cast<float>(p.grid_size) * 0.5F;
    const float scale = 2.F * kPiF / p.image_size;
    Offsets o;
    o.u = ((static_cast<float>(m.x) + half_s) - half_g) * scale;
    o.v = ((static_cast<float>(m.y) + half_s) - half_g) * scale;
    o.w = (2.F * kPiF) * w_lambda;
    return o;
}

// Degridder's subgrid offsets: float base times a double scale, rounded back to float
// (gridding/src/degridding_benchmark.cpp:72,123-129).
inline Offsets degridder_offsets(const oracle_params &p, const oracle_metadata &m)
{
    const float w_lambda = static_cast<float>(p.w_step_in_lambda * (m.z + 0.5));
    const double scale = 2 * M_PI / static_cast<double>(p.image_size);
    const float bu = (static_cast<float>(m.x) + static_cast<float>(p.subgrid_size) / 2) -
                     static_cast<float>(p.grid_size) / 2;
    const float bv = (static_cast<float>(m.y) + static_cast<float>(p.subgrid_size) / 2) -
                     static_cast<float>(p.grid_size) / 2;
    Offsets o;
    o.u = static_cast<float>(static_cast<double>(bu) * scale);
    o.v = static_cast<float>(static_cast<double>(bv) * scale);
    o.w = static_cast<float>(2 * M_PI * static_cast<double>(w_lambda));
    return o;
}

// 2x2 complex matrices are stored row-major {xx, xy, yx, yy}.  mat2 follows the product order of the
// reference's matmul (gridding/inc/math.hpp:39-50): first the a[0|2]*b[0|1] terms, then += a[1|3]*b[2|3].
template <typename C> inline void mat2(const C *a, const C *b, C *c)
{
    c[0] = a[0] * b[0];
    c[1] = a[0] * b[1];
    c[2] = a[2] * b[0];
    c[3] = a[2] * b[1];
    c[0] += a[1] * b[2];
    c[1] += a[1] * b[3];
    c[2] += a[3] * b[2];
    c[3] += a[3] * b[3];
}

template <typename C> inline void herm2(const C *a, C *h)
{
    h[0] = std::conj(a[0]);
    h[1] = std::conj(a[2]);
    h[2] = std::conj(a[1]);
    h[3] = std::conj(a[3]);
}

// P <- A1^H * P * A2 (apply_aterm_gridder, gridding/inc/math.hpp:87-100).
template <typename C> inline void aterm_a1h_p_a2(C *pix, const C *a1, const C *a2)
{
    C a1h[4], t[4];
    herm2(a1, a1h);
    mat2(a1h, pix, t);
    mat2(t, a2, pix);
}

// P <- A1 * P * A2^H (the adjoint; the reference's CUDA degridder applies this,
// gridding/inc/common_gridding.cuh:82-94).
template <typename C> inline void aterm_a1_p_a2h(C *pix, const C *a1, const C *a2)
{
    C a2h[4], t[4];
    mat2(a1, pix, t);
    herm2(a2, a2h);
    mat2(t, a2h, pix);
}

inline size_t aterm_index(const oracle_params &p, const oracle_metadata &m, unsigned station, int y, int x)
{
    const size_t S = p.subgrid_size;
    return ((static_cast<size_t>(m.aterm_index) * p.nr_stations + station) * S * S + y * S + x) * kCorr;
}

inline void channel_window(const oracle_params &p, const oracle_channels *ch, int s, int &c0, int &c1)
{
    if (ch)
    {
        c0 = ch[s].channel_begin;
        c1 = ch[s].channel_begin + ch[s].nr_channels;
    }
    else
    {
        c0 = 0;
        c1 = p.nr_channels;
    }
}

// Row of the first timestep: (baseline_offset - first subgrid's baseline_offset) + time_offset
// (gridding_benchmark.cpp:36-37,61).
inline long first_row(const oracle_metadata *md, int s)
{
    return static_cast<long>(md[s].baseline_offset - md[0].baseline_offset) + md[s].time_offset;
}

bool valid(const oracle_params *p)
{
    return p && p->nr_subgrids >= 0 && p->subgrid_size > 0 && p->grid_size > 0 && p->nr_channels > 0 &&
           p->nr_stations > 0;
}
} // namespace

extern "C" int oracle_grid_fp32(const oracle_params *pp, const oracle_metadata *md, const oracle_channels *ch,
                                const float *uvw, const float *wn, const float *vis_f, const float *sph,
                                const float *aterm_f, float *sub_f, int nthreads)
{
    if (!valid(pp))
        return 1;
    const oracle_params p = *pp;
    const int S = p.subgrid_size;
    const size_t C = p.nr_channels;
    const cf *vis = reinterpret_cast<const cf *>(vis_f);
    const cf *aterms = reinterpret_cast<const cf *>(aterm_f);
    cf *subgrids = reinterpret_cast<cf *>(sub_f);
    if (nthreads > 0)
        omp_set_num_threads(nthreads);

    // Serial subgrid loop, parallel over pixels (gridding_benchmark.cpp:56,87-91).
    for (int s = 0; s < p.nr_subgrids; ++s)
    {
        const oracle_metadata m = md[s];
        const long row0 = first_row(md, s);
        const Offsets off = gridder_offsets(p, m);
        int c0, c1;
        channel_window(p, ch, s, c0, c1);

#pragma omp parallel for collapse(2)
        for (int y = 0; y < S; ++y)
            for (int x = 0; x < S; ++x)
            {
                cf acc[kCorr] = {0.F, 0.F, 0.F, 0.F};
                const float l = pixel_lm(x, S, p.image_size);
                const float mm = pixel_lm(y, S, p.image_size);
                const float n = pixel_n(l, mm);
                for (int t = 0; t < m.nr_timesteps; ++t)
                {
                    const float *uv = uvw + 3 * (row0 + t);
                    // phase_index and phase_offset (gridding_benchmark.cpp:112,115), left-to-right sums.
                    const float index = (uv[0] * l + uv[1] * mm) + uv[2] * n;
                    const float offset = (off.u * l + off.v * mm) + off.w * n;
                    for (int c = c0; c < c1; ++c)
                    {
                        const float phase = offset - index * wn[c]; // :121
                        const cf phasor(std::cos(phase), std::sin(phase));
                        const cf *v = vis + ((row0 + t) * C + c) * kCorr;
                        for (int pol = 0; pol < kCorr; ++pol)
                            acc[pol] += v[pol] * phasor; // :132
                    }
                }
                aterm_a1h_p_a2(acc, aterms + aterm_index(p, m, m.station1, y, x),
                               aterms + aterm_index(p, m, m.station2, y, x));
                const float taper = sph[y * S + x];
                for (int pol = 0; pol < kCorr; ++pol)
                    subgrids[((static_cast<size_t>(s) * kCorr + pol) * S + y) * S + x] = acc[pol] * taper; // :162
            }
    }
    return 0;
}

extern "C" int oracle_degrid_fp32(const oracle_params *pp, const oracle_metadata *md, const oracle_channels *ch,
                                  const float *uvw, const float *wn, const float *sub_f, const float *sph,
                                  const float *aterm_f, float *vis_f, int convention, int nthreads)
{
    if (!valid(pp) || (convention != ORACLE_ATERM_ADJOINT && convention != ORACLE_ATERM_REFERENCE))
        return 1;
    const oracle_params p = *pp;
    const int S = p.subgrid_size;
    const size_t C = p.nr_channels;
    const cf *subgrids = reinterpret_cast<const cf *>(sub_f);
    const cf *aterms = reinterpret_cast<const cf *>(aterm_f);
    cf *vis = reinterpret_cast<cf *>(vis_f);
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
    std::vector<cf> pixels(static_cast<size_t>(S) * S * kCorr);

    for (int s = 0; s < p.nr_subgrids; ++s)
    {
        const oracle_metadata m = md[s];
        const long row0 = first_row(md, s);
        int c0, c1;
        channel_window(p, ch, s, c0, c1);

        // Taper, then A-term, per pixel (degridding_benchmark.cpp:78-120).
#pragma omp parallel for collapse(2)
        for (int y = 0; y < S; ++y)
            for (int x = 0; x < S; ++x)
            {
                const float taper = sph[y * S + x];
                cf px[kCorr];
                for (int pol = 0; pol < kCorr; ++pol)
                    px[pol] = taper * subgrids[((static_cast<size_t>(s) * kCorr + pol) * S + y) * S + x]; // :109
                const cf *a1 = aterms + aterm_index(p, m, m.station1, y, x);
                const cf *a2 = aterms + aterm_index(p, m, m.station2, y, x);
                if (convention == ORACLE_ATERM_REFERENCE)
                    aterm_a1h_p_a2(px, a1, a2); // :113
                else
                    aterm_a1_p_a2h(px, a1, a2);
                for (int pol = 0; pol < kCorr; ++pol)
                    pixels[(static_cast<size_t>(y) * S + x) * kCorr + pol] = px[pol];
            }

        const Offsets off = degridder_offsets(p, m);
        const int nch = c1 - c0;

        // Parallel over (time, channel), sum over every pixel (degridding_benchmark.cpp:136-181).
#pragma omp parallel for collapse(2)
        for (int t = 0; t < m.nr_timesteps; ++t)
            for (int cc = 0; cc < nch; ++cc)
            {
                const int c = c0 + cc;
                const float *uv = uvw + 3 * (row0 + t);
                cf sum[kCorr] = {0.F, 0.F, 0.F, 0.F};
                for (int y = 0; y < S; ++y)
                    for (int x = 0; x < S; ++x)
                    {
                        const float l = pixel_lm(x, S, p.image_size);
                        const float mm = pixel_lm(y, S, p.image_size);
                        const float n = pixel_n(l, mm);
                        const float index = (uv[0] * l + uv[1] * mm) + uv[2] * n;
                        const float offset = (off.u * l + off.v * mm) + off.w * n;
                        const float phase = index * wn[c] - offset; // :166
                        const cf phasor(std::cos(phase), std::sin(phase));
                        const cf *px = &pixels[(static_cast<size_t>(y) * S + x) * kCorr];
                        for (int pol = 0; pol < kCorr; ++pol)
                            sum[pol] += px[pol] * phasor; // :171
                    }
                cf *out = vis + ((row0 + t) * C + c) * kCorr;
                for (int pol = 0; pol < kCorr; ++pol)
                    out[pol] = sum[pol];
            }
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------------------------
// FP64 oracle: the mathematics of the same operators, evaluated in double from the float inputs.
// ---------------------------------------------------------------------------------------------------------------
namespace
{
inline double lm64(int x, int S, double image_size)
{
    return (x + 0.5 - S / 2.0) * image_size / S;
}

inline double n64(double l, double m)
{
    const double r2 = l * l + m * m;
    return r2 > 1.0 ? 1.0 : r2 / (1.0 + std::sqrt(1.0 - r2));
}

struct Off64
{
    double u, v, w;
};

inline Off64 offsets64(const oracle_params &p, const oracle_metadata &m)
{
    const double scale = 2 * M_PI / static_cast<double>(p.image_size);
    Off64 o;
    o.u = (m.x + p.subgrid_size / 2.0 - p.grid_size / 2.0) * scale;
    o.v = (m.y + p.subgrid_size / 2.0 - p.grid_size / 2.0) * scale;
    o.w = 2 * M_PI * static_cast<double>(p.w_step_in_lambda) * (m.z + 0.5);
    return o;
}

inline void load_aterm64(const float *aterm_f, size_t idx, cd *a)
{
    for (int k = 0; k < kCorr; ++k)
        a[k] = cd(aterm_f[2 * (idx + k)], aterm_f[2 * (idx + k) + 1]);
}
} // namespace

extern "C" int oracle_grid_fp64(const oracle_params *pp, const oracle_metadata *md, const oracle_channels *ch,
                                const float *uvw, const float *wn, const float *vis_f, const float *sph,
                                const float *aterm_f, double *sub_d, int nthreads)
{
    if (!valid(pp))
        return 1;
    const oracle_params p = *pp;
    const int S = p.subgrid_size;
    const size_t C = p.nr_channels;
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
    for (int s = 0; s < p.nr_subgrids; ++s)
    {
        const oracle_metadata m = md[s];
        const long row0 = first_row(md, s);
        const Off64 off = offsets64(p, m);
        int c0, c1;
        channel_window(p, ch, s, c0, c1);
#pragma omp parallel for collapse(2) schedule(static)
        for (int y = 0; y < S; ++y)
            for (int x = 0; x < S; ++x)
            {
                const double l = lm64(x, S, p.image_size), mm = lm64(y, S, p.image_size), n = n64(l, mm);
                const double poff = off.u * l + off.v * mm + off.w * n;
                cd acc[kCorr] = {};
                for (int t = 0; t < m.nr_timesteps; ++t)
                {
                    const float *uv = uvw + 3 * (row0 + t);
                    const double index = double(uv[0]) * l + double(uv[1]) * mm + double(uv[2]) * n;
                    for (int c = c0; c < c1; ++c)
                    {
                        const double phase = poff - index * double(wn[c]);
                        const cd phasor(std::cos(phase), std::sin(phase));
                        const float *v = vis_f + 2 * ((row0 + t) * C + c) * kCorr;
                        for (int pol = 0; pol < kCorr; ++pol)
                            acc[pol] += cd(v[2 * pol], v[2 * pol + 1]) * phasor;
                    }
                }
                cd a1[4], a2[4];
                load_aterm64(aterm_f, aterm_index(p, m, m.station1, y, x), a1);
                load_aterm64(aterm_f, aterm_index(p, m, m.station2, y, x), a2);
                aterm_a1h_p_a2(acc, a1, a2);
                const double taper = sph[y * S + x];
                for (int pol = 0; pol < kCorr; ++pol)
                {
                    const size_t o = 2 * (((static_cast<size_t>(s) * kCorr + pol) * S + y) * S + x);
                    sub_d[o] = acc[pol].real() * taper;
                    sub_d[o + 1] = acc[pol].imag() * taper;
                }
            }
    }
    return 0;
}

extern "C" int oracle_degrid_fp64(const oracle_params *pp, const oracle_metadata *md, const oracle_channels *ch,
                                  const float *uvw, const float *wn, const float *sub_f, const float *sph,
                                  const float *aterm_f, double *vis_d, int nthreads)
{
    if (!valid(pp))
        return 1;
    const oracle_params p = *pp;
    const int S = p.subgrid_size;
    const size_t C = p.nr_channels;
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
    const size_t P = static_cast<size_t>(S) * S;
    std::vector<cd> pixels(P * kCorr);
    std::vector<double> lmn(P * 3);
    for (int y = 0; y < S; ++y)
        for (int x = 0; x < S; ++x)
        {
            const double l = lm64(x, S, p.image_size), mm = lm64(y, S, p.image_size);
            lmn[3 * (y * S + x)] = l;
            lmn[3 * (y * S + x) + 1] = mm;
            lmn[3 * (y * S + x) + 2] = n64(l, mm);
        }
    for (int s = 0; s < p.nr_subgrids; ++s)
    {
        const oracle_metadata m = md[s];
        const long row0 = first_row(md, s);
        const Off64 off = offsets64(p, m);
        int c0, c1;
        channel_window(p, ch, s, c0, c1);
        for (int y = 0; y < S; ++y)
            for (int x = 0; x < S; ++x)
            {
                const double taper = sph[y * S + x];
                cd px[kCorr];
                for (int pol = 0; pol < kCorr; ++pol)
                {
                    const size_t o = 2 * (((static_cast<size_t>(s) * kCorr + pol) * S + y) * S + x);
                    px[pol] = cd(sub_f[o], sub_f[o + 1]) * taper;
                }
                cd a1[4], a2[4];
                load_aterm64(aterm_f, aterm_index(p, m, m.station1, y, x), a1);
                load_aterm64(aterm_f, aterm_index(p, m, m.station2, y, x), a2);
                aterm_a1_p_a2h(px, a1, a2);
                for (int pol = 0; pol < kCorr; ++pol)
                    pixels[(static_cast<size_t>(y) * S + x) * kCorr + pol] = px[pol];
            }
        const int nch = c1 - c0;
#pragma omp parallel for collapse(2) schedule(static)
        for (int t = 0; t < m.nr_timesteps; ++t)
            for (int cc = 0; cc < nch; ++cc)
            {
                const int c = c0 + cc;
                const float *uv = uvw + 3 * (row0 + t);
                cd sum[kCorr] = {};
                for (size_t q = 0; q < P; ++q)
                {
                    const double l = lmn[3 * q], mm = lmn[3 * q + 1], n = lmn[3 * q + 2];
                    const double poff = off.u * l + off.v * mm + off.w * n;
                    const double index = double(uv[0]) * l + double(uv[1]) * mm + double(uv[2]) * n;
                    const double phase = index * double(wn[c]) - poff;
                    const cd phasor(std::cos(phase), std::sin(phase));
                    for (int pol = 0; pol < kCorr; ++pol)
                        sum[pol] += pixels[q * kCorr + pol] * phasor;
                }
                double *out = vis_d + 2 * ((row0 + t) * C + c) * kCorr;
                for (int pol = 0; pol < kCorr; ++pol)
                {
                    out[2 * pol] = sum[pol].real();
                    out[2 * pol + 1] = sum[pol].imag();
                }
            }
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------------------------
// Stages after the gridder: brute-force 2-D DFT and plain adder/splitter placement, in double.
// ---------------------------------------------------------------------------------------------------------------
extern "C" int oracle_dft2d_fp64(double *data, int batch, int S, int sign, int nthreads)
{
    if (batch < 0 || S <= 0 || (sign != 1 && sign != -1))
        return 1;
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
    std::vector<cd> tw(S);
    for (int k = 0; k < S; ++k)
        tw[k] = std::polar(1.0, sign * 2 * M_PI * k / S);
#pragma omp parallel for schedule(static)
    for (int b = 0; b < batch; ++b)
    {
        cd *x = reinterpret_cast<cd *>(data) + static_cast<size_t>(b) * S * S;
        std::vector<cd> tmp(static_cast<size_t>(S) * S);
        // Rows then columns, each a direct O(S^2) DFT with exact integer twiddle indices.
        for (int r = 0; r < S; ++r)
            for (int k = 0; k < S; ++k)
            {
                cd acc = 0;
                for (int n = 0; n < S; ++n)
                    acc += x[r * S + n] * tw[(static_cast<long>(k) * n) % S];
                tmp[r * S + k] = acc;
            }
        for (int c = 0; c < S; ++c)
            for (int k = 0; k < S; ++k)
            {
                cd acc = 0;
                for (int n = 0; n < S; ++n)
                    acc += tmp[n * S + c] * tw[(static_cast<long>(k) * n) % S];
                x[k * S + c] = acc;
            }
    }
    return 0;
}

extern "C" int oracle_adder_fp64(const oracle_metadata *md, int n, int S, int G, const float *sub_f, double *grid)
{
    if (n < 0 || S <= 0 || G <= 0)
        return 1;
    for (int s = 0; s < n; ++s)
        for (int pol = 0; pol < kCorr; ++pol)
            for (int y = 0; y < S; ++y)
                for (int x = 0; x < S; ++x)
                {
                    const long gy = static_cast<long>(md[s].y) + y, gx = static_cast<long>(md[s].x) + x;
                    if (gy < 0 || gy >= G || gx < 0 || gx >= G)
                        continue;
                    const size_t si = 2 * (((static_cast<size_t>(s) * kCorr + pol) * S + y) * S + x);
                    const size_t gi = 2 * ((static_cast<size_t>(pol) * G + gy) * G + gx);
                    grid[gi] += sub_f[si];
                    grid[gi + 1] += sub_f[si + 1];
                }
    return 0;
}

extern "C" int oracle_splitter_fp64(const oracle_metadata *md, int n, int S, int G, const double *grid,
                                    double *sub_d)
{
    if (n < 0 || S <= 0 || G <= 0)
        return 1;
    for (int s = 0; s < n; ++s)
        for (int pol = 0; pol < kCorr; ++pol)
            for (int y = 0; y < S; ++y)
                for (int x = 0; x < S; ++x)
                {
                    const long gy = static_cast<long>(md[s].y) + y, gx = static_cast<long>(md[s].x) + x;
                    const size_t si = 2 * (((static_cast<size_t>(s) * kCorr + pol) * S + y) * S + x);
                    if (gy < 0 || gy >= G || gx < 0 || gx >= G)
                    {
                        sub_d[si] = sub_d[si + 1] = 0;
                        continue;
                    }
                    const size_t gi = 2 * ((static_cast<size_t>(pol) * G + gy) * G + gx);
                    sub_d[si] = grid[gi];
                    sub_d[si + 1] = grid[gi + 1];
                }
    return 0;
}
