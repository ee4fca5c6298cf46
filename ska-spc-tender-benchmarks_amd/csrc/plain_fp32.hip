// Plain FP32 gridder and degridder for gfx950.
//
// Phases are rounded exactly as in the reference CPU path (gridding/src/gridding_benchmark.cpp:99-121,
// gridding/src/degridding_benchmark.cpp:157-166): the whole file is compiled with contraction off, so
// `a * b + c` rounds twice; the complex multiply-accumulates are written as explicit fused multiply-adds.
//
// Gridder: one workgroup of 256 threads owns 1024 or 768 pixels of one subgrid (4 or 3 per thread, 4 correlations
//   each, held in VGPRs).  Visibilities and uvw of the subgrid stream through LDS in batches of 256 visibilities;
//   every lane reads the same LDS address per visibility (broadcast, conflict-free).  Taper and A-term are
//   applied in the epilogue, which overwrites the subgrid.
// Degridder: one workgroup owns 256 (timestep, channel-chunk) items of one subgrid, each thread one timestep x
//   up to 2 or 4 channels.  Tapered, A-term-corrected pixels plus (l, m, n, phase offset) are staged through LDS in
//   batches of 512 pixels; each thread sums its visibilities over all pixels in registers, so no cross-lane
//   reduction is needed.
#include <algorithm>

#include "idgx_device.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace idgx
{
namespace
{
constexpr int kThreads = 256;
constexpr int kGridBatchVis = 256; // visibilities per LDS stage
constexpr int kGridChunk = 64;     // channels per chunk (bounds the per-stage uvw count at 256 timesteps)
constexpr int kDegridPixBatch = 512;

// Visibility x phasor accumulated into acc[0..7] = {XX.re, XX.im, XY.re, ..., YY.im}.
__device__ __forceinline__ void cmac4(float *acc, float4 va, float4 vb, float cs, float sn)
{
    acc[0] = __builtin_fmaf(va.x, cs, acc[0]);
    acc[0] = __builtin_fmaf(-va.y, sn, acc[0]);
    acc[1] = __builtin_fmaf(va.x, sn, acc[1]);
    acc[1] = __builtin_fmaf(va.y, cs, acc[1]);
    acc[2] = __builtin_fmaf(va.z, cs, acc[2]);
    acc[2] = __builtin_fmaf(-va.w, sn, acc[2]);
    acc[3] = __builtin_fmaf(va.z, sn, acc[3]);
    acc[3] = __builtin_fmaf(va.w, cs, acc[3]);
    acc[4] = __builtin_fmaf(vb.x, cs, acc[4]);
    acc[4] = __builtin_fmaf(-vb.y, sn, acc[4]);
    acc[5] = __builtin_fmaf(vb.x, sn, acc[5]);
    acc[5] = __builtin_fmaf(vb.y, cs, acc[5]);
    acc[6] = __builtin_fmaf(vb.z, cs, acc[6]);
    acc[6] = __builtin_fmaf(-vb.w, sn, acc[6]);
    acc[7] = __builtin_fmaf(vb.z, sn, acc[7]);
    acc[7] = __builtin_fmaf(vb.w, cs, acc[7]);
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

// Five workgroups per CU (five waves per SIMD) with the hardware sin/cos.  The kernel needs 89 VGPRs with or
// without the bound, yet the bounded build measured 3 % faster on the SKA-Low-like workload (the compiler
// schedules for the tighter budget).
template <int NPIX, int SC>
__global__ __launch_bounds__(kThreads, SC == kSincosHw ? 5 : 1) void gridder_plain(KernelArgs a)
{
    __shared__ float4 s_vis[2 * kGridBatchVis];
    __shared__ float4 s_uvw[kGridBatchVis];
    __shared__ float s_wn[kGridChunk];

    const int s = blockIdx.x;
    const int tid = threadIdx.x;
    const int S = a.p.S;
    const int P = S * S;
    const idgx_metadata m = a.metadata[s];
    int c0, nc;
    channel_window(a, s, c0, nc);
    const long long row0 = first_row(a.metadata, m);
    const bool ok = subgrid_in_bounds(a.p, m, row0, c0, nc);
    const long long C = a.p.nr_channels;
    const int pbase = blockIdx.y * (kThreads * NPIX);

    const Offsets off = gridder_offsets(m, S, a.p.G, a.p.image_size, a.p.w_step);
    float l[NPIX], mm[NPIX], n[NPIX], poff[NPIX];
    float acc[NPIX][8];
#pragma unroll
    for (int i = 0; i < NPIX; ++i)
    {
        const int p = min(pbase + i * kThreads + tid, P - 1);
        const int x = p % S, y = p / S;
        l[i] = pixel_lm(x, S, a.p.image_size);
        mm[i] = pixel_lm(y, S, a.p.image_size);
        n[i] = pixel_n(l[i], mm[i]);
        poff[i] = (off.u * l[i] + off.v * mm[i]) + off.w * n[i];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            acc[i][k] = 0.F;
    }

    if (ok)
    {
        const float4 *vis = reinterpret_cast<const float4 *>(a.visibilities);
        for (int cb = 0; cb < nc; cb += kGridChunk)
        {
            const int cc = min(kGridChunk, nc - cb);
            const int tb_max = kGridBatchVis / cc;
            for (int t0 = 0; t0 < m.nr_timesteps; t0 += tb_max)
            {
                const int tb = min(tb_max, m.nr_timesteps - t0);
                __syncthreads(); // previous stage fully consumed
                if (tid < cc)
                    s_wn[tid] = a.wavenumbers[c0 + cb + tid];
                for (int q = tid; q < tb; q += kThreads)
                {
                    const float *uv = a.uvw + 3 * (row0 + t0 + q);
                    s_uvw[q] = make_float4(uv[0], uv[1], uv[2], 0.F);
                }
                for (int q = tid; q < 2 * tb * cc; q += kThreads)
                {
                    const int item = q >> 1;
                    const int t = item / cc, c = item - t * cc;
                    s_vis[q] = vis[((row0 + t0 + t) * C + c0 + cb + c) * 2 + (q & 1)];
                }
                __syncthreads();
                for (int t = 0; t < tb; ++t)
                {
                    const float4 uv = s_uvw[t];
                    float pidx[NPIX];
#pragma unroll
                    for (int i = 0; i < NPIX; ++i)
                        pidx[i] = (uv.x * l[i] + uv.y * mm[i]) + uv.z * n[i];
                    for (int c = 0; c < cc; ++c)
                    {
                        const float k = s_wn[c];
                        const float4 va = s_vis[2 * (t * cc + c)];
                        const float4 vb = s_vis[2 * (t * cc + c) + 1];
#pragma unroll
                        for (int i = 0; i < NPIX; ++i)
                        {
                            const float phase = poff[i] - pidx[i] * k;
                            float sn, cs;
                            sincos_rad<SC>(phase, sn, cs);
                            cmac4(acc[i], va, vb, cs, sn);
                        }
                    }
                }
            }
        }
    }

    // Epilogue: P <- taper * A1^H P A2, overwrite the subgrid (zeros for a subgrid rejected above).
    float2 *out = reinterpret_cast<float2 *>(a.subgrids_out) + static_cast<size_t>(s) * kCorr * P;
    const float4 *aterms = reinterpret_cast<const float4 *>(a.aterms);
#pragma unroll
    for (int i = 0; i < NPIX; ++i)
    {
        const int p = pbase + i * kThreads + tid;
        if (p >= P)
            continue;
        cfloat px[4] = {{acc[i][0], acc[i][1]}, {acc[i][2], acc[i][3]}, {acc[i][4], acc[i][5]}, {acc[i][6], acc[i][7]}};
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

// Workgroups per CU with the hardware sin/cos: two channels per thread fit six (the most the 24 KB of LDS allows)
// in 79 VGPRs without spills, four channels per thread fit five in 96 VGPRs.
template <int NCH, int SC> constexpr int degrid_plain_min_blocks()
{
    return SC == kSincosHw ? (NCH <= 2 ? 6 : 5) : 1;
}

template <int NCH, int SC>
__global__ __launch_bounds__(kThreads, (degrid_plain_min_blocks<NCH, SC>())) void degridder_plain(KernelArgs a,
                                                                                                  int per)
{
    // Per staged pixel: (l, m, n, phase offset), then the four correlations, side by side (one address register
    // serves the pixel loop's three reads).
    __shared__ float4 s_px[3 * kDegridPixBatch];

    // Workgroup b = vblock * n + s: a subgrid's workgroups are spread over the XCDs (balances uneven subgrids).
    const int vblock = blockIdx.x / a.p.nr_subgrids;
    const int s = blockIdx.x - vblock * a.p.nr_subgrids;
    if (vblock >= per)
        return; // uniform
    const int tid = threadIdx.x;
    const int S = a.p.S;
    const int P = S * S;
    const idgx_metadata m = a.metadata[s];
    int c0, nc;
    channel_window(a, s, c0, nc);
    const long long row0 = first_row(a.metadata, m);
    if (!subgrid_in_bounds(a.p, m, row0, c0, nc) || nc == 0)
        return; // uniform: nothing of this subgrid is written
    const int nchunks = (nc + NCH - 1) / NCH;
    if (static_cast<long long>(vblock) * kThreads >= static_cast<long long>(m.nr_timesteps) * nchunks)
        return; // uniform: no item of this block exists

    const int item = vblock * kThreads + tid;
    const int t = item / nchunks;
    const int cj = (item - t * nchunks) * NCH;
    const bool active = t < m.nr_timesteps;
    const long long C = a.p.nr_channels;

    float u = 0.F, v = 0.F, w = 0.F;
    float k[NCH];
#pragma unroll
    for (int j = 0; j < NCH; ++j)
        k[j] = 0.F;
    if (active)
    {
        const float *uv = a.uvw + 3 * (row0 + t);
        u = uv[0];
        v = uv[1];
        w = uv[2];
#pragma unroll
        for (int j = 0; j < NCH; ++j)
            if (cj + j < nc)
                k[j] = a.wavenumbers[c0 + cj + j];
    }
    float acc[NCH][8];
#pragma unroll
    for (int j = 0; j < NCH; ++j)
#pragma unroll
        for (int q = 0; q < 8; ++q)
            acc[j][q] = 0.F;

    const Offsets off = degridder_offsets(m, S, a.p.G, a.p.image_size, a.p.w_step);
    const float2 *sub = reinterpret_cast<const float2 *>(a.subgrids) + static_cast<size_t>(s) * kCorr * P;
    const float4 *aterms = reinterpret_cast<const float4 *>(a.aterms);
    const size_t abase = static_cast<size_t>(m.aterm_index) * a.p.nr_stations;

    for (int pb = 0; pb < P; pb += kDegridPixBatch)
    {
        const int np = min(kDegridPixBatch, P - pb);
        __syncthreads();
        for (int q = tid; q < np; q += kThreads)
        {
            const int p = pb + q;
            const int x = p % S, y = p / S;
            const float taper = a.spheroidal[p];
            cfloat px[4];
#pragma unroll
            for (int pol = 0; pol < 4; ++pol)
            {
                const float2 g = sub[pol * P + p];
                px[pol] = {taper * g.x, taper * g.y};
            }
            cfloat a1[4], a2[4];
            load_jones(aterms, (abase + m.station1) * P + p, a1);
            load_jones(aterms, (abase + m.station2) * P + p, a2);
            if (a.p.convention == IDGX_ATERM_REFERENCE)
                apply_a1h_p_a2(px, a1, a2);
            else
                apply_a1_p_a2h(px, a1, a2);
            s_px[3 * q + 1] = make_float4(px[0].re, px[0].im, px[1].re, px[1].im);
            s_px[3 * q + 2] = make_float4(px[2].re, px[2].im, px[3].re, px[3].im);
            const float l = pixel_lm(x, S, a.p.image_size);
            const float mm = pixel_lm(y, S, a.p.image_size);
            const float n = pixel_n(l, mm);
            s_px[3 * q] = make_float4(l, mm, n, (off.u * l + off.v * mm) + off.w * n);
        }
        __syncthreads();
        if (active)
        {
            for (int q = 0; q < np; ++q)
            {
                const float4 g = s_px[3 * q];
                const float4 pa = s_px[3 * q + 1];
                const float4 pbv = s_px[3 * q + 2];
                const float pidx = (u * g.x + v * g.y) + w * g.z;
#pragma unroll
                for (int j = 0; j < NCH; ++j)
                {
                    if (cj + j < nc)
                    {
                        const float phase = pidx * k[j] - g.w;
                        float sn, cs;
                        sincos_rad<SC>(phase, sn, cs);
                        cmac4(acc[j], pa, pbv, cs, sn);
                    }
                }
            }
        }
    }
    if (active)
    {
        float4 *out = reinterpret_cast<float4 *>(a.visibilities_out);
#pragma unroll
        for (int j = 0; j < NCH; ++j)
            if (cj + j < nc)
            {
                const size_t o = ((row0 + t) * C + c0 + cj + j) * 2;
                out[o] = make_float4(acc[j][0], acc[j][1], acc[j][2], acc[j][3]);
                out[o + 1] = make_float4(acc[j][4], acc[j][5], acc[j][6], acc[j][7]);
            }
    }
}

// ---- persistent work-queue forms (batches) ---------------------------------------------------------------------
//
// The same arithmetic as the kernels above, item by item from a device queue, over each subgrid's selected
// visibilities: a selection list (vis_index, flagged samples left out) or every (t, c) of its channel window.
// Flagged samples are skipped rather than multiplied by zero.  A skipped zero term leaves a running sum unchanged,
// so a gridder item equals the dense kernel on zeroed visibilities wherever both sum in (t, c) order (windows of at
// most 64 channels; wider windows are summed in 64-channel chunks by the dense kernel).

constexpr int kQueueVis = 256; // gridder entries per LDS stage (one per thread)

struct PlainGridShared
{
    // Per staged entry: its four correlations (two float4), then (k, 0, 0, 0), side by side, so that the entry loop
    // reads all three from one address register with constant offsets.
    float4 ent[3 * kQueueVis];
    float4 run_uvw[kQueueVis]; // uvw of each run of entries that share a timestep
    int run_start[kQueueVis + 1];
    int wave_runs[kThreads / 64];
};

// Per staged pixel: (l, m, n, phase offset), then the four correlations (two float4), side by side, so that the
// pixel loop reads all three from one address register with constant offsets.
struct PlainDegridShared
{
    float4 px[3 * kDegridPixBatch];
};

// Selected visibilities of subgrid s: the list entries (null = dense, every (t, c) of the window) and their count.
__device__ __forceinline__ int plain_selection(const KernelArgs &a, int s, int nr_timesteps, int nc,
                                               const uint32_t *&index)
{
    if (a.vis_index)
    {
        const long long o0 = a.vis_offset[s], o1 = a.vis_offset[s + 1];
        index = a.vis_index + o0;
        return static_cast<int>(o1 - o0);
    }
    index = nullptr;
    return nr_timesteps * nc;
}

// Entry v of a selection as (t, c) relative to the subgrid's first row and channel window; false for a list entry
// outside the window, which the kernels skip.
__device__ __forceinline__ bool plain_entry(const uint32_t *index, int v, int nr_timesteps, int nc, int &t, int &c)
{
    if (index)
    {
        const uint32_t e = index[v];
        t = static_cast<int>(e >> 16);
        c = static_cast<int>(e & 0xffffu);
        return t < nr_timesteps && c < nc;
    }
    t = v / nc;
    c = v - t * nc;
    return true;
}

// Gridder item: pixels [pblk * 256 * NPIX, ...) of subgrid s.  Entries are staged 256 at a time; the entries of a
// stage that share a timestep form a run (a workgroup scan over "timestep differs from the previous entry"), and
// each run forms the per-pixel index term (u l + v m) + w n once, as the dense kernel does per timestep.
template <int NPIX, int SC>
__device__ __forceinline__ void gridder_plain_item(const KernelArgs &a, int s, int pblk, PlainGridShared &sm)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int S = a.p.S;
    const int P = S * S;
    const idgx_metadata m = a.metadata[s];
    int c0, nc;
    channel_window(a, s, c0, nc);
    const long long row0 = first_row(a.metadata, m);
    const bool ok = subgrid_in_bounds(a.p, m, row0, c0, nc);
    const long long C = a.p.nr_channels;
    const int T = m.nr_timesteps;
    const int pbase = pblk * (kThreads * NPIX);

    const Offsets off = gridder_offsets(m, S, a.p.G, a.p.image_size, a.p.w_step);
    float l[NPIX], mm[NPIX], n[NPIX], poff[NPIX];
    float acc[NPIX][8];
#pragma unroll
    for (int i = 0; i < NPIX; ++i)
    {
        const int p = min(pbase + i * kThreads + tid, P - 1);
        const int x = p % S, y = p / S;
        l[i] = pixel_lm(x, S, a.p.image_size);
        mm[i] = pixel_lm(y, S, a.p.image_size);
        n[i] = pixel_n(l[i], mm[i]);
        poff[i] = (off.u * l[i] + off.v * mm[i]) + off.w * n[i];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            acc[i][k] = 0.F;
    }

    const uint32_t *index = nullptr;
    const int V = ok && T > 0 && nc > 0 ? plain_selection(a, s, T, nc, index) : 0;
    const float4 *vis = reinterpret_cast<const float4 *>(a.visibilities);
    for (int v0 = 0; v0 < V; v0 += kQueueVis)
    {
        const int nb = min(kQueueVis, V - v0);
        __syncthreads(); // the previous stage is consumed
        bool start = false;
        int t = 0;
        if (tid < nb)
        {
            int c = 0;
            float4 va = make_float4(0.F, 0.F, 0.F, 0.F), vb = va;
            float k = 0.F;
            if (plain_entry(index, v0 + tid, T, nc, t, c))
            {
                const size_t o = static_cast<size_t>((row0 + t) * C + c0 + c) * 2;
                va = vis[o];
                vb = vis[o + 1];
                k = a.wavenumbers[c0 + c];
            }
            t = min(t, T - 1); // a skipped entry still names a row inside the subgrid
            sm.ent[3 * tid] = va;
            sm.ent[3 * tid + 1] = vb;
            sm.ent[3 * tid + 2] = make_float4(k, 0.F, 0.F, 0.F);
            start = tid == 0;
            if (!start)
            {
                int tp, cp;
                (void)plain_entry(index, v0 + tid - 1, T, nc, tp, cp);
                start = min(tp, T - 1) != t;
            }
        }
        const unsigned long long mask = __ballot(start);
        const int below = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
            static_cast<unsigned>(mask >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u)));
        if (lane == 0)
            sm.wave_runs[wave] = __popcll(mask);
        __syncthreads();
        int before = 0, nruns = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w)
        {
            const int r = sm.wave_runs[w];
            before += w < wave ? r : 0;
            nruns += r;
        }
        if (start)
        {
            const int r = before + below;
            const float *uv = a.uvw + 3 * (row0 + t);
            sm.run_start[r] = tid;
            sm.run_uvw[r] = make_float4(uv[0], uv[1], uv[2], 0.F);
        }
        if (tid == 0)
            sm.run_start[nruns] = nb;
        __syncthreads();
        nruns = __builtin_amdgcn_readfirstlane(nruns);
        for (int r = 0; r < nruns; ++r)
        {
            const float4 uv = sm.run_uvw[r];
            const int e0 = __builtin_amdgcn_readfirstlane(sm.run_start[r]);
            const int e1 = __builtin_amdgcn_readfirstlane(sm.run_start[r + 1]);
            float pidx[NPIX];
#pragma unroll
            for (int i = 0; i < NPIX; ++i)
                pidx[i] = (uv.x * l[i] + uv.y * mm[i]) + uv.z * n[i];
            for (int e = e0; e < e1; ++e)
            {
                const float4 va = sm.ent[3 * e];
                const float4 vb = sm.ent[3 * e + 1];
                const float k = sm.ent[3 * e + 2].x;
#pragma unroll
                for (int i = 0; i < NPIX; ++i)
                {
                    const float phase = poff[i] - pidx[i] * k;
                    float sn, cs;
                    sincos_rad<SC>(phase, sn, cs);
                    cmac4(acc[i], va, vb, cs, sn);
                }
            }
        }
    }

    // Epilogue: P <- taper * A1^H P A2, overwrite the item's pixels (zeros for a subgrid rejected above).
    float2 *out = reinterpret_cast<float2 *>(a.subgrids_out) + static_cast<size_t>(s) * kCorr * P;
    const float4 *aterms = reinterpret_cast<const float4 *>(a.aterms);
#pragma unroll
    for (int i = 0; i < NPIX; ++i)
    {
        const int p = pbase + i * kThreads + tid;
        if (p >= P)
            continue;
        cfloat px[4] = {{acc[i][0], acc[i][1]}, {acc[i][2], acc[i][3]}, {acc[i][4], acc[i][5]}, {acc[i][6], acc[i][7]}};
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

// Degridder item: selected visibilities [vblk * 256 * NCH, ...) of subgrid s, NCH consecutive ones per thread.  When
// every thread of a wave holds visibilities of a single timestep (dense windows whose width NCH divides), the index
// term is formed once per pixel for all of them; otherwise once per visibility.
template <int NCH, int SC>
__device__ __forceinline__ void degridder_plain_item(const KernelArgs &a, int s, int vblk, PlainDegridShared &sm)
{
    const int tid = threadIdx.x;
    const int S = a.p.S;
    const int P = S * S;
    const idgx_metadata m = a.metadata[s];
    int c0, nc;
    channel_window(a, s, c0, nc);
    const long long row0 = first_row(a.metadata, m);
    const int T = m.nr_timesteps;
    if (!subgrid_in_bounds(a.p, m, row0, c0, nc) || nc == 0 || T == 0)
        return; // uniform: nothing of this subgrid is written
    const uint32_t *index = nullptr;
    const int V = plain_selection(a, s, T, nc, index);
    const int ebase = vblk * kThreads * NCH;
    if (ebase >= V)
        return; // uniform: no visibility of this item exists
    const long long C = a.p.nr_channels;

    float u[NCH], v[NCH], w[NCH], k[NCH];
    unsigned valid = 0u;
    int tfirst = -1;
    bool same = true;
#pragma unroll
    for (int j = 0; j < NCH; ++j)
    {
        u[j] = v[j] = w[j] = k[j] = 0.F;
        const int e = ebase + tid * NCH + j;
        int t, c;
        if (e < V && plain_entry(index, e, T, nc, t, c))
        {
            valid |= 1u << j;
            const float *uv = a.uvw + 3 * (row0 + t);
            u[j] = uv[0];
            v[j] = uv[1];
            w[j] = uv[2];
            k[j] = a.wavenumbers[c0 + c];
            same = same && (tfirst < 0 || t == tfirst);
            tfirst = tfirst < 0 ? t : tfirst;
        }
    }
    // The shared index term uses the thread's first visibility with data.
    float us = 0.F, vs = 0.F, ws = 0.F;
#pragma unroll
    for (int j = NCH - 1; j >= 0; --j)
        if (valid & (1u << j))
        {
            us = u[j];
            vs = v[j];
            ws = w[j];
        }
    const bool wave_live = __ballot(valid != 0u) != 0ull;   // wave-uniform
    const bool wave_shared = __ballot(!same) == 0ull;       // wave-uniform
    float acc[NCH][8];
#pragma unroll
    for (int j = 0; j < NCH; ++j)
#pragma unroll
        for (int q = 0; q < 8; ++q)
            acc[j][q] = 0.F;

    const Offsets off = degridder_offsets(m, S, a.p.G, a.p.image_size, a.p.w_step);
    const float2 *sub = reinterpret_cast<const float2 *>(a.subgrids) + static_cast<size_t>(s) * kCorr * P;
    const float4 *aterms = reinterpret_cast<const float4 *>(a.aterms);
    const size_t abase = static_cast<size_t>(m.aterm_index) * a.p.nr_stations;
    for (int pb = 0; pb < P; pb += kDegridPixBatch)
    {
        const int np = min(kDegridPixBatch, P - pb);
        __syncthreads();
        for (int q = tid; q < np; q += kThreads)
        {
            const int p = pb + q;
            const int x = p % S, y = p / S;
            const float taper = a.spheroidal[p];
            cfloat px[4];
#pragma unroll
            for (int pol = 0; pol < 4; ++pol)
            {
                const float2 g = sub[pol * P + p];
                px[pol] = {taper * g.x, taper * g.y};
            }
            cfloat a1[4], a2[4];
            load_jones(aterms, (abase + m.station1) * P + p, a1);
            load_jones(aterms, (abase + m.station2) * P + p, a2);
            if (a.p.convention == IDGX_ATERM_REFERENCE)
                apply_a1h_p_a2(px, a1, a2);
            else
                apply_a1_p_a2h(px, a1, a2);
            sm.px[3 * q + 1] = make_float4(px[0].re, px[0].im, px[1].re, px[1].im);
            sm.px[3 * q + 2] = make_float4(px[2].re, px[2].im, px[3].re, px[3].im);
            const float l = pixel_lm(x, S, a.p.image_size);
            const float mm = pixel_lm(y, S, a.p.image_size);
            const float n = pixel_n(l, mm);
            sm.px[3 * q] = make_float4(l, mm, n, (off.u * l + off.v * mm) + off.w * n);
        }
        __syncthreads();
        if (!wave_live)
            continue;
        if (NCH > 1 && wave_shared)
        {
            for (int q = 0; q < np; ++q)
            {
                const float4 g = sm.px[3 * q];
                const float4 pa = sm.px[3 * q + 1];
                const float4 pbv = sm.px[3 * q + 2];
                const float pidx = (us * g.x + vs * g.y) + ws * g.z;
#pragma unroll
                for (int j = 0; j < NCH; ++j)
                {
                    const float phase = pidx * k[j] - g.w;
                    float sn, cs;
                    sincos_rad<SC>(phase, sn, cs);
                    cmac4(acc[j], pa, pbv, cs, sn);
                }
            }
        }
        else
        {
            for (int q = 0; q < np; ++q)
            {
                const float4 g = sm.px[3 * q];
                const float4 pa = sm.px[3 * q + 1];
                const float4 pbv = sm.px[3 * q + 2];
#pragma unroll
                for (int j = 0; j < NCH; ++j)
                {
                    const float pidx = (u[j] * g.x + v[j] * g.y) + w[j] * g.z;
                    const float phase = pidx * k[j] - g.w;
                    float sn, cs;
                    sincos_rad<SC>(phase, sn, cs);
                    cmac4(acc[j], pa, pbv, cs, sn);
                }
            }
        }
    }
    float4 *out = reinterpret_cast<float4 *>(a.visibilities_out);
#pragma unroll
    for (int j = 0; j < NCH; ++j)
    {
        if (!(valid & (1u << j)))
            continue;
        int t, c;
        (void)plain_entry(index, ebase + tid * NCH + j, T, nc, t, c);
        const size_t o = static_cast<size_t>((row0 + t) * C + c0 + c) * 2;
        out[o] = make_float4(acc[j][0], acc[j][1], acc[j][2], acc[j][3]);
        out[o + 1] = make_float4(acc[j][4], acc[j][5], acc[j][6], acc[j][7]);
    }
}

template <int NPIX, int SC>
__global__ __launch_bounds__(kThreads, SC == kSincosHw ? 5 : 1) void gridder_plain_queue(QueueArgs q)
{
    __shared__ PlainGridShared sm;
    int cls, s, blk;
    while (next_queue_item(q, cls, s, blk))
        gridder_plain_item<NPIX, SC>(q.cls[cls], s, blk, sm);
}

// One launch over items of several subgrid classes, each with its class's pixels per thread (q.tiles[class]).
template <int SC>
__global__ __launch_bounds__(kThreads, SC == kSincosHw ? 5 : 1) void gridder_plain_queue_tiles(QueueArgs q)
{
    __shared__ PlainGridShared sm;
    int cls, s, blk;
    while (next_queue_item(q, cls, s, blk))
    {
        switch (q.tiles[cls]) // workgroup-uniform
        {
        case 2:
            gridder_plain_item<2, SC>(q.cls[cls], s, blk, sm);
            break;
        case 3:
            gridder_plain_item<3, SC>(q.cls[cls], s, blk, sm);
            break;
        default:
            gridder_plain_item<4, SC>(q.cls[cls], s, blk, sm);
            break;
        }
    }
}

// Workgroups per CU of the queue degridder with the hardware sin/cos: it also holds every visibility's uvw and, for
// several visibilities per thread, runs one of two inner loops, so it keeps fewer workgroups than the dense kernel:
// the most that leave no spills (96, 128 and 168 VGPRs for 1, 2 and 4 visibilities per thread).
template <int NCH, int SC> constexpr int degrid_plain_queue_min_blocks()
{
    return SC == kSincosHw ? (NCH == 1 ? 5 : NCH == 2 ? 4 : 3) : 1;
}

template <int NCH, int SC>
__global__ __launch_bounds__(kThreads, (degrid_plain_queue_min_blocks<NCH, SC>())) void degridder_plain_queue(
    QueueArgs q)
{
    __shared__ PlainDegridShared sm;
    int cls, s, blk;
    while (next_queue_item(q, cls, s, blk))
        degridder_plain_item<NCH, SC>(q.cls[cls], s, blk, sm);
}

// Shapes of the persistent plain kernels: {gridder pixels per thread, degridder visibilities per thread}.
constexpr PlainShape kPlainShapes[] = {{4, 2}, {3, 4}, {2, 1}};
constexpr int kNumPlainShapes = sizeof(kPlainShapes) / sizeof(kPlainShapes[0]);

template <int SC> void *plain_grid_queue_fn(int shape)
{
    if (shape == kShapeTilesPerClass)
        return reinterpret_cast<void *>(&gridder_plain_queue_tiles<SC>);
    if (shape < 0 || shape >= kNumPlainShapes)
        return nullptr;
    switch (kPlainShapes[shape].pixels)
    {
    case 2:
        return reinterpret_cast<void *>(&gridder_plain_queue<2, SC>);
    case 3:
        return reinterpret_cast<void *>(&gridder_plain_queue<3, SC>);
    default:
        return reinterpret_cast<void *>(&gridder_plain_queue<4, SC>);
    }
}

template <int SC> void *plain_degrid_queue_fn(int shape)
{
    if (shape < 0 || shape >= kNumPlainShapes)
        return nullptr;
    switch (kPlainShapes[shape].entries)
    {
    case 1:
        return reinterpret_cast<void *>(&degridder_plain_queue<1, SC>);
    case 2:
        return reinterpret_cast<void *>(&degridder_plain_queue<2, SC>);
    default:
        return reinterpret_cast<void *>(&degridder_plain_queue<4, SC>);
    }
}

void *plain_queue_fn(bool accurate_sincos, bool gridder, int shape)
{
    if (gridder)
        return accurate_sincos ? plain_grid_queue_fn<kSincosOcml>(shape) : plain_grid_queue_fn<kSincosHw>(shape);
    return accurate_sincos ? plain_degrid_queue_fn<kSincosOcml>(shape) : plain_degrid_queue_fn<kSincosHw>(shape);
}

hipError_t launch_plain_queue(const QueueArgs &q, bool accurate_sincos, bool gridder, int shape, int workgroups,
                              hipStream_t st)
{
    void *fn = plain_queue_fn(accurate_sincos, gridder, shape);
    if (!fn || workgroups <= 0)
        return hipErrorInvalidValue;
    QueueArgs arg = q;
    void *args[] = {&arg};
    return hipLaunchKernel(fn, dim3(workgroups), dim3(kThreads), args, 0, st);
}
} // namespace

template <int NPIX> hipError_t launch_grid_plain_pixels(const KernelArgs &a, bool accurate_sincos, hipStream_t stream)
{
    const int P = a.p.S * a.p.S;
    const dim3 grid(a.p.nr_subgrids, (P + kThreads * NPIX - 1) / (kThreads * NPIX));
    if (accurate_sincos)
        gridder_plain<NPIX, kSincosOcml><<<grid, kThreads, 0, stream>>>(a);
    else
        gridder_plain<NPIX, kSincosHw><<<grid, kThreads, 0, stream>>>(a);
    return hipGetLastError();
}

// Pixels per gridder thread: 4 (1024 per workgroup) unless 3 (768) leaves fewer lanes idle, as for S = 24 (576
// pixels: 75 % of the lanes busy with 3, 56 % with 4) and S = 48 (2304: 100 % with 3, 75 % with 4).  Lanes past
// the last pixel still run the visibility loop, so idle lanes cost time; the per-pixel cost is the same for both.
hipError_t launch_grid_plain(const KernelArgs &a, bool accurate_sincos, hipStream_t stream)
{
    if (plain_grid_pixels_for(a.p.S) == 3)
        return launch_grid_plain_pixels<3>(a, accurate_sincos, stream);
    return launch_grid_plain_pixels<4>(a, accurate_sincos, stream);
}

template <int NCH> hipError_t launch_degrid_plain_chunks(const KernelArgs &a, bool accurate_sincos, hipStream_t stream)
{
    const long long items = static_cast<long long>(a.p.max_nr_timesteps) * ((a.p.nr_channels + NCH - 1) / NCH);
    const long long per = items > 0 ? (items + kThreads - 1) / kThreads : 1;
    const long long blocks = static_cast<long long>(a.p.nr_subgrids) * per;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    const dim3 grid(static_cast<unsigned>(blocks));
    if (accurate_sincos)
        degridder_plain<NCH, kSincosOcml><<<grid, kThreads, 0, stream>>>(a, static_cast<int>(per));
    else
        degridder_plain<NCH, kSincosHw><<<grid, kThreads, 0, stream>>>(a, static_cast<int>(per));
    return hipGetLastError();
}

// Channels per degridder thread.  Four make each pixel's LDS reads and index term serve more visibilities, but a
// chunk that runs past the end of a subgrid's channel window idles its lanes: four channels when every subgrid
// takes all channels (no windows) and their count is a multiple of 4, two otherwise.  Measured: the reference
// generator's workload (16 channels, no windows) 17.1 ms with four, 18.3 ms with two; the SKA-Low-like workload
// (windows of 13.5 channels on average: 90 % of the lanes busy with four, 96.7 % with two) 82.3 and 79.6 ms.
hipError_t launch_degrid_plain(const KernelArgs &a, bool accurate_sincos, hipStream_t stream)
{
    if (plain_degrid_channels_for(a.channels != nullptr, a.p.nr_channels) == 4)
        return launch_degrid_plain_chunks<4>(a, accurate_sincos, stream);
    return launch_degrid_plain_chunks<2>(a, accurate_sincos, stream);
}

// Pixels per gridder thread for subgrid size S: 4 unless 3 leaves fewer lanes idle (see launch_grid_plain).
int plain_grid_pixels_for(int S)
{
    const long long P = static_cast<long long>(S) * S;
    const long long slots4 = (P + 1023) / 1024 * 1024, slots3 = (P + 767) / 768 * 768;
    return slots3 < slots4 ? 3 : 4;
}

// Channels per degridder thread (see launch_degrid_plain).
int plain_degrid_channels_for(bool channel_windows, int nr_channels)
{
    return !channel_windows && nr_channels % 4 == 0 ? 4 : 2;
}

int plain_shape_count()
{
    return kNumPlainShapes;
}

PlainShape plain_shape(int i)
{
    return kPlainShapes[i];
}

int plain_default_shape(bool gridder, int S)
{
    if (!gridder)
        return 0; // two visibilities per thread
    const int px = plain_grid_pixels_for(S);
    for (int i = 0; i < kNumPlainShapes; ++i)
        if (kPlainShapes[i].pixels == px)
            return i;
    return 0;
}

int plain_grid_items_per_subgrid(int shape, int S)
{
    const long long per = static_cast<long long>(kThreads) * kPlainShapes[shape].pixels;
    return static_cast<int>((static_cast<long long>(S) * S + per - 1) / per);
}

long long plain_grid_item_pixels(int shape, int S, int item)
{
    const long long per = static_cast<long long>(kThreads) * kPlainShapes[shape].pixels;
    const long long P = static_cast<long long>(S) * S;
    return std::max(0LL, std::min(per, P - item * per));
}

int plain_degrid_visibilities_per_item(int shape)
{
    return kThreads * kPlainShapes[shape].entries;
}

hipError_t launch_grid_plain_queue(const QueueArgs &q, bool accurate_sincos, int shape, int workgroups,
                                   hipStream_t st)
{
    return launch_plain_queue(q, accurate_sincos, true, shape, workgroups, st);
}

hipError_t launch_degrid_plain_queue(const QueueArgs &q, bool accurate_sincos, int shape, int workgroups,
                                     hipStream_t st)
{
    return launch_plain_queue(q, accurate_sincos, false, shape, workgroups, st);
}

int plain_queue_workgroups(bool accurate_sincos, bool gridder, int shape)
{
    void *fn = plain_queue_fn(accurate_sincos, gridder, shape);
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
