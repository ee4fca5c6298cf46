// Device-side IDG planner: time segments of baselines -> subgrid metadata + channel windows, for gfx950.
//
// The same plan as the host planner `plan_segments` of ska-spc-tender-benchmarks_amd/idgx/observation.py (the
// reference has no planner: its metadata are random, gridding/src/gridding_initialization.cpp:158-182), operation
// for operation in FP64 with contraction off, so the two are bit-identical:
//   1. per segment (baseline b, rows b * T + [t0, t1)): extent of u, v (and w) in metres;
//   2. channels grouped greedily: a group [a, c - 1] closes when adding channel c would make its uv footprint,
//      in grid cells (metres x f / c x image_size), wider than S - support in u or v;
//   3. each group becomes a subgrid centred on its footprint, kept if it lies inside the grid; with w_step its
//      w-layer is z = floor(middle of the w range x centre frequency / c / w_step).
// Subgrids come out in segment order, each segment's groups in channel order: with segments sorted by
// (baseline, t0) that is the host planner's (first row, first channel) order.
//
// Three passes: count the kept groups per segment (one thread per segment), an exclusive scan of the counts (one
// workgroup), write them.  The workspace holds the counts and offsets (16 B per segment).
#include "idgx_device.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace idgx
{
namespace
{
constexpr int kThreads = 256;
constexpr double kSpeedOfLight = 299792458.0;

struct SegmentExtent
{
    double umin, umax, vmin, vmax, wmin, wmax;
};

__device__ SegmentExtent segment_extent(const float *uvw, long long r0, long long r1)
{
    SegmentExtent e;
    e.umin = e.vmin = e.wmin = __builtin_inf();
    e.umax = e.vmax = e.wmax = -__builtin_inf();
    for (long long r = r0; r < r1; ++r)
    {
        const double u = uvw[3 * r], v = uvw[3 * r + 1], w = uvw[3 * r + 2];
        e.umin = fmin(e.umin, u);
        e.umax = fmax(e.umax, u);
        e.vmin = fmin(e.vmin, v);
        e.vmax = fmax(e.vmax, v);
        e.wmin = fmin(e.wmin, w);
        e.wmax = fmax(e.wmax, w);
    }
    return e;
}

struct Footprint
{
    double lo_u, hi_u, lo_v, hi_v;
};

// Grid-cell extent of the segment's samples with channels a .. z (scale = cells per metre of each channel).
__device__ __forceinline__ Footprint footprint(const SegmentExtent &e, double fa, double fz)
{
    Footprint f;
    f.lo_u = fmin(e.umin * fa, e.umin * fz);
    f.hi_u = fmax(e.umax * fa, e.umax * fz);
    f.lo_v = fmin(e.vmin * fa, e.vmin * fz);
    f.hi_v = fmax(e.vmax * fa, e.vmax * fz);
    return f;
}

struct PlannerArgs
{
    idgx_planner pl;
    const float *uvw;
    const double *frequencies;
    const uint32_t *station1, *station2;
    const int32_t *seg_bl, *seg_t0, *seg_t1;
    long long nr_segments;
    int32_t *counts;
    long long *offsets;
    idgx_metadata *metadata;
    idgx_channels *channels;
    long long capacity;
};

// Walks the channel groups of segment i; emit(a, z, x0, y0) is called for every kept group in channel order.
template <typename Emit> __device__ void plan_segment(const PlannerArgs &A, long long i, Emit emit)
{
    const idgx_planner &pl = A.pl;
    const long long r0 = static_cast<long long>(A.seg_bl[i]) * pl.nr_times;
    const SegmentExtent e = segment_extent(A.uvw, r0 + A.seg_t0[i], r0 + A.seg_t1[i]);
    const double width = pl.subgrid_size - pl.support;
    const int C = pl.nr_channels;
    int first = 0;
    for (int c = 1; c <= C; ++c)
    {
        bool close = true;
        if (c < C)
        {
            const Footprint f = footprint(e, A.frequencies[first] / kSpeedOfLight * pl.image_size,
                                          A.frequencies[c] / kSpeedOfLight * pl.image_size);
            close = !((f.hi_u - f.lo_u <= width) && (f.hi_v - f.lo_v <= width));
        }
        if (!close)
            continue;
        const int last = c - 1;
        const Footprint f = footprint(e, A.frequencies[first] / kSpeedOfLight * pl.image_size,
                                      A.frequencies[last] / kSpeedOfLight * pl.image_size);
        const double G = pl.grid_size, S = pl.subgrid_size;
        const long long x0 = static_cast<long long>(floor((f.lo_u + f.hi_u) / 2 + G / 2 - S / 2 + 0.5));
        const long long y0 = static_cast<long long>(floor((f.lo_v + f.hi_v) / 2 + G / 2 - S / 2 + 0.5));
        if (x0 >= 0 && x0 + pl.subgrid_size <= pl.grid_size && y0 >= 0 && y0 + pl.subgrid_size <= pl.grid_size)
            emit(first, last, static_cast<int>(x0), static_cast<int>(y0), e);
        first = c;
    }
}

__global__ __launch_bounds__(kThreads) void plan_count_kernel(PlannerArgs A)
{
    const long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= A.nr_segments)
        return;
    int n = 0;
    plan_segment(A, i, [&](int, int, int, int, const SegmentExtent &) { ++n; });
    A.counts[i] = n;
}

// Exclusive scan of the counts by one workgroup: each thread sums a contiguous chunk, the chunk sums are scanned in
// LDS, then each thread writes its chunk's offsets.  offsets[n] = total.
__global__ __launch_bounds__(kThreads) void plan_scan_kernel(const int32_t *counts, long long *offsets, long long n)
{
    __shared__ long long s_sum[kThreads];
    const int tid = threadIdx.x;
    const long long chunk = (n + kThreads - 1) / kThreads;
    const long long b = tid * chunk, e = min(n, b + chunk);
    long long sum = 0;
    for (long long i = b; i < e; ++i)
        sum += counts[i];
    s_sum[tid] = sum;
    __syncthreads();
    if (tid == 0)
    {
        long long run = 0;
        for (int t = 0; t < kThreads; ++t)
        {
            const long long v = s_sum[t];
            s_sum[t] = run;
            run += v;
        }
        offsets[n] = run;
    }
    __syncthreads();
    long long run = s_sum[tid];
    for (long long i = b; i < e; ++i)
    {
        offsets[i] = run;
        run += counts[i];
    }
}

__global__ __launch_bounds__(kThreads) void plan_write_kernel(PlannerArgs A)
{
    const long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= A.nr_segments)
        return;
    long long o = A.offsets[i];
    const idgx_planner &pl = A.pl;
    const int bl = A.seg_bl[i], t0 = A.seg_t0[i], t1 = A.seg_t1[i];
    plan_segment(A, i, [&](int first, int last, int x0, int y0, const SegmentExtent &e) {
        if (o < A.capacity)
        {
            idgx_metadata m;
            m.baseline_offset = 0;
            m.time_offset = bl * pl.nr_times + t0;
            m.nr_timesteps = t1 - t0;
            m.aterm_index = pl.aterm_interval > 0 ? t0 / pl.aterm_interval : 0;
            m.station1 = A.station1[bl];
            m.station2 = A.station2[bl];
            m.x = x0;
            m.y = y0;
            m.z = 0;
            if (pl.w_step_in_lambda > 0)
            {
                const double f_c = 0.5 * (A.frequencies[first] + A.frequencies[last]);
                const double w_mid = 0.5 * (e.wmin + e.wmax);
                m.z = static_cast<int>(floor(w_mid * f_c / kSpeedOfLight / pl.w_step_in_lambda));
            }
            A.metadata[o] = m;
            A.channels[o] = idgx_channels{first, last - first + 1};
        }
        ++o;
    });
}
} // namespace

size_t plan_workspace_bytes(long long nr_segments)
{
    return static_cast<size_t>(nr_segments) * sizeof(int32_t) +
           static_cast<size_t>(nr_segments + 1) * sizeof(long long) + 64;
}

hipError_t launch_plan(const idgx_planner &pl, const float *uvw, const double *frequencies, const uint32_t *station1,
                       const uint32_t *station2, const int32_t *seg_bl, const int32_t *seg_t0, const int32_t *seg_t1,
                       long long nr_segments, void *workspace, idgx_metadata *metadata, idgx_channels *channels,
                       long long capacity, long long *h_count, hipStream_t st)
{
    PlannerArgs A;
    A.pl = pl;
    A.uvw = uvw;
    A.frequencies = frequencies;
    A.station1 = station1;
    A.station2 = station2;
    A.seg_bl = seg_bl;
    A.seg_t0 = seg_t0;
    A.seg_t1 = seg_t1;
    A.nr_segments = nr_segments;
    A.offsets = reinterpret_cast<long long *>(workspace);
    A.counts = reinterpret_cast<int32_t *>(A.offsets + nr_segments + 1);
    A.metadata = metadata;
    A.channels = channels;
    A.capacity = capacity;
    const long long blocks = (nr_segments + kThreads - 1) / kThreads;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    hipError_t e = hipSuccess;
    if (blocks > 0)
    {
        plan_count_kernel<<<static_cast<unsigned>(blocks), kThreads, 0, st>>>(A);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
    {
        plan_scan_kernel<<<1, kThreads, 0, st>>>(A.counts, A.offsets, nr_segments);
        e = hipGetLastError();
    }
    if (e == hipSuccess && blocks > 0)
    {
        plan_write_kernel<<<static_cast<unsigned>(blocks), kThreads, 0, st>>>(A);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_count, A.offsets + nr_segments, sizeof(long long), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    return e;
}
} // namespace idgx
