// Host-buffer entry points: the reference launcher's whole contract in one C call, for callers that hold only
// host arrays (a cgo / ctypes / JNI binding, or the C++ drop-in launcher).  The reference's CUDA launcher
// (gridding/src/gridding_benchmark.cu:248-361, gridding/src/degridding_benchmark.cu:205-318) allocates device
// buffers, copies the inputs in, runs the kernel, copies back and reports benchmark_result
// (gridding/inc/common_gridding.h:5-12).  Here each call:
//   - validates the plan against the host metadata (and optional channel windows) before touching the device;
//   - allocates the device buffers (freed on every return path);
//   - for the mixed forms, creates a one-class batch (work queue, prepared-pixel buffers) outside the timed runs;
//   - times `iterations` uploads, `iterations` kernel runs and `iterations` downloads of the result with HIP
//     events on a private stream, and reports per-iteration microseconds and MB/s.
// Errors are status codes (no exit), so a binding can report them.
#include <hip/hip_runtime.h>

#include <cstddef>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "idgx.h"

namespace
{
int to_status(hipError_t e)
{
    if (e == hipSuccess)
        return IDGX_OK;
    if (e == hipErrorOutOfMemory)
        return IDGX_ERR_OUT_OF_MEMORY;
    return IDGX_ERR_DEVICE;
}

// Device allocations, stream, events and batch of one call, released in the destructor.
struct CallState
{
    static constexpr int kBuffers = 8;
    void *buf[kBuffers] = {};
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    idgx_batch *batch = nullptr;
    ~CallState()
    {
        idgx_batch_destroy(batch);
        for (hipEvent_t e : ev)
            if (e)
                (void)hipEventDestroy(e);
        if (stream)
            (void)hipStreamDestroy(stream);
        for (void *p : buf)
            if (p)
                (void)hipFree(p);
    }
};

enum Buffer
{
    kUvw = 0,
    kWavenumbers,
    kVisibilities,
    kSpheroidal,
    kAterms,
    kMetadata,
    kSubgrids,
    kChannels
};

double elapsed_us(hipEvent_t a, hipEvent_t b, int &rc)
{
    float ms = 0.F;
    if (rc == IDGX_OK && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess))
        rc = IDGX_ERR_DEVICE;
    return 1000.0 * ms;
}

int run_host(bool gridder, const idgx_plan *plan, const idgx_metadata *h_metadata, const idgx_channels *h_channels,
             const float *h_uvw, const float *h_wavenumbers, const void *h_in, const float *h_spheroidal,
             const void *h_aterms, void *h_out, int iterations, idgx_timing *timing)
{
    if (!plan || iterations < 1)
        return IDGX_ERR_INVALID_ARGUMENT;
    int rc = idgx_check_plan(plan, h_metadata, h_channels);
    if (rc != IDGX_OK)
        return rc;

    const size_t S = static_cast<size_t>(plan->subgrid_size);
    const size_t rows = static_cast<size_t>(plan->nr_rows);
    const size_t C = static_cast<size_t>(plan->nr_channels);
    const size_t n = static_cast<size_t>(plan->nr_subgrids);
    size_t bytes[CallState::kBuffers];
    bytes[kUvw] = rows * 3 * sizeof(float);
    bytes[kWavenumbers] = C * sizeof(float);
    bytes[kVisibilities] = rows * C * 4 * 8;
    bytes[kSpheroidal] = S * S * sizeof(float);
    bytes[kAterms] = static_cast<size_t>(plan->nr_aterm_slots) * plan->nr_stations * S * S * 4 * 8;
    bytes[kMetadata] = n * sizeof(idgx_metadata);
    bytes[kSubgrids] = n * 4 * S * S * 8;
    bytes[kChannels] = h_channels ? n * sizeof(idgx_channels) : 0;
    const void *host_in[CallState::kBuffers] = {h_uvw,       h_wavenumbers, gridder ? h_in : nullptr,
                                                h_spheroidal, h_aterms,      h_metadata,
                                                gridder ? nullptr : h_in,   h_channels};
    const int out_buffer = gridder ? kSubgrids : kVisibilities;
    // Every buffer the call moves must be given (an empty plan moves no metadata and no subgrids).
    const bool ins_given = h_uvw && h_wavenumbers && h_spheroidal && h_aterms && (n == 0 || (h_metadata && h_in));
    if (!ins_given || (bytes[out_buffer] > 0 && !h_out))
        return IDGX_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < CallState::kBuffers; ++i)
        if (bytes[i] == 0)
            host_in[i] = nullptr;

    CallState cs;
    hipError_t e = hipStreamCreateWithFlags(&cs.stream, hipStreamNonBlocking);
    for (int i = 0; i < CallState::kBuffers && e == hipSuccess; ++i)
        e = hipMalloc(&cs.buf[i], bytes[i] > 0 ? bytes[i] : 1);
    for (int i = 0; i < 6 && e == hipSuccess; ++i)
        e = hipEventCreate(&cs.ev[i]);
    // The degridder writes only the entries the plan covers: the rest of the output reads back as zeros.
    if (e == hipSuccess && !gridder)
        e = hipMemsetAsync(cs.buf[kVisibilities], 0, bytes[kVisibilities], cs.stream);
    if (e != hipSuccess)
        return to_status(e);

    idgx_batch_data bd{};
    if (plan->precision != IDGX_PLAIN_FP32)
    {
        idgx_batch_class cls{};
        cls.plan = *plan;
        cls.h_metadata = h_metadata;
        cls.h_channels = h_channels;
        rc = idgx_batch_create(&cls, 1, &cs.batch);
        if (rc != IDGX_OK)
            return rc;
        bd.d_uvw = static_cast<const float *>(cs.buf[kUvw]);
        bd.d_wavenumbers = static_cast<const float *>(cs.buf[kWavenumbers]);
        bd.d_visibilities = cs.buf[kVisibilities];
        bd.d_visibilities_out = cs.buf[kVisibilities];
        bd.cls[0].d_metadata = static_cast<const idgx_metadata *>(cs.buf[kMetadata]);
        bd.cls[0].d_channels = h_channels ? static_cast<const idgx_channels *>(cs.buf[kChannels]) : nullptr;
        bd.cls[0].d_spheroidal = static_cast<const float *>(cs.buf[kSpheroidal]);
        bd.cls[0].d_aterms = cs.buf[kAterms];
        bd.cls[0].d_subgrids = cs.buf[kSubgrids];
    }

    size_t in_bytes = 0;
    for (int i = 0; i < CallState::kBuffers; ++i)
        if (host_in[i])
            in_bytes += bytes[i];
    e = hipEventRecord(cs.ev[0], cs.stream);
    for (int it = 0; it < iterations && e == hipSuccess; ++it)
        for (int i = 0; i < CallState::kBuffers && e == hipSuccess; ++i)
            if (host_in[i])
                e = hipMemcpyAsync(cs.buf[i], host_in[i], bytes[i], hipMemcpyHostToDevice, cs.stream);
    if (e == hipSuccess)
        e = hipEventRecord(cs.ev[1], cs.stream);
    if (e != hipSuccess)
        return to_status(e);

    if (hipEventRecord(cs.ev[2], cs.stream) != hipSuccess)
        return IDGX_ERR_DEVICE;
    const auto *d_md = static_cast<const idgx_metadata *>(cs.buf[kMetadata]);
    const auto *d_uvw = static_cast<const float *>(cs.buf[kUvw]);
    const auto *d_wn = static_cast<const float *>(cs.buf[kWavenumbers]);
    const auto *d_sph = static_cast<const float *>(cs.buf[kSpheroidal]);
    const auto *d_ch = h_channels ? static_cast<const idgx_channels *>(cs.buf[kChannels]) : nullptr;
    for (int it = 0; it < iterations && rc == IDGX_OK; ++it)
    {
        if (cs.batch)
            rc = gridder ? idgx_batch_grid(cs.batch, &bd, cs.stream) : idgx_batch_degrid(cs.batch, &bd, cs.stream);
        else if (gridder)
            rc = idgx_grid(plan, d_md, d_ch, d_uvw, d_wn, cs.buf[kVisibilities], d_sph, cs.buf[kAterms],
                           cs.buf[kSubgrids], cs.stream);
        else
            rc = idgx_degrid(plan, d_md, d_ch, d_uvw, d_wn, cs.buf[kSubgrids], d_sph, cs.buf[kAterms],
                             cs.buf[kVisibilities], cs.stream);
    }
    if (rc != IDGX_OK)
        return rc;
    if (hipEventRecord(cs.ev[3], cs.stream) != hipSuccess)
        return IDGX_ERR_DEVICE;

    e = hipEventRecord(cs.ev[4], cs.stream);
    for (int it = 0; it < iterations && e == hipSuccess; ++it)
        e = hipMemcpyAsync(h_out, cs.buf[out_buffer], bytes[out_buffer], hipMemcpyDeviceToHost, cs.stream);
    if (e == hipSuccess)
        e = hipEventRecord(cs.ev[5], cs.stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(cs.stream);
    if (e != hipSuccess)
        return to_status(e);

    const double in_us = elapsed_us(cs.ev[0], cs.ev[1], rc);
    const double compute_us = elapsed_us(cs.ev[2], cs.ev[3], rc);
    const double out_us = elapsed_us(cs.ev[4], cs.ev[5], rc);
    if (rc == IDGX_OK && timing)
    {
        timing->in_transfer_time = in_us / iterations;
        timing->out_transfer_time = out_us / iterations;
        timing->compute_time = compute_us / iterations;
        // bytes per microsecond = MB/s
        timing->in_bandwidth = in_us > 0 ? static_cast<double>(in_bytes) * iterations / in_us : -1.;
        timing->out_bandwidth = out_us > 0 ? static_cast<double>(bytes[out_buffer]) * iterations / out_us : -1.;
    }
    return rc;
}
} // namespace

extern "C" int idgx_host_grid(const idgx_plan *plan, const idgx_metadata *h_metadata, const idgx_channels *h_channels,
                              const float *h_uvw, const float *h_wavenumbers, const void *h_visibilities,
                              const float *h_spheroidal, const void *h_aterms, void *h_subgrids, int iterations,
                              idgx_timing *timing)
{
    // A roctx range over the whole call (it returns synchronised): the counterpart of the reference CPU launcher's
    // VTune task (gridding/src/degridding_benchmark.cpp:41-45,191-193), seen by rocprofv3 --marker-trace.
    roctxRangePushA("idgx_host_grid");
    const int rc = run_host(true, plan, h_metadata, h_channels, h_uvw, h_wavenumbers, h_visibilities, h_spheroidal,
                            h_aterms, h_subgrids, iterations, timing);
    roctxRangePop();
    return rc;
}

extern "C" int idgx_host_degrid(const idgx_plan *plan, const idgx_metadata *h_metadata,
                                const idgx_channels *h_channels, const float *h_uvw, const float *h_wavenumbers,
                                const void *h_subgrids, const float *h_spheroidal, const void *h_aterms,
                                void *h_visibilities, int iterations, idgx_timing *timing)
{
    roctxRangePushA("idgx_host_degrid");
    const int rc = run_host(false, plan, h_metadata, h_channels, h_uvw, h_wavenumbers, h_subgrids, h_spheroidal,
                            h_aterms, h_visibilities, iterations, timing);
    roctxRangePop();
    return rc;
}
