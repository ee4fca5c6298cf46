// C ABI of libidgx: argument validation and dispatch to the gfx950 kernels.
#include "idgx.h"
#include "kernels.hpp"

namespace
{
using idgx::KernelArgs;
using idgx::PlanDev;

constexpr int kMinSubgrid = 8;
constexpr int kMaxSubgrid = 128;

int check_plan_struct(const idgx_plan *p)
{
    if (!p)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (p->nr_subgrids < 0 || p->nr_channels <= 0 || p->nr_stations <= 0 || p->nr_aterm_slots <= 0 ||
        p->nr_rows < 0 || p->max_nr_timesteps < 0 || p->grid_size <= 0 || !(p->image_size > 0.F))
        return IDGX_ERR_INVALID_ARGUMENT;
    if (p->subgrid_size < kMinSubgrid || p->subgrid_size > kMaxSubgrid || (p->subgrid_size & 1))
        return IDGX_ERR_UNSUPPORTED;
    if (p->aterm_convention != IDGX_ATERM_ADJOINT && p->aterm_convention != IDGX_ATERM_REFERENCE)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (p->precision < IDGX_PLAIN_FP32 || p->precision > IDGX_MIXED_FP16)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (p->flags & ~IDGX_FLAG_ACCURATE_SINCOS)
        return IDGX_ERR_INVALID_ARGUMENT;
    return IDGX_OK;
}

PlanDev to_dev(const idgx_plan *p)
{
    PlanDev d;
    d.nr_subgrids = p->nr_subgrids;
    d.S = p->subgrid_size;
    d.G = p->grid_size;
    d.nr_channels = p->nr_channels;
    d.nr_stations = p->nr_stations;
    d.nr_aterm_slots = p->nr_aterm_slots;
    d.max_nr_timesteps = p->max_nr_timesteps;
    d.nr_rows = p->nr_rows;
    d.image_size = p->image_size;
    d.w_step = p->w_step_in_lambda;
    d.convention = p->aterm_convention;
    return d;
}

bool accurate_sincos(const idgx_plan *p)
{
    return (p->flags & IDGX_FLAG_ACCURATE_SINCOS) != 0;
}

int to_status(hipError_t e)
{
    if (e == hipSuccess)
        return IDGX_OK;
    if (e == hipErrorOutOfMemory)
        return IDGX_ERR_OUT_OF_MEMORY;
    if (e == hipErrorInvalidConfiguration)
        return IDGX_ERR_UNSUPPORTED;
    return IDGX_ERR_DEVICE;
}
} // namespace

extern "C" int idgx_check_plan(const idgx_plan *plan, const idgx_metadata *md, const idgx_channels *ch)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (plan->nr_subgrids > 0 && !md)
        return IDGX_ERR_INVALID_ARGUMENT;
    for (int s = 0; s < plan->nr_subgrids; ++s)
    {
        const idgx_metadata &m = md[s];
        const long long row0 = static_cast<long long>(m.baseline_offset) - md[0].baseline_offset + m.time_offset;
        const int c0 = ch ? ch[s].channel_begin : 0;
        const int nc = ch ? ch[s].nr_channels : plan->nr_channels;
        if (m.nr_timesteps < 0 || m.nr_timesteps > plan->max_nr_timesteps || row0 < 0 ||
            row0 + m.nr_timesteps > plan->nr_rows || c0 < 0 || nc < 0 || c0 + nc > plan->nr_channels ||
            m.station1 >= static_cast<uint32_t>(plan->nr_stations) ||
            m.station2 >= static_cast<uint32_t>(plan->nr_stations) || m.aterm_index < 0 ||
            m.aterm_index >= plan->nr_aterm_slots)
            return IDGX_ERR_INVALID_METADATA;
    }
    return IDGX_OK;
}

extern "C" int idgx_grid(const idgx_plan *plan, const idgx_metadata *d_metadata, const idgx_channels *d_channels,
                         const float *d_uvw, const float *d_wavenumbers, const void *d_visibilities,
                         const float *d_spheroidal, const void *d_aterms, void *d_subgrids, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (plan->nr_subgrids == 0)
        return IDGX_OK;
    if (!d_metadata || !d_uvw || !d_wavenumbers || !d_visibilities || !d_spheroidal || !d_aterms || !d_subgrids)
        return IDGX_ERR_INVALID_ARGUMENT;
    KernelArgs a{};
    a.p = to_dev(plan);
    a.metadata = d_metadata;
    a.channels = d_channels;
    a.uvw = d_uvw;
    a.wavenumbers = d_wavenumbers;
    a.visibilities = d_visibilities;
    a.spheroidal = d_spheroidal;
    a.aterms = d_aterms;
    a.subgrids_out = d_subgrids;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    switch (plan->precision)
    {
    case IDGX_PLAIN_FP32:
        return to_status(idgx::launch_grid_plain(a, accurate_sincos(plan), st));
    default:
        return to_status(idgx::launch_grid_mixed(a, plan->precision, st));
    }
}

extern "C" int idgx_degrid(const idgx_plan *plan, const idgx_metadata *d_metadata, const idgx_channels *d_channels,
                           const float *d_uvw, const float *d_wavenumbers, const void *d_subgrids,
                           const float *d_spheroidal, const void *d_aterms, void *d_visibilities, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (plan->nr_subgrids == 0)
        return IDGX_OK;
    if (!d_metadata || !d_uvw || !d_wavenumbers || !d_visibilities || !d_spheroidal || !d_aterms || !d_subgrids)
        return IDGX_ERR_INVALID_ARGUMENT;
    KernelArgs a{};
    a.p = to_dev(plan);
    a.metadata = d_metadata;
    a.channels = d_channels;
    a.uvw = d_uvw;
    a.wavenumbers = d_wavenumbers;
    a.subgrids = d_subgrids;
    a.spheroidal = d_spheroidal;
    a.aterms = d_aterms;
    a.visibilities_out = d_visibilities;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    switch (plan->precision)
    {
    case IDGX_PLAIN_FP32:
        return to_status(idgx::launch_degrid_plain(a, accurate_sincos(plan), st));
    default:
        return to_status(idgx::launch_degrid_mixed(a, plan->precision, st));
    }
}

extern "C" const char *idgx_strerror(int status)
{
    switch (status)
    {
    case IDGX_OK:
        return "ok";
    case IDGX_ERR_INVALID_ARGUMENT:
        return "invalid argument";
    case IDGX_ERR_INVALID_METADATA:
        return "subgrid metadata reads outside the arrays";
    case IDGX_ERR_UNSUPPORTED:
        return "unsupported request";
    case IDGX_ERR_DEVICE:
        return "HIP runtime error";
    case IDGX_ERR_OUT_OF_MEMORY:
        return "out of device memory";
    default:
        return "unknown status";
    }
}

extern "C" const char *idgx_version(void)
{
    return "idgx 0.1 gfx950";
}
