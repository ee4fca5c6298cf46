// Internal kernel-launch interface of libidgx (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "idgx.h"
#include "idgx_device.hpp"

namespace idgx
{
// Arguments shared by the gridder and degridder kernels, passed by value.
struct KernelArgs
{
    PlanDev p;
    const idgx_metadata *metadata;
    const idgx_channels *channels; // may be null: every subgrid covers all channels
    const float *uvw;
    const float *wavenumbers;
    const void *visibilities; // gridder input
    void *visibilities_out;   // degridder output
    const float *spheroidal;
    const void *aterms;
    const void *subgrids; // degridder input
    void *subgrids_out;   // gridder output
};

hipError_t launch_grid_plain(const KernelArgs &a, bool accurate_sincos, hipStream_t stream);
hipError_t launch_degrid_plain(const KernelArgs &a, bool accurate_sincos, hipStream_t stream);
hipError_t launch_grid_mixed(const KernelArgs &a, int precision, hipStream_t stream);
hipError_t launch_degrid_mixed(const KernelArgs &a, int precision, hipStream_t stream);
} // namespace idgx
