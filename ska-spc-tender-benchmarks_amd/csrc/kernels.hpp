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

bool subgrid_fft_supported(int S);
hipError_t launch_subgrid_fft(void *subgrids, long long nr_subgrids, int S, bool forward, hipStream_t stream);
hipError_t launch_adder(const idgx_metadata *md, const int32_t *order, const int32_t *tile_start,
                        const int32_t *color_begin, int S, int G, const void *subgrids, void *grid, hipStream_t st);
hipError_t launch_splitter(const idgx_metadata *md, long long nr_subgrids, int S, int G, const void *grid,
                           void *subgrids, hipStream_t st);
} // namespace idgx
