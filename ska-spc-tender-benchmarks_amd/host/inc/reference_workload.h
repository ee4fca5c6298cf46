// The reference benchmark's synthetic workload (gridding/src/gridding_initialization.cpp), reproduced exactly.
#pragma once
#include <stdint.h>

#include "idgx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Shape keys of the reference's YAML configuration (gridding/src/gridding_configuration.cpp:237-284). */
typedef struct
{
    int32_t nstations;
    int32_t ntimeslots;
    int32_t ntimesteps_per_subgrid;
    int32_t nchannels;
    int32_t subgrid_size;
    int32_t grid_size;
} idgx_reference_shape;

/* Fills every input array the way launch_gridding does after srand(0) (gridding_benchmark.cpp:204-213).
 * Sizes: nsub = nstations*(nstations-1)/2 * ntimeslots; uvw float[nsub*T][3]; frequencies, wavenumbers
 * float[C]; visibilities complex64[nsub*T][C][4]; baselines uint32[nbl][2]; spheroidal float[S][S];
 * aterms complex64[ntimeslots][nstations][S][S][4]; metadata[nsub].  Returns 0, or 1 for a bad shape. */
int idgx_reference_workload(const idgx_reference_shape *shape, float *uvw, float *frequencies, float *wavenumbers,
                            float *visibilities, uint32_t *baselines, float *spheroidal, float *aterms,
                            idgx_metadata *metadata);

/* The reference's deterministic subgrid ramp (initialize_subgrids, gridding_initialization.cpp:184-206). */
void idgx_reference_subgrids(int nr_subgrids, int subgrid_size, float *subgrids);

#ifdef __cplusplus
}
#endif
