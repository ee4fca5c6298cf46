// Workload set-up and backend dispatch of the gridding benchmark (reference: launch_gridding,
// gridding/src/gridding_benchmark.cpp:184-228, and launch_degridding, gridding/src/degridding_benchmark.cpp:201-245).
// Inputs are the reference generator's, reproduced exactly by idgx_reference_workload (srand(0) and the same
// initialize_* order).  The degridder's input subgrids get the reference's initialize_subgrids ramp: the
// reference leaves them uninitialised (degridding_benchmark.cpp:221-230), which makes its output undefined.
#include <complex>

#include "degridding_benchmark.h"
#include "gridding_benchmark.h"
#include "reference_workload.h"

namespace gridding_benchmark
{
    namespace
    {
        struct Inputs
        {
            size_t nbl, nsub;
            Array2D<UVWCoordinate<float>> uvw;
            Array3D<Visibility<std::complex<float>>> visibilities;
            Array1D<Baseline> baselines;
            Array4D<Matrix2x2<std::complex<float>>> aterms;
            Array1D<float> frequencies, wavenumbers;
            Array2D<float> spheroidal;
            Array4D<std::complex<float>> subgrids;
            Array1D<Metadata> metadata;

            explicit Inputs(const configuration &c)
                : nbl(c.nstations * (c.nstations - 1) / 2), nsub(nbl * c.ntimeslots),
                  uvw(nsub, c.ntimesteps_per_subgrid), visibilities(nsub, c.ntimesteps_per_subgrid, c.nchannels),
                  baselines(nbl), aterms(c.ntimeslots, c.nstations, c.subgrid_size, c.subgrid_size),
                  frequencies(c.nchannels), wavenumbers(c.nchannels), spheroidal(c.subgrid_size, c.subgrid_size),
                  subgrids(nsub, 4, c.subgrid_size, c.subgrid_size), metadata(nsub)
            {
                idgx_reference_shape shape{static_cast<int32_t>(c.nstations), static_cast<int32_t>(c.ntimeslots),
                                           static_cast<int32_t>(c.ntimesteps_per_subgrid),
                                           static_cast<int32_t>(c.nchannels), static_cast<int32_t>(c.subgrid_size),
                                           static_cast<int32_t>(c.grid_size)};
                if (idgx_reference_workload(&shape, reinterpret_cast<float *>(uvw.data()), frequencies.data(),
                                            wavenumbers.data(), reinterpret_cast<float *>(visibilities.data()),
                                            reinterpret_cast<uint32_t *>(baselines.data()), spheroidal.data(),
                                            reinterpret_cast<float *>(aterms.data()),
                                            reinterpret_cast<idgx_metadata *>(metadata.data())) != 0)
                    benchmarks_common::log_and_abort("invalid workload shape (need nstations >= 2 and sizes >= 1)");
            }
        };

        template <class Launcher> benchmark_result dispatch(const configuration &c, Inputs &in)
        {
            return Launcher::launch(c, in.uvw, in.visibilities, in.baselines, in.aterms, in.frequencies,
                                    in.wavenumbers, in.spheroidal, in.subgrids, in.metadata);
        }
    } // namespace

    benchmark_result launch_gridding(const configuration &configuration)
    {
        Inputs in(configuration);
        if (configuration.htype == benchmarks_common::hardware_type::gpu)
            return dispatch<gridding_benchmark_launcher<benchmarks_common::hardware_type::gpu>>(configuration, in);
        benchmarks_common::log_and_abort("CPU support is disabled");
        return {};
    }

    benchmark_result launch_degridding(const configuration &configuration)
    {
        Inputs in(configuration);
        idgx_reference_subgrids(static_cast<int>(in.nsub), static_cast<int>(configuration.subgrid_size),
                                reinterpret_cast<float *>(in.subgrids.data()));
        if (configuration.htype == benchmarks_common::hardware_type::gpu)
            return dispatch<degridding_benchmark_launcher<benchmarks_common::hardware_type::gpu>>(configuration, in);
        benchmarks_common::log_and_abort("CPU support is disabled");
        return {};
    }
} // namespace gridding_benchmark
