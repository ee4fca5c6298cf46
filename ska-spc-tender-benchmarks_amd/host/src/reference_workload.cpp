// The reference benchmark's synthetic workload, reproduced value for value.
//
// launch_gridding / launch_degridding (gridding/src/gridding_benchmark.cpp:184-213,
// gridding/src/degridding_benchmark.cpp:201-230) seed glibc's rand() with srand(0) and then call the
// initialize_* routines of gridding/src/gridding_initialization.cpp in a fixed order.  The drop-in driver must
// feed the HIP backend exactly those inputs, so this file evaluates the same expressions with the same float /
// double promotions and the same rand() call order; tests/test_host.py pins every array against the golden
// vectors produced from the compiled reference.  Built with -ffp-contract=off.
#include "reference_workload.h"

#include <cmath>
#include <complex>
#include <cstdlib>

namespace
{
using cf = std::complex<float>;

// rand_normalized_float (gridding_initialization.cpp:7-10): float(rand()) / float(RAND_MAX).
float rand_unit_float()
{
    return static_cast<float>(rand()) / static_cast<float>(RAND_MAX);
}

// (double)rand() / (double)RAND_MAX, as used by initialize_aterms and initialize_metadata.
double rand_unit_double()
{
    return static_cast<double>(rand()) / static_cast<double>(RAND_MAX);
}
} // namespace

extern "C" int idgx_reference_workload(const idgx_reference_shape *shape, float *uvw, float *frequencies,
                                       float *wavenumbers, float *visibilities, uint32_t *baselines,
                                       float *spheroidal, float *aterms, idgx_metadata *metadata)
{
    if (!shape || shape->nstations < 2 || shape->ntimeslots < 1 || shape->ntimesteps_per_subgrid < 1 ||
        shape->nchannels < 1 || shape->subgrid_size < 1 || shape->grid_size < 1)
        return 1;
    const size_t nst = shape->nstations;
    const size_t nbl = nst * (nst - 1) / 2;
    const size_t nts = shape->ntimeslots;
    const size_t nsub = nbl * nts;
    const size_t T = shape->ntimesteps_per_subgrid;
    const size_t C = shape->nchannels;
    const size_t S = shape->subgrid_size;
    const unsigned G = shape->grid_size;
    const float image_size = 0.1F; // hard-coded by the reference (gridding_benchmark.cpp:188)

    srand(0);

    // initialize_uvw (:12-36): one random ellipse per subgrid row, w = 0.
    for (size_t bl = 0; bl < nsub; ++bl)
    {
        const float half_grid = 0.5F * static_cast<float>(G);
        const float radius_u = half_grid + rand_unit_float() * half_grid;
        const float radius_v = half_grid + rand_unit_float() * half_grid;
        const float angle_step = 1.F / static_cast<float>(T);
        for (size_t t = 0; t < T; ++t)
        {
            const float angle = static_cast<float>((static_cast<float>(t) + 0.5) * angle_step);
            float *o = uvw + 3 * (bl * T + t);
            o[0] = static_cast<float>(radius_u * std::cos(angle * M_PI));
            o[1] = static_cast<float>(radius_v * std::sin(angle * M_PI));
            o[2] = 0.F;
        }
    }

    // initialize_frequencies (:38-49): a size_t start plus a float step, evaluated in float.
    {
        const size_t start = 150e6;
        const float step = 0.7e6;
        for (size_t i = 0; i < C; ++i)
        {
            const float f = static_cast<float>(start) + step * static_cast<float>(i);
            frequencies[i] = static_cast<float>(static_cast<double>(f));
        }
    }

    // initialize_wavenumbers (:51-60): 2*pi*f/c in double.
    for (size_t i = 0; i < C; ++i)
        wavenumbers[i] = static_cast<float>(2 * M_PI * static_cast<double>(frequencies[i]) / 299792458.0);

    // initialize_visibilities (:62-94): a point source at (0.6 G, 0.7 G) pixels, scaled per correlation.
    {
        const float x_offset = static_cast<float>(0.6 * G);
        const float y_offset = static_cast<float>(0.7 * G);
        const float amplitude = 1.0F;
        const float l = x_offset * image_size / static_cast<float>(G);
        const float m = y_offset * image_size / static_cast<float>(G);
        const double c_light = 299792458.0;
        const float scale[4] = {1.01F, 1.02F, 1.03F, 1.04F};
        for (size_t bl = 0; bl < nsub; ++bl)
            for (size_t t = 0; t < T; ++t)
                for (size_t c = 0; c < C; ++c)
                {
                    const float *uv = uvw + 3 * (bl * T + t);
                    const float u = static_cast<float>(static_cast<double>(frequencies[c]) / c_light * uv[0]);
                    const float v = static_cast<float>(static_cast<double>(frequencies[c]) / c_light * uv[1]);
                    const float arg = static_cast<float>(-2 * M_PI * static_cast<double>(u * l + v * m));
                    const cf value = amplitude * std::exp(cf(0.F, arg));
                    cf *o = reinterpret_cast<cf *>(visibilities) + ((bl * T + t) * C + c) * 4;
                    for (int pol = 0; pol < 4; ++pol)
                        o[pol] = value * scale[pol];
                }
    }

    // initialize_baselines (:96-113): station pairs (s1 < s2) in lexicographic order.
    {
        size_t bl = 0;
        for (unsigned s1 = 0; s1 < nst; ++s1)
            for (unsigned s2 = s1 + 1; s2 < nst && bl < nbl; ++s2, ++bl)
            {
                baselines[2 * bl] = s1;
                baselines[2 * bl + 1] = s2;
            }
    }

    // initialize_spheroidal (:115-128): |-1 + 2y/S| * |-1 + 2x/S| (a tent, zero at the centre).
    for (size_t y = 0; y < S; ++y)
    {
        const float ty = std::fabs(-1 + static_cast<float>(y) * 2.0F / static_cast<float>(S));
        for (size_t x = 0; x < S; ++x)
        {
            const float tx = std::fabs(-1 + static_cast<float>(x) * 2.0F / static_cast<float>(S));
            spheroidal[y * S + x] = ty * tx;
        }
    }

    // initialize_aterms (:130-156): one rand() per (timeslot, station, y, x).
    for (size_t ts = 0; ts < nts; ++ts)
        for (size_t st = 0; st < nst; ++st)
            for (size_t y = 0; y < S; ++y)
                for (size_t x = 0; x < S; ++x)
                {
                    const float scale = static_cast<float>(0.8 + rand_unit_double() * 0.4);
                    const float value = spheroidal[y * S + x] * scale;
                    const cf diag(static_cast<float>(value + 0.1), static_cast<float>(-0.1));
                    const cf off(static_cast<float>(value - 0.2), static_cast<float>(0.1));
                    cf *o = reinterpret_cast<cf *>(aterms) + (((ts * nst + st) * S + y) * S + x) * 4;
                    o[0] = diag;
                    o[1] = off;
                    o[2] = off;
                    o[3] = diag;
                }

    // initialize_metadata (:158-182): two rand() per subgrid for its corner.
    for (size_t bl = 0; bl < nbl; ++bl)
        for (size_t ts = 0; ts < nts; ++ts)
        {
            idgx_metadata &md = metadata[bl * nts + ts];
            md.baseline_offset = 0;
            md.time_offset = static_cast<int32_t>(bl * nts * T + ts * T);
            md.nr_timesteps = static_cast<int32_t>(T);
            md.aterm_index = 0;
            md.station1 = baselines[2 * bl];
            md.station2 = baselines[2 * bl + 1];
            md.x = static_cast<int>(rand_unit_double() * G);
            md.y = static_cast<int>(rand_unit_double() * G);
            md.z = 0;
        }
    return 0;
}

extern "C" void idgx_reference_subgrids(int nr_subgrids, int subgrid_size, float *subgrids)
{
    // initialize_subgrids (:184-206): ((y*S + x + 1) / (100 S^2), corr / 10).
    const size_t S = subgrid_size;
    cf *o = reinterpret_cast<cf *>(subgrids);
    for (size_t s = 0; s < static_cast<size_t>(nr_subgrids); ++s)
        for (unsigned c = 0; c < 4; ++c)
            for (unsigned y = 0; y < S; ++y)
                for (unsigned x = 0; x < S; ++x)
                    o[((s * 4 + c) * S + y) * S + x] =
                        cf((y * S + x + 1) / (static_cast<float>(100) * S * S), c / 10.0F);
}
