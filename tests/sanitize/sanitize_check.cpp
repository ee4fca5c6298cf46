// Checking program of the sanitizer build (tests/sanitize/Makefile): exercises the host-side code under
// AddressSanitizer and UndefinedBehaviorSanitizer.
//   sanitize_check oracle            every oracle routine on reference-generator workloads of several shapes,
//                                    with channel windows, both A-term conventions, subgrids that overhang the grid
//                                    and an empty plan; prints "oracle: ok"
//   sanitize_check gridding <yaml>   the gridding driver's configuration reader; prints the configurations
//   sanitize_check fft <yaml>        the FFT driver's configuration reader
// A malformed file ends the program through the readers' own error path (exit status 255, as the reference's
// log_and_abort); any sanitizer report aborts it with a different status and its own message.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <vector>

#include "fft_configuration.h"
#include "gridding_configuration.h"
#include "oracle.h"
#include "reference_workload.h"

namespace
{
int fail(const char *what)
{
    std::fprintf(stderr, "sanitize_check: %s\n", what);
    return 1;
}

template <class T> bool finite(const std::vector<T> &v)
{
    for (const T x : v)
        if (!std::isfinite(static_cast<double>(x)))
            return false;
    return true;
}

int check_shape(int nstations, int ntimeslots, int T, int C, int S, int G, bool windows)
{
    idgx_reference_shape shape{nstations, ntimeslots, T, C, S, G};
    const int nbl = nstations * (nstations - 1) / 2, nsub = nbl * ntimeslots, rows = nsub * T;
    std::vector<float> uvw(3 * static_cast<size_t>(rows)), freq(C), wn(C), vis(8 * static_cast<size_t>(rows) * C);
    std::vector<uint32_t> baselines(2 * static_cast<size_t>(nbl));
    std::vector<float> sph(static_cast<size_t>(S) * S), aterms(8 * static_cast<size_t>(ntimeslots) * nstations * S * S);
    std::vector<idgx_metadata> md(nsub);
    if (idgx_reference_workload(&shape, uvw.data(), freq.data(), wn.data(), vis.data(), baselines.data(), sph.data(),
                                aterms.data(), md.data()) != 0)
        return fail("reference workload");
    std::vector<oracle_metadata> om(nsub);
    static_assert(sizeof(oracle_metadata) == sizeof(idgx_metadata), "36-byte plan records");
    std::memcpy(om.data(), md.data(), sizeof(idgx_metadata) * md.size());
    std::vector<oracle_channels> ch(nsub);
    for (int s = 0; s < nsub; ++s)
        ch[s] = {s % 2, C - (s % 2) - (s % 3 == 0 ? 1 : 0)}; // ragged windows inside [0, C)
    const oracle_channels *chp = windows ? ch.data() : nullptr;
    oracle_params p{nsub, S, G, C, nstations, 0.1F, windows ? 50.F : 0.F};
    const size_t nsubv = 8 * static_cast<size_t>(nsub) * S * S;
    std::vector<float> sub32(nsubv), vis32(vis.size(), 0.F);
    std::vector<double> sub64(nsubv), vis64(vis.size(), 0.0);
    if (oracle_grid_fp32(&p, om.data(), chp, uvw.data(), wn.data(), vis.data(), sph.data(), aterms.data(),
                         sub32.data(), 2) != 0 ||
        oracle_grid_fp64(&p, om.data(), chp, uvw.data(), wn.data(), vis.data(), sph.data(), aterms.data(),
                         sub64.data(), 2) != 0)
        return fail("gridder");
    for (const int conv : {ORACLE_ATERM_ADJOINT, ORACLE_ATERM_REFERENCE})
        if (oracle_degrid_fp32(&p, om.data(), chp, uvw.data(), wn.data(), sub32.data(), sph.data(), aterms.data(),
                               vis32.data(), conv, 2) != 0)
            return fail("degridder fp32");
    if (oracle_degrid_fp64(&p, om.data(), chp, uvw.data(), wn.data(), sub32.data(), sph.data(), aterms.data(),
                           vis64.data(), 2) != 0)
        return fail("degridder fp64");
    if (!finite(sub32) || !finite(sub64) || !finite(vis32) || !finite(vis64))
        return fail("non-finite oracle output");
    // Transform, add onto and split from a grid; the generator's corners can overhang the grid edge.
    if (oracle_dft2d_fp64(sub64.data(), 4 * nsub, S, -1, 2) != 0 || oracle_dft2d_fp64(sub64.data(), 4 * nsub, S, 1, 2))
        return fail("dft2d");
    std::vector<double> grid(8 * static_cast<size_t>(G) * G, 0.0), back(nsubv);
    if (oracle_adder_fp64(om.data(), nsub, S, G, sub32.data(), grid.data()) != 0 ||
        oracle_splitter_fp64(om.data(), nsub, S, G, grid.data(), back.data()) != 0)
        return fail("adder / splitter");
    // An empty plan touches nothing.
    oracle_params empty = p;
    empty.nr_subgrids = 0;
    if (oracle_grid_fp32(&empty, om.data(), nullptr, uvw.data(), wn.data(), vis.data(), sph.data(), aterms.data(),
                         sub32.data(), 1) != 0)
        return fail("empty plan");
    std::vector<float> ramp(nsubv);
    idgx_reference_subgrids(nsub, S, ramp.data());
    return finite(ramp) ? 0 : fail("subgrid ramp");
}
} // namespace

int main(int argc, char **argv)
{
    if (argc >= 2 && std::strcmp(argv[1], "oracle") == 0)
    {
        const int shapes[][7] = {{3, 2, 16, 8, 24, 512, 0}, {4, 1, 8, 4, 32, 256, 1}, {2, 3, 5, 3, 16, 64, 1},
                                 {5, 1, 4, 2, 48, 1024, 0}};
        for (const auto &sh : shapes)
            if (check_shape(sh[0], sh[1], sh[2], sh[3], sh[4], sh[5], sh[6] != 0) != 0)
                return 1;
        std::printf("oracle: ok\n");
        return 0;
    }
    if (argc == 3 && std::strcmp(argv[1], "gridding") == 0)
    {
        try
        {
            const auto cfgs = gridding_benchmark::read_configuration(argv[2]);
            for (const auto &c : cfgs)
                std::printf("gridding: %s G=%zu S=%zu stations=%zu channels=%zu slots=%zu T=%zu it=%zu %s\n",
                            gridding_benchmark::to_string(c.operation).c_str(), c.grid_size, c.subgrid_size,
                            c.nstations, c.nchannels, c.ntimeslots, c.ntimesteps_per_subgrid, c.niterations,
                            c.precision.c_str());
        }
        catch (const std::exception &e)
        {
            std::printf("gridding: error: %s\n", e.what());
            return 2;
        }
        return 0;
    }
    if (argc == 3 && std::strcmp(argv[1], "fft") == 0)
    {
        try
        {
            std::printf("fft: %zu configurations\n", fft_benchmark::read_configurations(argv[2]).size());
        }
        catch (const std::exception &e)
        {
            std::printf("fft: error: %s\n", e.what());
            return 2;
        }
        return 0;
    }
    std::fprintf(stderr, "usage: sanitize_check oracle | gridding <yaml> | fft <yaml>\n");
    return 64;
}
