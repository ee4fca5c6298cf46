// Golden-vector driver for the reference CPU gridding path.  TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile (target `ref`) together with the reference's own CPU translation units,
// straight from /root/reference, into oracle/_ref/ref_driver.  It replays the input set-up of
// `launch_gridding` (gridding/src/gridding_benchmark.cpp:184-213): srand(0) then the initialize_* calls
// in the same order, runs the reference CPU gridder and degridder launchers once each, and dumps every
// array as raw little-endian binary so tests/golden/make_golden.py can turn them into fixtures.
//
// The degridder's input subgrids are never initialised by the reference (degridding_benchmark.cpp:221-230);
// here they are filled with the reference's own `initialize_subgrids` ramp
// (gridding/src/gridding_initialization.cpp:184-206) so the output is defined.
//
// The GPU launcher specialisations are referenced unconditionally by launch_gridding/launch_degridding, so
// this file provides stubs for them (they are never called: only the CPU launchers run here).

#include "degridding_benchmark.h"
#include "gridding_benchmark.h"
#include "gridding_initialization.h"

#include <cstdio>
#include <cstdlib>
#include <string>

namespace gb = gridding_benchmark;
using benchmarks_common::hardware_type;

namespace gridding_benchmark
{
    template <>
    benchmark_result gridding_benchmark_launcher<hardware_type::gpu>::launch(
        const configuration &, Array2D<UVWCoordinate<float>> &, Array3D<Visibility<std::complex<float>>> &,
        Array1D<Baseline> &, Array4D<Matrix2x2<std::complex<float>>> &, Array1D<float> &, Array1D<float> &,
        Array2D<float> &, Array4D<std::complex<float>> &, Array1D<Metadata> &)
    {
        std::fprintf(stderr, "ref_driver: GPU launcher is not available\n");
        std::abort();
    }

    template <>
    benchmark_result degridding_benchmark_launcher<hardware_type::gpu>::launch(
        const configuration &, Array2D<UVWCoordinate<float>> &, Array3D<Visibility<std::complex<float>>> &,
        Array1D<Baseline> &, Array4D<Matrix2x2<std::complex<float>>> &, Array1D<float> &, Array1D<float> &,
        Array2D<float> &, Array4D<std::complex<float>> &, Array1D<Metadata> &)
    {
        std::fprintf(stderr, "ref_driver: GPU launcher is not available\n");
        std::abort();
    }
} // namespace gridding_benchmark

// out_dir "-" times the reference launchers only: nothing is written, one line with their compute times (the
// benchmark_result.compute_time of gridding/inc/common_gridding.h:5-12, microseconds) goes to stdout.
static void dump(const std::string &dir, const char *name, const void *data, size_t bytes)
{
    if (dir == "-")
        return;
    const std::string path = dir + "/" + name;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(data, 1, bytes, f) != bytes)
    {
        std::fprintf(stderr, "ref_driver: cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

struct Inputs
{
    gb::Array2D<gb::UVWCoordinate<float>> uvw;
    gb::Array3D<gb::Visibility<std::complex<float>>> visibilities;
    gb::Array1D<gb::Baseline> baselines;
    gb::Array4D<gb::Matrix2x2<std::complex<float>>> aterms;
    gb::Array1D<float> frequencies;
    gb::Array1D<float> wavenumbers;
    gb::Array2D<float> spheroidal;
    gb::Array4D<std::complex<float>> subgrids;
    gb::Array1D<gb::Metadata> metadata;

    explicit Inputs(const gb::configuration &c)
        : uvw(nsub(c), c.ntimesteps_per_subgrid), visibilities(nsub(c), c.ntimesteps_per_subgrid, c.nchannels),
          baselines(nbl(c)), aterms(c.ntimeslots, c.nstations, c.subgrid_size, c.subgrid_size),
          frequencies(c.nchannels), wavenumbers(c.nchannels), spheroidal(c.subgrid_size, c.subgrid_size),
          subgrids(nsub(c), 4, c.subgrid_size, c.subgrid_size), metadata(nsub(c))
    {
        // Same call order as launch_gridding (gridding_benchmark.cpp:204-213); the reference hard-codes
        // image_size = 0.1F there (:188).
        const float image_size = 0.1F;
        srand(0);
        gb::initialize_uvw(c.grid_size, uvw);
        gb::initialize_frequencies(frequencies);
        gb::initialize_wavenumbers(frequencies, wavenumbers);
        gb::initialize_visibilities(c.grid_size, image_size, frequencies, uvw, visibilities);
        gb::initialize_baselines(c.nstations, baselines);
        gb::initialize_spheroidal(spheroidal);
        gb::initialize_aterms(spheroidal, aterms);
        gb::initialize_metadata(c.grid_size, c.ntimeslots, c.ntimesteps_per_subgrid, baselines, metadata);
    }

    static size_t nbl(const gb::configuration &c)
    {
        return (c.nstations * (c.nstations - 1)) / 2;
    }
    static size_t nsub(const gb::configuration &c)
    {
        return nbl(c) * c.ntimeslots;
    }
};

int main(int argc, char **argv)
{
    if (argc != 8)
    {
        std::fprintf(stderr, "usage: ref_driver <out_dir> nstations ntimeslots ntimesteps nchannels subgrid grid\n");
        return 1;
    }
    const std::string dir = argv[1];
    gb::configuration c{};
    c.htype = hardware_type::cpu;
    c.niterations = 1;
    c.nstations = std::stoul(argv[2]);
    c.ntimeslots = std::stoul(argv[3]);
    c.ntimesteps_per_subgrid = std::stoul(argv[4]);
    c.nchannels = std::stoul(argv[5]);
    c.subgrid_size = std::stoul(argv[6]);
    c.grid_size = std::stoul(argv[7]);

    double grid_us = 0.0, degrid_us = 0.0;
    // Gridder: inputs as launch_gridding builds them, output subgrids.
    {
        Inputs in(c);
        c.operation = gb::operation_type::gridding;
        grid_us = gb::gridding_benchmark_launcher<hardware_type::cpu>::launch(
                      c, in.uvw, in.visibilities, in.baselines, in.aterms, in.frequencies, in.wavenumbers,
                      in.spheroidal, in.subgrids, in.metadata)
                      .compute_time;
        dump(dir, "uvw.bin", in.uvw.data(), in.uvw.size() * sizeof(in.uvw[0]));
        dump(dir, "frequencies.bin", in.frequencies.data(), in.frequencies.size() * sizeof(float));
        dump(dir, "wavenumbers.bin", in.wavenumbers.data(), in.wavenumbers.size() * sizeof(float));
        dump(dir, "visibilities.bin", in.visibilities.data(),
             in.visibilities.size() * sizeof(in.visibilities[0]));
        dump(dir, "baselines.bin", in.baselines.data(), in.baselines.size() * sizeof(in.baselines[0]));
        dump(dir, "spheroidal.bin", in.spheroidal.data(), in.spheroidal.size() * sizeof(float));
        dump(dir, "aterms.bin", in.aterms.data(), in.aterms.size() * sizeof(in.aterms[0]));
        dump(dir, "metadata.bin", in.metadata.data(), in.metadata.size() * sizeof(in.metadata[0]));
        dump(dir, "grid_subgrids.bin", in.subgrids.data(), in.subgrids.size() * sizeof(in.subgrids[0]));
    }

    // Degridder: same inputs, subgrids set by initialize_subgrids, output visibilities.
    {
        Inputs in(c);
        gb::initialize_subgrids(in.subgrids);
        dump(dir, "degrid_subgrids.bin", in.subgrids.data(), in.subgrids.size() * sizeof(in.subgrids[0]));
        c.operation = gb::operation_type::degridding;
        degrid_us = gb::degridding_benchmark_launcher<hardware_type::cpu>::launch(
                        c, in.uvw, in.visibilities, in.baselines, in.aterms, in.frequencies, in.wavenumbers,
                        in.spheroidal, in.subgrids, in.metadata)
                        .compute_time;
        dump(dir, "degrid_visibilities.bin", in.visibilities.data(),
             in.visibilities.size() * sizeof(in.visibilities[0]));
    }
    if (dir == "-")
        std::printf("{\"gridder_us\": %.3f, \"degridder_us\": %.3f}\n", grid_us, degrid_us);
    return 0;
}
