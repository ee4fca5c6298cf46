// gridding-benchmarks driver: reads a YAML configuration file, runs a warm-up gridding pass, runs every
// configuration, prints the result table and writes it as ';'-separated CSV to the output file name in the
// working directory (reference: gridding/src/main.cpp:15-163).  The first 12 columns are the reference's; the
// HIP backend appends the precision form and the throughput in million visibilities per second.
#include <chrono>
#include <cmath>
#include <iostream>
#include <string>

#include "benchmarks_common.h"
#include "degridding_benchmark.h"
#include "gridding_benchmark.h"
#include "gridding_configuration.h"

namespace
{
    using clk = std::chrono::steady_clock;

    std::string result_string(double x)
    {
        return x < 0. ? "N/A" : std::to_string(x);
    }

    void run(const std::vector<gridding_benchmark::configuration> &configurations, const std::filesystem::path &output)
    {
        std::cout << "##################################################\n";
        std::cout << "#### Gridding/degridding benchmarking utility ####\n";
        std::cout << "##################################################\n\n";
        std::cout << "##################################################\n";
        std::cout << "## Benchmarking configurations\n\n";

        std::vector<std::array<std::string, 7>> config_table;
        config_table.push_back({"Configuration ID", "  Operation", "  Grid size", "  Subgrid size",
                                "  Number of channels", "  Number of stations", "  Hardware"});
        int id = 0;
        for (const auto &c : configurations)
            config_table.push_back({std::to_string(id++), to_string(c.operation), std::to_string(c.grid_size),
                                    std::to_string(c.subgrid_size), std::to_string(c.nchannels),
                                    std::to_string(c.nstations), benchmarks_common::hardware_type_string(c.htype)});
        benchmarks_common::print_columns(config_table);
        std::cout << std::endl;

        std::cout << "##############\n## Warmup step\n\nWarming up...";
        const auto w0 = clk::now();
        gridding_benchmark::launch_gridding(configurations.front());
        const auto warmup_s = std::chrono::duration_cast<std::chrono::seconds>(clk::now() - w0).count();
        std::cout << " Done in " << warmup_s << " seconds.\n";
        std::cout << "==> Total benchmark should take a bit more than "
                  << std::round(static_cast<double>(configurations.size()) * warmup_s) << " seconds." << std::endl
                  << std::endl;

        std::cout << "###############\n## Benchmarking\n\n";
        std::vector<gridding_benchmark::benchmark_result> results;
        const auto b0 = clk::now();
        for (size_t i = 0; i < configurations.size(); ++i)
        {
            std::cout << "Running configuration " << i << "... " << std::flush;
            const auto t0 = clk::now();
            const auto &c = configurations[i];
            results.push_back(c.operation == gridding_benchmark::operation_type::gridding
                                  ? gridding_benchmark::launch_gridding(c)
                                  : gridding_benchmark::launch_degridding(c));
            std::cout << "Done in " << std::chrono::duration_cast<std::chrono::seconds>(clk::now() - t0).count()
                      << " seconds." << std::endl;
        }
        std::cout << "==> Total benchmarking time: "
                  << std::chrono::duration_cast<std::chrono::seconds>(clk::now() - b0).count() << " seconds."
                  << std::endl
                  << std::endl;

        std::vector<std::array<std::string, 14>> table;
        table.push_back({"  Configuration ID", "  Operation", "  Iterations", "  Grid size", "  Subgrid size",
                         "  Number of channels", "  Number of stations", "  Mean input transfer time (us)",
                         "  Achieved input bandwidth (MiB/s)", "  Mean compute time (us)",
                         "  Mean output transfer time (us)", "  Achieved output bandwidth (MiB/s)", "  Precision",
                         "  Throughput (Mvis/s)"});
        for (size_t i = 0; i < results.size(); ++i)
        {
            const auto &c = configurations[i];
            const auto &r = results[i];
            const double nvis = static_cast<double>(c.nstations * (c.nstations - 1) / 2) * c.ntimeslots *
                                c.ntimesteps_per_subgrid * c.nchannels;
            table.push_back({std::to_string(i), to_string(c.operation), std::to_string(c.niterations),
                             std::to_string(c.grid_size), std::to_string(c.subgrid_size), std::to_string(c.nchannels),
                             std::to_string(c.nstations), result_string(r.in_transfer_time),
                             result_string(r.in_bandwidth), result_string(r.compute_time),
                             result_string(r.out_transfer_time), result_string(r.out_bandwidth), c.precision,
                             result_string(r.compute_time > 0 ? nvis / r.compute_time : -1.)});
        }
        benchmarks_common::print_columns(table);
        benchmarks_common::write_csv(output, table);
    }
} // namespace

int main(int argc, char **argv)
{
    if (argc < 3)
    {
        std::cerr << "gridding-benchmarks <configuration file> <output file>" << std::endl;
        return -1;
    }
    std::vector<gridding_benchmark::configuration> configurations;
    try
    {
        configurations = gridding_benchmark::read_configuration(argv[1]);
    }
    catch (const std::exception &e)
    {
        std::cerr << "Error while reading configurations file:\n" << e.what() << std::endl;
        return -1;
    }
    run(configurations, argv[2]);
    return 0;
}
