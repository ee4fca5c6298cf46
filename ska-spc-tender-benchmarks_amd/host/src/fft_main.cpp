// fft-benchmarks driver: reads a YAML configuration file, runs a warm-up transform, runs every configuration,
// prints the result table and writes it as ';'-separated CSV to the output file name in the working directory
// (reference: fft/src/main.cpp:20-181).  The first 12 columns are the reference's; this backend appends the
// compute rate (5 N log2 N flops per N-point transform) and the transform's bandwidth (one read and one write of
// the plane).
#include <chrono>
#include <cmath>
#include <iostream>
#include <string>

#include "benchmarks_common.h"
#include "fft_configuration.h"

namespace
{
    using clk = std::chrono::steady_clock;

    std::string result_string(double x)
    {
        return x < 0. ? "N/A" : std::to_string(x);
    }

    void run(const std::vector<fft_benchmark::configuration> &configurations, const std::filesystem::path &output)
    {
        std::cout << "##################################\n";
        std::cout << "#### FFT benchmarking utility ####\n";
        std::cout << "##################################\n\n";
        std::cout << "##############################\n";
        std::cout << "## Benchmarking configurations\n\n";
        std::vector<std::array<std::string, 6>> config_table;
        config_table.push_back({"Configuration ID", "  Iterations", "  Dimensions", "  Max memory size",
                                "  Precision", "  Hardware"});
        int id = 0;
        for (const auto &c : configurations)
            config_table.push_back({std::to_string(id++), std::to_string(c.niterations),
                                    std::to_string(c.nx) + " * " + std::to_string(c.ny),
                                    std::to_string(c.memorysize) + " B",
                                    c.ftype == fft_benchmark::float_type::single_precision ? "single" : "double",
                                    benchmarks_common::hardware_type_string(c.htype)});
        benchmarks_common::print_columns(config_table);
        std::cout << std::endl;

        std::cout << "##############\n## Warmup step\n\nWarming up...";
        const auto w0 = clk::now();
        fft_benchmark::launch_benchmark(configurations.front());
        const auto warmup_s = std::chrono::duration_cast<std::chrono::seconds>(clk::now() - w0).count();
        std::cout << " Done in " << warmup_s << " seconds.\n";
        std::cout << "==> Total benchmark should take a bit more than "
                  << std::round(static_cast<double>(configurations.size()) * warmup_s) << " seconds." << std::endl
                  << std::endl;

        std::cout << "###############\n## Benchmarking\n\n";
        std::vector<fft_benchmark::benchmark_result> results;
        const auto b0 = clk::now();
        for (size_t i = 0; i < configurations.size(); ++i)
        {
            std::cout << "Running configuration " << i << "... " << std::flush;
            const auto t0 = clk::now();
            results.push_back(fft_benchmark::launch_benchmark(configurations[i]));
            std::cout << "Done in " << std::chrono::duration_cast<std::chrono::seconds>(clk::now() - t0).count()
                      << " seconds." << std::endl;
        }
        std::cout << "==> Total benchmarking time: "
                  << std::chrono::duration_cast<std::chrono::seconds>(clk::now() - b0).count() << " seconds."
                  << std::endl
                  << std::endl;

        std::vector<std::array<std::string, 14>> table;
        table.push_back({"  Configuration ID", "  Dimensions", "  Status", "  Max error", "  Number of iterations",
                         "  Batch size", "  Init time (us)", "  Mean input transfer time (us)",
                         "  Achieved input bandwidth (MiB/s)", "  Mean compute time (us)",
                         "  Mean output transfer time (us)", "  Achieved output bandwidth (MiB/s)",
                         "  Compute rate (GFLOP/s)", "  Transform bandwidth (GB/s)"});
        for (size_t i = 0; i < results.size(); ++i)
        {
            const auto &r = results[i];
            const auto &c = configurations[i];
            const bool ok = r.status == fft_benchmark::benchmark_result::status_t::success;
            const double n = static_cast<double>(c.nx) * static_cast<double>(c.ny);
            const double bytes = 2.0 * n * (c.ftype == fft_benchmark::float_type::single_precision ? 8.0 : 16.0);
            const double gflops = ok && r.compute_time > 0 ? 5.0 * n * std::log2(n) / r.compute_time / 1e3 : -1.;
            const double gbps = ok && r.compute_time > 0 ? bytes / r.compute_time / 1e3 : -1.;
            table.push_back({std::to_string(i), std::to_string(c.nx) + " * " + std::to_string(c.ny),
                             ok ? "OK" : "ERROR", result_string(r.max_error), std::to_string(r.niterations),
                             std::to_string(r.batch_size), result_string(r.init_time),
                             result_string(r.in_transfer_time), result_string(r.in_bandwidth),
                             result_string(r.compute_time), result_string(r.out_transfer_time),
                             result_string(r.out_bandwidth), result_string(gflops), result_string(gbps)});
        }
        benchmarks_common::print_columns(table);
        benchmarks_common::write_csv(output, table);
    }
} // namespace

int main(int argc, char **argv)
{
    if (argc < 3)
    {
        std::cerr << "fft-benchmarks <configuration file> <output file>" << std::endl;
        return -1;
    }
    std::vector<fft_benchmark::configuration> configurations;
    try
    {
        configurations = fft_benchmark::read_configurations(argv[1]);
    }
    catch (const std::exception &e)
    {
        std::cerr << "Error while reading configurations file:\n" << e.what() << std::endl;
        return -1;
    }
    run(configurations, argv[2]);
    return 0;
}
