// Shared helpers of the benchmark driver: hardware selection, table printing and the ';'-separated CSV writer.
// Same behaviour as the reference's benchmarks-common/inc/benchmarks_common.h:14-96 (right-aligned columns, a
// trailing ';' on every CSV row, the CSV written to the file name in the working directory, errors printed as
// "Error: <message>" before exiting with the status code).
#pragma once
#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <string_view>
#include <vector>

namespace benchmarks_common
{
    enum class hardware_type
    {
        cpu,
        gpu,
        err
    };

    inline std::string hardware_type_string(hardware_type h)
    {
        return h == hardware_type::cpu ? "cpu" : h == hardware_type::gpu ? "gpu" : "";
    }

    inline bool is_hardware_type_enabled(hardware_type h)
    {
#if defined(ENABLE_CPU) && ENABLE_CPU
        if (h == hardware_type::cpu)
            return true;
#endif
#if defined(ENABLE_GPU) && ENABLE_GPU
        if (h == hardware_type::gpu)
            return true;
#endif
        (void)h;
        return false;
    }

    template <size_t N> void print_columns(const std::vector<std::array<std::string, N>> &rows)
    {
        std::array<size_t, N> width{};
        for (const auto &r : rows)
            for (size_t i = 0; i < N; ++i)
                width[i] = std::max(width[i], r[i].size());
        for (const auto &r : rows)
        {
            for (size_t i = 0; i < N; ++i)
                std::cout << std::right << std::setw(static_cast<int>(width[i])) << r[i] << " ";
            std::cout << "\n";
        }
        std::cout.flush();
    }

    template <size_t N>
    void write_csv(const std::filesystem::path &path, const std::vector<std::array<std::string, N>> &rows)
    {
        std::ofstream out(path.filename());
        for (const auto &r : rows)
        {
            for (const auto &cell : r)
                out << cell << ";";
            out << "\n";
        }
    }

    inline void log_and_abort(std::string_view message, int status = -1)
    {
        std::cerr << "Error: " << message << "\n";
        std::exit(status);
    }
} // namespace benchmarks_common
