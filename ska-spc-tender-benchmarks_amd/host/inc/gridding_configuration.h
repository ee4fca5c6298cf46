// Benchmark configuration: the reference's flat YAML schema (gridding/src/gridding_configuration.cpp:10-144) plus
// optional keys of this backend.  read_configuration() expands hardware_types x operation_types into one
// configuration each, in that order.
#pragma once
#include <array>
#include <cstddef>
#include <filesystem>
#include <string>
#include <vector>

#include "benchmarks_common.h"

namespace gridding_benchmark
{
    enum class operation_type
    {
        gridding,
        degridding,
        err
    };

    inline operation_type to_operation_type(const std::string &s)
    {
        return s == "gridding" ? operation_type::gridding
                               : s == "degridding" ? operation_type::degridding : operation_type::err;
    }

    inline std::string to_string(operation_type o)
    {
        return o == operation_type::gridding ? "gridding" : o == operation_type::degridding ? "degridding" : "err";
    }

    struct configuration
    {
        benchmarks_common::hardware_type htype;
        operation_type operation;
        size_t niterations;
        size_t grid_size;
        size_t subgrid_size;
        size_t nstations;
        size_t nchannels;
        size_t ntimeslots;
        size_t ntimesteps_per_subgrid;
        // Optional keys (defaults are the reference's hard-coded values, gridding_benchmark.cpp:32-33).
        float image_size = 0.1F;
        float w_step_in_lambda = 0.F;
        std::string precision = "fp32"; // fp32 | bf16x3 | fp16x3 | bf16 | fp16
    };

    std::vector<configuration> read_configuration(std::filesystem::path path);
} // namespace gridding_benchmark
