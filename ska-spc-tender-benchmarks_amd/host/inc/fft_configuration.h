// Configuration of the FFT benchmark (the reference's fft/inc/fft_configuration.h:10-47): a batched 2-D complex
// transform of nx x ny points in single or double precision, forward or inverse, repeated niterations times over
// as many planes as fit in memorysize bytes.
#pragma once
#include <cstddef>
#include <filesystem>
#include <string>
#include <vector>

#include "benchmarks_common.h"

namespace fft_benchmark
{
    enum class float_type
    {
        single_precision,
        double_precision
    };

    // The exponent sign of the transform (heFFTe: forward = exp(-2 pi i ...), inverse = exp(+2 pi i ...)).
    enum class transform_type
    {
        forward = -1,
        inverse = 1
    };

    inline transform_type invert(transform_type t)
    {
        return t == transform_type::forward ? transform_type::inverse : transform_type::forward;
    }

    struct configuration
    {
        float_type ftype = float_type::single_precision;
        transform_type ttype = transform_type::forward;
        benchmarks_common::hardware_type htype = benchmarks_common::hardware_type::gpu;
        size_t nx = 0, ny = 0;
        size_t niterations = 1;
        size_t memorysize = 0;
    };

    struct benchmark_result
    {
        enum class status_t
        {
            failure,
            success
        };
        status_t status = status_t::failure;
        size_t batch_size = 0;
        size_t niterations = 0;
        double init_time = -1.;
        double in_transfer_time = -1.;
        double in_bandwidth = -1.;
        double out_bandwidth = -1.;
        double compute_time = -1.;
        double out_transfer_time = -1.;
        double max_error = -1.;
    };

    // Expansion order hardware types > float types > transform types > dimensions, as
    // fft/src/fft_configuration.cpp:142-166.
    std::vector<configuration> read_configurations(const std::filesystem::path &path);

    benchmark_result launch_benchmark(const configuration &configuration);
} // namespace fft_benchmark
