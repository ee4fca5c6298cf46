// Reader of the FFT benchmark's YAML files (fft/examples/*.yaml).  Keys, accepted tokens, error messages and the
// expansion order follow fft/src/fft_configuration.cpp:10-166; `dimensions` is a list of [nx, ny] pairs.
#include "fft_configuration.h"

#include <fstream>
#include <stdexcept>

#include "flat_yaml.hpp"

namespace fft_benchmark
{
    namespace
    {
        size_t to_size(const std::string &text, const std::string &key)
        {
            size_t pos = 0;
            unsigned long long v = 0;
            try
            {
                v = std::stoull(text, &pos);
            }
            catch (const std::exception &)
            {
                pos = std::string::npos;
            }
            if (pos != text.size())
                throw std::runtime_error("bad value for " + key + ": " + text);
            return static_cast<size_t>(v);
        }

        // "[[a, b], [c, d]]" -> {(a, b), (c, d)}; anything else is an error.
        std::vector<std::pair<size_t, size_t>> parse_dimensions(const std::string &raw)
        {
            std::vector<std::pair<size_t, size_t>> dims;
            const std::string s = benchmarks_common::trim(raw);
            if (s.size() < 2 || s.front() != '[' || s.back() != ']')
                return dims;
            size_t i = 1;
            while (i < s.size() - 1)
            {
                const size_t open = s.find('[', i);
                if (open == std::string::npos || open >= s.size() - 1)
                    break;
                const size_t close = s.find(']', open);
                if (close == std::string::npos)
                    throw std::runtime_error("dimensions: unterminated pair");
                const std::string inner = s.substr(open + 1, close - open - 1);
                const size_t comma = inner.find(',');
                if (comma == std::string::npos)
                    throw std::runtime_error("dimensions: expected [nx, ny]");
                dims.emplace_back(to_size(benchmarks_common::trim(inner.substr(0, comma)), "dimensions"),
                                  to_size(benchmarks_common::trim(inner.substr(comma + 1)), "dimensions"));
                i = close + 1;
            }
            return dims;
        }
    } // namespace

    std::vector<configuration> read_configurations(const std::filesystem::path &path)
    {
        if (!std::filesystem::exists(path))
            throw std::runtime_error{"File " + path.string() + " not found."};
        std::ifstream file(path);
        const benchmarks_common::Yaml y = benchmarks_common::parse(file);

        const auto dims = parse_dimensions(y.raw.count("dimensions") ? y.raw.at("dimensions") : std::string{});
        if (dims.empty())
            benchmarks_common::log_and_abort("dimensions section empty.");

        const auto fts = y.lists.count("float_types") ? y.lists.at("float_types") : std::vector<std::string>{};
        if (fts.empty())
            benchmarks_common::log_and_abort("float_types section empty.");
        std::vector<float_type> float_types;
        for (const auto &f : fts)
        {
            if (f == "float")
                float_types.push_back(float_type::single_precision);
            else if (f == "double")
                float_types.push_back(float_type::double_precision);
            else
                benchmarks_common::log_and_abort("Invalid float_type token " + f +
                                                 " was found. Must be either float or double.");
        }

        const auto tts = y.lists.count("transform_types") ? y.lists.at("transform_types") : std::vector<std::string>{};
        if (tts.empty())
            benchmarks_common::log_and_abort("transform_types section empty.");
        std::vector<transform_type> transform_types;
        for (const auto &t : tts)
        {
            if (t == "forward")
                transform_types.push_back(transform_type::forward);
            else if (t == "inverse")
                transform_types.push_back(transform_type::inverse);
            else
                benchmarks_common::log_and_abort("Invalid transform_type token " + t +
                                                 " was found. Must be either forward or inverse.");
        }

        const auto hws = y.lists.count("hardware_types") ? y.lists.at("hardware_types") : std::vector<std::string>{};
        if (hws.empty())
            benchmarks_common::log_and_abort("hardware_types section empty.");
        std::vector<benchmarks_common::hardware_type> hardware_types;
        for (const auto &h : hws)
        {
            if (h == "cpu")
            {
                if (!benchmarks_common::is_hardware_type_enabled(benchmarks_common::hardware_type::cpu))
                    benchmarks_common::log_and_abort("CPU support is disabled");
                hardware_types.push_back(benchmarks_common::hardware_type::cpu);
            }
            else if (h == "gpu")
            {
                if (!benchmarks_common::is_hardware_type_enabled(benchmarks_common::hardware_type::gpu))
                    benchmarks_common::log_and_abort("GPU support is disabled");
                hardware_types.push_back(benchmarks_common::hardware_type::gpu);
            }
            else
            {
                benchmarks_common::log_and_abort("Invalid hardware_type token " + h +
                                                 " was found. Must be either cpu or gpu.");
            }
        }

        if (!y.scalars.count("niterations") || y.scalars.at("niterations").empty())
            benchmarks_common::log_and_abort("Invalid niterations section. Should look like this:\n"
                                             "niterations: <number of iterations to run>");
        const size_t niterations = to_size(y.scalars.at("niterations"), "niterations");
        if (!y.scalars.count("memorysize") || y.scalars.at("memorysize").empty())
            benchmarks_common::log_and_abort("Invalid memorysize section. Should look like this:\n"
                                             "memorysize: <memory size to use in bytes>");
        const size_t memorysize = to_size(y.scalars.at("memorysize"), "memorysize");

        std::vector<configuration> configurations;
        for (const auto h : hardware_types)
            for (const auto f : float_types)
                for (const auto t : transform_types)
                    for (const auto &d : dims)
                    {
                        configuration c;
                        c.htype = h;
                        c.ftype = f;
                        c.ttype = t;
                        c.nx = d.first;
                        c.ny = d.second;
                        c.niterations = niterations;
                        c.memorysize = memorysize;
                        configurations.push_back(c);
                    }
        return configurations;
    }
} // namespace fft_benchmark
