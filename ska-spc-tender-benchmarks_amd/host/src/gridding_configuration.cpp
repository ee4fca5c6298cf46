// Configuration reader for the gridding benchmark's flat YAML files (gridding/examples/*.yaml): scalar
// `key: value` lines and flow lists `key: [a, b]`, '#' comments.  Keys, expansion order (hardware types outer,
// operations inner) and error messages follow gridding/src/gridding_configuration.cpp:10-144.  Optional keys of
// this backend: image_size, w_step_in_lambda, precision.
#include "gridding_configuration.h"

#include "flat_yaml.hpp"

#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace gridding_benchmark
{
    namespace
    {
        using benchmarks_common::Yaml;
        using benchmarks_common::parse;

        size_t required_size(const Yaml &y, const std::string &key, const std::string &what)
        {
            const auto it = y.scalars.find(key);
            if (it == y.scalars.end() || it->second.empty())
                benchmarks_common::log_and_abort("Invalid " + key + " section. Should look like this:\n  nbatches: <" +
                                                 what + ">");
            try
            {
                size_t pos = 0;
                const unsigned long long v = std::stoull(it->second, &pos);
                if (pos != it->second.size())
                    throw std::invalid_argument(key);
                return static_cast<size_t>(v);
            }
            catch (const std::exception &)
            {
                throw std::runtime_error("bad value for " + key + ": " + it->second);
            }
        }
    } // namespace

    std::vector<configuration> read_configuration(std::filesystem::path path)
    {
        if (!std::filesystem::exists(path))
            throw std::runtime_error{"File " + path.string() + " not found."};
        std::ifstream file(path);
        const Yaml y = parse(file);

        const auto hw = y.lists.count("hardware_types") ? y.lists.at("hardware_types") : std::vector<std::string>{};
        if (hw.empty())
            benchmarks_common::log_and_abort("hardware_types section empty.");
        std::vector<benchmarks_common::hardware_type> hardware_types;
        for (const auto &h : hw)
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

        const auto ops = y.lists.count("operation_types") ? y.lists.at("operation_types") : std::vector<std::string>{};
        if (ops.empty())
            benchmarks_common::log_and_abort("operation_types section empty.");
        std::vector<operation_type> operation_types;
        for (const auto &o : ops)
        {
            const auto op = to_operation_type(o);
            if (op == operation_type::err)
                benchmarks_common::log_and_abort("Invalid operation type token " + o +
                                                 " was found. Must be either gridding or degridding.");
            operation_types.push_back(op);
        }

        configuration base{};
        base.grid_size = required_size(y, "grid_size", "grid size");
        base.subgrid_size = required_size(y, "subgrid_size", "subgrid size");
        base.niterations = required_size(y, "niterations", "niterations");
        base.nstations = required_size(y, "nstations", "number of stations");
        base.nchannels = required_size(y, "nchannels", "number of channels");
        base.ntimeslots = required_size(y, "ntimeslots", "number of time slots");
        base.ntimesteps_per_subgrid = required_size(y, "ntimesteps_per_subgrid", "number of time steps per subgrid");
        if (y.scalars.count("image_size"))
            base.image_size = std::stof(y.scalars.at("image_size"));
        if (y.scalars.count("w_step_in_lambda"))
            base.w_step_in_lambda = std::stof(y.scalars.at("w_step_in_lambda"));
        if (y.scalars.count("precision"))
            base.precision = y.scalars.at("precision");

        std::vector<configuration> configurations;
        for (const auto h : hardware_types)
            for (const auto o : operation_types)
            {
                configuration c = base;
                c.htype = h;
                c.operation = o;
                configurations.push_back(c);
            }
        return configurations;
    }
} // namespace gridding_benchmark
