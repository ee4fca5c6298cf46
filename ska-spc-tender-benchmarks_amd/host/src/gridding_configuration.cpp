// Configuration reader for the gridding benchmark's flat YAML files (gridding/examples/*.yaml): scalar
// `key: value` lines and flow lists `key: [a, b]`, '#' comments.  Keys, expansion order (hardware types outer,
// operations inner) and error messages follow gridding/src/gridding_configuration.cpp:10-144.  Optional keys of
// this backend: image_size, w_step_in_lambda, precision.
#include "gridding_configuration.h"

#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace gridding_benchmark
{
    namespace
    {
        std::string trim(const std::string &s)
        {
            const auto b = s.find_first_not_of(" \t\r\n");
            if (b == std::string::npos)
                return "";
            const auto e = s.find_last_not_of(" \t\r\n");
            return s.substr(b, e - b + 1);
        }

        std::string unquote(const std::string &s)
        {
            if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
                return s.substr(1, s.size() - 2);
            return s;
        }

        struct Yaml
        {
            std::map<std::string, std::string> scalars;
            std::map<std::string, std::vector<std::string>> lists;
        };

        Yaml parse(std::istream &in)
        {
            Yaml y;
            std::string line;
            int lineno = 0;
            while (std::getline(in, line))
            {
                ++lineno;
                const auto hash = line.find('#');
                if (hash != std::string::npos)
                    line = line.substr(0, hash);
                line = trim(line);
                if (line.empty() || line == "---")
                    continue;
                const auto colon = line.find(':');
                if (colon == std::string::npos)
                    throw std::runtime_error("line " + std::to_string(lineno) + ": expected 'key: value'");
                const std::string key = trim(line.substr(0, colon));
                const std::string value = trim(line.substr(colon + 1));
                if (!value.empty() && value.front() == '[')
                {
                    if (value.back() != ']')
                        throw std::runtime_error("line " + std::to_string(lineno) + ": unterminated list");
                    std::vector<std::string> items;
                    std::stringstream ss(value.substr(1, value.size() - 2));
                    std::string item;
                    while (std::getline(ss, item, ','))
                    {
                        item = unquote(trim(item));
                        if (!item.empty())
                            items.push_back(item);
                    }
                    y.lists[key] = items;
                }
                else
                {
                    y.scalars[key] = unquote(value);
                }
            }
            return y;
        }

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
