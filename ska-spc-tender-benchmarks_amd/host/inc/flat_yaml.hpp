// Reader for the benchmarks' flat YAML files (gridding/examples/*.yaml, fft/examples/*.yaml): scalar `key: value`
// lines and flow lists `key: [a, b]` (nested lists such as `[[a, b], [c, d]]` are kept as raw text for the caller),
// '#' comments.
#pragma once
#include <istream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace benchmarks_common
{
    inline std::string trim(const std::string &s)
    {
        const auto b = s.find_first_not_of(" \t\r\n");
        if (b == std::string::npos)
            return "";
        const auto e = s.find_last_not_of(" \t\r\n");
        return s.substr(b, e - b + 1);
    }

    inline std::string unquote(const std::string &s)
    {
        if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
            return s.substr(1, s.size() - 2);
        return s;
    }

    struct Yaml
    {
        std::map<std::string, std::string> scalars;
        std::map<std::string, std::vector<std::string>> lists;
        std::map<std::string, std::string> raw; // value text of every key
    };

    inline Yaml parse(std::istream &in)
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
            y.raw[key] = value;
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
} // namespace benchmarks_common
