#include "params.hpp"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace spns_hit {

static std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

Params::Params(const std::string& path) : path_(path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Parameters file is missing!!! Include \"" + path + "\", please");
    std::string line;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ss(line);
        std::string key, value;
        if (!(ss >> key)) continue;
        ss >> value;
        key = lower(key);
        if (kv_.count(key)) throw std::runtime_error("GetParameter: double definition of parameter " + key + "!");
        kv_[key] = value;
    }
}

bool Params::has(const std::string& key) const { return kv_.count(lower(key)) != 0; }

std::string Params::word(const std::string& key, const std::string& def) const
{
    auto it = kv_.find(lower(key));
    return it == kv_.end() ? def : it->second;
}

long Params::integer(const std::string& key, long def) const
{
    auto it = kv_.find(lower(key));
    return it == kv_.end() ? def : std::strtol(it->second.c_str(), nullptr, 10);
}

double Params::real(const std::string& key, double def) const
{
    auto it = kv_.find(lower(key));
    return it == kv_.end() ? def : std::strtod(it->second.c_str(), nullptr);
}

long Params::need_int(const std::string& key) const
{
    if (!has(key)) throw std::runtime_error("Parameter " + key + " is missing in " + path_);
    return integer(key);
}

double Params::need_real(const std::string& key) const
{
    if (!has(key)) throw std::runtime_error("Parameter " + key + " is missing in " + path_);
    return real(key);
}

}  // namespace spns_hit
