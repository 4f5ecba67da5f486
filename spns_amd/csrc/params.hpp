// Parameters file reader with the reference's syntax (src/parser.cpp:90-102,
// 311-366, 451-482): one "key value" pair per line, keys case-insensitive,
// '#' starts a comment, unknown keys ignored, a key defined twice is an error.
#pragma once
#include <map>
#include <stdexcept>
#include <string>

namespace spns_hit {

class Params {
public:
    explicit Params(const std::string& path);
    bool has(const std::string& key) const;
    std::string word(const std::string& key, const std::string& def = "") const;
    long integer(const std::string& key, long def = 0) const;
    double real(const std::string& key, double def = 0.0) const;
    // required keys: throw with the reference's message when missing
    long need_int(const std::string& key) const;
    double need_real(const std::string& key) const;
    const std::string& path() const { return path_; }

private:
    std::string path_;
    std::map<std::string, std::string> kv_;
};

}  // namespace spns_hit
