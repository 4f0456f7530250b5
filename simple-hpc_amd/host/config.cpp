// Config file reader with the reference's format and messages (src/config.cpp:16-46).
#include "host.hpp"

#include <fstream>
#include <istream>

namespace sh {

namespace {
std::string strip(const std::string& s) {
    const char* ws = " \t\r\n";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos)
        return {};
    const size_t b = s.find_last_not_of(ws);
    return s.substr(a, b - a + 1);
}
} // namespace

ConfigMap parse_config(std::istream& in) {
    ConfigMap out;
    std::string raw;
    for (int line_no = 1; std::getline(in, raw); ++line_no) {
        const size_t hash = raw.find('#');
        const std::string line = strip(hash == std::string::npos ? raw : raw.substr(0, hash));
        if (line.empty())
            continue;
        const std::string at = "config line " + std::to_string(line_no) + ": ";
        const size_t eq = line.find('=');
        if (eq == std::string::npos)
            throw std::runtime_error(at + "malformed line (expected key = value)");
        const std::string key = strip(line.substr(0, eq));
        const std::string val = strip(line.substr(eq + 1));
        if (key.empty() || val.empty())
            throw std::runtime_error(at + "malformed line (empty key or value)");
        if (!out.emplace(key, val).second)
            throw std::runtime_error(at + "duplicate key '" + key + "'");
    }
    return out;
}

ConfigMap load_config(const std::string& path) {
    std::ifstream f(path);
    if (!f)
        throw std::runtime_error("cannot open config file: " + path);
    return parse_config(f);
}

} // namespace sh
