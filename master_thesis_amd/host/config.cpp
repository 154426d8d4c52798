// config.cpp — reader of the .scene / .renderer / .camera text format.
#include "config.hpp"

#include <cstdlib>
#include <fstream>
#include <sstream>

namespace mth {

const std::vector<std::string>& InputEntry::values(const std::string& key) const {
    auto it = data_.find(key);
    if (it == data_.end()) throw InitError("InputEntry", "missing key \"" + key + "\" in entry \"" + name + "\"");
    return it->second;
}

static const std::string& value_at(const InputEntry& e, const std::string& key, size_t i) {
    const auto& v = e.values(key);
    if (i >= v.size())
        throw InitError("InputEntry", "key \"" + key + "\" has no value " + std::to_string(i) + " in \"" + e.name + "\"");
    return v[i];
}

float InputEntry::get_float(const std::string& key, size_t i) const {
    float x = 0.0f;
    std::stringstream(value_at(*this, key, i)) >> x;
    return x;
}

unsigned InputEntry::get_uint(const std::string& key, size_t i) const {
    unsigned x = 0;
    std::stringstream(value_at(*this, key, i)) >> x;
    return x;
}

std::string InputEntry::get_string(const std::string& key, size_t i) const { return value_at(*this, key, i); }

Vec3 InputEntry::get_vec3(const std::string& key) const {
    return Vec3(get_float(key, 0), get_float(key, 1), get_float(key, 2));
}

Rotation InputEntry::get_rotation(const std::string& key) const {
    Vec3 axis(std::stof(value_at(*this, key, 0)), std::stof(value_at(*this, key, 1)),
              std::stof(value_at(*this, key, 2)));
    return Rotation::from_axis_angle(axis, std::stof(value_at(*this, key, 3)));
}

// Character-level reader with the reference's rules: '\n' skipped, '#'
// consumes the line, '\t' starts a key(values) line, any other non-space
// character starts a new entry whose name is the rest of the line.
static std::vector<InputEntry> parse_stream(std::istream& in) {
    std::vector<InputEntry> out;
    InputEntry cur("");
    while (in.good()) {
        int c = in.get();
        if (c == EOF) break;
        if (c == '\n') continue;
        if (c == '#') {
            std::string tmp;
            std::getline(in, tmp);
        } else if (c == '\t') {
            std::string key;
            std::getline(in, key, '(');
            std::vector<std::string> vals;
            std::string v;
            while (in.good()) {
                int ch = in.get();
                if (ch == EOF) break;
                if (ch == ')') {
                    in.get();  // the newline after ')'
                    break;
                } else if (ch == ',') {
                    vals.push_back(v);
                    v.clear();
                } else if (ch != ' ') {
                    v += (char)ch;
                }
            }
            vals.push_back(v);
            cur.insert(key, std::move(vals));
        } else if (c != ' ') {
            in.unget();
            if (!cur.name.empty()) out.push_back(cur);
            std::string name;
            std::getline(in, name);
            cur = InputEntry(name);
        }
    }
    if (!cur.name.empty()) out.push_back(cur);
    return out;
}

std::vector<InputEntry> parse_config_file(const std::string& path) {
    std::ifstream in(path);
    if (in.fail()) throw InitError("InputParser", "failed to open file \"" + path + "\"!");
    return parse_stream(in);
}

std::vector<InputEntry> parse_config_text(const std::string& text) {
    std::istringstream in(text);
    return parse_stream(in);
}

}  // namespace mth
