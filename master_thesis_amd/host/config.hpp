// config.hpp — the reference's text configuration format
// (software_renderer/input_parser.h:9-63, input_parser.cpp:38-93):
//
//   EntryName                  a line starting in column 0
//   <TAB>key(v1, v2, ...)      values of the current entry; spaces dropped
//   # comment                  whole-line comment
//
// InputEntry::get<T> converts with operator>> (floats via strtof semantics),
// as input_parser.h:20-24 does.  Errors are reported as InitError (the
// counterpart of InitException, init_exception.h:9-19).
#pragma once

#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "mth_math.hpp"

namespace mth {

class InitError : public std::runtime_error {
  public:
    InitError(const std::string& system, const std::string& msg)
        : std::runtime_error("Init exception: " + system + ": " + msg) {}
};

class InputEntry {
  public:
    explicit InputEntry(std::string n = "") : name(std::move(n)) {}
    std::string name;

    bool has(const std::string& key) const { return data_.count(key) > 0; }
    const std::vector<std::string>& values(const std::string& key) const;
    float get_float(const std::string& key, size_t i = 0) const;
    unsigned get_uint(const std::string& key, size_t i = 0) const;
    std::string get_string(const std::string& key, size_t i = 0) const;
    Vec3 get_vec3(const std::string& key) const;
    Rotation get_rotation(const std::string& key) const;  // axis x, y, z, angle (input_parser.cpp:20-26)
    void insert(const std::string& key, std::vector<std::string> vals) { data_.emplace(key, std::move(vals)); }

  private:
    std::unordered_map<std::string, std::vector<std::string>> data_;
};

// Parses a whole file; throws InitError if it cannot be opened.
std::vector<InputEntry> parse_config_file(const std::string& path);
// Parses text already in memory (same grammar).
std::vector<InputEntry> parse_config_text(const std::string& text);

}  // namespace mth
