// TOML-driven setup: a TOML subset parser and the reference's ArgParser
// interface (PLiB/Util/ArgParser.h:170-255).
//
// The reference parses its config with toml11 and lets `--section.key=value`
// (or `--section.key value`) on the command line override any scalar
// (HiDEM/Docs/Input.md:11-18); the last command-line argument is the config
// path.  toml11 is not available here, so this parser covers the subset the
// HiDEM configs use: [table], [[array.of.tables]], key = value with strings,
// integers, floats (incl. 1e+09, inf, nan), booleans, arrays and inline
// tables, and # comments.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace hdm {

struct TomlValue {
    enum Kind { None, Str, Int, Float, Bool, Array, Table } kind = None;
    std::string s;
    int64_t i = 0;
    double f = 0;
    bool b = false;
    std::vector<TomlValue> arr;
    std::map<std::string, TomlValue> table;

    bool is_table() const { return kind == Table; }
    bool contains(const std::string &k) const { return kind == Table && table.count(k) != 0; }
    const TomlValue &at(const std::string &k) const;
    TomlValue &operator[](const std::string &k);
    std::string dump(int indent = 0) const;
};

struct TomlError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

TomlValue toml_parse_string(const std::string &text, const std::string &name = "<string>");
TomlValue toml_parse_file(const std::string &path);

// Conversions used by ArgParser::getArg<T>; throw TomlError on a type mismatch.
template <class T> T toml_as(const TomlValue &v);
// Parses a command-line literal (Parsers::getAs in the reference).
template <class T> T cli_as(const std::string &s);

class ArgParser {
public:
    TomlValue table;

    ArgParser() { table.kind = TomlValue::Table; }
    ArgParser(int argc, const char *const *argv);
    static ArgParser fromString(const std::string &toml, const std::vector<std::string> &overrides = {});

    const std::string &getConfigPath() const { return configPath; }

    // TOML value of "section.key" (or "key"), overridden by a matching
    // --section.key command-line argument; std::nullopt when neither exists.
    template <class T> std::optional<T> getArg(const std::string &name);

    std::vector<std::string> queried() const { return cliArgs; }

private:
    std::string configPath;
    std::map<std::string, std::string> lineArguments;
    std::vector<std::string> cliArgs;
    void parseCommandLine(const std::vector<std::string> &args, bool lastIsConfig);
};

}  // namespace hdm
