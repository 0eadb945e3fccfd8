// TOML subset parser and ArgParser (see toml_lite.h).
#include "toml_lite.h"

#include <cctype>
#include <cmath>
#include <fstream>
#include <limits>
#include <sstream>

namespace hdm {

const TomlValue &TomlValue::at(const std::string &k) const {
    if (kind != Table) throw TomlError("not a table when looking up '" + k + "'");
    auto it = table.find(k);
    if (it == table.end()) throw TomlError("key '" + k + "' not found");
    return it->second;
}

TomlValue &TomlValue::operator[](const std::string &k) {
    if (kind == None) kind = Table;
    if (kind != Table) throw TomlError("not a table when setting '" + k + "'");
    return table[k];
}

static std::string quote(const std::string &s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        if (c == '\n') { o += "\\n"; continue; }
        o += c;
    }
    return o + "\"";
}

std::string TomlValue::dump(int indent) const {
    std::ostringstream o;
    o.precision(17);
    switch (kind) {
    case None: return "";
    case Str: return quote(s);
    case Int: o << i; return o.str();
    case Float:
        if (std::isinf(f)) return f > 0 ? "inf" : "-inf";
        if (std::isnan(f)) return "nan";
        o << f;
        if (o.str().find_first_of(".eE") == std::string::npos) o << ".0";
        return o.str();
    case Bool: return b ? "true" : "false";
    case Array: {
        o << "[";
        for (size_t k = 0; k < arr.size(); k++) o << (k ? ", " : "") << arr[k].dump(indent);
        o << "]";
        return o.str();
    }
    case Table: {
        // scalars first, then sub-tables, then arrays of tables
        for (auto &kv : table)
            if (kv.second.kind != Table && !(kv.second.kind == Array && !kv.second.arr.empty() && kv.second.arr[0].kind == Table))
                o << kv.first << " = " << kv.second.dump(indent) << "\n";
        for (auto &kv : table)
            if (kv.second.kind == Table) o << "\n[" << kv.first << "]\n" << kv.second.dump(indent + 1);
        for (auto &kv : table)
            if (kv.second.kind == Array && !kv.second.arr.empty() && kv.second.arr[0].kind == Table)
                for (auto &t : kv.second.arr) o << "\n[[" << kv.first << "]]\n" << t.dump(indent + 1);
        return o.str();
    }
    }
    return "";
}

namespace {

struct Parser {
    const std::string &t;
    std::string name;
    size_t pos = 0;
    int line = 1;

    Parser(const std::string &text, const std::string &n) : t(text), name(n) {}

    [[noreturn]] void error(const std::string &msg) {
        throw TomlError(name + ":" + std::to_string(line) + ": " + msg);
    }
    bool eof() const { return pos >= t.size(); }
    char peek(size_t k = 0) const { return pos + k < t.size() ? t[pos + k] : '\0'; }
    char get() {
        char c = t[pos++];
        if (c == '\n') line++;
        return c;
    }
    void skip_ws() {
        while (!eof() && (peek() == ' ' || peek() == '\t' || peek() == '\r')) pos++;
    }
    void skip_comment() {
        if (peek() == '#')
            while (!eof() && peek() != '\n') pos++;
    }
    void skip_ws_nl_comments() {
        for (;;) {
            skip_ws();
            skip_comment();
            if (peek() == '\n') {
                get();
                continue;
            }
            break;
        }
    }
    void expect_line_end() {
        skip_ws();
        skip_comment();
        if (eof()) return;
        if (peek() != '\n') error(std::string("unexpected character '") + peek() + "'");
        get();
    }

    std::string parse_simple_key() {
        skip_ws();
        if (peek() == '"' || peek() == '\'') return parse_string_literal();
        std::string k;
        while (!eof() && (std::isalnum((unsigned char)peek()) || peek() == '_' || peek() == '-')) k += get();
        if (k.empty()) error("expected a key");
        return k;
    }
    std::vector<std::string> parse_key() {
        std::vector<std::string> parts{parse_simple_key()};
        skip_ws();
        while (peek() == '.') {
            get();
            parts.push_back(parse_simple_key());
            skip_ws();
        }
        return parts;
    }

    std::string parse_string_literal() {
        char q = get();
        if (q == '"' && peek() == '"' && peek(1) == '"') error("multi-line strings are not supported");
        std::string s;
        while (!eof() && peek() != q) {
            char c = get();
            if (c == '\n') error("newline in string");
            if (q == '"' && c == '\\') {
                char e = get();
                switch (e) {
                case 'n': s += '\n'; break;
                case 't': s += '\t'; break;
                case 'r': s += '\r'; break;
                case '"': s += '"'; break;
                case '\\': s += '\\'; break;
                default: error(std::string("unsupported escape \\") + e);
                }
            } else {
                s += c;
            }
        }
        if (eof()) error("unterminated string");
        get();
        return s;
    }

    TomlValue parse_number_or_bool() {
        TomlValue v;
        std::string tok;
        while (!eof() && (std::isalnum((unsigned char)peek()) || peek() == '.' || peek() == '+' || peek() == '-' || peek() == '_'))
            tok += get();
        if (tok == "true" || tok == "false") {
            v.kind = TomlValue::Bool;
            v.b = tok == "true";
            return v;
        }
        std::string clean;
        for (char c : tok)
            if (c != '_') clean += c;
        std::string body = clean;
        if (!body.empty() && (body[0] == '+' || body[0] == '-')) body = body.substr(1);
        if (body == "inf" || body == "nan") {
            v.kind = TomlValue::Float;
            v.f = body == "inf" ? std::numeric_limits<double>::infinity() : std::numeric_limits<double>::quiet_NaN();
            if (clean[0] == '-') v.f = -v.f;
            return v;
        }
        if (body.empty()) error("expected a value");
        bool isFloat = clean.find_first_of(".eE") != std::string::npos;
        try {
            size_t used = 0;
            if (isFloat) {
                v.kind = TomlValue::Float;
                v.f = std::stod(clean, &used);
            } else {
                v.kind = TomlValue::Int;
                v.i = std::stoll(clean, &used, 10);
            }
            if (used != clean.size()) error("bad number '" + tok + "'");
        } catch (const std::logic_error &) {
            error("bad number '" + tok + "'");
        }
        return v;
    }

    TomlValue parse_value() {
        skip_ws();
        char c = peek();
        if (c == '"' || c == '\'') {
            TomlValue v;
            v.kind = TomlValue::Str;
            v.s = parse_string_literal();
            return v;
        }
        if (c == '[') {
            get();
            TomlValue v;
            v.kind = TomlValue::Array;
            skip_ws_nl_comments();
            while (peek() != ']') {
                v.arr.push_back(parse_value());
                skip_ws_nl_comments();
                if (peek() == ',') {
                    get();
                    skip_ws_nl_comments();
                } else if (peek() != ']') {
                    error("expected ',' or ']' in array");
                }
            }
            get();
            return v;
        }
        if (c == '{') {
            get();
            TomlValue v;
            v.kind = TomlValue::Table;
            skip_ws();
            while (peek() != '}') {
                auto key = parse_key();
                skip_ws();
                if (get() != '=') error("expected '=' in inline table");
                set_dotted(v, key, parse_value());
                skip_ws();
                if (peek() == ',') {
                    get();
                    skip_ws();
                } else if (peek() != '}') {
                    error("expected ',' or '}' in inline table");
                }
            }
            get();
            return v;
        }
        return parse_number_or_bool();
    }

    void set_dotted(TomlValue &root, const std::vector<std::string> &key, TomlValue val) {
        TomlValue *cur = &root;
        for (size_t k = 0; k + 1 < key.size(); k++) {
            TomlValue &nx = (*cur)[key[k]];
            if (nx.kind == TomlValue::None) nx.kind = TomlValue::Table;
            if (nx.kind != TomlValue::Table) error("key '" + key[k] + "' is not a table");
            cur = &nx;
        }
        if (cur->contains(key.back())) error("duplicate key '" + key.back() + "'");
        (*cur)[key.back()] = std::move(val);
    }

    TomlValue *descend(TomlValue &root, const std::vector<std::string> &key, bool arrayTable) {
        TomlValue *cur = &root;
        for (size_t k = 0; k < key.size(); k++) {
            bool last = k + 1 == key.size();
            TomlValue &nx = (*cur)[key[k]];
            if (last && arrayTable) {
                if (nx.kind == TomlValue::None) nx.kind = TomlValue::Array;
                if (nx.kind != TomlValue::Array) error("'" + key[k] + "' is not an array of tables");
                nx.arr.emplace_back();
                nx.arr.back().kind = TomlValue::Table;
                return &nx.arr.back();
            }
            if (nx.kind == TomlValue::None) nx.kind = TomlValue::Table;
            if (nx.kind == TomlValue::Array && !nx.arr.empty() && nx.arr.back().kind == TomlValue::Table) {
                cur = &nx.arr.back();
                continue;
            }
            if (nx.kind != TomlValue::Table) error("'" + key[k] + "' is not a table");
            cur = &nx;
        }
        return cur;
    }

    TomlValue parse_document() {
        TomlValue root;
        root.kind = TomlValue::Table;
        TomlValue *cur = &root;
        for (;;) {
            skip_ws_nl_comments();
            if (eof()) break;
            if (peek() == '[') {
                bool arr = peek(1) == '[';
                get();
                if (arr) get();
                auto key = parse_key();
                skip_ws();
                if (get() != ']') error("expected ']'");
                if (arr && get() != ']') error("expected ']]'");
                cur = descend(root, key, arr);
                expect_line_end();
                continue;
            }
            auto key = parse_key();
            skip_ws();
            if (get() != '=') error("expected '='");
            set_dotted(*cur, key, parse_value());
            expect_line_end();
        }
        return root;
    }
};

std::string strip_prefix(const std::string &s, char c) {
    size_t k = 0;
    while (k < s.size() && s[k] == c) k++;
    return s.substr(k);
}

std::string trim(const std::string &s) {
    size_t b = s.find_first_not_of(" \t"), e = s.find_last_not_of(" \t");
    return b == std::string::npos ? "" : s.substr(b, e - b + 1);
}

}  // namespace

TomlValue toml_parse_string(const std::string &text, const std::string &name) {
    Parser p(text, name);
    return p.parse_document();
}

TomlValue toml_parse_file(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw TomlError("cannot open config file '" + path + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    return toml_parse_string(ss.str(), path);
}

template <> double toml_as<double>(const TomlValue &v) {
    if (v.kind == TomlValue::Float) return v.f;
    if (v.kind == TomlValue::Int) return (double)v.i;
    throw TomlError("value is not a number");
}
template <> float toml_as<float>(const TomlValue &v) { return (float)toml_as<double>(v); }
template <> int64_t toml_as<int64_t>(const TomlValue &v) {
    if (v.kind == TomlValue::Int) return v.i;
    throw TomlError("value is not an integer");
}
template <> int toml_as<int>(const TomlValue &v) { return (int)toml_as<int64_t>(v); }
template <> unsigned int toml_as<unsigned int>(const TomlValue &v) { return (unsigned int)toml_as<int64_t>(v); }
template <> bool toml_as<bool>(const TomlValue &v) {
    if (v.kind == TomlValue::Bool) return v.b;
    throw TomlError("value is not a boolean");
}
template <> std::string toml_as<std::string>(const TomlValue &v) {
    if (v.kind == TomlValue::Str) return v.s;
    throw TomlError("value is not a string");
}
template <> std::vector<int> toml_as<std::vector<int>>(const TomlValue &v) {
    if (v.kind != TomlValue::Array) throw TomlError("value is not an array");
    std::vector<int> o;
    for (auto &e : v.arr) o.push_back(toml_as<int>(e));
    return o;
}
template <> std::vector<double> toml_as<std::vector<double>>(const TomlValue &v) {
    if (v.kind != TomlValue::Array) throw TomlError("value is not an array");
    std::vector<double> o;
    for (auto &e : v.arr) o.push_back(toml_as<double>(e));
    return o;
}

template <> double cli_as<double>(const std::string &s) { return std::stod(trim(s)); }
template <> float cli_as<float>(const std::string &s) { return std::stof(trim(s)); }
template <> int64_t cli_as<int64_t>(const std::string &s) { return std::stoll(trim(s)); }
template <> int cli_as<int>(const std::string &s) { return std::stoi(trim(s)); }
template <> unsigned int cli_as<unsigned int>(const std::string &s) { return (unsigned int)std::stoul(trim(s)); }
template <> bool cli_as<bool>(const std::string &s) {
    std::string t = trim(s);
    if (t == "true" || t == "1") return true;
    if (t == "false" || t == "0") return false;
    throw TomlError("bad boolean '" + s + "'");
}
template <> std::string cli_as<std::string>(const std::string &s) {
    std::string t = trim(s);
    if (t.size() >= 2 && t.front() == '"' && t.back() == '"') t = t.substr(1, t.size() - 2);
    return t;
}
template <> std::vector<int> cli_as<std::vector<int>>(const std::string &s) {
    return toml_as<std::vector<int>>(toml_parse_string("v = " + s).at("v"));
}
template <> std::vector<double> cli_as<std::vector<double>>(const std::string &s) {
    return toml_as<std::vector<double>>(toml_parse_string("v = " + s).at("v"));
}

static TomlValue to_toml(double v) { TomlValue t; t.kind = TomlValue::Float; t.f = v; return t; }
static TomlValue to_toml(float v) { return to_toml((double)v); }
static TomlValue to_toml(int64_t v) { TomlValue t; t.kind = TomlValue::Int; t.i = v; return t; }
static TomlValue to_toml(int v) { return to_toml((int64_t)v); }
static TomlValue to_toml(unsigned int v) { return to_toml((int64_t)v); }
static TomlValue to_toml(bool v) { TomlValue t; t.kind = TomlValue::Bool; t.b = v; return t; }
static TomlValue to_toml(const std::string &v) { TomlValue t; t.kind = TomlValue::Str; t.s = v; return t; }
template <class E> static TomlValue to_toml(const std::vector<E> &v) {
    TomlValue t;
    t.kind = TomlValue::Array;
    for (auto &e : v) t.arr.push_back(to_toml(e));
    return t;
}

// ArgParser::ArgParser (ArgParser.h:170-196): arguments after the program
// name are "--section.key=value" or "--section.key value..." (values that do
// not start with '-' are joined); the last argument is the config file.
void ArgParser::parseCommandLine(const std::vector<std::string> &args, bool lastIsConfig) {
    int argc = (int)args.size();
    int stop = lastIsConfig ? argc - 1 : argc;
    for (int i = 0; i < stop; ++i) {
        std::string argument = args[(size_t)i];
        while (i < stop - 1 && !args[(size_t)i + 1].empty() && args[(size_t)i + 1].front() != '-') {
            i++;
            argument += " " + args[(size_t)i];
        }
        size_t sp = argument.find(' '), eq = argument.find('=');
        size_t split = std::min(sp, eq);
        if (split != std::string::npos) lineArguments[strip_prefix(argument.substr(0, split), '-')] = argument.substr(split + 1);
        else lineArguments[strip_prefix(argument, '-')] = "true";
    }
}

ArgParser::ArgParser(int argc, const char *const *argv) {
    table.kind = TomlValue::Table;
    std::vector<std::string> args;
    for (int i = 1; i < argc; i++) args.emplace_back(argv[i]);
    if (args.empty()) throw TomlError("usage: <program> [--section.key=value ...] config.toml");
    parseCommandLine(args, true);
    configPath = args.back();
    table = toml_parse_file(configPath);
}

ArgParser ArgParser::fromString(const std::string &toml, const std::vector<std::string> &overrides) {
    ArgParser a;
    a.table = toml_parse_string(toml, "<config>");
    a.parseCommandLine(overrides, false);
    return a;
}

template <class T> std::optional<T> ArgParser::getArg(const std::string &name) {
    cliArgs.push_back(name);
    std::optional<T> tomlValue, lineValue;
    std::string first, second;
    size_t dot = name.find('.');
    if (dot == std::string::npos) {
        if (table.contains(name)) tomlValue = toml_as<T>(table.at(name));
    } else {
        first = name.substr(0, dot);
        second = name.substr(dot + 1);
        if (table.contains(first) && table.at(first).is_table() && table.at(first).contains(second))
            tomlValue = toml_as<T>(table.at(first).at(second));
    }
    auto it = lineArguments.find(name);
    if (it != lineArguments.end()) lineValue = cli_as<T>(it->second);
    if (lineValue) {
        // the reference writes command-line values back into the table (ArgParser.h:236-251)
        if (dot == std::string::npos) table[name] = to_toml(*lineValue);
        else table[first][second] = to_toml(*lineValue);
        return lineValue;
    }
    return tomlValue;
}

template std::optional<double> ArgParser::getArg<double>(const std::string &);
template std::optional<float> ArgParser::getArg<float>(const std::string &);
template std::optional<int> ArgParser::getArg<int>(const std::string &);
template std::optional<int64_t> ArgParser::getArg<int64_t>(const std::string &);
template std::optional<unsigned int> ArgParser::getArg<unsigned int>(const std::string &);
template std::optional<bool> ArgParser::getArg<bool>(const std::string &);
template std::optional<std::string> ArgParser::getArg<std::string>(const std::string &);
template std::optional<std::vector<int>> ArgParser::getArg<std::vector<int>>(const std::string &);
template std::optional<std::vector<double>> ArgParser::getArg<std::vector<double>>(const std::string &);

}  // namespace hdm
