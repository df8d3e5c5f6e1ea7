// input_variables.cpp -- the input-file reader.  Behaviour of the
// reference's parser (src/input_variables.cu:55-208) that input files rely
// on, kept here:
//   - text from '#' to the end of a line is dropped; a line contributes its
//     first two whitespace-separated tokens as key and value, and a line with
//     fewer is ignored; a later key replaces an earlier one;
//   - a value is every type it reads as, so "100" is an int and a double,
//     "1.5" an int (1) and a double, "true" a bool, "\"bec\"" a string (the
//     quotes removed); the int and double readings are those of std::stoi
//     and std::stod: a leading number is enough, and an out-of-range one
//     does not count;
//   - an unreadable file ends the program with status 2, a typed lookup of a
//     missing key with status 3.
#include "../hpp/input_variables.hpp"

#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

namespace {

// the number strtol reads from the start of s (base 10), if any and in int range
std::optional<int> read_int(const std::string& s)
{
    const char* b = s.c_str();
    char* e = nullptr;
    errno = 0;
    const long v = std::strtol(b, &e, 10);
    if (e == b || errno == ERANGE || v < INT_MIN || v > INT_MAX) return std::nullopt;
    return (int)v;
}

// the number strtod reads from the start of s, if any and representable
std::optional<double> read_double(const std::string& s)
{
    const char* b = s.c_str();
    char* e = nullptr;
    errno = 0;
    const double v = std::strtod(b, &e);
    if (e == b || errno == ERANGE) return std::nullopt;
    return v;
}

// the whitespace-separated tokens of a line, up to `most`
std::vector<std::string> tokens(const std::string& line, size_t most)
{
    std::vector<std::string> out;
    size_t i = 0;
    while (out.size() < most) {
        while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
        if (i == line.size()) break;
        const size_t j = i;
        while (i < line.size() && !std::isspace((unsigned char)line[i])) ++i;
        out.push_back(line.substr(j, i - j));
    }
    return out;
}

[[noreturn]] void missing(const char* kind, const std::string& key)
{
    std::cerr << kind << ": No such variable " << key << std::endl;
    std::exit(3);
}

}  // namespace

InputVariables::Value::Value(const std::string& t) : text(t)
{
    if (t == "true" || t == "false") b = (t == "true");
    i = read_int(t);
    if (t.front() == '"' && t.back() == '"') s = t.size() >= 2 ? t.substr(1, t.size() - 2) : std::string();
    d = read_double(t);
}

InputVariables::InputVariables(const std::string filename)
{
    std::cout << "#INFO   Reading " << filename << std::endl;
    readFile(filename);
}

void InputVariables::readFile(const std::string& filename)
{
    std::ifstream in(filename);
    if (!in.is_open()) {
        std::cerr << "Error opening file: " << filename << std::endl;
        std::exit(2);
    }
    std::string line;
    while (std::getline(in, line)) {
        const std::vector<std::string> kv = tokens(line.substr(0, line.find('#')), 2);
        if (kv.size() < 2) continue;
        entries.erase(kv[0]);
        entries.emplace(kv[0], Value(kv[1]));
    }
}

const InputVariables::Value* InputVariables::find(const std::string& key) const
{
    const auto it = entries.find(key);
    return it == entries.end() ? nullptr : &it->second;
}

std::string InputVariables::getVariable(const std::string& varName) const
{
    if (const Value* v = find(varName)) return v->text;
    std::cerr << "Variables: No such variable " << varName << std::endl;
    return "None";
}

int InputVariables::getInt(const std::string& varName) const
{
    const Value* v = find(varName);
    if (!v || !v->i) missing("Ints", varName);
    return *v->i;
}

double InputVariables::getDouble(const std::string& varName) const
{
    const Value* v = find(varName);
    if (!v || !v->d) missing("Doubles", varName);
    return *v->d;
}

std::string InputVariables::getString(const std::string& varName) const
{
    const Value* v = find(varName);
    if (!v || !v->s) missing("Strings", varName);
    return *v->s;
}

bool InputVariables::getBool(const std::string& varName) const
{
    const Value* v = find(varName);
    if (!v || !v->b) missing("Bools", varName);
    return *v->b;
}

template <class T, class Get> std::map<std::string, T> InputVariables::typed(Get get) const
{
    std::map<std::string, T> out;
    for (const auto& e : entries)
        if (const auto& x = get(e.second)) out.emplace(e.first, *x);
    return out;
}

std::map<std::string, std::string> InputVariables::getVariables() const
{
    std::map<std::string, std::string> out;
    for (const auto& e : entries) out.emplace(e.first, e.second.text);
    return out;
}
std::map<std::string, std::string> InputVariables::getStrings() const
{
    return typed<std::string>([](const Value& v) -> const std::optional<std::string>& { return v.s; });
}
std::map<std::string, double> InputVariables::getDoubles() const
{
    return typed<double>([](const Value& v) -> const std::optional<double>& { return v.d; });
}
std::map<std::string, bool> InputVariables::getBools() const
{
    return typed<bool>([](const Value& v) -> const std::optional<bool>& { return v.b; });
}
std::map<std::string, int> InputVariables::getInts() const
{
    return typed<int>([](const Value& v) -> const std::optional<int>& { return v.i; });
}

void InputVariables::printVariables(const std::string _name) const
{
    std::cout << _name << std::endl;
    auto show = [](const auto& m) {
        for (const auto& e : m) std::cout << e.first << "\t" << e.second << std::endl;
    };
    if (_name == "Bools") show(getBools());
    else if (_name == "Ints") show(getInts());
    else if (_name == "Doubles") show(getDoubles());
    else if (_name == "Strings") show(getStrings());
    else if (_name == "Variables") show(getVariables());
}
