// input_variables.cpp -- parser semantics of src/input_variables.cu:55-208:
// text after '#' is dropped, the first two whitespace-separated tokens are
// key and value, a value lands in every typed map it converts to (so "100"
// is an int and a double), quoted values are strings, and a missing typed
// key exits with status 3 (an unreadable file with status 2).
#include "../hpp/input_variables.hpp"

#include <fstream>
#include <iostream>
#include <sstream>

InputVariables::InputVariables(const std::string filename)
{
    std::cout << "#INFO   Reading " << filename << std::endl;
    readFile(filename);
    convertTypes();
}

bool InputVariables::isDouble(const std::string& str)
{
    try {
        std::stod(str);
        return true;
    } catch (const std::exception&) {
        return false;
    }
}

bool InputVariables::isInteger(const std::string& str)
{
    try {
        std::stoi(str);
        return true;
    } catch (const std::exception&) {
        return false;
    }
}

void InputVariables::convertTypes()
{
    for (const auto& entry : variables) {
        const std::string& v = entry.second;
        if (v == "false" || v == "true") bools[entry.first] = (v == "true");
        if (isInteger(v)) ints[entry.first] = std::stoi(v);
        if (v.front() == '"' && v.back() == '"') strings[entry.first] = v.substr(1, v.length() - 2);
        if (isDouble(v)) doubles[entry.first] = std::stod(v);
    }
}

std::string InputVariables::trim(const std::string& str)
{
    size_t first = str.find_first_not_of(" \t\n\r");
    if (first == std::string::npos) return "";
    size_t last = str.find_last_not_of(" \t\n\r");
    return str.substr(first, last - first + 1);
}

bool InputVariables::isCommentOrEmpty(const std::string& line)
{
    std::string t = trim(line);
    return t.empty() || t[0] == '#';
}

void InputVariables::processLine(const std::string& line)
{
    std::istringstream iss(line);
    std::string name, value;
    if (!(iss >> name >> value)) return;
    variables[name] = value;
}

void InputVariables::readFile(const std::string& filename)
{
    std::ifstream file(filename);
    if (!file.is_open()) {
        std::cerr << "Error opening file: " << filename << std::endl;
        exit(2);
    }
    std::string line;
    while (std::getline(file, line)) {
        size_t c = line.find('#');
        if (c != std::string::npos) line = line.substr(0, c);
        if (isCommentOrEmpty(line)) continue;
        processLine(line);
    }
}

void InputVariables::printVariables(const std::string _name) const
{
    std::cout << _name << std::endl;
    if (_name == "Bools")
        for (const auto& e : bools) std::cout << e.first << "\t" << e.second << std::endl;
    else if (_name == "Ints")
        for (const auto& e : ints) std::cout << e.first << "\t" << e.second << std::endl;
    else if (_name == "Doubles")
        for (const auto& e : doubles) std::cout << e.first << "\t" << e.second << std::endl;
    else if (_name == "Strings")
        for (const auto& e : strings) std::cout << e.first << "\t" << e.second << std::endl;
}

std::string InputVariables::getVariable(const std::string& varName) const
{
    auto it = variables.find(varName);
    if (it != variables.end()) return it->second;
    std::cerr << "Variables: No such variable " << varName << std::endl;
    return "None";
}

int InputVariables::getInt(const std::string& varName) const
{
    auto it = ints.find(varName);
    if (it != ints.end()) return it->second;
    std::cerr << "Ints: No such variable " << varName << std::endl;
    exit(3);
}

double InputVariables::getDouble(const std::string& varName) const
{
    auto it = doubles.find(varName);
    if (it != doubles.end()) return it->second;
    std::cerr << "Doubles: No such variable " << varName << std::endl;
    exit(3);
}

std::string InputVariables::getString(const std::string& varName) const
{
    auto it = strings.find(varName);
    if (it != strings.end()) return it->second;
    std::cerr << "Strings: No such variable " << varName << std::endl;
    exit(3);
}

bool InputVariables::getBool(const std::string& varName) const
{
    auto it = bools.find(varName);
    if (it != bools.end()) return it->second;
    std::cerr << "Bools: No such variable " << varName << std::endl;
    exit(3);
}
