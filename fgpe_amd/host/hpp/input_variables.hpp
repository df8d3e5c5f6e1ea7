// input_variables.hpp -- the input-file reader (the reference's
// InputVariables, hpp/input_variables.hpp): `key value` lines with '#'
// comments; each value is offered as every type it reads as (bool, int,
// double, quoted string), and a typed lookup of a missing key ends the
// program with status 3.
#ifndef FGPE_HOST_INPUT_VARIABLES_HPP
#define FGPE_HOST_INPUT_VARIABLES_HPP

#include <map>
#include <optional>
#include <string>

class InputVariables {
public:
    InputVariables() {}
    explicit InputVariables(const std::string filename);
    // adds the file's entries (a later key replaces an earlier one)
    void readFile(const std::string& filename);
    void printVariables(const std::string _name) const;

    std::string getVariable(const std::string& varName) const;   // the raw text, "None" if absent
    int getInt(const std::string& varName) const;
    double getDouble(const std::string& varName) const;
    std::string getString(const std::string& varName) const;
    bool getBool(const std::string& varName) const;

    std::map<std::string, std::string> getVariables() const;
    std::map<std::string, std::string> getStrings() const;
    std::map<std::string, double> getDoubles() const;
    std::map<std::string, bool> getBools() const;
    std::map<std::string, int> getInts() const;

private:
    // one value and the types it reads as
    struct Value {
        std::string text;
        std::optional<bool> b;
        std::optional<int> i;
        std::optional<double> d;
        std::optional<std::string> s;
        explicit Value(const std::string& t);
    };
    std::map<std::string, Value> entries;

    template <class T, class Get>
    std::map<std::string, T> typed(Get get) const;
    const Value* find(const std::string& key) const;
};

#endif
