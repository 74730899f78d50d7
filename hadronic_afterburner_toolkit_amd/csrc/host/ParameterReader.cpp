// ParameterReader.cpp -- see ParameterReader.h.  Follows
// ParameterReader.cpp:29-176 and arsenal.cpp:37-60 of the reference.
#include "ParameterReader.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

namespace {

// all blanks and tabs removed (arsenal.cpp trim)
std::string squeeze(const std::string &s) {
    std::string t;
    for (char c : s)
        if (c != ' ' && c != '\t') t.push_back(c);
    return t;
}

std::string lower(std::string s) {
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// first number of the string, as `stringstream >> double` reads it
double to_double(const std::string &s) {
    std::stringstream sst(s + " ");
    double v = 0.0;
    sst >> v;
    return v;
}

}  // namespace

long ParameterReader::find(const std::string &name) const {
    const std::string key = lower(squeeze(name));
    for (size_t i = 0; i < names.size(); i++)
        if (names[i] == key) return (long)i;
    return -1;
}

void ParameterReader::parseLine(std::string line, const std::string &commentSymbol) {
    if (squeeze(line).empty()) return;
    line = line.substr(0, line.find(commentSymbol));
    if (squeeze(line).empty()) return;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) {
        std::cout << "ParameterReader::phraseEquationWithoutComments error:"
                  << "\"=\" symbol not found in equation assignment " << line << std::endl;
        exit(-1);
    }
    setVal(line.substr(0, eq), to_double(squeeze(line.substr(eq + 1))));
}

void ParameterReader::readFromFile(std::string filename) {
    std::ifstream f(filename.c_str());
    if (!f) {
        std::cout << "ParameterReader::readFromFile error: file " << filename
                  << " does not exist." << std::endl;
        exit(-1);
    }
    std::string line;
    while (std::getline(f, line)) parseLine(line, "#");
}

void ParameterReader::readFromArguments(long argc, char *argv[], std::string commentSymbol,
                                        long start_from) {
    for (long i = start_from; i < argc; i++) parseLine(argv[i], commentSymbol);
}

bool ParameterReader::exist(std::string name) const { return find(name) != -1; }

void ParameterReader::setVal(std::string name, double value) {
    const long idx = find(name);
    if (idx == -1) {
        names.push_back(lower(squeeze(name)));
        values.push_back(value);
    } else {
        values[idx] = value;
    }
}

double ParameterReader::getVal(std::string name) const {
    const long idx = find(name);
    if (idx != -1) return values[idx];
    std::cout << "ParameterReader::getVal error: parameter with name " << name << " not found."
              << std::endl;
    exit(-1);
}

double ParameterReader::getVal(std::string name, double defaultValue) const {
    const long idx = find(name);
    if (idx != -1) return values[idx];
    std::cout << "ParameterReader::getVal error: parameter with name " << name << " not found."
              << std::endl;
    std::cout << "using defaultValue: " << defaultValue << std::endl;
    return defaultValue;
}

void ParameterReader::echo() const {
    if (names.empty()) return;
    for (size_t i = 0; i < names.size(); i++) std::cout << names[i] << "=" << values[i] << "  ";
    std::cout << std::endl;
}
