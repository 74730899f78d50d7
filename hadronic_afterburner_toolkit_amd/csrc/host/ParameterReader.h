// ParameterReader -- the reference's key = value parameter store
// (src/ParameterReader.h), same interface and semantics: keys are
// lower-cased with blanks removed, values are doubles, "#" starts a
// comment, command-line "key=value" overrides the file, getVal(name)
// exits(-1) when the key is missing and getVal(name, default) falls back.
#ifndef HBT_AMD_PARAMETERREADER_H
#define HBT_AMD_PARAMETERREADER_H

#include <string>
#include <vector>

class ParameterReader {
  public:
    ParameterReader() = default;
    void readFromFile(std::string filename);
    void readFromArguments(long argc, char *argv[], std::string commentSymbol = "#",
                           long start_from = 1);
    bool exist(std::string name) const;
    void setVal(std::string name, double value);
    double getVal(std::string name) const;
    double getVal(std::string name, double defaultValue) const;
    void echo() const;

  private:
    std::vector<std::string> names;
    std::vector<double> values;
    long find(const std::string &name) const;
    void parseLine(std::string line, const std::string &commentSymbol);
};

#endif
