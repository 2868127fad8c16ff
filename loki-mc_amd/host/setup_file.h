// setup_file.h -- reader of LoKI-MC setup files: the indentation key tree
// (FieldInfo.C:12-98), '%' comments (Parse.C:786-795), numeric expressions
// with the reference's 16-significant-digit rounding (Parse.C:716-742) and
// vector expressions (Parse.C:590-652).
#pragma once
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace lokimc {

// Errors of the setup and input files; the message is what the reference's
// Message::error would print.
struct InputError : std::runtime_error {
    explicit InputError(const std::string& m) : std::runtime_error(m) {}
};

// One line of the setup file: "name:", "name: value", "- name" or
// "- name = value" (FieldInfo.C:27-73).
struct SetupField {
    std::string name, value;
    bool isEnumeration = false;
    int spaces = 0;               // leading white space of the line
    int level = 0;
    int parent = -1;              // index of the parent field, -1 at level 0
    std::vector<int> children;
};

class SetupFile {
public:
    // Parse `text` (the content of a setup file).
    static SetupFile fromText(const std::string& text);
    static SetupFile fromFile(const std::string& path);

    // Field at a dotted chain of names ("electronKinetics.numericsMC.nElectrons"), or null.
    const SetupField* field(const std::string& chain) const;
    bool has(const std::string& chain) const { return field(chain) != nullptr; }
    // Value of a field ("" when absent or without value).
    std::string value(const std::string& chain) const;
    // Numeric value of a field, or `fallback` when absent.
    double number(const std::string& chain, double fallback) const;
    // Names of the children of a field (FieldInfo::getFieldChildNames).
    std::vector<std::string> childNames(const std::string& chain) const;
    // A property map (FieldInfo::getFieldMap): the field's value names a
    // "name value" file, or each child is either such a file or "name = value".
    // Files are looked up under `inputDir`.
    std::map<std::string, std::string> propertyMap(const std::string& chain, const std::string& inputDir) const;
    // The setup as the reference saves it (FieldInfo::saveSetupInfo).
    std::string normalized() const;

    const std::vector<SetupField>& fields() const { return fields_; }

private:
    std::vector<SetupField> fields_;
};

// '%' starts a comment (Parse::removeComments).
std::string removeComments(const std::string& line);
// Tokens separated by white space.
std::vector<std::string> splitSpaces(const std::string& s);
// Tokens separated by any of `delims` (empty tokens dropped, as strtok).
std::vector<std::string> splitChars(const std::string& s, const std::string& delims);

// Value of a numeric expression (+ - * / ^ %, unary signs, parentheses,
// ABS EXP SIGN SQRT LOG LOG10 SIN COS TAN ASIN ACOS ATAN FACTORIAL, constants
// E and PI), rounded to 16 significant digits as the reference's parser
// prints and re-reads it (Parse::str2value).  Throws InputError.
double evalNumber(const std::string& expression);
// linspace(a,b,n), logspace(a,b,n), a:b, a:step:b, [x,y,...] or a scalar
// (Parse::evalVectorExpress).
std::vector<double> evalVector(const std::string& expression);
std::vector<double> linSpace(double start, double end, double num);

}  // namespace lokimc
