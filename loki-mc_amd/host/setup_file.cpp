// setup_file.cpp -- see setup_file.h.
#include "setup_file.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace lokimc {

std::string removeComments(const std::string& line) {
    const size_t k = line.find('%');
    return k == std::string::npos ? line : line.substr(0, k);
}

std::vector<std::string> splitChars(const std::string& s, const std::string& delims) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : s) {
        if (delims.find(ch) != std::string::npos) {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(ch);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

std::vector<std::string> splitSpaces(const std::string& s) { return splitChars(s, " \t\r\n\v\f"); }

static bool blank(const std::string& s) {
    for (char ch : s)
        if (!std::isspace((unsigned char)ch)) return false;
    return true;
}

// ------------------------------------------------------------ key tree
SetupFile SetupFile::fromText(const std::string& text) {
    SetupFile sf;
    std::istringstream in(text);
    std::string raw;
    int lineNumber = 0;
    while (std::getline(in, raw)) {
        const std::string line = removeComments(raw);
        if (line.empty() || blank(line)) continue;
        SetupField f;
        const std::vector<std::string> tok = splitSpaces(line);
        if (tok[0] == "-") {
            f.isEnumeration = true;
            if (tok.size() == 2) {
                f.name = tok[1];
            } else if (tok.size() == 4) {
                f.name = tok[1];
                f.value = tok[3];
            } else {
                throw InputError("Could not parse line " + std::to_string(lineNumber) + " of the setup file:\n" + line);
            }
        } else {
            const size_t colon = tok[0].find(':');
            if (colon == std::string::npos || tok.size() > 2)
                throw InputError("Could not parse line " + std::to_string(lineNumber) + " of the setup file:\n" + line);
            f.name = tok[0].substr(0, colon);
            if (tok.size() == 2) f.value = tok[1];
        }
        for (char ch : line) {
            if (!std::isspace((unsigned char)ch)) break;
            ++f.spaces;
        }
        // parent: the closest earlier field without value, not an enumeration,
        // and indented less (FieldInfo.C:76-105)
        const int id = (int)sf.fields_.size();
        for (int j = id - 1; j >= 0; --j) {
            const SetupField& p = sf.fields_[j];
            if (f.spaces > p.spaces && p.value.empty() && !p.isEnumeration) {
                f.parent = j;
                f.level = p.level + 1;
                break;
            }
        }
        sf.fields_.push_back(f);
        if (f.parent >= 0) sf.fields_[f.parent].children.push_back(id);
        ++lineNumber;
    }
    return sf;
}

SetupFile SetupFile::fromFile(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw InputError("The setup file '" + path + "' could not be opened.");
    std::stringstream ss;
    ss << in.rdbuf();
    return fromText(ss.str());
}

const SetupField* SetupFile::field(const std::string& chain) const {
    const std::vector<std::string> names = splitChars(chain, ".");
    if (names.empty()) return nullptr;
    const SetupField* cur = nullptr;
    for (const SetupField& f : fields_)
        if (f.level == 0 && f.name == names[0]) { cur = &f; break; }
    for (size_t i = 1; i < names.size() && cur; ++i) {
        const SetupField* next = nullptr;
        for (int c : cur->children)
            if (fields_[c].name == names[i]) { next = &fields_[c]; break; }
        cur = next;
    }
    return cur;
}

std::string SetupFile::value(const std::string& chain) const {
    const SetupField* f = field(chain);
    return f ? f->value : std::string();
}

double SetupFile::number(const std::string& chain, double fallback) const {
    const SetupField* f = field(chain);
    if (!f || f->value.empty()) return fallback;
    return evalNumber(f->value);
}

std::vector<std::string> SetupFile::childNames(const std::string& chain) const {
    std::vector<std::string> out;
    const SetupField* f = field(chain);
    if (f && f->value.empty())
        for (int c : f->children) out.push_back(fields_[c].name);
    return out;
}

static void readPropertyFile(const std::string& path, std::map<std::string, std::string>& m) {
    std::ifstream in(path);
    if (!in) throw InputError("The file " + path + " could not be opened.");
    std::string raw;
    while (std::getline(in, raw)) {
        const std::vector<std::string> tok = splitSpaces(removeComments(raw));
        if (tok.empty()) continue;
        if (tok.size() != 2)
            throw InputError("Error in the parsing of the following property file: " + path +
                             ". Check the following line: \n" + raw);
        m[tok[0]] = tok[1];
    }
}

std::map<std::string, std::string> SetupFile::propertyMap(const std::string& chain, const std::string& inputDir) const {
    std::map<std::string, std::string> m;
    const SetupField* f = field(chain);
    if (!f) return m;
    if (!f->value.empty()) {
        readPropertyFile(inputDir + "/" + f->value, m);
        return m;
    }
    for (int c : f->children) {
        const SetupField& ch = fields_[c];
        if (ch.value.empty()) readPropertyFile(inputDir + "/" + ch.name, m);
        else m[ch.name] = ch.value;
    }
    return m;
}

std::string SetupFile::normalized() const {
    std::string out;
    for (const SetupField& f : fields_) {
        out += std::string(2 * f.level, ' ');
        if (f.isEnumeration) out += f.value.empty() ? "- " + f.name : "- " + f.name + " = " + f.value;
        else out += f.value.empty() ? f.name + ":" : f.name + ": " + f.value;
        out += "\n";
    }
    return out;
}

// ---------------------------------------------------------- expressions
namespace {

// Recursive descent with the precedence of the reference's vendored parser
// (lowest to highest): + -, * /, ^ (left to right), postfix !, unary minus,
// functions, then numbers, constants and parentheses.
struct ExprParser {
    const std::string& s;
    size_t i = 0;
    explicit ExprParser(const std::string& text) : s(text) {}

    [[noreturn]] void fail() const {
        throw InputError("Could not parse the mathematical expression: " + s + ". Please check the input files.");
    }
    void skip() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool eat(char c) {
        skip();
        if (i < s.size() && s[i] == c) { ++i; return true; }
        return false;
    }
    double parse() {
        const double v = sum();
        skip();
        if (i != s.size()) fail();
        return v;
    }
    double sum() {
        double v = product();
        for (;;) {
            if (eat('+')) v = v + product();
            else if (eat('-')) v = v - product();
            else return v;
        }
    }
    double product() {
        double v = power();
        for (;;) {
            if (eat('*')) v = v * power();
            else if (eat('/')) {
                const double d = power();
                if (d == 0) fail();
                v = v / d;
            } else return v;
        }
    }
    double power() {
        double v = postfix();
        while (eat('^')) v = std::pow(v, postfix());
        return v;
    }
    double postfix() {
        double v = unary();
        while (eat('!')) v = factorial(v);
        return v;
    }
    double unary() {
        if (eat('-')) return -function();
        return function();
    }
    static double factorial(double v) {
        if (v < 0 || v != std::floor(v)) throw InputError("factorial of a non-natural number");
        double r = 1;
        for (int k = 2; k <= (int)v; ++k) r *= k;
        return r;
    }
    double function() {
        skip();
        if (i < s.size() && std::isalpha((unsigned char)s[i])) {
            size_t j = i;
            while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) ++j;
            std::string name = s.substr(i, j - i);
            for (char& ch : name) ch = (char)std::toupper((unsigned char)ch);
            if (name == "E") { i = j; return std::exp(1.0); }
            if (name == "PI" || name == "M_PI") { i = j; return M_PI; }
            i = j;
            const double x = primary();
            if (name == "ABS") return std::fabs(x);
            if (name == "EXP") return std::exp(x);
            if (name == "SIGN") return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0);
            if (name == "SQRT") return std::sqrt(x);
            if (name == "LOG") return std::log(x);
            if (name == "LOG10") return std::log10(x);
            if (name == "SIN") return std::sin(x);
            if (name == "COS") return std::cos(x);
            if (name == "TAN") return std::tan(x);
            if (name == "ASIN") return std::asin(x);
            if (name == "ACOS") return std::acos(x);
            if (name == "ATAN") return std::atan(x);
            if (name == "FACTORIAL") return factorial(x);
            fail();
        }
        return primary();
    }
    double primary() {
        skip();
        if (eat('(')) {
            const double v = sum();
            if (!eat(')')) fail();
            return v;
        }
        if (i < s.size() && (std::isdigit((unsigned char)s[i]) || s[i] == '.')) {
            const char* b = s.c_str() + i;
            char* e = nullptr;
            const double v = std::strtod(b, &e);
            if (e == b) fail();
            i += (size_t)(e - b);
            return v;
        }
        if (i < s.size() && (std::isalpha((unsigned char)s[i]) || s[i] == '-')) return unary();
        fail();
    }
};

}  // namespace

double evalNumber(const std::string& expression) {
    ExprParser p(expression);
    const double v = p.parse();
    if (std::isnan(v)) p.fail();
    // the reference prints the answer with %.16g and reads it back
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.16g", v);
    return std::strtod(buf, nullptr);
}

std::vector<double> linSpace(double start, double end, double num) {
    std::vector<double> out;
    if (num == 0) return out;
    if (num == 1) { out.push_back(start); return out; }
    const double delta = (end - start) / (num - 1);
    for (double k = 0; k < num - 1; ++k) out.push_back(start + delta * k);
    out.push_back(end);
    return out;
}

std::vector<double> evalVector(const std::string& expression) {
    auto args = [&](const std::string& what) {
        const size_t b = expression.find('('), e = expression.find(')');
        if (b == std::string::npos || e == std::string::npos) throw InputError("Invalid use of '" + what + "' in the following expression: " + expression);
        const std::vector<std::string> tok = splitChars(expression.substr(b + 1, e - b - 1), ",");
        if (tok.size() != 3) throw InputError("Invalid use of '" + what + "' in the following expression: " + expression);
        return std::vector<double>{evalNumber(tok[0]), evalNumber(tok[1]), evalNumber(tok[2])};
    };
    if (expression.find("linspace") != std::string::npos) {
        const std::vector<double> a = args("linspace");
        return linSpace(a[0], a[1], a[2]);
    }
    if (expression.find("logspace") != std::string::npos) {
        const std::vector<double> a = args("logspace");
        std::vector<double> v = linSpace(a[0], a[1], a[2]);
        for (double& x : v) x = std::pow(10.0, x);
        return v;
    }
    const size_t colons = (size_t)std::count(expression.begin(), expression.end(), ':');
    if (colons == 1) {
        const std::vector<std::string> tok = splitChars(expression, ":");
        const int ini = (int)evalNumber(tok.at(0)), fin = (int)evalNumber(tok.at(1));
        if (ini > fin) throw InputError("Error while using 'evalVectorExpress'. When an expression like 'initial:final' is to be evaluated, 'final' cannot be smaller than 'initial'. Check this expression: " + expression);
        std::vector<double> v;
        for (int k = ini; k <= fin; ++k) v.push_back(k);
        return v;
    }
    if (colons == 2) {
        const std::vector<std::string> tok = splitChars(expression, ":");
        const double ini = evalNumber(tok.at(0)), step = evalNumber(tok.at(1)), fin = evalNumber(tok.at(2));
        if (ini > fin || step <= 0) throw InputError("Error while using 'evalVectorExpress'. When an expression like 'initial:step:final' is to be evaluated, 'final' cannot be smaller than 'initial' and 'step' must be larger than 0. Check this expression: " + expression);
        std::vector<double> v;
        for (double x = ini; x <= fin; x += step) v.push_back(x);
        return v;
    }
    if (expression.find('[') != std::string::npos && expression.find(']') != std::string::npos) {
        std::vector<double> v;
        for (const std::string& t : splitChars(expression, "[],")) v.push_back(evalNumber(t));
        return v;
    }
    return {evalNumber(expression)};
}

}  // namespace lokimc
