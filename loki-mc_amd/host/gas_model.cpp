// gas_model.cpp -- see gas_model.h.
#include "gas_model.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>

namespace lokimc {

std::string RawState::name() const {
    std::string s = gas + "(";
    if (!ionCharg.empty()) s += ionCharg + ",";
    s += eleLevel;
    if (!vibLevel.empty()) {
        s += ",v=" + vibLevel;
        if (!rotLevel.empty()) s += ",J=" + rotLevel;
    }
    return s + ")";
}

static std::string afterEquals(const std::string& s) {
    const std::vector<std::string> t = splitChars(s, "=");
    return t.size() > 1 ? t[1] : std::string();
}

RawState parseState(const std::string& text) {
    const std::vector<std::string> tok = splitChars(text, "()");
    if (tok.size() < 2)
        throw InputError("The states must have the electronic state defined between parentheses!\nPlease check '" + text + "'");
    RawState r;
    r.gas = tok[0];
    const std::vector<std::string> f = splitChars(tok[1], ",");
    auto vib = [&](const std::string& s) { return s.find("v=") != std::string::npos || s.find("w=") != std::string::npos; };
    switch (f.size()) {
        case 1: r.eleLevel = f[0]; break;
        case 2:
            if (vib(f[1])) { r.eleLevel = f[0]; r.vibLevel = afterEquals(f[1]); }
            else { r.ionCharg = f[0]; r.eleLevel = f[1]; }
            break;
        case 3:
            if (f[2].find("J=") != std::string::npos) {
                r.eleLevel = f[0];
                if (vib(f[1])) r.vibLevel = afterEquals(f[1]);
                r.rotLevel = afterEquals(f[2]);
            } else {
                r.ionCharg = f[0]; r.eleLevel = f[1]; r.vibLevel = f[2];
            }
            break;
        case 4:
            r.ionCharg = f[0]; r.eleLevel = f[1];
            if (vib(f[2])) r.vibLevel = afterEquals(f[2]);
            r.rotLevel = afterEquals(f[3]);
            break;
        default: throw InputError("Error! Check the state '" + text + "'.");
    }
    return r;
}

// Leading stoichiometric coefficient of "2N2(X)" (Parse::getStoiCoeff).
static double stoiCoeff(std::string& token) {
    size_t k = 0;
    while (k < token.size() && !std::isalpha((unsigned char)token[k])) ++k;
    if (k == 0) return 1;
    const double c = evalNumber(token.substr(0, k));
    token.erase(0, k);
    return c;
}

// "e + A(gnd)" -> electrons and states; '+' inside parentheses is part of a
// state (Parse.C:197-245).
static void parseSide(const std::string& side, double& electrons, std::vector<RawState>& states,
                      std::vector<double>& stoi) {
    std::string cur;
    bool inParen = false;
    auto flush = [&]() {
        double c = stoiCoeff(cur);
        if (cur == "e" || cur == "E") electrons += c;
        else if (!cur.empty()) {
            const RawState r = parseState(cur);
            bool merged = false;
            for (size_t j = 0; j < states.size(); ++j)
                if (states[j] == r) { stoi[j] += c; merged = true; break; }
            if (!merged) { states.push_back(r); stoi.push_back(c); }
        }
        cur.clear();
    };
    for (size_t i = 0; i < side.size(); ++i) {
        const char ch = side[i];
        if (ch == '+' && !inParen) { flush(); continue; }
        if (ch == '(') inParen = true;
        else if (ch == ')') inParen = false;
        cur.push_back(ch);
    }
    flush();
}

static std::string noSpaces(const std::string& s) {
    std::string o;
    for (const std::string& t : splitSpaces(s)) o += t;
    return o;
}

static LXCatEntry parseEntry(const std::string& descriptionLine, double threshold, std::vector<double>&& e,
                             std::vector<double>&& s) {
    LXCatEntry en;
    const std::string line = noSpaces(descriptionLine);
    const size_t b = line.find_first_of('['), c = line.find_last_of(']');
    if (b == std::string::npos || c == std::string::npos || c <= b)
        throw InputError("Error in the parsing of one LXCat file. Check the description line:\n" + descriptionLine);
    std::vector<std::string> parts = splitChars(line.substr(b + 1, c - b - 1), ",");
    if (parts.empty()) throw InputError("Error in the parsing of one LXCat file:\n" + descriptionLine);
    std::string last = parts.back();
    std::string lower = last;
    for (char& ch : lower) ch = (char)std::tolower((unsigned char)ch);
    if (lower == "momentum-transfer" || lower == "integral") {
        en.subtype = lower;
        parts.pop_back();
        last = parts.back();
    } else {
        en.subtype = (last == "Elastic" || last == "Effective") ? "momentum-transfer" : "integral";
    }
    static const char* kTypes[] = {"Elastic", "Effective", "Attachment", "Ionization", "Excitation", "Vibrational", "Rotational"};
    bool ok = false;
    for (const char* t : kTypes) ok |= last == t;
    if (!ok)
        throw InputError("Error in the parsing of one LXCat file. Invalid type of collision in the following line: \n" +
                         descriptionLine + "\nValid types are: Elastic, Effective, Attachment, Ionization, Excitation, Vibrational and Rotational.");
    en.type = last;
    parts.pop_back();
    std::string coll;
    for (const std::string& p : parts) coll += "," + p;
    if (!coll.empty()) coll.erase(0, 1);
    std::string left, right;
    size_t k;
    if ((k = coll.find("<->")) != std::string::npos) {
        en.isReverse = true;
        left = coll.substr(0, k);
        right = coll.substr(k + 3);
    } else if ((k = coll.find("->")) != std::string::npos) {
        left = coll.substr(0, k);
        right = coll.substr(k + 2);
    } else {
        throw InputError("Error in the parsing of one LXCat file. Check the direction of the collision presented in the following line: \n" + descriptionLine);
    }
    double eL = 0, eR = 0;
    std::vector<RawState> targets;
    std::vector<double> tStoi;
    parseSide(left, eL, targets, tStoi);
    if (targets.empty())
        throw InputError("Could not find a target in the collision presented in the following LXCat line:\n" + descriptionLine +
                         "\n Please check your LXCat files.");
    parseSide(right, eR, en.products, en.productStoi);
    en.target = targets[0];
    en.threshold = threshold;
    en.energy = std::move(e);
    en.sigma = std::move(s);
    return en;
}

std::vector<LXCatEntry> readLXCat(const std::vector<std::string>& files, const std::string& inputDir) {
    std::vector<LXCatEntry> out;
    for (const std::string& name : files) {
        std::ifstream in(inputDir + "/" + name);
        if (!in) throw InputError("The file '" + name + "' could not be opened.");
        std::string line;
        while (std::getline(in, line)) {
            if (line.find("PROCESS:") == std::string::npos) continue;
            const std::string processLine = line;
            double threshold = 0;
            std::getline(in, line);
            const size_t b = line.find("E ="), e = line.find("eV");
            if (b != std::string::npos && e != std::string::npos) {
                threshold = evalNumber(line.substr(b + 3, e - b - 3));
            } else if (processLine.find("Elastic") == std::string::npos && processLine.find("Effective") == std::string::npos) {
                throw InputError("Error! Could not find a threshold for the following process at file '" + name + "':\n" +
                                 processLine + "\nThe threshold must be in the line immediately after the definition of the process.");
            }
            std::string description;
            std::getline(in, description);
            while (std::getline(in, line))
                if (line.find("-----") != std::string::npos) break;
            std::vector<double> en, sg;
            while (std::getline(in, line) && line.find("-----") == std::string::npos) {
                const std::vector<std::string> tok = splitSpaces(line);
                if (tok.empty()) continue;
                if (tok.size() != 2)
                    throw InputError("Error when reading the cross section values of the LXCat collision described in the following line:\n" +
                                     description + "Check the LXCat file '" + name + "'");
                en.push_back(evalNumber(tok[0]));
                sg.push_back(evalNumber(tok[1]));
            }
            if (en.empty())
                throw InputError("Error when reading the cross section values of the LXCat collision described in the following line:\n" +
                                 description + "Check the LXCat file '" + name + "'");
            out.push_back(parseEntry(description, threshold, std::move(en), std::move(sg)));
        }
    }
    return out;
}

// ------------------------------------------------------------- mixture
namespace {

constexpr double kNonDef = -123456789.0;          // Constant::NON_DEF
constexpr double kElectronCharge = 1.6021766208e-19;  // Constant.h:7
constexpr double kBoltzmannJ = 1.38064852e-23;        // Constant.h:6
constexpr double kPlanck = 6.626070040e-34;           // Constant.h:13
constexpr double kPi = 3.14159265358979323846;
const double kBoltzmannInEV = kBoltzmannJ / kElectronCharge;        // Constant.h:19
const double kPlanckReducedInEV = kPlanck / (2.0 * kPi * kElectronCharge);  // Constant.h:20
constexpr double kEps = 2.220446049250313e-16;

// A state with its family (State.h:20-75, EedfState.C:8-95).
struct State {
    RawState raw;
    int gas = -1;
    std::string type;  // ele, vib, rot, ion
    int parent = -1;
    std::vector<int> siblings, children;
    double energy = kNonDef, statWeight = kNonDef, population = 0, density = 0;
    bool isTarget = false;
    std::vector<int> colls;  // collisions with this state as target or reverse product
};

struct XSec {
    std::vector<double> e, s;
    bool defined() const { return !e.empty(); }
};

// Collision.C:18-86
struct Collision {
    std::string type;
    int target = -1;
    std::vector<int> products;
    std::vector<double> stoi;
    bool isReverse = false, extra = false;
    double threshold = 0;
    XSec integral, momTransf;
    std::string scattering = "isotropic";
    std::vector<double> params;
};

struct Gas {
    std::string name;
    double mass = kNonDef, harmonicFrequency = kNonDef, anharmonicFrequency = kNonDef, rotationalConstant = kNonDef;
    double fraction = 0, opb = kNonDef;
    std::vector<int> states, colls;       // colls: non-extra, in creation order
    std::map<int, double> effPop;         // populations for an Elastic derived from an Effective (by state)
    bool effPopDefined = false;
};

int scatteringCode(const std::string& n) {
    if (n == "isotropic") return EKMC_SCAT_ISOTROPIC;
    if (n == "forward") return EKMC_SCAT_FORWARD;
    if (n == "bornDipole") return EKMC_SCAT_BORN_DIPOLE;
    if (n == "surendra") return EKMC_SCAT_SURENDRA;
    if (n == "coulombScreen") return EKMC_SCAT_COULOMB_SCREEN;
    if (n == "momentumConservationIonization") return EKMC_SCAT_MOMENTUM_CONSERVATION_IONIZATION;
    throw InputError("The angular scattering function '" + n + "' is not defined.");
}

int requiredParams(const std::string& n) { return n == "coulombScreen" ? 2 : 0; }

void warn(const std::string& msg) { std::fprintf(stderr, "%s\n", msg.c_str()); }

struct Model {
    std::vector<Gas> gases;
    std::vector<State> states;
    std::vector<Collision> colls;

    int gasId(const std::string& g) {
        for (size_t i = 0; i < gases.size(); ++i)
            if (gases[i].name == g) return (int)i;
        Gas x;
        x.name = g;
        gases.push_back(x);
        return (int)gases.size() - 1;
    }
    int findExact(const RawState& r) const {
        for (size_t i = 0; i < states.size(); ++i)
            if (states[i].raw == r) return (int)i;
        return -1;
    }
    // State::add with EedfState::addFamily: a new state's relatives among the
    // gas's existing states, scanned in creation order.
    int addState(const RawState& r) {
        const int found = findExact(r);
        if (found >= 0) return found;
        const int g = gasId(r.gas);
        State s;
        s.raw = r;
        s.gas = g;
        s.type = !r.ionCharg.empty() ? "ion" : (!r.rotLevel.empty() ? "rot" : (!r.vibLevel.empty() ? "vib" : "ele"));
        const int id = (int)states.size();
        for (int o : gases[g].states) {
            State& t = states[o];
            const bool sameEle = t.raw.eleLevel == r.eleLevel, sameVib = t.raw.vibLevel == r.vibLevel;
            if (s.type == "rot") {
                if (t.type == "vib" && sameEle && sameVib) { s.parent = o; t.children.push_back(id); }
                else if (t.type == "rot" && sameEle && sameVib) { s.siblings.push_back(o); t.siblings.push_back(id); }
            } else if (s.type == "vib") {
                if (t.type == "ele" && sameEle) { s.parent = o; t.children.push_back(id); }
                else if (t.type == "vib" && sameEle) { s.siblings.push_back(o); t.siblings.push_back(id); }
                else if (t.type == "rot" && sameEle && sameVib) { s.children.push_back(o); t.parent = id; }
            } else if (s.type == "ele") {
                if (t.type == "ele") { s.siblings.push_back(o); t.siblings.push_back(id); }
                else if (t.type == "vib" && sameEle) { s.children.push_back(o); t.parent = id; }
            } else if (t.type == "ion") {
                s.siblings.push_back(o);
                t.siblings.push_back(id);
            }
        }
        states.push_back(s);
        gases[g].states.push_back(id);
        return id;
    }
    // State::fixOrphanStates: parents of rotational, then vibrational states
    void fixOrphans() {
        for (size_t i = 0; i < states.size(); ++i)
            if (states[i].type == "rot" && states[i].parent < 0) {
                RawState p = states[i].raw;
                p.rotLevel.clear();
                addState(p);
            }
        for (size_t i = 0; i < states.size(); ++i)
            if (states[i].type == "vib" && states[i].parent < 0) {
                RawState p = states[i].raw;
                p.vibLevel.clear();
                p.rotLevel.clear();
                addState(p);
            }
    }
    // State::findPointer: '*' as the electronic level takes every electronic
    // state of the gas, as the vibrational level the children of the
    // electronic state, as the rotational level the children of the
    // vibrational state.
    std::vector<int> find(const RawState& r) const {
        std::vector<int> out;
        auto gasOf = [&](const State& s) { return gases[s.gas].name; };
        if (r.eleLevel == "*") {
            for (size_t i = 0; i < states.size(); ++i)
                if (gasOf(states[i]) == r.gas && states[i].type == "ele") {
                    out.push_back((int)i);
                    out.insert(out.end(), states[i].siblings.begin(), states[i].siblings.end());
                    break;
                }
        } else if (r.vibLevel == "*") {
            for (const State& s : states)
                if (gasOf(s) == r.gas && s.raw.eleLevel == r.eleLevel && s.type == "ele") { out = s.children; break; }
        } else if (r.rotLevel == "*") {
            for (const State& s : states)
                if (gasOf(s) == r.gas && s.raw.eleLevel == r.eleLevel && s.raw.vibLevel == r.vibLevel && s.type == "vib") {
                    out = s.children;
                    break;
                }
        } else {
            const int i = findExact(r);
            if (i >= 0) out.push_back(i);
        }
        return out;
    }
    std::string description(const Collision& c) const {  // Collision.C:88-122
        std::string s = "e+" + states[c.target].raw.name();
        s += c.isReverse ? "<->" : "->";
        if (c.type == "Ionization") s += "e+e+";
        else if (c.type != "Attachment") s += "e+";
        for (size_t i = 0; i < c.products.size(); ++i) {
            if (c.stoi[i] > 1) s += std::to_string((int)c.stoi[i]);
            s += states[c.products[i]].raw.name();
            if (i + 1 < c.products.size()) s += "+";
        }
        return s + "," + c.type;
    }
    // Collision::find: same type, threshold, direction and products among the
    // target's collisions.
    int findCollision(const Collision& c) const {
        for (int k : states[c.target].colls) {
            const Collision& o = colls[k];
            if (o.threshold != c.threshold || o.type != c.type || o.isReverse != c.isReverse ||
                o.products.size() != c.products.size())
                continue;
            bool equal = true;
            for (size_t j = 0; j < o.products.size() && equal; ++j) {
                bool any = false;
                for (size_t q = 0; q < c.products.size(); ++q)
                    if (o.products[j] == c.products[q] && o.stoi[j] == c.stoi[q]) { any = true; break; }
                equal = any;
            }
            if (equal) return k;
        }
        return -1;
    }
    // Collision::add: a collision met again supplies its missing cross section
    // (integral or momentum-transfer) or is dropped as a duplicate.
    void addCollision(Collision c) {
        const int k = findCollision(c);
        if (k >= 0) {
            Collision& o = colls[k];
            if (!o.integral.defined() && c.integral.defined()) o.integral = c.integral;
            else if (!o.momTransf.defined() && c.momTransf.defined()) o.momTransf = c.momTransf;
            else warn("Warning! Avoiding duplicated electron impact collision:\n\t " + description(o));
            return;
        }
        auto check = [&](XSec& x) {
            for (size_t i = 0; i + 1 < x.e.size(); ++i)
                if (x.e[i] >= x.e[i + 1])
                    throw InputError("The energy column of the cross section corresponding to the collision shown below is not strictly increasing.\n" +
                                     description(c) + "\nPlease fix this in the LXCat files.");
            if ((c.type == "Elastic" || c.type == "Effective") && x.e[0] != 0) {
                x.e.insert(x.e.begin(), 0.0);
                x.s.insert(x.s.begin(), x.s[0]);
            }
        };
        if (c.integral.defined()) check(c.integral);
        if (c.momTransf.defined()) check(c.momTransf);
        const int id = (int)colls.size();
        colls.push_back(c);
        if (!c.extra) {
            gases[states[c.target].gas].colls.push_back(id);
            states[c.target].colls.push_back(id);
            states[c.target].isTarget = true;
        }
        if (c.isReverse) {
            if (c.products.size() == 1 && c.stoi[0] == 1) {
                if (!c.extra) {
                    states[c.products[0]].colls.push_back(id);
                    states[c.products[0]].isTarget = true;
                }
            } else {
                throw InputError("Error while creating collision '" + description(c) +
                                 "'. Klein-Rosseland microreversibility relation valid only for binary collisions.\n");
            }
        }
        if ((c.type == "Effective" || c.type == "Elastic") && states[c.target].type != "ele")
            throw InputError("Found ''" + c.type + "'' collision with ''" + states[c.target].raw.name() + "'' as target [" +
                             description(c) + "].\n" + c.type +
                             " collisions are only allowed for electronic states. Please check LXCat files.\n");
    }
    void addEntries(const std::vector<LXCatEntry>& entries, bool extra) {  // Setup.h:292-349
        for (const LXCatEntry& e : entries) {
            Collision c;
            c.type = e.type;
            c.target = addState(e.target);
            for (const RawState& p : e.products) c.products.push_back(addState(p));
            c.stoi = e.productStoi;
            c.isReverse = e.isReverse;
            c.threshold = e.threshold;
            c.extra = extra;
            XSec x;
            x.e = e.energy;
            x.s = e.sigma;
            if (e.subtype == "momentum-transfer") c.momTransf = x;
            else c.integral = x;
            addCollision(c);
        }
    }
};

// Linear interpolation as gsl_interp_linear evaluates it: the interval with
// x[i] <= v < x[i+1], the last point in the last interval.
double linearAt(const std::vector<double>& x, const std::vector<double>& y, double v) {
    if (x.size() < 2) return x.size() == 1 && v == x[0] ? y[0] : 0.0;
    size_t lo = 0, hi = x.size() - 1;
    while (hi > lo + 1) {
        const size_t mid = (lo + hi) / 2;
        if (x[mid] > v) hi = mid; else lo = mid;
    }
    const double dx = x[lo + 1] - x[lo];
    return dx > 0.0 ? y[lo] + (v - x[lo]) / dx * (y[lo + 1] - y[lo]) : 0.0;
}

// Collision::interpolatedCrossSection (Collision.C:231-289): zero at or below
// the threshold of an inelastic collision and outside the table.
std::vector<double> interpolated(const Collision& c, bool momTransf, const std::vector<double>& energies) {
    std::vector<double> out(energies.size(), 0.0);
    size_t minIndex = 0;
    if (c.type != "Effective" && c.type != "Elastic") {
        minIndex = energies.size();
        for (size_t i = 0; i < energies.size(); ++i)
            if (energies[i] > c.threshold) { minIndex = i; break; }
    }
    const XSec& x = momTransf ? c.momTransf : c.integral;
    for (size_t i = minIndex; i < energies.size(); ++i) {
        const double v = energies[i];
        out[i] = (v >= x.e.front() && v <= x.e.back()) ? linearAt(x.e, x.s, v) : 0.0;
    }
    return out;
}

// Collision::superElasticCrossSection (Collision.C:291-353), Klein-Rosseland.
std::vector<double> superElastic(const Model& M, const Collision& c, bool momTransf, const std::vector<double>& energies) {
    const State& t = M.states[c.target];
    const State& p = M.states[c.products[0]];
    if (!c.isReverse) throw InputError("Collision '" + M.description(c) + " is not defined as bidirectional.\n");
    if (t.statWeight == kNonDef)
        throw InputError("The statistical weight of the state '" + t.raw.name() + "' is not defined.\nThe super elastic cross section of the collision '" +
                         M.description(c) + " cannot be evaluated.\n");
    if (p.statWeight == kNonDef)
        throw InputError("The statistical weight of the state '" + p.raw.name() + "' is not defined.\nThe super elastic cross section of the collision '" +
                         M.description(c) + " cannot be evaluated.\n");
    std::vector<double> out(energies.size(), 0.0);
    if (energies.empty()) return out;
    size_t minIndex = 0;
    if (energies[0] == 0) {
        if (energies.size() == 1) return out;
        minIndex = 1;
    }
    std::vector<double> shifted(energies.size());
    for (size_t i = 0; i < energies.size(); ++i) shifted[i] = energies[i] + c.threshold;
    std::vector<double> in = interpolated(c, momTransf, shifted);
    const double ratio = t.statWeight / p.statWeight;
    for (double& v : in) v *= ratio;
    for (size_t i = minIndex; i < energies.size(); ++i) out[i] = (1.0 + c.threshold / energies[i]) * in[i];
    return out;
}

// AngularDistributionFunctions.h:32-105: I(energy, angle), normalized as
// 2 pi Int_0^pi I sin(angle) d(angle) = 1.
double angularPdf(const Collision& c, double energy, double angle) {
    const double iso = 1.0 / (4.0 * kPi);
    if (c.scattering == "bornDipole") {
        const double after = energy - c.threshold;
        if (after <= 0) return iso;
        const double q2 = after + energy - 2.0 * std::sqrt(energy * after) * std::cos(angle);
        const double se = std::sqrt(energy), sa = std::sqrt(after);
        return iso * se * sa / q2 / std::log((sa + se) / std::sqrt(c.threshold));
    }
    if (c.scattering == "surendra") {
        if (energy == 0) return iso;
        return energy / (4.0 * kPi * (1.0 + energy * std::pow(std::sin(angle / 2.0), 2)) * std::log(1.0 + energy));
    }
    if (c.scattering == "coulombScreen") {
        const double eps = c.params[0] == 0 ? energy : energy - c.threshold;
        if (eps <= 0) return iso;
        const double sp = c.params[1] / eps;
        return (sp * (sp + 1.0)) / kPi / std::pow(2.0 * sp + 1.0 - std::cos(angle), 2);
    }
    if (c.scattering == "forward") return angle == 0 ? 1.0 : 0.0;
    return iso;  // isotropic, momentumConservationIonization
}

// Collision::evaluateRawMomTransfCrossSection (Collision.C:124-196): the
// missing one of the integral and momentum-transfer cross sections, from the
// angular distribution (trapezoidal rule over `angleNumber` angles).
void completeCrossSections(const Model& M, Collision& c, double angleNumber) {
    const bool hasI = c.integral.defined(), hasM = c.momTransf.defined();
    if (hasI && hasM) return;
    if (c.scattering == "forward") {
        if (!hasI)
            throw InputError("When choosing the ''forward'' angularScatteringType for the collision\n" + M.description(c) +
                             "\nthe integralCrossSection must be defined in the LXCat files, and not the momentum-transfer!");
        c.momTransf.e = c.integral.e;
        c.momTransf.s.assign(c.integral.e.size(), 0.0);
        return;
    }
    if (c.scattering == "isotropic") {
        if (hasI) c.momTransf = c.integral;
        else c.integral = c.momTransf;
        return;
    }
    const int n = (int)angleNumber;
    if (n < 2) throw InputError("electronKinetics.anisotropicScattering.angleNumber must be at least 2.");
    const double step = kPi / (angleNumber - 1.0);
    std::vector<double> angles(n), aux(n);
    for (int j = 0; j < n; ++j) {
        angles[j] = j == n - 1 ? kPi : (double)j * step;
        aux[j] = std::cos(angles[j]) * std::sin(angles[j]);
    }
    XSec& from = hasI ? c.integral : c.momTransf;
    XSec& to = hasI ? c.momTransf : c.integral;
    to.e = from.e;
    to.s.assign(from.e.size(), 0.0);
    for (size_t i = 0; i < from.e.size(); ++i) {
        const double energy = from.e[i];
        double a = aux[0] * angularPdf(c, energy, angles[0]);
        for (int j = 1; j < n - 1; ++j) a += 2.0 * aux[j] * angularPdf(c, energy, angles[j]);
        a += aux[n - 1] * angularPdf(c, energy, angles[n - 1]);
        a *= kPi * step;
        to.s[i] = hasI ? from.s[i] * (1.0 - a) : from.s[i] / (1.0 - a);
    }
}

// ----- state property functions (StatePropertyFunctions.h:22-866)

double temperatureArgument(const std::vector<std::string>& strArgs, const std::vector<double>& args, size_t i,
                           double gasTemperature) {
    if (strArgs[i] == "gasTemperature") return gasTemperature;
    if (strArgs[i] == "electronTemperature")
        throw InputError("Property functions of the electron temperature are not supported by this program (the Monte Carlo "
                         "electron kinetics has no prescribed electron temperature).");
    return args[i];
}

double levelNumber(const std::string& s) { return std::stod(s); }

void boltzmannLike(Model& M, const std::vector<int>& ids, double T, double cutoff, int cutoffLevel) {
    double ground = 1e20, norm = 0;
    for (int i : ids) ground = std::fmin(ground, M.states[i].energy);
    for (int i : ids) {
        State& s = M.states[i];
        const double level = cutoffLevel == 1 ? levelNumber(s.raw.vibLevel) : (cutoffLevel == 2 ? levelNumber(s.raw.rotLevel) : 0);
        if (cutoffLevel != 0 && level > cutoff) {
            s.population = 0;
        } else {
            s.population = s.statWeight * std::exp(-(s.energy - ground) / (kBoltzmannInEV * T));
            norm += s.population;
        }
    }
    for (int i : ids) M.states[i].population /= norm;
}

void treanor(Model& M, const std::vector<int>& ids, double T0, double T1, bool gordiets, const std::string& fname) {
    double groundE = 0, firstE = 0;
    bool g0 = false, g1 = false;
    for (int i : ids) {
        const State& s = M.states[i];
        if (s.raw.vibLevel == "0") { groundE = s.energy; g0 = true; }
        else if (s.raw.vibLevel == "1") { firstE = s.energy; g1 = true; }
    }
    if (!g0 || !g1)
        throw InputError("Unable to find groundEnergy or firstEnergy to populate state" + M.states[ids[0]].raw.name() +
                         " and its siblings with function " + fname + ". \nCheck input file.");
    double norm = 0;
    for (int i : ids) {
        State& s = M.states[i];
        const double v = levelNumber(s.raw.vibLevel);
        s.population = s.statWeight * std::exp(-1.0 / kBoltzmannInEV *
                                               (v * (firstE - groundE) * (1.0 / T1 - 1.0 / T0) + (s.energy - groundE) / T0));
        norm += s.population;
    }
    for (int i : ids) M.states[i].population /= norm;
    if (!gordiets) return;
    const double anh = M.gases[M.states[ids[0]].gas].anharmonicFrequency;
    const double vLimit = std::floor(0.5 * (1.0 + (firstE - groundE) * T0 / (kPlanckReducedInEV * anh * T1)));
    int atLimit = -1;
    for (size_t k = 0; k < ids.size(); ++k)
        if (levelNumber(M.states[ids[k]].raw.vibLevel) == vLimit) atLimit = (int)k;
    if (atLimit < 0)
        throw InputError("treanorGordietsPopulation of " + M.states[ids[0]].raw.name() +
                         ": no vibrational state at the Treanor minimum v = " + std::to_string((int)vLimit) + ".");
    std::vector<double> tg(ids.size(), 0.0);
    norm = 0;
    for (size_t k = 0; k < ids.size(); ++k) {
        const double v = levelNumber(M.states[ids[k]].raw.vibLevel);
        tg[k] = v <= vLimit ? M.states[ids[k]].population : M.states[ids[atLimit]].population * vLimit / v;
        norm += tg[k];
    }
    for (size_t k = 0; k < ids.size(); ++k) M.states[ids[k]].population = tg[k] / norm;
}

void evaluateStateProperty(Model& M, const std::string& property, const std::vector<int>& ids, const std::string& value,
                           const std::map<std::string, std::string>& workCond, double gasTemperature) {
    const std::string func = splitChars(value, "@").empty() ? value : splitChars(value, "@")[0];
    std::vector<double> args;
    std::vector<std::string> strArgs;
    if (value.find('@') != std::string::npos) {  // Setup.h:561-574
        const std::vector<std::string> parts = splitChars(value, "@");
        if (parts.size() > 1) strArgs = splitChars(parts[1], ",");
        for (const std::string& a : strArgs) {
            if (a == "gasTemperature") args.push_back(gasTemperature);
            else if (workCond.count(a)) args.push_back(evalNumber(workCond.at(a)));
            else args.push_back(evalNumber(a));
        }
    }
    State& first = M.states[ids[0]];
    const Gas& gas = M.gases[first.gas];
    auto need = [&](bool ok, const std::string& msg) { if (!ok) throw InputError(msg); };
    auto noArgs = [&](const std::string& f) {
        need(args.empty(), "Wrong number of arguments when evaluating " + f + " function. \nCheck input file");
    };
    auto forProperty = [&](const std::string& f, const char* prop) {
        need(property == prop, "Trying to use " + f + " function to set up property " + property + ". \nCheck input file");
    };
    auto checkEnergiesWeights = [&](const std::string& f) {
        for (int i : ids) {
            const State& s = M.states[i];
            need(s.energy != kNonDef, "Unable to find " + s.raw.name() + " energy for the evaluation of " + f + " function. \nCheck input file");
            need(s.statWeight != kNonDef, "Unable to find " + s.raw.name() + " statistical weight for the evaluation of " + f +
                                               " function. \nCheck input file\n");
        }
    };
    if (func.rfind("rotationalDegeneracy", 0) == 0 &&
        (func == "rotationalDegeneracy" || func == "rotationalDegeneracy_H2" || func == "rotationalDegeneracy_N2" ||
         func == "rotationalDegeneracy_H2O" || func == "rotationalDegeneracy_NO")) {
        forProperty(func, "statisticalWeight");
        noArgs(func);
        need(first.type == "rot", "Trying to assign " + func + " to non rotational state " + first.raw.name() + ". \nCheck input file");
        for (int i : ids) {
            State& s = M.states[i];
            if (func == "rotationalDegeneracy_H2O") {
                const double J = std::stod(s.raw.rotLevel.substr(0, 1)), Ka = std::stod(s.raw.rotLevel.substr(1, 1)),
                             Kc = std::stod(s.raw.rotLevel.substr(2, 1));
                double w = 2.0 * J + 1;
                if (std::fmod(std::fabs(Ka - Kc), 2) != 0) w *= 3.0;
                s.statWeight = w;
                continue;
            }
            const double J = levelNumber(s.raw.rotLevel);
            if (func == "rotationalDegeneracy") s.statWeight = 2.0 * J + 1.0;
            else if (func == "rotationalDegeneracy_H2") s.statWeight = (2.0 - std::pow(-1, J)) * (2.0 * J + 1.0);
            else if (func == "rotationalDegeneracy_N2") s.statWeight = 3.0 * (1.0 + 0.5 * (1.0 + std::pow(-1, J))) * (2.0 * J + 1.0);
            else s.statWeight = 3.0 * (2.0 * J + 1.0);  // NO
        }
    } else if (func == "harmonicOscillatorEnergy" || func == "morseOscillatorEnergy") {
        forProperty(func, "energy");
        noArgs(func);
        need(first.type == "vib", "Trying to assign " + func + " to non vibrational state " + first.raw.name() + ".\nCheck input file.");
        need(gas.harmonicFrequency != kNonDef, "Unable to find harmonicFrequency to evaluate the energy of the state " +
                                                   first.raw.name() + " with function " + func + ".\nCheck input file");
        if (func == "morseOscillatorEnergy")
            need(gas.anharmonicFrequency != kNonDef, "Unable to find anharmonicFrequency to evaluate the energy of the state " +
                                                         first.raw.name() + " with function morseOscillatorEnergy.\nCheck input file");
        for (int i : ids) {
            State& s = M.states[i];
            const double v = levelNumber(s.raw.vibLevel);
            const Gas& sg = M.gases[s.gas];
            s.energy = func == "harmonicOscillatorEnergy"
                           ? kPlanckReducedInEV * sg.harmonicFrequency * (v + 0.5)
                           : kPlanckReducedInEV * (sg.harmonicFrequency * (v + 0.5) - sg.anharmonicFrequency * std::pow((v + 0.5), 2));
        }
    } else if (func == "rigidRotorEnergy" || func == "rigidRotorEnergy_NO") {
        forProperty(func, "energy");
        noArgs(func);
        need(first.type == "rot", "Trying to assign " + func + " to non rotational state " + first.raw.name() + "\nCheck input file.");
        if (func == "rigidRotorEnergy")
            need(gas.rotationalConstant != kNonDef, "Unable to find rotationalConstant to evaluate the energy of the state " +
                                                        first.raw.name() + " with function rigidRotorEnergy.\nCheck input file");
        const bool x12 = first.raw.eleLevel == "X_1/2";
        for (int i : ids) {
            State& s = M.states[i];
            const double J = levelNumber(s.raw.rotLevel);
            if (func == "rigidRotorEnergy") {
                s.energy = M.gases[s.gas].rotationalConstant * J * (J + 1.0);
            } else {
                const double B = x12 ? 0.2073e-3 : 0.2133e-3, omega = x12 ? 0.5 : 1.5;
                s.energy = B * (J * (J + 1.0) - omega * omega);
            }
        }
    } else if (func == "boltzmannPopulation" || func == "boltzmannPopulationVibrationalCutoff" ||
               func == "boltzmannPopulationRotationalCutoff") {
        forProperty(func, "population");
        const int cut = func == "boltzmannPopulation" ? 0 : (func == "boltzmannPopulationVibrationalCutoff" ? 1 : 2);
        need(args.size() == (cut == 0 ? 1u : 2u), "Wrong number of arguments when evaluating " + func + " function. \nCheck input file");
        if (cut == 1) need(first.type == "vib", "Trying to assign " + func + " to non vibrational state" + first.raw.name() + ". \nCheck input file");
        if (cut == 2) need(first.type == "rot", "Trying to assign " + func + " to non rotational state" + first.raw.name() + ". \nCheck input file");
        if (cut != 0)
            need(std::fmod(args[1], 1) == 0 && args[1] >= 0,
                 "Error found when evaluating population of state " + first.raw.name() + ".\nMaximum " +
                     (cut == 1 ? "vibrational" : "rotational") + " level ''" + std::to_string(args[1]) +
                     "should be a non-negative integer.\nCheck input file");
        checkEnergiesWeights(func);
        boltzmannLike(M, ids, temperatureArgument(strArgs, args, 0, gasTemperature), cut ? args[1] : 0, cut);
    } else if (func == "treanorPopulation" || func == "treanorGordietsPopulation") {
        forProperty(func, "population");
        need(args.size() == 2, "Wrong number of arguments when evaluating " + func + " function. \nCheck input file");
        if (func == "treanorGordietsPopulation")
            need(gas.anharmonicFrequency != kNonDef, "Unable to find " + gas.name + " anharmonicFrequency to populate" +
                                                         first.raw.name() + " and its siblings with function treanorGordietsPopulation. \nCheck input file\n");
        need(first.type == "vib", "Trying to assign " + func + " to non vibrational state" + first.raw.name() + ". \nCheck input file");
        checkEnergiesWeights(func);
        treanor(M, ids, temperatureArgument(strArgs, args, 0, gasTemperature), temperatureArgument(strArgs, args, 1, gasTemperature),
                func == "treanorGordietsPopulation", func);
    } else if (value.find('@') != std::string::npos) {
        throw InputError("Error! Trying to use the property function '" + func + "' which is not defined in the code.\n");
    } else {
        const double v = evalNumber(value);
        for (int i : ids) {
            State& s = M.states[i];
            if (property == "energy") s.energy = v;
            else if (property == "statisticalWeight") s.statWeight = v;
            else s.population = v;
        }
    }
}

// EedfGas::checkPopulationNorms (EedfGas.C:29-101)
void checkPopulationNorms(const Model& M, const Gas& g) {
    if (g.fraction == 0) return;
    double gasNorm = 0;
    bool eleDone = false, ionDone = false;
    auto err = [](double x) { return std::fabs(x - 1) > 10 * kEps; };
    for (int si : g.states) {
        const State& st = M.states[si];
        if (st.type == "ele" && !eleDone) {
            std::vector<int> group(1, si);
            group.insert(group.end(), st.siblings.begin(), st.siblings.end());
            for (int ei : group) {
                const State& e = M.states[ei];
                if (!(e.isTarget && e.population != 0)) continue;
                gasNorm += e.population;
                if (e.children.empty()) continue;
                double vibNorm = 0;
                bool anyVib = false;
                for (int vi : e.children) {
                    const State& v = M.states[vi];
                    if (!(v.isTarget && v.population != 0)) continue;
                    anyVib = true;
                    vibNorm += v.population;
                    if (v.children.empty()) continue;
                    double rotNorm = 0;
                    for (int ri : v.children) {
                        const State& r = M.states[ri];
                        if (r.isTarget && r.population != 0) rotNorm += r.population;
                    }
                    if (err(rotNorm))
                        throw InputError("Rotational distribution " + v.raw.name().substr(0, v.raw.name().size() - 1) +
                                         ",J=*) is not properly normalized. (Error = " + std::to_string(rotNorm - 1) + ")\n");
                }
                if (anyVib && err(vibNorm))
                    throw InputError("Vibrational distribution " + e.raw.name().substr(0, e.raw.name().size() - 1) +
                                     ",v=*) is not properly normalized. (Error = " + std::to_string(vibNorm - 1) + ")\n");
            }
            eleDone = true;
        }
        if (st.type == "ion" && !ionDone) {
            std::vector<int> group(1, si);
            group.insert(group.end(), st.siblings.begin(), st.siblings.end());
            for (int ii : group)
                if (M.states[ii].population != 0) gasNorm += M.states[ii].population;
            ionDone = true;
        }
    }
    if (err(gasNorm))
        throw InputError("Electronic/ionic distribution " + g.name + "(*) is not properly normalized. (Error = " +
                         std::to_string(gasNorm - 1) + ")\n");
}

// EedfGas::elasticFromEffectiveCrossSection (EedfGas.C:166-301): the
// Effective cross section minus every other collision of the gas weighted by
// the population of its target (default: ground electronic state, Boltzmann
// vibrational and rotational distributions at 300 K).
XSec elasticFromEffective(Model& M, int gi) {
    Gas& g = M.gases[gi];
    int eff = -1;
    for (int k : g.colls)
        if (M.colls[k].type == "Effective") { eff = k; break; }
    if (eff < 0)
        throw InputError("Gas ''" + g.name + "'' does not have an ''Effective'' collision defined, so ''Elastic'' collisions cannot be evaluated from it.\nPlease, check the corresponding LXCat file.");
    XSec raw = M.colls[eff].momTransf;
    if (!g.effPopDefined) {
        g.effPopDefined = true;
        const int ground = M.colls[eff].target;
        g.effPop[ground] = 1;
        const State& gs = M.states[ground];
        int vibGround = -1;
        auto check = [&](const State& s) {
            if (s.energy == kNonDef)
                throw InputError("Unable to find " + s.raw.name() + " energy for the evaluation of ''Elastic'' cross section of " + g.name + ".\nCheck input file");
            if (s.statWeight == kNonDef)
                throw InputError("Unable to find " + s.raw.name() + " statistical weight for the evaluation of ''Elastic'' cross section of " + g.name + ".\nCheck input file");
        };
        if (!gs.children.empty()) {
            double norm = 0;
            vibGround = gs.children[0];
            for (int c : gs.children) {
                const State& s = M.states[c];
                check(s);
                if (s.energy < M.states[vibGround].energy) vibGround = c;
                g.effPop[c] = s.statWeight * std::exp(-s.energy / (kBoltzmannInEV * 300.0));
                norm += g.effPop[c];
            }
            for (int c : gs.children) g.effPop[c] = g.effPop[c] / norm;
        }
        if (vibGround >= 0 && !M.states[vibGround].children.empty()) {
            double norm = 0;
            for (int c : M.states[vibGround].children) {
                const State& s = M.states[c];
                check(s);
                g.effPop[c] = s.statWeight * std::exp(-s.energy / (kBoltzmannInEV * 300.0));
                norm += g.effPop[c];
            }
            for (int c : M.states[vibGround].children) g.effPop[c] = g.effPop[vibGround] * g.effPop[c] / norm;
        }
    }
    auto pop = [&](int state) {
        const auto it = g.effPop.find(state);
        return it == g.effPop.end() ? 0.0 : it->second;
    };
    for (int k : g.colls) {
        const Collision& c = M.colls[k];
        if (c.type == "Effective" || c.type == "Elastic") continue;
        const std::vector<double> in = interpolated(c, true, raw.e);
        for (size_t i = 0; i < raw.s.size(); ++i) raw.s[i] -= pop(c.target) * in[i];
        if (c.isReverse) {
            const std::vector<double> sup = superElastic(M, c, true, raw.e);
            for (size_t i = 0; i < raw.s.size(); ++i) raw.s[i] -= pop(c.products[0]) * sup[i];
        }
    }
    bool negative = false;
    for (double& v : raw.s)
        if (v < 0) { negative = true; v = 0; }
    if (negative)
        warn("Negative values obtained when evaluating an Elastic cross section from an Effective one (" + g.name +
             ").\nNegative values have been clipped to 0 and unreliable results may be obtained.\nPlease, carefully check inputs and outputs of your simulation");
    return raw;
}

// EedfGas::checkElasticCollisions (EedfGas.C:120-164)
void checkElasticCollisions(Model& M, int gi) {
    const Gas& g = M.gases[gi];
    if (g.colls.empty()) return;
    for (int si : g.states) {
        if (M.states[si].type != "ele") continue;
        std::vector<int> group(1, si);
        group.insert(group.end(), M.states[si].siblings.begin(), M.states[si].siblings.end());
        for (int ei : group) {
            if (!M.states[ei].isTarget) continue;
            bool hasElastic = false;
            for (int k : M.states[ei].colls) hasElastic |= M.colls[k].type == "Elastic";
            if (hasElastic) continue;
            const XSec raw = elasticFromEffective(M, gi);
            Collision c;
            c.type = "Elastic";
            c.target = ei;
            c.products.assign(1, ei);
            c.stoi.assign(1, 1.0);
            c.integral = raw;
            c.momTransf = raw;
            M.addCollision(c);
        }
        break;
    }
}

}  // namespace

GasMixture buildMixture(const SetupFile& setup, const std::string& inputDir, double gasTemperature) {
    Model M;
    // gases, states and collisions of the LXCat files (Setup.h:292-349)
    const std::vector<LXCatEntry> entries = readLXCat(setup.childNames("electronKinetics.LXCatFiles"), inputDir);
    if (entries.empty()) throw InputError("No LXCat collisions found in electronKinetics.LXCatFiles.");
    M.addEntries(entries, false);
    if (setup.has("electronKinetics.LXCatFilesExtra"))
        M.addEntries(readLXCat(setup.childNames("electronKinetics.LXCatFilesExtra"), inputDir), true);
    M.fixOrphans();
    // gas properties (Setup.h:462-511); heat capacities and conductivities do
    // not enter the electron kinetics
    auto gasProperty = [&](const char* key, double Gas::*field) {
        for (const auto& kv : setup.propertyMap(std::string("electronKinetics.gasProperties.") + key, inputDir)) {
            for (Gas& g : M.gases) {
                if (g.name != kv.first) continue;
                if (kv.second.find('@') != std::string::npos)
                    throw InputError("Error! Trying to use the property function '" + splitChars(kv.second, "@")[0] +
                                     "' which is not defined in the code.\n");
                g.*field = evalNumber(kv.second);
            }
        }
    };
    gasProperty("mass", &Gas::mass);
    gasProperty("harmonicFrequency", &Gas::harmonicFrequency);
    gasProperty("anharmonicFrequency", &Gas::anharmonicFrequency);
    gasProperty("rotationalConstant", &Gas::rotationalConstant);
    gasProperty("fraction", &Gas::fraction);
    gasProperty("OPBParameter", &Gas::opb);
    {
        double norm = 0;
        for (const Gas& g : M.gases) norm += g.fraction;
        if (std::fabs(norm - 1) > 10 * kEps)
            throw InputError("Gas fractions are not properly normalized (Error = " + std::to_string(norm - 1) +
                             "). Please, check input file.\n");
    }
    // state properties (Setup.h:513-552), each map in key order
    std::map<std::string, std::string> workCond;
    for (const std::string& k : setup.childNames("workingConditions")) workCond[k] = setup.value("workingConditions." + k);
    for (const char* prop : {"energy", "statisticalWeight", "population"}) {
        for (const auto& kv : setup.propertyMap(std::string("electronKinetics.stateProperties.") + prop, inputDir)) {
            const std::vector<int> ids = M.find(parseState(kv.first));
            if (ids.empty()) continue;
            evaluateStateProperty(M, prop, ids, kv.second, workCond, gasTemperature);
        }
    }
    // densities (State.h:72-85)
    for (State& s : M.states) {
        const double f = M.gases[s.gas].fraction;
        if (s.type == "rot") s.density = s.population * M.states[s.parent].population * M.states[M.states[s.parent].parent].population * f;
        else if (s.type == "vib") s.density = s.population * M.states[s.parent].population * f;
        else s.density = s.population * f;
    }
    // populations for an Elastic cross section derived from an Effective one
    // (Setup.h:250-269)
    {
        std::map<std::string, std::string> eff;
        if (setup.has("electronKinetics.effectiveCrossSectionPopulations"))
            eff = setup.propertyMap("electronKinetics.effectiveCrossSectionPopulations", inputDir);
        if (!eff.empty())
            for (Gas& g : M.gases) g.effPopDefined = true;
        for (const auto& kv : eff) {
            const double v = evalNumber(kv.second);
            for (int id : M.find(parseState(kv.first))) M.gases[M.states[id].gas].effPop[id] = v;
        }
    }
    // anisotropic scattering (Setup.h:351-460)
    double angleNumber = 0;
    if (setup.has("electronKinetics.anisotropicScattering") &&
        (setup.value("electronKinetics.anisotropicScattering.isOn") == "true" ||
         setup.value("electronKinetics.anisotropicScattering.isOn") == "True" ||
         setup.value("electronKinetics.anisotropicScattering.isOn") == "1")) {
        angleNumber = evalNumber(setup.value("electronKinetics.anisotropicScattering.angleNumber"));
        std::vector<std::string> lines;
        for (const std::string& n : setup.childNames("electronKinetics.anisotropicScattering.collisions")) {
            if (n.find(';') != std::string::npos) { lines.push_back(n); continue; }
            std::ifstream in(inputDir + "/" + n);
            if (!in) throw InputError("The file " + n + " could not be opened.");
            std::string l;
            while (std::getline(in, l)) {
                const std::string c = noSpaces(removeComments(l));
                if (!c.empty()) lines.push_back(c);
            }
        }
        for (const std::string& l : lines) {
            const std::vector<std::string> p = splitChars(l, ";");
            bool found = false;
            auto params = [&](size_t at) {
                std::vector<double> v;
                if (p.size() > at)
                    for (const std::string& t : splitChars(p[at], ",")) v.push_back(evalNumber(t));
                return v;
            };
            if (!p.empty() && p[0] == "group") {
                if (p.size() < 4)
                    throw InputError("Error while reading the following setup line of electronKinetics.anisotropicScattering.collisions:\n" + l);
                for (const Gas& g : M.gases) {
                    if (g.name != p[1]) continue;
                    for (int k : g.colls)
                        if (M.colls[k].type == p[2]) { M.colls[k].scattering = p[3]; M.colls[k].params = params(4); found = true; }
                    break;
                }
            } else if (!p.empty() && p[0] == "single") {
                if (p.size() < 3)
                    throw InputError("Error while reading the following setup line of electronKinetics.anisotropicScattering.collisions:\n" + l);
                for (Collision& c : M.colls)
                    if (M.description(c) == p[1]) { c.scattering = p[2]; c.params = params(3); found = true; break; }
            } else {
                throw InputError("Error while reading the following setup line of electronKinetics.anisotropicScattering.collisions:\n" + l);
            }
            if (!found)
                throw InputError("Error while reading the following setup line of electronKinetics.anisotropicScattering.collisions:\n" +
                                 l + "\nThe collision (group or single) was not found.");
        }
    }
    for (const Gas& g : M.gases)
        for (int k : g.colls) {
            Collision& c = M.colls[k];
            scatteringCode(c.scattering);
            if ((int)c.params.size() != requiredParams(c.scattering))
                throw InputError("The angular scattering function '" + c.scattering + "' requires exactly " +
                                 std::to_string(requiredParams(c.scattering)) + " parameters, and not " + std::to_string(c.params.size()));
            completeCrossSections(M, c, angleNumber);
            if (c.scattering != "isotropic" && (c.type == "Effective" || c.type == "Attachment"))
                throw InputError("Error in the following collision:\n" + M.description(c) + "\n'" + c.type +
                                 "' collisions cannot have a user-prescribed angular scattering model.");
        }
    // per-gas checks (Setup.h:271-289)
    for (int gi = 0; gi < (int)M.gases.size(); ++gi) {
        if (M.gases[gi].colls.empty()) continue;
        if (M.gases[gi].mass == kNonDef)
            throw InputError("Mass of gas " + M.gases[gi].name +
                             " not found.\nNeeded for the evaluation of the elastic collision operator (Boltzmann).\nCheck input file.");
        checkPopulationNorms(M, M.gases[gi]);
        checkElasticCollisions(M, gi);
    }
    // flatten (BoltzmannMC.C:29-271): gases with collisions in order, their
    // non-Effective collisions in order, a superelastic entry after its partner
    GasMixture m;
    for (const Gas& g : M.gases) m.allGasNames.push_back(g.name);
    m.tableOffset.push_back(0);
    for (const Gas& g : M.gases) {
        if (g.colls.empty()) continue;
        const int slot = (int)m.gasNames.size();
        m.gasNames.push_back(g.name);
        m.gasMass.push_back(g.mass);
        m.gasFraction.push_back(g.fraction);
        for (int k : g.colls) {
            const Collision& c = M.colls[k];
            if (c.type == "Effective") continue;
            const int cid = (int)m.collisionDescription.size();
            const State& tgt = M.states[c.target];
            const int prodId = c.products.empty() ? -1 : c.products[0];
            const int code = c.type == "Ionization" ? EKMC_IONIZATION : (c.type == "Attachment" ? EKMC_ATTACHMENT : EKMC_CONSERVATIVE);
            const std::string desc = M.description(c);
            double ratio = 0;
            if (c.isReverse) {
                const double gt = tgt.statWeight, gp = M.states[prodId].statWeight;
                if (gt == kNonDef || gp == kNonDef || gp == 0)
                    throw InputError("Statistical weights of " + tgt.raw.name() + " and " + M.states[prodId].raw.name() +
                                     " are needed for the superelastic collision of\n" + desc);
                ratio = gt / gp;
            }
            auto push = [&](bool super) {
                m.type.push_back(code);
                m.isElastic.push_back(c.type == "Elastic" ? 1 : 0);
                m.isSuperelastic.push_back(super ? 1 : 0);
                m.gasIndex.push_back(slot);
                m.threshold.push_back(c.threshold);
                m.targetMass.push_back(g.mass);
                m.relDensity.push_back(super ? M.states[prodId].density : tgt.density);
                m.statWeightRatio.push_back(super ? ratio : 0.0);
                m.wParameter.push_back(code == EKMC_IONIZATION ? (g.opb != kNonDef ? g.opb : c.threshold) : 0.0);
                m.scattering.push_back(scatteringCode(c.scattering));
                m.scatteringParams.push_back(c.params.size() > 0 ? c.params[0] : 0.0);
                m.scatteringParams.push_back(c.params.size() > 1 ? c.params[1] : 0.0);
                if (!super) {
                    m.tableEnergy.insert(m.tableEnergy.end(), c.integral.e.begin(), c.integral.e.end());
                    m.tableSigma.insert(m.tableSigma.end(), c.integral.s.begin(), c.integral.s.end());
                }
                m.tableOffset.push_back((int32_t)m.tableEnergy.size());
                m.lxcatType.push_back(c.type);
                m.collisionId.push_back(cid);
                m.description.push_back(desc);
            };
            m.collisionFirstProcess.push_back(m.nProcesses());
            push(false);
            if (c.isReverse) push(true);
            m.collisionIsReverse.push_back(c.isReverse ? 1 : 0);
            m.collisionDescription.push_back(desc);
            m.collisionTargetDensity.push_back(tgt.density);
            m.collisionProductDensity.push_back(prodId >= 0 ? M.states[prodId].density : 0.0);
            m.collisionEnergy.push_back(c.integral.e);
            m.collisionSigmaIntegral.push_back(c.integral.s);
            m.collisionEnergyMomTransf.push_back(c.momTransf.e);
            m.collisionSigmaMomTransf.push_back(c.momTransf.s);
            m.collisionThreshold.push_back(c.threshold);
            m.collisionStatWeightRatio.push_back(ratio);
            m.collisionType.push_back(c.type);
            m.collisionGas.push_back(slot);
            m.collisionKind.push_back(0);
            m.collisionRefId.push_back(k);
        }
    }
    // collisions listed in the rate-coefficient files only (BoltzmannMC.C:
    // 2361-2382): the Effective ones of every gas, then the extra ones
    for (int kind = 1; kind <= 2; ++kind) {
        for (size_t gi = 0; gi < M.gases.size(); ++gi) {
            int slot = -1;
            for (size_t q = 0; q < m.gasNames.size(); ++q)
                if (m.gasNames[q] == M.gases[gi].name) slot = (int)q;
            for (size_t k = 0; k < M.colls.size(); ++k) {
                const Collision& c = M.colls[k];
                if (M.states[c.target].gas != (int)gi) continue;
                if (kind == 1 ? (c.extra || c.type != "Effective") : !c.extra) continue;
                const XSec& x = c.integral.defined() ? c.integral : c.momTransf;
                const XSec& mt = c.momTransf.defined() ? c.momTransf : c.integral;
                double ratio = 0;
                if (c.isReverse) {
                    const State& t = M.states[c.target];
                    const State& pr = M.states[c.products[0]];
                    if (t.statWeight == kNonDef || pr.statWeight == kNonDef)
                        throw InputError("Unable to find '" + (t.statWeight == kNonDef ? t : pr).raw.name() +
                                         "' statistical weight for the evaluation of superelastic rate coefficient of " +
                                         M.description(c) + "\n");
                    ratio = t.statWeight / pr.statWeight;
                }
                m.collisionFirstProcess.push_back(-1);
                m.collisionIsReverse.push_back(c.isReverse ? 1 : 0);
                m.collisionDescription.push_back(M.description(c));
                m.collisionTargetDensity.push_back(M.states[c.target].density);
                m.collisionProductDensity.push_back(c.products.empty() ? 0.0 : M.states[c.products[0]].density);
                m.collisionEnergy.push_back(x.e);
                m.collisionSigmaIntegral.push_back(x.s);
                m.collisionEnergyMomTransf.push_back(mt.e);
                m.collisionSigmaMomTransf.push_back(mt.s);
                m.collisionThreshold.push_back(c.threshold);
                m.collisionStatWeightRatio.push_back(ratio);
                m.collisionType.push_back(c.type);
                m.collisionGas.push_back(slot);
                m.collisionKind.push_back(kind);
                m.collisionRefId.push_back((int)k);
            }
        }
    }
    return m;
}

ekmc_mixture GasMixture::view() const {
    ekmc_mixture x{};
    x.n_processes = nProcesses();
    x.n_gases = (int32_t)gasNames.size();
    x.type = type.data();
    x.is_elastic = isElastic.data();
    x.is_superelastic = isSuperelastic.data();
    x.gas_index = gasIndex.data();
    x.threshold = threshold.data();
    x.target_mass = targetMass.data();
    x.rel_density = relDensity.data();
    x.stat_weight_ratio = statWeightRatio.data();
    x.w_parameter = wParameter.data();
    x.scattering = scattering.data();
    x.scattering_params = scatteringParams.data();
    x.table_offset = tableOffset.data();
    x.table_energy = tableEnergy.data();
    x.table_sigma = tableSigma.data();
    x.gas_fraction = gasFraction.data();
    return x;
}

}  // namespace lokimc
