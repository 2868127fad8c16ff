// gas_model.cpp -- see gas_model.h.
#include "gas_model.h"

#include <cctype>
#include <cmath>
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

struct State {
    RawState raw;
    int gas = -1;
    double population = 0, statWeight = -1;
    bool popSet = false;
};

struct Collision {
    LXCatEntry e;
    int gas = -1, target = -1, product = -1;
    std::string scattering = "isotropic";
    std::vector<double> params;
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

double propertyNumber(const std::string& key, const std::string& value, const std::string& what) {
    if (value.find('@') != std::string::npos)
        throw InputError("The " + what + " of '" + key + "' is given by a property function ('" + value +
                         "'); property functions are not supported by this reader.");
    return evalNumber(value);
}

std::string descriptionOf(const Collision& c, const std::vector<State>& states) {
    std::string s = "e+" + states[c.target].raw.name();
    s += c.e.isReverse ? "<->" : "->";
    if (c.e.type == "Ionization") s += "e+e+";
    else if (c.e.type != "Attachment") s += "e+";
    for (size_t i = 0; i < c.e.products.size(); ++i) {
        if (c.e.productStoi[i] > 1) s += std::to_string((int)c.e.productStoi[i]);
        s += c.e.products[i].name();
        if (i + 1 < c.e.products.size()) s += "+";
    }
    return s + "," + c.e.type;
}

}  // namespace

GasMixture buildMixture(const SetupFile& setup, const std::string& inputDir) {
    const std::vector<LXCatEntry> entries = readLXCat(setup.childNames("electronKinetics.LXCatFiles"), inputDir);
    if (entries.empty()) throw InputError("No LXCat collisions found in electronKinetics.LXCatFiles.");
    // gases and states in order of first appearance (Setup.h:300-320)
    std::vector<std::string> gases;
    std::vector<State> states;
    auto gasId = [&](const std::string& g) {
        for (size_t i = 0; i < gases.size(); ++i)
            if (gases[i] == g) return (int)i;
        gases.push_back(g);
        return (int)gases.size() - 1;
    };
    auto stateId = [&](const RawState& r) {
        const int g = gasId(r.gas);
        for (size_t i = 0; i < states.size(); ++i)
            if (states[i].raw == r) return (int)i;
        State s;
        s.raw = r;
        s.gas = g;
        states.push_back(s);
        return (int)states.size() - 1;
    };
    std::vector<Collision> colls;
    for (const LXCatEntry& e : entries) {
        Collision c;
        c.e = e;
        c.target = stateId(e.target);
        c.gas = states[c.target].gas;
        for (const RawState& p : e.products) {
            const int id = stateId(p);
            if (c.product < 0) c.product = id;
        }
        if (e.isReverse && (e.products.size() != 1 || e.productStoi[0] != 1))
            throw InputError("Error while creating collision '" + descriptionOf(c, states) +
                             "'. Klein-Rosseland microreversibility relation valid only for binary collisions.\n");
        if ((e.type == "Effective" || e.type == "Elastic") && !(e.target.vibLevel.empty() && e.target.rotLevel.empty()))
            throw InputError("Found ''" + e.type + "'' collision with ''" + e.target.name() + "'' as target [" +
                             descriptionOf(c, states) + "].\n" + e.type + " collisions are only allowed for electronic states. Please check LXCat files.\n");
        colls.push_back(c);
    }
    // gas properties
    const int nG = (int)gases.size();
    std::vector<double> mass(nG, -1), fraction(nG, 0), opb(nG, -1);
    std::vector<bool> hasColl(nG, false);
    for (const Collision& c : colls) hasColl[c.gas] = true;
    auto gasMap = [&](const std::string& key, std::vector<double>& dst, const char* what) {
        for (const auto& kv : setup.propertyMap("electronKinetics.gasProperties." + key, inputDir))
            for (int g = 0; g < nG; ++g)
                if (gases[g] == kv.first) dst[g] = propertyNumber(kv.first, kv.second, what);
    };
    gasMap("mass", mass, "mass");
    gasMap("fraction", fraction, "fraction");
    gasMap("OPBParameter", opb, "OPB parameter");
    // state properties
    for (const auto& kv : setup.propertyMap("electronKinetics.stateProperties.population", inputDir)) {
        if (kv.first.find('*') != std::string::npos)
            throw InputError("Population of '" + kv.first + "': wildcard states are not supported by this reader.");
        const RawState r = parseState(kv.first);
        for (State& s : states)
            if (s.raw == r) { s.population = propertyNumber(kv.first, kv.second, "population"); s.popSet = true; }
    }
    for (const auto& kv : setup.propertyMap("electronKinetics.stateProperties.statisticalWeight", inputDir)) {
        const RawState r = parseState(kv.first);
        for (State& s : states)
            if (s.raw == r) s.statWeight = propertyNumber(kv.first, kv.second, "statistical weight");
    }
    // densities (State.h:72-85): electronic states carry the gas fraction,
    // vibrational and rotational ones their parents' populations
    auto parentPop = [&](const RawState& r, bool vibParent) {
        RawState p = r;
        p.rotLevel.clear();
        if (!vibParent) p.vibLevel.clear();
        for (const State& s : states)
            if (s.raw == p) return s.population;
        return 0.0;
    };
    std::vector<double> density(states.size(), 0);
    for (size_t i = 0; i < states.size(); ++i) {
        const State& s = states[i];
        double d = s.population * fraction[s.gas];
        if (!s.raw.rotLevel.empty()) d *= parentPop(s.raw, true) * parentPop(s.raw, false);
        else if (!s.raw.vibLevel.empty()) d *= parentPop(s.raw, false);
        density[i] = d;
    }
    for (int g = 0; g < nG; ++g) {
        if (!hasColl[g]) continue;
        if (mass[g] < 0)
            throw InputError("Mass of gas " + gases[g] +
                             " not found.\nNeeded for the evaluation of the elastic collision operator (Boltzmann).\nCheck input file.");
        double norm = 0;
        bool anyPop = false;
        for (const State& s : states)
            if (s.gas == g && s.raw.vibLevel.empty() && s.raw.rotLevel.empty() && s.raw.ionCharg.empty()) {
                norm += s.population;
                anyPop |= s.popSet;
            }
        if (anyPop && std::fabs(norm - 1.0) > 10 * 2.220446049250313e-16)
            throw InputError("Error: the electronic populations of gas " + gases[g] + " are not normalized.");
        bool elastic = false, effective = false;
        for (const Collision& c : colls)
            if (c.gas == g) { elastic |= c.e.type == "Elastic"; effective |= c.e.type == "Effective"; }
        if (!elastic && effective)
            throw InputError("Gas " + gases[g] + " has an Effective cross section and no Elastic one; deriving the elastic cross section from the effective one is not supported by this reader.");
    }
    // anisotropic scattering (Setup.h:351-451)
    if (setup.has("electronKinetics.anisotropicScattering") &&
        (setup.value("electronKinetics.anisotropicScattering.isOn") == "true" ||
         setup.value("electronKinetics.anisotropicScattering.isOn") == "True" ||
         setup.value("electronKinetics.anisotropicScattering.isOn") == "1")) {
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
            if (!p.empty() && p[0] == "group" && p.size() >= 4) {
                for (Collision& c : colls)
                    if (gases[c.gas] == p[1] && c.e.type == p[2]) { c.scattering = p[3]; c.params = params(4); found = true; }
            } else if (!p.empty() && p[0] == "single" && p.size() >= 3) {
                for (Collision& c : colls)
                    if (descriptionOf(c, states) == p[1]) { c.scattering = p[2]; c.params = params(3); found = true; break; }
            } else {
                throw InputError("Error while reading the following setup line of electronKinetics.anisotropicScattering.collisions:\n" + l);
            }
            if (!found)
                throw InputError("Error while reading the following setup line of electronKinetics.anisotropicScattering.collisions:\n" +
                                 l + "\nThe collision (group or single) was not found.");
        }
    }
    for (const Collision& c : colls) {
        scatteringCode(c.scattering);
        if ((int)c.params.size() != requiredParams(c.scattering))
            throw InputError("The angular scattering function '" + c.scattering + "' requires exactly " +
                             std::to_string(requiredParams(c.scattering)) + " parameters, and not " + std::to_string(c.params.size()));
        if (c.scattering != "isotropic" && (c.e.type == "Effective" || c.e.type == "Attachment"))
            throw InputError("'" + c.e.type + "' collisions cannot have a user-prescribed angular scattering model.");
        if (c.scattering != "isotropic" && c.e.subtype == "momentum-transfer")
            throw InputError("Collision '" + descriptionOf(c, states) +
                             "' has only a momentum-transfer cross section and an anisotropic scattering model; the angular conversion to an integral cross section is not supported by this reader.");
    }
    // flatten (BoltzmannMC.C:29-271): gases with collisions in order, their
    // non-Effective collisions in order, a superelastic entry after its partner
    GasMixture m;
    m.allGasNames = gases;
    m.tableOffset.push_back(0);
    std::vector<int> gasSlot(nG, -1);
    for (int g = 0; g < nG; ++g) {
        if (!hasColl[g]) continue;
        gasSlot[g] = (int)m.gasNames.size();
        m.gasNames.push_back(gases[g]);
        m.gasMass.push_back(mass[g]);
        m.gasFraction.push_back(fraction[g]);
        for (const Collision& c : colls) {
            if (c.gas != g || c.e.type == "Effective") continue;
            const int cid = (int)m.collisionDescription.size();
            std::vector<double> en = c.e.energy, sg = c.e.sigma;
            for (size_t i = 0; i + 1 < en.size(); ++i)
                if (en[i] >= en[i + 1])
                    throw InputError("The energy column of the cross section corresponding to the collision shown below is not strictly increasing.\n" +
                                     descriptionOf(c, states) + "\nPlease fix this in the LXCat files.");
            if (c.e.type == "Elastic" && en[0] != 0) {  // Collision.C:40-55
                en.insert(en.begin(), 0.0);
                sg.insert(sg.begin(), sg[0]);
            }
            const int code = c.e.type == "Ionization" ? EKMC_IONIZATION : (c.e.type == "Attachment" ? EKMC_ATTACHMENT : EKMC_CONSERVATIVE);
            const std::string desc = descriptionOf(c, states);
            auto push = [&](bool super) {
                m.type.push_back(code);
                m.isElastic.push_back(c.e.type == "Elastic" ? 1 : 0);
                m.isSuperelastic.push_back(super ? 1 : 0);
                m.gasIndex.push_back(gasSlot[g]);
                m.threshold.push_back(c.e.threshold);
                m.targetMass.push_back(mass[g]);
                m.relDensity.push_back(super ? density[c.product] : density[c.target]);
                double ratio = 0;
                if (super) {
                    const double gt = states[c.target].statWeight, gp = states[c.product].statWeight;
                    if (gt < 0 || gp <= 0)
                        throw InputError("Statistical weights of " + states[c.target].raw.name() + " and " +
                                         states[c.product].raw.name() + " are needed for the superelastic collision of\n" + desc);
                    ratio = gt / gp;
                }
                m.statWeightRatio.push_back(ratio);
                m.wParameter.push_back(code == EKMC_IONIZATION ? (opb[g] >= 0 ? opb[g] : c.e.threshold) : 0.0);
                m.scattering.push_back(scatteringCode(c.scattering));
                m.scatteringParams.push_back(c.params.size() > 0 ? c.params[0] : 0.0);
                m.scatteringParams.push_back(c.params.size() > 1 ? c.params[1] : 0.0);
                if (!super) {
                    m.tableEnergy.insert(m.tableEnergy.end(), en.begin(), en.end());
                    m.tableSigma.insert(m.tableSigma.end(), sg.begin(), sg.end());
                }
                m.tableOffset.push_back((int32_t)m.tableEnergy.size());
                m.lxcatType.push_back(c.e.type);
                m.collisionId.push_back(cid);
                m.description.push_back(desc);
            };
            m.collisionFirstProcess.push_back(m.nProcesses());
            push(false);
            if (c.e.isReverse) push(true);
            m.collisionIsReverse.push_back(c.e.isReverse ? 1 : 0);
            m.collisionDescription.push_back(desc);
            m.collisionTargetDensity.push_back(density[c.target]);
            m.collisionProductDensity.push_back(c.product >= 0 ? density[c.product] : 0.0);
            m.collisionEnergy.push_back(en);
            m.collisionSigmaIntegral.push_back(sg);
            // isotropic scattering: the momentum-transfer and integral cross sections agree
            m.collisionSigmaMomTransf.push_back(sg);
            m.collisionThreshold.push_back(c.e.threshold);
            m.collisionStatWeightRatio.push_back(c.e.isReverse ? m.statWeightRatio.back() : 0.0);
            m.collisionType.push_back(c.e.type);
            m.collisionGas.push_back(gasSlot[g]);
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
