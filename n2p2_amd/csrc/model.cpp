// Model loader: input.nn subset of a 2G-HDNNP, scaling.data and weight files
// into the flat tables of model.hpp.  Citations are to n2p2's src/libnnp.
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <utility>

#include "nnpx.h"

namespace nnpx {
namespace {

[[noreturn]] void fail(int status, const std::string& msg) { throw Error{status, msg}; }

// Whitespace tokenizer equivalent to split(reduce(line)) (utility.cpp:33-77).
std::vector<std::string> tokens(const std::string& s)
{
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
        size_t j = i;
        while (j < s.size() && s[j] != ' ' && s[j] != '\t') ++j;
        if (j > i) out.emplace_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

double num(const std::string& s) { return std::atof(s.c_str()); }

// Atomic numbers by symbol (ElementMap::knownElements order, Z = index + 1).
int atomic_number(const std::string& sym)
{
    static const char* const table[] = {
        "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg",
        "Al", "Si", "P", "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr",
        "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br",
        "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd",
        "Ag", "Cd", "In", "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La",
        "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er",
        "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au",
        "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th",
        "Pa", "U", "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No"};
    for (int z = 0; z < (int)(sizeof table / sizeof table[0]); ++z)
        if (sym == table[z]) return z + 1;
    fail(NNPX_ERR_MODEL, "unknown element symbol \"" + sym + "\"");
}

// Settings file: key -> values in file order (Settings::parseLines,
// Settings.cpp:324-392; "#" and "!" start comments).
struct Settings {
    std::vector<std::pair<std::string, std::string>> kv;
    bool has(const std::string& k) const
    {
        for (auto& p : kv) if (p.first == k) return true;
        return false;
    }
    std::string get(const std::string& k) const
    {
        for (auto& p : kv) if (p.first == k) return p.second;
        fail(NNPX_ERR_MODEL, "keyword \"" + k + "\" not found");
    }
    std::vector<std::string> all(const std::string& k) const
    {
        std::vector<std::string> v;
        for (auto& p : kv) if (p.first == k) v.push_back(p.second);
        return v;
    }
};

Settings read_settings(const std::string& path)
{
    std::ifstream in(path);
    if (!in) fail(NNPX_ERR_FILE, "could not open file: \"" + path + "\"");
    Settings s;
    std::string line;
    while (std::getline(in, line)) {
        size_t p = line.find('#');
        if (p != std::string::npos) line.erase(p);
        p = line.find('!');
        if (p != std::string::npos) line.erase(p);
        std::vector<std::string> w = tokens(line);
        if (w.empty() || w[0] == "\r") continue;
        std::string val;
        for (size_t i = 1; i < w.size(); ++i) val += (i > 1 ? " " : "") + w[i];
        std::string key = w[0] == "symfunction" ? "symfunction_short" : w[0];
        s.kv.emplace_back(key, val);
    }
    return s;
}

int activation(const std::string& c)  // activationFromString (NeuralNetwork.cpp:1078-1097)
{
    static const char* const names[] = {"l", "t", "s", "p", "r", "g", "c", "S", "e", "h"};
    for (int a = 0; a < 10; ++a)
        if (c == names[a]) return a;
    fail(NNPX_ERR_MODEL, "unknown activation function \"" + c + "\"");
}

struct SF {
    int type, ec, e1, e2;
    double eta, rs, rc, lambda, zeta;
    int zetaInt;
    bool useIntPow;
    int index;
    int slot[kMaxElem];
    // compact types: radial window (rl, rc), angle window in degrees, subtype
    double rl, angleLeft, angleRight;
    std::string subtype;
    int core, asym;
};

// Order of SymFncExpRad::operator< / SymFncExpAngn::operator<
// (SymFncExpRad.cpp:57-75, SymFncExpAngn.cpp:51-77); cutoff type and alpha
// are shared by all functions of a model.
bool sf_before(const SF& a, const SF& b)
{
    if (a.type != b.type) return a.type < b.type;
    if (a.type >= 23) {  // SymFncCompRadWeighted / SymFncCompAng{n,w}Weighted operator<
        if (a.subtype != b.subtype) return a.subtype < b.subtype;
        if (a.rc != b.rc) return a.rc < b.rc;
        if (a.rl != b.rl) return a.rl < b.rl;
        if (a.type == 23) return false;
        if (a.angleLeft != b.angleLeft) return a.angleLeft < b.angleLeft;
        return a.angleRight < b.angleRight;
    }
    if (a.type == 12 || a.type == 13) {  // SymFncExpRadWeighted / SymFncExpAngnWeighted operator<
        if (a.rc != b.rc) return a.rc < b.rc;
        if (a.eta != b.eta) return a.eta < b.eta;
        if (a.rs != b.rs) return a.rs < b.rs;
        if (a.type == 12) return false;
        if (a.zeta != b.zeta) return a.zeta < b.zeta;
        return a.lambda < b.lambda;
    }
    if (sf_compact(a.type)) {  // SymFncCompRad.cpp:48-68, SymFncCompAngn/SymFncCompAngw.cpp:56-82
        if (a.subtype != b.subtype) return a.subtype < b.subtype;
        if (a.e1 != b.e1) return a.e1 < b.e1;
        if (a.type != 20 && a.e2 != b.e2) return a.e2 < b.e2;
        if (a.rc != b.rc) return a.rc < b.rc;
        if (a.rl != b.rl) return a.rl < b.rl;
        if (a.type == 20) return false;
        if (a.angleLeft != b.angleLeft) return a.angleLeft < b.angleLeft;
        return a.angleRight < b.angleRight;
    }
    if (a.rc != b.rc) return a.rc < b.rc;
    if (a.eta != b.eta) return a.eta < b.eta;
    if (a.rs != b.rs) return a.rs < b.rs;
    if (a.type == 2) return a.e1 < b.e1;
    if (a.zeta != b.zeta) return a.zeta < b.zeta;
    if (a.lambda != b.lambda) return a.lambda < b.lambda;
    if (a.e1 != b.e1) return a.e1 < b.e1;
    return a.e2 < b.e2;
}

// SymFncBaseComp::setCompactFunction (SymFncBaseComp.cpp:40-88): "p1".."p4"
// polynomials, a trailing "a" for the asymmetric form, "e" exponential.
void compact_subtype(SF& f)
{
    const std::string& t = f.subtype;
    if (t.empty() || t.size() > 3) fail(NNPX_ERR_MODEL, "invalid compact function type \"" + t + "\"");
    f.asym = 0;
    if (t[0] == 'p') {
        if (t.size() < 2 || t[1] < '1' || t[1] > '4') fail(NNPX_ERR_MODEL, "invalid polynom type \"" + t + "\"");
        f.core = t[1] - '0';
        if (t.size() == 3) {
            if (t[2] != 'a') fail(NNPX_ERR_MODEL, "invalid polynom specifier \"" + t + "\"");
            f.asym = 1;
        }
    } else if (t[0] == 'e') {
        f.core = 0;
    } else {
        fail(NNPX_ERR_MODEL, "unknown compact SF type \"" + t + "\"");
    }
}

}  // namespace

void load_model(const std::string& dir, HostModel& hm)
{
    Settings s = read_settings(dir + "/input.nn");
    ModelT& t = hm.t;
    t = ModelT{};

    if (s.has("nnp_type")) {
        std::string v = s.get("nnp_type");
        if (v != "2G-HDNNP" && v != "2")
            fail(NNPX_ERR_UNSUPPORTED, "only 2G-HDNNP models are supported");
    }
    // Normalisation (Mode::setupNormalization, Mode.cpp:239-300).
    int nk = s.has("mean_energy") + s.has("conv_energy") + s.has("conv_length");
    if (nk == 3) {
        hm.normalize = true;
        hm.meanEnergy = num(s.get("mean_energy"));
        hm.convEnergy = num(s.get("conv_energy"));
        hm.convLength = num(s.get("conv_length"));
    } else if (nk != 0) {
        fail(NNPX_ERR_MODEL, "use all or none of mean_energy, conv_energy, conv_length");
    }
    // Element map sorted by atomic number (ElementMap.cpp:36-50).
    std::vector<std::string> syms = tokens(s.get("elements"));
    std::sort(syms.begin(), syms.end(), [](const std::string& a, const std::string& b) {
        return atomic_number(a) < atomic_number(b);
    });
    const int ne = (int)syms.size();
    if (ne < 1 || ne > kMaxElem) fail(NNPX_ERR_UNSUPPORTED, "1..8 elements supported");
    if (std::atoi(s.get("number_of_elements").c_str()) != ne)
        fail(NNPX_ERR_MODEL, "inconsistent number of elements");
    hm.symbols = syms;
    hm.Z.clear();
    for (auto& x : syms) hm.Z.push_back(atomic_number(x));
    for (size_t e = 0; e < syms.size() && e < (size_t)kMaxElem; ++e) t.Zel[e] = (double)hm.Z[e];
    auto eidx = [&](const std::string& x) {
        for (int i = 0; i < ne; ++i) if (syms[i] == x) return i;
        fail(NNPX_ERR_MODEL, "element \"" + x + "\" not in element map");
    };
    hm.energyOffset.assign(ne, 0.0);
    for (auto& v : s.all("atom_energy")) {
        auto w = tokens(v);
        if (w.size() < 2) fail(NNPX_ERR_MODEL, "bad atom_energy line");
        hm.energyOffset[eidx(w[0])] = num(w[1]);
    }
    t.nelem = ne;
    // Cutoff (Mode::setupCutoff, Mode.cpp:589-682).
    auto ct = tokens(s.get("cutoff_type"));
    t.cutoffType = std::atoi(ct.at(0).c_str());
    t.alpha = ct.size() > 1 ? num(ct[1]) : 0.0;
    if (t.cutoffType < 0 || t.cutoffType > 8) fail(NNPX_ERR_ARGUMENT, "unknown cutoff type");
    if (t.alpha < 0.0 || t.alpha >= 1.0) fail(NNPX_ERR_ARGUMENT, "0 <= alpha < 1.0 is required");
    {
        double const E = std::exp(1.0);
        t.tanhPre = std::pow((E + 1 / E) / (E - 1 / E), 3);
    }

    // Symmetry functions (SymFncExpRad.cpp:77-96, SymFncBaseExpAng.cpp:50-93).
    std::vector<std::vector<SF>> sfs(ne);
    for (auto& v : s.all("symfunction_short")) {
        auto w = tokens(v);
        SF f{};
        f.ec = eidx(w.at(0));
        f.type = std::atoi(w.at(1).c_str());
        if (f.type == 2) {
            f.e1 = f.e2 = eidx(w.at(2));
            f.eta = num(w.at(3));
            f.rs = num(w.at(4));
            f.rc = num(w.at(5));
        } else if (f.type == 3 || f.type == 9) {  // narrow / wide angular, same parameters
            f.e1 = eidx(w.at(2));
            f.e2 = eidx(w.at(3));
            f.eta = num(w.at(4));
            f.lambda = num(w.at(5));
            f.zeta = num(w.at(6));
            f.rc = num(w.at(7));
            if (w.size() > 8) f.rs = num(w[8]);
            if (f.e1 > f.e2) std::swap(f.e1, f.e2);
            f.zetaInt = (int)std::round(f.zeta);
            f.useIntPow = std::fabs(f.zeta - f.zetaInt) <= std::numeric_limits<double>::min();
        } else if (f.type == 12) {  // SymFncExpRadWeighted::setParameters
            f.e1 = f.e2 = -1;
            f.eta = num(w.at(2));
            f.rs = num(w.at(3));
            f.rc = num(w.at(4));
        } else if (f.type == 13) {  // SymFncExpAngnWeighted::setParameters
            f.e1 = f.e2 = -1;
            f.eta = num(w.at(2));
            f.rs = num(w.at(3));
            f.lambda = num(w.at(4));
            f.zeta = num(w.at(5));
            f.rc = num(w.at(6));
            f.zetaInt = (int)std::round(f.zeta);
            f.useIntPow = std::fabs(f.zeta - f.zetaInt) <= std::numeric_limits<double>::min();
        } else if (f.type == 23) {  // SymFncCompRadWeighted::setParameters
            f.e1 = f.e2 = -1;
            f.rl = num(w.at(2));
            f.rc = num(w.at(3));
            f.subtype = w.at(4);
            if (f.rl >= f.rc) fail(NNPX_ERR_MODEL, "lower radial boundary >= upper radial boundary");
            compact_subtype(f);
        } else if (f.type == 24 || f.type == 25) {  // SymFncBaseCompAngWeighted::setParameters
            f.e1 = f.e2 = -1;
            f.rl = num(w.at(2));
            f.rc = num(w.at(3));
            f.angleLeft = num(w.at(4));
            f.angleRight = num(w.at(5));
            f.subtype = w.at(6);
            if (f.rl >= f.rc) fail(NNPX_ERR_MODEL, "lower radial boundary >= upper radial boundary");
            compact_subtype(f);
            if (f.angleLeft >= f.angleRight)
                fail(NNPX_ERR_MODEL, "left angle boundary right of or equal to right angle boundary");
            const double center = (f.angleLeft + f.angleRight) / 2.0;
            if ((f.angleLeft < 0.0 && center != 0.0) || (f.angleRight > 180.0 && center != 180.0))
                fail(NNPX_ERR_MODEL, "angle boundary out of [0,180] and center of angular function /= 0 or /= 180");
            if (f.angleRight - f.angleLeft > 360.0)
                fail(NNPX_ERR_MODEL, "periodic symmetry function cannot spread over domain > 360 degrees");
        } else if (f.type == 20) {  // SymFncCompRad.cpp:85-112
            f.e1 = f.e2 = eidx(w.at(2));
            f.rl = num(w.at(3));
            f.rc = num(w.at(4));
            f.subtype = w.at(5);
            if (f.rl > f.rc) fail(NNPX_ERR_MODEL, "lower radial boundary >= upper radial boundary");
            compact_subtype(f);
        } else if (f.type == 21 || f.type == 22) {  // SymFncBaseCompAng.cpp:47-110
            f.e1 = eidx(w.at(2));
            f.e2 = eidx(w.at(3));
            f.rl = num(w.at(4));
            f.rc = num(w.at(5));
            f.angleLeft = num(w.at(6));
            f.angleRight = num(w.at(7));
            f.subtype = w.at(8);
            if (f.e1 > f.e2) std::swap(f.e1, f.e2);
            if (f.rl > f.rc) fail(NNPX_ERR_MODEL, "lower radial boundary >= upper radial boundary");
            compact_subtype(f);
            if (f.angleLeft >= f.angleRight)
                fail(NNPX_ERR_MODEL, "left angle boundary right of or equal to right angle boundary");
            const double center = (f.angleLeft + f.angleRight) / 2.0;
            if ((f.angleLeft < 0.0 && center != 0.0) || (f.angleRight > 180.0 && center != 180.0))
                fail(NNPX_ERR_MODEL, "angle boundary out of [0,180] and center of angular function /= 0 or /= 180");
            if (f.angleRight - f.angleLeft > 360.0)
                fail(NNPX_ERR_MODEL, "periodic symmetry function cannot spread over domain > 360 degrees");
        } else {
            fail(NNPX_ERR_UNSUPPORTED, "symmetry function type " + w.at(1) + " not supported");
        }
        if (hm.normalize) {  // SymFnc*::changeLengthUnit
            f.eta /= hm.convLength * hm.convLength;
            f.rs *= hm.convLength;
            f.rc *= hm.convLength;
            f.rl *= hm.convLength;
        }
        sfs[f.ec].push_back(f);
    }

    t.maxCutoff = 0.0;
    int gcount = 0, mcount = 0, sfcount = 0;
    for (int e = 0; e < ne; ++e) {
        auto& v = sfs[e];
        if (v.empty()) fail(NNPX_ERR_RANGE, "element " + syms[e] + " has no symmetry functions");
        if ((int)v.size() > kMaxSF) fail(NNPX_ERR_UNSUPPORTED, "too many symmetry functions");
        std::sort(v.begin(), v.end(), sf_before);  // Element::sortSymmetryFunctions
        ElementT& E = t.el[e];
        E.nsf = (int)v.size();
        E.sf0 = sfcount;
        for (int i = 0; i < E.nsf; ++i) {
            v[i].index = i;
            t.maxCutoff = std::max(t.maxCutoff, v[i].rc);
        }
        // Derivative memory table (Element::setupSymmetryFunctionMemory,
        // Element.cpp:289-311).
        int tab = e * kMaxElem * kMaxSF;
        for (int n = 0; n < ne; ++n) {
            E.tab0[n] = tab + n * kMaxSF;
            E.nsfd[n] = 0;
            for (auto& f : v) {
                bool rel = sf_weighted(f.type) || (sf_radial(f.type) ? f.e1 == n : (f.e1 == n || f.e2 == n));
                f.slot[n] = -1;
                if (rel) {
                    f.slot[n] = E.nsfd[n];
                    t.table[E.tab0[n] + E.nsfd[n]++] = f.index;
                }
            }
            t.max_sfd = std::max(t.max_sfd, E.nsfd[n]);
        }
        // Groups (Element::setupSymmetryFunctionGroups, Element.cpp:194-287):
        // radial groups first, then narrow and wide angular; members in SF
        // order.  Compact groups do not key on rc.
        std::vector<std::vector<int>> groups;
        for (int pass : {2, 12, 20, 23, 3, 9, 13, 21, 22, 24, 25}) {
            for (auto& f : v) {
                if (f.type != pass) continue;
                bool added = false;
                for (auto& g : groups) {
                    const SF& h = v[g[0]];
                    if (h.type == f.type && (sf_compact(f.type) || h.rc == f.rc) && h.e1 == f.e1 &&
                        h.e2 == f.e2) {
                        g.push_back(f.index);
                        added = true;
                        break;
                    }
                }
                if (!added) groups.push_back({f.index});
            }
        }
        if ((int)groups.size() > kMaxGroups) fail(NNPX_ERR_UNSUPPORTED, "too many SF groups");
        E.grp0 = gcount;
        E.ngrp = (int)groups.size();
        E.nrad = 0;
        E.ucut = 1;
        for (auto& g : groups)
            if (v[g[0]].rc != v[groups[0][0]].rc) E.ucut = 0;
        for (auto& g : groups) {
            const SF& h = v[g[0]];
            const bool comp = sf_compact(h.type);
            if ((int)g.size() > (comp || sf_radial(h.type) ? kMaxMembers : kMaxMembersExp))
                fail(NNPX_ERR_UNSUPPORTED, "SF group too large");
            GroupT& G = t.grp[gcount++];
            G.type = h.type;
            G.e1 = h.e1;
            G.e2 = h.e2;
            G.nmem = (int)g.size();
            G.mem0 = mcount;
            double rc = h.rc;
            if (comp) {  // SymGrpCompRad/SymGrpCompAngn/SymGrpCompAngw::addMember
                G.rmin = 0.0;
                rc = 0.0;
                for (int idx : g) {
                    const SF& f = v[idx];
                    if (f.rl <= 0.0) G.rmin = 0.0;
                    else G.rmin = std::min(G.rmin, f.rl);
                    rc = std::max(rc, f.rc);
                }
                t.has_comp = 1;
            } else {
                t.has_exp = 1;
            }
            G.rc = rc;
            G.rc2 = rc * rc;
            G.rcinv = 1.0 / rc;
            G.rci = rc * t.alpha;
            G.iw = 1.0 / (rc - G.rci);
            if (sf_radial(h.type)) {
                E.nrad++;
                if (!comp) t.max_mem_rad = std::max(t.max_mem_rad, G.nmem);
            } else if (comp) {
                t.max_mem_comp = std::max(t.max_mem_comp, G.nmem);
            } else {
                t.max_mem_ang = std::max(t.max_mem_ang, G.nmem);
            }
            for (size_t q = 0; q < g.size(); ++q) {
                const SF& f = v[g[q]];
                MemberT& M = t.mem[mcount++];
                M.sf = f.index;
                M.slot1 = f.e1 >= 0 ? f.slot[f.e1] : -1;
                M.slot2 = f.e2 >= 0 ? f.slot[f.e2] : -1;
                for (int n = 0; n < kMaxElem; ++n) M.slotE[n] = n < ne ? f.slot[n] : -1;
                if (sf_weighted(f.type)) t.has_weighted = 1;
                const SF* prev = q ? &v[g[q - 1]] : nullptr;
                M.calcExp = (!prev || prev->eta != f.eta || prev->rs != f.rs) ? 1 : 0;
                M.useIntPow = f.useIntPow ? 1 : 0;
                M.zetaInt = f.zetaInt;
                M.eta = f.eta;
                M.rs = f.rs;
                M.lambda = f.lambda;
                M.zeta = f.zeta;
                if (f.type == 3 || f.type == 9 || f.type == 13) {  // SymGrpExpAngn/-Angw/-AngnWeighted::sortMembers (SymGrpExpAngn.cpp:108-148)
                    M.factorNorm = std::pow(2.0, 1.0 - f.zeta);
                    M.factorDeriv = 2.0 * f.eta / f.zeta / f.lambda;
                    M.zetaLambda = f.zeta * f.lambda;
                }
                M.sfac = 1.0;
                if (comp) {
                    // CompactFunction::setLeftRight (CompactFunction.cpp:58-77) of
                    // the radial window and of the angle window in radians
                    // (SymFncBaseCompAng.cpp:96-100, SymGrpCompAngw.cpp:88-91).
                    M.calcExp = 0;
                    M.core = f.core;
                    M.asym = f.asym;
                    M.rl = f.rl;
                    M.rcm = f.rc;
                    M.cen = (f.rc + f.rl) / 2.0;
                    M.scl = 1.0 / ((f.rc - f.rl) / 2.0);
                    M.al = f.angleLeft * M_PI / 180.0;
                    M.ar = f.angleRight * M_PI / 180.0;
                    M.acen = (M.ar + M.al) / 2.0;
                    M.ascl = 1.0 / ((M.ar - M.al) / 2.0);
                    // Narrow groups re-evaluate the r_jk factor only when
                    // (rl, rc) change or the previous member was skipped
                    // (SymGrpCompAngn.cpp:108-126, 288-296).  `walk` marks
                    // members whose reused factor may be another subtype's.
                    M.calcComp =
                        ((f.type != 21 && f.type != 24) || !prev || prev->rc != f.rc || prev->rl != f.rl) ? 1 : 0;
                    M.walk = 0;
                    for (int r = (int)q - 1; !M.calcComp && r >= 0; --r) {
                        const SF& p = v[g[r]];
                        if (p.core != f.core || p.asym != f.asym) M.walk = 1;
                        if (p.rc != f.rc || p.rl != f.rl || r == 0 || v[g[r - 1]].rc != p.rc || v[g[r - 1]].rl != p.rl)
                            break;
                    }
                }
            }
        }
        sfcount += E.nsf;
        t.max_sf = std::max(t.max_sf, E.nsf);
    }
    t.ngroups = gcount;
    t.nmembers = mcount;

    // Neural networks (Mode::setupNeuralNetwork, Mode.cpp:1227-1512):
    // global topology keywords, per-element node counts.
    int nh = std::atoi(s.get("global_hidden_layers_short").c_str());
    auto nodes = tokens(s.get("global_nodes_short"));
    auto acts = tokens(s.get("global_activation_short"));
    if ((int)nodes.size() != nh || (int)acts.size() != nh + 1)
        fail(NNPX_ERR_MODEL, "inconsistent global NN topology");
    if (nh + 2 > kMaxLayers) fail(NNPX_ERR_UNSUPPORTED, "too many hidden layers");
    std::vector<std::vector<int>> width(ne, std::vector<int>(nh + 2));
    for (int e = 0; e < ne; ++e) {
        width[e][0] = t.el[e].nsf;
        for (int l = 1; l <= nh; ++l) width[e][l] = std::atoi(nodes[l - 1].c_str());
        width[e][nh + 1] = 1;
    }
    for (auto& v : s.all("element_nodes_short")) {
        auto w = tokens(v);
        int e = eidx(w.at(0));
        if ((int)w.size() - 1 != nh) fail(NNPX_ERR_MODEL, "inconsistent per-element NN topology");
        for (int l = 1; l <= nh; ++l) width[e][l] = std::atoi(w[l].c_str());
    }
    bool normNodes = s.has("normalize_nodes");
    int woff = 0;
    for (int e = 0; e < ne; ++e) {
        NNT& N = t.el[e].nn;
        N.nl = nh + 2;
        N.normalize = normNodes ? 1 : 0;
        for (int l = 0; l < N.nl; ++l) {
            N.n[l] = width[e][l];
            // the input layer holds the element's SFs (up to kMaxSF)
            if (N.n[l] < 1 || (l > 0 && N.n[l] > kMaxWidth))
                fail(NNPX_ERR_UNSUPPORTED, "NN layer width out of range");
            N.act[l] = l == 0 ? AF_IDENTITY : activation(acts[l - 1]);
            t.max_width = std::max(t.max_width, N.n[l]);
            if (l > 0) t.max_hidden = std::max(t.max_hidden, N.n[l]);
        }
        for (int l = 1; l < N.nl; ++l) {
            N.woff[l] = woff;
            woff += N.n[l] * N.n[l - 1];
            N.boff[l] = woff;
            woff += N.n[l];
        }
    }
    t.nweights = woff;

    // Scaling (Mode::setupSymmetryFunctionScaling, Mode.cpp:781-909;
    // SymFnc::setScalingType, SymFnc.cpp:52-98).
    int st = ST_NONE;
    bool sc = s.has("scale_symmetry_functions"), ce = s.has("center_symmetry_functions");
    if (sc && !ce) st = ST_SCALE;
    else if (!sc && ce) st = ST_CENTER;
    else if (sc && ce) st = ST_SCALECENTER;
    else if (s.has("scale_symmetry_functions_sigma")) st = ST_SCALESIGMA;
    double Smin = 0.0, Smax = 0.0;
    if (st == ST_SCALE || st == ST_SCALECENTER || st == ST_SCALESIGMA) {
        Smin = s.has("scale_min_short") ? num(s.get("scale_min_short")) : 0.0;
        Smax = s.has("scale_max_short") ? num(s.get("scale_max_short")) : 1.0;
    }
    {
        std::string path = dir + "/scaling.data";
        std::ifstream in(path);
        if (!in) fail(NNPX_ERR_FILE, "could not open file: \"" + path + "\"");
        std::vector<std::string> lines;
        std::string line;
        while (std::getline(in, line))
            if (!line.empty() && line[0] != '#') lines.push_back(line);
        size_t off = 0;
        for (int e = 0; e < ne; ++e) {
            const ElementT& E = t.el[e];
            for (int i = 0; i < E.nsf; ++i) {
                if (off + i >= lines.size()) fail(NNPX_ERR_RANGE, "scaling.data has too few lines");
                auto w = tokens(lines[off + i]);
                if (w.size() < 5) fail(NNPX_ERR_MODEL, "inconsistent scaling statistics");
                if (std::atoi(w[0].c_str()) != e + 1 && std::atoi(w[1].c_str()) != i + 1)
                    fail(NNPX_ERR_MODEL, "inconsistent scaling statistics");
                ScaleT& S = t.scale[E.sf0 + i];
                S.type = st;
                S.Smin = Smin;
                S.Gmin = num(w[2]);
                const double Gmax = num(w[3]);
                S.Gmax = Gmax;
                S.Gmean = num(w[4]);
                double Gsigma = w.size() > 5 ? num(w[5]) : 0.0;
                if (st == ST_SCALE || st == ST_SCALECENTER) S.factor = (Smax - Smin) / (Gmax - S.Gmin);
                else if (st == ST_SCALESIGMA) S.factor = (Smax - Smin) / Gsigma;
                else S.factor = 1.0;
            }
            off += E.nsf;
        }
    }
    // Scaling factors into the group members (SymGrp*::setScalingFactors,
    // SymGrpExpRad.cpp:127-136, SymGrpBaseExpAng.cpp:45-57).
    for (int e = 0; e < ne; ++e) {
        const ElementT& E = t.el[e];
        for (int g = E.grp0; g < E.grp0 + E.ngrp; ++g) {
            const GroupT& G = t.grp[g];
            for (int q = G.mem0; q < G.mem0 + G.nmem; ++q) {
                MemberT& M = t.mem[q];
                M.sfac = t.scale[E.sf0 + M.sf].factor;
                if (G.type == 3 || G.type == 9 || G.type == 13) {
                    M.factorNorm *= M.sfac;
                    M.gnorm = M.factorNorm / M.sfac;
                }
            }
        }
    }
    // Weights (Mode::readNeuralNetworkWeights, Mode.cpp:2331-2350;
    // NeuralNetwork::setConnections, NeuralNetwork.cpp:147-169).
    hm.weights.assign(t.nweights, 0.0);
    for (int e = 0; e < ne; ++e) {
        char name[64];
        std::snprintf(name, sizeof name, "/weights.%03d.data", hm.Z[e]);
        std::string path = dir + name;
        std::ifstream in(path);
        if (!in) fail(NNPX_ERR_FILE, "could not open file: \"" + path + "\"");
        std::vector<double> w;
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#') continue;
            auto x = tokens(line);
            w.push_back(x.empty() ? std::numeric_limits<double>::quiet_NaN() : num(x[0]));
        }
        const NNT& N = t.el[e].nn;
        size_t c = 0;
        for (int l = 1; l < N.nl; ++l) {
            for (int j = 0; j < N.n[l - 1]; ++j)
                for (int k = 0; k < N.n[l]; ++k) {
                    if (c >= w.size()) fail(NNPX_ERR_RANGE, "too few weights in " + path);
                    hm.weights[N.woff[l] + k * N.n[l - 1] + j] = w[c++];
                }
            for (int k = 0; k < N.n[l]; ++k) {
                if (c >= w.size()) fail(NNPX_ERR_RANGE, "too few weights in " + path);
                hm.weights[N.boff[l] + k] = w[c++];
            }
        }
    }
}

}  // namespace nnpx
