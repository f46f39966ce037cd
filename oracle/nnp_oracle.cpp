// CPU FP64 oracle of the 2G-HDNNP energy/force path -- TEST INFRASTRUCTURE.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library; it is the checker, never the product.  It restates n2p2's
// algorithm from scratch; every block cites the reference file:line it
// follows (paths relative to the n2p2 source tree).  Expressions keep the
// reference's association order and the library is compiled with
// -ffp-contract=off so that, with glibc libm, results reproduce the
// reference's golden vectors bit for bit (test/cpp/ExamplePrediction.h).
#include "nnp_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using std::string;
using std::vector;

thread_local string g_err;

// ---------------------------------------------------------------- strings --
// utility.cpp:33-77 (split(reduce(line))): tokens separated by runs of
// blanks/tabs.
vector<string> words(const string& line)
{
    vector<string> out;
    size_t i = 0, n = line.size();
    while (i < n) {
        while (i < n && (line[i] == ' ' || line[i] == '\t')) ++i;
        if (i >= n) break;
        size_t j = i;
        while (j < n && line[j] != ' ' && line[j] != '\t') ++j;
        out.push_back(line.substr(i, j - i));
        i = j;
    }
    return out;
}

string fmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
string fmt(const char* f, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof buf, f, ap);
    va_end(ap);
    return buf;
}

// Element symbols ordered by atomic number (Z = position + 1).
const char* kSymbols[] = {
    "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al",
    "Si", "P", "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe",
    "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr",
    "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn",
    "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm",
    "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",
    "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn",
    "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am", "Cm", "Bk", "Cf",
    "Es", "Fm", "Md", "No"};

int atomic_number(const string& s)
{
    for (size_t i = 0; i < sizeof kSymbols / sizeof kSymbols[0]; ++i)
        if (s == kSymbols[i]) return (int)i + 1;
    throw std::runtime_error("unknown element symbol \"" + s + "\"");
}

// ------------------------------------------------------- core functions --
// CoreFunction.h:214-319 (fdfEXP, fdfPOLY1..4) on x in [0, 1):
// core 0 exponential, 1..4 the polynomials of increasing order.
enum Core { CORE_EXP = 0, CORE_POLY1, CORE_POLY2, CORE_POLY3, CORE_POLY4 };

void core_fdf(int core, double x, double& fx, double& dfx)
{
    static const double E = exp(1.0);
    if (core == CORE_EXP) {
        double const t = 1.0 / (x * x - 1.0);
        double const t2 = exp(t);
        fx = E * t2;
        dfx = -2.0 * E * x * t * t * t2;
    } else if (core == CORE_POLY1) {
        fx = (2.0 * x - 3.0) * x * x + 1.0;
        dfx = x * (6.0 * x - 6.0);
    } else if (core == CORE_POLY2) {
        double const x2 = x * x;
        fx = ((15.0 - 6.0 * x) * x - 10.0) * x * x2 + 1.0;
        dfx = x2 * ((60.0 - 30.0 * x) * x - 30.0);
    } else if (core == CORE_POLY3) {
        double const x2 = x * x;
        fx = (x * (x * (20.0 * x - 70.0) + 84.0) - 35.0) * x2 * x2 + 1.0;
        dfx = x2 * x * (x * (x * (140.0 * x - 420.0) + 420.0) - 140.0);
    } else {
        double x4 = x * x;
        x4 *= x4;
        fx = (x * (x * ((315.0 - 70.0 * x) * x - 540.0) + 420.0) - 126.0) *
             x * x4 + 1.0;
        dfx = x4 *
              (x * (x * ((2520.0 - 630.0 * x) * x - 3780.0) + 2520.0) - 630.0);
    }
}

// A compact function on [left, right]: the core evaluated at
// |x - center| / width, optionally through the asymmetric map x -> (2 - x) x
// (CompactFunction.cpp:58-77 setLeftRight, CompactFunction.h:150-172 fdf,
// CoreFunction.h:197-212 asymmetric fdf).
struct Compact {
    int core = CORE_POLY2;
    bool asym = false;
    double center = 0, width = 0, left = 0, right = 0, scale = 0;
    void set_left_right(double l, double r)
    {
        if (l >= r) throw std::invalid_argument("compact function: left >= right");
        left = l;
        right = r;
        center = (r + l) / 2.0;
        width = (r - l) / 2.0;
        scale = 1.0 / width;
    }
    void fdf(double a, double& fa, double& dfa) const
    {
        a = (a - center) * scale;
        double const x = std::abs(a);
        if (asym) {
            core_fdf(core, (2.0 - x) * x, fa, dfa);
            dfa *= 2.0 - 2.0 * x;
        } else {
            core_fdf(core, x, fa, dfa);
        }
        dfa *= std::copysign(scale, a);
    }
};

// ---------------------------------------------------------------- cutoff --
// CutoffFunction.cpp:25-28 constants, 95-112 set*, 114-213 f/df,
// CutoffFunction.h:210-220 fdf guard, CoreFunction.h:214-319 cores.
struct Cutoff {
    int type = 0;
    double rc = 0, rcinv = 0, rci = 0, alpha = 0, iw = 0;
    void set(int t, double r, double a)
    {
        if (t < 0 || t > 8) throw std::invalid_argument("unknown cutoff type");
        if (a < 0.0 || a >= 1.0)
            throw std::invalid_argument("0 <= alpha < 1.0 is required");
        type = t;
        rc = r;
        rcinv = 1.0 / r;
        alpha = a;
        rci = rc * alpha;
        iw = 1.0 / (rc - rci);
    }
    void fdf(double r, double& fc, double& dfc) const
    {
        static const double PI = 4.0 * atan(1.0);
        static const double E = exp(1.0);
        static const double TANH_PRE = pow((E + 1 / E) / (E - 1 / E), 3);
        if (r >= rc) { fc = 0.0; dfc = 0.0; return; }
        switch (type) {
        case 0: fc = 1.0; dfc = 0.0; return;
        case 1: {
            if (r < rci) { fc = 1.0; dfc = 0.0; return; }
            double const x = (r - rci) * iw;
            double const t = cos(PI * x);
            fc = 0.5 * (t + 1.0);
            dfc = -0.5 * iw * PI * sqrt(1.0 - t * t);
            return;
        }
        case 2: {
            double const t = tanh(1.0 - r * rcinv);
            double const t2 = t * t;
            fc = t * t2;
            dfc = 3.0 * t2 * (t2 - 1.0) * rcinv;
            return;
        }
        case 3: {
            double const t = tanh(1.0 - r * rcinv);
            double const t2 = t * t;
            fc = TANH_PRE * t * t2;
            dfc = 3.0 * TANH_PRE * t2 * (t2 - 1.0) * rcinv;
            return;
        }
        default: break;
        }
        // Polynomial / exponential cores on x in [0, 1).
        if (r < rci) { fc = 1.0; dfc = 0.0; return; }
        double const x = (r - rci) * iw;
        core_fdf(type - 4, x, fc, dfc);
        dfc *= iw;
    }
};

// ------------------------------------------------------- symmetry functions --
enum Scaling { ST_NONE, ST_SCALE, ST_CENTER, ST_SCALECENTER, ST_SCALESIGMA };

// Types: 2 radial, 3 angular (narrow), 9 angular (wide) with Gaussian terms
// and a cutoff function; 20 radial, 21 angular (narrow), 22 angular (wide)
// built from compact functions (SymFncCompRad, SymFncCompAngn,
// SymFncCompAngw).
inline bool is_compact(int type) { return type >= 20; }
inline bool is_radial(int type) { return type == 2 || type == 12 || type == 20 || type == 23; }
// Weighted types (SymFncExpRadWeighted, SymFncExpAngnWeighted,
// SymFncCompRadWeighted, SymFncCompAngnWeighted, SymFncCompAngwWeighted):
// every neighbour element contributes, weighted by its atomic number.
inline bool is_weighted(int type) { return type == 12 || type == 13 || type >= 23; }

struct SF {
    int type = 0;
    int ec = 0, e1 = 0, e2 = 0;
    double eta = 0, rs = 0, rc = 0, lambda = 0, zeta = 0;
    // compact types: lower radial bound, angle window (degrees), subtype
    double rl = 0, angleLeft = 0, angleRight = 0;
    string subtype;
    Compact cr, ca;  // radial and angular compact functions
    int zetaInt = 0;
    bool useIntPow = false;
    int index = 0;                // sorted index within the element
    vector<int> indexPerElement;  // compact derivative slot per neighbour element
    // scaling (SymFnc.cpp:52-98)
    int scalingType = ST_NONE;
    double Smin = 0, Smax = 0, Gmin = 0, Gmax = 0, Gmean = 0, Gsigma = 0;
    double scalingFactor = 1.0;

    bool relevant(int e) const  // SymFnc*::checkRelevantElement
    {
        if (is_weighted(type)) return true;
        return is_radial(type) ? e == e1 : (e == e1 || e == e2);
    }
    // SymFnc.cpp:169-195
    double scale(double v) const
    {
        switch (scalingType) {
        case ST_NONE: return v;
        case ST_SCALE: return Smin + scalingFactor * (v - Gmin);
        case ST_CENTER: return v - Gmean;
        case ST_SCALECENTER: return Smin + scalingFactor * (v - Gmean);
        case ST_SCALESIGMA: return Smin + scalingFactor * (v - Gmean);
        }
        return 0.0;
    }
    // SymFnc::unscale (SymFnc.cpp:197-223)
    double unscale(double v) const
    {
        switch (scalingType) {
        case ST_NONE: return v;
        case ST_SCALE: return (v - Smin) / scalingFactor + Gmin;
        case ST_CENTER: return v + Gmean;
        case ST_SCALECENTER: return (v - Smin) / scalingFactor + Gmean;
        case ST_SCALESIGMA: return (v - Smin) / scalingFactor + Gmean;
        }
        return 0.0;
    }
};

// Sort key of SymFncExpRad::operator< (SymFncExpRad.cpp:57-75) and
// SymFncExpAngn::operator< (SymFncExpAngn.cpp:51-77).  The cutoff type and
// alpha are equal for all functions of one model.
bool sf_less(const SF& a, const SF& b)
{
    if (a.ec != b.ec) return a.ec < b.ec;
    if (a.type != b.type) return a.type < b.type;
    if (a.type >= 23) {
        // SymFncCompRadWeighted.cpp, SymFncCompAng{n,w}Weighted.cpp operator<
        if (a.subtype != b.subtype) return a.subtype < b.subtype;
        if (a.rc != b.rc) return a.rc < b.rc;
        if (a.rl != b.rl) return a.rl < b.rl;
        if (a.type == 23) return false;
        if (a.angleLeft != b.angleLeft) return a.angleLeft < b.angleLeft;
        return a.angleRight < b.angleRight;
    }
    if (a.type == 12 || a.type == 13) {
        // SymFncExpRadWeighted.cpp, SymFncExpAngnWeighted.cpp operator<
        if (a.rc != b.rc) return a.rc < b.rc;
        if (a.eta != b.eta) return a.eta < b.eta;
        if (a.rs != b.rs) return a.rs < b.rs;
        if (a.type == 12) return false;
        if (a.zeta != b.zeta) return a.zeta < b.zeta;
        return a.lambda < b.lambda;
    }
    if (is_compact(a.type)) {
        // SymFncCompRad.cpp:48-68, SymFncCompAngn/SymFncCompAngw.cpp:56-82
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

// One SF group: members sharing (ec, type, rc, e1[, e2])
// (Element.cpp:194-287, SymGrpExpRad.cpp:110-136, SymGrpExpAngn.cpp:108-148,
// SymGrpBaseExpAng.cpp:45-57).
// Compact groups share (ec, type, e1[, e2]) only; their radial range is
// (rmin, rmax) over the members (SymGrpCompRad.cpp:45-77,
// SymGrpCompAngn/SymGrpCompAngw.cpp:46-80, SymGrpBaseComp.cpp:24-31).
struct Group {
    int type = 0, ec = 0, e1 = 0, e2 = 0;
    double rc = 0;
    Cutoff fc;
    vector<int> mem;  // SF indices, sorted like the SF order
    // angular per-member constants
    vector<double> factorNorm, factorDeriv, zetaLambda, scalingFactors;
    vector<char> calcExp;
    // compact groups
    double rmin = 0, rmax = 0;
    vector<double> mrl, mrc, mal, mar;  // member ranges (angles in radians)
    vector<char> calcComp;  // narrow: r_jk term recomputed (SymGrpCompAngn.cpp:108-126)
    vector<double> Z;       // weighted types: atomic number by element index
};

// ------------------------------------------------------------ neural net --
enum Act { AF_IDENTITY, AF_TANH, AF_LOGISTIC, AF_SOFTPLUS, AF_RELU, AF_GAUSSIAN,
           AF_COS, AF_REVLOGISTIC, AF_EXP, AF_HARMONIC };

int act_from(const string& c)  // NeuralNetwork.cpp:1078-1097
{
    if (c == "l") return AF_IDENTITY;
    if (c == "t") return AF_TANH;
    if (c == "s") return AF_LOGISTIC;
    if (c == "p") return AF_SOFTPLUS;
    if (c == "r") return AF_RELU;
    if (c == "g") return AF_GAUSSIAN;
    if (c == "c") return AF_COS;
    if (c == "S") return AF_REVLOGISTIC;
    if (c == "e") return AF_EXP;
    if (c == "h") return AF_HARMONIC;
    throw std::runtime_error("unknown activation function \"" + c + "\"");
}

struct NN {
    vector<int> n;    // neurons per layer (layer 0 = input)
    vector<int> act;  // per layer
    vector<vector<double>> W;  // W[l][k*n[l-1]+j], l >= 1
    vector<vector<double>> b;  // b[l][k]
    bool normalize = false;
    size_t numConnections() const
    {
        size_t c = 0;
        for (size_t l = 1; l < n.size(); ++l) c += (size_t)n[l] * (n[l - 1] + 1);
        return c;
    }
    // NeuralNetwork.cpp:147-169 (setConnections order).
    void setConnections(const vector<double>& c)
    {
        size_t count = 0;
        W.assign(n.size(), {});
        b.assign(n.size(), {});
        for (size_t l = 1; l < n.size(); ++l) {
            W[l].assign((size_t)n[l] * n[l - 1], 0.0);
            b[l].assign(n[l], 0.0);
            for (int j = 0; j < n[l - 1]; ++j)
                for (int k = 0; k < n[l]; ++k) W[l][(size_t)k * n[l - 1] + j] = c.at(count++);
            for (int k = 0; k < n[l]; ++k) b[l][k] = c.at(count++);
        }
    }
};

// Activation (NeuralNetwork.cpp:778-889); returns value, sets dfdx.
double activate(int a, double x, double& dfdx)
{
    const double EXP_LIMIT = 35.0;
    double v;
    switch (a) {
    case AF_IDENTITY: dfdx = 1.0; return x;
    case AF_TANH: v = tanh(x); dfdx = 1.0 - v * v; return v;
    case AF_LOGISTIC:
        if (x > EXP_LIMIT) { dfdx = 0.0; return 1.0; }
        if (x < -EXP_LIMIT) { dfdx = 0.0; return 0.0; }
        v = 1.0 / (1.0 + exp(-x)); dfdx = v * (1.0 - v); return v;
    case AF_SOFTPLUS:
        if (x > EXP_LIMIT) { dfdx = 1.0; return x; }
        if (x < -EXP_LIMIT) { dfdx = 0.0; return 0.0; }
        {
            double const t = exp(x);
            v = log(1.0 + t);
            dfdx = 1.0 / (1.0 + 1.0 / t);
            return v;
        }
    case AF_RELU:
        if (x > 0.0) { dfdx = 1.0; return x; }
        dfdx = 0.0; return 0.0;
    case AF_GAUSSIAN: { double const t = exp(-0.5 * x * x); dfdx = -x * t; return t; }
    case AF_COS: dfdx = -sin(x); return cos(x);
    case AF_REVLOGISTIC:
        v = 1.0 / (1.0 + exp(-x)); dfdx = v * (v - 1.0); return 1.0 - v;
    case AF_EXP: v = exp(-x); dfdx = -v; return v;
    case AF_HARMONIC: dfdx = 2.0 * x; return x * x;
    }
    dfdx = 0.0;
    return 0.0;
}

// --------------------------------------------------------------- elements --
struct Element {
    string symbol;
    int Z = 0;
    double energyOffset = 0.0;
    vector<SF> sfs;      // sorted
    vector<Group> grps;  // sorted
    vector<vector<int>> table;  // symmetryFunctionTable[e]
    NN nn;
};

struct Vec { double x = 0, y = 0, z = 0; };

struct Neighbor {
    int j;
    int bc[3];
    int element;
    double d;
    double dr[3];
    vector<double> dGdr;  // 3 * n_SFD(element)
};

struct Atom {
    int element = 0;
    double r[3] = {0, 0, 0};
    vector<Neighbor> nbr;
    vector<int> unique;  // neighborsUnique (first entry = self)
    vector<double> G, dEdG, dGdr;  // dGdr: 3 * n_SF
    double energy = 0.0;
    double f[3] = {0, 0, 0};
};

}  // namespace

struct oracle_ctx {
    vector<Element> el;
    int cutoffType = 0;
    double cutoffAlpha = 0.0;
    double maxCutoff = 0.0;
    bool normalize = false;
    double meanEnergy = 0, convEnergy = 1, convLength = 1;
    // structure
    vector<Atom> atoms;
    bool periodic = false;
    double box[3][3] = {{0}}, invbox[3][3] = {{0}};
    double energy = 0.0;
    double phase[5] = {0, 0, 0, 0, 0};
};

namespace {

// ----------------------------------------------------------- model setup --
// SymFncBaseComp::setCompactFunction (SymFncBaseComp.cpp:40-88): "p1".."p4"
// polynomials, a trailing "a" for the asymmetric form, "e" exponential.
void set_compact_subtype(SF& s)
{
    const string& t = s.subtype;
    if (t.size() < 1 || t.size() > 3)
        throw std::runtime_error("invalid compact function type specification: \"" + t + "\"");
    if (t[0] == 'p') {
        if (t.size() < 2 || t[1] < '1' || t[1] > '4')
            throw std::runtime_error("invalid polynom type: \"" + t + "\"");
        s.cr.core = CORE_POLY1 + (t[1] - '1');
        if (t.size() == 3) {
            if (t[2] != 'a') throw std::runtime_error("invalid polynom specifier: \"" + t + "\"");
            s.cr.asym = true;
        }
    } else if (t[0] == 'e') {
        s.cr.core = CORE_EXP;
    } else {
        throw std::runtime_error("unknown compact SF type: \"" + t + "\"");
    }
}

void load_model(oracle_ctx& c, const string& dir)
{
    // Settings::readFile / parseLines (Settings.cpp:188-392).
    std::ifstream in(dir + "/input.nn");
    if (!in) throw std::runtime_error("could not open " + dir + "/input.nn");
    std::multimap<string, string> kv;
    vector<std::pair<string, string>> ordered;
    string line;
    while (std::getline(in, line)) {
        if (line.empty() || line == "\r") continue;
        size_t p = line.find('#');
        if (p != string::npos) line.erase(p);
        p = line.find('!');
        if (p != string::npos) line.erase(p);
        vector<string> w = words(line);
        if (w.empty()) continue;
        string key = w[0], val;
        for (size_t i = 1; i < w.size(); ++i) val += (i > 1 ? " " : "") + w[i];
        if (key == "symfunction") key = "symfunction_short";
        kv.emplace(key, val);
        ordered.emplace_back(key, val);
    }
    auto has = [&](const char* k) { return kv.count(k) > 0; };
    auto get = [&](const char* k) {
        auto it = kv.find(k);
        if (it == kv.end()) throw std::runtime_error(string("keyword \"") + k + "\" not found");
        return it->second;
    };
    if (has("nnp_type")) {
        string t = get("nnp_type");
        if (t != "2G-HDNNP" && t != "2")
            throw std::runtime_error("only 2G-HDNNP models are supported");
    }
    // Mode::setupNormalization (Mode.cpp:239-300).
    int nn_keys = has("mean_energy") + has("conv_energy") + has("conv_length");
    if (nn_keys == 3) {
        c.normalize = true;
        c.meanEnergy = atof(get("mean_energy").c_str());
        c.convEnergy = atof(get("conv_energy").c_str());
        c.convLength = atof(get("conv_length").c_str());
    } else if (nn_keys != 0) {
        throw std::runtime_error("incorrect usage of normalization keywords");
    }
    // ElementMap::registerElements (ElementMap.cpp:36-50): sort by Z.
    vector<string> syms = words(get("elements"));
    std::sort(syms.begin(), syms.end(), [](const string& a, const string& b) {
        return atomic_number(a) < atomic_number(b);
    });
    if ((size_t)atoi(get("number_of_elements").c_str()) != syms.size())
        throw std::runtime_error("inconsistent number of elements");
    auto eidx = [&](const string& s) {
        for (size_t i = 0; i < syms.size(); ++i) if (syms[i] == s) return (int)i;
        throw std::runtime_error("element \"" + s + "\" not in element map");
    };
    c.el.assign(syms.size(), Element());
    for (size_t i = 0; i < syms.size(); ++i) {
        c.el[i].symbol = syms[i];
        c.el[i].Z = atomic_number(syms[i]);
    }
    for (auto it = kv.lower_bound("atom_energy"); it != kv.upper_bound("atom_energy"); ++it) {
        vector<string> a = words(it->second);
        c.el.at(eidx(a.at(0))).energyOffset = atof(a.at(1).c_str());
    }
    // Mode::setupCutoff (Mode.cpp:589-682).
    vector<string> ct = words(get("cutoff_type"));
    c.cutoffType = atoi(ct.at(0).c_str());
    if (ct.size() > 1) c.cutoffAlpha = atof(ct.at(1).c_str());
    // Mode::setupSymmetryFunctions (Mode.cpp:684-759), parameter parsing
    // SymFncExpRad.cpp:77-96, SymFncBaseExpAng.cpp:50-93.
    for (auto& kvp : ordered) {
        if (kvp.first != "symfunction_short") continue;
        vector<string> a = words(kvp.second);
        SF s;
        s.ec = eidx(a.at(0));
        s.type = atoi(a.at(1).c_str());
        if (s.type == 2) {
            s.e1 = eidx(a.at(2));
            s.eta = atof(a.at(3).c_str());
            s.rs = atof(a.at(4).c_str());
            s.rc = atof(a.at(5).c_str());
        } else if (s.type == 3 || s.type == 9) {  // SymFncExpAngw.cpp parameters match type 3
            s.e1 = eidx(a.at(2));
            s.e2 = eidx(a.at(3));
            s.eta = atof(a.at(4).c_str());
            s.lambda = atof(a.at(5).c_str());
            s.zeta = atof(a.at(6).c_str());
            s.rc = atof(a.at(7).c_str());
            if (a.size() > 8) s.rs = atof(a.at(8).c_str());
            if (s.e1 > s.e2) std::swap(s.e1, s.e2);
            s.zetaInt = (int)std::round(s.zeta);
            s.useIntPow = fabs(s.zeta - s.zetaInt) <= std::numeric_limits<double>::min();
        } else if (s.type == 12) {  // SymFncExpRadWeighted.cpp setParameters
            s.eta = atof(a.at(2).c_str());
            s.rs = atof(a.at(3).c_str());
            s.rc = atof(a.at(4).c_str());
        } else if (s.type == 13) {  // SymFncExpAngnWeighted.cpp setParameters
            s.eta = atof(a.at(2).c_str());
            s.rs = atof(a.at(3).c_str());
            s.lambda = atof(a.at(4).c_str());
            s.zeta = atof(a.at(5).c_str());
            s.rc = atof(a.at(6).c_str());
            s.zetaInt = (int)std::round(s.zeta);
            s.useIntPow = fabs(s.zeta - s.zetaInt) <= std::numeric_limits<double>::min();
        } else if (s.type == 23) {  // SymFncCompRadWeighted.cpp setParameters
            s.rl = atof(a.at(2).c_str());
            s.rc = atof(a.at(3).c_str());
            s.subtype = a.at(4);
            if (s.rl > s.rc) throw std::runtime_error("lower radial boundary >= upper radial boundary");
            set_compact_subtype(s);
        } else if (s.type == 24 || s.type == 25) {  // SymFncBaseCompAngWeighted.cpp setParameters
            s.rl = atof(a.at(2).c_str());
            s.rc = atof(a.at(3).c_str());
            s.angleLeft = atof(a.at(4).c_str());
            s.angleRight = atof(a.at(5).c_str());
            s.subtype = a.at(6);
            if (s.rl > s.rc) throw std::runtime_error("lower radial boundary >= upper radial boundary");
            set_compact_subtype(s);
            if (s.angleLeft >= s.angleRight)
                throw std::runtime_error("left angle boundary right of or equal to right angle boundary");
            double const center = (s.angleLeft + s.angleRight) / 2.0;
            if ((s.angleLeft < 0.0 && center != 0.0) || (s.angleRight > 180.0 && center != 180.0))
                throw std::runtime_error("angle boundary out of [0,180] and center of angular function /= 0 or /= 180");
            if (s.angleRight - s.angleLeft > 360.0)
                throw std::runtime_error("periodic symmetry function cannot spread over domain > 360 degrees");
            s.ca.core = s.cr.core;
            s.ca.set_left_right(s.angleLeft * M_PI / 180.0, s.angleRight * M_PI / 180.0);
        } else if (s.type == 20) {  // SymFncCompRad.cpp:85-112
            s.e1 = eidx(a.at(2));
            s.rl = atof(a.at(3).c_str());
            s.rc = atof(a.at(4).c_str());
            s.subtype = a.at(5);
            if (s.rl > s.rc) throw std::runtime_error("lower radial boundary >= upper radial boundary");
            set_compact_subtype(s);
        } else if (s.type == 21 || s.type == 22) {  // SymFncBaseCompAng.cpp:47-110
            s.e1 = eidx(a.at(2));
            s.e2 = eidx(a.at(3));
            s.rl = atof(a.at(4).c_str());
            s.rc = atof(a.at(5).c_str());
            s.angleLeft = atof(a.at(6).c_str());
            s.angleRight = atof(a.at(7).c_str());
            s.subtype = a.at(8);
            if (s.e1 > s.e2) std::swap(s.e1, s.e2);
            if (s.rl > s.rc) throw std::runtime_error("lower radial boundary >= upper radial boundary");
            set_compact_subtype(s);
            if (s.angleLeft >= s.angleRight)
                throw std::runtime_error("left angle boundary right of or equal to right angle boundary");
            double const center = (s.angleLeft + s.angleRight) / 2.0;
            if ((s.angleLeft < 0.0 && center != 0.0) || (s.angleRight > 180.0 && center != 180.0))
                throw std::runtime_error("angle boundary out of [0,180] and center of angular function /= 0 or /= 180");
            if (s.angleRight - s.angleLeft > 360.0)
                throw std::runtime_error("periodic symmetry function cannot spread over domain > 360 degrees");
            // the angular core has the radial core's type, never asymmetric
            s.ca.core = s.cr.core;
            s.ca.set_left_right(s.angleLeft * M_PI / 180.0, s.angleRight * M_PI / 180.0);
        } else {
            throw std::runtime_error(fmt("symmetry function type %d not supported", s.type));
        }
        c.el.at(s.ec).sfs.push_back(s);
    }
    c.maxCutoff = 0.0;
    for (auto& e : c.el) {
        if (c.normalize) {  // SymFnc*::changeLengthUnit
            for (auto& s : e.sfs) {
                s.eta /= c.convLength * c.convLength;
                s.rs *= c.convLength;
                s.rc *= c.convLength;
                s.rl *= c.convLength;
            }
        }
        for (auto& s : e.sfs)
            if (is_compact(s.type)) s.cr.set_left_right(s.rl, s.rc);
        std::sort(e.sfs.begin(), e.sfs.end(), sf_less);
        for (size_t i = 0; i < e.sfs.size(); ++i) {
            e.sfs[i].index = (int)i;
            c.maxCutoff = std::max(c.maxCutoff, e.sfs[i].rc);
        }
        // Element::setupSymmetryFunctionMemory (Element.cpp:289-311).
        e.table.assign(c.el.size(), {});
        for (auto& s : e.sfs) {
            s.indexPerElement.assign(c.el.size(), -1);
            for (size_t ne = 0; ne < c.el.size(); ++ne)
                if (s.relevant((int)ne)) {
                    s.indexPerElement[ne] = (int)e.table[ne].size();
                    e.table[ne].push_back(s.index);
                }
        }
        // Element::setupSymmetryFunctionGroups (Element.cpp:194-287).
        for (auto& s : e.sfs) {
            bool added = false;
            for (auto& g : e.grps) {
                if (g.type == s.type && (is_compact(s.type) || g.rc == s.rc) && g.e1 == s.e1 &&
                    (is_radial(s.type) || g.e2 == s.e2)) {
                    g.mem.push_back(s.index);
                    added = true;
                    break;
                }
            }
            if (!added) {
                Group g;
                g.type = s.type; g.ec = s.ec; g.e1 = s.e1; g.e2 = s.e2; g.rc = s.rc;
                if (!is_compact(s.type)) g.fc.set(c.cutoffType, s.rc, c.cutoffAlpha);
                g.mem.push_back(s.index);
                e.grps.push_back(g);
            }
        }
        // Compact groups: member ranges and the r_jk recompute flags
        // (SymGrpCompRad.cpp:79-107, SymGrpCompAngn.cpp:77-126,
        // SymGrpCompAngw.cpp:82-115).
        for (auto& g : e.grps) {
            if (!is_compact(g.type)) continue;
            for (size_t i = 0; i < g.mem.size(); ++i) {
                const SF& s = e.sfs[g.mem[i]];
                if (s.rl <= 0.0) g.rmin = 0.0;
                else g.rmin = std::min(g.rmin, s.rl);
                g.rmax = std::max(g.rmax, s.rc);
                g.mrl.push_back(s.rl);
                g.mrc.push_back(s.rc);
                g.mal.push_back(s.angleLeft * M_PI / 180.0);
                g.mar.push_back(s.angleRight * M_PI / 180.0);
                const SF* prev = i ? &e.sfs[g.mem[i - 1]] : nullptr;
                g.calcComp.push_back((g.type != 21 && g.type != 24) || !prev || prev->rc != s.rc ||
                                     prev->rl != s.rl);
            }
        }
        // Members are in SF order already (SF sort refines the member sort).
        for (auto& g : e.grps) {
            for (auto& el : c.el) g.Z.push_back((double)el.Z);
            // SymGrpExpAngn/SymGrpExpAngw/SymGrpExpAngnWeighted::sortMembers
            // (SymGrpExpAngn.cpp:108-148): the same member set-up
            if (g.type != 3 && g.type != 9 && g.type != 13) continue;
            for (size_t i = 0; i < g.mem.size(); ++i) {
                const SF& s = e.sfs[g.mem[i]];
                g.factorNorm.push_back(pow(2.0, 1.0 - s.zeta));
                g.factorDeriv.push_back(2.0 * s.eta / s.zeta / s.lambda);
                g.zetaLambda.push_back(s.zeta * s.lambda);
                const SF* prev = i ? &e.sfs[g.mem[i - 1]] : nullptr;
                g.calcExp.push_back(!prev || prev->eta != s.eta || prev->rs != s.rs);
            }
        }
    }
    // Mode::setupNeuralNetwork (Mode.cpp:1227-1512), 2G, global keywords.
    int nh = atoi(get("global_hidden_layers_short").c_str());
    vector<string> nodes = words(get("global_nodes_short"));
    vector<string> acts = words(get("global_activation_short"));
    if ((int)nodes.size() != nh || (int)acts.size() != nh + 1)
        throw std::runtime_error("inconsistent global NN topology");
    for (auto& e : c.el) {
        e.nn.n.assign(nh + 2, 0);
        e.nn.act.assign(nh + 2, AF_IDENTITY);
        e.nn.n[0] = (int)e.sfs.size();
        for (int l = 1; l <= nh; ++l) {
            e.nn.n[l] = atoi(nodes[l - 1].c_str());
            e.nn.act[l] = act_from(acts[l - 1]);
        }
        e.nn.n[nh + 1] = 1;
        e.nn.act[nh + 1] = act_from(acts[nh]);
        e.nn.normalize = has("normalize_nodes");
    }
    for (auto it = kv.lower_bound("element_nodes_short"); it != kv.upper_bound("element_nodes_short"); ++it) {
        vector<string> a = words(it->second);
        Element& e = c.el.at(eidx(a.at(0)));
        if ((int)a.size() - 1 != nh) throw std::runtime_error("inconsistent per-element NN topology");
        for (int l = 1; l <= nh; ++l) e.nn.n[l] = atoi(a.at(l).c_str());
    }
    // Mode::setupSymmetryFunctionScaling (Mode.cpp:781-909).
    int st = ST_NONE;
    bool sc = has("scale_symmetry_functions"), ce = has("center_symmetry_functions");
    if (sc && !ce) st = ST_SCALE;
    else if (!sc && ce) st = ST_CENTER;
    else if (sc && ce) st = ST_SCALECENTER;
    else if (has("scale_symmetry_functions_sigma")) st = ST_SCALESIGMA;
    double Smin = 0.0, Smax = 0.0;
    if (st == ST_SCALE || st == ST_SCALECENTER || st == ST_SCALESIGMA) {
        Smin = has("scale_min_short") ? atof(get("scale_min_short").c_str()) : 0.0;
        Smax = has("scale_max_short") ? atof(get("scale_max_short").c_str()) : 1.0;
    }
    std::ifstream sf(dir + "/scaling.data");
    if (!sf) throw std::runtime_error("could not open " + dir + "/scaling.data");
    vector<string> lines;
    while (std::getline(sf, line))
        if (!line.empty() && line[0] != '#') lines.push_back(line);
    size_t off = 0;
    for (auto& e : c.el) {
        for (auto& s : e.sfs) {  // SymFnc::setScalingType (SymFnc.cpp:52-98)
            vector<string> a = words(lines.at(off + s.index));
            if ((size_t)atoi(a.at(0).c_str()) != (size_t)s.ec + 1 &&
                (size_t)atoi(a.at(1).c_str()) != (size_t)s.index + 1)
                throw std::runtime_error("inconsistent scaling statistics");
            s.scalingType = st;
            s.Gmin = atof(a.at(2).c_str());
            s.Gmax = atof(a.at(3).c_str());
            s.Gmean = atof(a.at(4).c_str());
            if (a.size() > 5) s.Gsigma = atof(a.at(5).c_str());
            s.Smin = Smin;
            s.Smax = Smax;
            if (st == ST_SCALE || st == ST_SCALECENTER) s.scalingFactor = (Smax - Smin) / (s.Gmax - s.Gmin);
            else if (st == ST_SCALESIGMA) s.scalingFactor = (Smax - Smin) / s.Gsigma;
            else s.scalingFactor = 1.0;
        }
        off += e.sfs.size();
        for (auto& g : e.grps) {  // SymGrp*::setScalingFactors
            g.scalingFactors.clear();
            for (size_t i = 0; i < g.mem.size(); ++i) {
                g.scalingFactors.push_back(e.sfs[g.mem[i]].scalingFactor);
                if (g.type == 3 || g.type == 9 || g.type == 13) g.factorNorm[i] *= g.scalingFactors[i];
            }
        }
    }
    // Mode::readNeuralNetworkWeights (Mode.cpp:2331-2350),
    // readColumnsFromFile (utility.cpp:247-290), column 0.
    for (auto& e : c.el) {
        string fn = dir + fmt("/weights.%03d.data", e.Z);
        std::ifstream wf(fn);
        if (!wf) throw std::runtime_error("could not open " + fn);
        vector<double> w;
        while (std::getline(wf, line)) {
            if (line.empty() || line[0] == '#') continue;
            vector<string> a = words(line);
            w.push_back(a.empty() ? std::numeric_limits<double>::quiet_NaN() : atof(a[0].c_str()));
        }
        if (w.size() < e.nn.numConnections())
            throw std::runtime_error("too few weights in " + fn);
        e.nn.setConnections(w);
    }
}

// ------------------------------------------------------------- structure --
// Structure::calculateInverseBox (Structure.cpp:518-544).
void inverse_box(oracle_ctx& c)
{
    double (&b)[3][3] = c.box;
    double (&ib)[3][3] = c.invbox;
    double invdet = b[0][0] * b[1][1] * b[2][2] + b[1][0] * b[2][1] * b[0][2] +
                    b[2][0] * b[0][1] * b[1][2] - b[2][0] * b[1][1] * b[0][2] -
                    b[0][0] * b[2][1] * b[1][2] - b[1][0] * b[0][1] * b[2][2];
    invdet = 1.0 / invdet;
    ib[0][0] = b[1][1] * b[2][2] - b[2][1] * b[1][2];
    ib[0][1] = b[2][1] * b[0][2] - b[0][1] * b[2][2];
    ib[0][2] = b[0][1] * b[1][2] - b[1][1] * b[0][2];
    ib[1][0] = b[2][0] * b[1][2] - b[1][0] * b[2][2];
    ib[1][1] = b[0][0] * b[2][2] - b[2][0] * b[0][2];
    ib[1][2] = b[1][0] * b[0][2] - b[0][0] * b[1][2];
    ib[2][0] = b[1][0] * b[2][1] - b[2][0] * b[1][1];
    ib[2][1] = b[2][0] * b[0][1] - b[0][0] * b[2][1];
    ib[2][2] = b[0][0] * b[1][1] - b[1][0] * b[0][1];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) ib[i][k] *= invdet;
}

// Structure::remap (Structure.cpp:1224-1248).
void remap(oracle_ctx& c, Atom& a)
{
    double f[3];
    for (int k = 0; k < 3; ++k)
        f[k] = a.r[0] * c.invbox[0][k] + a.r[1] * c.invbox[1][k] + a.r[2] * c.invbox[2][k];
    for (int k = 0; k < 3; ++k) if (f[k] > 1.0) f[k] -= (int)f[k];
    for (int k = 0; k < 3; ++k) if (f[k] < 0.0) f[k] += 1.0 - (int)f[k];
    for (int k = 0; k < 3; ++k) if (f[k] == 1.0) f[k] = 0.0;
    for (int k = 0; k < 3; ++k)
        a.r[k] = f[0] * c.box[0][k] + f[1] * c.box[1][k] + f[2] * c.box[2][k];
}

// Structure::toNormalizedUnits (Structure.cpp:1250-1277) for positions/box.
void to_normalized(oracle_ctx& c)
{
    if (!c.normalize) return;
    if (c.periodic) {
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) c.box[i][k] *= c.convLength;
        double const inv = 1.0 / c.convLength;  // Vec3D::operator/=
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) c.invbox[i][k] *= inv;
    }
    for (auto& a : c.atoms)
        for (int k = 0; k < 3; ++k) a.r[k] *= c.convLength;
}

// Structure::calculatePbcCopies (Structure.cpp:485-516).
void pbc_copies(const oracle_ctx& c, double rc, int pbc[3])
{
    auto cross = [](const double* u, const double* v, double* w) {
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
    };
    auto normalize = [](double* w) {
        double n = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        w[0] /= n; w[1] /= n; w[2] /= n;
    };
    auto dot = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    double axb[3], axc[3], bxc[3];
    cross(c.box[0], c.box[1], axb); normalize(axb);
    cross(c.box[0], c.box[2], axc); normalize(axc);
    cross(c.box[1], c.box[2], bxc); normalize(bxc);
    double proj[3] = {fabs(dot(c.box[0], bxc)), fabs(dot(c.box[1], axc)), fabs(dot(c.box[2], axb))};
    for (int k = 0; k < 3; ++k) {
        pbc[k] = 0;
        while (pbc[k] * proj[k] <= rc) pbc[k]++;
    }
}

// Distance vector of the (j, image) pair, in the reference's evaluation order
// (Structure.cpp:322-325): ((r_i - r_j) + bc0*a) + bc1*b + bc2*c.
inline void pair_dr(const oracle_ctx& c, const Atom& ai, const Atom& aj, int b0, int b1, int b2, double* dr)
{
    for (int k = 0; k < 3; ++k)
        dr[k] = ai.r[k] - aj.r[k] + b0 * c.box[0][k] + b1 * c.box[1][k] + b2 * c.box[2][k];
}

void push_neighbor(Atom& ai, int i, int j, const int* bc, int ej, const double* dr)
{
    Neighbor n;
    n.j = j;
    n.bc[0] = bc[0]; n.bc[1] = bc[1]; n.bc[2] = bc[2];
    n.element = ej;
    n.d = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    n.dr[0] = dr[0]; n.dr[1] = dr[1]; n.dr[2] = dr[2];
    ai.nbr.push_back(n);
    if (ai.unique.back() != j && i != j) ai.unique.push_back(j);
}

// Structure::calculateNeighborList (Structure.cpp:287-398), brute force.
void neighbors_brute(oracle_ctx& c, double rc)
{
    const int N = (int)c.atoms.size();
    int pbc[3] = {0, 0, 0};
    if (c.periodic) pbc_copies(c, rc, pbc);
    double const rc2 = rc * rc;
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < N; ++i) {
        Atom& ai = c.atoms[i];
        ai.nbr.clear();
        ai.unique.assign(1, i);
        for (int j = 0; j < N; ++j) {
            const Atom& aj = c.atoms[j];
            for (int b0 = -pbc[0]; b0 <= pbc[0]; ++b0)
                for (int b1 = -pbc[1]; b1 <= pbc[1]; ++b1)
                    for (int b2 = -pbc[2]; b2 <= pbc[2]; ++b2) {
                        if (i == j && b0 == 0 && b1 == 0 && b2 == 0) continue;
                        double dr[3];
                        pair_dr(c, ai, aj, b0, b1, b2, dr);
                        if (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] <= rc2) {
                            int bc[3] = {b0, b1, b2};
                            push_neighbor(ai, i, j, bc, aj.element, dr);
                        }
                    }
        }
    }
}

// Cell list producing the same (j, image) set in the same order.  Cells are
// taken in fractional coordinates; each axis has nc_k >= 3 cells of
// perpendicular width >= rc.  Candidates from the 27 surrounding cells are
// tested with the same expression and collected, then sorted by
// (j, bc0, bc1, bc2), which is the brute-force loop order.  Requires
// positions inside the cell (fractional coordinates in [0, 1)), which
// Structure::remap guarantees; falls back to brute force otherwise.
void neighbors_cell(oracle_ctx& c, double rc)
{
    const int N = (int)c.atoms.size();
    if (!c.periodic) { neighbors_brute(c, rc); return; }
    int pbc[3];
    pbc_copies(c, rc, pbc);
    // perpendicular widths: proj_k from the same construction
    int nc[3];
    {
        auto cross = [](const double* u, const double* v, double* w) {
            w[0] = u[1] * v[2] - u[2] * v[1];
            w[1] = u[2] * v[0] - u[0] * v[2];
            w[2] = u[0] * v[1] - u[1] * v[0];
        };
        double n[3][3];
        cross(c.box[1], c.box[2], n[0]);
        cross(c.box[2], c.box[0], n[1]);
        cross(c.box[0], c.box[1], n[2]);
        for (int k = 0; k < 3; ++k) {
            double len = sqrt(n[k][0] * n[k][0] + n[k][1] * n[k][1] + n[k][2] * n[k][2]);
            double proj = fabs(c.box[k][0] * n[k][0] + c.box[k][1] * n[k][1] + c.box[k][2] * n[k][2]) / len;
            nc[k] = (int)floor(proj / (rc * (1.0 + 1e-12) + 1e-12));
        }
    }
    if (nc[0] < 3 || nc[1] < 3 || nc[2] < 3) { neighbors_brute(c, rc); return; }
    vector<int> cell(N);
    vector<double> frac(3 * (size_t)N);
    for (int i = 0; i < N; ++i) {
        const Atom& a = c.atoms[i];
        int ci[3];
        for (int k = 0; k < 3; ++k) {
            double f = a.r[0] * c.invbox[0][k] + a.r[1] * c.invbox[1][k] + a.r[2] * c.invbox[2][k];
            if (!(f >= 0.0 && f < 1.0)) { neighbors_brute(c, rc); return; }
            frac[3 * i + k] = f;
            ci[k] = std::min((int)(f * nc[k]), nc[k] - 1);
        }
        cell[i] = (ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
    }
    const int ncell = nc[0] * nc[1] * nc[2];
    vector<int> start(ncell + 1, 0), order(N);
    for (int i = 0; i < N; ++i) start[cell[i] + 1]++;
    for (int k = 0; k < ncell; ++k) start[k + 1] += start[k];
    {
        vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < N; ++i) order[fill[cell[i]]++] = i;
    }
    double const rc2 = rc * rc;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < N; ++i) {
        Atom& ai = c.atoms[i];
        int ci[3] = {cell[i] / (nc[1] * nc[2]), (cell[i] / nc[2]) % nc[1], cell[i] % nc[2]};
        struct Cand { long key; int j; int bc[3]; double dr[3]; };
        vector<Cand> cand;
        for (int d0 = -1; d0 <= 1; ++d0)
            for (int d1 = -1; d1 <= 1; ++d1)
                for (int d2 = -1; d2 <= 1; ++d2) {
                    int cc[3] = {ci[0] + d0, ci[1] + d1, ci[2] + d2};
                    int bc[3];
                    for (int k = 0; k < 3; ++k) {
                        // neighbour cell wraps: image of j is r_j + m*box,
                        // i.e. bc = -m in the reference's dr expression
                        int m = cc[k] < 0 ? -1 : (cc[k] >= nc[k] ? 1 : 0);
                        cc[k] -= m * nc[k];
                        bc[k] = -m;
                    }
                    int id = (cc[0] * nc[1] + cc[1]) * nc[2] + cc[2];
                    for (int p = start[id]; p < start[id + 1]; ++p) {
                        int j = order[p];
                        if (i == j && bc[0] == 0 && bc[1] == 0 && bc[2] == 0) continue;
                        double dr[3];
                        pair_dr(c, ai, c.atoms[j], bc[0], bc[1], bc[2], dr);
                        if (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] <= rc2) {
                            Cand x;
                            x.j = j;
                            x.bc[0] = bc[0]; x.bc[1] = bc[1]; x.bc[2] = bc[2];
                            x.dr[0] = dr[0]; x.dr[1] = dr[1]; x.dr[2] = dr[2];
                            x.key = (((long)j * 3 + (bc[0] + 1)) * 3 + (bc[1] + 1)) * 3 + (bc[2] + 1);
                            cand.push_back(x);
                        }
                    }
                }
        std::sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) { return a.key < b.key; });
        ai.nbr.clear();
        ai.unique.assign(1, i);
        for (auto& x : cand) push_neighbor(ai, i, x.j, x.bc, c.atoms[x.j].element, x.dr);
    }
    (void)pbc;
}

// ------------------------------------------------------ SF group kernels --
// SymGrpExpRad::calculate (SymGrpExpRad.cpp:144-239).
void calc_radial(const Element& e, const Group& g, Atom& a)
{
    const size_t M = g.mem.size();
    vector<double> result(M, 0.0);
    for (auto& n : a.nbr) {
        if (!(g.e1 == n.element && n.d < g.rc)) continue;
        double const rij = n.d;
        double pfc, pdfc;
        g.fc.fdf(rij, pfc, pdfc);
        for (size_t k = 0; k < M; ++k) {
            const SF& s = e.sfs[g.mem[k]];
            double const pexp = exp(-s.eta * (rij - s.rs) * (rij - s.rs));
            result[k] += pexp * pfc;
            double const p1 = g.scalingFactors[k] * (pdfc - 2.0 * s.eta * (rij - s.rs) * pfc) * pexp / rij;
            double const px = p1 * n.dr[0], py = p1 * n.dr[1], pz = p1 * n.dr[2];
            double* d = &a.dGdr[3 * s.index];
            d[0] += px; d[1] += py; d[2] += pz;
            d = &n.dGdr[3 * s.indexPerElement[g.e1]];
            d[0] -= px; d[1] -= py; d[2] -= pz;
        }
    }
    for (size_t k = 0; k < M; ++k) {
        const SF& s = e.sfs[g.mem[k]];
        a.G[s.index] = s.scale(result[k]);
    }
}

// SymGrpExpRadWeighted::calculate (SymGrpExpRadWeighted.cpp:137-231): every
// neighbour inside rc, each term weighted by its atomic number.
void calc_radial_weighted(const Element& e, const Group& g, Atom& a)
{
    const size_t M = g.mem.size();
    vector<double> result(M, 0.0);
    for (auto& n : a.nbr) {
        if (!(n.d < g.rc)) continue;
        double const rij = n.d;
        int const nej = n.element;
        double pfc, pdfc;
        g.fc.fdf(rij, pfc, pdfc);
        for (size_t k = 0; k < M; ++k) {
            const SF& s = e.sfs[g.mem[k]];
            double const pexp = g.Z[nej] * exp(-s.eta * (rij - s.rs) * (rij - s.rs));
            result[k] += pexp * pfc;
            double const p1 = g.scalingFactors[k] * (pdfc - 2.0 * s.eta * (rij - s.rs) * pfc) * pexp / rij;
            double const px = p1 * n.dr[0], py = p1 * n.dr[1], pz = p1 * n.dr[2];
            double* d = &a.dGdr[3 * s.index];
            d[0] += px; d[1] += py; d[2] += pz;
            d = &n.dGdr[3 * s.indexPerElement[nej]];
            d[0] -= px; d[1] -= py; d[2] -= pz;
        }
    }
    for (size_t k = 0; k < M; ++k) {
        const SF& s = e.sfs[g.mem[k]];
        a.G[s.index] = s.scale(result[k]);
    }
}

double pow_int(double x, int n);

// SymGrpExpAngnWeighted::calculate (SymGrpExpAngnWeighted.cpp:165-395): all
// neighbour pairs j < k inside rc with r_jk < rc, the cutoff product
// weighted by Z_j Z_k, the exponent always in shifted form.
void calc_angular_weighted(const Element& e, const Group& g, Atom& a)
{
    const size_t M = g.mem.size();
    vector<double> result(M, 0.0);
    double const rc = g.rc, rc2 = rc * rc;
    size_t nn = a.nbr.size();
    if (nn == 0) nn = 1;
    for (size_t j = 0; j + 1 < nn; ++j) {
        Neighbor& nj = a.nbr[j];
        int const nej = nj.element;
        double const rij = nj.d;
        if (!(rij < rc)) continue;
        double const r2ij = rij * rij;
        double pfcij, pdfcij;
        g.fc.fdf(rij, pfcij, pdfcij);
        const double* dr1 = nj.dr;
        for (size_t k = j + 1; k < nn; ++k) {
            Neighbor& nk = a.nbr[k];
            int const nek = nk.element;
            double const rik = nk.d;
            if (!(rik < rc)) continue;
            const double* dr2 = nk.dr;
            double dr3[3] = {dr2[0] - dr1[0], dr2[1] - dr1[1], dr2[2] - dr1[2]};
            double rjk = dr3[0] * dr3[0] + dr3[1] * dr3[1] + dr3[2] * dr3[2];
            if (!(rjk < rc2)) continue;
            double pfcik, pdfcik;
            g.fc.fdf(rik, pfcik, pdfcik);
            rjk = sqrt(rjk);
            double pfcjk, pdfcjk;
            g.fc.fdf(rjk, pfcjk, pdfcjk);
            double const rinvijik = 1.0 / rij / rik;
            double const costijk = (dr1[0] * dr2[0] + dr1[1] * dr2[1] + dr1[2] * dr2[2]) * rinvijik;
            double const z = g.Z[nej] * g.Z[nek];
            double const pfc = z * pfcij * pfcik * pfcjk;
            double const r2ik = rik * rik;
            double const pr1 = z * pfcik * pfcjk * pdfcij / rij;
            double const pr2 = z * pfcij * pfcjk * pdfcik / rik;
            double const pr3 = z * pfcij * pfcik * pdfcjk / rjk;
            double vexp = 0.0, rijs = 0.0, riks = 0.0, rjks = 0.0;
            for (size_t l = 0; l < M; ++l) {
                const SF& s = e.sfs[g.mem[l]];
                if (g.calcExp[l]) {
                    rijs = rij - s.rs;
                    riks = rik - s.rs;
                    rjks = rjk - s.rs;
                    double const r2sum = rijs * rijs + riks * riks + rjks * rjks;
                    vexp = exp(-s.eta * r2sum);
                }
                double const plambda = 1.0 + s.lambda * costijk;
                double fg = vexp;
                if (plambda <= 0.0) fg = 0.0;
                else if (s.useIntPow) fg *= pow_int(plambda, s.zetaInt - 1);
                else fg *= pow(plambda, s.zeta - 1.0);
                result[l] += fg * plambda * pfc;
                fg *= g.factorNorm[l];
                double const pfczl = pfc * g.zetaLambda[l];
                double const p2etapl = plambda * g.factorDeriv[l];
                double const p1 = fg * (pfczl * (rinvijik - costijk / r2ij - p2etapl * rijs / rij) + pr1 * plambda);
                double const p2 = fg * (pfczl * (rinvijik - costijk / r2ik - p2etapl * riks / rik) + pr2 * plambda);
                double const p3 = fg * (pfczl * (rinvijik + p2etapl * rjks / rjk) - pr3 * plambda);
                double const p1x = p1 * dr1[0], p1y = p1 * dr1[1], p1z = p1 * dr1[2];
                double const p2x = p2 * dr2[0], p2y = p2 * dr2[1], p2z = p2 * dr2[2];
                double const p3x = p3 * dr3[0], p3y = p3 * dr3[1], p3z = p3 * dr3[2];
                double* d = &a.dGdr[3 * s.index];
                d[0] += p1x + p2x; d[1] += p1y + p2y; d[2] += p1z + p2z;
                d = &nj.dGdr[3 * s.indexPerElement[nej]];
                d[0] -= p1x + p3x; d[1] -= p1y + p3y; d[2] -= p1z + p3z;
                d = &nk.dGdr[3 * s.indexPerElement[nek]];
                d[0] -= p2x - p3x; d[1] -= p2y - p3y; d[2] -= p2z - p3z;
            }
        }
    }
    for (size_t l = 0; l < M; ++l) {
        const SF& s = e.sfs[g.mem[l]];
        result[l] *= g.factorNorm[l] / g.scalingFactors[l];
        a.G[s.index] = s.scale(result[l]);
    }
}

// pow_int (utility.cpp:292-316): exponentiation by squaring.
double pow_int(double x, int n)
{
    unsigned int m;
    if (n < 0) { x = 1.0 / x; m = -n; }
    else m = n;
    double r = 1.0;
    do {
        if (m & 1) r *= x;
        m >>= 1;
        x *= x;
    } while (m);
    return r;
}

// SymGrpExpAngn::calculate (SymGrpExpAngn.cpp:156-403).
void calc_angular(const Element& e, const Group& g, Atom& a)
{
    const size_t M = g.mem.size();
    vector<double> result(M, 0.0);
    double const rc = g.rc, rc2 = rc * rc;
    size_t nn = a.nbr.size();
    if (nn == 0) nn = 1;
    for (size_t j = 0; j + 1 < nn; ++j) {
        Neighbor& nj = a.nbr[j];
        int const nej = nj.element;
        double const rij = nj.d;
        if (!((g.e1 == nej || g.e2 == nej) && rij < rc)) continue;
        double const r2ij = rij * rij;
        double pfcij, pdfcij;
        g.fc.fdf(rij, pfcij, pdfcij);
        const double* dr1 = nj.dr;
        for (size_t k = j + 1; k < nn; ++k) {
            Neighbor& nk = a.nbr[k];
            int const nek = nk.element;
            if (!((g.e1 == nej && g.e2 == nek) || (g.e2 == nej && g.e1 == nek))) continue;
            double const rik = nk.d;
            if (!(rik < rc)) continue;
            const double* dr2 = nk.dr;
            double dr3[3] = {dr2[0] - dr1[0], dr2[1] - dr1[1], dr2[2] - dr1[2]};
            double rjk = dr3[0] * dr3[0] + dr3[1] * dr3[1] + dr3[2] * dr3[2];
            if (!(rjk < rc2)) continue;
            double pfcik, pdfcik;
            g.fc.fdf(rik, pfcik, pdfcik);
            rjk = sqrt(rjk);
            double pfcjk, pdfcjk;
            g.fc.fdf(rjk, pfcjk, pdfcjk);
            double const rinvijik = 1.0 / rij / rik;
            double const costijk = (dr1[0] * dr2[0] + dr1[1] * dr2[1] + dr1[2] * dr2[2]) * rinvijik;
            double const pfc = pfcij * pfcik * pfcjk;
            double const r2ik = rik * rik;
            double const r2sum = r2ij + r2ik + rjk * rjk;
            double const pr1 = pfcik * pfcjk * pdfcij / rij;
            double const pr2 = pfcij * pfcjk * pdfcik / rik;
            double const pr3 = pfcij * pfcik * pdfcjk / rjk;
            double vexp = 0.0, rijs = 0.0, riks = 0.0, rjks = 0.0;
            for (size_t l = 0; l < M; ++l) {
                const SF& s = e.sfs[g.mem[l]];
                if (g.calcExp[l]) {
                    if (s.rs > 0.0) {
                        rijs = rij - s.rs;
                        riks = rik - s.rs;
                        rjks = rjk - s.rs;
                        vexp = exp(-s.eta * (rijs * rijs + riks * riks + rjks * rjks));
                    } else {
                        vexp = exp(-s.eta * r2sum);
                    }
                }
                double const plambda = 1.0 + s.lambda * costijk;
                double fg = vexp;
                if (plambda <= 0.0) fg = 0.0;
                else if (s.useIntPow) fg *= pow_int(plambda, s.zetaInt - 1);
                else fg *= pow(plambda, s.zeta - 1.0);
                result[l] += fg * plambda * pfc;
                fg *= g.factorNorm[l];
                double const pfczl = pfc * g.zetaLambda[l];
                double const p2etapl = plambda * g.factorDeriv[l];
                double p1, p2, p3;
                if (s.rs > 0.0) {
                    p1 = fg * (pfczl * (rinvijik - costijk / r2ij - p2etapl * rijs / rij) + pr1 * plambda);
                    p2 = fg * (pfczl * (rinvijik - costijk / r2ik - p2etapl * riks / rik) + pr2 * plambda);
                    p3 = fg * (pfczl * (rinvijik + p2etapl * rjks / rjk) - pr3 * plambda);
                } else {
                    p1 = fg * (pfczl * (rinvijik - costijk / r2ij - p2etapl) + pr1 * plambda);
                    p2 = fg * (pfczl * (rinvijik - costijk / r2ik - p2etapl) + pr2 * plambda);
                    p3 = fg * (pfczl * (rinvijik + p2etapl) - pr3 * plambda);
                }
                double const p1x = p1 * dr1[0], p1y = p1 * dr1[1], p1z = p1 * dr1[2];
                double const p2x = p2 * dr2[0], p2y = p2 * dr2[1], p2z = p2 * dr2[2];
                double const p3x = p3 * dr3[0], p3y = p3 * dr3[1], p3z = p3 * dr3[2];
                double* d = &a.dGdr[3 * s.index];
                d[0] += p1x + p2x; d[1] += p1y + p2y; d[2] += p1z + p2z;
                d = &nj.dGdr[3 * s.indexPerElement[nej]];
                d[0] -= p1x + p3x; d[1] -= p1y + p3y; d[2] -= p1z + p3z;
                d = &nk.dGdr[3 * s.indexPerElement[nek]];
                d[0] -= p2x - p3x; d[1] -= p2y - p3y; d[2] -= p2z - p3z;
            }
        }
    }
    for (size_t l = 0; l < M; ++l) {
        const SF& s = e.sfs[g.mem[l]];
        result[l] *= g.factorNorm[l] / g.scalingFactors[l];
        a.G[s.index] = s.scale(result[l]);
    }
}

// SymGrpExpAngw::calculate (SymGrpExpAngw.cpp:156-380): every pair j < k of
// the group's elements inside rc, no r_jk cutoff and no r_jk term.
void calc_angular_wide(const Element& e, const Group& g, Atom& a)
{
    const size_t M = g.mem.size();
    vector<double> result(M, 0.0);
    double const rc = g.rc;
    size_t nn = a.nbr.size();
    if (nn == 0) nn = 1;
    for (size_t j = 0; j + 1 < nn; ++j) {
        Neighbor& nj = a.nbr[j];
        int const nej = nj.element;
        double const rij = nj.d;
        if (!((g.e1 == nej || g.e2 == nej) && rij < rc)) continue;
        double const r2ij = rij * rij;
        double pfcij, pdfcij;
        g.fc.fdf(rij, pfcij, pdfcij);
        const double* dr1 = nj.dr;
        for (size_t k = j + 1; k < nn; ++k) {
            Neighbor& nk = a.nbr[k];
            int const nek = nk.element;
            if (!((g.e1 == nej && g.e2 == nek) || (g.e2 == nej && g.e1 == nek))) continue;
            double const rik = nk.d;
            if (!(rik < rc)) continue;
            const double* dr2 = nk.dr;
            double const dr3[3] = {dr2[0] - dr1[0], dr2[1] - dr1[1], dr2[2] - dr1[2]};
            double pfcik, pdfcik;
            g.fc.fdf(rik, pfcik, pdfcik);
            double const rinvijik = 1.0 / rij / rik;
            double const costijk = (dr1[0] * dr2[0] + dr1[1] * dr2[1] + dr1[2] * dr2[2]) * rinvijik;
            double const pfc = pfcij * pfcik;
            double const r2ik = rik * rik;
            double const r2sum = r2ij + r2ik;
            double const pr1 = pfcik * pdfcij / rij;
            double const pr2 = pfcij * pdfcik / rik;
            double vexp = 0.0, rijs = 0.0, riks = 0.0;
            for (size_t l = 0; l < M; ++l) {
                const SF& s = e.sfs[g.mem[l]];
                if (g.calcExp[l]) {
                    if (s.rs > 0.0) {
                        rijs = rij - s.rs;
                        riks = rik - s.rs;
                        vexp = exp(-s.eta * (rijs * rijs + riks * riks));
                    } else {
                        vexp = exp(-s.eta * r2sum);
                    }
                }
                double const plambda = 1.0 + s.lambda * costijk;
                double fg = vexp;
                if (plambda <= 0.0) fg = 0.0;
                else if (s.useIntPow) fg *= pow_int(plambda, s.zetaInt - 1);
                else fg *= pow(plambda, s.zeta - 1.0);
                result[l] += fg * plambda * pfc;
                fg *= g.factorNorm[l];
                double const pfczl = pfc * g.zetaLambda[l];
                double const p2etapl = plambda * g.factorDeriv[l];
                double p1, p2;
                if (s.rs > 0.0) {
                    p1 = fg * (pfczl * (rinvijik - costijk / r2ij - p2etapl * rijs / rij) + pr1 * plambda);
                    p2 = fg * (pfczl * (rinvijik - costijk / r2ik - p2etapl * riks / rik) + pr2 * plambda);
                } else {
                    p1 = fg * (pfczl * (rinvijik - costijk / r2ij - p2etapl) + pr1 * plambda);
                    p2 = fg * (pfczl * (rinvijik - costijk / r2ik - p2etapl) + pr2 * plambda);
                }
                double const p3 = fg * pfczl * rinvijik;
                double const p1x = p1 * dr1[0], p1y = p1 * dr1[1], p1z = p1 * dr1[2];
                double const p2x = p2 * dr2[0], p2y = p2 * dr2[1], p2z = p2 * dr2[2];
                double const p3x = p3 * dr3[0], p3y = p3 * dr3[1], p3z = p3 * dr3[2];
                double* d = &a.dGdr[3 * s.index];
                d[0] += p1x + p2x; d[1] += p1y + p2y; d[2] += p1z + p2z;
                d = &nj.dGdr[3 * s.indexPerElement[nej]];
                d[0] -= p1x + p3x; d[1] -= p1y + p3y; d[2] -= p1z + p3z;
                d = &nk.dGdr[3 * s.indexPerElement[nek]];
                d[0] -= p2x - p3x; d[1] -= p2y - p3y; d[2] -= p2z - p3z;
            }
        }
    }
    for (size_t l = 0; l < M; ++l) {
        const SF& s = e.sfs[g.mem[l]];
        result[l] *= g.factorNorm[l] / g.scalingFactors[l];
        a.G[s.index] = s.scale(result[l]);
    }
}

// SymGrpCompRad::calculate (SymGrpCompRad.cpp:129-199).
// With `weighted` (type 23), SymGrpCompRadWeighted::calculate
// (SymGrpCompRadWeighted.cpp:120-200): every neighbour element, each term
// weighted by the neighbour's atomic number.
void calc_comp_radial(const Element& e, const Group& g, Atom& a, bool weighted)
{
    const size_t M = g.mem.size();
    vector<double> result(M, 0.0);
    for (auto& n : a.nbr) {
        double const rij = n.d;
        if (!((weighted || g.e1 == n.element) && rij < g.rmax && rij > g.rmin)) continue;
        double const Z = g.Z[n.element];
        for (size_t k = 0; k < M; ++k) {
            if (rij <= g.mrl[k] || rij >= g.mrc[k]) continue;
            const SF& s = e.sfs[g.mem[k]];
            double rad, drad;
            s.cr.fdf(rij, rad, drad);
            if (weighted) result[k] += Z * rad;
            else result[k] += rad;
            if (drad == 0.0) continue;
            if (weighted) drad *= Z * g.scalingFactors[k] / rij;
            else drad *= g.scalingFactors[k] / rij;
            double const px = drad * n.dr[0], py = drad * n.dr[1], pz = drad * n.dr[2];
            double* d = &a.dGdr[3 * s.index];
            d[0] += px; d[1] += py; d[2] += pz;
            d = &n.dGdr[3 * s.indexPerElement[n.element]];
            d[0] -= px; d[1] -= py; d[2] -= pz;
        }
    }
    for (size_t k = 0; k < M; ++k) {
        const SF& s = e.sfs[g.mem[k]];
        a.G[s.index] = s.scale(result[k]);
    }
}

// SymGrpCompAngw::calculate (SymGrpCompAngw.cpp:117-341) and, with
// `narrow`, SymGrpCompAngn::calculate (SymGrpCompAngn.cpp:128-390): the
// product of compact radial functions of r_ij, r_ik (narrow: and r_jk) with
// a compact function of the angle acos(cos theta_ijk).  In the narrow form
// the r_jk factor is only re-evaluated when a member's (rl, rc) differ from
// the previous member's or the previous member was skipped for this
// triplet; otherwise the previous member's value is reused.
// With `weighted` (types 24, 25), SymGrpCompAngnWeighted::calculate and
// SymGrpCompAngwWeighted::calculate (SymGrpCompAngnWeighted.cpp:131-390,
// SymGrpCompAngwWeighted.cpp:116-345): all neighbour pairs, the angular
// factor weighted by Z_j Z_k.
void calc_comp_angular(const Element& e, const Group& g, Atom& a, bool narrow, bool weighted = false)
{
    const size_t M = g.mem.size();
    vector<double> result(M, 0.0), radij(M, 0.0), dradij(M, 0.0);
    double const rmin = g.rmin, rmax = g.rmax;
    double const r2min = rmin * rmin, r2max = rmax * rmax;
    size_t nn = a.nbr.size();
    if (nn == 0) nn = 1;
    for (size_t j = 0; j + 1 < nn; ++j) {
        Neighbor& nj = a.nbr[j];
        int const nej = nj.element;
        double const rij = nj.d;
        if (!((weighted || g.e1 == nej || g.e2 == nej) && rij < rmax && rij > rmin)) continue;
        for (size_t l = 0; l < M; ++l) {
            if (rij > g.mrl[l] && rij < g.mrc[l]) e.sfs[g.mem[l]].cr.fdf(rij, radij[l], dradij[l]);
            else radij[l] = 0.0;
        }
        const double* dr1 = nj.dr;
        for (size_t k = j + 1; k < nn; ++k) {
            Neighbor& nk = a.nbr[k];
            int const nek = nk.element;
            if (!(weighted || (g.e1 == nej && g.e2 == nek) || (g.e2 == nej && g.e1 == nek))) continue;
            double const rik = nk.d;
            if (!(rik < rmax && rik > rmin)) continue;
            const double* dr2 = nk.dr;
            double const dr3[3] = {dr2[0] - dr1[0], dr2[1] - dr1[1], dr2[2] - dr1[2]};
            double rjk = 0.0;
            if (narrow) {
                rjk = dr3[0] * dr3[0] + dr3[1] * dr3[1] + dr3[2] * dr3[2];
                if (rjk >= r2max || rjk <= r2min) continue;
                rjk = sqrt(rjk);
            }
            double const rinvijik = 1.0 / (rij * rik);
            double const costijk = (dr1[0] * dr2[0] + dr1[1] * dr2[1] + dr1[2] * dr2[2]) * rinvijik;
            // the angular polynomial vanishes at 0 and 180 degrees
            if (costijk <= -1.0 || costijk >= 1.0) continue;
            double const acostijk = acos(costijk);
            double const rinvij = rinvijik * rik;
            double const rinvik = rinvijik * rij;
            double const rinvjk = narrow ? 1.0 / rjk : 0.0;
            double const phiijik0 = rinvij * (rinvik - rinvij * costijk);
            double const phiikij0 = rinvik * (rinvij - rinvik * costijk);
            double const dacostijk = -1.0 / sqrt(1.0 - costijk * costijk);
            bool skip = false;
            double radjk = 0.0, dradjk = 0.0;
            for (size_t l = 0; l < M; ++l) {
                if (radij[l] == 0.0 || rik <= g.mrl[l] || rik >= g.mrc[l] ||
                    (narrow && (rjk <= g.mrl[l] || rjk >= g.mrc[l])) ||
                    acostijk <= g.mal[l] || acostijk >= g.mar[l]) {
                    skip = true;
                    continue;
                }
                const SF& s = e.sfs[g.mem[l]];
                double radik, dradik, ang, dang;
                s.cr.fdf(rik, radik, dradik);
                if (narrow && (g.calcComp[l] || skip)) {
                    s.cr.fdf(rjk, radjk, dradjk);
                    skip = false;
                }
                s.ca.fdf(acostijk, ang, dang);
                double const z = g.Z[nej] * g.Z[nek];
                if (weighted) ang *= z;
                double rad = narrow ? radij[l] * radik * radjk : radij[l] * radik;
                result[l] += rad * ang;
                dang *= dacostijk;
                double const phiijik = phiijik0 * dang;
                double const phiikij = phiikij0 * dang;
                double const psiijik = rinvijik * dang;
                double chiij, chiik, chijk = 0.0;
                if (narrow) {
                    chiij = rinvij * dradij[l] * radik * radjk;
                    chiik = rinvik * radij[l] * dradik * radjk;
                    chijk = -rinvjk * radij[l] * radik * dradjk;
                } else {
                    // the weighted group multiplies in the other order
                    // (SymGrpCompAngwWeighted.cpp:257)
                    chiij = weighted ? rinvij * dradij[l] * radik : rinvij * radik * dradij[l];
                    chiik = rinvik * radij[l] * dradik;
                }
                double p1, p2, p3;
                double const sfac = g.scalingFactors[l];
                if (dang != 0.0 && rad != 0.0) {
                    if (weighted) rad *= sfac * z;
                    else rad *= sfac;
                    ang *= sfac;
                    p1 = rad * phiijik + ang * chiij;
                    p2 = rad * phiikij + ang * chiik;
                    p3 = narrow ? rad * psiijik + ang * chijk : rad * psiijik;
                } else if (ang != 0.0) {
                    ang *= sfac;
                    p1 = ang * chiij;
                    p2 = ang * chiik;
                    p3 = narrow ? ang * chijk : 0.0;
                } else {
                    continue;
                }
                double const p1x = p1 * dr1[0], p1y = p1 * dr1[1], p1z = p1 * dr1[2];
                double const p2x = p2 * dr2[0], p2y = p2 * dr2[1], p2z = p2 * dr2[2];
                double const p3x = p3 * dr3[0], p3y = p3 * dr3[1], p3z = p3 * dr3[2];
                double* d = &a.dGdr[3 * s.index];
                d[0] += p1x + p2x; d[1] += p1y + p2y; d[2] += p1z + p2z;
                d = &nj.dGdr[3 * s.indexPerElement[nej]];
                d[0] -= p1x + p3x; d[1] -= p1y + p3y; d[2] -= p1z + p3z;
                d = &nk.dGdr[3 * s.indexPerElement[nek]];
                d[0] -= p2x - p3x; d[1] -= p2y - p3y; d[2] -= p2z - p3z;
            }
        }
    }
    for (size_t l = 0; l < M; ++l) {
        const SF& s = e.sfs[g.mem[l]];
        a.G[s.index] = s.scale(result[l]);
    }
}

// Mode::calculateSymmetryFunctionGroups (Mode.cpp:1630-1709) with
// Atom::allocate (Atom.cpp:168-240).
void calc_sf(oracle_ctx& c)
{
    const int N = (int)c.atoms.size();
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < N; ++i) {
        Atom& a = c.atoms[i];
        const Element& e = c.el[a.element];
        const size_t ns = e.sfs.size();
        a.G.assign(ns, 0.0);
        a.dEdG.assign(ns, 0.0);
        a.dGdr.assign(3 * ns, 0.0);
        for (auto& n : a.nbr) n.dGdr.assign(3 * e.table[n.element].size(), 0.0);
        for (auto& g : e.grps) {
            if (g.type == 2) calc_radial(e, g, a);
            else if (g.type == 9) calc_angular_wide(e, g, a);
            else if (g.type == 12) calc_radial_weighted(e, g, a);
            else if (g.type == 13) calc_angular_weighted(e, g, a);
            else if (g.type == 20 || g.type == 23) calc_comp_radial(e, g, a, g.type == 23);
            else if (g.type == 21 || g.type == 22) calc_comp_angular(e, g, a, g.type == 21);
            else if (g.type == 24 || g.type == 25) calc_comp_angular(e, g, a, g.type == 24, true);
            else calc_angular(e, g, a);
        }
    }
}

// Mode::calculateAtomicNeuralNetworks (Mode.cpp:1711-1732) with
// NeuralNetwork::propagateLayer (NeuralNetwork.cpp:761-896) and the
// forward-mode NeuralNetwork::calculateDEdG (NeuralNetwork.cpp:396-442).
void calc_nn(oracle_ctx& c)
{
    const int N = (int)c.atoms.size();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        Atom& a = c.atoms[i];
        const NN& nn = c.el[a.element].nn;
        const int L = (int)nn.n.size();
        vector<vector<double>> val(L), dfdx(L);
        val[0] = a.G;
        for (int l = 1; l < L; ++l) {
            val[l].assign(nn.n[l], 0.0);
            dfdx[l].assign(nn.n[l], 0.0);
            for (int k = 0; k < nn.n[l]; ++k) {
                double t = 0.0;
                const double* w = &nn.W[l][(size_t)k * nn.n[l - 1]];
                for (int j = 0; j < nn.n[l - 1]; ++j) t += w[j] * val[l - 1][j];
                t += nn.b[l][k];
                if (nn.normalize) t /= nn.n[l - 1];
                val[l][k] = activate(nn.act[l], t, dfdx[l][k]);
            }
        }
        a.energy = val[L - 1][0];
        const int nh = L - 2;
        vector<vector<double>> inner(nh), outer(nh);
        for (int l = 0; l < nh; ++l) {
            inner[l].assign(nn.n[l + 1], 0.0);
            outer[l].assign(nn.n[l + 2], 0.0);
        }
        for (int k = 0; k < nn.n[0]; ++k) {
            for (int q = 0; q < nn.n[1]; ++q) {
                inner[0][q] = nn.W[1][(size_t)q * nn.n[0] + k] * dfdx[1][q];
                if (nn.normalize) inner[0][q] /= nn.n[0];
            }
            for (int l = 1; l < nh + 1; ++l) {
                for (int i2 = 0; i2 < nn.n[l + 1]; ++i2) {
                    outer[l - 1][i2] = 0.0;
                    for (int i1 = 0; i1 < nn.n[l]; ++i1)
                        outer[l - 1][i2] += nn.W[l + 1][(size_t)i2 * nn.n[l] + i1] * inner[l - 1][i1];
                    outer[l - 1][i2] *= dfdx[l + 1][i2];
                    if (nn.normalize) outer[l - 1][i2] /= nn.n[l];
                    if (l < nh) inner[l][i2] = outer[l - 1][i2];
                }
            }
            a.dEdG[k] = outer[nh - 1][0];
        }
    }
}

// Mode::calculateForces (Mode.cpp:1922-1984), Atom::calculateSelfForceShort
// and calculatePairForceShort (Atom.cpp:371-401).
void calc_forces_pull(oracle_ctx& c)
{
    const int N = (int)c.atoms.size();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        Atom& ai = c.atoms[i];
        double f[3] = {0.0, 0.0, 0.0};
        double s[3] = {0.0, 0.0, 0.0};
        for (size_t q = 0; q < ai.dEdG.size(); ++q)
            for (int k = 0; k < 3; ++k) s[k] -= ai.dEdG[q] * ai.dGdr[3 * q + k];
        for (int k = 0; k < 3; ++k) f[k] += s[k];
        for (size_t u = 1; u < ai.unique.size(); ++u) {
            const Atom& aj = c.atoms[ai.unique[u]];
            const vector<vector<int>>& table = c.el[aj.element].table;
            for (auto& n : aj.nbr) {
                if (n.j != i) continue;
                const vector<int>& t = table[n.element];
                double p[3] = {0.0, 0.0, 0.0};
                for (size_t q = 0; q < t.size(); ++q)
                    for (int k = 0; k < 3; ++k) p[k] -= aj.dEdG[t[q]] * n.dGdr[3 * q + k];
                for (int k = 0; k < 3; ++k) f[k] += p[k];
            }
        }
        for (int k = 0; k < 3; ++k) ai.f[k] = f[k];
    }
}

// InterfaceLammps::getForces (InterfaceLammps.cpp:679-740) without unit
// factors: every neighbour entry, self images included, scatters.
void calc_forces_push(oracle_ctx& c)
{
    for (auto& a : c.atoms) a.f[0] = a.f[1] = a.f[2] = 0.0;
    for (size_t i = 0; i < c.atoms.size(); ++i) {
        Atom& a = c.atoms[i];
        const vector<vector<int>>& table = c.el[a.element].table;
        for (auto& n : a.nbr) {
            const vector<int>& t = table[n.element];
            Atom& b = c.atoms[n.j];
            for (size_t q = 0; q < t.size(); ++q)
                for (int k = 0; k < 3; ++k) b.f[k] -= a.dEdG[t[q]] * n.dGdr[3 * q + k];
        }
        for (size_t q = 0; q < a.dEdG.size(); ++q)
            for (int k = 0; k < 3; ++k) a.f[k] -= a.dEdG[q] * a.dGdr[3 * q + k];
    }
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void set_box(oracle_ctx& c, const double* box)
{
    c.periodic = box != nullptr;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) c.box[i][k] = box ? box[3 * i + k] : 0.0;
    if (c.periodic) inverse_box(c);
}

}  // namespace

// ==================================================================== C API
extern "C" {

// ----------------------------------------------------------------------
// Training gradients (config "force-training gradient"): dE/dc and dF/dc of
// the connections c, restated from NeuralNetwork::calculateDEdc
// (NeuralNetwork.cpp:444-488), calculateDFdc (490-527), calculateDEdb
// (593-626), calculateDxdG (628-657), calculateD2EdGdc (659-719) and the
// activation derivatives of propagateLayer (761-896); the per-component
// Jacobian dG_i/dx_{a,c} is gathered as Training::collectDGdxia does
// (Training.cpp:3586-3625) and contracted with the force residual field.
namespace {

struct NNState {
    vector<vector<double>> val, dfdx, d2fdx2, dxdG;
};

// Value and first two derivatives of an activation (propagateLayer).
double activate2(int a, double x, double& d1, double& d2)
{
    const double EXP_LIMIT = 35.0;
    double v;
    switch (a) {
    case AF_IDENTITY: d1 = 1.0; d2 = 0.0; return x;
    case AF_TANH: v = tanh(x); d1 = 1.0 - v * v; d2 = -2.0 * v * (1.0 - v * v); return v;
    case AF_LOGISTIC:
        if (x > EXP_LIMIT) { d1 = 0.0; d2 = 0.0; return 1.0; }
        if (x < -EXP_LIMIT) { d1 = 0.0; d2 = 0.0; return 0.0; }
        v = 1.0 / (1.0 + exp(-x)); d1 = v * (1.0 - v); d2 = v * (1.0 - v) * (1.0 - 2.0 * v); return v;
    case AF_SOFTPLUS:
        if (x > EXP_LIMIT) { d1 = 1.0; d2 = 0.0; return x; }
        if (x < -EXP_LIMIT) { d1 = 0.0; d2 = 0.0; return 0.0; }
        {
            const double t = exp(x);
            v = log(1.0 + t);
            const double s = 1.0 / (1.0 + 1.0 / t);
            d1 = s; d2 = s * (1.0 - s);
            return v;
        }
    case AF_RELU:
        d2 = 0.0;
        if (x > 0.0) { d1 = 1.0; return x; }
        d1 = 0.0; return 0.0;
    case AF_GAUSSIAN: { const double t = exp(-0.5 * x * x); d1 = -x * t; d2 = (x * x - 1.0) * t; return t; }
    case AF_COS: { const double t = cos(x); d1 = -sin(x); d2 = -t; return t; }
    case AF_REVLOGISTIC:
        v = 1.0 / (1.0 + exp(-x)); d1 = v * (v - 1.0); d2 = v * (v - 1.0) * (1.0 - 2.0 * v); return 1.0 - v;
    case AF_EXP: v = exp(-x); d1 = -v; d2 = v; return v;
    case AF_HARMONIC: d1 = 2.0 * x; d2 = 2.0; return x * x;
    }
    throw std::runtime_error("activation without second derivative");
}

struct ConnOffsets {
    vector<size_t> w, b;   // connection offsets of weights / biases of layer l
    vector<size_t> bo;     // offsets of layer l in the bias-only array
    size_t n = 0, nb = 0;
};

ConnOffsets conn_offsets(const NN& nn)
{
    ConnOffsets o;
    const size_t L = nn.n.size();
    o.w.assign(L, 0); o.b.assign(L, 0); o.bo.assign(L, 0);
    for (size_t l = 1; l < L; ++l) {
        o.w[l] = o.n;
        o.n += (size_t)nn.n[l - 1] * nn.n[l];
        o.b[l] = o.n;
        o.n += nn.n[l];
        o.bo[l] = o.nb;
        o.nb += nn.n[l];
    }
    return o;
}

void nn_forward2(const NN& nn, const vector<double>& G, NNState& st)
{
    const int L = (int)nn.n.size();
    st.val.assign(L, {}); st.dfdx.assign(L, {}); st.d2fdx2.assign(L, {}); st.dxdG.assign(L, {});
    st.val[0] = G;
    for (int l = 1; l < L; ++l) {
        st.val[l].assign(nn.n[l], 0.0);
        st.dfdx[l].assign(nn.n[l], 0.0);
        st.d2fdx2[l].assign(nn.n[l], 0.0);
        st.dxdG[l].assign(nn.n[l], 0.0);
        for (int k = 0; k < nn.n[l]; ++k) {
            double t = 0.0;
            const double* w = &nn.W[l][(size_t)k * nn.n[l - 1]];
            for (int j = 0; j < nn.n[l - 1]; ++j) t += w[j] * st.val[l - 1][j];
            t += nn.b[l][k];
            if (nn.normalize) t /= nn.n[l - 1];
            st.val[l][k] = activate2(nn.act[l], t, st.dfdx[l][k], st.d2fdx2[l][k]);
        }
    }
}

// NeuralNetwork::calculateDEdc: dE/dc in connection order.
void nn_dEdc(const NN& nn, const NNState& st, const ConnOffsets& o, double* d)
{
    const int L = (int)nn.n.size();
    for (size_t q = 0; q < o.n; ++q) d[q] = 0.0;
    for (int i = 0; i < nn.n[L - 1]; ++i) {
        d[o.b[L - 1] + i] = st.dfdx[L - 1][i];
        if (nn.normalize) d[o.b[L - 1] + i] /= nn.n[L - 2];
    }
    for (int i = L - 2; i >= 0; --i) {
        for (int j = 0; j < nn.n[i]; ++j) {
            for (int k = 0; k < nn.n[i + 1]; ++k) {
                d[o.w[i + 1] + (size_t)j * nn.n[i + 1] + k] = d[o.b[i + 1] + k] * st.val[i][j];
                if (i >= 1)
                    d[o.b[i] + j] += d[o.b[i + 1] + k] * nn.W[i + 1][(size_t)k * nn.n[i] + j] * st.dfdx[i][j];
            }
            if (nn.normalize && i >= 1) d[o.b[i] + j] /= nn.n[i - 1];
        }
    }
}

// NeuralNetwork::calculateDEdb: dE/db in bias-only order.
void nn_dEdb(const NN& nn, const NNState& st, const ConnOffsets& o, vector<double>& d)
{
    const int L = (int)nn.n.size();
    d.assign(o.nb, 0.0);
    for (int i = 0; i < nn.n[L - 1]; ++i) {
        d[o.bo[L - 1] + i] = st.dfdx[L - 1][i];
        if (nn.normalize) d[o.bo[L - 1] + i] /= nn.n[L - 2];
    }
    for (int i = L - 2; i >= 0; --i) {
        for (int j = 0; j < nn.n[i]; ++j) {
            for (int k = 0; k < nn.n[i + 1]; ++k)
                if (i >= 1)
                    d[o.bo[i] + j] += d[o.bo[i + 1] + k] * nn.W[i + 1][(size_t)k * nn.n[i] + j] * st.dfdx[i][j];
            if (nn.normalize && i >= 1) d[o.bo[i] + j] /= nn.n[i - 1];
        }
    }
}

// NeuralNetwork::calculateDxdG for input `index`.
void nn_dxdG(const NN& nn, NNState& st, int index)
{
    const int L = (int)nn.n.size();
    for (int i = 0; i < nn.n[1]; ++i) {
        st.dxdG[1][i] = nn.W[1][(size_t)i * nn.n[0] + index];
        if (nn.normalize) st.dxdG[1][i] /= nn.n[0];
    }
    for (int l = 2; l < L; ++l)
        for (int j = 0; j < nn.n[l]; ++j) {
            double t = 0.0;
            for (int k = 0; k < nn.n[l - 1]; ++k)
                t += nn.W[l][(size_t)j * nn.n[l - 1] + k] * st.dfdx[l - 1][k] * st.dxdG[l - 1][k];
            if (nn.normalize) t /= nn.n[l - 1];
            st.dxdG[l][j] = t;
        }
}

// NeuralNetwork::calculateD2EdGdc for input `index`.
void nn_d2EdGdc(const NN& nn, const NNState& st, const ConnOffsets& o, int index, const vector<double>& dEdb,
                double* d2)
{
    const int L = (int)nn.n.size();
    for (int i = 0; i < nn.n[L - 1]; ++i) {
        d2[o.b[L - 1] + i] = st.d2fdx2[L - 1][i] * st.dxdG[L - 1][i];
        if (nn.normalize) d2[o.b[L - 1] + i] /= nn.n[L - 2];
    }
    for (int i = L - 2; i >= 0; --i) {
        for (int j = 0; j < nn.n[i]; ++j) {
            for (int k = 0; k < nn.n[i + 1]; ++k) {
                double& w = d2[o.w[i + 1] + (size_t)j * nn.n[i + 1] + k];
                if (i == 0) {
                    w = d2[o.b[i + 1] + k] * st.val[i][j];
                    if (j == index) w += dEdb[o.bo[i + 1] + k];
                } else {
                    w = d2[o.b[i + 1] + k] * st.val[i][j] + dEdb[o.bo[i + 1] + k] * st.dfdx[i][j] * st.dxdG[i][j];
                }
                if (i >= 1)
                    d2[o.b[i] + j] += nn.W[i + 1][(size_t)k * nn.n[i] + j] *
                                      (d2[o.b[i + 1] + k] * st.dfdx[i][j] +
                                       dEdb[o.bo[i + 1] + k] * st.d2fdx2[i][j] * st.dxdG[i][j]);
            }
            if (nn.normalize && i >= 1) d2[o.b[i] + j] /= nn.n[i - 1];
        }
    }
}

// NeuralNetwork::calculateDFdc: dF/dc = -sum_i d2E/dG_i dc * dGdxyz_i.
void nn_dFdc(const NN& nn, NNState& st, const ConnOffsets& o, const double* dGdxyz, double* dFdc)
{
    vector<double> dEdb;
    nn_dEdb(nn, st, o, dEdb);
    vector<double> d2(o.n);
    for (size_t q = 0; q < o.n; ++q) dFdc[q] = 0.0;
    for (int i = 0; i < nn.n[0]; ++i) {
        std::fill(d2.begin(), d2.end(), 0.0);
        nn_dxdG(nn, st, i);
        nn_d2EdGdc(nn, st, o, i, dEdb, d2.data());
        for (size_t q = 0; q < o.n; ++q) dFdc[q] -= d2[q] * dGdxyz[i];
    }
}

}  // namespace

const char* oracle_last_error(void) { return g_err.c_str(); }

#define ORACLE_TRY(...)                                   \
    try { __VA_ARGS__ } catch (const std::exception& ex) { \
        g_err = ex.what(); return 1; }

oracle_ctx* oracle_create(const char* model_dir)
{
    oracle_ctx* c = new oracle_ctx();
    try {
        load_model(*c, model_dir);
    } catch (const std::exception& ex) {
        g_err = ex.what();
        delete c;
        return nullptr;
    }
    return c;
}

void oracle_destroy(oracle_ctx* c) { delete c; }

int oracle_num_elements(const oracle_ctx* c) { return (int)c->el.size(); }
int oracle_num_sf(const oracle_ctx* c, int e) { return (int)c->el.at(e).sfs.size(); }
int oracle_atomic_number(const oracle_ctx* c, int e) { return c->el.at(e).Z; }
double oracle_max_cutoff(const oracle_ctx* c) { return c->maxCutoff; }
int oracle_normalize(const oracle_ctx* c) { return c->normalize ? 1 : 0; }

// Structure::readFromLines (Structure.cpp:142-271) for the first structure,
// then Prediction::readStructureFromFile (Prediction.cpp:50-65).
int oracle_read_structure(oracle_ctx* c, const char* path)
{
    ORACLE_TRY({
        std::ifstream in(path);
        if (!in) throw std::runtime_error(string("could not open ") + path);
        string line;
        vector<double> pos, box;
        vector<int> elem;
        bool begun = false;
        while (std::getline(in, line)) {
            vector<string> w = words(line);
            if (w.empty()) continue;
            if (w[0] == "begin") { begun = true; continue; }
            if (!begun) throw std::runtime_error("expected \"begin\"");
            if (w[0] == "lattice") {
                for (int k = 1; k <= 3; ++k) box.push_back(atof(w.at(k).c_str()));
            } else if (w[0] == "atom") {
                for (int k = 1; k <= 3; ++k) pos.push_back(atof(w.at(k).c_str()));
                int e = -1;
                for (size_t q = 0; q < c->el.size(); ++q) if (c->el[q].symbol == w.at(4)) e = (int)q;
                if (e < 0) throw std::runtime_error("unknown element " + w.at(4));
                elem.push_back(e);
            } else if (w[0] == "end") {
                break;
            }
        }
        if (!(box.empty() || box.size() == 9)) throw std::runtime_error("strange number of box vectors");
        int rc = oracle_set_structure(c, (int)elem.size(), pos.data(), elem.data(), box.empty() ? nullptr : box.data());
        if (rc) return rc;
    })
    return 0;
}

int oracle_set_structure(oracle_ctx* c, int n, const double* pos, const int* elem, const double* box)
{
    ORACLE_TRY({
        c->atoms.assign(n, Atom());
        set_box(*c, box);
        for (int i = 0; i < n; ++i) {
            if (elem[i] < 0 || elem[i] >= (int)c->el.size()) throw std::runtime_error("element index out of range");
            c->atoms[i].element = elem[i];
            for (int k = 0; k < 3; ++k) c->atoms[i].r[k] = pos[3 * i + k];
            if (c->periodic) remap(*c, c->atoms[i]);
        }
        to_normalized(*c);
    })
    return 0;
}

long oracle_build_neighbors(oracle_ctx* c, int nbr_mode, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    try {
        if (nbr_mode == ORACLE_NBR_CELL) neighbors_cell(*c, c->maxCutoff);
        else neighbors_brute(*c, c->maxCutoff);
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
    long tot = 0;
    for (auto& a : c->atoms) tot += (long)a.nbr.size();
    return tot;
}

// Prediction::predict (Prediction.cpp:67-78) -> Mode::evaluateNNP
// (Mode.cpp:2029-2077), then toPhysicalUnits and energy offsets.
int oracle_predict(oracle_ctx* c, int force_mode, int nbr_mode, int nthreads)
{
    ORACLE_TRY({
#ifdef _OPENMP
        if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
        double t0 = now();
        if (nbr_mode == ORACLE_NBR_CELL) neighbors_cell(*c, c->maxCutoff);
        else neighbors_brute(*c, c->maxCutoff);
        double t1 = now();
        calc_sf(*c);
        double t2 = now();
        calc_nn(*c);
        double t3 = now();
        double E = 0.0;  // Mode::calculateEnergy (Mode.cpp:1876-1901)
        for (auto& a : c->atoms) E += a.energy;
        c->energy = E;
        double t4 = now();
        if (force_mode == ORACLE_FORCES_PUSH) calc_forces_push(*c);
        else calc_forces_pull(*c);
        double t5 = now();
        c->phase[0] = t1 - t0; c->phase[1] = t2 - t1; c->phase[2] = t3 - t2;
        c->phase[3] = t4 - t3; c->phase[4] = t5 - t4;
        // Structure::toPhysicalUnits (Structure.cpp:1279-1306),
        // Atom::toPhysicalUnits (Atom.cpp:123-150).
        const size_t N = c->atoms.size();
        if (c->normalize) {
            c->energy = c->energy / c->convEnergy + N * c->meanEnergy;
            double const ff = 1.0 / (c->convEnergy / c->convLength);
            double const fr = 1.0 / c->convLength;
            for (auto& a : c->atoms)
                for (int k = 0; k < 3; ++k) { a.f[k] *= ff; a.r[k] *= fr; }
            if (c->periodic) {
                double const inv = 1.0 / c->convLength;
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < 3; ++k) { c->box[i][k] *= inv; c->invbox[i][k] *= c->convLength; }
            }
        }
        // Mode::addEnergyOffset (Mode.cpp:2079-2096).
        for (size_t e = 0; e < c->el.size(); ++e) {
            size_t cnt = 0;
            for (auto& a : c->atoms) cnt += (a.element == (int)e);
            c->energy += cnt * c->el[e].energyOffset;
        }
    })
    return 0;
}

int oracle_num_atoms(const oracle_ctx* c) { return (int)c->atoms.size(); }
double oracle_energy(const oracle_ctx* c) { return c->energy; }

int oracle_get_forces(const oracle_ctx* c, double* f)
{
    for (size_t i = 0; i < c->atoms.size(); ++i)
        for (int k = 0; k < 3; ++k) f[3 * i + k] = c->atoms[i].f[k];
    return 0;
}

int oracle_get_atomic_energies(const oracle_ctx* c, double* e)
{
    for (size_t i = 0; i < c->atoms.size(); ++i) e[i] = c->atoms[i].energy;
    return 0;
}

int oracle_get_positions(const oracle_ctx* c, double* r)
{
    for (size_t i = 0; i < c->atoms.size(); ++i)
        for (int k = 0; k < 3; ++k) r[3 * i + k] = c->atoms[i].r[k];
    return 0;
}

int oracle_get_G(const oracle_ctx* c, double* G, int ld)
{
    for (size_t i = 0; i < c->atoms.size(); ++i)
        for (size_t s = 0; s < c->atoms[i].G.size() && (int)s < ld; ++s) G[i * ld + s] = c->atoms[i].G[s];
    return 0;
}

int oracle_get_dEdG(const oracle_ctx* c, double* d, int ld)
{
    for (size_t i = 0; i < c->atoms.size(); ++i)
        for (size_t s = 0; s < c->atoms[i].dEdG.size() && (int)s < ld; ++s) d[i * ld + s] = c->atoms[i].dEdG[s];
    return 0;
}

// Atom::dGdr and Atom::Neighbor::dGdr (Atom.h:35-131) of the last predict:
// self[N][ld_self][3]; nbr[entry][ld_nbr][3] for every neighbour entry in
// list order, slot q of an entry with neighbour element e being SF
// table[e][q] of the centre element (Element::setupSymmetryFunctionMemory,
// Element.cpp:289-311).  Unused slots are zero; either pointer may be NULL.
int oracle_get_dGdr(const oracle_ctx* c, double* self, int ld_self, double* nbr, int ld_nbr)
{
    size_t p = 0;
    for (size_t i = 0; i < c->atoms.size(); ++i) {
        const Atom& a = c->atoms[i];
        if (a.dGdr.empty()) {
            g_err = "dGdr: run oracle_predict first";
            return 1;
        }
        if (self)
            for (int q = 0; q < ld_self; ++q)
                for (int k = 0; k < 3; ++k)
                    self[(i * ld_self + q) * 3 + k] = 3 * (size_t)q < a.dGdr.size() ? a.dGdr[3 * q + k] : 0.0;
        for (const Neighbor& n : a.nbr) {
            if (nbr)
                for (int q = 0; q < ld_nbr; ++q)
                    for (int k = 0; k < 3; ++k)
                        nbr[(p * ld_nbr + q) * 3 + k] = 3 * (size_t)q < n.dGdr.size() ? n.dGdr[3 * q + k] : 0.0;
            ++p;
        }
    }
    return 0;
}

// Element::updateSymmetryFunctionStatistics (Element.cpp:462-541) over the
// atoms in order: an SF whose unscaled value lies below Gmin - 1000 eps or
// above Gmax + 1000 eps.  Records (atom, SF, type, value, Gmin, Gmax) in
// (atom, SF) order; returns their number, writes at most max.
long oracle_extrapolation_warnings(const oracle_ctx* c, long max, int* atom, int* sf, int* type, double* value,
                                   double* gmin, double* gmax)
{
    const double eps = 1000.0 * std::numeric_limits<double>::epsilon();
    long n = 0;
    for (size_t i = 0; i < c->atoms.size(); ++i) {
        const Atom& a = c->atoms[i];
        const Element& e = c->el[a.element];
        for (size_t q = 0; q < a.G.size(); ++q) {
            const SF& f = e.sfs[q];
            const double v = f.unscale(a.G[q]);
            if (!(v + eps < f.Gmin || v - eps > f.Gmax)) continue;
            if (n < max) {
                atom[n] = (int)i;
                sf[n] = (int)q;
                type[n] = f.type;
                value[n] = v;
                gmin[n] = f.Gmin;
                gmax[n] = f.Gmax;
            }
            ++n;
        }
    }
    return n;
}

// The low-neighbour check of Mode::calculateSymmetryFunctionGroups
// (Mode.cpp:1667-1679): Atom::calculateNumNeighbors(Element::
// getMinCutoffRadius()) (entries with d <= rmin, Atom.cpp:307-327) below
// Element::getMinNeighbors() (Element.cpp:384-411: 1 for radial, 2 for
// angular SFs).  Records (atom, count); returns their number.
long oracle_low_neighbor_atoms(const oracle_ctx* c, long max, int* atom, int* num)
{
    long n = 0;
    for (size_t i = 0; i < c->atoms.size(); ++i) {
        const Atom& a = c->atoms[i];
        const Element& e = c->el[a.element];
        double rmin = std::numeric_limits<double>::max();
        size_t need = 0;
        for (const SF& f : e.sfs) {
            rmin = std::min(rmin, f.rc);
            need = std::max<size_t>(need, is_radial(f.type) ? 1 : 2);
        }
        size_t cnt = 0;
        for (const Neighbor& nb : a.nbr)
            if (nb.d <= rmin) ++cnt;
        if (cnt < need) {
            if (n < max) {
                atom[n] = (int)i;
                num[n] = (int)cnt;
            }
            ++n;
        }
    }
    return n;
}

// KalmanFilter::update (KalmanFilter.cpp:119-185) on caller arrays, FP64
// loops: P [n][n] row-major (only its lower triangle is read, as
// selfAdjointView<Lower> does; the whole matrix is written), w [n],
// H [n][m] (column q = Jacobian of observation q), xi [m]; par (in/out) =
// {type (0 standard, 1 fading memory), eta, etatau, etamax, lambda, nu,
// gamma, q, qtau, qmin}.  The m x m system is solved by Gauss-Jordan
// elimination with partial pivoting (A^-1, then K = X A^-1).
int oracle_kalman_update(int n, int m, double* P, double* w, const double* H, const double* xi, double* par)
{
    const int type = (int)par[0];
    double& eta = par[1];
    const double etatau = par[2], etamax = par[3];
    double& lambda = par[4];
    const double nu = par[5];
    double& gamma = par[6];
    double& q = par[7];
    const double qtau = par[8], qmin = par[9];
    auto Pl = [&](int i, int j) { return i >= j ? P[(size_t)i * n + j] : P[(size_t)j * n + i]; };
    // X = P H
    vector<double> X((size_t)n * m, 0.0);
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < m; ++c) {
            double v = 0.0;
            for (int j = 0; j < n; ++j) v += Pl(i, j) * H[(size_t)j * m + c];
            X[(size_t)i * m + c] = v;
        }
    // A = H^T X (+ R)
    vector<double> A((size_t)m * m, 0.0);
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b) {
            double v = 0.0;
            for (int j = 0; j < n; ++j) v += H[(size_t)j * m + a] * X[(size_t)j * m + b];
            A[(size_t)a * m + b] = v;
        }
    if (type == 0 && eta < etamax) eta *= std::exp(etatau);
    for (int a = 0; a < m; ++a) A[(size_t)a * m + a] += type == 0 ? 1.0 / eta : lambda;
    // A^-1
    vector<double> Ai((size_t)m * m, 0.0);
    for (int a = 0; a < m; ++a) Ai[(size_t)a * m + a] = 1.0;
    for (int col = 0; col < m; ++col) {
        int piv = col;
        for (int r = col + 1; r < m; ++r)
            if (std::fabs(A[(size_t)r * m + col]) > std::fabs(A[(size_t)piv * m + col])) piv = r;
        if (A[(size_t)piv * m + col] == 0.0) {
            g_err = "kalman: singular innovation matrix";
            return 1;
        }
        if (piv != col)
            for (int k = 0; k < m; ++k) {
                std::swap(A[(size_t)piv * m + k], A[(size_t)col * m + k]);
                std::swap(Ai[(size_t)piv * m + k], Ai[(size_t)col * m + k]);
            }
        const double d = A[(size_t)col * m + col];
        for (int k = 0; k < m; ++k) {
            A[(size_t)col * m + k] /= d;
            Ai[(size_t)col * m + k] /= d;
        }
        for (int r = 0; r < m; ++r)
            if (r != col) {
                const double f = A[(size_t)r * m + col];
                for (int k = 0; k < m; ++k) {
                    A[(size_t)r * m + k] -= f * A[(size_t)col * m + k];
                    Ai[(size_t)r * m + k] -= f * Ai[(size_t)col * m + k];
                }
            }
    }
    // K = X A^-1
    vector<double> K((size_t)n * m, 0.0);
    for (int i = 0; i < n; ++i)
        for (int b = 0; b < m; ++b) {
            double v = 0.0;
            for (int a = 0; a < m; ++a) v += X[(size_t)i * m + a] * Ai[(size_t)a * m + b];
            K[(size_t)i * m + b] = v;
        }
    // P = (P - K X^T) (/ lambda) + q I, symmetric from the lower triangle
    const double inv = type == 1 ? 1.0 / lambda : 1.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double v = 0.0;
            for (int b = 0; b < m; ++b) v += K[(size_t)i * m + b] * X[(size_t)j * m + b];
            double p = P[(size_t)i * n + j] - v;
            if (type == 1) p *= inv;
            if (i == j) p += q;
            P[(size_t)i * n + j] = p;
            P[(size_t)j * n + i] = p;
        }
    // w = w + K xi
    for (int i = 0; i < n; ++i) {
        double v = 0.0;
        for (int b = 0; b < m; ++b) v += K[(size_t)i * m + b] * xi[b];
        w[i] += v;
    }
    if (q > qmin) q *= std::exp(-qtau);
    if (type == 1) {
        lambda = nu * lambda + 1.0 - nu;
        gamma = 1.0 / (1.0 + lambda / gamma);
    }
    return 0;
}

int oracle_num_neighbors(const oracle_ctx* c, int i) { return (int)c->atoms.at(i).nbr.size(); }

int oracle_get_neighbors(const oracle_ctx* c, int i, int* j, int* bc)
{
    const Atom& a = c->atoms.at(i);
    for (size_t q = 0; q < a.nbr.size(); ++q) {
        j[q] = a.nbr[q].j;
        for (int k = 0; k < 3; ++k) bc[3 * q + k] = a.nbr[q].bc[k];
    }
    return 0;
}

double oracle_phase_time(const oracle_ctx* c, int p) { return (p >= 0 && p < 5) ? c->phase[p] : 0.0; }

int oracle_cutoff_fdf(int type, double rc, double alpha, double r, double* fc, double* dfc)
{
    ORACLE_TRY({
        Cutoff f;
        f.set(type, rc, alpha);
        f.fdf(r, *fc, *dfc);
    })
    return 0;
}

}  // extern "C"

int oracle_num_connections(const oracle_ctx* c, int e)
{
    if (e < 0 || e >= (int)c->el.size()) return -1;
    return (int)c->el[e].nn.numConnections();
}

int oracle_get_weights(const oracle_ctx* c, double* w)
{
    ORACLE_TRY({
        size_t p = 0;
        for (auto const& e : c->el) {
            const NN& nn = e.nn;
            for (size_t l = 1; l < nn.n.size(); ++l) {
                for (int j = 0; j < nn.n[l - 1]; ++j)
                    for (int k = 0; k < nn.n[l]; ++k) w[p++] = nn.W[l][(size_t)k * nn.n[l - 1] + j];
                for (int k = 0; k < nn.n[l]; ++k) w[p++] = nn.b[l][k];
            }
        }
    });
    return 0;
}

int oracle_set_weights(oracle_ctx* c, const double* w)
{
    ORACLE_TRY({
        size_t p = 0;
        for (auto& e : c->el) {
            const size_t n = e.nn.numConnections();
            e.nn.setConnections(vector<double>(w + p, w + p + n));
            p += n;
        }
    });
    return 0;
}

// Training::dPdc (Training.cpp:3319-3389) after oracle_predict:
// property 0 ("energy"): out[c] = sum_a dE_a/dw_c (NeuralNetwork::calculateDEdc);
// property 1 ("force"): out[(3 ia + k) * n_conn + c] = sum_a dF_{ia,k}/dw_c,
// NeuralNetwork::calculateDFdc with dG_a/dx_{ia,k} from
// Training::collectDGdxia (Training.cpp:3586-3625): the entries of atom a's
// neighbours that are atom ia (self images included) plus, for a == ia, the
// atom's own derivative.  Connections of all elements in element order.
int oracle_dPdc(oracle_ctx* c, int property, double* out)
{
    ORACLE_TRY({
        if (c->normalize) throw std::runtime_error("dPdc: normalised models are not supported");
        if (property != 0 && property != 1) throw std::runtime_error("dPdc: property must be 0 or 1");
        const size_t N = c->atoms.size();
        for (auto const& a : c->atoms)
            if (a.dEdG.empty() || a.dGdr.empty()) throw std::runtime_error("dPdc: run oracle_predict first");
        vector<size_t> off(c->el.size() + 1, 0);
        for (size_t e = 0; e < c->el.size(); ++e) off[e + 1] = off[e] + c->el[e].nn.numConnections();
        const size_t nc = off.back();
        const size_t rows = property == 0 ? 1 : 3 * N;
        for (size_t q = 0; q < rows * nc; ++q) out[q] = 0.0;
        vector<NNState> st(N);
        for (size_t i = 0; i < N; ++i) nn_forward2(c->el[c->atoms[i].element].nn, c->atoms[i].G, st[i]);
        if (property == 0) {
            for (size_t i = 0; i < N; ++i) {
                const NN& nn = c->el[c->atoms[i].element].nn;
                const ConnOffsets o = conn_offsets(nn);
                vector<double> tmp(o.n);
                nn_dEdc(nn, st[i], o, tmp.data());
                double* g = out + off[c->atoms[i].element];
                for (size_t q = 0; q < o.n; ++q) g[q] += tmp[q];
            }
            return 0;
        }
        for (size_t ia = 0; ia < N; ++ia)
            for (int k = 0; k < 3; ++k) {
                double* row = out + (3 * ia + k) * nc;
                for (size_t i = 0; i < N; ++i) {
                    const Atom& a = c->atoms[i];
                    const size_t ns = a.G.size();
                    vector<double> dGdxia(ns, 0.0);
                    for (const Neighbor& n : a.nbr) {
                        if ((size_t)n.j != ia) continue;
                        const vector<int>& t = c->el[a.element].table[n.element];
                        for (size_t q = 0; q < t.size(); ++q) dGdxia[t[q]] += n.dGdr[3 * q + k];
                    }
                    if (i == ia)
                        for (size_t q = 0; q < ns; ++q) dGdxia[q] += a.dGdr[3 * q + k];
                    const NN& nn = c->el[a.element].nn;
                    const ConnOffsets o = conn_offsets(nn);
                    vector<double> tmp(o.n);
                    nn_dFdc(nn, st[i], o, dGdxia.data(), tmp.data());
                    double* g = row + off[a.element];
                    for (size_t q = 0; q < o.n; ++q) g[q] += tmp[q];
                }
            }
    });
    return 0;
}

int oracle_loss_gradient(oracle_ctx* c, double E_ref, const double* F_ref, double wE, double wF, double* loss,
                         double* grad)
{
    ORACLE_TRY({
        if (c->normalize) throw std::runtime_error("loss gradient: normalised models are not supported");
        const size_t N = c->atoms.size();
        for (auto const& a : c->atoms)
            if (a.dEdG.empty() || a.dGdr.empty()) throw std::runtime_error("loss gradient: run oracle_predict first");
        vector<size_t> off(c->el.size() + 1, 0);
        for (size_t e = 0; e < c->el.size(); ++e) off[e + 1] = off[e] + c->el[e].nn.numConnections();
        for (size_t q = 0; q < off.back(); ++q) grad[q] = 0.0;
        const double r = c->energy - E_ref;
        double lf = 0.0;
        vector<double> u(3 * N);
        for (size_t a = 0; a < N; ++a)
            for (int k = 0; k < 3; ++k) {
                const double d = c->atoms[a].f[k] - F_ref[3 * a + k];
                lf += d * d;
                u[3 * a + k] = 2.0 * wF * d;
            }
        loss[1] = wE * r * r;
        loss[2] = wF * lf;
        loss[0] = loss[1] + loss[2];
        // Training::collectDGdxia contracted with u: every neighbour entry of
        // atom i (self images included) and atom i's own derivative.
        vector<vector<double>> gdot(N);
        for (size_t i = 0; i < N; ++i) {
            const Atom& a = c->atoms[i];
            const size_t ns = a.G.size();
            vector<double>& g = gdot[i];
            g.assign(ns, 0.0);
            for (const Neighbor& n : a.nbr) {
                const vector<int>& t = c->el[a.element].table[n.element];
                for (size_t q = 0; q < t.size(); ++q)
                    for (int k = 0; k < 3; ++k) g[t[q]] += n.dGdr[3 * q + k] * u[3 * n.j + k];
            }
            for (size_t q = 0; q < ns; ++q)
                for (int k = 0; k < 3; ++k) g[q] += a.dGdr[3 * q + k] * u[3 * i + k];
        }
        // Training::dPdc (Training.cpp:3319-3389) weighted by the loss adjoints.
        const double cE = 2.0 * wE * r;
        for (size_t i = 0; i < N; ++i) {
            const Atom& a = c->atoms[i];
            const NN& nn = c->el[a.element].nn;
            const ConnOffsets o = conn_offsets(nn);
            NNState st;
            nn_forward2(nn, a.G, st);
            vector<double> dEdc(o.n), dFdc(o.n);
            nn_dEdc(nn, st, o, dEdc.data());
            nn_dFdc(nn, st, o, gdot[i].data(), dFdc.data());
            double* g = grad + off[a.element];
            for (size_t q = 0; q < o.n; ++q) g[q] += cE * dEdc[q] + dFdc[q];
        }
    });
    return 0;
}
