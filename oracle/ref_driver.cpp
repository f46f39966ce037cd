// Driver for the reference's own compute classes -- TEST INFRASTRUCTURE.
//
// Built by oracle/Makefile into oracle/_ref/libnnpref.so together with the
// n2p2 sources that compile without Eigen (Element, SymFnc*, SymGrp*,
// CutoffFunction, CoreFunction, NeuralNetwork, Atom, Settings, ...).  The
// Eigen-dependent Structure/Mode classes are not built; this file restates
// only their orchestration (neighbour list, setup order, units) and calls the
// reference classes for everything on the hot path:
//   Element::calculateSymmetryFunctionGroups   (Element.cpp:449-460)
//   NeuralNetwork::propagate / calculateDEdG   (NeuralNetwork.cpp:386-442)
//   Atom::calculateSelfForceShort / calculatePairForceShort (Atom.cpp:371-401)
// It exists to pin the from-scratch oracle (nnp_oracle.cpp) against the
// reference code itself; only tests load it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <map>
#include <omp.h>
#include <stdexcept>
#include <string>
#include <vector>

#include "Atom.h"
#include "Element.h"
#include "ElementMap.h"
#include "NeuralNetwork.h"
#include "Settings.h"
#include "SymFnc.h"
#include "utility.h"

using namespace nnp;
using std::string;
using std::vector;

namespace {
thread_local string g_err;
}

struct ref_ctx {
    settings::Settings settings;
    ElementMap em;
    vector<Element> elements;
    double maxCutoff = 0.0;
    bool normalize = false;
    double meanEnergy = 0, convEnergy = 1, convLength = 1;
    vector<Atom> atoms;
    bool periodic = false;
    double box[3][3] = {{0}}, invbox[3][3] = {{0}};
    double energy = 0.0;
    double tphase[5] = {0, 0, 0, 0, 0};
};

namespace {

// Restates the setup order of Prediction::setup (Prediction.cpp:36-48) /
// Mode::setupGeneric (Mode.cpp:213-237) using the reference classes.
void setup(ref_ctx& c, const string& dir)
{
    c.settings.loadFile(dir + "/input.nn");
    settings::Settings& s = c.settings;
    if (s.keywordExists("mean_energy") && s.keywordExists("conv_energy") &&
        s.keywordExists("conv_length")) {
        c.normalize = true;
        c.meanEnergy = atof(s["mean_energy"].c_str());
        c.convEnergy = atof(s["conv_energy"].c_str());
        c.convLength = atof(s["conv_length"].c_str());
    }
    c.em.registerElements(s["elements"]);
    for (size_t i = 0; i < c.em.size(); ++i) c.elements.push_back(Element(i, c.em));
    if (s.keywordExists("atom_energy")) {
        auto r = s.getValues("atom_energy");
        for (auto it = r.first; it != r.second; ++it) {
            vector<string> a = split(reduce(it->second.first));
            c.elements.at(c.em[a.at(0)]).setAtomicEnergyOffset(atof(a.at(1).c_str()));
        }
    }
    vector<string> ct = split(s["cutoff_type"]);
    auto cutoffType = (CutoffFunction::CutoffType)atoi(ct.at(0).c_str());
    double alpha = ct.size() > 1 ? atof(ct.at(1).c_str()) : 0.0;
    auto r = s.getValues("symfunction_short");
    for (auto it = r.first; it != r.second; ++it) {
        vector<string> a = split(reduce(it->second.first));
        c.elements.at(c.em[a.at(0)]).addSymmetryFunction(it->second.first, it->second.second);
    }
    for (auto& e : c.elements) {
        if (c.normalize) e.changeLengthUnitSymmetryFunctions(c.convLength);
        e.sortSymmetryFunctions();
        c.maxCutoff = std::max(e.getMaxCutoffRadius(), c.maxCutoff);
        e.setCutoffFunction(cutoffType, alpha);
    }
    for (auto& e : c.elements) e.setupSymmetryFunctionMemory();
    for (auto& e : c.elements) e.setupSymmetryFunctionGroups();
    // Neural networks (global topology keywords only).
    int nh = atoi(s["global_hidden_layers_short"].c_str());
    vector<string> nodes = split(reduce(s["global_nodes_short"]));
    vector<string> acts = split(reduce(s["global_activation_short"]));
    for (auto& e : c.elements) {
        vector<int> n(nh + 2);
        vector<NeuralNetwork::ActivationFunction> af(nh + 2);
        n[0] = (int)e.numSymmetryFunctions();
        af[0] = NeuralNetwork::AF_IDENTITY;
        for (int l = 1; l <= nh; ++l) {
            n[l] = atoi(nodes.at(l - 1).c_str());
            af[l] = activationFromString(acts.at(l - 1));
        }
        n[nh + 1] = 1;
        af[nh + 1] = activationFromString(acts.at(nh));
        const int* np = n.data();
        const NeuralNetwork::ActivationFunction* ap = af.data();
        e.neuralNetworks.emplace(std::piecewise_construct, std::forward_as_tuple("short"),
                                 std::forward_as_tuple(nh + 2, np, ap));
        e.neuralNetworks.at("short").setNormalizeNeurons(s.keywordExists("normalize_nodes"));
    }
    // Scaling (Mode::setupSymmetryFunctionScaling, Mode.cpp:781-909).
    SymFnc::ScalingType st = SymFnc::ST_NONE;
    bool sc = s.keywordExists("scale_symmetry_functions");
    bool ce = s.keywordExists("center_symmetry_functions");
    if (sc && !ce) st = SymFnc::ST_SCALE;
    else if (!sc && ce) st = SymFnc::ST_CENTER;
    else if (sc && ce) st = SymFnc::ST_SCALECENTER;
    else if (s.keywordExists("scale_symmetry_functions_sigma")) st = SymFnc::ST_SCALESIGMA;
    double Smin = 0.0, Smax = 0.0;
    if (st == SymFnc::ST_SCALE || st == SymFnc::ST_SCALECENTER || st == SymFnc::ST_SCALESIGMA) {
        Smin = s.keywordExists("scale_min_short") ? atof(s["scale_min_short"].c_str()) : 0.0;
        Smax = s.keywordExists("scale_max_short") ? atof(s["scale_max_short"].c_str()) : 1.0;
    }
    std::ifstream f((dir + "/scaling.data").c_str());
    if (!f) throw std::runtime_error("could not open scaling.data");
    string line;
    vector<string> lines;
    while (std::getline(f, line))
        if (!line.empty() && line.at(0) != '#') lines.push_back(line);
    for (auto& e : c.elements) {
        e.setScaling(st, lines, Smin, Smax);
        lines.erase(lines.begin(), lines.begin() + e.numSymmetryFunctions());
    }
    // Weights (Mode::readNeuralNetworkWeights, Mode.cpp:2331-2350).
    for (auto& e : c.elements) {
        string fn = dir + strpr("/weights.%03zu.data", e.getAtomicNumber());
        vector<double> w = readColumnsFromFile(fn, vector<size_t>(1, 0)).at(0);
        e.neuralNetworks.at("short").setConnections(&(w.front()));
    }
}

// Structure::calculateInverseBox / remap / calculatePbcCopies /
// calculateNeighborList restated with Vec3D (Structure.cpp:287-544, 1224-1248).
void inverse_box(ref_ctx& c)
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

Vec3D row(const double (&m)[3][3], int i) { return Vec3D(m[i][0], m[i][1], m[i][2]); }

void remap(ref_ctx& c, Atom& a)
{
    Vec3D f = a.r[0] * row(c.invbox, 0) + a.r[1] * row(c.invbox, 1) + a.r[2] * row(c.invbox, 2);
    for (int k = 0; k < 3; ++k) if (f[k] > 1.0) f[k] -= (int)f[k];
    for (int k = 0; k < 3; ++k) if (f[k] < 0.0) f[k] += 1.0 - (int)f[k];
    for (int k = 0; k < 3; ++k) if (f[k] == 1.0) f[k] = 0.0;
    a.r = f[0] * row(c.box, 0) + f[1] * row(c.box, 1) + f[2] * row(c.box, 2);
}

void neighbors(ref_ctx& c)
{
    const size_t N = c.atoms.size();
    int pbc[3] = {0, 0, 0};
    double rc = c.maxCutoff;
    Vec3D b0 = row(c.box, 0), b1 = row(c.box, 1), b2 = row(c.box, 2);
    if (c.periodic) {
        Vec3D axb = b0.cross(b1).normalize();
        Vec3D axc = b0.cross(b2).normalize();
        Vec3D bxc = b1.cross(b2).normalize();
        double proj[3] = {fabs(b0 * bxc), fabs(b1 * axc), fabs(b2 * axb)};
        for (int k = 0; k < 3; ++k) while (pbc[k] * proj[k] <= rc) pbc[k]++;
    }
    double const rc2 = rc * rc;
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t i = 0; i < N; ++i) {
        Atom& ai = c.atoms[i];
        ai.neighbors.clear();
        ai.neighborsUnique.assign(1, i);
        ai.numNeighborsUnique = 1;
        ai.numNeighbors = 0;
        ai.numNeighborsPerElement.assign(c.em.size(), 0);
        for (size_t j = 0; j < N; ++j)
            for (int q0 = -pbc[0]; q0 <= pbc[0]; q0++)
                for (int q1 = -pbc[1]; q1 <= pbc[1]; q1++)
                    for (int q2 = -pbc[2]; q2 <= pbc[2]; q2++) {
                        if (i == j && q0 == 0 && q1 == 0 && q2 == 0) continue;
                        Vec3D dr = ai.r - c.atoms[j].r + q0 * b0 + q1 * b1 + q2 * b2;
                        if (dr.norm2() <= rc2) {
                            ai.neighbors.push_back(Atom::Neighbor());
                            Atom::Neighbor& n = ai.neighbors.back();
                            n.index = j;
                            n.tag = j;
                            n.element = c.atoms[j].element;
                            n.d = dr.norm();
                            n.dr = dr;
                            ai.numNeighborsPerElement[c.atoms[j].element]++;
                            ai.numNeighbors++;
                            if (ai.neighborsUnique.back() != j && i != j) {
                                ai.neighborsUnique.push_back(j);
                                ai.numNeighborsUnique++;
                            }
                        }
                    }
        ai.hasNeighborList = true;
    }
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }

// OpenMP threads of the parallel loops (the reference's OpenMP sites); the
// CPU baseline times both the full core count and one thread.
void ref_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
}

ref_ctx* ref_create(const char* dir)
{
    ref_ctx* c = new ref_ctx();
    try {
        setup(*c, dir);
    } catch (const std::exception& ex) {
        g_err = ex.what();
        delete c;
        return nullptr;
    }
    return c;
}

void ref_destroy(ref_ctx* c) { delete c; }

int ref_set_structure(ref_ctx* c, int n, const double* pos, const int* elem, const double* box)
{
    try {
        c->atoms.assign(n, Atom());
        c->periodic = box != nullptr;
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) c->box[i][k] = box ? box[3 * i + k] : 0.0;
        if (c->periodic) inverse_box(*c);
        for (int i = 0; i < n; ++i) {
            Atom& a = c->atoms[i];
            a.index = i;
            a.tag = i;
            a.element = elem[i];
            a.r = Vec3D(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
            a.numNeighborsPerElement.resize(c->em.size(), 0);
            if (c->periodic) remap(*c, a);
        }
        if (c->normalize) {  // Structure::toNormalizedUnits
            if (c->periodic) {
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < 3; ++k) c->box[i][k] *= c->convLength;
                double const inv = 1.0 / c->convLength;
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < 3; ++k) c->invbox[i][k] *= inv;
            }
            for (auto& a : c->atoms) a.r *= c->convLength;
        }
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
    return 0;
}

// Mode::evaluateNNP (Mode.cpp:2029-2077) + Prediction::predict tail.
int ref_predict(ref_ctx* c)
{
    try {
        double t0 = now();
        neighbors(*c);
        double t1 = now();
        const size_t N = c->atoms.size();
        // Mode::calculateSymmetryFunctionGroups (Mode.cpp:1630-1709)
#pragma omp parallel for schedule(dynamic, 4)
        for (size_t i = 0; i < N; ++i) {
            Atom& a = c->atoms[i];
            Element& e = c->elements.at(a.element);
            a.numSymmetryFunctions = e.numSymmetryFunctions();
            a.numSymmetryFunctionDerivatives = e.getSymmetryFunctionNumTable();
            a.allocate(true, c->maxCutoff);
            e.calculateSymmetryFunctionGroups(a, true);
            a.hasSymmetryFunctions = true;
            a.hasSymmetryFunctionDerivatives = true;
        }
        double t2 = now();
        // Mode::calculateAtomicNeuralNetworks (Mode.cpp:1711-1732), serial
        for (auto& a : c->atoms) {
            NeuralNetwork& nn = c->elements.at(a.element).neuralNetworks.at("short");
            nn.setInput(&(a.G.front()));
            nn.propagate();
            nn.calculateDEdG(&(a.dEdG.front()));
            nn.getOutput(&(a.energy));
        }
        double t3 = now();
        c->energy = 0.0;
        for (auto& a : c->atoms) c->energy += a.energy;
        double t4 = now();
        // Mode::calculateForces (Mode.cpp:1922-1984)
#pragma omp parallel for
        for (size_t i = 0; i < N; ++i) {
            Atom& ai = c->atoms.at(i);
            ai.f = Vec3D{};
            ai.f += ai.calculateSelfForceShort();
            for (auto it = ai.neighborsUnique.begin() + 1; it != ai.neighborsUnique.end(); ++it) {
                Atom& aj = c->atoms.at(*it);
                vector<vector<size_t>> const& tableFull = c->elements.at(aj.element).getSymmetryFunctionTable();
                for (size_t k = 0; k < aj.numNeighbors; ++k) {
                    Atom::Neighbor const& n = aj.neighbors[k];
                    if (n.index == ai.index) ai.f += aj.calculatePairForceShort(n, &tableFull);
                }
            }
        }
        double t5 = now();
        c->tphase[0] = t1 - t0; c->tphase[1] = t2 - t1; c->tphase[2] = t3 - t2;
        c->tphase[3] = t4 - t3; c->tphase[4] = t5 - t4;
        if (c->normalize) {  // Structure::toPhysicalUnits
            c->energy = c->energy / c->convEnergy + N * c->meanEnergy;
            for (auto& a : c->atoms) {
                a.f /= c->convEnergy / c->convLength;
                a.r /= c->convLength;
            }
        }
        for (size_t e = 0; e < c->elements.size(); ++e) {  // Mode::addEnergyOffset
            size_t cnt = 0;
            for (auto& a : c->atoms) cnt += (a.element == e);
            c->energy += cnt * c->elements[e].getAtomicEnergyOffset();
        }
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
    return 0;
}

int ref_num_sf(const ref_ctx* c, int e) { return (int)c->elements.at(e).numSymmetryFunctions(); }
double ref_energy(const ref_ctx* c) { return c->energy; }
int ref_get_forces(const ref_ctx* c, double* f)
{
    for (size_t i = 0; i < c->atoms.size(); ++i)
        for (int k = 0; k < 3; ++k) f[3 * i + k] = c->atoms[i].f[k];
    return 0;
}
int ref_get_G(const ref_ctx* c, double* G, int ld)
{
    for (size_t i = 0; i < c->atoms.size(); ++i)
        for (size_t s = 0; s < c->atoms[i].G.size() && (int)s < ld; ++s) G[i * ld + s] = c->atoms[i].G[s];
    return 0;
}
int ref_get_dEdG(const ref_ctx* c, double* d, int ld)
{
    for (size_t i = 0; i < c->atoms.size(); ++i)
        for (size_t s = 0; s < c->atoms[i].dEdG.size() && (int)s < ld; ++s) d[i * ld + s] = c->atoms[i].dEdG[s];
    return 0;
}
// Atom::dGdr and Atom::Neighbor::dGdr of the reference's Atom objects after
// ref_predict, in the layout of oracle_get_dGdr.
int ref_get_dGdr(const ref_ctx* c, double* self, int ld_self, double* nbr, int ld_nbr)
{
    size_t p = 0;
    for (size_t i = 0; i < c->atoms.size(); ++i) {
        const Atom& a = c->atoms[i];
        if (self)
            for (int q = 0; q < ld_self; ++q)
                for (int k = 0; k < 3; ++k)
                    self[(i * ld_self + q) * 3 + k] = (size_t)q < a.dGdr.size() ? a.dGdr[q][k] : 0.0;
        for (const Atom::Neighbor& n : a.neighbors) {
            if (nbr)
                for (int q = 0; q < ld_nbr; ++q)
                    for (int k = 0; k < 3; ++k)
                        nbr[(p * ld_nbr + q) * 3 + k] = (size_t)q < n.dGdr.size() ? n.dGdr[q][k] : 0.0;
            ++p;
        }
    }
    return 0;
}
// Extrapolation warnings through the reference's own
// Element::updateSymmetryFunctionStatistics (Element.cpp:462-541) with
// collectExtrapolationWarnings on, atoms in order (tags = atom indices);
// records gathered from SymFncStatistics and put in (atom, SF) order.
long ref_extrapolation_warnings(ref_ctx* c, long max, int* atom, int* sf, int* type, double* value, double* gmin,
                                double* gmax)
{
    try {
        for (auto& e : c->elements) {
            e.statistics.collectStatistics = false;
            e.statistics.collectExtrapolationWarnings = true;
            e.statistics.writeExtrapolationWarnings = false;
            e.statistics.stopOnExtrapolationWarnings = false;
            e.statistics.resetExtrapolationWarnings();
        }
        for (size_t i = 0; i < c->atoms.size(); ++i) {
            Atom& a = c->atoms[i];
            a.tag = (int64_t)i;
            c->elements.at(a.element).updateSymmetryFunctionStatistics(a);
        }
        struct Rec { int atom, sf, type; double value, gmin, gmax; };
        vector<Rec> rec;
        for (auto const& e : c->elements)
            for (auto const& kv : e.statistics.data) {
                auto const& d = kv.second;
                for (size_t k = 0; k < d.valueEW.size(); ++k)
                    rec.push_back(Rec{(int)d.indexAtomEW[k], (int)kv.first, (int)d.type, d.valueEW[k], d.Gmin, d.Gmax});
            }
        std::sort(rec.begin(), rec.end(), [](Rec const& x, Rec const& y) {
            return x.atom != y.atom ? x.atom < y.atom : x.sf < y.sf;
        });
        for (size_t r = 0; r < rec.size() && (long)r < max; ++r) {
            atom[r] = rec[r].atom;
            sf[r] = rec[r].sf;
            type[r] = rec[r].type;
            value[r] = rec[r].value;
            gmin[r] = rec[r].gmin;
            gmax[r] = rec[r].gmax;
        }
        return (long)rec.size();
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// The low-neighbour check (Mode.cpp:1667-1679) with the reference's
// Atom::calculateNumNeighbors and Element::getMinCutoffRadius /
// getMinNeighbors.
long ref_low_neighbor_atoms(const ref_ctx* c, long max, int* atom, int* num)
{
    long n = 0;
    for (size_t i = 0; i < c->atoms.size(); ++i) {
        Atom const& a = c->atoms[i];
        Element const& e = c->elements.at(a.element);
        const size_t cnt = a.calculateNumNeighbors(e.getMinCutoffRadius());
        if (cnt < e.getMinNeighbors()) {
            if (n < max) {
                atom[n] = (int)i;
                num[n] = (int)cnt;
            }
            ++n;
        }
    }
    return n;
}

int ref_get_atomic_energies(const ref_ctx* c, double* e)
{
    for (size_t i = 0; i < c->atoms.size(); ++i) e[i] = c->atoms[i].energy;
    return 0;
}
double ref_phase_time(const ref_ctx* c, int p) { return (p >= 0 && p < 5) ? c->tphase[p] : 0.0; }

// Training::dPdc (Training.cpp:3319-3389) with the reference's
// NeuralNetwork::calculateDEdc / calculateDFdc; the dG/dx_{ia,k} gather of
// Training::collectDGdxia (Training.cpp:3586-3625, not part of the compiled
// subset) is restated here.  After ref_predict of a non-normalised model.
// property 0: out[n_conn]; property 1: out[3N][n_conn].
int ref_dPdc(ref_ctx* c, int property, double* out)
{
    try {
        if (c->normalize) throw std::runtime_error("dPdc: normalised models are not supported");
        vector<size_t> off(c->elements.size() + 1, 0);
        for (size_t e = 0; e < c->elements.size(); ++e)
            off[e + 1] = off[e] + c->elements[e].neuralNetworks.at("short").getNumConnections();
        const size_t nc = off.back(), N = c->atoms.size();
        const size_t rows = property == 0 ? 1 : 3 * N;
        for (size_t q = 0; q < rows * nc; ++q) out[q] = 0.0;
        for (size_t r = 0; r < rows; ++r) {
            const size_t ia = r / 3, k = r % 3;
            for (auto& a : c->atoms) {
                NeuralNetwork& nn = c->elements.at(a.element).neuralNetworks.at("short");
                vector<double> tmp(nn.getNumConnections(), 0.0);
                nn.setInput(a.G.data());
                nn.propagate();
                if (property == 0) {
                    nn.calculateDEdc(tmp.data());
                } else {
                    vector<double> dGdxia(a.numSymmetryFunctions, 0.0);
                    vector<vector<size_t>> const& tableFull = c->elements.at(a.element).getSymmetryFunctionTable();
                    for (size_t i = 0; i < a.numNeighbors; ++i) {
                        Atom::Neighbor const& n = a.neighbors[i];
                        if (n.index != ia) continue;
                        vector<size_t> const& table = tableFull.at(n.element);
                        for (size_t j = 0; j < n.dGdr.size(); ++j) dGdxia[table.at(j)] += n.dGdr[j][k];
                    }
                    if (a.index == ia)
                        for (size_t i = 0; i < a.numSymmetryFunctions; ++i) dGdxia[i] += a.dGdr[i][k];
                    nn.calculateDEdG(a.dEdG.data());
                    nn.calculateDFdc(tmp.data(), dGdxia.data());
                }
                for (size_t j = 0; j < tmp.size(); ++j) out[r * nc + off[a.element] + j] += tmp[j];
            }
        }
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
    return 0;
}

}  // extern "C"
