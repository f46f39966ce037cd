// nnpx::Mode and nnpx::Prediction: the n2p2 Mode/Prediction compute surface
// (Mode.cpp:1630-1984, 2029-2130; Prediction.cpp:36-78) over the nnpx C ABI.
// The device holds one structure at a time; a Structure object becomes
// resident when calculateSymmetryFunctionGroups is called on it, and the
// later calls of the same evaluation copy their results back into it.
#include <cinttypes>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "nnpx_mode.hpp"

namespace nnpx {

namespace {

// Map a C-ABI status to the exception type n2p2 throws for that condition.
[[noreturn]] void raise(nnpx_status st, const char* msg)
{
    std::string m = msg ? msg : "";
    switch (st) {
    case NNPX_ERR_ARGUMENT: throw std::invalid_argument(m);
    case NNPX_ERR_RANGE: throw std::range_error(m);
    default: throw std::runtime_error(m);
    }
}

}  // namespace

Mode::Mode(int device)
{
    nnpx_status st = nnpx_create(&ctx_, device);
    if (st != NNPX_OK) raise(st, nnpx_last_error(ctx_));
}

Mode::~Mode() { nnpx_destroy(ctx_); }

void Mode::check(nnpx_status st) const
{
    if (st != NNPX_OK) raise(st, nnpx_last_error(ctx_));
}

void Mode::setupFromDirectory(std::string const& directory)
{
    check(nnpx_load_model(ctx_, directory.c_str()));
    int ne = 0;
    check(nnpx_model_info(ctx_, &ne, &maxSF_, &maxCutoffRadius_));
    std::string syms;
    for (int e = 0; e < ne; ++e) {
        char b[8];
        check(nnpx_element_symbol(ctx_, e, b, sizeof b));
        syms += std::string(b) + " ";
    }
    elementMap.registerElements(syms);
    int norm = 0;
    atomicEnergyOffset_.assign(ne, 0.0);
    check(nnpx_model_units(ctx_, &norm, &meanEnergy_, &convEnergy_, &convLength_, atomicEnergyOffset_.data()));
    normalize_ = norm != 0;
    resident_ = nullptr;
}

void Mode::setForceMode(ForceMode m) { check(nnpx_set_option(ctx_, NNPX_OPT_FORCE_MODE, (int64_t)m)); }

void Mode::setupSymmetryFunctionStatistics(bool collectStatistics, bool collectExtrapolationWarnings,
                                           bool writeExtrapolationWarnings, bool stopOnExtrapolationWarnings)
{
    if (collectStatistics)
        throw std::runtime_error("ERROR: symmetry function statistics (min/max/mean) are not collected by this "
                                 "library.\n");
    collectEW_ = collectExtrapolationWarnings;
    writeEW_ = writeExtrapolationWarnings;
    stopEW_ = stopOnExtrapolationWarnings;
}

// The checks of Mode::calculateSymmetryFunctionGroups on the resident
// structure: low-neighbour atoms to the log (Mode.cpp:1667-1679) and, when
// enabled, extrapolation warnings (Element.cpp:462-541, Mode.cpp:1694-1702).
void Mode::checkStructure(Structure const& s)
{
    if (log_) {
        int64_t n = 0;
        check(nnpx_low_neighbor_atoms(ctx_, 0, &n, nullptr, nullptr));
        if (n > 0) {
            std::vector<int> a(n), c(n);
            check(nnpx_low_neighbor_atoms(ctx_, n, &n, a.data(), c.data()));
            for (int64_t k = 0; k < n; ++k)
                *log_ << strpr("WARNING: Structure %6zu Atom %6zu : %zu neighbors.\n", s.index,
                               s.atoms.at(a[k]).index, (size_t)c[k]);
        }
    }
    if (!collectEW_ && !writeEW_ && !stopEW_) return;
    int64_t n = 0;
    check(nnpx_extrapolation_warnings(ctx_, 0, &n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
    if (n == 0) return;
    std::vector<int> a(n), q(n), t(n);
    std::vector<double> v(n), lo(n), hi(n);
    check(nnpx_extrapolation_warnings(ctx_, n, &n, a.data(), q.data(), t.data(), v.data(), lo.data(), hi.data()));
    const double cl = normalize_ ? convLength_ : 1.0;
    (void)cl;
    for (int64_t k = 0; k < n; ++k) {
        Atom const& at = s.atoms.at(a[k]);
        const std::string line =
            strpr("### NNP EXTRAPOLATION WARNING ### STRUCTURE: %6zu ATOM: %9" PRId64 " ELEMENT: %2s SYMFUNC: %4zu "
                  "TYPE: %2zu VALUE: %10.3E MIN: %10.3E MAX: %10.3E\n",
                  at.indexStructure, at.tag, elementMap[at.element].c_str(), (size_t)q[k] + 1, (size_t)t[k], v[k],
                  lo[k], hi[k]);
        if (collectEW_) ewLines_.push_back(line);
        if (writeEW_) std::cerr << line;
        if (stopEW_) throw std::out_of_range(line + "ERROR: Symmetry function value out of range.\n");
    }
}

void Mode::upload(Structure& s)
{
    if (s.numElements != elementMap.size())
        throw std::runtime_error("ERROR: Structure element map does not match the model.\n");
    const size_t n = s.atoms.size();
    std::vector<double> pos(3 * n);
    std::vector<int> elem(n);
    for (size_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) pos[3 * i + c] = s.atoms[i].r[c];
        elem[i] = (int)s.atoms[i].element;
    }
    double box[9] = {0};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) box[3 * r + c] = s.box[r][c];
    const int64_t off[2] = {0, (int64_t)n};
    const int periodic = s.isPeriodic ? 1 : 0;
    check(nnpx_set_structures(ctx_, 1, off, pos.data(), elem.data(), box, &periodic));
    resident_ = &s;
    residentAtoms_ = n;
}

void Mode::requireResident(Structure const& s) const
{
    if (resident_ != &s || residentAtoms_ != s.atoms.size())
        throw std::runtime_error("ERROR: Symmetry functions of this structure were not calculated by this Mode "
                                 "(call calculateSymmetryFunctionGroups first).\n");
}

void Mode::calculateSymmetryFunctionGroups(Structure& s, bool const derivatives)
{
    // Same skip rules as Mode.cpp:1633-1635.
    if (s.hasSymmetryFunctionDerivatives) return;
    if (s.hasSymmetryFunctions && !derivatives) return;
    upload(s);
    if (precision_ == Precision::MIXED) {
        check(nnpx_evaluate(ctx_, NNPX_MIXED, derivatives ? NNPX_EVAL_FORCES : 0u));
        mixedForces_ = derivatives;
    } else {
        check(nnpx_calculate_neighbor_list(ctx_));
        check(nnpx_calculate_symmetry_function_groups(ctx_, derivatives ? 1 : 0));
    }
    const size_t n = s.atoms.size();
    std::vector<double> G(n * maxSF_);
    check(nnpx_get_G(ctx_, G.data(), maxSF_));
    int64_t total = 0;
    std::vector<int> counts(n);
    check(nnpx_get_neighbors(ctx_, &total, counts.data(), nullptr, nullptr));
    std::vector<int> jj, sh;
    if (copyNeighbors_) {
        jj.resize(total);
        sh.resize(3 * total);
        check(nnpx_get_neighbors(ctx_, &total, counts.data(), jj.data(), sh.data()));
    }
    std::vector<double> dself, dnbr;
    int ldn = 0;
    const bool dgdr = derivatives && copyDerivatives_ && precision_ == Precision::PLAIN_FP64;
    if (dgdr) {
        ldn = maxSF_;
        dself.resize(n * maxSF_ * 3);
        dnbr.resize((size_t)total * ldn * 3);
        check(nnpx_get_dGdr(ctx_, dself.data(), maxSF_, dnbr.data(), ldn));
    }
    size_t p = 0;
    for (size_t i = 0; i < n; ++i) {
        Atom& a = s.atoms[i];
        int nsf = 0;
        check(nnpx_num_sf(ctx_, (int)a.element, &nsf));
        a.numSymmetryFunctions = nsf;
        a.G.assign(G.begin() + i * maxSF_, G.begin() + i * maxSF_ + nsf);
        a.numNeighbors = counts[i];
        a.hasNeighborList = true;
        if (copyNeighbors_) {
            a.neighbors.resize(counts[i]);
            for (int k = 0; k < counts[i]; ++k) {
                Atom::Neighbor& nb = a.neighbors[k];
                const size_t e = p + k;
                nb.index = jj[e];
                nb.tag = jj[e];
                nb.element = s.atoms[jj[e]].element;
                // dr = ((ri - rj) + bc0*a) + bc1*b + bc2*c, Structure.cpp:318-335
                for (int c = 0; c < 3; ++c) {
                    double d = s.atoms[i].r[c] - s.atoms[jj[e]].r[c];
                    d += sh[3 * e] * s.box[0][c];
                    d += sh[3 * e + 1] * s.box[1][c];
                    d += sh[3 * e + 2] * s.box[2][c];
                    nb.dr[c] = d;
                }
                nb.d = std::sqrt(nb.dr.norm2());
            }
        }
        if (dgdr) {
            a.dGdr.resize(nsf);
            for (int q = 0; q < nsf; ++q)
                for (int c = 0; c < 3; ++c) a.dGdr[q][c] = dself[(i * maxSF_ + q) * 3 + c];
            if (copyNeighbors_) {
                for (int k = 0; k < counts[i]; ++k) {
                    int nsd = 0;
                    check(nnpx_num_sf_derivatives(ctx_, (int)a.element, (int)a.neighbors[k].element, &nsd));
                    a.neighbors[k].dGdr.resize(nsd);
                    for (int q = 0; q < nsd; ++q)
                        for (int c = 0; c < 3; ++c)
                            a.neighbors[k].dGdr[q][c] = dnbr[((p + k) * ldn + q) * 3 + c];
                }
            }
        }
        p += counts[i];
        a.hasSymmetryFunctions = true;
        a.hasSymmetryFunctionDerivatives = derivatives;
    }
    s.hasNeighborList = true;
    s.hasSymmetryFunctions = true;
    if (derivatives) s.hasSymmetryFunctionDerivatives = true;
    checkStructure(s);
}

void Mode::calculateAtomicNeuralNetworks(Structure& s, bool const derivatives, std::string)
{
    requireResident(s);
    if (precision_ == Precision::PLAIN_FP64) check(nnpx_calculate_atomic_neural_networks(ctx_, derivatives ? 1 : 0));
    const size_t n = s.atoms.size();
    std::vector<double> e(n);
    check(nnpx_get_atomic_energies(ctx_, e.data()));
    std::vector<double> d;
    if (derivatives) {
        d.resize(n * maxSF_);
        check(nnpx_get_dEdG(ctx_, d.data(), maxSF_));
    }
    for (size_t i = 0; i < n; ++i) {
        Atom& a = s.atoms[i];
        a.energy = e[i];
        if (derivatives) a.dEdG.assign(d.begin() + i * maxSF_, d.begin() + i * maxSF_ + a.numSymmetryFunctions);
    }
}

void Mode::calculateEnergy(Structure& s) const
{
    requireResident(s);
    if (precision_ == Precision::PLAIN_FP64) check(nnpx_calculate_energy(ctx_));
    double e = 0.0;
    check(nnpx_get_energy(ctx_, &e));
    s.energyShort = e;
    s.energy = s.energyShort + s.energyElec;
}

void Mode::calculateForces(Structure& s) const
{
    requireResident(s);
    if (precision_ == Precision::PLAIN_FP64) {
        check(nnpx_calculate_forces(ctx_));
    } else if (!mixedForces_) {
        check(nnpx_evaluate(ctx_, NNPX_MIXED, NNPX_EVAL_FORCES));
    }
    std::vector<double> f(3 * s.atoms.size());
    check(nnpx_get_forces(ctx_, f.data()));
    for (size_t i = 0; i < s.atoms.size(); ++i)
        for (int c = 0; c < 3; ++c) s.atoms[i].f[c] = f[3 * i + c];
}

void Mode::evaluateNNP(Structure& s, bool useForces, bool useDEdG)
{
    useDEdG = useForces || useDEdG;
    s.calculateNeighborList(maxCutoffRadius_, false);
    calculateSymmetryFunctionGroups(s, useForces);
    calculateAtomicNeuralNetworks(s, useDEdG);
    calculateEnergy(s);
    if (useForces) calculateForces(s);
}

void Mode::addEnergyOffset(Structure& s, bool ref)
{
    for (size_t e = 0; e < elementMap.size(); ++e) {
        if (ref) s.energyRef += s.numAtomsPerElement.at(e) * atomicEnergyOffset_.at(e);
        else s.energy += s.numAtomsPerElement.at(e) * atomicEnergyOffset_.at(e);
    }
}

void Mode::removeEnergyOffset(Structure& s, bool ref)
{
    for (size_t e = 0; e < elementMap.size(); ++e) {
        if (ref) s.energyRef -= s.numAtomsPerElement.at(e) * atomicEnergyOffset_.at(e);
        else s.energy -= s.numAtomsPerElement.at(e) * atomicEnergyOffset_.at(e);
    }
}

double Mode::getEnergyOffset(Structure const& s) const
{
    double r = 0.0;
    for (size_t e = 0; e < elementMap.size(); ++e) r += s.numAtomsPerElement.at(e) * atomicEnergyOffset_.at(e);
    return r;
}

// Prediction::setup (Prediction.cpp:36-48): extrapolation warnings written
// to stderr as they occur.
void Prediction::setup()
{
    setupFromDirectory(directory);
    setupSymmetryFunctionStatistics(false, false, true, false);
}

void Prediction::readStructureFromFile(std::string const& fileName)
{
    structure.reset();
    structure.setElementMap(elementMap);
    structure.readFromFile(fileName);
    removeEnergyOffset(structure);
    if (normalize_) structure.toNormalizedUnits(meanEnergy_, convEnergy_, convLength_, 1.0);
}

void Prediction::predict()
{
    evaluateNNP(structure);
    if (normalize_) structure.toPhysicalUnits(meanEnergy_, convEnergy_, convLength_, 1.0);
    addEnergyOffset(structure, false);
    addEnergyOffset(structure, true);
}

}  // namespace nnpx
