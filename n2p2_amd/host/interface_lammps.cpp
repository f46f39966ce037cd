// nnpx::InterfaceLammps: n2p2's LAMMPS interface (src/libnnpif/LAMMPS/
// InterfaceLammps.cpp) over the nnpx C ABI's caller-provided neighbour lists.
#include "nnpx_lammps.hpp"

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace nnpx {

namespace {

std::string trim_spaces(std::string s)
{
    std::string out;
    for (char c : s)
        if (c != ' ' && c != '\t') out += c;
    return out;
}

std::vector<std::string> split_on(std::string const& s, char sep)
{
    std::vector<std::string> v;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, sep)) v.push_back(item);
    return v;
}

}  // namespace

InterfaceLammps::InterfaceLammps(int device) : Mode(device) {}

void InterfaceLammps::initialize(char const* const& directory, char const* const& emap, bool showew, bool resetew,
                                 int showewsum, int maxew, double cflength, double cfenergy, double lammpsCutoff,
                                 int lammpsNtypes, int myRank)
{
    emap_ = emap ? emap : "";
    showew_ = showew;
    resetew_ = resetew;
    showewsum_ = showewsum;
    maxew_ = maxew;
    cflength_ = cflength;
    cfenergy_ = cfenergy;
    myRank_ = myRank;
    std::string dir(directory);
    if (dir.empty() || dir.back() != '/') dir += '/';
    setupFromDirectory(dir);
    // Warnings are collected for the pair style's summary and limit
    // (InterfaceLammps.cpp:76-90).
    setupSymmetryFunctionStatistics(false, showew || showewsum > 0 || maxew >= 0, false, false);
    const double sfCutoff = getMaxCutoffRadius();
    if (lammpsCutoff < sfCutoff)
        throw std::runtime_error("ERROR: LAMMPS cutoff via pair_coeff keyword is smaller than maximum symmetry "
                                 "function cutoff.\n");
    mapTypeToElement_.clear();
    mapElementToType_.clear();
    ignoreType_.clear();
    if (emap_.empty()) {
        if (elementMap.size() != (size_t)lammpsNtypes)
            throw std::runtime_error(strpr("ERROR: No element mapping given and number of LAMMPS atom types (%d) "
                                           "and NNP elements (%zu) does not match.\n",
                                           lammpsNtypes, elementMap.size()));
        for (int i = 0; i < lammpsNtypes; ++i) {
            mapTypeToElement_[i + 1] = i;
            mapElementToType_[i] = i + 1;
        }
    } else {
        const std::vector<std::string> parts = split_on(trim_spaces(emap_), ',');
        if (elementMap.size() < parts.size())
            throw std::runtime_error(strpr("ERROR: Element mapping is inconsistent, NNP elements: %zu, emap "
                                           "elements: %zu.\n", elementMap.size(), parts.size()));
        for (std::string const& s : parts) {
            const std::vector<std::string> ts = split_on(s, ':');
            if (ts.size() != 2)
                throw std::runtime_error(strpr("ERROR: Invalid element mapping string: \"%s\".\n", s.c_str()));
            const int t = std::stoi(ts[0]);
            if (t > lammpsNtypes)
                throw std::runtime_error(strpr("ERROR: LAMMPS type \"%d\" not present, there are only %d types "
                                               "defined.\n", t, lammpsNtypes));
            size_t e = elementMap.size();
            for (size_t q = 0; q < elementMap.size(); ++q)
                if (elementMap[q] == ts[1]) e = q;
            if (e == elementMap.size()) throw std::out_of_range("ERROR: element \"" + ts[1] + "\" not in the model.\n");
            mapTypeToElement_[t] = e;
            mapElementToType_[e] = t;
        }
    }
    for (int i = 1; i <= lammpsNtypes; ++i) ignoreType_[i] = mapTypeToElement_.find(i) == mapTypeToElement_.end();
    structure.setElementMap(elementMap);
    initialized_ = true;
}

void InterfaceLammps::setLocalAtoms(int numAtomsLocal, int const* const atomType)
{
    for (size_t e = 0; e < structure.numAtomsPerElement.size(); ++e) structure.numAtomsPerElement[e] = 0;
    structure.numAtomsPerElement.resize(elementMap.size(), 0);
    structure.index = myRank_;
    structure.numAtoms = 0;
    structure.hasNeighborList = false;
    structure.hasSymmetryFunctions = false;
    structure.hasSymmetryFunctionDerivatives = false;
    structure.energy = 0.0;
    structure.atoms.clear();
    indexMap_.assign(numAtomsLocal, std::numeric_limits<size_t>::max());
    structure.atoms.reserve(numAtomsLocal);
    for (int i = 0; i < numAtomsLocal; ++i) {
        if (ignoreType_[atomType[i]]) continue;
        indexMap_[i] = structure.numAtoms++;
        structure.atoms.push_back(Atom());
        Atom& a = structure.atoms.back();
        a.index = i;
        a.indexStructure = myRank_;
        a.tag = i;
        a.element = mapTypeToElement_[atomType[i]];
        structure.numAtomsPerElement[a.element]++;
    }
}

void InterfaceLammps::setLocalAtomPositions(double const* const* const atomPos)
{
    for (Atom& a : structure.atoms)
        for (int c = 0; c < 3; ++c) {
            a.r[c] = atomPos[a.index][c] * cflength_;
            if (normalize_) a.r[c] *= convLength_;
        }
}

void InterfaceLammps::setLocalTags(int const* const atomTag)
{
    for (Atom& a : structure.atoms) a.tag = atomTag[a.index];
}

void InterfaceLammps::setLocalTags(int64_t const* const atomTag)
{
    for (Atom& a : structure.atoms) a.tag = atomTag[a.index];
}

void InterfaceLammps::setBoxVectors(double const* boxlo, double const* boxhi, double const xy, double const xz,
                                    double const yz)
{
    structure.isPeriodic = true;
    structure.box[0] = Vec3D(boxhi[0] - boxlo[0], 0.0, 0.0);
    structure.box[1] = Vec3D(xy, boxhi[1] - boxlo[1], 0.0);
    structure.box[2] = Vec3D(xz, yz, boxhi[2] - boxlo[2]);
    structure.isTriclinic = xy != 0.0 || xz != 0.0 || yz != 0.0;
    for (int i = 0; i < 3; ++i) {
        structure.box[i] *= cflength_;
        if (normalize_) structure.box[i] *= convLength_;
    }
    structure.calculateInverseBox();
    structure.calculateVolume();
}

void InterfaceLammps::allocateNeighborlists(int const* const numneigh)
{
    for (Atom& a : structure.atoms) a.neighbors.reserve(numneigh[a.index]);
}

void InterfaceLammps::addNeighbor(int i, int j, int64_t tag, int type, double dx, double dy, double dz, double d2)
{
    if (ignoreType_[type] || indexMap_.at(i) == std::numeric_limits<size_t>::max()) return;
    Atom& a = structure.atoms[indexMap_[i]];
    a.numNeighbors++;
    a.neighbors.push_back(Atom::Neighbor());
    Atom::Neighbor& n = a.neighbors.back();
    n.index = j;
    n.tag = tag;
    n.element = mapTypeToElement_[type];
    // the reference's order: times cflength, then times conv_length
    n.dr = Vec3D(dx * cflength_, dy * cflength_, dz * cflength_);
    n.d = std::sqrt(d2) * cflength_;
    if (normalize_) {
        n.dr *= convLength_;
        n.d *= convLength_;
    }
}

void InterfaceLammps::process()
{
    nnpx_ctx* ctx = context();
    const size_t n = structure.atoms.size();
    // C-ABI atom indices: the structure's atoms first, then every other
    // LAMMPS index a list refers to (ghosts)
    lammpsIndex_.assign(n, 0);
    std::map<int64_t, int> other;
    for (size_t k = 0; k < n; ++k) lammpsIndex_[k] = (int64_t)structure.atoms[k].index;
    std::vector<int> elem(n);
    std::vector<int64_t> off(n + 1, 0);
    for (size_t k = 0; k < n; ++k) {
        elem[k] = (int)structure.atoms[k].element;
        off[k + 1] = off[k] + (int64_t)structure.atoms[k].neighbors.size();
    }
    std::vector<int> idx(off[n]), nel(off[n]);
    std::vector<double> dr(3 * (size_t)off[n]);
    for (size_t k = 0; k < n; ++k) {
        const Atom& a = structure.atoms[k];
        for (size_t q = 0; q < a.neighbors.size(); ++q) {
            const Atom::Neighbor& nb = a.neighbors[q];
            const int64_t p = off[k] + (int64_t)q;
            const int64_t j = (int64_t)nb.index;
            int abi;
            if (j >= 0 && (size_t)j < indexMap_.size() && indexMap_[j] != std::numeric_limits<size_t>::max()) {
                abi = (int)indexMap_[j];
            } else {
                auto it = other.find(j);
                if (it == other.end()) {
                    abi = (int)lammpsIndex_.size();
                    other[j] = abi;
                    lammpsIndex_.push_back(j);
                } else {
                    abi = it->second;
                }
            }
            idx[p] = abi;
            nel[p] = (int)nb.element;
            for (int c = 0; c < 3; ++c) dr[3 * p + c] = nb.dr[c];
        }
    }
    const int ntotal = (int)lammpsIndex_.size();
    if (n == 0) {
        structure.energy = 0.0;
        force_.clear();
        return;
    }
    check(nnpx_set_neighbors_external(ctx, (int)n, elem.data(), ntotal, off.data(), idx.data(), nel.data(),
                                      dr.data()));
    check(nnpx_evaluate(ctx, (int)getPrecision(), NNPX_EVAL_FORCES));
    double e = 0.0;
    check(nnpx_get_energy(ctx, &e));
    std::vector<double> ea(n);
    check(nnpx_get_atomic_energies(ctx, ea.data()));
    for (size_t k = 0; k < n; ++k) structure.atoms[k].energy = ea[k];
    force_.assign(3 * (size_t)ntotal, 0.0);
    check(nnpx_get_forces(ctx, force_.data()));
    structure.energyShort = e;
    structure.energy = e;
    structure.hasSymmetryFunctions = true;
    structure.hasSymmetryFunctionDerivatives = true;
    // Mode::physicalEnergy + addEnergyOffset (InterfaceLammps.cpp:471-478)
    if (normalize_) structure.energy = structure.energy / convEnergy_ + (double)n * meanEnergy_;
    addEnergyOffset(structure, false);
    // extrapolation warnings of this step (Element.cpp:462-541 with
    // collectExtrapolationWarnings; resetew clears them every step)
    if (resetew_) ew_.clear();
    if (showew_ || showewsum_ > 0 || maxew_ >= 0) {
        int64_t cnt = 0;
        check(nnpx_extrapolation_warnings(ctx, 0, &cnt, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
        if (cnt > 0) {
            std::vector<int> a(cnt), q(cnt), t(cnt);
            std::vector<double> v(cnt), lo(cnt), hi(cnt);
            check(nnpx_extrapolation_warnings(ctx, cnt, &cnt, a.data(), q.data(), t.data(), v.data(), lo.data(),
                                              hi.data()));
            for (int64_t r = 0; r < cnt; ++r) {
                const Atom& at = structure.atoms.at(a[r]);
                ew_.push_back(EW{(int)at.element, q[r], t[r], at.tag, v[r], lo[r], hi[r]});
            }
        }
    }
}

double InterfaceLammps::getEnergy() const { return structure.energy / cfenergy_; }

double InterfaceLammps::getAtomicEnergy(int index) const
{
    const Atom& a = structure.atoms.at(index);
    const double off = atomicEnergyOffset_.at(a.element);
    if (normalize_) return (a.energy / convEnergy_ + meanEnergy_ + off) / cfenergy_;
    return (a.energy + off) / cfenergy_;
}

void InterfaceLammps::getForces(double* const* const& atomF) const
{
    const double cfforce = cflength_ / cfenergy_;
    const double convForce = normalize_ ? convLength_ / convEnergy_ : 1.0;
    for (size_t k = 0; k < lammpsIndex_.size(); ++k) {
        double* f = atomF[lammpsIndex_[k]];
        for (int c = 0; c < 3; ++c) f[c] += force_[3 * k + c] * cfforce * convForce;
    }
}

double InterfaceLammps::getMaxCutoffRadius() const
{
    if (normalize_) return maxCutoffRadius_ / convLength_ / cflength_;
    return maxCutoffRadius_ / cflength_;
}

void InterfaceLammps::writeExtrapolationWarnings()
{
    std::vector<EW> v = ew_;
    std::stable_sort(v.begin(), v.end(), [](EW const& x, EW const& y) {
        return x.element != y.element ? x.element < y.element : x.sf < y.sf;
    });
    std::ostream* log = logStream();
    if (!log) return;
    for (EW const& w : v)
        *log << strpr("### NNP EXTRAPOLATION WARNING ### STRUCTURE: %6zu ATOM: %9zu ELEMENT: %2s SYMFUNC: %4zu TYPE: "
                      "%2zu VALUE: %10.3E MIN: %10.3E MAX: %10.3E\n",
                      (size_t)myRank_, (size_t)w.tag, elementMap[w.element].c_str(), (size_t)w.sf + 1, (size_t)w.type,
                      w.value, w.gmin, w.gmax);
}

void InterfaceLammps::clearExtrapolationWarnings() { ew_.clear(); }

void InterfaceLammps::writeToFile(std::string const fileName, bool const append)
{
    structure.toPhysicalUnits(meanEnergy_, convEnergy_, convLength_, 1.0);
    structure.writeToFile(fileName, false, append);
    structure.toNormalizedUnits(meanEnergy_, convEnergy_, convLength_, 1.0);
}

}  // namespace nnpx
