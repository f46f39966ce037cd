// Host-side structure container of the nnpx C++ façade: input.data parsing,
// periodic remap, unit conversion and output.data writing with the file
// formats and rules of n2p2's Structure (Structure.cpp:142-271, 518-544,
// 1224-1306, 1475-1532) and ElementMap (ElementMap.cpp:36-50).
// Built with -ffp-contract=off so remap and unit conversion round exactly as
// the reference does.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>

#include "nnpx_mode.hpp"

namespace nnpx {

namespace {

const char* const kSymbols[] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",
    "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge",
    "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd",
    "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd",
    "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg",
    "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U",  "Np", "Pu", "Am", "Cm",
    "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr", "Rf", "Db", "Sg", "Bh", "Hs", "Mt", "Ds", "Rg", "Cn",
    "Nh", "Fl", "Mc", "Lv", "Ts", "Og"};

std::vector<std::string> words(std::string const& line)
{
    std::istringstream is(line);
    std::vector<std::string> w;
    std::string t;
    while (is >> t) w.push_back(t);
    return w;
}

std::string fmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char* f, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof buf, f, ap);
    va_end(ap);
    return buf;
}

Vec3D add(Vec3D const& a, Vec3D const& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
Vec3D scale(double s, Vec3D const& a) { return {s * a[0], s * a[1], s * a[2]}; }

}  // namespace

size_t ElementMap::atomicNumber(std::string const& symbol)
{
    for (size_t z = 0; z < sizeof kSymbols / sizeof kSymbols[0]; ++z)
        if (symbol == kSymbols[z]) return z + 1;
    throw std::runtime_error("ERROR: Unknown element symbol \"" + symbol + "\".\n");
}

void ElementMap::registerElements(std::string const& symbols)
{
    std::vector<std::string> s = words(symbols);
    std::vector<std::pair<size_t, std::string>> zs;
    for (auto const& x : s) zs.emplace_back(atomicNumber(x), x);
    std::sort(zs.begin(), zs.end());
    sym_.clear();
    z_.clear();
    for (auto const& p : zs) {
        if (!z_.empty() && z_.back() == p.first) throw std::runtime_error("ERROR: Element added more than once.\n");
        z_.push_back(p.first);
        sym_.push_back(p.second);
    }
}

size_t ElementMap::operator[](std::string const& symbol) const
{
    for (size_t i = 0; i < sym_.size(); ++i)
        if (sym_[i] == symbol) return i;
    throw std::out_of_range("ERROR: Element \"" + symbol + "\" not in element map.\n");
}

void Structure::setElementMap(ElementMap const& map)
{
    elementMap = map;
    numElements = map.size();
    numAtomsPerElement.assign(numElements, 0);
}

void Structure::reset()
{
    ElementMap m = elementMap;
    *this = Structure();
    setElementMap(m);
}

void Structure::readFromFile(std::string const& fileName)
{
    std::ifstream file(fileName.c_str());
    if (!file.is_open()) throw std::runtime_error("ERROR: Could not open file: \"" + fileName + "\".\n");
    std::string line;
    std::vector<std::string> lines;
    if (!std::getline(file, line)) throw std::runtime_error("ERROR: Unexpected file content, expected \"begin\" keyword.\n");
    lines.push_back(line);
    std::vector<std::string> w = words(line);
    if (w.empty() || w[0] != "begin")
        throw std::runtime_error("ERROR: Unexpected file content, expected \"begin\" keyword.\n");
    while (std::getline(file, line)) {
        lines.push_back(line);
        w = words(line);
        if (!w.empty() && w[0] == "end") break;
    }
    readFromLines(lines);
}

void Structure::readFromLines(std::vector<std::string> const& lines)
{
    size_t iBox = 0;
    if (lines.empty() || words(lines[0]).empty() || words(lines[0])[0] != "begin")
        throw std::runtime_error("ERROR: Unexpected line content, expected \"begin\" keyword.\n");
    if (numAtomsPerElement.size() != numElements) numAtomsPerElement.assign(numElements, 0);
    for (auto const& line : lines) {
        std::vector<std::string> w = words(line);
        if (w.empty()) throw std::runtime_error("ERROR: Unexpected empty line in structure.\n");
        const std::string& k = w[0];
        if (k == "begin") {
            for (size_t i = 1; i < w.size(); ++i)
                if (w[i] != "set=train" && w[i] != "set=test")
                    throw std::runtime_error("ERROR: Unknown keyword in structure specification, check \"begin\" "
                                             "arguments.\n");
        } else if (k == "comment") {
            std::string t = line;
            t.erase(t.find("comment"), 8);
            comment = t;
        } else if (k == "lattice") {
            if (iBox > 2) throw std::runtime_error("ERROR: Too many box vectors.\n");
            if (w.size() < 4) throw std::runtime_error("ERROR: Bad lattice line.\n");
            for (int c = 0; c < 3; ++c) box[iBox][c] = std::atof(w[1 + c].c_str());
            ++iBox;
            if (iBox == 3) {
                isPeriodic = true;
                const double tiny = std::numeric_limits<double>::min();
                isTriclinic = box[0][1] > tiny || box[0][2] > tiny || box[1][0] > tiny || box[1][2] > tiny ||
                              box[2][0] > tiny || box[2][1] > tiny;
                calculateInverseBox();
                calculateVolume();
            }
        } else if (k == "atom") {
            if (w.size() < 10) throw std::runtime_error("ERROR: Bad atom line.\n");
            Atom a;
            a.index = numAtoms;
            a.indexStructure = index;
            a.tag = (int64_t)numAtoms;
            for (int c = 0; c < 3; ++c) a.r[c] = std::atof(w[1 + c].c_str());
            a.element = elementMap[w[4]];
            a.chargeRef = std::atof(w[5].c_str());
            for (int c = 0; c < 3; ++c) a.fRef[c] = std::atof(w[7 + c].c_str());
            atoms.push_back(a);
            ++numAtoms;
            ++numAtomsPerElement.at(a.element);
        } else if (k == "energy") {
            energyRef = std::atof(w.at(1).c_str());
        } else if (k == "charge") {
            chargeRef = std::atof(w.at(1).c_str());
        } else if (k == "end") {
            if (!(iBox == 0 || iBox == 3)) throw std::runtime_error("ERROR: Strange number of box vectors.\n");
            if (isPeriodic && std::abs(chargeRef) > 10 * std::numeric_limits<double>::epsilon())
                throw std::runtime_error("ERROR: In structure with index " + std::to_string(index) +
                                         "; if PBCs are applied, the\nsimulation cell has to be neutrally charged.\n");
            break;
        } else {
            throw std::runtime_error("ERROR: Unexpected file content, unknown keyword \"" + k + "\".\n");
        }
    }
    numElementsPresent = 0;
    for (size_t e = 0; e < numElements; ++e)
        if (numAtomsPerElement[e] > 0) ++numElementsPresent;
    if (isPeriodic) remap();
}

void Structure::calculateInverseBox()
{
    double invdet = box[0][0] * box[1][1] * box[2][2] + box[1][0] * box[2][1] * box[0][2] +
                    box[2][0] * box[0][1] * box[1][2] - box[2][0] * box[1][1] * box[0][2] -
                    box[0][0] * box[2][1] * box[1][2] - box[1][0] * box[0][1] * box[2][2];
    invdet = 1.0 / invdet;
    invbox[0] = {box[1][1] * box[2][2] - box[2][1] * box[1][2], box[2][1] * box[0][2] - box[0][1] * box[2][2],
                 box[0][1] * box[1][2] - box[1][1] * box[0][2]};
    invbox[1] = {box[2][0] * box[1][2] - box[1][0] * box[2][2], box[0][0] * box[2][2] - box[2][0] * box[0][2],
                 box[1][0] * box[0][2] - box[0][0] * box[1][2]};
    invbox[2] = {box[1][0] * box[2][1] - box[2][0] * box[1][1], box[2][0] * box[0][1] - box[0][0] * box[2][1],
                 box[0][0] * box[1][1] - box[1][0] * box[0][1]};
    for (auto& v : invbox) v *= invdet;
}

void Structure::calculateVolume()
{
    // |a . (b x c)|
    const Vec3D& a = box[0];
    const Vec3D& b = box[1];
    const Vec3D& c = box[2];
    Vec3D bxc{b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    volume = std::fabs(a[0] * bxc[0] + a[1] * bxc[1] + a[2] * bxc[2]);
}

void Structure::remap()
{
    for (auto& at : atoms) remap(at);
}

void Structure::remap(Atom& at)
{
    Vec3D f = add(add(scale(at.r[0], invbox[0]), scale(at.r[1], invbox[1])), scale(at.r[2], invbox[2]));
    for (int c = 0; c < 3; ++c) {
        if (f[c] > 1.0) f[c] -= (int)f[c];
        if (f[c] < 0.0) f[c] += 1.0 - (int)f[c];
        if (f[c] == 1.0) f[c] = 0.0;
    }
    at.r = add(add(scale(f[0], box[0]), scale(f[1], box[1])), scale(f[2], box[2]));
}

void Structure::calculateNeighborList(double cutoffRadius, bool)
{
    maxCutoffRadius = cutoffRadius;
    hasNeighborList = true;
}

void Structure::toNormalizedUnits(double meanEnergy, double convEnergy, double convLength, double convCharge)
{
    if (isPeriodic) {
        for (int i = 0; i < 3; ++i) {
            box[i] *= convLength;
            invbox[i] /= convLength;
        }
    }
    energyRef = (energyRef - numAtoms * meanEnergy) * convEnergy;
    energy = (energy - numAtoms * meanEnergy) * convEnergy;
    chargeRef *= convCharge;
    charge *= convCharge;
    volume *= convLength * convLength * convLength;
    for (auto& a : atoms) {
        a.energy *= convEnergy;
        a.r *= convLength;
        a.f *= convEnergy / convLength;
        a.fRef *= convEnergy / convLength;
        a.charge *= convCharge;
        a.chargeRef *= convCharge;
        if (a.hasSymmetryFunctionDerivatives) {
            for (auto& x : a.dEdG) x *= convEnergy;
            for (auto& v : a.dGdr) v /= convLength;
        }
        for (auto& n : a.neighbors) {
            n.d *= convLength;
            n.dr *= convLength;
            for (auto& v : n.dGdr) v /= convLength;
        }
    }
}

void Structure::toPhysicalUnits(double meanEnergy, double convEnergy, double convLength, double convCharge)
{
    if (isPeriodic) {
        for (int i = 0; i < 3; ++i) {
            box[i] /= convLength;
            invbox[i] *= convLength;
        }
    }
    energyRef = energyRef / convEnergy + numAtoms * meanEnergy;
    energy = energy / convEnergy + numAtoms * meanEnergy;
    chargeRef /= convCharge;
    charge /= convCharge;
    volume /= convLength * convLength * convLength;
    for (auto& a : atoms) {
        a.energy /= convEnergy;
        a.r /= convLength;
        a.f /= convEnergy / convLength;
        a.fRef /= convEnergy / convLength;
        a.charge /= convCharge;
        a.chargeRef /= convCharge;
        if (a.hasSymmetryFunctionDerivatives) {
            for (auto& x : a.dEdG) x /= convEnergy;
            for (auto& v : a.dGdr) v *= convLength;
        }
        for (auto& n : a.neighbors) {
            n.d /= convLength;
            n.dr /= convLength;
            for (auto& v : n.dGdr) v *= convLength;
        }
    }
}

void Structure::writeToFile(std::ofstream* const& file, bool const ref) const
{
    if (!file->is_open()) throw std::runtime_error("ERROR: Cannot write to file, file is not open.\n");
    std::ofstream& o = *file;
    o << "begin\n";
    o << fmt("comment %s\n", comment.c_str());
    if (isPeriodic)
        for (int i = 0; i < 3; ++i) o << fmt("lattice %24.16E %24.16E %24.16E\n", box[i][0], box[i][1], box[i][2]);
    for (auto const& a : atoms) {
        const Vec3D& f = ref ? a.fRef : a.f;
        o << fmt("atom %24.16E %24.16E %24.16E %2s %24.16E %24.16E %24.16E %24.16E %24.16E\n", a.r[0], a.r[1],
                 a.r[2], elementMap[a.element].c_str(), ref ? a.chargeRef : a.charge, 0.0, f[0], f[1], f[2]);
    }
    o << fmt("energy %24.16E\n", ref ? energyRef : energy);
    o << fmt("charge %24.16E\n", ref ? chargeRef : charge);
    o << "end\n";
}

void Structure::writeToFile(std::string const fileName, bool const ref, bool const append) const
{
    std::ofstream file(fileName.c_str(), append ? std::ofstream::app : std::ofstream::out);
    if (!file.is_open()) throw std::runtime_error("ERROR: Could not open file: \"" + fileName + "\".\n");
    std::ofstream* p = &file;
    writeToFile(p, ref);
}

}  // namespace nnpx
