/*
 * C++ host façade over the nnpx C ABI: the nnp::Mode / nnp::Prediction /
 * nnp::Structure surface that n2p2's callers use for the energy/force path,
 * with the same method names, argument meaning, result fields and exception
 * types, backed by one GPU context.
 *
 * Mirrors (citations relative to the n2p2 source tree):
 *   Mode::calculateSymmetryFunctionGroups   Mode.h:313-315, Mode.cpp:1630-1709
 *   Mode::calculateAtomicNeuralNetworks     Mode.h:341-344, Mode.cpp:1711-1732
 *   Mode::calculateEnergy                   Mode.h:362,     Mode.cpp:1876-1901
 *   Mode::calculateForces                   Mode.h:378,     Mode.cpp:1922-1984
 *   Mode::evaluateNNP                       Mode.h:385-387, Mode.cpp:2029-2077
 *   Mode::add/removeEnergyOffset, getEnergyOffset   Mode.cpp:2079-2130
 *   Prediction::setup/readStructureFromFile/predict Prediction.cpp:36-78
 *   Structure::readFromFile/readFromLines/remap/writeToFile/unit conversion
 *                                           Structure.cpp:142-271, 1224-1306, 1475-1532
 *
 * Differences a caller can observe:
 *   - the neighbour list is built on the device inside
 *     calculateSymmetryFunctionGroups; Structure::calculateNeighborList only
 *     records the cutoff.  Atom::numNeighbors is filled; Atom::neighbors is
 *     filled only with setCopyNeighbors(true);
 *   - Atom::dGdr and Atom::Neighbor::dGdr are filled only in the plain form
 *     with setCopyDerivatives(true) (the mixed form never stores dG/dr);
 *   - in the mixed form calculateSymmetryFunctionGroups runs the whole fused
 *     evaluation; the later calls copy its results into the structure.
 * One Mode drives one GPU from one host thread.
 */
#ifndef NNPX_MODE_HPP
#define NNPX_MODE_HPP

#include <cstddef>
#include <cstdint>
#include <fstream>
#include <ostream>
#include <string>
#include <vector>

#include "nnpx.h"

namespace nnpx {

struct Vec3D {
    double r[3] = {0.0, 0.0, 0.0};
    Vec3D() = default;
    Vec3D(double x, double y, double z) : r{x, y, z} {}
    double& operator[](size_t i) { return r[i]; }
    double operator[](size_t i) const { return r[i]; }
    Vec3D& operator*=(double a) { r[0] *= a; r[1] *= a; r[2] *= a; return *this; }
    Vec3D& operator/=(double a) { r[0] /= a; r[1] /= a; r[2] /= a; return *this; }
    double norm2() const { return r[0] * r[0] + r[1] * r[1] + r[2] * r[2]; }
};

/* Element symbols sorted by atomic number (ElementMap.cpp:36-50). */
class ElementMap {
public:
    void registerElements(std::string const& symbols);
    size_t size() const { return sym_.size(); }
    std::string operator[](size_t index) const { return sym_.at(index); }
    size_t operator[](std::string const& symbol) const;
    size_t atomicNumber(size_t index) const { return z_.at(index); }
    static size_t atomicNumber(std::string const& symbol);

private:
    std::vector<std::string> sym_;
    std::vector<size_t> z_;
};

struct Atom {
    struct Neighbor {
        size_t index = 0;
        int64_t tag = 0;
        size_t element = 0;
        double d = 0.0;
        Vec3D dr;
        std::vector<Vec3D> dGdr;
    };
    bool hasNeighborList = false;
    bool hasSymmetryFunctions = false;
    bool hasSymmetryFunctionDerivatives = false;
    size_t index = 0;
    size_t indexStructure = 0;
    int64_t tag = 0;
    size_t element = 0;
    size_t numNeighbors = 0;
    size_t numSymmetryFunctions = 0;
    double energy = 0.0;
    double charge = 0.0;
    double chargeRef = 0.0;
    Vec3D r, f, fRef;
    std::vector<double> G;
    std::vector<double> dEdG;
    std::vector<Vec3D> dGdr;
    std::vector<Neighbor> neighbors;
};

class Structure {
public:
    ElementMap elementMap;
    bool isPeriodic = false;
    bool isTriclinic = false;
    bool hasNeighborList = false;
    bool hasSymmetryFunctions = false;
    bool hasSymmetryFunctionDerivatives = false;
    size_t index = 0;
    size_t numAtoms = 0;
    size_t numElements = 0;
    size_t numElementsPresent = 0;
    double energy = 0.0;
    double energyRef = 0.0;
    double energyShort = 0.0;
    double energyElec = 0.0;
    double charge = 0.0;
    double chargeRef = 0.0;
    double volume = 0.0;
    double maxCutoffRadius = 0.0;
    Vec3D box[3];
    Vec3D invbox[3];
    std::string comment;
    std::vector<size_t> numAtomsPerElement;
    std::vector<Atom> atoms;

    void setElementMap(ElementMap const& map);
    void reset();
    /* First configuration of an input.data file (Structure::readFromFile). */
    void readFromFile(std::string const& fileName = "input.data");
    void readFromLines(std::vector<std::string> const& lines);
    void remap();
    /* One atom (Structure::remap(Atom&), Structure.cpp:1224-1248). */
    void remap(Atom& atom);
    void calculateInverseBox();
    void calculateVolume();
    /* Records the cutoff; the list itself is built on the device. */
    void calculateNeighborList(double cutoffRadius, bool sortByDistance = false);
    void toNormalizedUnits(double meanEnergy, double convEnergy, double convLength, double convCharge);
    void toPhysicalUnits(double meanEnergy, double convEnergy, double convLength, double convCharge);
    void writeToFile(std::ofstream* const& file, bool const ref = true) const;
    void writeToFile(std::string const fileName = "output.data", bool const ref = true,
                     bool const append = false) const;
};

class Mode {
public:
    enum class Precision { PLAIN_FP64 = NNPX_PLAIN_FP64, MIXED = NNPX_MIXED };
    enum class ForceMode { PULL = NNPX_FORCES_PULL, PUSH = NNPX_FORCES_PUSH };

    explicit Mode(int device = 0);
    virtual ~Mode();
    Mode(Mode const&) = delete;
    Mode& operator=(Mode const&) = delete;

    /* input.nn, scaling.data and weights.%03d.data from one directory
     * (Prediction::setup, Prediction.cpp:36-48). */
    void setupFromDirectory(std::string const& directory);
    void setPrecision(Precision p) { precision_ = p; }
    Precision getPrecision() const { return precision_; }
    void setForceMode(ForceMode m);
    void setCopyNeighbors(bool on) { copyNeighbors_ = on; }
    void setCopyDerivatives(bool on) { copyDerivatives_ = on; }
    /* Mode::setupSymmetryFunctionStatistics (Mode.cpp:1172-1205): after each
     * calculateSymmetryFunctionGroups, extrapolation warnings
     * (Element::updateSymmetryFunctionStatistics, Element.cpp:462-541) are
     * collected, written to stderr in the reference's format, or raised as
     * std::out_of_range.  Collecting min/max/mean statistics is not provided
     * (collectStatistics must be false). */
    void setupSymmetryFunctionStatistics(bool collectStatistics, bool collectExtrapolationWarnings,
                                         bool writeExtrapolationWarnings, bool stopOnExtrapolationWarnings);
    /* Collected extrapolation warnings: count, the reference's summary lines
     * (SymFncStatistics::getExtrapolationWarningLines format, one per
     * warning) and reset (SymFncStatistics::resetExtrapolationWarnings). */
    size_t getNumExtrapolationWarnings() const { return ewLines_.size(); }
    std::vector<std::string> const& getExtrapolationWarningLines() const { return ewLines_; }
    void resetExtrapolationWarnings() { ewLines_.clear(); }
    /* Where log lines go (the low-neighbour warnings of
     * calculateSymmetryFunctionGroups, Mode.cpp:1667-1679); none by default. */
    void setLogStream(std::ostream* log) { log_ = log; }

    void calculateSymmetryFunctionGroups(Structure& structure, bool const derivatives);
    void calculateAtomicNeuralNetworks(Structure& structure, bool const derivatives, std::string id = "");
    void calculateEnergy(Structure& structure) const;
    void calculateForces(Structure& structure) const;
    void evaluateNNP(Structure& structure, bool useForces = true, bool useDEdG = true);

    void addEnergyOffset(Structure& structure, bool ref = true);
    void removeEnergyOffset(Structure& structure, bool ref = true);
    double getEnergyOffset(Structure const& structure) const;

    size_t getNumElements() const { return elementMap.size(); }
    double getMaxCutoffRadius() const { return maxCutoffRadius_; }
    bool useNormalization() const { return normalize_; }
    double getMeanEnergy() const { return meanEnergy_; }
    double getConvEnergy() const { return convEnergy_; }
    double getConvLength() const { return convLength_; }
    nnpx_ctx* context() const { return ctx_; }

    ElementMap elementMap;

protected:
    bool normalize_ = false;
    double meanEnergy_ = 0.0;
    double convEnergy_ = 1.0;
    double convLength_ = 1.0;
    double maxCutoffRadius_ = 0.0;
    std::vector<double> atomicEnergyOffset_;
    /* Raises the exception n2p2 throws for a non-OK status. */
    void check(nnpx_status st) const;
    std::ostream* logStream() const { return log_; }

private:
    void upload(Structure& s);
    void requireResident(Structure const& s) const;

    nnpx_ctx* ctx_ = nullptr;
    Precision precision_ = Precision::PLAIN_FP64;
    bool copyNeighbors_ = false;
    bool copyDerivatives_ = false;
    int maxSF_ = 0;
    Structure const* resident_ = nullptr;
    size_t residentAtoms_ = 0;
    bool mixedForces_ = false;
    bool collectEW_ = false, writeEW_ = false, stopEW_ = false;
    std::vector<std::string> ewLines_;
    std::ostream* log_ = nullptr;
    void checkStructure(Structure const& s);
};

class Prediction : public Mode {
public:
    explicit Prediction(int device = 0) : Mode(device) {}
    void setup();
    void readStructureFromFile(std::string const& fileName = "input.data");
    void predict();

    std::string directory = ".";
    Structure structure;
};

/* Column header block of n2p2's output files (createFileHeader +
 * appendLinesToFile, utility.cpp:111-204): title lines, column table, column
 * numbers and names right-aligned to the given widths. */
void writeFileHeader(std::ofstream& o, std::vector<std::string> const& title, std::vector<size_t> const& size,
                     std::vector<std::string> const& name, std::vector<std::string> const& info);

/* printf into a std::string (n2p2's strpr, utility.cpp:247-264). */
std::string strpr(const char* format, ...) __attribute__((format(printf, 1, 2)));

/* All configurations of an input.data file as line blocks begin..end (the
 * reading loop of Dataset::loadStructuresFromFile); throws when the file
 * cannot be opened. */
std::vector<std::vector<std::string>> readStructureBlocks(std::string const& fileName);

/* Keyword present in a settings file (input.nn), comments after '#'
 * ignored (Settings::keywordExists on the file's contents). */
bool settingsFileHasKeyword(std::string const& fileName, std::string const& keyword);

/* Positions (scaled by a length factor), element indices, boxes and
 * periodicity of a list of structures, in the packed layout of
 * nnpx_set_structures. */
struct PackedStructures {
    std::vector<int64_t> offset{0};
    std::vector<double> position, box;
    std::vector<int> element, periodic;
    void add(Structure const& s, double convLength);
    int count() const { return (int)periodic.size(); }
    void clear();
};

}  // namespace nnpx

#endif /* NNPX_MODE_HPP */
