/*
 * nnpx::InterfaceLammps -- the interface LAMMPS' pair_style hdnnp drives
 * (n2p2 src/libnnpif/LAMMPS/InterfaceLammps.h/.cpp), over the nnpx C ABI:
 * the local atoms and LAMMPS' full neighbour list (local and ghost atoms,
 * distance vectors) are handed to nnpx_set_neighbors_external and the push
 * force form scatters onto locals and ghosts, as InterfaceLammps::getForces
 * does (InterfaceLammps.cpp:679-740).  2G-HDNNP only: the 4G methods
 * (charges, QEq, electrostatics) are not provided.
 *
 * Units: LAMMPS lengths times cflength (and conv_length when input.nn
 * normalises) are model lengths; energies returned are the model's physical
 * energies (offsets included) divided by cfenergy; forces are multiplied by
 * cflength / cfenergy (and conv_length / conv_energy when normalised),
 * exactly the reference's conversions.  Without MPI (one rank), the
 * extrapolation-warning buffers are empty as in an N2P2_NO_MPI build.
 */
#ifndef NNPX_LAMMPS_HPP
#define NNPX_LAMMPS_HPP

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "nnpx_mode.hpp"

namespace nnpx {

class InterfaceLammps : public Mode {
public:
    explicit InterfaceLammps(int device = 0);

    /* InterfaceLammps::initialize (InterfaceLammps.cpp:51-256): model from
     * directory, element mapping "type:Symbol,..." (empty: types 1..n map to
     * the elements in order), extrapolation-warning handling, unit factors,
     * cutoff consistency check against LAMMPS' pair cutoff. */
    void initialize(char const* const& directory, char const* const& emap, bool showew, bool resetew,
                    int showewsum, int maxew, double cflength, double cfenergy, double lammpsCutoff,
                    int lammpsNtypes, int myRank);
    void setGlobalStructureStatus(bool const status) { hasGlobalStructure_ = status; }
    bool getGlobalStructureStatus() { return hasGlobalStructure_; }
    /* Local atoms by LAMMPS type (1-based); atoms of unmapped types are
     * ignored (InterfaceLammps.cpp:267-302). */
    void setLocalAtoms(int numAtomsLocal, int const* const atomType);
    void setLocalAtomPositions(double const* const* const atomPos);
    void setLocalTags(int const* const atomTag);
    void setLocalTags(int64_t const* const atomTag);
    void setBoxVectors(double const* boxlo, double const* boxhi, double const xy, double const xz, double const yz);
    void allocateNeighborlists(int const* const numneigh);
    /* One entry of atom i's list: LAMMPS index j (local or ghost), its tag,
     * type and r_i - r_j with squared length d2 (InterfaceLammps.cpp:410-441);
     * entries of ignored types or of ignored atoms are dropped. */
    void addNeighbor(int i, int j, int64_t tag, int type, double dx, double dy, double dz, double d2);
    void finalizeNeighborList() {}
    /* SF groups, NNs and energy of the current atoms and lists
     * (InterfaceLammps::process, InterfaceLammps.cpp:463-500). */
    void process();
    double getEnergy() const;
    /* Atomic energy of structure atom `index` (not the LAMMPS index). */
    double getAtomicEnergy(int index) const;
    /* Adds the forces onto LAMMPS' force array (local + ghost rows). */
    void getForces(double* const* const& atomF) const;
    bool isInitialized() const { return initialized_; }
    double getMaxCutoffRadius() const;
    double getMaxCutoffRadiusOverall() { return getMaxCutoffRadius(); }
    long getEWBufferSize() const { return 0; }
    void fillEWBuffer(char* const&, int) const {}
    void extractEWBuffer(char const* const&, int) {}
    /* The collected extrapolation warnings to the log, in the order of
     * SymFncStatistics::getExtrapolationWarningLines (element, SF, occurrence),
     * and their count; clear empties them. */
    void writeExtrapolationWarnings();
    void clearExtrapolationWarnings();
    size_t getNumExtrapolationWarnings() const { return ew_.size(); }
    /* The structure in input.data format, physical units. */
    void writeToFile(std::string const fileName, bool const append);

    Structure structure;

private:
    struct EW {
        int element, sf, type;
        int64_t tag;
        double value, gmin, gmax;
    };
    std::string emap_;
    bool showew_ = true, resetew_ = false;
    int showewsum_ = 0, maxew_ = 0, myRank_ = 0;
    double cflength_ = 1.0, cfenergy_ = 1.0;
    bool initialized_ = false, hasGlobalStructure_ = false;
    std::map<int, size_t> mapTypeToElement_;
    std::map<size_t, int> mapElementToType_;
    std::map<int, bool> ignoreType_;
    std::vector<size_t> indexMap_;  // LAMMPS local index -> structure atom
    std::vector<int64_t> lammpsIndex_;  // C-ABI atom index -> LAMMPS index (locals, then others)
    std::vector<double> force_;     // forces of the last process, C-ABI order, model units
    std::vector<EW> ew_;
};

}  // namespace nnpx

#endif /* NNPX_LAMMPS_HPP */
