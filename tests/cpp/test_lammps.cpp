// Drives nnpx::InterfaceLammps the way LAMMPS' pair_style hdnnp does
// (pair_hdnnp.cpp compute(): setLocalAtoms, setLocalTags, setBoxVectors,
// allocateNeighborlists, addNeighbor for every entry of the full list,
// finalizeNeighborList, process, getEnergy / getAtomicEnergy, getForces into
// the local+ghost force array).  Input and output are binary files written
// and read by tests/test_gpu_lammps.py:
//   test_lammps <model_dir> <in.bin> <out.bin> [mixed]
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "nnpx_lammps.hpp"

namespace {

struct Reader {
    std::FILE* f;
    template <typename T>
    T get()
    {
        T v;
        if (std::fread(&v, sizeof v, 1, f) != 1) throw std::runtime_error("short input");
        return v;
    }
    template <typename T>
    std::vector<T> vec(size_t n)
    {
        std::vector<T> v(n);
        if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short input");
        return v;
    }
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: test_lammps <model_dir> <in.bin> <out.bin> [mixed]\n");
        return 1;
    }
    try {
        Reader r{std::fopen(argv[2], "rb")};
        if (!r.f) throw std::runtime_error("cannot open input");
        const int64_t nlocal = r.get<int64_t>(), nrows = r.get<int64_t>(), ntypes = r.get<int64_t>();
        const double cflength = r.get<double>(), cfenergy = r.get<double>(), cutoff = r.get<double>();
        const int64_t elen = r.get<int64_t>();
        std::vector<char> em = r.vec<char>(elen);
        const std::string emap(em.begin(), em.end());
        std::vector<int32_t> type = r.vec<int32_t>(nlocal);
        std::vector<int64_t> tag = r.vec<int64_t>(nlocal);
        std::vector<double> pos = r.vec<double>(3 * nlocal);
        std::vector<double> lohi = r.vec<double>(9);
        std::vector<int32_t> numneigh = r.vec<int32_t>(nlocal);

        nnpx::InterfaceLammps ifc;
        const char* dir = argv[1];
        const char* emc = emap.c_str();
        ifc.initialize(dir, emc, false, false, 0, 0, cflength, cfenergy, cutoff, (int)ntypes, 0);
        if (argc > 4 && std::strcmp(argv[4], "mixed") == 0) ifc.setPrecision(nnpx::Mode::Precision::MIXED);
        ifc.setLocalAtoms((int)nlocal, type.data());
        ifc.setLocalTags(tag.data());
        std::vector<double const*> rows(nlocal);
        for (int64_t i = 0; i < nlocal; ++i) rows[i] = &pos[3 * i];
        ifc.setLocalAtomPositions(rows.data());
        ifc.setBoxVectors(&lohi[0], &lohi[3], lohi[6], lohi[7], lohi[8]);
        ifc.allocateNeighborlists(numneigh.data());
        for (int64_t i = 0; i < nlocal; ++i)
            for (int k = 0; k < numneigh[i]; ++k) {
                const int32_t j = r.get<int32_t>();
                const int64_t t = r.get<int64_t>();
                const int32_t ty = r.get<int32_t>();
                const double dx = r.get<double>(), dy = r.get<double>(), dz = r.get<double>(), d2 = r.get<double>();
                ifc.addNeighbor((int)i, j, t, ty, dx, dy, dz, d2);
            }
        std::fclose(r.f);
        ifc.finalizeNeighborList();
        ifc.process();
        std::vector<double> F(3 * nrows, 0.0);
        std::vector<double*> frows(nrows);
        for (int64_t q = 0; q < nrows; ++q) frows[q] = &F[3 * q];
        ifc.getForces(frows.data());
        std::FILE* o = std::fopen(argv[3], "wb");
        if (!o) throw std::runtime_error("cannot write output");
        const double E = ifc.getEnergy();
        const int64_t n = (int64_t)ifc.structure.atoms.size();
        std::fwrite(&E, sizeof E, 1, o);
        std::fwrite(&n, sizeof n, 1, o);
        for (int64_t q = 0; q < n; ++q) {
            const double ea = ifc.getAtomicEnergy((int)q);
            std::fwrite(&ea, sizeof ea, 1, o);
        }
        std::fwrite(F.data(), sizeof(double), F.size(), o);
        const int64_t new_ = (int64_t)ifc.getNumExtrapolationWarnings();
        std::fwrite(&new_, sizeof new_, 1, o);
        const double rc = ifc.getMaxCutoffRadius();
        std::fwrite(&rc, sizeof rc, 1, o);
        std::fclose(o);
        return 0;
    } catch (std::exception const& e) {
        std::fprintf(stderr, "%s", e.what());
        return 2;
    }
}
