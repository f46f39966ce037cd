// Drives the nnpx C++ façade the way n2p2 code drives nnp::Mode
// (usage of Mode.h:72-80 and Prediction.cpp:50-78) and prints the results as
// JSON for tests/test_facade.py to check.
//
//   test_facade structure <model_dir> <input.data> <output.data>
//       CPU only: parse, remap and normalise a structure, write it back.
//   test_facade steps <model_dir> <input.data> [mixed]
//       GPU: neighbour list, SF groups, NNs, energy and forces step by step.
//   test_facade dgdr <model_dir> <input.data> <out.bin>
//       GPU: SF groups with setCopyNeighbors/setCopyDerivatives(true); writes
//       (model units) for every atom its Atom::dGdr rows, then for every
//       neighbour entry in list order its index and Neighbor::dGdr rows, as
//       int64 counts followed by doubles (layout in the code below).
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "nnpx_mode.hpp"

namespace {

void print_structure(nnpx::Structure const& s, bool results)
{
    std::printf("{\"numAtoms\": %zu, \"isPeriodic\": %d, \"energy\": %.17g, \"energyRef\": %.17g,\n", s.numAtoms,
                (int)s.isPeriodic, s.energy, s.energyRef);
    std::printf(" \"box\": [");
    for (int i = 0; i < 3; ++i)
        std::printf("%s[%.17g, %.17g, %.17g]", i ? ", " : "", s.box[i][0], s.box[i][1], s.box[i][2]);
    std::printf("],\n \"r\": [");
    for (size_t i = 0; i < s.atoms.size(); ++i)
        std::printf("%s[%.17g, %.17g, %.17g]", i ? ", " : "", s.atoms[i].r[0], s.atoms[i].r[1], s.atoms[i].r[2]);
    std::printf("]");
    if (results) {
        std::printf(",\n \"f\": [");
        for (size_t i = 0; i < s.atoms.size(); ++i)
            std::printf("%s[%.17g, %.17g, %.17g]", i ? ", " : "", s.atoms[i].f[0], s.atoms[i].f[1], s.atoms[i].f[2]);
        std::printf("],\n \"G0\": [");
        for (size_t q = 0; q < s.atoms[0].G.size(); ++q) std::printf("%s%.17g", q ? ", " : "", s.atoms[0].G[q]);
        std::printf("],\n \"dEdG0\": [");
        for (size_t q = 0; q < s.atoms[0].dEdG.size(); ++q) std::printf("%s%.17g", q ? ", " : "", s.atoms[0].dEdG[q]);
        std::printf("],\n \"numNeighbors0\": %zu, \"numNeighbors\": [", s.atoms[0].numNeighbors);
        for (size_t i = 0; i < s.atoms.size(); ++i) std::printf("%s%zu", i ? ", " : "", s.atoms[i].numNeighbors);
        std::printf("],\n \"flags\": [%d, %d, %d]", (int)s.hasNeighborList, (int)s.hasSymmetryFunctions,
                    (int)s.hasSymmetryFunctionDerivatives);
    }
    std::printf("}\n");
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        if (argc >= 5 && std::strcmp(argv[1], "structure") == 0) {
            // The element map and units come from input.nn only; no device is touched.
            nnpx::ElementMap em;
            std::ifstream in(std::string(argv[2]) + "/input.nn");
            std::string line, elems;
            double mean = 0.0, ce = 1.0, cl = 1.0;
            bool norm = false;
            while (std::getline(in, line)) {
                char key[64] = {0};
                if (std::sscanf(line.c_str(), "%63s", key) != 1) continue;
                if (!std::strcmp(key, "elements")) elems = line.substr(line.find("elements") + 8);
                if (!std::strcmp(key, "mean_energy")) { std::sscanf(line.c_str(), "%*s %lf", &mean); norm = true; }
                if (!std::strcmp(key, "conv_energy")) std::sscanf(line.c_str(), "%*s %lf", &ce);
                if (!std::strcmp(key, "conv_length")) std::sscanf(line.c_str(), "%*s %lf", &cl);
            }
            if (elems.find('#') != std::string::npos) elems.erase(elems.find('#'));
            em.registerElements(elems);
            nnpx::Structure s;
            s.setElementMap(em);
            s.readFromFile(argv[3]);
            s.writeToFile(std::string(argv[4]), true, false);
            if (norm) s.toNormalizedUnits(mean, ce, cl, 1.0);
            print_structure(s, false);
            return 0;
        }
        if (argc >= 4 && std::strcmp(argv[1], "steps") == 0) {
            nnpx::Prediction p;
            p.directory = argv[2];
            p.setup();
            if (argc >= 5 && std::strcmp(argv[4], "mixed") == 0) p.setPrecision(nnpx::Mode::Precision::MIXED);
            p.readStructureFromFile(argv[3]);
            nnpx::Structure& s = p.structure;
            s.calculateNeighborList(p.getMaxCutoffRadius());
            p.calculateSymmetryFunctionGroups(s, true);
            p.calculateAtomicNeuralNetworks(s, true);
            p.calculateEnergy(s);
            p.calculateForces(s);
            if (p.useNormalization())
                s.toPhysicalUnits(p.getMeanEnergy(), p.getConvEnergy(), p.getConvLength(), 1.0);
            p.addEnergyOffset(s, false);
            p.addEnergyOffset(s, true);
            print_structure(s, true);
            return 0;
        }
        if (argc >= 5 && std::strcmp(argv[1], "dgdr") == 0) {
            nnpx::Prediction p;
            p.directory = argv[2];
            p.setup();
            p.setCopyNeighbors(true);
            p.setCopyDerivatives(true);
            p.readStructureFromFile(argv[3]);
            nnpx::Structure& s = p.structure;
            s.calculateNeighborList(p.getMaxCutoffRadius());
            p.calculateSymmetryFunctionGroups(s, true);
            // per atom: int64 nsf, nsf*3 doubles, int64 numNeighbors; per
            // neighbour: int64 index, int64 n, n*3 doubles
            std::FILE* f = std::fopen(argv[4], "wb");
            if (!f) throw std::runtime_error("cannot write output");
            auto put_i = [&](int64_t v) { std::fwrite(&v, sizeof v, 1, f); };
            auto put_v = [&](std::vector<nnpx::Vec3D> const& v) {
                put_i((int64_t)v.size());
                for (auto const& x : v) {
                    double const w[3] = {x[0], x[1], x[2]};
                    std::fwrite(w, sizeof(double), 3, f);
                }
            };
            for (auto const& a : s.atoms) {
                put_v(a.dGdr);
                put_i((int64_t)a.neighbors.size());
                for (auto const& n : a.neighbors) {
                    put_i((int64_t)n.index);
                    put_v(n.dGdr);
                }
            }
            std::fclose(f);
            std::printf("{\"numAtoms\": %zu, \"flags\": [%d, %d]}\n", s.numAtoms, (int)s.hasSymmetryFunctions,
                        (int)s.hasSymmetryFunctionDerivatives);
            return 0;
        }
    } catch (std::exception const& e) {
        std::fprintf(stderr, "%s", e.what());
        return 2;
    }
    std::fprintf(stderr, "usage: test_facade structure|steps ...\n");
    return 1;
}
