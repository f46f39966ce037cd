// nnpx-checkdw: command-line equivalent of n2p2's nnp-checkdw
// (src/application/nnp-checkdw.cpp) on one GPU: analytic derivatives of the
// energy and of every force component with respect to every network
// connection (Training::dPdc, Training.cpp:3319-3389) against central
// differences of the same quantities (Training::dPdcN, Training.cpp:3391-3471).
//
// Analytic rows come from nnpx_weight_vjp: a structure is packed 1 + 3N times,
// copy 0 seeded on its energy and copy 1 + 3a + k on force component (a, k),
// so one device call returns the energy row and the whole force Jacobian.
// Numeric rows perturb one connection at a time (w + delta, then
// (w + delta) - 2 delta, as dPdcN does) and evaluate every structure of a
// batch at once; each perturbation gives the energy and all force columns of
// that connection.  Always the plain FP64 form.  Model (normalised) units, as
// the reference works after Training::toNormalizedUnits.
//
// Writes checkdw-summary.out, checkdw-weights.energy.out,
// checkdw-weights.force.out (with use_short_forces) and nnp-checkdw.log.0000
// in the reference's formats; a property of a structure is "WARN" when its
// largest error exceeds 100 delta^2 (nnp-checkdw.cpp:237).
//
//   nnpx-checkdw <stage> [<delta>] [--device N] [--batch-atoms M]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "nnpx_mode.hpp"

namespace {

using nnpx::strpr;

int usage(const char* argv0)
{
    std::cout << "USAGE: " << argv0 << " <stage> <<delta>> [--device N] [--batch-atoms M]\n"
              << "       <stage> ..... Training stage (irrelevant for NNPs with only one stage).\n"
              << "       <<delta>> ... (optional) Displacement for central difference (default: 1.0e-4).\n"
              << "       Execute in directory with these NNP files present:\n"
              << "       - input.data (structure file)\n"
              << "       - input.nn (NNP settings)\n"
              << "       - scaling.data (symmetry function scaling data)\n"
              << "       - \"weights.%03d.data\" (weights files)\n";
    return 1;
}

void check(nnpx_ctx* ctx, nnpx_status st)
{
    if (st != NNPX_OK) throw std::runtime_error(std::string("ERROR: ") + nnpx_last_error(ctx) + "\n");
}

}  // namespace

int main(int argc, char* argv[])
{
    int device = 0;
    long batchAtoms = 1000000;
    std::vector<std::string> pos;
    for (int a = 1; a < argc; ++a) {
        std::string s = argv[a];
        if (s == "--device" && a + 1 < argc) device = std::atoi(argv[++a]);
        else if (s == "--batch-atoms" && a + 1 < argc) batchAtoms = std::max(1L, std::atol(argv[++a]));
        else if (s.rfind("--", 0) != 0) pos.push_back(s);
        else return usage(argv[0]);
    }
    if (pos.empty() || pos.size() > 2) return usage(argv[0]);
    // stage: only one stage for 2G networks (Training::setStage is a no-op)
    const double delta = pos.size() == 2 ? std::atof(pos[1].c_str()) : 1.0e-4;

    std::ofstream log("nnp-checkdw.log.0000");
    try {
        nnpx::Mode mode(device);
        mode.setupFromDirectory(".");
        nnpx_ctx* ctx = mode.context();
        const bool normalize = mode.useNormalization();
        const double cL = normalize ? mode.getConvLength() : 1.0;
        check(ctx, nnpx_set_option(ctx, NNPX_OPT_FORCE_MODE, NNPX_FORCES_PULL));
        const bool useForces = nnpx::settingsFileHasKeyword("input.nn", "use_short_forces");
        std::vector<std::string> pk{"energy"};
        if (useForces) pk.push_back("force");

        int nconnInt = 0;
        check(ctx, nnpx_num_connections(ctx, -1, &nconnInt));
        const size_t nconn = (size_t)nconnInt;
        std::vector<double> w0(nconn);
        check(ctx, nnpx_get_weights(ctx, w0.data()));

        log << "\n*** SETUP WEIGHT DERIVATIVES CHECK ********************************************\n\n";
        log << "Weight derivatives will be checked for these properties:\n";
        for (auto const& k : pk) log << " - " << (k == "energy" ? "energies" : "forces") << "\n";
        log << "\n";
        log << "\n*** ANALYTIC/NUMERIC WEIGHT DERIVATIVES CHECK *********************************\n\n";
        log << strpr("Delta for symmetric difference quotient: %11.3E\n", delta);

        std::ofstream fs("checkdw-summary.out");
        log << "Per-structure summary of analytic/numeric weight derivative comparison\n"
               "will be written to \"checkdw-summary.out\"\n";
        {
            std::vector<size_t> size{10, 10};
            std::vector<std::string> name{"struct", "numAtoms"};
            std::vector<std::string> info{"Structure index.", "Number of atoms in structure."};
            for (auto const& k : pk) {
                size.push_back(16);
                name.push_back("MAE_" + k);
                info.push_back("Mean over all absolute differences between analytic and numeric weight "
                               "derivatives for property \"" + k + "\".");
                size.push_back(16);
                name.push_back("maxAE_" + k);
                info.push_back("Maximum over all absolute differences between analytic and numeric weight "
                               "derivatives for property \"" + k + "\".");
            }
            nnpx::writeFileHeader(
                fs, {strpr("Per-structure summary of analytic vs. numeric weight derivative comparison (delta = %11.3E).",
                           delta)},
                size, name, info);
        }
        std::map<std::string, std::ofstream> fw;
        for (auto const& k : pk) {
            const std::string fn = "checkdw-weights." + k + ".out";
            log << "Individual analytic/numeric weight derivatives for property \"" + k +
                       "\"\nwill be written to \"" + fn + "\"\n";
            fw[k].open(fn);
            nnpx::writeFileHeader(fw[k],
                                  {strpr("Comparison of analytic and numeric weight derivatives for property \"%s\" "
                                         "(delta = %11.3E).",
                                         k.c_str(), delta)},
                                  {10, 10, 10, 24, 24}, {"struct", "index", "weight", "analytic", "numeric"},
                                  {"Structure index.", "Property index.", "Weight index.",
                                   "Analytic weight derivatives.", "Numeric weight derivatives."});
        }
        log << "\n";
        log << "                      |";
        for (auto const& k : pk) log << strpr(" %33s |", k.c_str());
        log << "\n";
        log << "                      |";
        for (size_t q = 0; q < pk.size(); ++q) log << " meanAbsError  maxAbsError verdict |";
        log << "\n";
        log << "-----------------------";
        for (size_t q = 0; q < pk.size(); ++q) log << "------------------------------------";
        log << "\n";

        std::vector<nnpx::Structure> structures;
        for (auto const& lines : nnpx::readStructureBlocks("input.data")) {
            structures.emplace_back();
            structures.back().setElementMap(mode.elementMap);
            structures.back().readFromLines(lines);
            structures.back().index = structures.size() - 1;
        }

        // Numeric rows of a batch are held for every connection: bound the
        // batch so that they stay within 2^27 doubles (1 GiB).
        const long numericAtoms =
            std::min(batchAtoms, std::max(1L, (long)((size_t(1) << 27) / (3 * std::max<size_t>(nconn, 1)))));
        int warning = 0;
        for (size_t b0 = 0; b0 < structures.size();) {
            size_t b1 = b0;
            long atoms = 0;
            while (b1 < structures.size() && (b1 == b0 || atoms + (long)structures[b1].numAtoms <= numericAtoms))
                atoms += (long)structures[b1++].numAtoms;
            const int ns = (int)(b1 - b0);

            // numeric: dPdcN for all structures of the batch
            nnpx::PackedStructures pb;
            for (size_t s = b0; s < b1; ++s) pb.add(structures[s], cL);
            check(ctx, nnpx_set_structures(ctx, ns, pb.offset.data(), pb.position.data(), pb.element.data(),
                                           pb.box.data(), pb.periodic.data()));
            std::vector<double> numE((size_t)ns * nconn), numF(useForces ? (size_t)3 * atoms * nconn : 0);
            std::vector<double> w = w0, Eh(ns), El(ns), Fh(useForces ? 3 * atoms : 0), Fl(Fh.size());
            const unsigned flags = useForces ? NNPX_EVAL_FORCES : 0u;
            for (size_t o = 0; o < nconn; ++o) {
                w[o] += delta;
                check(ctx, nnpx_set_weights(ctx, w.data()));
                check(ctx, nnpx_evaluate(ctx, NNPX_PLAIN_FP64, flags));
                check(ctx, nnpx_get_energy(ctx, Eh.data()));
                if (useForces) check(ctx, nnpx_get_forces(ctx, Fh.data()));
                w[o] -= 2.0 * delta;
                check(ctx, nnpx_set_weights(ctx, w.data()));
                check(ctx, nnpx_evaluate(ctx, NNPX_PLAIN_FP64, flags));
                check(ctx, nnpx_get_energy(ctx, El.data()));
                if (useForces) check(ctx, nnpx_get_forces(ctx, Fl.data()));
                w[o] = w0[o];
                for (int s = 0; s < ns; ++s) numE[(size_t)s * nconn + o] = (Eh[s] - El[s]) / (2.0 * delta);
                for (long r = 0; r < (useForces ? 3 * atoms : 0); ++r)
                    numF[(size_t)r * nconn + o] = (Fh[r] - Fl[r]) / (2.0 * delta);
            }
            check(ctx, nnpx_set_weights(ctx, w0.data()));

            // analytic: dPdc per structure from seeded copies
            for (size_t s = b0; s < b1; ++s) {
                nnpx::Structure const& st = structures[s];
                const size_t N = st.numAtoms;
                const size_t copies = 1 + (useForces ? 3 * N : 0);
                std::vector<double> ana(copies * nconn);
                const size_t perCall = std::max<size_t>(
                    1, std::min((size_t)batchAtoms / std::max<size_t>(N, 1),
                                (size_t(1) << 24) / std::max<size_t>(nconn, 1)));
                for (size_t c0 = 0; c0 < copies; c0 += perCall) {
                    const size_t c1 = std::min(copies, c0 + perCall);
                    nnpx::PackedStructures pc;
                    std::vector<double> vE(c1 - c0, 0.0), vF(3 * N * (c1 - c0), 0.0);
                    for (size_t c = c0; c < c1; ++c) {
                        pc.add(st, cL);
                        if (c == 0) vE[0] = 1.0;
                        else vF[3 * N * (c - c0) + (c - 1)] = 1.0;  // atom (c-1)/3, component (c-1)%3
                    }
                    check(ctx, nnpx_set_structures(ctx, pc.count(), pc.offset.data(), pc.position.data(),
                                                   pc.element.data(), pc.box.data(), pc.periodic.data()));
                    check(ctx, nnpx_weight_vjp(ctx, NNPX_PLAIN_FP64, vE.data(), vF.data(), ana.data() + c0 * nconn));
                }

                const size_t row0 = (size_t)(pb.offset[s - b0]) * 3;  // first force row of s in numF
                log << strpr("Configuration %6zu: ", st.index + 1);
                fs << strpr("%10zu %10zu", st.index + 1, N);
                for (auto const& k : pk) {
                    const bool en = k == "energy";
                    const size_t rows = en ? 1 : 3 * N;
                    double mean = 0.0, mx = 0.0;
                    size_t count = 0;
                    for (size_t i = 0; i < rows; ++i)
                        for (size_t j = 0; j < nconn; ++j) {
                            const double a = ana[(en ? 0 : 1 + i) * nconn + j];
                            const double n = en ? numE[(s - b0) * nconn + j] : numF[(row0 + i) * nconn + j];
                            fw[k] << strpr("%10zu %10zu %10zu %24.16E %24.16E\n", st.index + 1, i + 1, j + 1, a, n);
                            const double error = std::fabs(a - n);
                            mean += error;
                            mx = std::max(error, mx);
                            ++count;
                        }
                    mean /= (double)count;
                    std::string swarn = "OK";
                    if (mx > 100 * delta * delta) {
                        swarn = "WARN";
                        ++warning;
                    }
                    log << strpr("  %12.3E %12.3E %7s ", mean, mx, swarn.c_str());
                    fs << strpr(" %16.8E %16.8E", mean, mx);
                }
                log << "\n";
                fs << "\n";
            }
            b0 = b1;
        }
        if (warning > 0) {
            log << "\n";
            log << "IMPORTANT: Some warnings were issued. By default, this happens if the maximum\n"
                   "           absolute error (\"maxAbsError\") is higher than 100 * delta\xc2\xb2.\n"
                   "           However, this does NOT mean that analytic derivatives are incorrect!\n"
                   "           Repeat this analysis with a different delta and check whether\n"
                   "           the error scales with O(delta\xc2\xb2). The prefactor for your system\n"
                   "           could be higher than 10, hence, as long as there is a O(delta\xc2\xb2)\n"
                   "           scaling the analytic derivatives are probably correct.\n";
        }
        log << "*******************************************************************************\n";
    } catch (std::exception const& e) {
        log << e.what();
        std::cerr << e.what();
        return 1;
    }
    return 0;
}
