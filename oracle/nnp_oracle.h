/*
 * CPU FP64 oracle for the Behler-Parrinello (2G-HDNNP) energy/force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (n2p2_amd, libnnpx) never links, loads or calls it.
 *
 * It restates, from scratch and expression for expression, the reference
 * algorithm of n2p2 (citations are relative to the n2p2 source tree):
 *   - input.nn subset, element map, SF parsing/sorting/grouping, scaling,
 *     NN topology and weights     Mode.cpp:213-237, 589-909, 1227-1512, 2331-2350
 *   - input.data reading + remap  Structure.cpp:142-271, 518-544, 1215-1248
 *   - unit normalisation          Structure.cpp:1250-1306, Atom.cpp:78-150
 *   - neighbour list (O(N^2))     Structure.cpp:287-398, 485-516
 *   - radial / angular SF groups  SymGrpExpRad.cpp:144-239, SymGrpExpAngn.cpp:156-403
 *   - cutoff functions            CutoffFunction.cpp:102-163, CutoffFunction.h:210-220
 *   - NN forward + dE/dG          NeuralNetwork.cpp:386-442, 761-896
 *   - energy / forces (pull)      Mode.cpp:1876-1901, 1922-1984
 *   - forces (push, LAMMPS form)  InterfaceLammps.cpp:679-740
 *   - Prediction::predict         Prediction.cpp:50-78, Mode.cpp:2029-2096
 *
 * All functions return 0 on success and a non-zero code on error; the error
 * text is available from oracle_last_error().
 */
#ifndef NNP_ORACLE_H
#define NNP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

/* Force semantics. */
enum { ORACLE_FORCES_PULL = 0, /* Mode::calculateForces (Mode.cpp:1922-1984) */
       ORACLE_FORCES_PUSH = 1  /* InterfaceLammps::getForces (InterfaceLammps.cpp:679-740) */ };

/* Neighbour list algorithm. */
enum { ORACLE_NBR_BRUTE = 0,   /* Structure::calculateNeighborList */
       ORACLE_NBR_CELL  = 1 }; /* set-identical cell list, same order */

const char* oracle_last_error(void);

/* Setup from a model directory holding input.nn, scaling.data and
 * weights.%03d.data (Prediction::setup, Prediction.cpp:36-48). */
oracle_ctx* oracle_create(const char* model_dir);
void        oracle_destroy(oracle_ctx* c);

int    oracle_num_elements(const oracle_ctx* c);
int    oracle_num_sf(const oracle_ctx* c, int element);
int    oracle_atomic_number(const oracle_ctx* c, int element);
double oracle_max_cutoff(const oracle_ctx* c);
int    oracle_normalize(const oracle_ctx* c);

/* Read the first structure of an input.data file
 * (Prediction::readStructureFromFile, Prediction.cpp:50-65): remap,
 * remove energy offsets, convert to normalised units. */
int oracle_read_structure(oracle_ctx* c, const char* path);

/* Set a structure from arrays in physical units.  box is 9 doubles (three
 * lattice vectors as rows) or NULL for a non-periodic structure.  Positions
 * are remapped when periodic, exactly as Structure::readFromLines does.
 * elem holds element indices (order of atomic number). */
int oracle_set_structure(oracle_ctx* c, int n, const double* pos,
                         const int* elem, const double* box);

/* Full prediction (Prediction::predict): neighbour list, SF groups, NNs,
 * energy, forces, back to physical units, energy offsets added.
 * nthreads <= 0 leaves the OpenMP default. */
int oracle_predict(oracle_ctx* c, int force_mode, int nbr_mode, int nthreads);

/* Results (physical units after oracle_predict). */
int    oracle_num_atoms(const oracle_ctx* c);
double oracle_energy(const oracle_ctx* c);
int    oracle_get_forces(const oracle_ctx* c, double* f);          /* 3N */
int    oracle_get_atomic_energies(const oracle_ctx* c, double* e); /* N (normalised units) */
int    oracle_get_positions(const oracle_ctx* c, double* r);       /* 3N, after remap (physical) */
/* Per-atom scaled SFs and dE/dG (normalised units), padded to ld columns. */
int    oracle_get_G(const oracle_ctx* c, double* G, int ld);
int    oracle_get_dEdG(const oracle_ctx* c, double* dEdG, int ld);

/* Stored derivatives of the last oracle_predict (normalised units):
 * self[N][ld_self][3] = Atom::dGdr; nbr[entry][ld_nbr][3] = Neighbor::dGdr of
 * every neighbour entry in list order, slots in Element::symmetryFunctionTable
 * order (Element.cpp:289-311), unused slots zero.  Either may be NULL. */
int    oracle_get_dGdr(const oracle_ctx* c, double* self, int ld_self,
                       double* nbr, int ld_nbr);

/* Checks of the SF driver after oracle_predict: extrapolation warnings
 * (Element::updateSymmetryFunctionStatistics, Element.cpp:462-541) as
 * (atom, SF, type, unscaled value, Gmin, Gmax) records in (atom, SF) order,
 * and low-neighbour atoms (Mode.cpp:1667-1679) as (atom, count) records.
 * Both return the number of records and write at most max. */
long   oracle_extrapolation_warnings(const oracle_ctx* c, long max, int* atom, int* sf,
                                     int* type, double* value, double* gmin,
                                     double* gmax);
long   oracle_low_neighbor_atoms(const oracle_ctx* c, long max, int* atom, int* num);

/* KalmanFilter::update (KalmanFilter.cpp:119-185) on caller arrays: P [n][n]
 * (lower triangle read, symmetric result written), w [n], H [n][m],
 * xi [m]; par (in/out) = {type (0 standard, 1 fading memory), eta, etatau,
 * etamax, lambda, nu, gamma, q, qtau, qmin}. */
int    oracle_kalman_update(int n, int m, double* P, double* w, const double* H,
                            const double* xi, double* par);

/* Neighbour list of atom i: count, then indices and image shifts
 * (bc0,bc1,bc2 per entry) in list order. */
int    oracle_num_neighbors(const oracle_ctx* c, int i);
int    oracle_get_neighbors(const oracle_ctx* c, int i, int* j, int* bc);

/* Wall time in seconds of the phases of the last oracle_predict:
 * 0 neighbour list, 1 SF, 2 NN, 3 energy, 4 forces. */
double oracle_phase_time(const oracle_ctx* c, int phase);

/* Only the neighbour list (for set-identity checks); returns total pairs. */
long   oracle_build_neighbors(oracle_ctx* c, int nbr_mode, int nthreads);

/* Cutoff function value and derivative (CutoffFunction::fdf). */
int    oracle_cutoff_fdf(int type, double rc, double alpha, double r,
                         double* fc, double* dfc);

/* ---- training gradient (restated from NeuralNetwork::calculateDEdc /
 * calculateDFdc, Training::dPdc and Training::collectDGdxia) ---- */
/* Connections of element e's network (setConnections order). */
int oracle_num_connections(const oracle_ctx* c, int e);
/* All networks' connections, elements in order (weights files' order). */
int oracle_get_weights(const oracle_ctx* c, double* w);
int oracle_set_weights(oracle_ctx* c, const double* w);
/* After oracle_predict (forces, pull form) of a non-normalised model:
 * loss = wE (E - E_ref)^2 + wF sum_{a,c} (F_ac - F_ref,ac)^2 as
 * loss[0] total, loss[1] energy part, loss[2] force part, and its gradient
 * with respect to all connections (layout of oracle_get_weights). */
int oracle_loss_gradient(oracle_ctx* c, double E_ref, const double* F_ref, double wE, double wF,
                         double* loss, double* grad);
/* After oracle_predict (forces, pull form) of a non-normalised model:
 * Training::dPdc rows.  property 0: out[n_conn] = dE/dw; property 1:
 * out[3N][n_conn], row 3a+k = dF_ak/dw. */
int oracle_dPdc(oracle_ctx* c, int property, double* out);

#ifdef __cplusplus
}
#endif

#endif
