/*
 * nnpx -- MI355X (gfx950) Behler-Parrinello (2G-HDNNP) energy/force path.
 *
 * C ABI of libnnpx.so.  Plain pointers and sizes only; every entry point
 * returns an nnpx_status and never throws.  One context drives one GPU on
 * one host thread (HIP stream per context).  Citations refer to the n2p2
 * sources the entry point stands in for (src/libnnp/...).
 *
 * Units: positions, box vectors, energies and forces are in the model's
 * internal units -- the same units nnp::Mode's calculate* functions see
 * (normalised units when input.nn holds mean_energy/conv_energy/conv_length,
 * Structure::toNormalizedUnits, Structure.cpp:1250-1277).  nnpx_model_units
 * returns the conversion constants so a caller can convert exactly as
 * Prediction::readStructureFromFile / Prediction::predict do.
 */
#ifndef NNPX_H
#define NNPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nnpx_ctx nnpx_ctx;

typedef enum {
    NNPX_OK = 0,
    NNPX_ERR_ARGUMENT = 1,   /* std::invalid_argument in the reference      */
    NNPX_ERR_FILE = 2,       /* runtime_error "Could not open file"         */
    NNPX_ERR_MODEL = 3,      /* runtime_error from setup (bad input.nn ...)  */
    NNPX_ERR_RANGE = 4,      /* std::range_error / out_of_range             */
    NNPX_ERR_STATE = 5,      /* call order violated (no model, no structure) */
    NNPX_ERR_HIP = 6,        /* HIP runtime failure                         */
    NNPX_ERR_UNSUPPORTED = 7 /* model feature outside the supported subset  */
} nnpx_status;

/* Precision forms. */
typedef enum {
    NNPX_PLAIN_FP64 = 0, /* FP64 throughout, stored dG/dr (reference layout) */
    NNPX_MIXED = 1       /* FP32 SF (compensated), BF16 MFMA NN, fused
                            recomputing force kernel, FP64 reductions        */
} nnpx_precision;

/* Force semantics. */
typedef enum {
    NNPX_FORCES_PULL = 0, /* nnp::Mode::calculateForces, Mode.cpp:1922-1984:
                             contributions of an atom's own periodic images
                             are not added (loop starts at neighborsUnique+1) */
    NNPX_FORCES_PUSH = 1  /* InterfaceLammps::getForces,
                             InterfaceLammps.cpp:679-740: every neighbour
                             entry scatters, self images included            */
} nnpx_force_mode;

/* Options for nnpx_set_option. */
typedef enum {
    NNPX_OPT_FORCE_MODE = 0,  /* nnpx_force_mode, default NNPX_FORCES_PULL  */
    NNPX_OPT_TIMING = 1,      /* 1: record HIP events around every kernel   */
    NNPX_OPT_NEIGHBOR_ALGO = 2, /* 0 auto (cell list where every axis holds
                                  >= 3 cells), 1 brute force, 2 cell list   */
    NNPX_OPT_BF16_TERMS = 3,    /* mixed form NN: 1 = plain bf16 operands,
                                  3 = hi/lo split, three bf16 MFMA products
                                  (default 3)                              */
    NNPX_OPT_SF_GATHER = 4,     /* mixed SF kernel: copy each group's partner
                                  geometry into list order; -1 auto (when
                                  it costs no occupancy, default), 0 off,
                                  1 on (where it fits in LDS); results are
                                  identical either way                     */
    NNPX_OPT_PLAIN_FORCES = 5,  /* forces of nnpx_evaluate in the plain form:
                                  0 (default) recompute every SF derivative
                                  in FP64 and contract it with dE/dG on the
                                  fly (no dG/dr stored; lists too long for
                                  the kernel's LDS fall back to 1); 1 store
                                  dG/dr in the reference layout and contract
                                  it (Mode::calculateForces).  The step calls
                                  (nnpx_calculate_*) always store dG/dr.   */
    NNPX_OPT_ANGULAR_MERGE_SF = 6,  /* mixed SF kernel: walk all angular
                                  groups of an element in one triplet pass
                                  (models whose groups share type, cutoff
                                  and term constants); -1 auto (where the
                                  elements average >= 4 angular groups,
                                  default), 0 off, 1 on; results agree to
                                  FP32 summation order                     */
    NNPX_OPT_ANGULAR_MERGE_FORCE = 7, /* the same for the mixed force kernel */
    NNPX_OPT_MD_STAGING = 8,    /* MD (read by nnpx_md_init): LDS slots of
                                  the fused kernels for the in-cutoff
                                  neighbours of an atom; -1 (default) the
                                  most met at nnpx_md_init plus an eighth
                                  (at least 16), 0 the whole skin-extended
                                  list, n > 0 that many (rounded up to 32).
                                  A run of steps that meets more is
                                  repeated from its start with the whole
                                  list staged.                             */
    NNPX_OPT_PAIR_LISTS = 9     /* mixed form: the SF kernel keeps each
                                  atom's angular pairs in device memory
                                  and the force kernel reads them instead
                                  of testing the candidates again (both
                                  walk row x column blocks); off, the
                                  force kernel walks rows.  -1 auto (off
                                  until nnpx_autotune keeps the faster,
                                  default), 0 off, 1 on; results agree to
                                  FP32 summation order                     */
} nnpx_option;

/* Flags for nnpx_evaluate. */
#define NNPX_EVAL_FORCES 1u  /* compute forces (else energy + dE/dG only) */

/* ---- context ---------------------------------------------------------- */
/* No counterpart: creates the HIP stream and device state of one Mode. */
nnpx_status nnpx_create(nnpx_ctx** ctx, int device);
nnpx_status nnpx_destroy(nnpx_ctx* ctx);
/* Last error message of this context (or of the last failed create). */
const char* nnpx_last_error(const nnpx_ctx* ctx);
nnpx_status nnpx_set_option(nnpx_ctx* ctx, int option, int64_t value);
/* Current value of an option (as set, or as the autotuner chose it). */
nnpx_status nnpx_get_option(const nnpx_ctx* ctx, int option, int64_t* value);
nnpx_status nnpx_synchronize(nnpx_ctx* ctx);

/* ---- model ------------------------------------------------------------ */
/* Replaces Prediction::setup (Prediction.cpp:36-48): Mode::loadSettingsFile,
 * Mode::setupGeneric (Mode.cpp:213-237), setupSymmetryFunctionScaling
 * (Mode.cpp:781-909), setupNeuralNetworkWeights (Mode.cpp:1514-1547).
 * dir holds input.nn, scaling.data and weights.%03d.data. */
nnpx_status nnpx_load_model(nnpx_ctx* ctx, const char* dir);
/* Mode::getNumElements / Element::numSymmetryFunctions /
 * Mode::getMaxCutoffRadius (Mode.h). */
nnpx_status nnpx_model_info(const nnpx_ctx* ctx, int* num_elements,
                            int* max_sf, double* max_cutoff);
nnpx_status nnpx_num_sf(const nnpx_ctx* ctx, int element, int* num_sf);
/* Atomic number of element index e (ElementMap::atomicNumber). */
nnpx_status nnpx_atomic_number(const nnpx_ctx* ctx, int element, int* z);
/* Element index of a symbol (ElementMap::operator[]). */
nnpx_status nnpx_element_index(const nnpx_ctx* ctx, const char* symbol, int* element);
/* Number of SFs of a centre atom of element e that depend on a neighbour of
 * element e_nbr: the length of Atom::Neighbor::dGdr
 * (Element::getSymmetryFunctionNumTable, Element.cpp:289-311). */
nnpx_status nnpx_num_sf_derivatives(const nnpx_ctx* ctx, int element, int element_nbr, int* n);
/* Symbol of element index e (ElementMap::operator[](size_t)), written with
 * its terminating NUL into buf of len bytes. */
nnpx_status nnpx_element_symbol(const nnpx_ctx* ctx, int element, char* buf, int len);
/* normalize flag, mean_energy, conv_energy, conv_length and the per-element
 * atom_energy offsets (Mode.cpp:239-300, 2079-2096); offsets has
 * num_elements entries. */
nnpx_status nnpx_model_units(const nnpx_ctx* ctx, int* normalize,
                             double* mean_energy, double* conv_energy,
                             double* conv_length, double* offsets);

/* ---- structures ------------------------------------------------------- */
/* Replaces filling nnp::Structure (Structure::readFromLines,
 * Structure.cpp:142-271) for a batch of structures packed back to back.
 *   atom_offsets[n_struct+1]: first atom of each structure
 *   pos[3*N]: positions, elem[N]: element indices (order of atomic number)
 *   box[9*n_struct]: lattice vectors as rows (ignored where !periodic)
 *   periodic[n_struct]: 1 periodic, 0 isolated
 * Positions of periodic structures are used as given (callers remap, as
 * Structure::remap does at read time).  Data is copied to device memory. */
nnpx_status nnpx_set_structures(nnpx_ctx* ctx, int n_struct,
                                const int64_t* atom_offsets,
                                const double* pos, const int* elem,
                                const double* box, const int* periodic);
/* Overwrite positions only (same atoms, same boxes), e.g. for finite
 * differences or MD driven from the host. */
nnpx_status nnpx_update_positions(nnpx_ctx* ctx, const double* pos);

/* Caller-provided neighbour lists, the LAMMPS path
 * (InterfaceLammps::setLocalAtoms / addNeighbor, InterfaceLammps.cpp:
 * 410-441): nlocal atoms (elem[nlocal]) are evaluated; their neighbour
 * entries p in [nb_offsets[i], nb_offsets[i+1]) give the index of the
 * neighbour in the caller's atom array (local or ghost, 0..ntotal-1), its
 * element and the distance vector r_i - r_j (nb_dr[3p..3p+2], model length
 * units).  Periodicity is carried by the ghosts.  Replaces the structures of
 * the context; forces then use the push form and cover all ntotal atoms
 * (nnpx_get_forces returns 3*ntotal values, ghosts included, as
 * InterfaceLammps::getForces fills LAMMPS' local+ghost force array,
 * InterfaceLammps.cpp:679-740).  No neighbour list is built on the device. */
nnpx_status nnpx_set_neighbors_external(nnpx_ctx* ctx, int nlocal, const int* elem,
                                        int ntotal, const int64_t* nb_offsets,
                                        const int* nb_index, const int* nb_elem,
                                        const double* nb_dr);

/* ---- compute (the nnp::Mode boundary) --------------------------------- */
/* Structure::calculateNeighborList(maxCutoffRadius, false)
 * (Structure.cpp:287-398): same (j, image) set, same order. */
nnpx_status nnpx_calculate_neighbor_list(nnpx_ctx* ctx);
/* Mode::calculateSymmetryFunctionGroups (Mode.cpp:1630-1709), FP64.
 * derivatives != 0 also stores dG/dr in the reference layout
 * (Atom::dGdr, Atom::Neighbor::dGdr). */
nnpx_status nnpx_calculate_symmetry_function_groups(nnpx_ctx* ctx, int derivatives);
/* Mode::calculateAtomicNeuralNetworks (Mode.cpp:1711-1732), FP64;
 * derivatives != 0 also computes dE/dG (adjoint). */
nnpx_status nnpx_calculate_atomic_neural_networks(nnpx_ctx* ctx, int derivatives);
/* Mode::calculateEnergy (Mode.cpp:1876-1901): per-structure sums. */
nnpx_status nnpx_calculate_energy(nnpx_ctx* ctx);
/* Mode::calculateForces (Mode.cpp:1922-1984) with the force mode option. */
nnpx_status nnpx_calculate_forces(nnpx_ctx* ctx);
/* Mode::evaluateNNP (Mode.cpp:2029-2077) in either precision form:
 * neighbour list, SF, NN, energy and (flag NNPX_EVAL_FORCES) forces.
 * The plain form's forces follow NNPX_OPT_PLAIN_FORCES (nnpx_get_dGdr needs
 * the stored form).  Asynchronous on the context's stream; getters
 * synchronise. */
nnpx_status nnpx_evaluate(nnpx_ctx* ctx, int precision, unsigned flags);

/* ---- results ---------------------------------------------------------- */
nnpx_status nnpx_get_energy(nnpx_ctx* ctx, double* energy /* n_struct */);
nnpx_status nnpx_get_atomic_energies(nnpx_ctx* ctx, double* e /* N */);
nnpx_status nnpx_get_forces(nnpx_ctx* ctx, double* f /* 3N */);
/* Scaled symmetry functions and dE/dG of every atom, row stride ld >= max_sf
 * (Atom::G, Atom::dEdG). */
nnpx_status nnpx_get_G(nnpx_ctx* ctx, double* G, int ld);
nnpx_status nnpx_get_dEdG(nnpx_ctx* ctx, double* dEdG, int ld);
/* Neighbour list: total entry count, then per-atom counts and, for every
 * entry in list order, the neighbour index and its image shift
 * (bc0, bc1, bc2).  Pass NULL arrays to query the total only. */
nnpx_status nnpx_get_neighbors(nnpx_ctx* ctx, int64_t* total, int* counts,
                               int* j, int* shift);
/* Stored dG/dr of the plain form (after SF with derivatives):
 *   self[N*ld_self*3] -- Atom::dGdr, row stride ld_self >= max_sf
 *   nbr[total*ld_nbr*3] -- Atom::Neighbor::dGdr per neighbour entry, slots
 *   in Element::symmetryFunctionTable order (Element.cpp:289-311).
 * Either pointer may be NULL. */
nnpx_status nnpx_get_dGdr(nnpx_ctx* ctx, double* self, int ld_self,
                          double* nbr, int ld_nbr);

/* ---- checks of the SF driver ------------------------------------------ */
/* Extrapolation warnings of the last evaluation's symmetry functions
 * (Element::updateSymmetryFunctionStatistics, Element.cpp:462-541, run by
 * Mode::calculateSymmetryFunctionGroups when enabled, Mode.cpp:1694-1702, as
 * Prediction::setup and InterfaceLammps::initialize enable it): every
 * (atom, SF) whose unscaled value (SymFnc::unscale) lies below Gmin - 1000 eps
 * or above Gmax + 1000 eps of scaling.data.  *count receives their number;
 * the first min(max, count) in (atom, SF) order are written: atom index,
 * SF index (0-based, the element's sorted order), SF type, unscaled value,
 * Gmin and Gmax (model units).  max = 0 counts only; the optional arrays
 * may be NULL. */
nnpx_status nnpx_extrapolation_warnings(nnpx_ctx* ctx, int64_t max, int64_t* count,
                                        int* atom, int* sf, int* type,
                                        double* value, double* gmin, double* gmax);
/* Atoms with fewer neighbours within their element's smallest SF cutoff than
 * their SFs need (1 radial, 2 angular: Element::getMinNeighbors /
 * getMinCutoffRadius, Element.cpp:384-411) -- the check of
 * Mode::calculateSymmetryFunctionGroups (Mode.cpp:1667-1679).  Same
 * count/max convention; records are (atom index, neighbour count), in atom
 * order. */
nnpx_status nnpx_low_neighbor_atoms(nnpx_ctx* ctx, int64_t max, int64_t* count,
                                    int* atom, int* num_neighbors);

/* ---- training gradient ------------------------------------------------ */
/* Connections (weights and biases) of element e's network, or of all
 * networks for e = -1 (Training::getNumConnections). */
nnpx_status nnpx_num_connections(const nnpx_ctx* ctx, int element, int* n);
/* All connections, elements in order, each network in
 * NeuralNetwork::setConnections order (NeuralNetwork.cpp:147-169) -- the
 * order of the weights.%03d.data files. */
nnpx_status nnpx_get_weights(const nnpx_ctx* ctx, double* w);
nnpx_status nnpx_set_weights(nnpx_ctx* ctx, const double* w);
/* Energy+force loss of the current structures and its gradient with respect
 * to every connection (layout of nnpx_get_weights):
 *   L = sum_s (E_s - E_ref[s])^2 / N_s + beta/(3N) sum_{a,c} (F_ac - F_ref)^2
 * loss[0] = L, loss[1] = energy term, loss[2] = force term.  The force term
 * is differentiated through the forces (second order) as Training::dPdc
 * with NeuralNetwork::calculateDFdc and Training::collectDGdxia do
 * (Training.cpp:3319-3389, 3586-3625).  NNPX_PLAIN_FP64: FP64 from the stored
 * dG/dr; NNPX_MIXED: FP32 recomputed tangents and FP32 network adjoints,
 * gradient summed in FP64.  Model units; normalize_nodes is not supported. */
nnpx_status nnpx_loss_gradient(nnpx_ctx* ctx, int precision, const double* E_ref,
                               const double* F_ref, double beta, double* loss,
                               double* grad);
/* Weight derivatives of energies and forces, one gradient per structure
 * (replaces Training::dPdc, Training.cpp:3319-3389, for the "energy" and
 * "force" properties).  With seeds vE[n_struct] and vF[3N], structure s
 * receives
 *   grad[s * n_conn + c] = vE[s] dE_s/dw_c + sum_{a in s, k} vF[3a+k] dF_ak/dw_c
 * (n_conn = nnpx_num_connections(ctx, -1, ...), layout of nnpx_get_weights).
 * vE = 1, vF = 0 gives the "energy" row; a unit vF on (atom a, component k)
 * gives the "force" row of that component, so packing copies of a structure
 * with different seeds yields the whole Jacobian in one call.  Evaluates the
 * structures first (forces included).  Model units; normalize_nodes is not
 * supported. */
nnpx_status nnpx_weight_vjp(nnpx_ctx* ctx, int precision, const double* vE,
                            const double* vF, double* grad);

/* ---- training update (Kalman filter) ---------------------------------- */
/* The weight updaters of Training with updater_type 1 (Training.cpp:
 * 1040-1175; KalmanFilter constructor and setParametersStandard /
 * setParametersFadingMemory, KalmanFilter.cpp:27-250): one filter over all
 * connections (update_strategy 0) or one per element (1), P = I / epsilon
 * in FP64 on the device.  type 0 (standard): par = {epsilon, q0, qtau,
 * qmin, eta0, etatau, etamax}; type 1 (fading memory): {epsilon, q0, qtau,
 * qmin, lambda, nu, unused}.  qtau and etatau are per update (the reference
 * divides the input.nn values by the projected updates per epoch). */
nnpx_status nnpx_kalman_init(nnpx_ctx* ctx, int type, int per_element, const double* par);
/* One Training::update (Training.cpp:2256-2697) of the weights for the
 * property "energy" (0) or "force" (1), for a batch of update candidates of
 * the current structures chosen by the caller (the reference draws them by
 * its selection mode): structure[b]; for forces also atom[b] (index in the
 * packed structures) and component[b].  E_ref[n_struct] / F_ref[3N] in
 * model units.  jacobian_full = 0 sums the batch into one observation
 * (jacobian_mode 0, and 1 on one process), 1 keeps one per candidate
 * (jacobian_mode 2).  Per-element filters scale each candidate's error by
 * the element's share (atoms of the structure, or neighbours of the atom).
 * Forces: error and Jacobian times force_weight.  The Jacobian rows come from
 * the weight-derivative kernels in `precision`; the filter runs in FP64.
 * Afterwards the model holds the new weights; error[b] (optional) receives
 * each candidate's error before the update. */
nnpx_status nnpx_kalman_update(nnpx_ctx* ctx, int precision, int property, int batch,
                               const int* structure, const int* atom, const int* component,
                               const double* E_ref, const double* F_ref, double force_weight,
                               int jacobian_full, double* error);
/* State of updater u: P[n*n] (optional) and scalars[7] = {n, first
 * connection, eta, lambda, gamma, q, updates}. */
nnpx_status nnpx_kalman_state(nnpx_ctx* ctx, int updater, double* P, double* scalars);

/* ---- molecular dynamics ----------------------------------------------- */
/* No counterpart in nnp::Mode: the velocity-Verlet loop a driver runs around
 * the energy/force call (LAMMPS fix nve with pair_style hdnnp,
 * InterfaceLammps.cpp:463-500 per step).  One structure per context.
 *   masses[num_elements]: amu; vel[3N]: A/fs (physical); dt: fs;
 *   skin: A (physical), added to the list cutoff; the list is rebuilt on the
 *   device when an atom moved more than skin/2 since the last build.
 * Computes the forces of the initial positions. */
nnpx_status nnpx_md_init(nnpx_ctx* ctx, int precision, const double* masses,
                         const double* vel, double dt, double skin);
/* Runs nsteps steps.  graph_steps > 0 captures that many steps in a HIP
 * graph (once, re-captured when graph_steps changes or a buffer grows) and
 * replays it; the remainder runs eagerly.  Fails with NNPX_ERR_RANGE if the
 * neighbour capacity overflowed during the run. */
nnpx_status nnpx_md_run(nnpx_ctx* ctx, int nsteps, int graph_steps);
/* Current positions (model units, unwrapped since the last rebuild) and
 * velocities (A/fs); either pointer may be NULL. */
nnpx_status nnpx_md_state(nnpx_ctx* ctx, double* pos, double* vel);
/* Potential and kinetic energy (model energy units) of the first
 * min(max, steps) steps since nnpx_md_init, the step count and the number
 * of list rebuilds. */
nnpx_status nnpx_md_energies(nnpx_ctx* ctx, double* epot, double* ekin,
                             int64_t max, int64_t* nsteps, int64_t* rebuilds);

/* ---- launch-shape autotuner ------------------------------------------- */
/* No counterpart in the reference.  Times candidate launch shapes of the
 * mixed-form kernels on the current structures (HIP events, 3 launches
 * each) and keeps the fastest in the context: the triplet-buffer size of the
 * SF, force and tangent kernels (LDS per wave against batching) and the
 * atoms per block of the gradient kernel.  Results are not changed by the
 * choice beyond summation order.  values[4] = {trip_sf, trip_force,
 * trip_tangent, grad_atoms}; triplet buffers are multiples of 64 in
 * [64, 4096]. */
nnpx_status nnpx_autotune(nnpx_ctx* ctx);
nnpx_status nnpx_get_tuning(const nnpx_ctx* ctx, int* values);
nnpx_status nnpx_set_tuning(nnpx_ctx* ctx, const int* values);

/* ---- evaluation without device allocation ----------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): nnpx_set_structures + nnpx_evaluate + nnpx_get_energy
 * (+ nnpx_get_forces) in one call, with every intermediate buffer (neighbour
 * and cell lists, symmetry functions, dE/dG, energies, forces) carved from a
 * file-scope device array of 64 MiB that each GPU holds until the process
 * ends; the call itself allocates no device memory.  Same kernels and results
 * as the allocating calls.  Inputs and outputs are host arrays as in those
 * calls: energy[n_struct], forces[3N] (with NNPX_EVAL_FORCES).  An input whose
 * buffers exceed the array fails with NNPX_ERR_RANGE; a context set up for MD
 * is refused (NNPX_ERR_STATE).  Afterwards the context holds no structures.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (from several contexts or threads) race. */
nnpx_status nnpx_predict_noalloc(nnpx_ctx* ctx, int precision, unsigned flags,
                                 int n_struct, const int64_t* atom_offsets,
                                 const double* pos, const int* elem,
                                 const double* box, const int* periodic,
                                 double* energy, double* forces);

/* ---- measurement ------------------------------------------------------ */
/* No counterpart in the reference: multiplies two known 16x16 matrices with
 * the bf16 MFMA operand layout the NN kernels use and returns the max abs
 * error against the exact product (a lane-layout check of the device). */
nnpx_status nnpx_mfma_selftest(nnpx_ctx* ctx, double* max_abs_err);
/* With NNPX_OPT_TIMING set, summed event time and launch count of the named
 * kernel since the last reset ("nbr", "sf64", "nn64", "force64", "energy",
 * "sf32", "nnbf16", "force32", ...). */
nnpx_status nnpx_kernel_time(nnpx_ctx* ctx, const char* name, double* ms,
                             int64_t* launches);
nnpx_status nnpx_reset_kernel_times(nnpx_ctx* ctx);
/* Algorithmic work counters of the current structures (after a neighbour
 * list): P pairs within the SF cutoff, T triplets, member evaluations. */
nnpx_status nnpx_work_counts(nnpx_ctx* ctx, int64_t* pairs,
                             int64_t* triplets, int64_t* pair_members,
                             int64_t* triplet_members, int64_t* triplet_exps);

#ifdef __cplusplus
}
#endif

#endif /* NNPX_H */
