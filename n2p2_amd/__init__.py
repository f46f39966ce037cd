"""n2p2_amd -- MI355X (gfx950) Behler-Parrinello energy/force path.

Python mirror of the ``nnp::Mode`` compute boundary over the C ABI of
``libnnpx.so`` (``include/nnpx.h``).  The compute runs only in the HIP
library; importing this package never falls back to a CPU path, and every
call raises :class:`NNPError` when the library or a GPU is missing.
"""
import ctypes
import os

import numpy as np

__all__ = ["NNPError", "Mode", "library", "PLAIN_FP64", "MIXED", "FORCES_PULL", "FORCES_PUSH",
           "read_input_data", "remap", "write_output_data"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnnpx.so")

PLAIN_FP64 = 0
MIXED = 1
FORCES_PULL = 0
FORCES_PUSH = 1
OPT_FORCE_MODE = 0
OPT_TIMING = 1
OPT_NEIGHBOR_ALGO = 2
OPT_BF16_TERMS = 3
OPT_SF_GATHER = 4
OPT_PLAIN_FORCES = 5
OPT_ANGULAR_MERGE_SF = 6
OPT_ANGULAR_MERGE_FORCE = 7
OPT_MD_STAGING = 8
OPT_PAIR_LISTS = 9
EVAL_FORCES = 1

_STATUS = {0: "OK", 1: "ARGUMENT", 2: "FILE", 3: "MODEL", 4: "RANGE", 5: "STATE", 6: "HIP", 7: "UNSUPPORTED"}

_lib = None


class NNPError(RuntimeError):
    """A non-zero nnpx_status; ``status`` holds the code."""

    def __init__(self, status, msg):
        super().__init__("%s: %s" % (_STATUS.get(status, status), msg))
        self.status = status


_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_c_i64_p = ctypes.POINTER(ctypes.c_int64)

_SIGS = {
    "nnpx_create": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int],
    "nnpx_destroy": [ctypes.c_void_p],
    "nnpx_set_option": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64],
    "nnpx_get_option": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
    "nnpx_synchronize": [ctypes.c_void_p],
    "nnpx_load_model": [ctypes.c_void_p, ctypes.c_char_p],
    "nnpx_model_info": [ctypes.c_void_p, _c_int_p, _c_int_p, _c_dbl_p],
    "nnpx_num_sf": [ctypes.c_void_p, ctypes.c_int, _c_int_p],
    "nnpx_atomic_number": [ctypes.c_void_p, ctypes.c_int, _c_int_p],
    "nnpx_element_index": [ctypes.c_void_p, ctypes.c_char_p, _c_int_p],
    "nnpx_model_units": [ctypes.c_void_p, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p],
    "nnpx_set_structures": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_update_positions": [ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_calculate_neighbor_list": [ctypes.c_void_p],
    "nnpx_calculate_symmetry_function_groups": [ctypes.c_void_p, ctypes.c_int],
    "nnpx_calculate_atomic_neural_networks": [ctypes.c_void_p, ctypes.c_int],
    "nnpx_calculate_energy": [ctypes.c_void_p],
    "nnpx_calculate_forces": [ctypes.c_void_p],
    "nnpx_evaluate": [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint],
    "nnpx_get_energy": [ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_get_atomic_energies": [ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_get_forces": [ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_get_G": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int],
    "nnpx_get_dEdG": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int],
    "nnpx_get_neighbors": [ctypes.c_void_p, _c_i64_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_get_dGdr": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int],
    "nnpx_kernel_time": [ctypes.c_void_p, ctypes.c_char_p, _c_dbl_p, _c_i64_p],
    "nnpx_reset_kernel_times": [ctypes.c_void_p],
    "nnpx_work_counts": [ctypes.c_void_p, _c_i64_p, _c_i64_p, _c_i64_p, _c_i64_p, _c_i64_p],
    "nnpx_mfma_selftest": [ctypes.c_void_p, _c_dbl_p],
    "nnpx_element_symbol": [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int],
    "nnpx_num_sf_derivatives": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_int_p],
    "nnpx_md_init": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                     ctypes.c_double],
    "nnpx_md_run": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int],
    "nnpx_num_connections": [ctypes.c_void_p, ctypes.c_int, _c_int_p],
    "nnpx_autotune": [ctypes.c_void_p],
    "nnpx_get_tuning": [ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_set_tuning": [ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_set_neighbors_external": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_get_weights": [ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_extrapolation_warnings": [ctypes.c_void_p, ctypes.c_int64, _c_i64_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_low_neighbor_atoms": [ctypes.c_void_p, ctypes.c_int64, _c_i64_p, ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_kalman_init": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
    "nnpx_kalman_state": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_kalman_update": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                           ctypes.c_int, ctypes.c_void_p],
    "nnpx_set_weights": [ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_loss_gradient": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                           ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_weight_vjp": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_md_state": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "nnpx_md_energies": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, _c_i64_p, _c_i64_p],
}


def library():
    """Load libnnpx.so (built by ``__graft_entry__.build()``); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NNPError(6, "HIP library %s is not built (run __graft_entry__.build())" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        lib.nnpx_last_error.argtypes = [ctypes.c_void_p]
        lib.nnpx_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Mode:
    """One GPU context with a loaded model: the nnp::Mode compute calls.

    Method names follow nnp::Mode (Mode.h:313-387): calculate_neighbor_list,
    calculate_symmetry_function_groups, calculate_atomic_neural_networks,
    calculate_energy, calculate_forces and evaluate (evaluateNNP).
    """

    def __init__(self, model_dir=None, device=0):
        self._lib = library()
        h = ctypes.c_void_p()
        st = self._lib.nnpx_create(ctypes.byref(h), device)
        if st:
            raise NNPError(st, self._lib.nnpx_last_error(None).decode())
        self._h = h
        self.n_atoms = 0
        self.n_struct = 0
        if model_dir is not None:
            self.load_model(model_dir)

    def _ck(self, st):
        if st:
            raise NNPError(st, self._lib.nnpx_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nnpx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- model ---------------------------------------------------------
    def load_model(self, model_dir):
        self._ck(self._lib.nnpx_load_model(self._h, os.fsencode(model_dir)))
        ne, ms, rc = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._ck(self._lib.nnpx_model_info(self._h, ctypes.byref(ne), ctypes.byref(ms), ctypes.byref(rc)))
        self.num_elements, self.max_sf, self.max_cutoff = ne.value, ms.value, rc.value
        self.num_sf = []
        self.atomic_numbers = []
        for e in range(self.num_elements):
            v = ctypes.c_int()
            self._ck(self._lib.nnpx_num_sf(self._h, e, ctypes.byref(v)))
            self.num_sf.append(v.value)
            self._ck(self._lib.nnpx_atomic_number(self._h, e, ctypes.byref(v)))
            self.atomic_numbers.append(v.value)
        norm = ctypes.c_int()
        me, ce, cl = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        off = np.zeros(self.num_elements)
        self._ck(self._lib.nnpx_model_units(self._h, ctypes.byref(norm), ctypes.byref(me), ctypes.byref(ce),
                                            ctypes.byref(cl), off.ctypes.data_as(_c_dbl_p)))
        self.normalize = bool(norm.value)
        self.mean_energy, self.conv_energy, self.conv_length = me.value, ce.value, cl.value
        self.energy_offsets = off

    def element_index(self, symbol):
        v = ctypes.c_int()
        self._ck(self._lib.nnpx_element_index(self._h, symbol.encode(), ctypes.byref(v)))
        return v.value

    def set_option(self, option, value):
        self._ck(self._lib.nnpx_set_option(self._h, option, int(value)))

    def get_option(self, option):
        v = np.zeros(1, np.int64)
        self._ck(self._lib.nnpx_get_option(self._h, option, _ptr(v)))
        return int(v[0])

    # -- structures ----------------------------------------------------
    def set_structures(self, positions, elements, boxes=None, atom_offsets=None, periodic=None):
        """Set one or many structures (model units, already remapped).

        positions: (N, 3) float64; elements: (N,) element indices;
        boxes: (S, 3, 3) lattice vectors as rows or None (isolated);
        atom_offsets: (S+1,) first atom of each structure (default one structure).
        """
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        el = np.ascontiguousarray(elements, dtype=np.int32).reshape(-1)
        n = pos.shape[0]
        if atom_offsets is None:
            atom_offsets = [0, n]
        off = np.ascontiguousarray(atom_offsets, dtype=np.int64)
        ns = len(off) - 1
        if boxes is None:
            box = np.zeros((ns, 9))
            per = np.zeros(ns, dtype=np.int32) if periodic is None else np.asarray(periodic, np.int32)
        else:
            box = np.ascontiguousarray(boxes, dtype=np.float64).reshape(ns, 9)
            per = np.ones(ns, dtype=np.int32) if periodic is None else np.ascontiguousarray(periodic, np.int32)
        self._ck(self._lib.nnpx_set_structures(self._h, ns, _ptr(off), _ptr(pos), _ptr(el), _ptr(box), _ptr(per)))
        self.n_atoms, self.n_struct, self.n_force = n, ns, None
        self._pos_shape = pos.shape

    def set_neighbors_external(self, elem_local, n_total, offsets, index, nb_elem, dr):
        """Caller-provided neighbour lists (the LAMMPS path): local atoms'
        elements, total atom count (local + ghosts), per-atom entry offsets,
        neighbour indices, their elements and distance vectors r_i - r_j."""
        e = np.ascontiguousarray(elem_local, dtype=np.int32)
        o = np.ascontiguousarray(offsets, dtype=np.int64)
        j = np.ascontiguousarray(index, dtype=np.int32)
        ne = np.ascontiguousarray(nb_elem, dtype=np.int32)
        d = np.ascontiguousarray(dr, dtype=np.float64).reshape(-1)
        self._ck(self._lib.nnpx_set_neighbors_external(self._h, len(e), _ptr(e), int(n_total), _ptr(o), _ptr(j),
                                                       _ptr(ne), _ptr(d)))
        self.n_atoms, self.n_struct, self.n_force = len(e), 1, int(n_total)

    def update_positions(self, positions):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(self._pos_shape)
        self._ck(self._lib.nnpx_update_positions(self._h, _ptr(pos)))

    # -- compute (nnp::Mode) -------------------------------------------
    def calculate_neighbor_list(self):
        self._ck(self._lib.nnpx_calculate_neighbor_list(self._h))

    def calculate_symmetry_function_groups(self, derivatives=True):
        self._ck(self._lib.nnpx_calculate_symmetry_function_groups(self._h, int(derivatives)))

    def calculate_atomic_neural_networks(self, derivatives=True):
        self._ck(self._lib.nnpx_calculate_atomic_neural_networks(self._h, int(derivatives)))

    def calculate_energy(self):
        self._ck(self._lib.nnpx_calculate_energy(self._h))

    def calculate_forces(self):
        self._ck(self._lib.nnpx_calculate_forces(self._h))

    def evaluate(self, precision=PLAIN_FP64, forces=True):
        self._ck(self._lib.nnpx_evaluate(self._h, precision, EVAL_FORCES if forces else 0))

    def synchronize(self):
        self._ck(self._lib.nnpx_synchronize(self._h))

    # -- results -------------------------------------------------------
    def energies(self):
        e = np.zeros(self.n_struct)
        self._ck(self._lib.nnpx_get_energy(self._h, _ptr(e)))
        return e

    def atomic_energies(self):
        e = np.zeros(self.n_atoms)
        self._ck(self._lib.nnpx_get_atomic_energies(self._h, _ptr(e)))
        return e

    def forces(self):
        f = np.zeros((getattr(self, "n_force", None) or self.n_atoms, 3))
        self._ck(self._lib.nnpx_get_forces(self._h, _ptr(f)))
        return f

    def G(self):
        g = np.zeros((self.n_atoms, self.max_sf))
        self._ck(self._lib.nnpx_get_G(self._h, _ptr(g), self.max_sf))
        return g

    def dEdG(self):
        g = np.zeros((self.n_atoms, self.max_sf))
        self._ck(self._lib.nnpx_get_dEdG(self._h, _ptr(g), self.max_sf))
        return g

    def neighbors(self):
        """(counts, j, shifts) of the neighbour list in list order."""
        tot = ctypes.c_int64()
        self._ck(self._lib.nnpx_get_neighbors(self._h, ctypes.byref(tot), None, None, None))
        counts = np.zeros(self.n_atoms, np.int32)
        j = np.zeros(tot.value, np.int32)
        sh = np.zeros((tot.value, 3), np.int32)
        self._ck(self._lib.nnpx_get_neighbors(self._h, ctypes.byref(tot), _ptr(counts), _ptr(j), _ptr(sh)))
        return counts, j, sh

    def dGdr(self, ld_nbr):
        """Stored dG/dr: (self [N, max_sf, 3], neighbour entries [P, ld_nbr, 3])."""
        counts, _, _ = self.neighbors()
        s = np.zeros((self.n_atoms, self.max_sf, 3))
        nb = np.zeros((int(counts.sum()), ld_nbr, 3))
        self._ck(self._lib.nnpx_get_dGdr(self._h, _ptr(s), self.max_sf, _ptr(nb), ld_nbr))
        return s, nb

    def extrapolation_warnings(self):
        """Extrapolation warnings of the last evaluation
        (Element::updateSymmetryFunctionStatistics): a structured array of
        (atom, sf, type, value, gmin, gmax) in (atom, SF) order."""
        n = ctypes.c_int64()
        self._ck(self._lib.nnpx_extrapolation_warnings(self._h, 0, ctypes.byref(n), None, None, None, None, None,
                                                         None))
        k = n.value
        a, q, t = (np.zeros(k, np.int32) for _ in range(3))
        v, lo, hi = (np.zeros(k) for _ in range(3))
        if k:
            self._ck(self._lib.nnpx_extrapolation_warnings(self._h, k, ctypes.byref(n), _ptr(a), _ptr(q), _ptr(t),
                                                             _ptr(v), _ptr(lo), _ptr(hi)))
        out = np.zeros(k, dtype=[("atom", np.int32), ("sf", np.int32), ("type", np.int32), ("value", float),
                                 ("gmin", float), ("gmax", float)])
        out["atom"], out["sf"], out["type"], out["value"], out["gmin"], out["gmax"] = a, q, t, v, lo, hi
        return out

    def low_neighbor_atoms(self):
        """(atom indices, neighbour counts) of the atoms with fewer neighbours
        within their element's smallest SF cutoff than their SFs need."""
        n = ctypes.c_int64()
        self._ck(self._lib.nnpx_low_neighbor_atoms(self._h, 0, ctypes.byref(n), None, None))
        a = np.zeros(n.value, np.int32)
        c = np.zeros(n.value, np.int32)
        if n.value:
            self._ck(self._lib.nnpx_low_neighbor_atoms(self._h, n.value, ctypes.byref(n), _ptr(a), _ptr(c)))
        return a, c

    # -- Kalman filter training updates (Training::update) ------------------
    def kalman_init(self, kalman_type=0, per_element=False, epsilon=1e-2, q0=0.0, qtau=0.0, qmin=0.0,
                    eta=1e-2, etatau=0.0, etamax=1e-2, lam=0.99, nu=0.995):
        """Kalman filter updaters: type 0 standard (eta, etatau, etamax), 1
        fading memory (lambda, nu); one filter or one per element."""
        par = np.array([epsilon, q0, qtau, qmin] + ([eta, etatau, etamax] if kalman_type == 0 else [lam, nu, 0.0]),
                       np.float64)
        self._ck(self._lib.nnpx_kalman_init(self._h, kalman_type, 1 if per_element else 0, _ptr(par)))

    def kalman_state(self, updater=0, with_P=True):
        sc = np.zeros(7)
        self._ck(self._lib.nnpx_kalman_state(self._h, updater, None, _ptr(sc)))
        n = int(sc[0])
        P = None
        if with_P:
            P = np.zeros((n, n))
            self._ck(self._lib.nnpx_kalman_state(self._h, updater, _ptr(P), _ptr(sc)))
        return P, {"n": n, "offset": int(sc[1]), "eta": sc[2], "lambda": sc[3], "gamma": sc[4], "q": sc[5],
                   "updates": int(sc[6])}

    def kalman_update(self, prop, structures, E_ref=None, F_ref=None, atoms=None, components=None,
                      force_weight=1.0, full=False, precision=PLAIN_FP64):
        """One Training::update for "energy" or "force" with the given batch
        of candidates; returns the candidates' errors before the update."""
        b = len(structures)
        st = np.ascontiguousarray(structures, np.int32)
        at = None if atoms is None else np.ascontiguousarray(atoms, np.int32)
        co = None if components is None else np.ascontiguousarray(components, np.int32)
        er = None if E_ref is None else np.ascontiguousarray(E_ref, np.float64)
        fr = None if F_ref is None else np.ascontiguousarray(F_ref, np.float64).reshape(-1)
        err = np.zeros(b)
        self._ck(self._lib.nnpx_kalman_update(self._h, precision, 0 if prop == "energy" else 1, b, _ptr(st), _ptr(at),
                                              _ptr(co), _ptr(er), _ptr(fr), float(force_weight), 1 if full else 0,
                                              _ptr(err)))
        return err

    def kernel_time(self, name):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        self._ck(self._lib.nnpx_kernel_time(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def reset_kernel_times(self):
        self._ck(self._lib.nnpx_reset_kernel_times(self._h))

    def mfma_selftest(self):
        v = ctypes.c_double()
        self._ck(self._lib.nnpx_mfma_selftest(self._h, ctypes.byref(v)))
        return v.value

    # ------------------------------------------------------ autotuner --
    def autotune(self):
        """Time candidate launch shapes on the current structures and keep the
        fastest; returns the chosen {trip_sf, trip_force, trip_tangent, grad_atoms}."""
        self._ck(self._lib.nnpx_autotune(self._h))
        return self.tuning()

    def tuning(self):
        """Launch shapes in use: the triplet buffers, the gradient kernel's
        atoms per block, the merged angular walk and pair list settings
        (-1 auto)."""
        v = np.zeros(4, np.int32)
        self._ck(self._lib.nnpx_get_tuning(self._h, _ptr(v)))
        out = dict(zip(("trip_sf", "trip_force", "trip_tangent", "grad_atoms"), (int(x) for x in v)))
        out["merge_sf"] = self.get_option(OPT_ANGULAR_MERGE_SF)
        out["merge_force"] = self.get_option(OPT_ANGULAR_MERGE_FORCE)
        out["pair_lists"] = self.get_option(OPT_PAIR_LISTS)
        return out

    def set_tuning(self, trip_sf, trip_force, trip_tangent, grad_atoms):
        v = np.array([trip_sf, trip_force, trip_tangent, grad_atoms], np.int32)
        self._ck(self._lib.nnpx_set_tuning(self._h, _ptr(v)))

    # ------------------------------------------------------ training --
    def num_connections(self, element=-1):
        n = ctypes.c_int()
        self._ck(self._lib.nnpx_num_connections(self._h, element, ctypes.byref(n)))
        return n.value

    def weights(self):
        w = np.zeros(self.num_connections())
        self._ck(self._lib.nnpx_get_weights(self._h, _ptr(w)))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        if len(w) != self.num_connections():
            raise NNPError(1, "wrong number of connections")
        self._ck(self._lib.nnpx_set_weights(self._h, _ptr(w)))

    def loss_gradient(self, E_ref, F_ref, beta, precision=PLAIN_FP64):
        """(loss[3], gradient) of sum_s (E_s - E_ref)^2/N_s + beta/(3N) sum (F - F_ref)^2."""
        E = np.ascontiguousarray(E_ref, dtype=np.float64).reshape(-1)
        F = np.ascontiguousarray(F_ref, dtype=np.float64).reshape(-1)
        if len(E) != self.n_struct or len(F) != 3 * self.n_atoms:
            raise NNPError(1, "targets have the wrong size")
        loss = np.zeros(3)
        g = np.zeros(self.num_connections())
        self._ck(self._lib.nnpx_loss_gradient(self._h, precision, _ptr(E), _ptr(F), float(beta), _ptr(loss),
                                              _ptr(g)))
        return loss, g

    def weight_vjp(self, vE, vF, precision=PLAIN_FP64):
        """Per-structure weight derivatives [n_struct, n_conn]:
        vE_s dE_s/dw + sum_{a in s} vF_a . dF_a/dw (Training::dPdc rows)."""
        E = np.ascontiguousarray(vE, dtype=np.float64).reshape(-1)
        F = np.ascontiguousarray(vF, dtype=np.float64).reshape(-1)
        if len(E) != self.n_struct or len(F) != 3 * self.n_atoms:
            raise NNPError(1, "seeds have the wrong size")
        g = np.zeros((self.n_struct, self.num_connections()))
        self._ck(self._lib.nnpx_weight_vjp(self._h, precision, _ptr(E), _ptr(F), _ptr(g)))
        return g

    # ------------------------------------------------------------ MD --
    def md_init(self, masses, velocities, dt, skin, precision=MIXED):
        """Velocity Verlet set-up: masses per element (amu), velocities
        [N, 3] (A/fs), dt (fs), list skin (A); computes the initial forces."""
        m = np.ascontiguousarray(masses, dtype=np.float64)
        v = np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1)
        if len(m) != self.num_elements or len(v) != 3 * self.n_atoms:
            raise NNPError(1, "masses or velocities have the wrong size")
        self._ck(self._lib.nnpx_md_init(self._h, precision, _ptr(m), _ptr(v), float(dt), float(skin)))

    def md_run(self, nsteps, graph_steps=0):
        self._ck(self._lib.nnpx_md_run(self._h, int(nsteps), int(graph_steps)))

    def md_state(self):
        p = np.zeros((self.n_atoms, 3))
        v = np.zeros((self.n_atoms, 3))
        self._ck(self._lib.nnpx_md_state(self._h, _ptr(p), _ptr(v)))
        return p, v

    def md_energies(self, max_steps=1 << 16):
        ep = np.zeros(max_steps)
        ek = np.zeros(max_steps)
        n, r = ctypes.c_int64(), ctypes.c_int64()
        self._ck(self._lib.nnpx_md_energies(self._h, _ptr(ep), _ptr(ek), max_steps, ctypes.byref(n),
                                            ctypes.byref(r)))
        k = min(n.value, max_steps)
        return ep[:k], ek[:k], r.value

    def work_counts(self):
        v = [ctypes.c_int64() for _ in range(5)]
        self._ck(self._lib.nnpx_work_counts(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(["pairs", "triplets", "pair_members", "triplet_members", "triplet_exps"],
                        [x.value for x in v]))


# ---------------------------------------------------------------- files --
def read_structures(path, symbols):
    """All structures of an n2p2 input.data file (Structure::readFromLines,
    Structure.cpp:142-271) without the remap: a list of (pos, elem, box or
    None)."""
    out = []
    index = {s: i for i, s in enumerate(symbols)}
    pos = el = box = None
    with open(path) as f:
        for line in f:
            w = line.split()
            if not w:
                continue
            if w[0] == "begin":
                pos, el, box = [], [], []
            elif pos is None:
                raise ValueError("expected 'begin'")
            elif w[0] == "lattice":
                box.append([float(x) for x in w[1:4]])
            elif w[0] == "atom":
                pos.append([float(x) for x in w[1:4]])
                el.append(index[w[4]])
            elif w[0] == "end":
                if box and len(box) != 3:
                    raise ValueError("strange number of box vectors")
                out.append((np.array(pos), np.array(el, np.int32), (np.array(box) if box else None)))
                pos = None
    return out


def read_input_data(path, symbols):
    """First structure of an n2p2 input.data file: (pos, elem, box or None)."""
    s = read_structures(path, symbols)
    if not s:
        raise ValueError("no structure in %s" % path)
    return s[0]


def inverse_box(box):
    """Structure::calculateInverseBox (Structure.cpp:518-544), same operation order."""
    b = box
    invdet = (b[0][0] * b[1][1] * b[2][2] + b[1][0] * b[2][1] * b[0][2] + b[2][0] * b[0][1] * b[1][2]
              - b[2][0] * b[1][1] * b[0][2] - b[0][0] * b[2][1] * b[1][2] - b[1][0] * b[0][1] * b[2][2])
    invdet = 1.0 / invdet
    ib = np.zeros((3, 3))
    ib[0, 0] = b[1][1] * b[2][2] - b[2][1] * b[1][2]
    ib[0, 1] = b[2][1] * b[0][2] - b[0][1] * b[2][2]
    ib[0, 2] = b[0][1] * b[1][2] - b[1][1] * b[0][2]
    ib[1, 0] = b[2][0] * b[1][2] - b[1][0] * b[2][2]
    ib[1, 1] = b[0][0] * b[2][2] - b[2][0] * b[0][2]
    ib[1, 2] = b[1][0] * b[0][2] - b[0][0] * b[1][2]
    ib[2, 0] = b[1][0] * b[2][1] - b[2][0] * b[1][1]
    ib[2, 1] = b[2][0] * b[0][1] - b[0][0] * b[2][1]
    ib[2, 2] = b[0][0] * b[1][1] - b[1][0] * b[0][1]
    return ib * invdet


def remap(pos, box):
    """Structure::remap (Structure.cpp:1224-1248) elementwise, same rounding;
    positions of an isolated structure (box None) are returned unchanged."""
    if box is None:
        return np.array(pos, dtype=np.float64)
    ib = inverse_box(box)
    r = np.asarray(pos, dtype=np.float64)
    f = r[:, 0:1] * ib[0] + r[:, 1:2] * ib[1] + r[:, 2:3] * ib[2]
    f = np.where(f > 1.0, f - np.trunc(f), f)
    f = np.where(f < 0.0, f + (1.0 - np.trunc(f)), f)
    f = np.where(f == 1.0, 0.0, f)
    return f[:, 0:1] * box[0] + f[:, 1:2] * box[1] + f[:, 2:3] * box[2]


def write_output_data(path, pos, elem, symbols, forces, energy, box=None, comment=""):
    """Structure::writeToFile (Structure.cpp:1475-1532) for predicted results."""
    with open(path, "w") as f:
        f.write("begin\ncomment %s\n" % comment)
        if box is not None:
            for v in box:
                f.write("lattice %24.16E %24.16E %24.16E\n" % tuple(v))
        for r, e, fo in zip(pos, elem, forces):
            f.write("atom %24.16E %24.16E %24.16E %2s %24.16E %24.16E %24.16E %24.16E %24.16E\n"
                    % (r[0], r[1], r[2], symbols[e], 0.0, 0.0, fo[0], fo[1], fo[2]))
        f.write("energy %24.16E\ncharge %24.16E\nend\n" % (energy, 0.0))
