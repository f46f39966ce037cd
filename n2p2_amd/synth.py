"""Seeded synthetic workloads and model directories for the benchmark configs.

Every generator is deterministic for its seed.  Model directories are written
in n2p2's file formats (input.nn, scaling.data, weights.%03d.data) so the same
directory loads in the GPU library and in n2p2-compatible tools.

Workloads (sizes as named by the benchmark configurations):
  water_box(4)        64 H2O, 192 atoms, 12.42 A cell         (water-192)
  water_box(16)       4,096 H2O, 12,288 atoms, 49.68 A cell    (water-12288)
  random_batch()      512 periodic structures, 16..512 atoms   (batch-512)
  fcc_lattice()       32^3 FCC cells, 131,072 atoms, 4 species (fcc-131072)
  water_set()         256 water-like structures, 66..255 atoms (water-set-256)

The scaling statistics of a model (nnp-scaling semantics: per SF min, max,
mean, sigma of the unscaled values) are computed here in numpy over a seeded
sample structure; this is model preparation, not part of the timed path.
"""
import math
import os

import numpy as np

from . import remap

SEED_WATER = 20240601
SEED_BATCH = 20240602
SEED_FCC = 20240603
SEED_WATERSET = 20240604
SEED_WEIGHTS = 20240610
SEED_TEACHER = 20240611
SEED_VELOCITY = 20240612

ETA_RAD = np.geomspace(0.001, 0.5, 10)


# ------------------------------------------------------------ structures --
def _rotations(rng, n):
    """Uniform random rotation matrices (from uniform unit quaternions)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], 1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)], 1)


def _water_molecules(rng, centers, jitter=0.1):
    """O,H,H per molecule: O at the site, standard geometry (O-H 0.9572 A,
    HOH 104.52 deg) rotated uniformly, N(0, jitter^2) noise on every coordinate.
    Elements: H = 0, O = 1."""
    roh, ang = 0.9572, math.radians(104.52)
    local = np.array([[0.0, 0.0, 0.0],
                      [roh * math.sin(ang / 2), 0.0, roh * math.cos(ang / 2)],
                      [-roh * math.sin(ang / 2), 0.0, roh * math.cos(ang / 2)]])
    R = _rotations(rng, len(centers))
    pos = centers[:, None, :] + np.einsum("mij,aj->mai", R, local)
    pos = pos + rng.normal(0.0, jitter, pos.shape)
    elem = np.tile(np.array([1, 0, 0], np.int32), len(centers))
    return pos.reshape(-1, 3), elem


def water_box(n_side=4, spacing=3.105, seed=SEED_WATER):
    """n_side^3 molecules on a simple-cubic lattice, cubic periodic cell."""
    rng = np.random.default_rng(seed)
    g = np.arange(n_side) * spacing
    centers = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos, elem = _water_molecules(rng, centers)
    box = np.eye(3) * (n_side * spacing)
    return remap(pos, box), elem, box


def random_batch(n_struct=512, nmin=16, nmax=512, density=0.1, dmin=0.9, seed=SEED_BATCH):
    """Uneven periodic structures: N ~ U{nmin..nmax}, floor(N/3) atoms of
    element 1 (O) and the rest element 0 (H), cubic L = (N/density)^(1/3),
    uniform positions with periodic minimum-image rejection at dmin."""
    rng = np.random.default_rng(seed)
    pos_all, el_all, boxes, off = [], [], [], [0]
    for _ in range(n_struct):
        n = int(rng.integers(nmin, nmax + 1))
        L = (n / density) ** (1.0 / 3.0)
        pts = np.zeros((0, 3))
        while len(pts) < n:
            p = rng.random(3) * L
            if len(pts):
                d = pts - p
                d -= L * np.round(d / L)
                if np.min(np.einsum("ij,ij->i", d, d)) < dmin * dmin:
                    continue
            pts = np.vstack([pts, p])
        el = np.zeros(n, np.int32)
        el[: n // 3] = 1
        box = np.eye(3) * L
        pos_all.append(remap(pts, box))
        el_all.append(el)
        boxes.append(box)
        off.append(off[-1] + n)
    return np.vstack(pos_all), np.concatenate(el_all), np.array(boxes), np.array(off, np.int64)


def fcc_lattice(n_cells=32, a=3.6, n_species=4, jitter=0.05, seed=SEED_FCC):
    """n_cells^3 conventional FCC cells, species uniform at random."""
    rng = np.random.default_rng(seed)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.arange(n_cells)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 1, 3)
    pos = ((cells + basis[None]) * a).reshape(-1, 3)
    pos = pos + rng.normal(0.0, jitter, pos.shape)
    elem = rng.integers(0, n_species, len(pos)).astype(np.int32)
    box = np.eye(3) * (n_cells * a)
    return remap(pos, box), elem, box


def water_set(n_struct=256, mmin=22, mmax=85, density=0.1, seed=SEED_WATERSET):
    """Water-like structures: N_mol ~ U{mmin..mmax} molecules on the sites of
    a k^3 grid (k = ceil(N_mol^(1/3))) in a cubic cell at `density` atoms/A^3."""
    rng = np.random.default_rng(seed)
    pos_all, el_all, boxes, off = [], [], [], [0]
    for _ in range(n_struct):
        m = int(rng.integers(mmin, mmax + 1))
        L = (3 * m / density) ** (1.0 / 3.0)
        k = int(math.ceil(m ** (1.0 / 3.0) - 1e-9))
        g = (np.arange(k) + 0.5) * (L / k)
        sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
        centers = sites[np.sort(rng.choice(len(sites), m, replace=False))]
        p, el = _water_molecules(rng, centers)
        box = np.eye(3) * L
        pos_all.append(remap(p, box))
        el_all.append(el)
        boxes.append(box)
        off.append(off[-1] + len(el))
    return np.vstack(pos_all), np.concatenate(el_all), np.array(boxes), np.array(off, np.int64)


# ------------------------------------------------------------- SF models --
def water_sf_lines(rc=6.35):
    """Per element 20 G2 (2 neighbour elements x 10 eta, rs = 0) and 30 G3
    (HH, HO, OO x [lambda +-1 x zeta 1,2,4,16 at eta 0.001; lambda +-1,
    zeta 1 at eta 0.01])."""
    lines = []
    for c in ("H", "O"):
        for n in ("H", "O"):
            for eta in ETA_RAD:
                lines.append("%s 2 %s %.10e 0.0 %.4f" % (c, n, eta, rc))
        for a, b in (("H", "H"), ("H", "O"), ("O", "O")):
            for lam in (1.0, -1.0):
                for zeta in (1.0, 2.0, 4.0, 16.0):
                    lines.append("%s 3 %s %s 0.001 %.1f %.1f %.4f" % (c, a, b, lam, zeta, rc))
            for lam in (1.0, -1.0):
                lines.append("%s 3 %s %s 0.01 %.1f 1.0 %.4f" % (c, a, b, lam, rc))
    return lines


FCC_SYMBOLS = ("Al", "Ni", "Cu", "Pd")


def fcc_sf_lines(rc=6.0, symbols=FCC_SYMBOLS):
    """Per element 40 G2 (4 neighbour elements x 10 eta) and 40 G3 (10
    unordered element pairs x lambda +-1 x zeta 1,4 at eta 0.001)."""
    lines = []
    for c in symbols:
        for n in symbols:
            for eta in ETA_RAD:
                lines.append("%s 2 %s %.10e 0.0 %.4f" % (c, n, eta, rc))
        for ia in range(len(symbols)):
            for ib in range(ia, len(symbols)):
                for lam in (1.0, -1.0):
                    for zeta in (1.0, 4.0):
                        lines.append("%s 3 %s %s 0.001 %.1f %.1f %.4f" % (c, symbols[ia], symbols[ib], lam, zeta, rc))
    return lines


ATOMIC_NUMBER = {"H": 1, "O": 8, "Al": 13, "Ni": 28, "Cu": 29, "Pd": 46}
# standard atomic weights (amu) used as MD masses
MASS = {"H": 1.008, "O": 15.999, "Al": 26.9815, "Ni": 58.6934, "Cu": 63.546, "Pd": 106.42}
KB_EV = 8.617333262e-5          # Boltzmann constant, eV/K
ACCEL = 9.64853321e-3           # (eV/A)/amu in A/fs^2


def maxwell_velocities(masses, temperature=300.0, seed=SEED_VELOCITY):
    """Seeded Maxwell-Boltzmann velocities (A/fs) for per-atom masses (amu),
    net momentum removed."""
    rng = np.random.default_rng(seed)
    m = np.asarray(masses, dtype=np.float64)
    v = rng.normal(size=(len(m), 3)) * np.sqrt(KB_EV * temperature * ACCEL / m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    return v


def _parse(line, symbols):
    w = line.split()
    idx = {s: i for i, s in enumerate(symbols)}
    t = int(w[1])
    if t == 12:
        return dict(type=12, ec=idx[w[0]], e1=-1, e2=-1, eta=float(w[2]), rs=float(w[3]), rc=float(w[4]),
                    lam=0.0, zeta=0.0)
    if t == 13:
        return dict(type=13, ec=idx[w[0]], e1=-1, e2=-1, eta=float(w[2]), rs=float(w[3]), lam=float(w[4]),
                    zeta=float(w[5]), rc=float(w[6]))
    if t == 23:
        return dict(type=23, ec=idx[w[0]], e1=-1, e2=-1, rl=float(w[2]), rc=float(w[3]), sub=w[4], al=0.0, ar=0.0)
    if t in (24, 25):
        return dict(type=t, ec=idx[w[0]], e1=-1, e2=-1, rl=float(w[2]), rc=float(w[3]), al=float(w[4]),
                    ar=float(w[5]), sub=w[6])
    if t == 20:
        return dict(type=20, ec=idx[w[0]], e1=idx[w[2]], e2=idx[w[2]], rl=float(w[3]), rc=float(w[4]),
                    sub=w[5], al=0.0, ar=0.0)
    if t in (21, 22):
        e1, e2 = sorted((idx[w[2]], idx[w[3]]))
        return dict(type=t, ec=idx[w[0]], e1=e1, e2=e2, rl=float(w[4]), rc=float(w[5]), al=float(w[6]),
                    ar=float(w[7]), sub=w[8])
    if t == 2:
        return dict(type=2, ec=idx[w[0]], e1=idx[w[2]], e2=idx[w[2]], eta=float(w[3]), rs=float(w[4]),
                    rc=float(w[5]), lam=0.0, zeta=0.0)
    e1, e2 = sorted((idx[w[2]], idx[w[3]]))
    return dict(type=3, ec=idx[w[0]], e1=e1, e2=e2, eta=float(w[4]), lam=float(w[5]), zeta=float(w[6]),
                rc=float(w[7]), rs=float(w[8]) if len(w) > 8 else 0.0)


def _sort_key(s):
    if s["type"] >= 23:  # weighted compact: subtype, rc, rl, angles
        return (s["type"], s["sub"], s["rc"], s["rl"], s["al"], s["ar"])
    if s["type"] in (12, 13):  # weighted Gaussian: rc, eta, rs, zeta, lambda
        return (s["type"], s["rc"], s["eta"], s["rs"], s["zeta"], s["lam"])
    if s["type"] >= 20:  # SymFncCompRad.cpp:48-68, SymFncCompAngw.cpp:56-82
        return (s["type"], s["sub"], s["e1"], s["e2"], s["rc"], s["rl"], s["al"], s["ar"])
    if s["type"] == 2:
        return (s["type"], s["rc"], s["eta"], s["rs"], s["e1"])
    return (s["type"], s["rc"], s["eta"], s["rs"], s["zeta"], s["lam"], s["e1"], s["e2"])


def sorted_sfs(lines, symbols):
    """SFs per element in n2p2's sorted order (SymFncExpRad/SymFncExpAngn
    operator<, SymFncExpRad.cpp:57-75, SymFncExpAngn.cpp:51-77; compact
    types by subtype, elements, rc, rl, angles)."""
    sfs = [_parse(l, symbols) for l in lines]
    out = []
    for e in range(len(symbols)):
        mine = [s for s in sfs if s["ec"] == e]
        mine.sort(key=_sort_key)
        out.append(mine)
    return out


def _cutoff_cos(r, rc):
    return np.where(r < rc, 0.5 * (np.cos(np.pi * r / rc) + 1.0), 0.0)


def sf_values(sfs, pos, elem, box, atoms=None):
    """Unscaled G2/G3 values (CT_COS, alpha = 0) for `atoms` (default all),
    numpy, brute-force images.  Used only to derive scaling statistics."""
    rcmax = max(s["rc"] for e in sfs for s in e)
    ib = np.linalg.inv(box)
    n = len(pos)
    atoms = range(n) if atoms is None else atoms
    proj = [abs(np.dot(box[k], np.cross(box[(k + 1) % 3], box[(k + 2) % 3])))
            / np.linalg.norm(np.cross(box[(k + 1) % 3], box[(k + 2) % 3])) for k in range(3)]
    p = [int(math.floor(rcmax / proj[k])) + 1 for k in range(3)]
    shifts = np.array([[a, b, c] for a in range(-p[0], p[0] + 1) for b in range(-p[1], p[1] + 1)
                       for c in range(-p[2], p[2] + 1)], dtype=float) @ box
    out = {}
    del ib
    for i in atoms:
        dr = (pos[i] - pos)[None, :, :] + shifts[:, None, :]
        dr = dr.reshape(-1, 3)
        ej = np.tile(elem, len(shifts))
        d = np.sqrt(np.einsum("ij,ij->i", dr, dr))
        keep = (d <= rcmax) & (d > 0)
        dr, d, ej = dr[keep], d[keep], ej[keep]
        vals = []
        for s in sfs[elem[i]]:
            fc = _cutoff_cos(d, s["rc"])
            if s["type"] == 2:
                m = ej == s["e1"]
                vals.append(np.sum(np.exp(-s["eta"] * (d[m] - s["rs"]) ** 2) * fc[m]))
                continue
            mj = ((ej == s["e1"]) | (ej == s["e2"])) & (d < s["rc"])
            idx = np.nonzero(mj)[0]
            a, b = np.triu_indices(len(idx), 1)
            ja, kb = idx[a], idx[b]
            ok = ((ej[ja] == s["e1"]) & (ej[kb] == s["e2"])) | ((ej[ja] == s["e2"]) & (ej[kb] == s["e1"]))
            ja, kb = ja[ok], kb[ok]
            d3 = dr[kb] - dr[ja]
            rjk = np.sqrt(np.einsum("ij,ij->i", d3, d3))
            ok = rjk < s["rc"]
            ja, kb, rjk = ja[ok], kb[ok], rjk[ok]
            cos = np.einsum("ij,ij->i", dr[ja], dr[kb]) / d[ja] / d[kb]
            pl = 1.0 + s["lam"] * cos
            term = np.where(pl > 0, np.maximum(pl, 0) ** s["zeta"], 0.0)
            rsum = (d[ja] - s["rs"]) ** 2 + (d[kb] - s["rs"]) ** 2 + (rjk - s["rs"]) ** 2
            g = 2.0 ** (1 - s["zeta"]) * np.sum(term * np.exp(-s["eta"] * rsum) * fc[ja] * fc[kb]
                                                * _cutoff_cos(rjk, s["rc"]))
            vals.append(g)
        out[i] = np.array(vals)
    return out


def scaling_rows(sfs, pos, elem, box, atoms=None):
    """nnp-scaling statistics rows (e_index, sf_index, min, max, mean, sigma)."""
    vals = sf_values(sfs, pos, elem, box, atoms)
    return rows_from_values([np.array([v for i, v in vals.items() if elem[i] == e]) for e in range(len(sfs))])


def rows_from_values(mats):
    """Scaling rows from per-element matrices of unscaled SF values
    (atoms x functions)."""
    rows = []
    for e, mat in enumerate(mats):
        for k in range(mat.shape[1]):
            col = mat[:, k]
            mx = col.max()
            mn = col.min()
            if mx - mn <= 1e-12 * max(abs(mx), 1.0):  # avoid a zero scaling range
                mx = mn + max(abs(mn), 1.0) * 1e-3
            rows.append((e + 1, k + 1, mn, mx, col.mean(), col.std()))
    return rows


def write_model(directory, symbols, sf_lines, hidden, activations, rows, seed=SEED_WEIGHTS, cutoff_type="1 0.0",
                weights=None, scaling=True):
    """Write input.nn, scaling.data and weights.%03d.data.

    Weights and biases ~ U(-1/sqrt(fan_in), +1/sqrt(fan_in)) drawn per element
    in NeuralNetwork::setConnections order (NeuralNetwork.cpp:147-169)."""
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "input.nn"), "w") as f:
        f.write("# synthetic 2G-HDNNP model (seeded)\n")
        f.write("number_of_elements %d\n" % len(symbols))
        f.write("elements %s\n" % " ".join(symbols))
        f.write("cutoff_type %s\n" % cutoff_type)
        if scaling:
            f.write("scale_symmetry_functions\ncenter_symmetry_functions\n")
        f.write("scale_min_short 0.0\nscale_max_short 1.0\n")
        f.write("global_hidden_layers_short %d\n" % len(hidden))
        f.write("global_nodes_short %s\n" % " ".join(str(h) for h in hidden))
        f.write("global_activation_short %s\n" % " ".join(activations))
        for l in sf_lines:
            f.write("symfunction_short %s\n" % l)
    with open(os.path.join(directory, "scaling.data"), "w") as f:
        f.write("#" * 80 + "\n# Symmetry function scaling data.\n" + "#" * 80 + "\n")
        for r in rows:
            f.write("%10d %10d %24.16E %24.16E %24.16E %24.16E\n" % r)
    rng = np.random.default_rng(seed)
    sfs = sorted_sfs(sf_lines, symbols)
    all_w = {}
    for e, sym in enumerate(symbols):
        layers = [len(sfs[e])] + list(hidden) + [1]
        if weights is not None:
            w = weights[sym]
        else:
            w = []
            for l in range(1, len(layers)):
                lim = 1.0 / math.sqrt(layers[l - 1])
                w.extend(rng.uniform(-lim, lim, layers[l - 1] * layers[l]))
                w.extend(rng.uniform(-lim, lim, layers[l]))
            w = np.array(w)
        all_w[sym] = w
        with open(os.path.join(directory, "weights.%03d.data" % ATOMIC_NUMBER[sym]), "w") as f:
            f.write("#" * 80 + "\n# Neural network connection values (weights and biases).\n" + "#" * 80 + "\n")
            for v in w:
                f.write("%24.16E\n" % v)
    return all_w


def water_model(directory, seed=SEED_WEIGHTS, cutoff_type="1 0.0", activations=("t", "t", "l"), hidden=(25, 25)):
    """The 50-25-25-1 water model (hidden layer widths ``hidden``), scaling
    statistics from the seeded 192-atom box (computed with the cosine cutoff
    whatever ``cutoff_type`` says: they are only the model's scaling
    constants)."""
    symbols = ("H", "O")
    lines = water_sf_lines()
    if os.path.exists(os.path.join(directory, "weights.008.data")):
        return directory
    pos, elem, box = water_box(4)
    rows = scaling_rows(sorted_sfs(lines, symbols), pos, elem, box)
    write_model(directory, symbols, lines, tuple(hidden), activations, rows, seed=seed, cutoff_type=cutoff_type)
    return directory


def compact_sf_lines():
    """Water SFs built from compact functions (types 20, 21, 22): every
    polynomial order, the asymmetric and exponential subtypes, inner bounds
    below and above zero, angle windows centred at 0 and 180 degrees, and
    narrow (type 21) members of different subtypes sharing (rl, rc) next to
    each other in sort order."""
    lines = []
    for c, o in (("H", "O"), ("O", "H")):
        lines += ["%s 20 H -6.0 6.0 p2" % c, "%s 20 H 0.0 3.0 p2a" % c, "%s 20 O 0.5 4.0 p3" % c,
                  "%s 20 O -6.0 6.0 e" % c, "%s 20 H 1.0 5.0 p4a" % c, "%s 20 O 0.0 6.0 p1" % c,
                  "%s 20 %s 2.0 5.5 p4" % (c, o)]
        lines += ["%s 21 H O -6.0 6.0 0.0 180.0 p2" % c, "%s 21 H O -6.0 6.0 60.0 150.0 p2" % c,
                  "%s 21 H O -6.0 6.0 -90.0 90.0 p2a" % c, "%s 21 H O -6.0 6.0 0.0 180.0 p2a" % c,
                  "%s 21 H H 0.0 6.0 90.0 270.0 p3" % c, "%s 21 H H 0.0 6.0 0.0 120.0 p3a" % c,
                  "%s 21 O O -6.0 6.0 30.0 160.0 e" % c]
        lines += ["%s 22 H O -6.0 6.0 0.0 180.0 p2" % c, "%s 22 H O -6.0 6.0 0.0 180.0 p2a" % c,
                  "%s 22 O O 1.0 6.0 40.0 140.0 e" % c, "%s 22 H H -6.0 6.0 -90.0 90.0 p4" % c,
                  "%s 22 H H 0.5 5.0 20.0 100.0 p1" % c, "%s 22 H O -5.0 5.0 90.0 270.0 p3a" % c]
    return lines


def weighted_sf_lines(rc=6.0):
    """Water SFs mixing element-resolved Gaussian functions (types 2, 3) with
    the element-weighted ones (12, 13, 23, 24, 25), so that derivative slots
    of both kinds interleave."""
    lines = []
    for c in ("H", "O"):
        lines += ["%s 2 H 0.01 0.0 %.1f" % (c, rc), "%s 3 H O 0.005 1.0 2.0 %.1f" % (c, rc)]
        lines += ["%s 12 %s %s %.1f" % (c, eta, rs, rc) for eta, rs in ((0.001, 0.0), (0.02, 1.5), (0.1, 3.0))]
        lines += ["%s 13 0.002 0.0 %s %s %.1f" % (c, lam, zeta, rc) for lam in (1.0, -1.0) for zeta in (1.0, 4.0)]
        lines += ["%s 13 0.01 1.0 1.0 2.5 %.1f" % (c, rc)]
        lines += ["%s 23 -6.0 6.0 p2" % c, "%s 23 0.5 4.5 p3a" % c, "%s 23 1.0 5.5 e" % c]
        lines += ["%s 24 -6.0 6.0 0.0 180.0 p2" % c, "%s 24 -6.0 6.0 60.0 150.0 p2" % c,
                  "%s 24 -6.0 6.0 -90.0 90.0 p2a" % c, "%s 24 0.0 5.0 90.0 270.0 p4" % c]
        lines += ["%s 25 -6.0 6.0 0.0 180.0 p2" % c, "%s 25 0.5 5.0 40.0 140.0 p1" % c,
                  "%s 25 -5.0 5.0 90.0 270.0 p3a" % c]
    return lines


def weighted_water_model(directory, rows=None, seed=SEED_WEIGHTS, cutoff_type="2 0.0"):
    """A water model on weighted_sf_lines() (30-30-1 networks, tanh), scaled
    and centred with `rows`, unscaled without (used to measure the rows)."""
    symbols = ("H", "O")
    lines = weighted_sf_lines()
    sfs = sorted_sfs(lines, symbols)
    scaled = rows is not None
    if not scaled:
        rows = [(e + 1, k + 1, 0.0, 1.0, 0.0, 1.0) for e in range(2) for k in range(len(sfs[e]))]
    write_model(directory, symbols, lines, (30, 30), ("t", "t", "l"), rows, seed=seed, scaling=scaled,
                cutoff_type=cutoff_type)
    return directory


def compact_water_model(directory, rows=None, seed=SEED_WEIGHTS):
    """A water model on compact_sf_lines() (30-30-1 networks, tanh).  With
    `rows` (nnp-scaling rows for the sorted functions) the functions are
    scaled and centred; without, unscaled (used to measure the rows)."""
    symbols = ("H", "O")
    lines = compact_sf_lines()
    sfs = sorted_sfs(lines, symbols)
    scaled = rows is not None
    if not scaled:
        rows = [(e + 1, k + 1, 0.0, 1.0, 0.0, 1.0) for e in range(2) for k in range(len(sfs[e]))]
    write_model(directory, symbols, lines, (30, 30), ("t", "t", "l"), rows, seed=seed, scaling=scaled)
    return directory


def fcc_model(directory, seed=SEED_WEIGHTS):
    """The 80-30-30-1 four-species model, cutoff 6.0 A; scaling statistics
    from a seeded 6^3-cell sample of the same lattice law."""
    symbols = FCC_SYMBOLS
    lines = fcc_sf_lines()
    if os.path.exists(os.path.join(directory, "weights.046.data")):
        return directory
    pos, elem, box = fcc_lattice(n_cells=6, seed=SEED_FCC + 1)
    rows = scaling_rows(sorted_sfs(lines, symbols), pos, elem, box)
    write_model(directory, symbols, lines, (30, 30), ("t", "t", "l"), rows, seed=seed)
    return directory
