"""Parity of the plain FP64 GPU form against the CPU oracle and the
reference's golden vectors (run on an MI355X: ``pytest -m gpu``).

Tolerances: energy 1e-10 relative, forces 1e-8 (model units), symmetry
functions 1e-12 relative; neighbour lists bit-exact (same entries, same order).
"""
import json
import os

import numpy as np
import pytest

import n2p2_amd
from n2p2_amd import read_input_data, remap, synth
from oracle_lib import BRUTE, CELL, PULL, PUSH, Oracle, compact_model, kat_model, weighted_model

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "H2O_RPBE-D3")
E_REL = 1e-10
F_ABS = 1e-8


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def water_dir(tmp_path_factory):
    return synth.water_model(str(tmp_path_factory.mktemp("water")))


GOLDEN = {"H2O_RPBE-D3": ["H", "O"], "Cu2S_PBE": ["S", "Cu"], "Anisole_SCAN": ["H", "C", "O"],
          "DMABN_SCAN": ["H", "C", "N"], "Ethylbenzene_SCAN": ["H", "C"]}


def predict_golden(mode, gdir=GOLD, symbols=("H", "O"), precision=n2p2_amd.PLAIN_FP64):
    """Prediction::readStructureFromFile + predict (Prediction.cpp:50-78) on the GPU."""
    pos, elem, box = read_input_data(os.path.join(gdir, "input.data"), list(symbols))
    pos = remap(pos, box)
    cl, ce = mode.conv_length, mode.conv_energy
    mode.set_structures(pos * cl, elem, None if box is None else (box * cl)[None])
    mode.evaluate(precision, forces=True)
    n = len(elem)
    E = mode.energies()[0] / ce + n * mode.mean_energy
    E += sum(np.count_nonzero(elem == e) * mode.energy_offsets[e] for e in range(mode.num_elements))
    F = mode.forces() * (1.0 / (ce / cl))
    return E, F


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_plain_fp64(name):
    """test/cpp/test_Prediction.cpp examples: H2O_RPBE-D3 (SF types 2/3,
    CT_TANHU), Cu2S_PBE (types 2/3/9, CT_POLY2, SCALESIGMA, softplus) and the
    three SCAN models (compact types 20/22, normalised units)."""
    say("case: %s golden, plain FP64" % name)
    gdir = os.path.join(HERE, "golden", name)
    exp = json.load(open(os.path.join(gdir, "expected.json")))
    m = n2p2_amd.Mode(gdir)
    E, F = predict_golden(m, gdir, GOLDEN[name])
    Fg = np.array(exp["forces"])
    dE = abs(E - exp["energy"])
    dF = np.abs(F - Fg).max()
    say("  |dE| = %.3e (rel %.3e), max|dF| = %.3e" % (dE, dE / abs(exp["energy"]), dF))
    assert dE / abs(exp["energy"]) <= E_REL
    assert dF <= F_ABS


def compare_neighbors(m, o):
    c1, j1, s1 = m.neighbors()
    c2, j2, s2 = o.neighbors()
    assert (c1 == c2).all()
    assert (j1 == j2).all()
    assert (s1 == s2).all()
    return int(c1.sum())


@pytest.mark.parametrize("case", ["water192", "water1536_cells", "golden_cells", "small_cells",
                                  "triclinic_cells"])
def test_neighbor_list_identical(case, water_dir):
    say("case: neighbour list", case)
    if case == "golden_cells":
        d = GOLD
        pos, elem, box = read_input_data(os.path.join(GOLD, "input.data"), ["H", "O"])
        o = Oracle(d)
        o.set_structure(pos, elem, box)
        m = n2p2_amd.Mode(d)
        cl = m.conv_length
        m.set_structures(remap(pos, box) * cl, elem, None if box is None else (box * cl)[None])
        o.build_neighbors(BRUTE)
    else:
        if case == "water192":
            pos, elem, box = synth.water_box(4)
        elif case == "water1536_cells":
            pos, elem, box = synth.water_box(8)
        elif case == "triclinic_cells":
            # skewed cell with >= 3 cells per axis (perpendicular widths
            # 38-41 bohr against rc = 12): the cell search's skipped cells
            # are bounded by their largest per-axis gap here
            rng = np.random.default_rng(13)
            box = np.array([[41.0, 0.0, 0.0], [9.0, 39.5, 0.0], [-6.0, 5.0, 40.5]])
            pos = remap(rng.random((900, 3)) @ box, box)
            elem = (rng.random(900) < 0.35).astype(np.int32)
        else:
            rng = np.random.default_rng(11)
            box = np.eye(3) * 4.2
            pos = remap(rng.random((9, 3)) * 4.2, box)
            elem = np.array([1, 1, 1, 0, 0, 0, 0, 0, 0], np.int32)
        o = Oracle(water_dir)
        o.set_structure(pos, elem, box)
        o.build_neighbors(BRUTE)
        m = n2p2_amd.Mode(water_dir)
        m.set_structures(remap(pos, box), elem, box[None])
    m.calculate_neighbor_list()
    n = compare_neighbors(m, o)
    say("  %d entries identical" % n)


@pytest.mark.parametrize("case", ["batch24", "triclinic"])
def test_neighbor_list_images(case, water_dir):
    """Brute-force lists of cells smaller than 3 rc (images pruned by slab
    distance): every structure of an uneven batch, and a skewed triclinic
    cell with several image shells, entry for entry against the oracle."""
    say("case: neighbour list with images,", case)
    if case == "batch24":
        pos, elem, boxes, off = synth.random_batch(n_struct=24, nmin=8, nmax=160)
        pos = np.vstack([remap(pos[off[s]:off[s + 1]], boxes[s]) for s in range(24)])
    else:
        rng = np.random.default_rng(12)
        boxes = np.array([[[5.0, 0.0, 0.0], [1.7, 4.6, 0.0], [-1.2, 0.9, 4.1]]])
        f = rng.random((20, 3))
        pos = remap(f @ boxes[0], boxes[0])
        elem = np.array([1] * 7 + [0] * 13, np.int32)
        off = np.array([0, 20])
    m = n2p2_amd.Mode(water_dir)
    m.set_structures(pos, elem, boxes, off)
    m.calculate_neighbor_list()
    c1, j1, s1 = m.neighbors()
    e0 = np.concatenate([[0], np.cumsum(c1)])
    total = 0
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        o = Oracle(water_dir)
        o.set_structure(pos[a:b], elem[a:b], boxes[s])
        o.build_neighbors(BRUTE)
        c2, j2, s2 = o.neighbors()
        assert (c1[a:b] == c2).all(), s
        assert (j1[e0[a]:e0[b]] - a == j2).all(), s
        assert (s1[e0[a]:e0[b]] == s2).all(), s
        total += int(c2.sum())
    say("  %d entries identical over %d structures" % (total, len(off) - 1))


def check_against_oracle(m, o, force_mode=PULL):
    o.predict(force_mode, BRUTE)
    G, Go = m.G(), o.G()
    # scaled G are O(1) (Smin + factor * (G - mean)); relative to max(1, |G|)
    gerr = (np.abs(G - Go) / np.maximum(np.abs(Go), 1.0)).max()
    E, Eo = m.energies()[0], o.energy()
    F, Fo = m.forces(), o.forces()
    dEdG, dEdGo = m.dEdG(), o.dEdG()
    derr = np.abs(dEdG - dEdGo).max() / max(np.abs(dEdGo).max(), 1e-30)
    ferr = np.abs(F - Fo).max()
    say("  G rel %.2e  dEdG rel %.2e  E rel %.2e  F abs %.2e" % (gerr, derr, abs(E - Eo) / abs(Eo), ferr))
    assert gerr <= 1e-12
    assert derr <= 1e-10
    assert abs(E - Eo) <= E_REL * abs(Eo)
    assert ferr <= F_ABS


def test_water192_plain_vs_oracle(water_dir):
    say("case: water-192 plain FP64 vs oracle")
    pos, elem, box = synth.water_box(4)
    m = n2p2_amd.Mode(water_dir)
    m.set_structures(remap(pos, box), elem, box[None])
    m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
    o = Oracle(water_dir)
    o.set_structure(pos, elem, box)
    check_against_oracle(m, o)


def test_water12288_plain_vs_oracle(water_dir):
    say("case: water-12288 plain FP64 vs oracle (cell list)")
    pos, elem, box = synth.water_box(16)
    m = n2p2_amd.Mode(water_dir)
    m.set_structures(remap(pos, box), elem, box[None])
    m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
    o = Oracle(water_dir)
    o.set_structure(pos, elem, box)
    o.predict(PULL, CELL)
    E, Eo = m.energies()[0], o.energy()
    F, Fo = m.forces(), o.forces()
    say("  E rel %.2e  F abs %.2e" % (abs(E - Eo) / abs(Eo), np.abs(F - Fo).max()))
    assert abs(E - Eo) <= E_REL * abs(Eo)
    assert np.abs(F - Fo).max() <= F_ABS


def test_fcc_four_species_plain(tmp_path_factory):
    """Four-species FCC model (80 SFs, 80-30-30-1 NN, rc 6.0): 5^3 cells."""
    say("case: FCC 500 atoms plain FP64")
    d = synth.fcc_model(str(tmp_path_factory.mktemp("fcc")))
    pos, elem, box = synth.fcc_lattice(n_cells=5)
    m = n2p2_amd.Mode(d)
    m.set_structures(remap(pos, box), elem, box[None])
    m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE)
    E, Eo = m.energies()[0], o.energy()
    dF = float(np.abs(m.forces() - o.forces()).max())
    say("  E rel %.2e  max|dF| %.2e" % (abs(E - Eo) / abs(Eo), dF))
    assert abs(E - Eo) / abs(Eo) <= E_REL
    assert dF <= F_ABS


def test_step_api_equals_evaluate(water_dir):
    """calculate_symmetry_function_groups / _atomic_neural_networks / _energy /
    _forces in sequence give the evaluate results."""
    say("case: step-by-step Mode calls")
    pos, elem, box = synth.water_box(4)
    m = n2p2_amd.Mode(water_dir)
    m.set_structures(pos, elem, box[None])
    m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
    E1, F1 = m.energies(), m.forces()
    m.calculate_neighbor_list()
    m.calculate_symmetry_function_groups(True)
    m.calculate_atomic_neural_networks(True)
    m.calculate_energy()
    m.calculate_forces()
    # evaluate recomputes the derivatives (fused), the step calls store them:
    # equal up to summation order
    assert np.abs(m.energies() - E1).max() <= 1e-14 * np.abs(E1).max()
    assert np.abs(m.forces() - F1).max() <= 1e-12


@pytest.mark.parametrize("force_mode", [PULL, PUSH])
def test_batched_small_cells(force_mode, water_dir):
    """Several uneven structures in one call, including cells smaller than rc
    (self images); pull form reproduces Mode::calculateForces, push form the
    InterfaceLammps scatter."""
    say("case: batch of 12 structures, force mode", force_mode)
    pos, elem, boxes, off = synth.random_batch(n_struct=12, nmin=16, nmax=90)
    m = n2p2_amd.Mode(water_dir)
    m.set_option(n2p2_amd.OPT_FORCE_MODE, force_mode)
    gpos = np.vstack([remap(pos[off[s]:off[s + 1]], boxes[s]) for s in range(12)])
    m.set_structures(gpos, elem, boxes, off)
    m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
    E, F = m.energies(), m.forces()
    for s in range(12):
        a, b = off[s], off[s + 1]
        o = Oracle(water_dir)
        o.set_structure(pos[a:b], elem[a:b], boxes[s])
        o.predict(force_mode, BRUTE)
        assert abs(E[s] - o.energy()) <= E_REL * abs(o.energy()), s
        assert np.abs(F[a:b] - o.forces()).max() <= F_ABS, s


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("force_mode", [PULL, PUSH])
def test_plain_force_paths(path, force_mode, water_dir, tmp_path_factory):
    """Both force paths of the plain form (NNPX_OPT_PLAIN_FORCES: 0 the fused
    FP64 recompute, 1 the stored dG/dr) against the oracle: a batch of
    uneven cells including ones smaller than rc (self images), and the
    weighted (types 12/13 with 2/3 and 23-25) and compact (20-22) models."""
    say("case: plain force path %d, force mode %d" % (path, force_mode))
    pos, elem, boxes, off = synth.random_batch(n_struct=8, nmin=16, nmax=90, seed=synth.SEED_BATCH + 9)
    models = [("water", water_dir),
              ("weighted", weighted_model(str(tmp_path_factory.mktemp("w") / "m"))),
              ("compact", compact_model(str(tmp_path_factory.mktemp("c") / "m")))]
    for name, d in models:
        m = n2p2_amd.Mode(d)
        m.set_option(n2p2_amd.OPT_PLAIN_FORCES, path)
        m.set_option(n2p2_amd.OPT_FORCE_MODE, force_mode)
        m.set_structures(pos, elem, boxes, off)
        m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
        E, F = m.energies(), m.forces()
        worst_e = worst_f = 0.0
        o = Oracle(d)
        for s in range(len(off) - 1):
            a, b = off[s], off[s + 1]
            o.set_structure(pos[a:b], elem[a:b], boxes[s])
            o.predict(force_mode, BRUTE)
            worst_e = max(worst_e, abs(E[s] - o.energy()) / abs(o.energy()))
            worst_f = max(worst_f, float(np.abs(F[a:b] - o.forces()).max()))
        say("  %s: max E rel %.2e  max |dF| %.2e" % (name, worst_e, worst_f))
        assert worst_e <= E_REL
        assert worst_f <= F_ABS


def test_update_positions(water_dir):
    say("case: update_positions")
    pos, elem, box = synth.water_box(4)
    m = n2p2_amd.Mode(water_dir)
    m.set_structures(pos, elem, box[None])
    rng = np.random.default_rng(5)
    p2 = remap(pos + rng.normal(0, 0.02, pos.shape), box)
    m.update_positions(remap(p2, box))
    m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
    o = Oracle(water_dir)
    o.set_structure(p2, elem, box)
    o.predict(PULL, BRUTE)
    assert abs(m.energies()[0] - o.energy()) <= E_REL * abs(o.energy())
    assert np.abs(m.forces() - o.forces()).max() <= F_ABS


@pytest.mark.parametrize("name", ["Mono", "Dual", "Mixed"])
def test_sf_known_answers_gpu(name, tmp_path):
    """test/cpp/fixtures.h known answers of every SF type (2, 3, 9, 12, 13,
    20-25) on the fixture clusters through the plain GPU form: G of atom 0 at
    1000*eps (test_SymmetryFunctions.cpp:36), forces equal to the oracle's."""
    k = json.load(open(os.path.join(HERE, "golden", "sf_kat.json")))["sets"][name]
    say("case: SF known answers on the GPU, %s set (%d functions)" % (name, len(k["functions"])))
    worst = 0.0
    for q, row in enumerate(k["functions"]):
        d = str(tmp_path / ("m%d" % q))
        kat_model(d, k["elements"], [row["line"]])
        pos = np.array(k["pos"], float)
        elem = np.array(k["elem"], np.int32)
        m = n2p2_amd.Mode(d)
        m.set_structures(pos, elem, None)
        m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
        g = m.G()[0, 0]
        o = Oracle(d)
        o.set_structure(pos, elem, None)
        o.predict(PULL, BRUTE)
        dF = float(np.abs(m.forces() - o.forces()).max())
        worst = max(worst, abs(g - row["value"]))
        assert abs(g - row["value"]) <= 1000 * np.finfo(float).eps, (row["line"], g, row["value"])
        assert dF <= F_ABS, (row["line"], dF)
    say("  max |G - KAT| %.2e" % worst)


@pytest.mark.parametrize("n_side", [4, 8])
def test_weighted_model_plain(n_side, tmp_path_factory):
    """Water model mixing element-resolved types 2/3 with the element-weighted
    types 12, 13, 23, 24, 25: G, dE/dG, energy and forces against the oracle;
    192 atoms (brute-force list) and 1,536 atoms (cell list)."""
    say("case: weighted water model, %d atoms, plain FP64" % (3 * n_side ** 3))
    d = weighted_model(str(tmp_path_factory.mktemp("weighted") / "m"))
    pos, elem, box = synth.water_box(n_side)
    m = n2p2_amd.Mode(d)
    m.set_structures(remap(pos, box), elem, box[None])
    m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE if n_side == 4 else CELL)
    G, Go = m.G(), o.G()
    gerr = float((np.abs(G - Go) / np.maximum(np.abs(Go), 1.0)).max())
    E, Eo = m.energies()[0], o.energy()
    dF = float(np.abs(m.forces() - o.forces()).max())
    say("  G err %.2e  E rel %.2e  max|dF| %.2e (|F| rms %.3f)"
        % (gerr, abs(E - Eo) / abs(Eo), dF, float(np.sqrt(np.mean(o.forces() ** 2)))))
    assert gerr <= 1e-12
    assert abs(E - Eo) <= E_REL * abs(Eo)
    assert dF <= F_ABS


@pytest.mark.parametrize("n_side", [4, 8])
def test_compact_model_plain(n_side, tmp_path_factory):
    """Compact-function water model (types 20/21/22, every subtype, narrow
    members of different subtypes sharing (rl, rc)): G, dE/dG, energy and
    forces against the oracle; 192 atoms (brute-force list) and 1,536 atoms
    (cell list)."""
    say("case: compact water model, %d atoms, plain FP64" % (3 * n_side ** 3))
    d = compact_model(str(tmp_path_factory.mktemp("compact") / "m"))
    pos, elem, box = synth.water_box(n_side)
    m = n2p2_amd.Mode(d)
    m.set_structures(remap(pos, box), elem, box[None])
    m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE if n_side == 4 else CELL)
    G, Go = m.G(), o.G()
    gerr = float((np.abs(G - Go) / np.maximum(np.abs(Go), 1.0)).max())
    derr = float(np.abs(m.dEdG() - o.dEdG()).max())
    E, Eo = m.energies()[0], o.energy()
    dF = float(np.abs(m.forces() - o.forces()).max())
    say("  G err %.2e  dEdG err %.2e  E rel %.2e  max|dF| %.2e (|F| rms %.3f)"
        % (gerr, derr, abs(E - Eo) / abs(Eo), dF, float(np.sqrt(np.mean(o.forces() ** 2)))))
    assert gerr <= 1e-12
    assert derr <= 1e-10
    assert abs(E - Eo) <= E_REL * abs(Eo)
    assert dF <= F_ABS
