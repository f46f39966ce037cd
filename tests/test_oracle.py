"""Pins of the CPU oracle (CPU only): the reference's own golden vectors and
known answers, the reference compute classes where they build, and
closed-form / invariance checks chosen so that a dropped term, a sign error or
a wrong index fails at least one of them."""
import json
import os

import numpy as np
import pytest

from oracle_lib import (BRUTE, CELL, PULL, PUSH, Oracle, RefLib, compact_model, cutoff_fdf, kat_model,
                        ref_available, weighted_model)

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
EPS = np.finfo(float).eps


@pytest.mark.parametrize("name", ["H2O_RPBE-D3", "Cu2S_PBE", "Anisole_SCAN", "DMABN_SCAN", "Ethylbenzene_SCAN"])
def test_golden_energy_and_forces(name):
    """test/cpp/test_Prediction.cpp:25,44-56: E and all forces at 1000*eps
    (H2O_RPBE-D3: SF types 2/3, CT_TANHU; Cu2S_PBE: types 2/3/9, CT_POLY2,
    SCALESIGMA, softplus; the three SCAN models: compact types 20/22 with
    polynomial and asymmetric polynomial subtypes, normalised units)."""
    d = os.path.join(GOLD, name)
    exp = json.load(open(os.path.join(d, "expected.json")))
    o = Oracle(d)
    o.read_structure(os.path.join(d, "input.data"))
    o.predict(PULL, BRUTE)
    assert abs(o.energy() - exp["energy"]) <= 1000 * EPS
    F = np.array(exp["forces"])
    assert np.abs(o.forces() - F).max() <= 1000 * EPS


def test_cutoff_known_answers():
    """test/cpp/test_CutoffFunction.cpp:17-50 (100*eps), and zero beyond rc."""
    k = json.load(open(os.path.join(GOLD, "cutoff_kat.json")))
    for t, (f, df) in enumerate(zip(k["f"], k["df"])):
        fc, dfc = cutoff_fdf(t, k["rc"], k["alpha"], k["r"])
        assert abs(fc - f) <= 100 * EPS, (k["types"][t], fc, f)
        assert abs(dfc - df) <= 100 * EPS, (k["types"][t], dfc, df)
        assert cutoff_fdf(t, k["rc"], k["alpha"], k["rc"]) == (0.0, 0.0)
        assert cutoff_fdf(t, k["rc"], k["alpha"], 1.01 * k["rc"]) == (0.0, 0.0)
    # CT_COS closed form fc = (cos(pi r/rc)+1)/2, dfc = -pi/(2 rc) sin(pi r/rc)
    for r in (0.3, 1.7, 2.9):
        fc, dfc = cutoff_fdf(1, 3.0, 0.0, r)
        assert abs(fc - 0.5 * (np.cos(np.pi * r / 3) + 1)) < 1e-15
        assert abs(dfc + np.pi / 6 * np.sin(np.pi * r / 3)) < 1e-14


@pytest.mark.parametrize("name", ["Mono", "Dual", "Mixed"])
def test_sf_known_answers(tmp_path, name):
    """test/cpp/fixtures.h:151-258 values of atom 0 for types 2, 3, 9, 12, 13, 20-25 at 1000*eps
    (test/cpp/test_SymmetryFunctions.cpp:36, 162-190)."""
    from n2p2_amd.synth import sorted_sfs
    k = json.load(open(os.path.join(GOLD, "sf_kat.json")))["sets"][name]
    for row in k["functions"]:
        d = str(tmp_path / ("m%d" % abs(hash(row["line"]))))
        lines, order = kat_model(d, k["elements"], [row["line"]])
        o = Oracle(d)
        # element indices of the fixture are in ElementMap (Z) order already
        o.set_structure(np.array(k["pos"], float), np.array(k["elem"]), None)
        o.predict(PULL, BRUTE)
        sfs = sorted_sfs(lines, order)
        s0 = sfs[0]
        parsed = [l for l in lines if l.split()[0] == order[0]]
        # index of the KAT function among element 0's sorted functions
        from n2p2_amd.synth import _parse
        target = _parse(row["line"], order)
        idx = [i for i, s in enumerate(s0) if s == target][0]
        assert parsed
        g = o.G()[0, idx]
        assert abs(g - row["value"]) <= 1000 * EPS, (row["line"], g, row["value"])


def test_forces_match_finite_differences(water_model_dir):
    """Analytic forces vs central differences (nnp-checkf threshold 10*h^2)."""
    from n2p2_amd import synth
    pos, elem, box = synth.water_box(4)
    o = Oracle(water_model_dir)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE)
    F = o.forces()
    assert np.abs(F.sum(axis=0)).max() < 1e-12  # translation invariance
    sub = [0, 1, 2, 95, 191]
    o2 = Oracle(water_model_dir)
    h = 1e-4
    for i in sub:
        for c in range(3):
            e = []
            for s in (-1, 1):
                p = pos.copy()
                p[i, c] += s * h
                o2.set_structure(p, elem, box)
                o2.predict(PULL, BRUTE)
                e.append(o2.energy())
            fd = -(e[1] - e[0]) / (2 * h)
            assert abs(fd - F[i, c]) <= 10 * h * h, (i, c, fd, F[i, c])


def test_self_image_semantics(water_model_dir):
    """Cell edge < rc: push form (every entry scatters) equals finite
    differences; the pull form of Mode::calculateForces (loop from
    neighborsUnique+1, Mode.cpp:1957) differs by the self-image terms."""
    rng = np.random.default_rng(7)
    L = 5.5
    box = np.eye(3) * L
    pos = rng.random((8, 3)) * L
    elem = np.array([1, 1, 0, 0, 0, 0, 0, 0], np.int32)
    o = Oracle(water_model_dir)
    o.set_structure(pos, elem, box)
    o.predict(PUSH, BRUTE)
    Fpush = o.forces()
    o.predict(PULL, BRUTE)
    Fpull = o.forces()
    assert np.abs(Fpush - Fpull).max() > 1e-6
    h = 1e-5
    for i in range(3):
        for c in range(3):
            e = []
            for s in (-1, 1):
                p = pos.copy()
                p[i, c] += s * h
                o.set_structure(p, elem, box)
                o.predict(PULL, BRUTE)
                e.append(o.energy())
            fd = -(e[1] - e[0]) / (2 * h)
            assert abs(fd - Fpush[i, c]) < 1e-6 * max(1.0, abs(fd)), (i, c, fd, Fpush[i, c])


def test_cell_list_equals_brute_force():
    """Cell-list neighbour list: same (j, image) entries in the same order as
    the O(N^2) reference loop (1,536-atom water box, 3 cells per axis)."""
    from n2p2_amd import synth
    pos, elem, box = synth.water_box(8)
    o = Oracle(_water_dir())
    o.set_structure(pos, elem, box)
    n1 = o.build_neighbors(BRUTE)
    c1, j1, s1 = o.neighbors()
    n2 = o.build_neighbors(CELL)
    c2, j2, s2 = o.neighbors()
    assert n1 == n2
    assert (c1 == c2).all() and (j1 == j2).all() and (s1 == s2).all()


_WATER = {}


def _water_dir():
    if "d" not in _WATER:
        import tempfile
        from n2p2_amd import synth
        _WATER["d"] = synth.water_model(os.path.join(tempfile.mkdtemp(), "water"))
    return _WATER["d"]


@pytest.mark.skipif(not ref_available(), reason="reference compute classes not built (no reference tree)")
@pytest.mark.parametrize("case", ["golden", "water192", "small_cell", "fcc108", "cu2s", "dmabn", "compact",
                                  "weighted"])
def test_oracle_bitwise_equals_reference_classes(case, water_model_dir, tmp_path):
    """The oracle reproduces n2p2's own SymGrp/NeuralNetwork/Atom code bit for bit."""
    from n2p2_amd import read_input_data, synth
    if case == "golden":
        d = os.path.join(GOLD, "H2O_RPBE-D3")
        pos, elem, box = read_input_data(os.path.join(d, "input.data"), ["H", "O"])
    elif case == "water192":
        d = water_model_dir
        pos, elem, box = synth.water_box(4)
    elif case == "cu2s":
        d = os.path.join(GOLD, "Cu2S_PBE")
        pos, elem, box = read_input_data(os.path.join(d, "input.data"), ["S", "Cu"])
    elif case == "dmabn":
        d = os.path.join(GOLD, "DMABN_SCAN")
        pos, elem, box = read_input_data(os.path.join(d, "input.data"), ["H", "C", "N"])
    elif case == "compact":
        # every compact subtype, and narrow members of different subtypes
        # sharing (rl, rc), whose r_jk factor the reference reuses
        d = compact_model(str(tmp_path / "compact"))
        pos, elem, box = synth.water_box(4)
    elif case == "weighted":
        # element-weighted types 12/13/23/24/25 interleaved with types 2/3
        d = weighted_model(str(tmp_path / "weighted"))
        pos, elem, box = synth.water_box(4)
    elif case == "fcc108":
        d = synth.fcc_model(str(tmp_path / "fcc"))
        pos, elem, box = synth.fcc_lattice(n_cells=3)
    else:
        d = water_model_dir
        rng = np.random.default_rng(3)
        box = np.eye(3) * 5.0
        pos = rng.random((6, 3)) * 5.0
        elem = np.array([1, 0, 0, 1, 0, 0], np.int32)
    o = Oracle(d)
    r = RefLib(d)
    o.set_structure(pos, elem, box)
    r.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE)
    r.predict()
    assert o.energy() == r.energy()
    assert (o.forces() == r.forces()).all()
    assert (o.G() == r.G(o.max_sf)).all()
    assert (o.dEdG() == r.dEdG(o.max_sf)).all()


@pytest.mark.parametrize("part", ["energy", "force"])
def test_loss_gradient_matches_finite_differences(tmp_path, part):
    """Training gradient of wE (E - E_ref)^2 + wF sum (F - F_ref)^2 (the
    reference's calculateDEdc / calculateDFdc / collectDGdxia, restated)
    against central differences with h = 1e-5 on 48 connections spread over
    both networks and all layers, for the energy and the force term
    separately; 81-atom water box (no self images), targets from a second
    (teacher) network."""
    from n2p2_amd import synth
    d = synth.water_model(str(tmp_path / "student"))
    t = synth.water_model(str(tmp_path / "teacher"), seed=synth.SEED_TEACHER)
    pos, elem, box = synth.water_box(3)
    teach = Oracle(t)
    teach.set_structure(pos, elem, box)
    teach.predict(PULL, BRUTE)
    E_ref, F_ref = teach.energy(), teach.forces()
    n = len(elem)
    wE, wF = (1.0 / n, 0.0) if part == "energy" else (0.0, 0.7 / (3 * n))
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE)
    loss, g = o.loss_gradient(E_ref, F_ref, wE, wF)
    assert loss[0] > 0
    w0 = o.weights()
    rng = np.random.default_rng(5)
    idx = np.unique(np.concatenate([rng.choice(len(w0), 40, replace=False),
                                    [0, 1, 1250, 1275, 1950, len(w0) // 2, len(w0) - 1, len(w0) - 26]]))

    def L(w):
        o.set_weights(w)
        o.set_structure(pos, elem, box)
        o.predict(PULL, BRUTE)
        return o.loss_gradient(E_ref, F_ref, wE, wF)[0][0]

    h = 1e-5
    worst = 0.0
    for k in idx:
        wp, wm = w0.copy(), w0.copy()
        wp[k] += h
        wm[k] -= h
        fd = (L(wp) - L(wm)) / (2 * h)
        worst = max(worst, abs(fd - g[k]))
    o.set_weights(w0)
    scale = np.abs(g).max()
    assert worst <= 1e-6 * scale, (worst, scale)


def _dPdc_case(case, tmp_path):
    """(model dir, structures) of the weight-derivative pins: three dimers of
    the nnp-checkdw LJ example, or a 24-atom water box whose 6.21 A cell is
    smaller than the cutoff (self images in collectDGdxia)."""
    from n2p2_amd import read_structures, synth
    if case == "lj":
        d = os.path.join(GOLD, "LJ_checkdw")
        s = read_structures(os.path.join(d, "input.data"), ["Ne"])
        return d, [s[0], s[50], s[100]]
    d = synth.water_model(str(tmp_path / "water"))
    return d, [synth.water_box(2)]


@pytest.mark.parametrize("case", ["lj", "water24"])
def test_dPdc_matches_finite_differences(case, tmp_path):
    """Training::dPdc restated (oracle_dPdc): the energy row and every force
    row against central differences of the oracle's energy and forces,
    h = 1e-5 (Training::dPdcN, Training.cpp:3391-3471); every connection of
    the LJ network, 40 connections over both water networks.  collectDGdxia
    keeps the entries of an atom's own images, so in the self-image cell the
    rows are derivatives of the push-form forces (InterfaceLammps::getForces
    semantics), not of Mode::calculateForces, which skips them; the pull form
    differs there by O(1e-4) and is not the function differentiated."""
    d, structs = _dPdc_case(case, tmp_path)
    o = Oracle(d)
    w0 = o.weights()
    rng = np.random.default_rng(11)
    idx = np.arange(len(w0)) if len(w0) <= 64 else np.unique(np.concatenate(
        [rng.choice(len(w0), 36, replace=False), [0, len(w0) // 2, len(w0) - 26, len(w0) - 1]]))
    h = 1e-5
    for pos, elem, box in structs:
        o.set_weights(w0)
        o.set_structure(pos, elem, box)
        o.predict(PULL, BRUTE)
        dE = o.dPdc("energy")
        dF = o.dPdc("force")
        assert dF.shape == (3 * len(elem), len(w0))
        worstE = worstF = 0.0
        for k in idx:
            res = []
            for sgn in (1.0, -1.0):
                w = w0.copy()
                w[k] += sgn * h
                o.set_weights(w)
                o.set_structure(pos, elem, box)
                o.predict(PUSH, BRUTE)
                res.append((o.energy(), o.forces().reshape(-1)))
            worstE = max(worstE, abs((res[0][0] - res[1][0]) / (2 * h) - dE[k]))
            worstF = max(worstF, float(np.abs((res[0][1] - res[1][1]) / (2 * h) - dF[:, k]).max()))
        assert worstE <= 1e-7 * max(np.abs(dE).max(), 1.0), (worstE, np.abs(dE).max())
        assert worstF <= 1e-7 * max(np.abs(dF).max(), 1.0), (worstF, np.abs(dF).max())
    o.set_weights(w0)


@pytest.mark.skipif(not ref_available(), reason="reference compute classes not built (no reference tree)")
@pytest.mark.parametrize("case", ["lj", "water24"])
def test_dPdc_equals_reference_classes(case, tmp_path):
    """oracle_dPdc against Training::dPdc driven through the reference's own
    NeuralNetwork::calculateDEdc / calculateDFdc and Atom/Element code."""
    d, structs = _dPdc_case(case, tmp_path)
    o = Oracle(d)
    r = RefLib(d)
    n = o.num_connections()
    for pos, elem, box in structs:
        o.set_structure(pos, elem, box)
        r.set_structure(pos, elem, box)
        o.predict(PULL, BRUTE)
        r.predict()
        for prop in ("energy", "force"):
            a, b = o.dPdc(prop), r.dPdc(prop, n)
            assert np.abs(a - b).max() <= 1e3 * EPS * max(np.abs(b).max(), 1.0), (prop, np.abs(a - b).max())
