"""Pins of the CPU oracle (CPU only): the reference's own golden vectors and
known answers, the reference compute classes where they build, and
closed-form / invariance checks chosen so that a dropped term, a sign error or
a wrong index fails at least one of them."""
import json
import os

import numpy as np
import pytest

from oracle_lib import (BRUTE, CELL, PULL, PUSH, Oracle, RefLib, compact_model, cutoff_fdf, kalman_update,
                        kat_model, ref_available, weighted_model)

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


@pytest.mark.parametrize("name", ["Mono", "Dual", "Mixed"])
def test_dGdr_matches_finite_differences(tmp_path, name):
    """test/cpp/test_SymmetryFunctions.cpp:101-155 procedure on the fixtures.h
    clusters, every SF type: the stored Atom::dGdr of atom 0 and
    Neighbor::dGdr of each of its neighbour entries against central
    differences of G with h = 1e-7, at 1e-6 (accuracyNumeric, :37)."""
    k = json.load(open(os.path.join(GOLD, "sf_kat.json")))["sets"][name]
    pos0 = np.array(k["pos"], float)
    elem = np.array(k["elem"])
    n = len(elem)
    h = 1e-7
    for q, row in enumerate(k["functions"]):
        d = str(tmp_path / ("m%d" % q))
        kat_model(d, k["elements"], [row["line"]])
        o = Oracle(d)
        o.set_structure(pos0, elem, None)
        o.predict(PULL, BRUTE)
        counts, js, _ = o.neighbors()
        js0 = list(js[:counts[0]])  # atom 0's entries (non-periodic: one per atom within rc)
        s_an, n_an = o.dGdr(o.max_sf)
        num = np.zeros((n, 3))
        for ia in range(n):
            for ic in range(3):
                g = []
                for sg in (1, -1):
                    p = pos0.copy()
                    p[ia, ic] += sg * h
                    o.set_structure(p, elem, None)
                    o.predict(PULL, BRUTE)
                    g.append(o.G()[0, 0])
                num[ia, ic] = (g[0] - g[1]) / (2 * h)
        assert np.abs(s_an[0, 0] - num[0]).max() <= 1e-6, (row["line"], s_an[0, 0], num[0])
        for ia in range(1, n):
            an = n_an[js0.index(ia), 0] if ia in js0 else np.zeros(3)  # beyond rc: dG/dr = 0
            assert np.abs(an - num[ia]).max() <= 1e-6, (row["line"], ia, an, num[ia])
        assert np.abs(num).max() > 1e-6, row["line"]


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
    # stored derivatives Atom::dGdr / Neighbor::dGdr, slot for slot
    # (Element::setupSymmetryFunctionMemory order, Element.cpp:289-311)
    counts, _, _ = o.neighbors()
    ldn = o.max_sf
    s_o, n_o = o.dGdr(ldn)
    s_r, n_r = r.dGdr(o.max_sf, ldn, int(counts.sum()))
    assert (s_o == s_r).all()
    assert (n_o == n_r).all()
    assert np.abs(n_o).max() > 0


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


def ew_structures():
    """Water boxes outside the range the scaling data was measured on: one
    compressed to 0.85 of its spacing (values above Gmax), one expanded to 1.3
    (values below Gmin); and an isolated dilute H/O cluster whose atoms have
    fewer neighbours than their SFs need."""
    from n2p2_amd import synth
    out = []
    for f in (0.85, 1.3):
        pos, elem, box = synth.water_box(4, spacing=3.105 * f)
        out.append(("water x%.2f" % f, pos, elem, box))
    pos = np.array([[0.0, 0, 0], [9.0, 0, 0], [0, 9.0, 0], [0.5, 0.4, 0.3], [20.0, 20, 20]])
    out.append(("dilute", pos, np.array([1, 0, 0, 0, 1], np.int32), None))
    return out


@pytest.mark.skipif(not ref_available(), reason="reference compute classes not built (no reference tree)")
def test_extrapolation_warnings_equal_reference_classes(water_model_dir):
    """Extrapolation warnings (Element::updateSymmetryFunctionStatistics,
    Element.cpp:462-541) and the low-neighbour check (Mode.cpp:1667-1679):
    the oracle's records equal those of the reference's own Element,
    SymFncStatistics and Atom code, bit for bit, on structures outside the
    scaling range and a dilute cluster."""
    total_ew = total_low = 0
    for name, pos, elem, box in ew_structures():
        o = Oracle(water_model_dir)
        r = RefLib(water_model_dir)
        o.set_structure(pos, elem, box)
        r.set_structure(pos, elem, box)
        o.predict(PULL, BRUTE)
        r.predict()
        eo, er = o.extrapolation_warnings(), r.extrapolation_warnings()
        assert len(eo) == len(er), name
        for f in ("atom", "sf", "type", "value", "gmin", "gmax"):
            assert (eo[f] == er[f]).all(), (name, f)
        lo, lr = o.low_neighbor_atoms(), r.low_neighbor_atoms()
        assert (lo[0] == lr[0]).all() and (lo[1] == lr[1]).all(), name
        total_ew += len(eo)
        total_low += len(lo[0])
    assert total_ew > 100 and total_low >= 3


def test_extrapolation_warning_rule(water_model_dir):
    """The rule restated: a record exactly where the unscaled value leaves
    [Gmin - 1000 eps, Gmax + 1000 eps], with unscale(scale(G)) == G to
    rounding (SCALECENTER: G = (S - Smin)/factor + Gmean)."""
    from n2p2_amd import synth
    pos, elem, box = synth.water_box(4, spacing=3.105 * 0.85)
    o = Oracle(water_model_dir)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE)
    rows = [l.split() for l in open(os.path.join(water_model_dir, "scaling.data")) if l.strip() and l[0] != "#"]
    stats = {(int(w[0]) - 1, int(w[1]) - 1): [float(x) for x in w[2:6]] for w in rows}
    G = o.G()
    want = set()
    for i in range(len(elem)):
        for q in range(o.nsf[elem[i]]):
            gmin, gmax, gmean, _ = stats[(elem[i], q)]
            v = G[i, q] / (1.0 / (gmax - gmin)) + gmean  # Smin 0, Smax 1
            if v + 1000 * EPS < gmin or v - 1000 * EPS > gmax:
                want.add((i, q))
    ew = o.extrapolation_warnings()
    assert {(int(a), int(q)) for a, q in zip(ew["atom"], ew["sf"])} == want
    assert len(want) > 50
    assert ((ew["value"] > ew["gmax"]) | (ew["value"] < ew["gmin"])).all()


def _spd(n, rng):
    a = rng.normal(size=(n, n))
    return a @ a.T / n + np.eye(n)


@pytest.mark.parametrize("m", [1, 3])
def test_kalman_standard_information_form(m):
    """KalmanFilter::update, standard type (KalmanFilter.cpp:119-185), pinned
    by the information form of the Kalman update rather than by its own
    formula: with R = I/eta and q = 0, P+^-1 = P^-1 + eta H H^T, the gain
    is K = eta P+ H and w+ = w + K xi; the process noise then adds q to the
    diagonal and anneals q by exp(-qtau) while q > qmin; eta grows by
    exp(etatau) up to etamax before it is used."""
    rng = np.random.default_rng(7 + m)
    n = 12
    P, w, H, xi = _spd(n, rng), rng.normal(size=n), rng.normal(size=(n, m)), rng.normal(size=m)
    eta = 0.7
    par = [0, eta, 0.0, eta, 0, 0, 1, 0.0, 0.0, 0.0]
    P1, w1, par1 = kalman_update(P, w, H, xi, par)
    info = np.linalg.inv(P) + eta * H @ H.T
    assert np.allclose(np.linalg.inv(P1), info, rtol=1e-10, atol=1e-10)
    K = eta * P1 @ H
    assert np.allclose(w1, w + K @ xi, rtol=1e-10, atol=1e-12)
    assert (P1 == P1.T).all()
    # process noise, its annealing and the growing learning rate
    q, qtau, qmin = 0.01, 0.5, 1e-4
    P2, _, par2 = kalman_update(P, w, H, xi, [0, eta, 0.3, 2.0, 0, 0, 1, q, qtau, qmin])
    eta2 = eta * np.exp(0.3)
    P2ref = np.linalg.inv(np.linalg.inv(P) + eta2 * H @ H.T) + q * np.eye(n)
    assert np.allclose(P2, P2ref, rtol=1e-10, atol=1e-12)
    assert abs(par2[1] - eta2) <= 1e-15 and abs(par2[7] - q * np.exp(-qtau)) <= 1e-17


def test_kalman_fading_memory_information_form():
    """Fading-memory type: with A = H^T P H + lambda I, P+ = (P - K X^T) /
    lambda equals (P^-1 + H H^T / lambda)^-1 / lambda, and lambda moves to
    nu lambda + 1 - nu."""
    rng = np.random.default_rng(3)
    n, m = 10, 2
    P, w, H, xi = _spd(n, rng), rng.normal(size=n), rng.normal(size=(n, m)), rng.normal(size=m)
    lam, nu = 0.98, 0.99
    P1, w1, par1 = kalman_update(P, w, H, xi, [1, 0, 0, 0, lam, nu, 1.0, 0.0, 0.0, 0.0])
    ref = np.linalg.inv(np.linalg.inv(P) + H @ H.T / lam) / lam
    assert np.allclose(P1, ref, rtol=1e-10, atol=1e-12)
    K = P @ H @ np.linalg.inv(H.T @ P @ H + lam * np.eye(m))
    assert np.allclose(w1, w + K @ xi, rtol=1e-10, atol=1e-12)
    assert abs(par1[4] - (nu * lam + 1 - nu)) <= 1e-15
    assert abs(par1[6] - 1.0 / (1.0 + par1[4] / 1.0)) <= 1e-15


def test_kalman_reads_lower_triangle():
    """Only the lower triangle of P is read (selfadjointView<Lower>): junk in
    the upper triangle does not change the update."""
    rng = np.random.default_rng(5)
    n = 8
    P, w, H, xi = _spd(n, rng), rng.normal(size=n), rng.normal(size=(n, 1)), rng.normal(size=1)
    par = [0, 1.0, 0.0, 1.0, 0, 0, 1, 0.0, 0.0, 0.0]
    Pj = P.copy()
    Pj[np.triu_indices(n, 1)] = 123.0
    a = kalman_update(P, w, H, xi, par)
    b = kalman_update(Pj, w, H, xi, par)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()
