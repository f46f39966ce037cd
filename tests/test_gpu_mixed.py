"""Mixed-precision form (FP32 symmetry functions, BF16 MFMA networks, fused
force kernel, FP64 reductions) against the FP64 oracle on the GPU.

Stated bar: force RMS error <= 1e-3 eV/A against the FP64 oracle; the errors
of each case are printed so the test log states them per configuration.
"""
import os

import numpy as np
import pytest

import n2p2_amd
from n2p2_amd import read_input_data, remap, synth
from oracle_lib import BRUTE, CELL, PULL, PUSH, Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "H2O_RPBE-D3")
F_RMS = 1e-3


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def water_dir(tmp_path_factory):
    return synth.water_model(str(tmp_path_factory.mktemp("water")))


def test_mfma_lane_layout():
    m = n2p2_amd.Mode()
    err = m.mfma_selftest()
    say("mfma self-test max abs err %.3e" % err)
    assert err <= 1e-6


def errors(m, o):
    F, Fo = m.forces(), o.forces()
    E, Eo = m.energies(), np.atleast_1d(o.energy())
    rms = float(np.sqrt(np.mean((F - Fo) ** 2)))
    mx = float(np.abs(F - Fo).max())
    erel = float(np.abs(E[0] - Eo[0]) / abs(Eo[0]))
    return rms, mx, erel


@pytest.mark.parametrize("terms", [3, 1])
def test_water192_mixed(terms, water_dir):
    say("case: water-192 mixed, bf16 terms", terms)
    pos, elem, box = synth.water_box(4)
    m = n2p2_amd.Mode(water_dir)
    m.set_option(n2p2_amd.OPT_BF16_TERMS, terms)
    m.set_structures(remap(pos, box), elem, box[None])
    m.evaluate(n2p2_amd.MIXED, forces=True)
    o = Oracle(water_dir)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE)
    G, Go = m.G(), o.G()
    gerr = float((np.abs(G - Go) / np.maximum(np.abs(Go), 1.0)).max())
    dEdG, dEdGo = m.dEdG(), o.dEdG()
    derr = float(np.abs(dEdG - dEdGo).max() / np.abs(dEdGo).max())
    rms, mx, erel = errors(m, o)
    say("  G err %.2e  dEdG rel %.2e  E rel %.2e  F rms %.2e max %.2e (|F| rms %.3f)"
        % (gerr, derr, erel, rms, mx, np.sqrt(np.mean(o.forces() ** 2))))
    assert gerr <= 1e-5
    assert rms <= F_RMS
    if terms == 3:
        # regression bounds near the achieved errors (force RMS ~1.4e-6)
        assert rms <= 1e-5
        assert derr <= 1e-4
        assert erel <= 1e-5


def test_water12288_mixed(water_dir):
    say("case: water-12288 mixed")
    pos, elem, box = synth.water_box(16)
    m = n2p2_amd.Mode(water_dir)
    m.set_structures(remap(pos, box), elem, box[None])
    m.evaluate(n2p2_amd.MIXED, forces=True)
    o = Oracle(water_dir)
    o.set_structure(pos, elem, box)
    o.predict(PULL, CELL)
    rms, mx, erel = errors(m, o)
    say("  E rel %.2e  F rms %.2e max %.2e" % (erel, rms, mx))
    assert rms <= F_RMS
    assert rms <= 1e-5  # regression bound (achieved ~1.4e-6)
    assert erel <= 1e-5


def test_golden_cu2s_mixed():
    """Cu2S_PBE golden (wide angular type 9, CT_POLY2, SCALESIGMA, softplus)
    in the mixed form against the reference's forces."""
    import json
    say("case: Cu2S_PBE mixed")
    gdir = os.path.join(HERE, "golden", "Cu2S_PBE")
    pos, elem, box = read_input_data(os.path.join(gdir, "input.data"), ["S", "Cu"])
    m = n2p2_amd.Mode(gdir)
    cl, ce = m.conv_length, m.conv_energy
    m.set_structures(remap(pos, box) * cl, elem, None if box is None else (box * cl)[None])
    m.evaluate(n2p2_amd.MIXED, forces=True)
    F = m.forces() * (1.0 / (ce / cl))
    Fg = np.array(json.load(open(os.path.join(gdir, "expected.json")))["forces"])
    rms = float(np.sqrt(np.mean((F - Fg) ** 2)))
    say("  F rms %.2e max %.2e (|F| rms %.3e)" % (rms, np.abs(F - Fg).max(), np.sqrt(np.mean(Fg ** 2))))
    assert rms <= F_RMS


def test_golden_h2o_mixed():
    """The H2O_RPBE-D3 model (CT_TANHU, rs > 0 radial terms, normalised units)."""
    say("case: H2O_RPBE-D3 mixed")
    pos, elem, box = read_input_data(os.path.join(GOLD, "input.data"), ["H", "O"])
    m = n2p2_amd.Mode(GOLD)
    cl, ce = m.conv_length, m.conv_energy
    m.set_structures(remap(pos, box) * cl, elem, None if box is None else (box * cl)[None])
    m.evaluate(n2p2_amd.MIXED, forces=True)
    o = Oracle(GOLD)
    o.read_structure(os.path.join(GOLD, "input.data"))
    o.predict(PULL, BRUTE)
    F = m.forces() * (1.0 / (ce / cl))  # physical units (eV/A here: Ha/bohr scaled)
    Fo = o.forces()
    rms = float(np.sqrt(np.mean((F - Fo) ** 2)))
    say("  F rms %.2e max %.2e (|F| rms %.3e)" % (rms, np.abs(F - Fo).max(), np.sqrt(np.mean(Fo ** 2))))
    assert rms <= F_RMS


@pytest.mark.parametrize("force_mode", [PULL, PUSH])
def test_batched_mixed(force_mode, water_dir):
    say("case: mixed batch of 24 structures, force mode", force_mode)
    pos, elem, boxes, off = synth.random_batch(n_struct=24, nmin=16, nmax=200)
    gpos = np.vstack([remap(pos[off[s]:off[s + 1]], boxes[s]) for s in range(24)])
    m = n2p2_amd.Mode(water_dir)
    m.set_option(n2p2_amd.OPT_FORCE_MODE, force_mode)
    m.set_structures(gpos, elem, boxes, off)
    m.evaluate(n2p2_amd.MIXED, forces=True)
    E, F = m.energies(), m.forces()
    Fo = np.zeros_like(F)
    for s in range(24):
        a, b = off[s], off[s + 1]
        o = Oracle(water_dir)
        o.set_structure(pos[a:b], elem[a:b], boxes[s])
        o.predict(force_mode, BRUTE)
        Fo[a:b] = o.forces()
        assert abs(E[s] - o.energy()) <= 1e-5 * max(1.0, abs(o.energy())), s
    rms = float(np.sqrt(np.mean((F - Fo) ** 2)))
    say("  F rms %.2e max %.2e" % (rms, np.abs(F - Fo).max()))
    assert rms <= F_RMS


def test_fcc_four_species_mixed(tmp_path_factory):
    """Four-species FCC model (80 SFs: 10-member radial groups, 4-member
    angular groups, cutoff 6.0 A): 5^3 cells, 500 atoms, cell list."""
    say("case: FCC 500 atoms mixed")
    d = synth.fcc_model(str(tmp_path_factory.mktemp("fcc")))
    pos, elem, box = synth.fcc_lattice(n_cells=5)
    m = n2p2_amd.Mode(d)
    m.set_structures(remap(pos, box), elem, box[None])
    m.evaluate(n2p2_amd.MIXED, forces=True)
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE)
    G, Go = m.G(), o.G()
    gerr = float((np.abs(G - Go) / np.maximum(np.abs(Go), 1.0)).max())
    rms, mx, erel = errors(m, o)
    say("  G err %.2e  E rel %.2e  F rms %.2e max %.2e" % (gerr, erel, rms, mx))
    assert gerr <= 1e-5
    assert rms <= F_RMS
    assert erel <= 1e-5


def test_plain_forces_match_finite_differences(water_dir):
    """FP64 GPU forces vs central differences of the FP64 GPU energy
    (nnp-checkf.cpp:173-295 semantics, threshold 10*h^2)."""
    say("case: plain FP64 forces vs finite differences")
    pos, elem, box = synth.water_box(4)
    pos = remap(pos, box)
    m = n2p2_amd.Mode(water_dir)
    m.set_structures(pos, elem, box[None])
    m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
    F = m.forces()
    h = 1e-4
    worst = 0.0
    for i in (0, 1, 2, 100):
        for c in range(3):
            e = []
            for sgn in (-1, 1):
                p = pos.copy()
                p[i, c] += sgn * h
                m.update_positions(remap(p, box))
                m.evaluate(n2p2_amd.PLAIN_FP64, forces=False)
                e.append(m.energies()[0])
            fd = -(e[1] - e[0]) / (2 * h)
            worst = max(worst, abs(fd - F[i, c]))
    say("  max |F - F_fd| = %.2e (threshold %.1e)" % (worst, 10 * h * h))
    assert worst <= 10 * h * h


def test_mixed_forces_match_finite_differences(water_dir):
    """The mixed form's fused backward is the derivative of the mixed form's
    own energy: its forces against central differences of the mixed energy
    (h = 5e-3 A, where the FP32 symmetry functions' rounding noise, ~1e-6 eV
    in E, is small against E(x + h) - E(x - h); at h = 1e-3 the measured
    deviation was 3.5e-3), 12 components of the 192-atom box."""
    say("case: mixed forces vs finite differences of the mixed energy")
    pos, elem, box = synth.water_box(4)
    pos = remap(pos, box)
    m = n2p2_amd.Mode(water_dir)
    m.set_structures(pos, elem, box[None])
    m.evaluate(n2p2_amd.MIXED, forces=True)
    F = m.forces()
    h = 5e-3
    worst = 0.0
    for i in (0, 1, 2, 100):
        for c in range(3):
            e = []
            for sgn in (-1, 1):
                p = pos.copy()
                p[i, c] += sgn * h
                m.update_positions(remap(p, box))
                m.evaluate(n2p2_amd.MIXED, forces=False)
                e.append(m.energies()[0])
            fd = -(e[1] - e[0]) / (2 * h)
            worst = max(worst, abs(fd - F[i, c]))
    say("  max |F - F_fd| = %.2e (|F| max %.3f)" % (worst, float(np.abs(F).max())))
    assert worst <= 5e-3


def test_autotune_keeps_results(water_dir):
    """The autotuner only changes launch shapes: forces before and after
    agree to summation-order noise, and explicit shapes round-trip."""
    say("case: autotune on water-1536")
    pos, elem, box = synth.water_box(8)
    m = n2p2_amd.Mode(water_dir)
    m.set_structures(remap(pos, box), elem, box[None])
    m.evaluate(n2p2_amd.MIXED, forces=True)
    F0 = m.forces()
    tu = m.autotune()
    say("  chosen", tu)
    assert all(v % 64 == 0 for v in (tu["trip_sf"], tu["trip_force"], tu["trip_tangent"]))
    m.evaluate(n2p2_amd.MIXED, forces=True)
    d = float(np.abs(m.forces() - F0).max())
    say("  max|dF| after tuning %.2e" % d)
    assert d <= 1e-4  # bf16 operand splits amplify FP32-ulp input changes to ~1e-5 relative
    m.set_tuning(128, 320, 512, 16)
    tu2 = m.tuning()
    assert {k: tu2[k] for k in ("trip_sf", "trip_force", "trip_tangent", "grad_atoms")} == \
        {"trip_sf": 128, "trip_force": 320, "trip_tangent": 512, "grad_atoms": 16}
    m.evaluate(n2p2_amd.MIXED, forces=True)
    assert float(np.abs(m.forces() - F0).max()) <= 1e-4
    with pytest.raises(n2p2_amd.NNPError):
        m.set_tuning(100, 256, 512, 16)
    with pytest.raises(n2p2_amd.NNPError):
        m.set_tuning(0, 256, 512, 16)


@pytest.mark.parametrize("name,symbols", [("Anisole_SCAN", ["H", "C", "O"]), ("DMABN_SCAN", ["H", "C", "N"]),
                                          ("Ethylbenzene_SCAN", ["H", "C"])])
def test_golden_scan_mixed(name, symbols):
    """The compact-function goldens (types 20 and 22, ~350 SFs per element,
    polynomial cores, normalised units) in the mixed form: FP32 compact terms
    (values kernel, then the contracted force kernel), the network input
    wider than 128; forces against the reference's golden forces."""
    import json
    say("case: %s mixed" % name)
    gdir = os.path.join(HERE, "golden", name)
    pos, elem, box = read_input_data(os.path.join(gdir, "input.data"), symbols)
    m = n2p2_amd.Mode(gdir)
    cl, ce = m.conv_length, m.conv_energy
    m.set_structures(remap(pos, box) * cl, elem, None if box is None else (box * cl)[None])
    m.evaluate(n2p2_amd.MIXED, forces=True)
    F = m.forces() * (1.0 / (ce / cl))
    exp = json.load(open(os.path.join(gdir, "expected.json")))
    Fg = np.array(exp["forces"])
    rms = float(np.sqrt(np.mean((F - Fg) ** 2)))
    E1 = m.energies()[0]
    m.evaluate(n2p2_amd.PLAIN_FP64, forces=False)
    dE = abs(E1 - m.energies()[0]) / ce / len(elem)  # physical units per atom
    say("  F rms %.2e max %.2e (|F| rms %.3e); |dE| per atom vs plain %.2e"
        % (rms, np.abs(F - Fg).max(), np.sqrt(np.mean(Fg ** 2)), dE))
    assert rms <= F_RMS
    assert dE <= 1e-4
    # the values path alone (no forces) gives the same energy
    m.evaluate(n2p2_amd.MIXED, forces=False)
    assert m.energies()[0] == E1


@pytest.mark.parametrize("nbox", [4, 5])
def test_compact_model_mixed(nbox, tmp_path):
    """A synthetic compact model with every subtype (types 20, 21, 22;
    polynomial, asymmetric and exponential cores; narrow members sharing
    (rl, rc)) on periodic water boxes, mixed form against the FP64 oracle,
    pull and push forces."""
    from oracle_lib import compact_model
    d = compact_model(str(tmp_path / "compact"))
    pos, elem, box = synth.water_box(nbox)
    pos = remap(pos, box)
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    for mode, name in ((PULL, "pull"), (PUSH, "push")):
        o.predict(mode, BRUTE)
        Fo, Eo = o.forces(), o.energy()
        m = n2p2_amd.Mode(d)
        m.set_option(n2p2_amd.OPT_FORCE_MODE, mode)
        m.set_structures(pos, elem, box[None])
        m.evaluate(n2p2_amd.MIXED, forces=True)
        F, E = m.forces(), m.energies()[0]
        rms = float(np.sqrt(np.mean((F - Fo) ** 2)))
        dE = abs(E - Eo) / len(elem)
        say("  %d atoms, %s: E %.6e, |dE| per atom %.2e  F rms %.2e max %.2e (|F| rms %.3e)"
            % (len(elem), name, Eo, dE, rms, np.abs(F - Fo).max(), np.sqrt(np.mean(Fo ** 2))))
        assert rms <= F_RMS
        assert dE <= 1e-4


@pytest.mark.parametrize("nbox", [4, 5])
def test_weighted_model_mixed(nbox, tmp_path):
    """Element-weighted types 12 and 13 (Gaussian, every neighbour element,
    terms times Z_j and Z_j Z_k) interleaved with types 2/3 and the weighted
    compact types 23/24/25, mixed form against the FP64 oracle, pull and
    push forces."""
    from oracle_lib import weighted_model
    d = weighted_model(str(tmp_path / "weighted"))
    pos, elem, box = synth.water_box(nbox)
    pos = remap(pos, box)
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    for mode, name in ((PULL, "pull"), (PUSH, "push")):
        o.predict(mode, BRUTE)
        Fo, Eo = o.forces(), o.energy()
        m = n2p2_amd.Mode(d)
        m.set_option(n2p2_amd.OPT_FORCE_MODE, mode)
        m.set_structures(pos, elem, box[None])
        m.evaluate(n2p2_amd.MIXED, forces=True)
        F, E = m.forces(), m.energies()[0]
        rms = float(np.sqrt(np.mean((F - Fo) ** 2)))
        dE = abs(E - Eo) / len(elem)
        say("  %d atoms, %s: E %.6e, |dE| per atom %.2e  F rms %.2e max %.2e (|F| rms %.3e)"
            % (len(elem), name, Eo, dE, rms, np.abs(F - Fo).max(), np.sqrt(np.mean(Fo ** 2))))
        assert rms <= F_RMS
        assert dE <= 1e-4


@pytest.mark.parametrize("remapped", [True, False])
def test_position_update_regrows_lists(water_dir, remapped):
    """nnpx_update_positions followed by an evaluation must give the results
    of a fresh context on the new positions, (1) when one atom gains more
    neighbours than the largest list of the last checked build (the mixed
    kernels stage that many slots; the overflow is flagged and the
    evaluation repeated) and (2) when atoms leave the cell of a small
    periodic structure (the brute-force image ranges are pruned only while
    every atom is inside).  192-atom water cell (12.42 A < 2 rc, brute
    force); both forms."""
    say("case: position update, water-192, remapped", remapped)
    pos, elem, box = synth.water_box(4)
    pos = remap(pos, box)
    o = Oracle(water_dir)
    o.set_structure(pos, elem, box)
    o.build_neighbors(BRUTE)
    counts = o.neighbors()[0]
    a = int(np.argmax(counts))
    rng = np.random.default_rng(17)
    new = pos.copy()
    L = box[0, 0]
    # move 24 atoms that are not neighbours of atom a into a 5.5 A shell around it
    d = new - new[a]
    d -= L * np.round(d / L)
    far = np.where(np.linalg.norm(d, axis=1) > 6.5)[0][:24]
    for q in far:
        for _ in range(1000):
            v = rng.normal(size=3)
            v *= rng.uniform(2.0, 5.5) / np.linalg.norm(v)
            p = new[a] + v
            dd = new - p
            dd -= L * np.round(dd / L)
            if np.linalg.norm(dd, axis=1).min() > 1.2:
                new[q] = p
                break
    if remapped:
        new = remap(new, box)
    else:
        new[::7] += box[0]  # a seventh of the atoms one cell over
    o.set_structure(new, elem, box)
    o.build_neighbors(BRUTE)
    grown = int(o.neighbors()[0].max())
    say("  largest list %d -> %d" % (counts.max(), grown))
    assert grown >= ((counts.max() + 31) // 32) * 32
    for precision in (n2p2_amd.PLAIN_FP64, n2p2_amd.MIXED):
        m = n2p2_amd.Mode(water_dir)
        m.set_structures(pos, elem, box[None])
        m.evaluate(precision, forces=True)
        m.update_positions(new)
        m.evaluate(precision, forces=True)
        F1, E1 = m.forces(), m.energies()[0]
        f = n2p2_amd.Mode(water_dir)
        f.set_structures(new, elem, box[None])
        f.evaluate(precision, forces=True)
        F2, E2 = f.forces(), f.energies()[0]
        dF = float(np.abs(F1 - F2).max())
        say("  precision %d: max |dF| %.2e  dE %.2e" % (precision, dF, abs(E1 - E2)))
        assert dF <= 1e-9 * max(np.abs(F2).max(), 1.0)
        assert abs(E1 - E2) <= 1e-9 * abs(E2)


@pytest.mark.parametrize("name", ["Mono", "Dual", "Mixed"])
def test_sf_known_answers_mixed(name, tmp_path):
    """test/cpp/fixtures.h known answers of every SF type (2, 3, 9, 12, 13,
    20-25) on the fixture clusters through the mixed form: G of atom 0 to
    FP32 accuracy (relative 1e-5 of max(|G|, 1)), forces within the force
    bar of the oracle's."""
    import json
    from oracle_lib import kat_model
    k = json.load(open(os.path.join(HERE, "golden", "sf_kat.json")))["sets"][name]
    say("case: SF known answers, mixed form, %s set (%d functions)" % (name, len(k["functions"])))
    worst = worstF = 0.0
    for q, row in enumerate(k["functions"]):
        d = str(tmp_path / ("m%d" % q))
        kat_model(d, k["elements"], [row["line"]])
        pos = np.array(k["pos"], float)
        elem = np.array(k["elem"], np.int32)
        m = n2p2_amd.Mode(d)
        m.set_structures(pos, elem, None)
        m.evaluate(n2p2_amd.MIXED, forces=True)
        g = m.G()[0, 0]
        o = Oracle(d)
        o.set_structure(pos, elem, None)
        o.predict(PULL, BRUTE)
        dF = float(np.abs(m.forces() - o.forces()).max())
        err = abs(g - row["value"]) / max(abs(row["value"]), 1.0)
        worst = max(worst, err)
        worstF = max(worstF, dF)
        assert err <= 1e-5, (row["line"], g, row["value"])
        assert dF <= F_RMS, (row["line"], dF)
    say("  max rel |G - KAT| %.2e, max |dF| %.2e" % (worst, worstF))


@pytest.mark.parametrize("kind", ["water", "weighted"])
def test_sf_gather_identical(kind, water_dir, tmp_path):
    """The SF kernel with and without the per-group copy of the partner
    geometry in list order (forced on, forced off, automatic): the same
    values loaded another way, so G agrees bit for bit; energies and forces
    differ at most by the summation order of the force kernel's FP64 device
    atomics (any two evaluations do)."""
    from oracle_lib import weighted_model
    d = water_dir if kind == "water" else weighted_model(str(tmp_path / "weighted"))
    say("case: sf gather on/off/auto, %s model, water-192" % kind)
    pos, elem, box = synth.water_box(4)
    m = n2p2_amd.Mode(d)
    m.set_structures(remap(pos, box), elem, box[None])
    out = []
    for g in (0, 1, -1):
        m.set_option(n2p2_amd.OPT_SF_GATHER, g)
        m.evaluate(n2p2_amd.MIXED, forces=True)
        out.append((m.G().copy(), m.energies().copy(), m.forces().copy()))
    for g, o in zip((1, -1), out[1:]):
        assert np.array_equal(out[0][0], o[0]), "gather %d: G differs" % g
        assert np.allclose(out[0][1], o[1], rtol=1e-13, atol=0.0), "gather %d: energies differ" % g
        assert np.abs(out[0][2] - o[2]).max() <= 1e-13 * np.abs(o[2]).max(), "gather %d: forces differ" % g


@pytest.mark.parametrize("kind", ["fcc", "water"])
def test_angular_merge(kind, water_dir, tmp_path):
    """All angular groups of an element in one triplet walk (merged) against
    the per-group walks and the oracle, with and without the partner gather:
    the same terms summed in another order, so G, energies and forces agree
    to FP32 summation order and both meet the oracle bounds.  FCC: 10
    four-member groups per element (merged by default); water: 3 ten-member
    groups (per-group walks by default)."""
    if kind == "fcc":
        d = synth.fcc_model(str(tmp_path / "fcc"))
        pos, elem, box = synth.fcc_lattice(n_cells=5)
    else:
        d = water_dir
        pos, elem, box = synth.water_box(4)
    say("case: angular merge on/off, %s, %d atoms" % (kind, len(elem)))
    m = n2p2_amd.Mode(d)
    m.set_structures(remap(pos, box), elem, box[None])
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE)
    Go = o.G()
    out = {}
    for mg in (0, 1):
        for gather in (0, 1):
            m.set_option(n2p2_amd.OPT_ANGULAR_MERGE_SF, mg)
            m.set_option(n2p2_amd.OPT_ANGULAR_MERGE_FORCE, mg)
            m.set_option(n2p2_amd.OPT_SF_GATHER, gather)
            m.evaluate(n2p2_amd.MIXED, forces=True)
            G = m.G().copy()
            gerr = float((np.abs(G - Go) / np.maximum(np.abs(Go), 1.0)).max())
            rms, mx, erel = errors(m, o)
            say("  merge %d gather %d: G err %.2e  E rel %.2e  F rms %.2e max %.2e" % (mg, gather, gerr, erel, rms, mx))
            assert gerr <= 1e-5 and erel <= 1e-5 and rms <= 1e-5
            out[(mg, gather)] = (G, m.energies().copy(), m.forces().copy())
    G0, E0, F0 = out[(0, 0)]
    for key, (G, E, F) in out.items():
        dg = float((np.abs(G - G0) / np.maximum(np.abs(G0), 1.0)).max())
        df = float(np.abs(F - F0).max() / np.abs(F0).max())
        say("  %s vs per-group walk: G %.2e  F %.2e" % (key, dg, df))
        assert dg <= 2e-6 and df <= 2e-5
    m.set_option(n2p2_amd.OPT_ANGULAR_MERGE_SF, -1)
    m.set_option(n2p2_amd.OPT_ANGULAR_MERGE_FORCE, -1)
    assert m.get_option(n2p2_amd.OPT_ANGULAR_MERGE_SF) == -1


@pytest.mark.parametrize("kind", ["water", "fcc"])
def test_pair_lists(kind, water_dir, tmp_path):
    """Angular pair lists kept by the SF kernel for the force kernel.  With
    lists on, the first evaluation has none yet (the force kernel tests its
    candidates in the same block order and buffers), reading the results
    sizes them, the next evaluation reads them: G and energies agree bit for
    bit, forces to the order of the FP64 device atomics on F (any two
    evaluations do).  A denser structure under the lists sized for the first
    overflows some walks (they test again) and must still agree with an
    evaluation without lists.  Off, the force kernel walks rows: the same
    terms in another order, within FP32 summation order; every variant
    meets the oracle bounds."""
    if kind == "fcc":
        d = synth.fcc_model(str(tmp_path / "fcc"))
        pos, elem, box = synth.fcc_lattice(n_cells=5)
    else:
        d = water_dir
        pos, elem, box = synth.water_box(4)
    say("case: pair lists on/off, %s, %d atoms" % (kind, len(elem)))

    def run(m, p, b):
        m.set_structures(remap(p, b), elem, b[None])
        m.evaluate(n2p2_amd.MIXED, forces=True)
        return m.G().copy(), m.energies().copy(), m.forces().copy()

    def same(a, b, what):
        assert np.array_equal(a[0], b[0]), "%s: G differs" % what
        assert np.array_equal(a[1], b[1]), "%s: energies differ" % what
        df = float(np.abs(a[2] - b[2]).max() / np.abs(b[2]).max())
        say("  %s: forces %.2e" % (what, df))
        assert df <= 1e-13, "%s: forces differ" % what

    o = Oracle(d)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE)
    m = n2p2_amd.Mode(d)
    m.set_option(n2p2_amd.OPT_PAIR_LISTS, 1)
    first = run(m, pos, box)      # no lists yet
    second = run(m, pos, box)     # the results read above sized them
    same(first, second, "lists read vs tested")
    rms, mx, erel = errors(m, o)
    say("  lists on: E rel %.2e  F rms %.2e max %.2e" % (erel, rms, mx))
    assert erel <= 1e-5 and rms <= 1e-5
    # denser: more pairs per atom than the lists were sized for
    dense = (pos * 0.9, box * 0.9)
    partly = run(m, *dense)
    m2 = n2p2_amd.Mode(d)
    m2.set_option(n2p2_amd.OPT_PAIR_LISTS, 1)
    tested = run(m2, *dense)
    same(partly, tested, "dense, lists partly kept vs tested")
    refit = run(m, *dense)
    same(refit, tested, "dense, lists refitted vs tested")
    # off: rows walk
    m.set_option(n2p2_amd.OPT_PAIR_LISTS, 0)
    rows = run(m, pos, box)
    assert m.get_option(n2p2_amd.OPT_PAIR_LISTS) == 0
    dg = float(np.abs(rows[0] - second[0]).max())
    df = float(np.abs(rows[2] - second[2]).max() / np.abs(second[2]).max())
    say("  rows walk vs lists: G %.2e  F %.2e" % (dg, df))
    assert dg == 0.0 and df <= 2e-5
    rms, mx, erel = errors(m, o)
    say("  lists off: E rel %.2e  F rms %.2e max %.2e" % (erel, rms, mx))
    assert erel <= 1e-5 and rms <= 1e-5
