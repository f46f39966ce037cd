"""Stored dG/dr of the plain FP64 form (Atom::dGdr, Atom::Neighbor::dGdr):
the public derivative layout that downstream code reads
(InterfaceLammps::getForces, InterfaceLammps.cpp:710-722;
Training::collectDGdxia, Training.cpp:3603-3621).

* nnpx_get_dGdr against the oracle, entry for entry and slot for slot
  (Element::setupSymmetryFunctionMemory order, Element.cpp:289-311), within
  1e-11 of max(|value|, 1) (the self derivative is minus a sum over ~100
  neighbour terms, summed in another order), on the 192-atom water box, a 6-atom cell with
  self images, the Cu2S golden (type 9), the compact and the element-weighted
  water models; the oracle's stored derivatives equal the reference classes'
  bit for bit (tests/test_oracle.py).
* The fixtures.h finite-difference procedure of test_SymmetryFunctions.cpp:
  101-155 (h = 1e-7, 1e-6) run through the GPU: all displaced copies of a
  cluster are one batched evaluation.
* The C++ façade's copy (setCopyNeighbors / setCopyDerivatives) on the
  H2O_RPBE-D3 golden.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import n2p2_amd
from n2p2_amd import read_input_data, remap, synth
from oracle_lib import BRUTE, PULL, Oracle, compact_model, kat_model, weighted_model

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
DRIVER = os.path.join(HERE, "cpp", "bin", "test_facade")
TOL = 1e-11


def say(*a):
    print(*a, flush=True)


def compare(m, o, label):
    ld = m.max_sf
    s_g, n_g = m.dGdr(ld)
    s_o, n_o = o.dGdr(ld)
    assert s_g.shape == s_o.shape and n_g.shape == n_o.shape
    es = float((np.abs(s_g - s_o) / np.maximum(np.abs(s_o), 1.0)).max())
    en = float((np.abs(n_g - n_o) / np.maximum(np.abs(n_o), 1.0)).max())
    say("  %s: %d atoms, %d entries x %d slots; self err %.2e, neighbour err %.2e (max |dG/dr| %.3f)"
        % (label, s_g.shape[0], n_g.shape[0], ld, es, en, float(np.abs(n_o).max())))
    # unused slots are zero on both sides
    assert ((n_o == 0) <= (np.abs(n_g) <= TOL)).all()
    assert es <= TOL
    assert en <= TOL
    assert np.abs(n_o).max() > 0


@pytest.mark.parametrize("case", ["water192", "small_cell", "cu2s", "compact", "weighted"])
def test_dGdr_vs_oracle(case, tmp_path):
    say("case: stored dG/dr vs oracle, %s" % case)
    if case == "cu2s":
        d = os.path.join(GOLD, "Cu2S_PBE")
        pos, elem, box = read_input_data(os.path.join(d, "input.data"), ["S", "Cu"])
    elif case == "compact":
        d = compact_model(str(tmp_path / "compact"))
        pos, elem, box = synth.water_box(4)
    elif case == "weighted":
        d = weighted_model(str(tmp_path / "weighted"))
        pos, elem, box = synth.water_box(4)
    elif case == "water192":
        d = synth.water_model(str(tmp_path / "water"))
        pos, elem, box = synth.water_box(4)
    else:
        d = synth.water_model(str(tmp_path / "water"))
        rng = np.random.default_rng(3)
        box = np.eye(3) * 5.0
        pos = rng.random((6, 3)) * 5.0
        elem = np.array([1, 0, 0, 1, 0, 0], np.int32)
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE)
    m = n2p2_amd.Mode(d)
    cl = m.conv_length
    p = remap(pos, box) if box is not None else pos
    m.set_structures(p * cl, elem, None if box is None else (box * cl)[None])
    m.calculate_neighbor_list()
    m.calculate_symmetry_function_groups(True)
    compare(m, o, case)


@pytest.mark.parametrize("name", ["Mono", "Dual", "Mixed"])
def test_dGdr_finite_differences_gpu(name, tmp_path):
    """Every SF type of fixtures.h: dG/dr of atom 0 (self and each neighbour
    entry) against central differences of G computed on the GPU, the 1 + 6n
    displaced copies of the cluster packed into one batched evaluation."""
    k = json.load(open(os.path.join(GOLD, "sf_kat.json")))["sets"][name]
    pos0 = np.array(k["pos"], float)
    elem = np.array(k["elem"], np.int32)
    n = len(elem)
    h = 1e-7
    say("case: dG/dr finite differences on the GPU, %s set (%d functions)" % (name, len(k["functions"])))
    copies = [pos0]
    for ia in range(n):
        for ic in range(3):
            for sg in (1, -1):
                p = pos0.copy()
                p[ia, ic] += sg * h
                copies.append(p)
    P = np.concatenate(copies)
    E = np.tile(elem, len(copies))
    off = np.arange(len(copies) + 1) * n
    worst = 0.0
    for q, row in enumerate(k["functions"]):
        d = str(tmp_path / ("m%d" % q))
        kat_model(d, k["elements"], [row["line"]])
        m = n2p2_amd.Mode(d)
        m.set_structures(P, E, None, atom_offsets=off, periodic=np.zeros(len(copies), np.int32))
        m.calculate_neighbor_list()
        m.calculate_symmetry_function_groups(True)
        G = m.G()[:, 0]
        counts, js, _ = m.neighbors()
        js0 = list(js[:counts[0]])
        s_an, n_an = m.dGdr(m.max_sf)
        num = np.zeros((n, 3))
        c = 1
        for ia in range(n):
            for ic in range(3):
                num[ia, ic] = (G[c * n] - G[(c + 1) * n]) / (2 * h)
                c += 2
        err = np.abs(s_an[0, 0] - num[0]).max()
        for ia in range(1, n):
            an = n_an[js0.index(ia), 0] if ia in js0 else np.zeros(3)
            err = max(err, np.abs(an - num[ia]).max())
        worst = max(worst, float(err))
        assert err <= 1e-6, (row["line"], err)
        assert np.abs(num).max() > 1e-6, row["line"]
    say("  max |analytic - numeric| %.2e" % worst)


def read_facade_dump(path):
    raw = open(path, "rb").read()
    pos = 0

    def i64():
        nonlocal pos
        v = int(np.frombuffer(raw, np.int64, 1, pos)[0])
        pos += 8
        return v

    def vecs():
        nonlocal pos
        n = i64()
        v = np.frombuffer(raw, np.float64, 3 * n, pos).reshape(n, 3)
        pos += 24 * n
        return v

    atoms = []
    while pos < len(raw):
        s = vecs()
        nb = [(i64(), vecs()) for _ in range(i64())]
        atoms.append((s, nb))
    return atoms


def test_facade_copies_dGdr(tmp_path):
    """nnpx::Mode with setCopyNeighbors/setCopyDerivatives fills
    Atom::dGdr and every Neighbor::dGdr (length n_SFD of the neighbour's
    element) as the oracle's Atom objects hold them, H2O_RPBE-D3 golden."""
    if not os.path.exists(DRIVER):
        pytest.fail("%s is not built; run __graft_entry__.build()" % DRIVER)
    d = os.path.join(GOLD, "H2O_RPBE-D3")
    out = str(tmp_path / "dgdr.bin")
    say("case: facade dG/dr copy, H2O_RPBE-D3 golden")
    r = subprocess.run([DRIVER, "dgdr", d, os.path.join(d, "input.data"), out], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout)["flags"] == [1, 1]
    atoms = read_facade_dump(out)
    o = Oracle(d)
    o.read_structure(os.path.join(d, "input.data"))
    o.predict(PULL, BRUTE)
    counts, js, _ = o.neighbors()
    s_o, n_o = o.dGdr(o.max_sf)
    assert len(atoms) == o.n
    p = 0
    worst = 0.0
    for i, (s, nb) in enumerate(atoms):
        ns = s.shape[0]
        worst = max(worst, float((np.abs(s - s_o[i, :ns]) / np.maximum(np.abs(s_o[i, :ns]), 1.0)).max()))
        assert not s_o[i, ns:].any()
        assert len(nb) == counts[i]
        for k, (j, v) in enumerate(nb):
            assert j == js[p + k]
            nsd = v.shape[0]
            worst = max(worst, float((np.abs(v - n_o[p + k, :nsd]) / np.maximum(np.abs(n_o[p + k, :nsd]), 1.0)).max())
                        if nsd else 0.0)
            assert not n_o[p + k, nsd:].any()
        p += counts[i]
    say("  max rel err %.2e over %d atoms, %d entries" % (worst, len(atoms), p))
    assert worst <= TOL
