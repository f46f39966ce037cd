"""C++ façade (nnpx::Mode / Prediction / Structure) and the nnpx-predict
command: the host side a C++ caller of n2p2 links against.

CPU tests: structure parsing, remap and unit conversion match the Python host
code (itself checked against the oracle), output.data round trip, usage text.
GPU tests: nnpx-predict on the H2O_RPBE-D3 golden (the test_Prediction.cpp
case) in both forms, and the step-by-step Mode API.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from n2p2_amd import read_input_data, remap

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden", "H2O_RPBE-D3")
PREDICT = os.path.join(ROOT, "n2p2_amd", "bin", "nnpx-predict")
DRIVER = os.path.join(HERE, "cpp", "bin", "test_facade")
E_REL = 1e-10
F_ABS = 1e-8
F_RMS_MIXED = 1e-3


def need(path):
    if not os.path.exists(path):
        pytest.fail("%s is not built; run __graft_entry__.build()" % path)


def run(args, cwd=None, timeout=120):
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout, r.stderr


def test_structure_parse_remap_and_units(tmp_path):
    need(DRIVER)
    out = str(tmp_path / "out.data")
    rc, so, se = run([DRIVER, "structure", GOLD, os.path.join(GOLD, "input.data"), out])
    assert rc == 0, se
    d = json.loads(so)
    pos, elem, box = read_input_data(os.path.join(GOLD, "input.data"), ["H", "O"])
    cl = [float(w[1]) for w in (line.split() for line in open(os.path.join(GOLD, "input.nn")))
          if w and w[0] == "conv_length"][0]
    assert d["numAtoms"] == len(elem) and d["isPeriodic"] == 1
    # remap in the façade equals the host remap bit for bit; units scale exactly
    assert (np.array(d["r"]) == remap(pos, box) * cl).all()
    assert (np.array(d["box"]) == box * cl).all()
    # output.data (ref=true) round trip: same atoms and box as the input
    pos2, elem2, box2 = read_input_data(out, ["H", "O"])
    assert (elem2 == elem).all()
    assert np.abs(pos2 - remap(pos, box)).max() == 0.0
    assert (box2 == box).all()


def test_predict_usage_without_arguments():
    need(PREDICT)
    rc, so, _ = run([PREDICT])
    assert rc == 1 and "USAGE" in so


def read_table(path, ncol):
    rows = [line.split() for line in open(path) if line.strip() and not line.startswith("#")]
    assert all(len(r) == ncol for r in rows)
    return np.array(rows, dtype=float)


def golden_run_dir(tmp_path, gdir=GOLD):
    d = tmp_path / "run"
    d.mkdir()
    for f in os.listdir(gdir):
        if f != "expected.json":
            shutil.copy(os.path.join(gdir, f), d / f)
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("name,natoms,z", [("H2O_RPBE-D3", 1080, {1, 8}), ("Cu2S_PBE", 144, {16, 29}),
                                           ("Anisole_SCAN", 256, {1, 6, 8})])
def test_nnpx_predict_golden_plain(tmp_path, name, natoms, z):
    need(PREDICT)
    print("case: nnpx-predict %s plain" % name, flush=True)
    gdir = os.path.join(HERE, "golden", name)
    d = golden_run_dir(tmp_path, gdir)
    rc, so, se = run([PREDICT, "0"], cwd=str(d))
    assert rc == 0, se + open(d / "nnp-predict.log").read()
    exp = json.load(open(os.path.join(gdir, "expected.json")))
    e = read_table(d / "energy.out", 6)
    f = read_table(d / "nnforces.out", 8)
    a = read_table(d / "nnatoms.out", 7)
    assert e.shape[0] == 1 and int(e[0, 1]) == natoms
    dE = abs(e[0, 3] - exp["energy"]) / abs(exp["energy"])
    dF = np.abs(f[:, 5:8] - np.array(exp["forces"])).max()
    print("  E rel %.2e  max|dF| %.2e" % (dE, dF), flush=True)
    assert dE <= E_REL
    assert dF <= F_ABS
    assert a.shape == (natoms, 7) and set(a[:, 2].astype(int)) == z
    assert sum(1 for line in open(d / "output.data") if line.startswith("atom")) == natoms


@pytest.mark.gpu
def test_nnpx_predict_golden_mixed(tmp_path):
    need(PREDICT)
    print("case: nnpx-predict H2O_RPBE-D3 mixed", flush=True)
    d = golden_run_dir(tmp_path)
    rc, so, se = run([PREDICT, "0", "--mixed"], cwd=str(d))
    assert rc == 0, se
    exp = json.load(open(os.path.join(GOLD, "expected.json")))
    f = read_table(d / "nnforces.out", 8)
    rms = float(np.sqrt(np.mean((f[:, 5:8] - np.array(exp["forces"])) ** 2)))
    print("  F rms %.2e" % rms, flush=True)
    assert rms <= F_RMS_MIXED


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_mode_step_api_golden(form):
    need(DRIVER)
    print("case: façade step API, H2O_RPBE-D3", form, flush=True)
    args = [DRIVER, "steps", GOLD, os.path.join(GOLD, "input.data")] + (["mixed"] if form == "mixed" else [])
    rc, so, se = run(args)
    assert rc == 0, se
    d = json.loads(so)
    exp = json.load(open(os.path.join(GOLD, "expected.json")))
    F = np.array(d["f"])
    Fg = np.array(exp["forces"])
    assert d["flags"] == [1, 1, 1]
    assert len(d["G0"]) == len(d["dEdG0"]) > 0
    assert min(d["numNeighbors"]) > 0
    if form == "plain":
        assert abs(d["energy"] - exp["energy"]) / abs(exp["energy"]) <= E_REL
        assert np.abs(F - Fg).max() <= F_ABS
    else:
        assert float(np.sqrt(np.mean((F - Fg) ** 2))) <= F_RMS_MIXED
