"""nnpx-dataset, the batched data-set evaluation of nnp-dataset
(src/application/nnp-dataset.cpp): every structure of a multi-structure
input.data evaluated on the GPU in packed batches, energy.comp /
forces.comp / output.data / sensitivity files and the RMSE/MAE table in the
reference's formats.

CPU test: usage text.  GPU tests: a 12-structure water set whose reference
energies and forces come from a second ("teacher") network through the
oracle, with atomic energy offsets; the tool's predictions equal the oracle's
for the evaluated network (to the printed precision), its error metrics equal
Structure::updateError / Dataset::collectError recomputed from the printed
columns, batching does not change results, and the mixed form stays within
its force bar.  The H2O_RPBE-D3 golden (normalised units, energies only)
reproduces the golden energy."""
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from n2p2_amd import synth, write_output_data
from oracle_lib import BRUTE, PULL, Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TOOL = os.path.join(ROOT, "n2p2_amd", "bin", "nnpx-dataset")
OFFSETS = {"H": -0.5, "O": -75.0}


def say(*a):
    print(*a, flush=True)


def need(path):
    if not os.path.exists(path):
        pytest.fail("%s is not built; run __graft_entry__.build()" % path)


def run(args, cwd=None, timeout=300):
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout, r.stderr


def read_table(path, ncol):
    rows = [line.split() for line in open(path) if line.strip() and not line.startswith("#")]
    assert all(len(r) == ncol for r in rows), path
    return np.array(rows, dtype=float)


def error_table(path):
    """ENERGY and FORCES rows of the log's error table (physical units)."""
    out = {}
    for line in open(path):
        w = line.split()
        if w and w[0] in ("ENERGY", "FORCES"):
            out[w[0]] = [float(x) for x in w[1:] if x != "|"]
    return out


def test_dataset_usage_without_arguments():
    need(TOOL)
    rc, so, _ = run([TOOL])
    assert rc == 1 and "USAGE" in so


@pytest.fixture(scope="module")
def water_set_dir(tmp_path_factory):
    """Student model + input.data with teacher energies/forces (offsets added)."""
    base = tmp_path_factory.mktemp("dataset")
    d = synth.water_model(str(base / "run"))
    with open(os.path.join(d, "input.nn"), "a") as f:
        f.write("use_short_forces\n")
        for s, v in OFFSETS.items():
            f.write("atom_energy %s %.10f\n" % (s, v))
    teacher = synth.water_model(str(base / "teacher"), seed=synth.SEED_TEACHER)
    pos, elem, boxes, off = synth.water_set(n_struct=12, mmin=12, mmax=40)
    t = Oracle(teacher)
    parts = []
    ref = []
    for s in range(12):
        a, b = off[s], off[s + 1]
        t.set_structure(pos[a:b], elem[a:b], boxes[s])
        t.predict(PULL, BRUTE)
        offset = sum(OFFSETS["HO"[e]] for e in elem[a:b])
        p = os.path.join(str(base), "s%d.data" % s)
        write_output_data(p, pos[a:b], elem[a:b], ("H", "O"), t.forces(), t.energy() + offset, boxes[s],
                          comment="structure %d" % s)
        parts.append(open(p).read())
        ref.append((t.energy() + offset, t.forces()))
    open(os.path.join(d, "input.data"), "w").write("".join(parts))
    return d, (pos, elem, boxes, off), ref


def student_predictions(d, data):
    pos, elem, boxes, off = data
    o = Oracle(d)
    E, F = [], []
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        o.set_structure(pos[a:b], elem[a:b], boxes[s])
        o.predict(PULL, BRUTE)
        E.append(o.energy())  # Prediction::predict semantics: atomic offsets included
        F.append(o.forces())
    return np.array(E), np.vstack(F)


@pytest.mark.gpu
@pytest.mark.parametrize("batch_atoms", [100000, 200])
def test_dataset_water_set_plain(water_set_dir, batch_atoms, tmp_path):
    need(TOOL)
    d0, data, ref = water_set_dir
    d = tmp_path / "run"
    shutil.copytree(d0, d)
    say("case: nnpx-dataset, 12 water structures, plain, batch-atoms %d" % batch_atoms)
    rc, so, se = run([TOOL, "0", "--batch-atoms", str(batch_atoms)], cwd=str(d))
    assert rc == 0, se + open(d / "nnp-dataset.log.0000").read()
    En, Fn = student_predictions(str(d), data)
    e = read_table(d / "energy.comp", 4)
    f = read_table(d / "forces.comp", 4)
    pos, elem, boxes, off = data
    assert (e[:, 0] == np.arange(12)).all() and (e[:, 1] == np.diff(off)).all()
    Eref = np.array([r[0] for r in ref])
    Fref = np.vstack([r[1] for r in ref])
    dE = float(np.abs(e[:, 3] - En).max() / np.abs(En).max())
    dF = float(np.abs(f[:, 3] - Fn.reshape(-1)).max())
    say("  energies rel %.2e  forces abs %.2e (printed to 9 digits)" % (dE, dF))
    assert np.abs(e[:, 2] - Eref).max() <= 1e-8 * np.abs(Eref).max()
    assert dE <= 1e-8
    assert np.abs(f[:, 2] - Fref.reshape(-1)).max() <= 1e-8 + 1e-8 * np.abs(Fref).max()
    assert dF <= 1e-8 + 1e-8 * np.abs(Fn).max()
    # error metrics of Structure::updateError / Dataset::collectError
    n = np.diff(off)
    diff = Eref - En
    rmsepa = np.sqrt(np.mean((diff / n) ** 2))
    rmse = np.sqrt(np.mean(diff ** 2))
    maepa = np.mean(np.abs(diff) / n)
    mae = np.mean(np.abs(diff))
    frmse = np.sqrt(np.mean((Fref - Fn) ** 2))
    fmae = np.mean(np.abs(Fref - Fn))
    tab = error_table(d / "nnp-dataset.log.0000")
    say("  log ENERGY", tab["ENERGY"], "FORCES", tab["FORCES"])
    assert np.allclose(tab["ENERGY"], [rmsepa, rmse, maepa, mae], rtol=1e-4)
    assert np.allclose(tab["FORCES"], [frmse, fmae], rtol=1e-4)
    # output.data holds the predictions (offsets added), sensitivity files exist
    out = [line for line in open(d / "output.data") if line.startswith("energy")]
    assert len(out) == 12
    assert np.allclose([float(x.split()[1]) for x in out], En, rtol=1e-12, atol=0)
    for z in (1, 8):
        s = read_table(d / ("sensitivity.%03d.out" % z), 5)
        assert s.shape == (50, 5) and abs(s[:, 1].sum() - 100.0) < 1e-4


@pytest.mark.gpu
def test_dataset_water_set_mixed(water_set_dir, tmp_path):
    need(TOOL)
    d0, data, ref = water_set_dir
    d = tmp_path / "run"
    shutil.copytree(d0, d)
    say("case: nnpx-dataset, 12 water structures, mixed")
    rc, so, se = run([TOOL, "0", "--mixed"], cwd=str(d))
    assert rc == 0, se
    En, Fn = student_predictions(str(d), data)
    e = read_table(d / "energy.comp", 4)
    f = read_table(d / "forces.comp", 4)
    rms = float(np.sqrt(np.mean((f[:, 3] - Fn.reshape(-1)) ** 2)))
    de = float(np.abs((e[:, 3] - En) / np.diff(data[3])).max())
    say("  force rms vs FP64 %.2e  energy per atom max %.2e" % (rms, de))
    assert rms <= 1e-3
    assert de <= 1e-4


@pytest.mark.gpu
def test_dataset_golden_normalised(tmp_path):
    """H2O_RPBE-D3 golden: normalised units, energies only (no
    use_short_forces); Ennp_phys equals the golden energy."""
    need(TOOL)
    gdir = os.path.join(HERE, "golden", "H2O_RPBE-D3")
    d = tmp_path / "run"
    d.mkdir()
    for fn in os.listdir(gdir):
        if fn != "expected.json":
            shutil.copy(os.path.join(gdir, fn), d / fn)
    say("case: nnpx-dataset on the H2O_RPBE-D3 golden")
    rc, so, se = run([TOOL, "0"], cwd=str(d))
    assert rc == 0, se
    e = read_table(d / "energy.comp", 6)
    exp = json.load(open(os.path.join(gdir, "expected.json")))["energy"]
    say("  Ennp_phys %.10e  golden %.10e" % (e[0, 3], exp))
    assert abs(e[0, 3] - exp) <= 1e-8 * abs(exp)
    assert not os.path.exists(d / "forces.comp")


CHECKF = os.path.join(ROOT, "n2p2_amd", "bin", "nnpx-checkf")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["Cu2S_PBE", "DMABN_SCAN"])
def test_checkf_goldens(name, tmp_path):
    """nnpx-checkf (nnp-checkf semantics, delta 1e-4): every structure OK
    (max |F_analytic - F_numeric| <= 10 delta^2), analytic forces equal the
    golden forces; Cu2S_PBE periodic, DMABN_SCAN isolated in normalised
    units."""
    need(CHECKF)
    gdir = os.path.join(HERE, "golden", name)
    d = tmp_path / "run"
    d.mkdir()
    for fn in os.listdir(gdir):
        if fn != "expected.json":
            shutil.copy(os.path.join(gdir, fn), d / fn)
    say("case: nnpx-checkf on %s" % name)
    rc, so, se = run([CHECKF], cwd=str(d))
    assert rc == 0, se
    f = read_table(d / "checkf-forces.out", 5)
    s = read_table(d / "checkf-summary.out", 4)
    exp = np.array(json.load(open(os.path.join(gdir, "expected.json")))["forces"]).reshape(-1)
    err = float(np.abs(f[:, 3] - f[:, 4]).max())
    say("  %d forces, max |analytic - numeric| %.2e, max |analytic - golden| %.2e"
        % (len(f), err, float(np.abs(f[:, 3] - exp).max())))
    assert len(f) == exp.size and s.shape == (1, 4)
    assert np.abs(f[:, 3] - exp).max() <= 1e-8
    assert err <= 10 * 1e-4 ** 2 and abs(s[0, 3] - err) <= 1e-12
    assert "WARNING" not in open(d / "nnp-checkf.log.0000").read()


CHECKDW = os.path.join(ROOT, "n2p2_amd", "bin", "nnpx-checkdw")


def test_checkdw_usage_without_arguments():
    need(CHECKDW)
    rc, so, _ = run([CHECKDW])
    assert rc == 1 and "USAGE" in so


@pytest.mark.gpu
@pytest.mark.parametrize("batch_atoms", [1000000, 5])
def test_checkdw_lj_example(batch_atoms, tmp_path):
    """nnpx-checkdw on the reference's nnp-checkdw LJ example (101 Ne dimers,
    22 connections, use_short_forces): the output files of the reference's
    tool test (test_nnp-checkdw.cpp:17-21), every property of every structure
    OK (max |analytic - numeric| <= 100 delta^2), analytic rows equal to the
    oracle's Training::dPdc rows, numeric rows to the oracle's central
    differences; batch-atoms 5 splits the numeric batches and the seeded
    copies."""
    need(CHECKDW)
    gdir = os.path.join(HERE, "golden", "LJ_checkdw")
    d = tmp_path / "run"
    shutil.copytree(gdir, d)
    say("case: nnpx-checkdw on the LJ example, batch-atoms %d" % batch_atoms)
    rc, so, se = run([CHECKDW, "0", "--batch-atoms", str(batch_atoms)], cwd=str(d))
    assert rc == 0, se
    for fn in ("nnp-checkdw.log.0000", "checkdw-summary.out", "checkdw-weights.energy.out",
               "checkdw-weights.force.out"):
        assert os.path.exists(d / fn), fn
    we = read_table(d / "checkdw-weights.energy.out", 5)
    wf = read_table(d / "checkdw-weights.force.out", 5)
    summ = read_table(d / "checkdw-summary.out", 6)
    assert we.shape == (101 * 22, 5) and wf.shape == (101 * 6 * 22, 5) and summ.shape == (101, 6)
    log = open(d / "nnp-checkdw.log.0000").read()
    assert len(re.findall(r"^Configuration .* OK .* OK", log, re.M)) == 101 and "WARN" not in log
    errE, errF = np.abs(we[:, 3] - we[:, 4]).max(), np.abs(wf[:, 3] - wf[:, 4]).max()
    say("  max |analytic - numeric|: energy %.2e  force %.2e" % (errE, errF))
    assert max(errE, errF) <= 100 * 1e-4 ** 2
    from n2p2_amd import read_structures
    structs = read_structures(os.path.join(gdir, "input.data"), ["Ne"])
    o = Oracle(str(d))
    w0 = o.weights()
    for s in (0, 50, 100):
        pos, elem, box = structs[s]
        o.set_weights(w0)
        o.set_structure(pos, elem, box)
        o.predict(PULL, BRUTE)
        aE, aF = o.dPdc("energy"), o.dPdc("force")
        rE = we[we[:, 0] == s + 1]
        rF = wf[wf[:, 0] == s + 1]
        assert (rE[:, 1] == 1).all() and (rE[:, 2] == np.arange(1, 23)).all()
        scale = max(np.abs(aF).max(), np.abs(aE).max())
        assert np.abs(rE[:, 3] - aE).max() <= 1e-10 * scale
        assert np.abs(rF[:, 3] - aF.reshape(-1)).max() <= 1e-10 * scale
        # numeric: w + delta, then (w + delta) - 2 delta, as Training::dPdcN
        k = 7
        res = []
        for step in (1e-4, -2e-4):
            w = w0.copy() if step > 0 else w
            w[k] += step
            o.set_weights(w)
            o.set_structure(pos, elem, box)
            o.predict(PULL, BRUTE)
            res.append((o.energy(), o.forces().reshape(-1)))
        nE = (res[0][0] - res[1][0]) / 2e-4
        nF = (res[0][1] - res[1][1]) / 2e-4
        assert abs(rE[k, 4] - nE) <= 1e-9 * max(abs(nE), 1.0)
        assert np.abs(rF[k::22, 4] - nF).max() <= 1e-9 * max(np.abs(nF).max(), 1.0)
    o.set_weights(w0)


@pytest.mark.gpu
@pytest.mark.parametrize("name,symbols", [("LJ_checkdw", ["Ne"]), ("H2O_RPBE-D3_checkf", ["H", "O"])])
def test_checkf_reference_examples(name, symbols, tmp_path):
    """nnpx-checkf on the examples the reference's nnp-checkf tool test runs
    (test_nnp-checkf.cpp:17-21, Example_nnp_checkf.h:18-24: the LJ dimers --
    the same files as the nnp-checkdw LJ example -- and three normalised
    48-atom H2O_RPBE-D3 cells).  The reference test only asks for the output
    files; here in addition every structure's verdict follows the rule
    (WARNING iff max |F_analytic - F_numeric| > 10 delta^2; the steepest LJ
    dimers miss it, as they would in the reference), the analytic column
    equals the oracle's forces and the numeric column the oracle's central
    differences (a few structures, a few coordinates each)."""
    need(CHECKF)
    from n2p2_amd import read_structures
    gdir = os.path.join(HERE, "golden", name)
    d = tmp_path / "run"
    shutil.copytree(gdir, d)
    say("case: nnpx-checkf on the %s example" % name)
    rc, so, se = run([CHECKF], cwd=str(d))
    assert rc == 0, se
    for fn in ("nnp-checkf.log.0000", "checkf-forces.out", "checkf-summary.out"):
        assert os.path.exists(d / fn), fn
    structs = read_structures(os.path.join(gdir, "input.data"), symbols)
    f = read_table(d / "checkf-forces.out", 5)
    summ = read_table(d / "checkf-summary.out", 4)
    assert len(f) == sum(3 * len(s[1]) for s in structs) and len(summ) == len(structs)
    verdicts = re.findall(r"^Configuration\s+(\d+):.*?(OK\.|WARNING!)", open(d / "nnp-checkf.log.0000").read(), re.M)
    assert len(verdicts) == len(structs)
    for (idx, v), row in zip(verdicts, summ):
        assert int(idx) == int(row[0]) and (v == "WARNING!") == (row[3] > 10 * 1e-4 ** 2)
    o = Oracle(str(d))
    wa = wn = 0.0
    for s in range(0, len(structs), max(1, len(structs) // 4)):
        pos, elem, box = structs[s]
        o.set_structure(pos, elem, box)
        o.predict(PULL, BRUTE)
        rows = f[f[:, 0] == s + 1]
        wa = max(wa, float(np.abs(rows[:, 3] - o.forces().reshape(-1)).max()))
        for q in (0, len(rows) // 2, len(rows) - 1):
            a, c = q // 3, q % 3
            E = []
            for sign in (-1.0, 1.0):
                p = pos.copy()
                p[a, c] += sign * 1e-4
                o.set_structure(p, elem, box)
                o.predict(PULL, BRUTE)
                E.append(o.energy())
            fn = -(E[1] - E[0]) / 2e-4
            wn = max(wn, abs(rows[q, 4] - fn) / max(abs(fn), 1.0))
    ok = sum(v == "OK." for _, v in verdicts)
    say("  %d structures (%d OK), max |analytic - oracle| %.2e, numeric vs oracle rel %.2e" % (len(structs), ok, wa, wn))
    assert wa <= 1e-8
    assert wn <= 1e-6
    if name == "H2O_RPBE-D3_checkf":
        assert ok == len(structs)
