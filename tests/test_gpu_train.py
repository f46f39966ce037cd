"""Energy+force training gradient on the device (config "force-training
gradient step"): FP64 and mixed (FP32) forms against the FP64 oracle's
restatement of Training::dPdc / NeuralNetwork::calculateDFdc, which is itself
checked against central finite differences on the CPU."""
import numpy as np
import pytest

import n2p2_amd
from n2p2_amd import remap, synth
from oracle_lib import BRUTE, PULL, Oracle

pytestmark = pytest.mark.gpu
BETA = 0.7


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def models(tmp_path_factory):
    d = tmp_path_factory.mktemp("train")
    return synth.water_model(str(d / "student")), synth.water_model(str(d / "teacher"), seed=synth.SEED_TEACHER)


def targets(teacher, pos, elem, boxes, off):
    E, F = [], []
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        o = Oracle(teacher)
        o.set_structure(pos[a:b], elem[a:b], boxes[s])
        o.predict(PULL, BRUTE)
        E.append(o.energy())
        F.append(o.forces())
    return np.array(E), np.vstack(F)


def oracle_gradient(student, pos, elem, boxes, off, E_ref, F_ref, energy_term=True):
    N = off[-1]
    loss, grad = np.zeros(3), None
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        o = Oracle(student)
        o.set_structure(pos[a:b], elem[a:b], boxes[s])
        o.predict(PULL, BRUTE)
        l, g = o.loss_gradient(E_ref[s], F_ref[a:b], 1.0 / (b - a) if energy_term else 0.0, BETA / (3.0 * N))
        loss += l
        grad = g if grad is None else grad + g
    return loss, grad


@pytest.mark.parametrize("precision", [n2p2_amd.PLAIN_FP64, n2p2_amd.MIXED])
def test_loss_gradient_vs_oracle(models, precision):
    say("case: loss gradient, 6 water structures (66-90 atoms), precision", precision)
    student, teacher = models
    pos, elem, boxes, off = synth.water_set(n_struct=6, mmin=22, mmax=30)
    E_ref, F_ref = targets(teacher, pos, elem, boxes, off)
    lo, go = oracle_gradient(student, pos, elem, boxes, off, E_ref, F_ref)
    m = n2p2_amd.Mode(student)
    m.set_structures(pos, elem, boxes, off)
    loss, g = m.loss_gradient(E_ref, F_ref, BETA, precision)
    rel = float(np.abs(g - go).max() / np.abs(go).max())
    lrel = np.abs(loss - lo) / np.abs(lo)
    say("  loss %s (oracle %s)  loss rel %s  grad max rel %.2e" % (loss, lo, lrel, rel))
    if precision == n2p2_amd.PLAIN_FP64:
        assert lrel.max() <= 1e-10
        assert rel <= 1e-9
    else:
        assert lrel.max() <= 1e-4
        assert rel <= 2e-3


def test_force_term_gradient_vs_oracle(models):
    """Force term alone: each form gets its own energies as energy targets
    (zero energy residual) and the oracle drops the energy term, so the
    second-order path carries the whole gradient."""
    say("case: force-term gradient, FP64 and mixed")
    student, teacher = models
    pos, elem, boxes, off = synth.water_set(n_struct=4, mmin=22, mmax=26, seed=synth.SEED_WATERSET + 1)
    _, F_ref = targets(teacher, pos, elem, boxes, off)
    lo, go = oracle_gradient(student, pos, elem, boxes, off, np.zeros(len(off) - 1), F_ref, energy_term=False)
    m = n2p2_amd.Mode(student)
    m.set_structures(pos, elem, boxes, off)
    for precision, tol in ((n2p2_amd.PLAIN_FP64, 1e-9), (n2p2_amd.MIXED, 2e-4)):
        m.evaluate(precision, forces=True)
        E_own = m.energies()
        loss, g = m.loss_gradient(E_own, F_ref, BETA, precision)
        rel = float(np.abs(g - go).max() / np.abs(go).max())
        say("  precision %d: energy loss %.1e  force loss %.6e (oracle %.6e)  grad max rel %.2e"
            % (precision, loss[1], loss[2], lo[2], rel))
        assert loss[1] <= 1e-20
        assert rel <= tol


def test_gradient_step_reduces_loss(models):
    """A small step against the FP32 gradient lowers the FP64 loss."""
    say("case: gradient step")
    student, teacher = models
    pos, elem, boxes, off = synth.water_set(n_struct=8, mmin=22, mmax=40, seed=synth.SEED_WATERSET + 2)
    E_ref, F_ref = targets(teacher, pos, elem, boxes, off)
    m = n2p2_amd.Mode(student)
    m.set_structures(pos, elem, boxes, off)
    l0, g = m.loss_gradient(E_ref, F_ref, BETA, n2p2_amd.MIXED)
    w = m.weights()
    m.set_weights(w - 1e-3 * g / np.abs(g).max())
    l1, _ = m.loss_gradient(E_ref, F_ref, BETA, n2p2_amd.PLAIN_FP64)
    say("  loss %.6e -> %.6e" % (l0[0], l1[0]))
    assert l1[0] < l0[0]
    m.set_weights(w)
    l2, _ = m.loss_gradient(E_ref, F_ref, BETA, n2p2_amd.PLAIN_FP64)
    assert abs(l2[0] - l0[0]) <= 1e-3 * l0[0]


@pytest.mark.parametrize("precision", [n2p2_amd.PLAIN_FP64, n2p2_amd.MIXED])
def test_weight_vjp_vs_oracle_dPdc(models, precision):
    """nnpx_weight_vjp: the per-structure rows vE_s dE_s/dw + sum vF . dF/dw
    of three structures in one call -- a 24-atom water box whose 6.21 A cell
    is smaller than the cutoff (self images) and two 66-78 atom boxes --
    against the oracle's Training::dPdc rows contracted with the same random
    seeds."""
    say("case: weight VJP, 3 structures (24 atoms self-image + 2 water boxes), precision", precision)
    student, _ = models
    p0, e0, b0 = synth.water_box(2)
    p1, e1, b1, o1 = synth.water_set(n_struct=2, mmin=22, mmax=26, seed=synth.SEED_WATERSET + 2)
    pos = np.vstack([p0, p1])
    elem = np.concatenate([e0, e1]).astype(np.int32)
    boxes = [b0] + list(b1)
    off = np.concatenate([[0], len(e0) + np.asarray(o1)])
    rng = np.random.default_rng(21)
    vE = rng.normal(size=3)
    vF = rng.normal(size=3 * len(elem))
    m = n2p2_amd.Mode(student)
    m.set_structures(pos, elem, boxes, off)
    g = m.weight_vjp(vE, vF, precision)
    assert g.shape == (3, m.num_connections())
    worst = 0.0
    for s in range(3):
        a, b = off[s], off[s + 1]
        o = Oracle(student)
        o.set_structure(pos[a:b], elem[a:b], boxes[s])
        o.predict(PULL, BRUTE)
        ref = vE[s] * o.dPdc("energy") + vF[3 * a:3 * b] @ o.dPdc("force")
        rel = float(np.abs(g[s] - ref).max() / np.abs(ref).max())
        say("  structure %d (%d atoms): max |row| %.3e  max rel %.2e" % (s, b - a, np.abs(ref).max(), rel))
        worst = max(worst, rel)
    assert worst <= (1e-9 if precision == n2p2_amd.PLAIN_FP64 else 2e-3)


@pytest.mark.parametrize("kind", ["compact", "weighted"])
@pytest.mark.parametrize("precision", [n2p2_amd.PLAIN_FP64, n2p2_amd.MIXED])
def test_weight_vjp_all_sf_types(kind, precision, tmp_path):
    """Weight derivatives through every SF type: the compact model (types
    20/21/22, the compact kernel's tangent mode in the mixed form) and the
    weighted model (types 12/13 with 2/3 and 23/24/25), per-structure rows
    vE dE/dw + vF . dF/dw of two structures against the oracle's
    Training::dPdc rows."""
    from oracle_lib import compact_model, weighted_model
    d = (compact_model if kind == "compact" else weighted_model)(str(tmp_path / kind))
    say("case: weight VJP, %s model, precision %d" % (kind, precision))
    p0, e0, b0 = synth.water_box(3)
    p1, e1, b1 = synth.water_box(4, seed=synth.SEED_WATER + 3)
    pos = np.vstack([remap(p0, b0), remap(p1, b1)])
    elem = np.concatenate([e0, e1]).astype(np.int32)
    off = np.array([0, len(e0), len(e0) + len(e1)])
    boxes = [b0, b1]
    rng = np.random.default_rng(31)
    vE = rng.normal(size=2)
    vF = rng.normal(size=3 * len(elem))
    m = n2p2_amd.Mode(d)
    m.set_structures(pos, elem, boxes, off)
    g = m.weight_vjp(vE, vF, precision)
    worst = 0.0
    for s in range(2):
        a, b = off[s], off[s + 1]
        o = Oracle(d)
        o.set_structure(pos[a:b], elem[a:b], boxes[s])
        o.predict(PULL, BRUTE)
        ref = vE[s] * o.dPdc("energy") + vF[3 * a:3 * b] @ o.dPdc("force")
        rel = float(np.abs(g[s] - ref).max() / np.abs(ref).max())
        say("  structure %d (%d atoms): max rel %.2e" % (s, b - a, rel))
        worst = max(worst, rel)
    assert worst <= (1e-9 if precision == n2p2_amd.PLAIN_FP64 else 2e-3)



@pytest.fixture(scope="module")
def wide_case(tmp_path_factory):
    """A 50-100-100-1 network on two 24-atom water boxes (6.21 A cells, self
    images) and the oracle's Training::dPdc rows of both."""
    d = synth.water_model(str(tmp_path_factory.mktemp("wide")), hidden=(100, 100))
    p0, e0, b0 = synth.water_box(2)
    p1, e1, b1 = synth.water_box(2, seed=synth.SEED_WATER + 7)
    pos = np.vstack([remap(p0, b0), remap(p1, b1)])
    elem = np.concatenate([e0, e1]).astype(np.int32)
    off = np.array([0, len(e0), len(e0) + len(e1)])
    boxes = [b0, b1]
    refs = []
    for s in range(2):
        a, b = off[s], off[s + 1]
        o = Oracle(d)
        o.set_structure(pos[a:b], elem[a:b], boxes[s])
        o.predict(PULL, BRUTE)
        refs.append((o.dPdc("energy"), o.dPdc("force")))
    return d, pos, elem, boxes, off, refs


@pytest.mark.parametrize("precision", [n2p2_amd.PLAIN_FP64, n2p2_amd.MIXED])
def test_weight_vjp_wide_network(wide_case, precision):
    """15301 connections per element, more than the 64 KB of LDS holds as an
    FP64 gradient: the gradient kernel's global-atomic variant, per-structure
    rows vE dE/dw + vF . dF/dw against the oracle's Training::dPdc rows."""
    d, pos, elem, boxes, off, refs = wide_case
    say("case: weight VJP, 50-100-100-1 network, 2 x 24 atoms, precision %d" % precision)
    rng = np.random.default_rng(41)
    vE = rng.normal(size=2)
    vF = rng.normal(size=3 * len(elem))
    m = n2p2_amd.Mode(d)
    m.set_structures(pos, elem, boxes, off)
    assert m.num_connections() == 2 * 15301
    g = m.weight_vjp(vE, vF, precision)
    worst = 0.0
    for s in range(2):
        a, b = off[s], off[s + 1]
        ref = vE[s] * refs[s][0] + vF[3 * a:3 * b] @ refs[s][1]
        rel = float(np.abs(g[s] - ref).max() / np.abs(ref).max())
        say("  structure %d (%d atoms): max rel %.2e" % (s, b - a, rel))
        worst = max(worst, rel)
    assert worst <= (1e-9 if precision == n2p2_amd.PLAIN_FP64 else 2e-3)


@pytest.mark.parametrize("precision", [n2p2_amd.PLAIN_FP64, n2p2_amd.MIXED])
def test_loss_gradient_wide_network(precision, tmp_path):
    """nnpx_loss_gradient (one summed gradient) with the 50-100-100-1 network
    whose FP64 connection gradient exceeds the LDS accumulator: every atom's
    connections go into one row through the global-atomic variant; three
    66-78 atom water structures against the oracle's loss gradient."""
    d = synth.water_model(str(tmp_path / "wide"), hidden=(100, 100))
    t = synth.water_model(str(tmp_path / "wide_teacher"), hidden=(100, 100), seed=synth.SEED_TEACHER)
    say("case: loss gradient, 50-100-100-1 network, 3 water structures, precision %d" % precision)
    pos, elem, boxes, off = synth.water_set(n_struct=3, mmin=22, mmax=26, seed=synth.SEED_WATERSET + 5)
    E_ref, F_ref = targets(t, pos, elem, boxes, off)
    lo, go = oracle_gradient(d, pos, elem, boxes, off, E_ref, F_ref)
    m = n2p2_amd.Mode(d)
    m.set_structures(pos, elem, boxes, off)
    loss, g = m.loss_gradient(E_ref, F_ref, BETA, precision)
    rel = float(np.abs(g - go).max() / np.abs(go).max())
    lrel = np.abs(loss - lo) / np.abs(lo)
    say("  loss rel %s  grad max rel %.2e" % (lrel, rel))
    if precision == n2p2_amd.PLAIN_FP64:
        assert lrel.max() <= 1e-10
        assert rel <= 1e-9
    else:
        assert lrel.max() <= 1e-4
        assert rel <= 2e-3
