"""Kalman-filter training updates on the device (Training::update with
updater_type 1, Training.cpp:2256-2697; KalmanFilter::update,
KalmanFilter.cpp:119-185) against the oracle: the oracle's Training::dPdc
rows and errors of the same candidates, fed to the oracle's KalmanFilter
restatement (itself pinned by the information form of the update,
tests/test_oracle.py), give the same new weights and covariance."""
import numpy as np
import pytest

import n2p2_amd
from n2p2_amd import synth
from oracle_lib import BRUTE, PULL, Oracle, kalman_update

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def setup(tmp_path_factory):
    d = tmp_path_factory.mktemp("kf")
    student = synth.water_model(str(d / "student"))
    teacher = synth.water_model(str(d / "teacher"), seed=synth.SEED_TEACHER)
    pos, elem, boxes, off = synth.water_set(n_struct=4, mmin=22, mmax=26, seed=synth.SEED_WATERSET + 11)
    E_ref, F_ref = [], []
    t = Oracle(teacher)
    for s in range(4):
        a, b = off[s], off[s + 1]
        t.set_structure(pos[a:b], elem[a:b], boxes[s])
        t.predict(PULL, BRUTE)
        E_ref.append(t.energy())
        F_ref.append(t.forces())
    return student, pos, elem, boxes, off, np.array(E_ref), np.vstack(F_ref)


def oracle_rows(model, pos, elem, boxes, off, s):
    o = Oracle(model)
    a, b = off[s], off[s + 1]
    o.set_structure(pos[a:b], elem[a:b], boxes[s])
    o.predict(PULL, BRUTE)
    return o


def conn_slices(m):
    out, p = [], 0
    for e in range(m.num_elements):
        n = m.num_connections(e)
        out.append((p, n))
        p += n
    return out


def test_kalman_energy_update_combined(setup):
    """Energy update, two candidates summed into one observation (jacobian
    mode 0), one standard filter over all connections."""
    model, pos, elem, boxes, off, E_ref, F_ref = setup
    say("case: Kalman energy update, 2 candidates summed, combined standard filter")
    m = n2p2_amd.Mode(model)
    m.set_structures(pos, elem, boxes, off)
    eps, eta = 0.02, 0.03
    m.kalman_init(0, False, epsilon=eps, eta=eta, etamax=eta)
    w0 = m.weights()
    cand = [0, 2]
    err = m.kalman_update("energy", cand, E_ref=E_ref)
    w1 = m.weights()
    P1, st = m.kalman_state(0)
    H = np.zeros(len(w0))
    xi = 0.0
    for s in cand:
        o = oracle_rows(model, pos, elem, boxes, off, s)
        H += o.dPdc("energy")
        xi += E_ref[s] - o.energy()
    Pref, wref, _ = kalman_update(np.eye(len(w0)) / eps, w0, H[:, None], [xi],
                                  [0, eta, 0.0, eta, 0, 0, 1, 0.0, 0.0, 0.0])
    dw = float(np.abs(w1 - wref).max() / np.abs(wref - w0).max())
    dP = float(np.abs(P1 - Pref).max() / np.abs(Pref).max())
    say("  errors %s  |dw| %.3e  rel err of the step %.2e  P rel err %.2e" % (err, np.abs(w1 - w0).max(), dw, dP))
    assert np.abs(w1 - w0).max() > 0
    assert dw <= 1e-9 and dP <= 1e-12
    assert st["updates"] == 1


@pytest.mark.parametrize("kalman_type", [0, 1])
def test_kalman_force_update_per_element(setup, kalman_type):
    """Force update, three candidates kept as three observations (jacobian
    mode 2), force weight 0.5, one filter per element (errors scaled by the
    atom's neighbours of that element over all its neighbours)."""
    model, pos, elem, boxes, off, E_ref, F_ref = setup
    say("case: Kalman force update, 3 candidates, per-element filters, type %d" % kalman_type)
    m = n2p2_amd.Mode(model)
    m.set_structures(pos, elem, boxes, off)
    eps, eta, lam, nu, fw = 0.05, 0.02, 0.97, 0.99, 0.5
    m.kalman_init(kalman_type, True, epsilon=eps, eta=eta, etamax=eta, lam=lam, nu=nu, q0=1e-4, qtau=0.1, qmin=1e-6)
    w0 = m.weights()
    cand_s = [1, 3, 1]
    cand_a = [off[1] + 5, off[3] + 17, off[1] + 30]
    cand_c = [0, 2, 1]
    err = m.kalman_update("force", cand_s, F_ref=F_ref, atoms=cand_a, components=cand_c, force_weight=fw, full=True)
    w1 = m.weights()
    rows, xs, fr = [], [], []
    for s, a, c in zip(cand_s, cand_a, cand_c):
        o = oracle_rows(model, pos, elem, boxes, off, s)
        la = a - off[s]
        rows.append(o.dPdc("force")[3 * la + c])
        xs.append(F_ref[a, c] - o.forces()[la, c])
        counts, js, _ = o.neighbors()
        p0 = int(counts[:la].sum())
        nb = elem[off[s] + js[p0:p0 + counts[la]]]
        fr.append([np.count_nonzero(nb == e) / counts[la] for e in range(2)])
    say("  errors %s (oracle %s)" % (err, xs))
    assert np.allclose(err, xs, rtol=1e-9, atol=1e-12)
    par0 = [kalman_type, eta, 0.0, eta, lam, nu, 1.0, 1e-4, 0.1, 1e-6]
    worst_w = worst_p = 0.0
    for u, (p, n) in enumerate(conn_slices(m)):
        H = np.array([r[p:p + n] for r in rows]).T * fw
        xi = np.array([x * fw * f[u] for x, f in zip(xs, fr)])
        Pref, wref, _ = kalman_update(np.eye(n) / eps, w0[p:p + n], H, xi, par0)
        P1, st = m.kalman_state(u)
        worst_w = max(worst_w, float(np.abs(w1[p:p + n] - wref).max() / np.abs(wref - w0[p:p + n]).max()))
        worst_p = max(worst_p, float(np.abs(P1 - Pref).max() / np.abs(Pref).max()))
        assert st["updates"] == 1 and st["n"] == n and st["offset"] == p
    say("  rel err of the weight step %.2e  P rel err %.2e" % (worst_w, worst_p))
    assert worst_w <= 1e-9 and worst_p <= 1e-12


@pytest.mark.parametrize("precision", [n2p2_amd.PLAIN_FP64, n2p2_amd.MIXED])
def test_kalman_training_lowers_the_loss(setup, precision):
    """Twenty updates alternating energy and force candidates lower the
    energy and force errors of the four structures."""
    model, pos, elem, boxes, off, E_ref, F_ref = setup
    say("case: 20 Kalman updates, precision %d" % precision)
    m = n2p2_amd.Mode(model)
    m.set_structures(pos, elem, boxes, off)

    def rmse():
        m.evaluate(precision, forces=True)
        n = np.diff(off)
        return (float(np.sqrt(np.mean(((m.energies() - E_ref) / n) ** 2))),
                float(np.sqrt(np.mean((m.forces() - F_ref) ** 2))))

    e0, f0 = rmse()
    m.kalman_init(0, False, epsilon=0.01, eta=0.01, etamax=1.0, etatau=0.05, q0=0.01, qtau=0.05, qmin=1e-6)
    rng = np.random.default_rng(1)
    for k in range(20):
        s = int(rng.integers(4))
        if k % 2 == 0:
            m.kalman_update("energy", [s], E_ref=E_ref, precision=precision)
        else:
            a = int(off[s] + rng.integers(off[s + 1] - off[s]))
            m.kalman_update("force", [s], F_ref=F_ref, atoms=[a], components=[int(rng.integers(3))],
                            force_weight=1.0, precision=precision)
    e1, f1 = rmse()
    say("  energy RMSE/atom %.4e -> %.4e, force RMSE %.4e -> %.4e" % (e0, e1, f0, f1))
    assert e1 < e0 and f1 < f0
