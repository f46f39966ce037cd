"""Velocity-Verlet MD on the device: trajectory parity with a host
integrator driven by the FP64 oracle's forces, HIP-graph replays against
eager steps, energy conservation and the skin-triggered list rebuild."""
import numpy as np
import pytest

import n2p2_amd
from n2p2_amd import remap, synth
from oracle_lib import BRUTE, PULL, Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def fcc_dir(tmp_path_factory):
    return synth.fcc_model(str(tmp_path_factory.mktemp("fcc")))


def masses_fcc():
    return np.array([synth.MASS[s] for s in synth.FCC_SYMBOLS])


def min_image(d, box):
    L = np.diag(box)
    return d - L * np.round(d / L)


def setup_md(d, n_cells, precision, skin, temperature=300.0, dt=1.0, seed=synth.SEED_VELOCITY, pair_lists=-1):
    pos, elem, box = synth.fcc_lattice(n_cells=n_cells)
    m = n2p2_amd.Mode(d)
    m.set_structures(remap(pos, box), elem, box[None])
    m.set_option(n2p2_amd.OPT_PAIR_LISTS, pair_lists)
    mass = masses_fcc()
    vel = synth.maxwell_velocities(mass[elem], temperature, seed)
    m.md_init(mass, vel, dt, skin, precision)
    return m, pos, elem, box, vel, mass


def test_md_plain_matches_host_verlet(fcc_dir):
    """5 steps of plain-FP64 device MD equal a host velocity-Verlet loop
    whose forces come from the oracle (108 atoms, 3^3 cells)."""
    say("case: MD plain FP64 vs host Verlet + oracle forces, FCC 108")
    m, pos, elem, box, vel, mass = setup_md(fcc_dir, 3, n2p2_amd.PLAIN_FP64, skin=0.5)
    dt = 1.0
    hk = 0.5 * dt * synth.ACCEL / mass[elem]
    o = Oracle(fcc_dir)
    x, v = pos.copy(), vel.copy()
    o.set_structure(x, elem, box)
    o.predict(PULL, BRUTE)
    F = o.forces()
    for _ in range(5):
        v = v + hk[:, None] * F
        x = x + dt * v
        o.set_structure(x, elem, box)
        o.predict(PULL, BRUTE)
        F = o.forces()
        v = v + hk[:, None] * F
    m.md_run(5)
    xg, vg = m.md_state()
    dx = float(np.abs(min_image(xg - x, box)).max())
    dv = float(np.abs(vg - v).max())
    dF = float(np.abs(m.forces() - F).max())
    say("  max|dx| %.2e  max|dv| %.2e  max|dF| %.2e" % (dx, dv, dF))
    assert dx <= 1e-10 and dv <= 1e-12 and dF <= 1e-8


@pytest.mark.parametrize("precision,pair_lists", [(n2p2_amd.MIXED, 0), (n2p2_amd.MIXED, 1), (n2p2_amd.PLAIN_FP64, -1)])
def test_md_graph_replays_equal_eager(fcc_dir, precision, pair_lists):
    """Graph-captured steps (5 per graph, four replays with host reads and
    eager work between them) against the same steps run eagerly; the mixed
    form with and without the angular pair lists (sized at nnpx_md_init,
    read by the captured force kernels)."""
    say("case: MD graph vs eager, FCC 500, precision", precision, "pair lists", pair_lists)
    a = setup_md(fcc_dir, 5, precision, skin=0.4, temperature=600.0, pair_lists=pair_lists)[0]
    b = setup_md(fcc_dir, 5, precision, skin=0.4, temperature=600.0, pair_lists=pair_lists)[0]
    for rep in range(4):
        a.md_run(5)
        b.md_run(5, graph_steps=5)
        xa, va = a.md_state()
        xb, vb = b.md_state()
        ea, _, ra = a.md_energies()
        eb, _, rb = b.md_energies()
        dx = float(np.abs(xa - xb).max())
        say("  replay %d: max|dx| %.2e  max|dv| %.2e  |dE| %.2e  rebuilds %d/%d"
            % (rep, dx, float(np.abs(va - vb).max()), float(np.abs(ea - eb).max()), ra, rb))
        assert dx <= 1e-9
        assert float(np.abs(ea - eb).max()) <= 1e-8 * max(1.0, float(np.abs(ea).max()))
        assert ra == rb
        b.energies()  # eager host work on the graph context between replays


def test_md_graph_dropped_on_option_change(fcc_dir):
    """Changing the BF16 term count between graph-captured runs re-captures
    the step graph: graph runs before and after the change equal the same
    eager runs (the one-term forces differ from the three-term ones by far
    more than the tolerance, so a stale graph would show)."""
    say("case: MD graph, BF16 terms 3 -> 1 between runs, FCC 500")
    a = setup_md(fcc_dir, 5, n2p2_amd.MIXED, skin=0.4)[0]
    b = setup_md(fcc_dir, 5, n2p2_amd.MIXED, skin=0.4)[0]
    c = setup_md(fcc_dir, 5, n2p2_amd.MIXED, skin=0.4)[0]
    for m in (a, b, c):
        m.set_option(n2p2_amd.OPT_BF16_TERMS, 3)
    a.md_run(5, graph_steps=5)
    b.md_run(5)
    c.md_run(5)
    a.set_option(n2p2_amd.OPT_BF16_TERMS, 1)
    b.set_option(n2p2_amd.OPT_BF16_TERMS, 1)
    a.md_run(5, graph_steps=5)
    b.md_run(5)
    c.md_run(5)
    xa, _ = a.md_state()
    xb, _ = b.md_state()
    xc, _ = c.md_state()
    dab = float(np.abs(xa - xb).max())
    dbc = float(np.abs(xb - xc).max())
    say("  graph vs eager max|dx| %.2e; one vs three terms max|dx| %.2e" % (dab, dbc))
    assert dab <= 1e-9
    assert dbc > 100 * max(dab, 1e-12)


@pytest.mark.parametrize("precision", [n2p2_amd.MIXED, n2p2_amd.PLAIN_FP64])
@pytest.mark.parametrize("staging", [-1, 32])
def test_md_fitted_staging(fcc_dir, staging, precision):
    """MD with the in-cutoff staging fitted at nnpx_md_init (-1) and with a
    capacity below the 78 in-cutoff neighbours of the lattice (32): the
    too-small run overflows, is repeated from its start with the whole list
    staged, and both end where the run with the whole list staged (0) ends;
    later runs of the repeated context keep the whole list."""
    say("case: MD staging %d vs whole list, FCC 500, 2 x 10 steps in graphs of 5, precision %d"
        % (staging, precision))

    def run(opt):
        pos, elem, box = synth.fcc_lattice(n_cells=5)
        m = n2p2_amd.Mode(fcc_dir)
        m.set_option(n2p2_amd.OPT_MD_STAGING, opt)
        m.set_structures(remap(pos, box), elem, box[None])
        mass = masses_fcc()
        m.md_init(mass, synth.maxwell_velocities(mass[elem], 600.0, synth.SEED_VELOCITY), 1.0, 1.0, precision)
        out = []
        for _ in range(2):
            m.md_run(10, graph_steps=5)
            x, v = m.md_state()
            ep, ek, nr = m.md_energies()
            out.append((x, v, ep.copy(), ek.copy(), nr))
        return out
    ref, got = run(0), run(staging)
    for k, (a, b) in enumerate(zip(ref, got)):
        dx = float(np.abs(a[0] - b[0]).max())
        dv = float(np.abs(a[1] - b[1]).max())
        de = float(np.abs(a[2] - b[2]).max())
        say("  run %d: max|dx| %.2e  max|dv| %.2e  max|dEpot| %.2e  steps %d/%d" % (k, dx, dv, de, len(a[2]), len(b[2])))
        assert len(a[2]) == len(b[2]) == 10 * (k + 1)
        assert dx <= 1e-9 and dv <= 1e-9 and de <= 1e-8 * max(1.0, float(np.abs(a[2]).max()))


def test_md_energy_conservation_and_rebuilds(fcc_dir):
    """Mixed form, 500 atoms, 200 steps at 600 K with a 0.3 A skin: the list
    is rebuilt on the device, total energy is conserved, and the final forces
    equal a fresh evaluation at the final positions."""
    say("case: MD mixed energy conservation, FCC 500")
    m, pos, elem, box, vel, mass = setup_md(fcc_dir, 5, n2p2_amd.MIXED, skin=0.3, temperature=600.0)
    m.md_run(200, graph_steps=20)
    ep, ek, rebuilds = m.md_energies()
    et = ep + ek
    drift = float(np.abs(et - et[0]).max())
    say("  steps %d  rebuilds %d  Ekin %.4f..%.4f  Etot drift %.3e eV (%.2e eV/atom)"
        % (len(ep), rebuilds, ek.min(), ek.max(), drift, drift / len(elem)))
    assert len(ep) == 200 and rebuilds >= 1
    assert drift <= 1e-3 * len(elem) * synth.KB_EV * 600.0 * 10
    F = m.forces()
    x, _ = m.md_state()
    f = n2p2_amd.Mode(fcc_dir)
    f.set_structures(remap(x, box), elem, box[None])
    f.evaluate(n2p2_amd.MIXED, forces=True)
    dF = float(np.abs(F - f.forces()).max())
    say("  final forces vs fresh evaluation: max|dF| %.2e" % dF)
    assert dF <= 1e-6
