"""Every cutoff type of CutoffFunction (CutoffFunction.h:36-58, with the
inner cutoff alpha) and every activation of NeuralNetwork
(NeuralNetwork.cpp:1078-1097) through both GPU forms, water-192, against the
FP64 oracle.  Bars: plain FP64 energy 1e-10 relative and forces 1e-8; mixed
force RMS 1e-3."""
import numpy as np
import pytest

import n2p2_amd
from n2p2_amd import remap, synth
from oracle_lib import BRUTE, PULL, Oracle

pytestmark = pytest.mark.gpu

# (cutoff_type line, hidden activations): each cutoff type once, each
# activation letter at least once.
CASES = [
    ("0 0.0", ("t", "t")),   # CT_HARD
    ("1 0.3", ("s", "t")),   # CT_COS, inner cutoff
    ("2 0.0", ("p", "p")),   # CT_TANHU
    ("3 0.2", ("g", "t")),   # CT_TANH
    ("4 0.1", ("c", "t")),   # CT_EXP
    ("5 0.0", ("S", "t")),   # CT_POLY1
    ("6 0.25", ("e", "t")),  # CT_POLY2
    ("7 0.0", ("h", "t")),   # CT_POLY3
    ("8 0.4", ("r", "t")),   # CT_POLY4
]


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("cutoff,acts", CASES, ids=["ct" + c.split()[0] for c, _ in CASES])
def test_cutoff_and_activation(cutoff, acts, tmp_path):
    say("case: water-192 cutoff_type %s, activations %s" % (cutoff, "".join(acts)))
    d = synth.water_model(str(tmp_path / "m"), cutoff_type=cutoff, activations=acts + ("l",))
    pos, elem, box = synth.water_box(4)
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    o.predict(PULL, BRUTE)
    Eo, Fo = o.energy(), o.forces()
    m = n2p2_amd.Mode(d)
    m.set_structures(remap(pos, box), elem, box[None])
    m.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
    E, F = m.energies()[0], m.forces()
    erel = abs(E - Eo) / abs(Eo)
    dF = float(np.abs(F - Fo).max())
    m.evaluate(n2p2_amd.MIXED, forces=True)
    Fm = m.forces()
    rms = float(np.sqrt(np.mean((Fm - Fo) ** 2)))
    say("  plain E rel %.2e max|dF| %.2e; mixed F rms %.2e (|F| rms %.3f)"
        % (erel, dF, rms, float(np.sqrt(np.mean(Fo ** 2)))))
    assert erel <= 1e-10
    assert dF <= 1e-8
    assert rms <= 1e-3
