"""Checks of the SF driver on the GPU against the oracle (itself equal to the
reference's Element / SymFncStatistics / Atom code, tests/test_oracle.py):
extrapolation warnings (Element::updateSymmetryFunctionStatistics,
Element.cpp:462-541) and low-neighbour atoms (Mode.cpp:1667-1679), on water
boxes outside the scaling range and a dilute cluster, in both forms."""
import numpy as np
import pytest

import n2p2_amd
from n2p2_amd import remap, synth
from oracle_lib import BRUTE, PULL, Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def structures():
    out = []
    for f in (0.85, 1.3):
        pos, elem, box = synth.water_box(4, spacing=3.105 * f)
        out.append(("water x%.2f" % f, pos, elem, box))
    pos = np.array([[0.0, 0, 0], [9.0, 0, 0], [0, 9.0, 0], [0.5, 0.4, 0.3], [20.0, 20, 20]])
    out.append(("dilute", pos, np.array([1, 0, 0, 0, 1], np.int32), None))
    return out


@pytest.fixture(scope="module")
def water_dir(tmp_path_factory):
    return synth.water_model(str(tmp_path_factory.mktemp("water")))


@pytest.mark.parametrize("precision", [n2p2_amd.PLAIN_FP64, n2p2_amd.MIXED])
def test_extrapolation_warnings_and_low_neighbors(water_dir, precision):
    for name, pos, elem, box in structures():
        say("case: SF driver checks, %s, precision %d" % (name, precision))
        o = Oracle(water_dir)
        o.set_structure(pos, elem, box)
        o.predict(PULL, BRUTE)
        eo = o.extrapolation_warnings()
        m = n2p2_amd.Mode(water_dir)
        m.set_structures(pos if box is None else remap(pos, box), elem, None if box is None else box[None])
        m.evaluate(precision, forces=False)
        eg = m.extrapolation_warnings()
        lo, lg = o.low_neighbor_atoms(), m.low_neighbor_atoms()
        say("  %d warnings (oracle %d), %d low-neighbour atoms (oracle %d)" % (len(eg), len(eo), len(lg[0]),
                                                                               len(lo[0])))
        assert (lg[0] == lo[0]).all() and (lg[1] == lo[1]).all()
        if precision == n2p2_amd.PLAIN_FP64:
            assert len(eg) == len(eo)
            for f in ("atom", "sf", "type", "gmin", "gmax"):
                assert (eg[f] == eo[f]).all(), f
            assert np.abs(eg["value"] - eo["value"]).max(initial=0.0) <= 1e-10 * max(1.0, np.abs(eo["value"]).max(
                initial=0.0))
        else:
            # FP32 symmetry functions: the same records except ones whose value
            # lies within 1e-5 (relative) of a bound on the side that has them
            def near(r):
                return min(abs(r["value"] - r["gmin"]), abs(r["value"] - r["gmax"])) <= 1e-5 * max(1.0, abs(r["value"]))
            kg = {(int(r["atom"]), int(r["sf"])): r for r in eg}
            ko = {(int(r["atom"]), int(r["sf"])): r for r in eo}
            assert all(near(kg[k]) for k in set(kg) - set(ko))
            assert all(near(ko[k]) for k in set(ko) - set(kg))
            common = sorted(set(kg) & set(ko))
            assert len(common) >= 0.99 * len(ko)
