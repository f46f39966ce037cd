"""Caller-provided neighbour lists (the LAMMPS path, InterfaceLammps::
addNeighbor / getForces): a periodic box is handed over as local atoms plus
ghost images with explicit distance vectors; forces land on locals and
ghosts (push form) and, folded back onto the owners as LAMMPS' reverse
communication does, equal the periodic push-form forces of the oracle."""
import numpy as np
import pytest

import n2p2_amd
from n2p2_amd import remap, synth
from oracle_lib import BRUTE, PUSH, Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def lammps_like(pos, elem, box, model_dir):
    """Local atoms + one ghost per (atom, non-zero image) entry of the
    reference neighbour list."""
    o = Oracle(model_dir)
    o.set_structure(pos, elem, box)
    o.build_neighbors(BRUTE)
    counts, js, shifts = o.neighbors()
    n = len(elem)
    ghost = {}
    owner = []
    index, nb_elem, dr = [], [], []
    p = 0
    for i in range(n):
        for _ in range(counts[i]):
            j, b = int(js[p]), tuple(int(x) for x in shifts[p])
            if b == (0, 0, 0):
                idx = j
            else:
                if (j, b) not in ghost:
                    ghost[(j, b)] = n + len(owner)
                    owner.append(j)
                idx = ghost[(j, b)]
            d = pos[i] - pos[j]
            d = d + b[0] * box[0]
            d = d + b[1] * box[1]
            d = d + b[2] * box[2]
            index.append(idx)
            nb_elem.append(elem[j])
            dr.append(d)
            p += 1
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return offsets, np.array(index), np.array(nb_elem), np.array(dr), np.array(owner, dtype=np.int64)


@pytest.mark.parametrize("precision", [n2p2_amd.PLAIN_FP64, n2p2_amd.MIXED])
def test_external_list_equals_periodic_push(precision):
    say("case: external neighbour list (LAMMPS path), water-192, precision", precision)
    import tempfile
    d = synth.water_model(tempfile.mkdtemp())
    pos, elem, box = synth.water_box(4)
    pos = remap(pos, box)
    offsets, index, nb_elem, dr, owner = lammps_like(pos, elem, box, d)
    n = len(elem)
    m = n2p2_amd.Mode(d)
    m.set_neighbors_external(elem, n + len(owner), offsets, index, nb_elem, dr)
    m.evaluate(precision, forces=True)
    E = m.energies()[0]
    F = m.forces()
    assert F.shape == (n + len(owner), 3)
    Ff = F[:n].copy()
    np.add.at(Ff, owner, F[n:])
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    o.predict(PUSH, BRUTE)
    Eo, Fo = o.energy(), o.forces()
    erel = abs(E - Eo) / abs(Eo)
    dF = float(np.abs(Ff - Fo).max())
    rms = float(np.sqrt(np.mean((Ff - Fo) ** 2)))
    say("  ghosts %d  E rel %.2e  max|dF| %.2e  rms %.2e" % (len(owner), erel, dF, rms))
    if precision == n2p2_amd.PLAIN_FP64:
        assert erel <= 1e-12 and dF <= 1e-10
    else:
        assert erel <= 1e-5 and rms <= 1e-3


def test_external_list_after_periodic_structure():
    """A context that first evaluated a sparse periodic box (small fitted
    neighbour staging) and is then handed a denser external list: the
    staging is refitted to the external list and the mixed result equals
    the oracle's push form, as on a fresh context."""
    say("case: set_structures (sparse box) then external list (water-192), mixed")
    import tempfile
    d = synth.water_model(tempfile.mkdtemp())
    pos, elem, box = synth.water_box(4)
    pos = remap(pos, box)
    m = n2p2_amd.Mode(d)
    sparse = 1.6 * box
    m.set_structures(remap(1.6 * pos, sparse), elem, sparse[None])
    m.evaluate(n2p2_amd.MIXED, forces=True)
    offsets, index, nb_elem, dr, owner = lammps_like(pos, elem, box, d)
    n = len(elem)
    m.set_neighbors_external(elem, n + len(owner), offsets, index, nb_elem, dr)
    m.evaluate(n2p2_amd.MIXED, forces=True)
    E = m.energies()[0]
    F = m.forces()
    Ff = F[:n].copy()
    np.add.at(Ff, owner, F[n:])
    o = Oracle(d)
    o.set_structure(pos, elem, box)
    o.predict(PUSH, BRUTE)
    Eo, Fo = o.energy(), o.forces()
    erel = abs(E - Eo) / abs(Eo)
    rms = float(np.sqrt(np.mean((Ff - Fo) ** 2)))
    say("  dense list: max neighbours %d  E rel %.2e  F rms %.2e" % (int(np.diff(offsets).max()), erel, rms))
    assert erel <= 1e-5 and rms <= 1e-5
    # a second read returns the same evaluation
    assert m.energies()[0] == E
