"""nnpx_predict_noalloc: the evaluation entry point that allocates no device
memory (its buffers come from the library's file-scope scratch array),
called through the C ABI and checked against the allocating entry points
(nnpx_set_structures + nnpx_evaluate + nnpx_get_energy / nnpx_get_forces) on
two calls in a row, the second with a smaller input, in both forms; and the
clear error for an input beyond the array."""
import ctypes

import numpy as np
import pytest

import n2p2_amd
from n2p2_amd import remap, synth

pytestmark = pytest.mark.gpu

NNPX_ERR_RANGE = 4


def say(*a):
    print(*a, flush=True)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Ctx:
    """A bare C-ABI context with a loaded model."""

    def __init__(self, lib, model_dir):
        self.lib = lib
        self.h = ctypes.c_void_p()
        assert lib.nnpx_create(ctypes.byref(self.h), 0) == 0, lib.nnpx_last_error(None)
        assert lib.nnpx_load_model(self.h, model_dir.encode()) == 0, lib.nnpx_last_error(self.h)

    def error(self):
        return self.lib.nnpx_last_error(self.h).decode()

    def predict_noalloc(self, precision, pos, elem, boxes, off):
        ns = len(off) - 1
        E = np.zeros(ns)
        F = np.zeros((off[-1], 3))
        st = self.lib.nnpx_predict_noalloc(
            self.h, precision, 1, ns, _p(np.ascontiguousarray(off, np.int64)),
            _p(np.ascontiguousarray(pos, np.float64)), _p(np.ascontiguousarray(elem, np.int32)),
            _p(np.ascontiguousarray(boxes, np.float64).reshape(ns, 9)), _p(np.ones(ns, np.int32)), _p(E), _p(F))
        return st, E, F

    def close(self):
        self.lib.nnpx_destroy(self.h)


def reference(model_dir, precision, pos, elem, boxes, off):
    m = n2p2_amd.Mode(model_dir)
    m.set_structures(pos, elem, boxes, off)
    m.evaluate(precision, forces=True)
    return m.energies(), m.forces()


@pytest.fixture(scope="module")
def water_dir(tmp_path_factory):
    return synth.water_model(str(tmp_path_factory.mktemp("water")))


def inputs(precision):
    """A water box (1,536 atoms in the mixed form; 192 in the plain form,
    whose stored dG/dr takes ~90 KB per atom), then a smaller batch."""
    pos, elem, box = synth.water_box(8 if precision == n2p2_amd.MIXED else 4)
    big = (remap(pos, box), elem, box[None], np.array([0, len(elem)]))
    p, e, b, o = synth.random_batch(n_struct=2, nmin=30, nmax=80)
    assert o[-1] < len(elem)
    return [big, (p, e, b, o)]


@pytest.mark.parametrize("precision", [n2p2_amd.MIXED, n2p2_amd.PLAIN_FP64])
def test_noalloc_equals_allocating_calls(water_dir, precision):
    lib = n2p2_amd.library()
    lib.nnpx_predict_noalloc.restype = ctypes.c_int
    c = Ctx(lib, water_dir)
    try:
        for call, (pos, elem, boxes, off) in enumerate(inputs(precision)):
            say("case: no-alloc call %d, %d atoms in %d structures, precision %d"
                % (call, off[-1], len(off) - 1, precision))
            st, E, F = c.predict_noalloc(precision, pos, elem, boxes, off)
            assert st == 0, c.error()
            Er, Fr = reference(water_dir, precision, pos, elem, boxes, off)
            de = float(np.abs(E - Er).max() / np.abs(Er).max())
            df = float(np.abs(F - Fr).max())
            say("  max E rel %.2e  max |dF| %.2e" % (de, df))
            assert de <= 1e-12
            assert df <= 1e-10
        # the context holds no structures afterwards
        out = np.zeros(1)
        assert lib.nnpx_get_energy(c.h, _p(out)) != 0
    finally:
        c.close()


def test_noalloc_rejects_oversize(water_dir):
    lib = n2p2_amd.library()
    c = Ctx(lib, water_dir)
    try:
        pos, elem, box = synth.water_box(24)  # 41,472 atoms: beyond 64 MiB
        say("case: no-alloc oversize input, %d atoms" % len(elem))
        st, _, _ = c.predict_noalloc(n2p2_amd.MIXED, remap(pos, box), elem, box[None], np.array([0, len(elem)]))
        say("  status %d: %s" % (st, c.error()))
        assert st == NNPX_ERR_RANGE
        assert "scratch" in c.error()
        # a later call that fits still works
        pos, elem, box = synth.water_box(4)
        st, E, F = c.predict_noalloc(n2p2_amd.MIXED, remap(pos, box), elem, box[None], np.array([0, len(elem)]))
        assert st == 0, c.error()
        Er, Fr = reference(water_dir, n2p2_amd.MIXED, remap(pos, box), elem, box[None], np.array([0, len(elem)]))
        assert abs(E[0] - Er[0]) <= 1e-12 * abs(Er[0])
        assert np.abs(F - Fr).max() <= 1e-10
    finally:
        c.close()
