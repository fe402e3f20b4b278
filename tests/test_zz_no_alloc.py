"""The allocation-free diagnostics entry point (bhs_run_diagnostics_static),
called through the C ABI, against bhs_run_diagnostics: the same eight words,
bit for bit, on two calls in a row, the second on a smaller run."""
import ctypes

import numpy as np
import pytest

from bhfluid import sph

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _both(eng, run):
    L = sph.lib()
    a = np.zeros(8)
    b = np.zeros(8)
    assert L.bhs_run_diagnostics_static(eng._h, run.id, a.ctypes.data_as(_dp)) == 0
    assert L.bhs_run_diagnostics(eng._h, run.id, b.ctypes.data_as(_dp)) == 0
    return a, b


def test_static_diagnostics_match_allocating_ones():
    big = dict(sph.c2_grid(), nr=64, nth=32, nph=32)
    small = dict(sph.c2_grid(), nr=32, nth=32, nph=32)
    bb, bs = sph.uniform_blocks(big, 16), sph.uniform_blocks(small, 16)
    assert len(bs) < len(bb)
    eng = sph.Engine(16, len(bb) + len(bs))
    r1 = eng.add_run(sph.PHYSICS, big, bb)
    # tables not yet on the device: refused, nothing launched
    out = np.zeros(8)
    assert sph.lib().bhs_run_diagnostics_static(eng._h, r1.id, out.ctypes.data_as(_dp)) == -3
    r1.init_torus(sph.TORUS, 42)
    eng.step([r1], 3)
    a, b = _both(eng, r1)
    assert np.array_equal(a, b), (a, b)
    assert a[0] > 0 and a[3] > 0
    r2 = eng.add_run(dict(sph.PHYSICS, a=0.5), small, bs)
    r2.init_torus(dict(sph.TORUS, beta=30.0), 7)
    eng.step([r1, r2], 2)
    a2, b2 = _both(eng, r2)          # second call, smaller run
    assert np.array_equal(a2, b2), (a2, b2)
    assert not np.array_equal(a2, a)
    a3, b3 = _both(eng, r1)          # and the larger one again
    assert np.array_equal(a3, b3)
