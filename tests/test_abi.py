"""The C ABI: the library loads, exports every entry point include/bhfluid.h
declares, and the host-side part of the API (tree construction, initial
data, regrid, state exchange, error codes) works without a GPU.  The step
itself must fail loudly when no device is present; there is no CPU path."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import bhfluid
import configs

HDR = Path(__file__).resolve().parents[1] / "include" / "bhfluid.h"


def declared_symbols():
    text = HDR.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(bh_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "bh_rk2_step" in syms and "bh_tree_create" in syms and len(syms) >= 25


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(str(bhfluid.library_path()))
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_host_api_roundtrip():
    lo, hi, nqlt, regrid, a, q = configs.R3
    t = bhfluid.AmrTree(lo, hi, nqlt, 1.0, a, q, configs.GAMMA)
    assert t.nblocks == nqlt ** 3
    rho_max = bhfluid.torus_setup(t, regrid=True)
    assert rho_max > 0
    meta, edges = t.blocks()
    assert meta.shape == (t.nblocks, 14) and edges.shape == (t.nblocks, 3, 17)
    levels = set(meta[meta[:, 4] == 1, 0])
    assert levels == {0, 1}  # two refinement levels among the leaves
    # parents and children are consistent
    for b, m in enumerate(meta):
        for c in m[6:]:
            if c >= 0:
                assert meta[c, 5] == b and meta[c, 0] == m[0] + 1
    U, W, Bf = t.state()
    U2 = U.copy()
    U2[0, 0, 0] = 123.0
    t.set_state(U=U2)
    assert t.state()[0][0, 0, 0] == 123.0
    with pytest.raises(ValueError):
        t.set_state(U=U[:1])


def test_invalid_arguments_rejected():
    with pytest.raises(ValueError):
        bhfluid.AmrTree((0, 0), (1, 1, 1), 1)
    with pytest.raises(ValueError):
        bhfluid.AmrTree((0, 0, 0), (1, 1, 1), 0)
    L = bhfluid.lib()
    lo = np.zeros(3)
    h = ctypes.c_void_p()
    dp = ctypes.POINTER(ctypes.c_double)
    assert L.bh_tree_create(lo.ctypes.data_as(dp), lo.ctypes.data_as(dp), 0, 1.0, 0.0, 0.0, 5 / 3,
                            ctypes.byref(h)) < 0
    assert L.bh_last_error(None)


def test_rkstg_rejects_bad_stage():
    t = bhfluid.AmrTree(*configs.R0[:3], 1.0, 0.9, 0.1, configs.GAMMA)
    with pytest.raises(bhfluid.BhError):
        bhfluid.RK2Integrator().rkstg(t, 0.1, 2)


@pytest.mark.skipif(Path("/dev/kfd").exists(), reason="a GPU is present")
def test_step_without_gpu_fails_loudly():
    t = bhfluid.AmrTree(*configs.R0[:3], 1.0, 0.9, 0.1, configs.GAMMA)
    bhfluid.torus_setup(t)
    with pytest.raises(bhfluid.BhError):
        bhfluid.RK2Integrator(0.4).step(t)


def test_fm_init_without_torus_returns_false():
    # a box above the torus holds no torus cells: the reference prints an
    # error and returns (initconds.cpp:483-486)
    t = bhfluid.AmrTree((-5, -5, 30), (5, 5, 40), 1, 1.0, 0.9, 0.1, configs.GAMMA)
    assert bhfluid.init.fm_init(t) is False


def test_residual_sign_proof_on_host():
    """Host build of the division-free residual sign (exact reciprocal and
    square root there): every proved sign agrees with the residual, and most
    points away from the root get one."""
    import c2p_sign_cases as cs
    L = bhfluid.lib()
    pts, near = cs.points(L, np.random.default_rng(5), n_sets=600)
    sign, f = cs.residual(L, pts, 0)
    assert cs.check(sign, f) == 0
    far = ~near & np.isfinite(f)
    assert (sign[far] != 0).mean() > 0.9


def test_dump_roundtrip_from_host_state(tmp_path):
    """The dump of a tree whose current state is on the host holds every
    leaf's active-zone W and the leaf table, as read_dump returns them."""
    lo, hi, nqlt, regrid, a, Q = configs.R3
    t = bhfluid.AmrTree(lo, hi, nqlt, 1.0, a, Q, configs.GAMMA)
    bhfluid.torus_setup(t, regrid=regrid)
    path = tmp_path / "state.bhd"
    t.write_dump(path, time=1.25, step=7)
    d = bhfluid.read_dump(path)
    meta, edges = t.blocks()
    leaf = meta[:, 4] == 1
    _, W, _ = t.state()
    Wl = W[leaf].reshape(-1, 8, 20, 20, 20)[:, :, 2:18, 2:18, 2:18]
    assert d["version"] == 1 and d["step"] == 7 and d["time"] == 1.25
    assert (d["M"], d["a"], d["Q"], d["gamma"]) == (1.0, a, Q, configs.GAMMA)
    assert np.array_equal(d["level"], meta[leaf, 0]) and np.array_equal(d["ijk"], meta[leaf, 1:4])
    e = edges.reshape(-1, 3, 17)[leaf]
    assert np.array_equal(d["lo"], e[:, :, 0]) and np.array_equal(d["hi"], e[:, :, 16])
    assert np.array_equal(d["W"].view(np.int64), Wl.view(np.int64))
    assert path.stat().st_size == 128 + 64 * leaf.sum() + leaf.sum() * 8 * 4096 * 8


def test_set_blocks_reproduces_a_regridded_tree():
    """bh_tree_set_blocks: a fresh tree given the block table of a regridded
    one (R3: 227 patches on three levels) has the same blocks, edges and cell
    geometry bit for bit, takes its state, and regrids to the same tree."""
    import configs
    lo, hi, nqlt, _, a, q = configs.R3
    src = bhfluid.AmrTree(lo, hi, nqlt, 1.0, a, q, configs.GAMMA)
    bhfluid.torus_setup(src, regrid=True)
    meta, edges = src.blocks()
    dst = bhfluid.AmrTree(lo, hi, nqlt, 1.0, a, q, configs.GAMMA)
    dst.set_blocks(meta)
    m2, e2 = dst.blocks()
    assert np.array_equal(m2, meta) and np.array_equal(e2.view(np.int64), edges.view(np.int64))
    dst.set_state(*src.state())
    for x, y in zip(src.state(), dst.state()):
        assert np.array_equal(x.view(np.int64), y.view(np.int64))
    src.regrid()
    dst.regrid()
    assert np.array_equal(src.blocks()[0], dst.blocks()[0])
    for x, y in zip(src.state(), dst.state()):
        assert np.array_equal(x.view(np.int64), y.view(np.int64))


def test_set_blocks_rejects_inconsistent_tables():
    import configs
    lo, hi, nqlt, _, a, q = configs.R2
    src = bhfluid.AmrTree(lo, hi, nqlt, 1.0, a, q, configs.GAMMA)
    bhfluid.torus_setup(src, regrid=True)
    meta, _ = src.blocks()
    dst = bhfluid.AmrTree(lo, hi, nqlt, 1.0, a, q, configs.GAMMA)
    child = int(np.flatnonzero(meta[:, 5] >= 0)[0])
    bad = meta.copy()
    bad[child, 1] += 2                      # ijk outside its parent
    with pytest.raises(bhfluid.BhError, match="block"):
        dst.set_blocks(bad)
    bad = meta.copy()
    bad[int(meta[child, 5]), 4] = 1          # a parent marked leaf
    with pytest.raises(bhfluid.BhError, match="leaf flag"):
        dst.set_blocks(bad)
    bad = meta.copy()
    bad[0, 5] = -1
    bad[0, 1] = 99                          # not a root of the grid
    with pytest.raises(bhfluid.BhError, match="root"):
        dst.set_blocks(bad)
    dst.set_blocks(meta)                    # the tree is still usable
    assert dst.nblocks == len(meta)


def test_params_limiter_roundtrip_and_validation():
    import configs
    lo, hi, nqlt, _, a, q = configs.R1
    t = bhfluid.AmrTree(lo, hi, nqlt, 1.0, a, q, configs.GAMMA)
    assert t.limiter == "vanleer"               # the reference's
    for name in ("minmod", "mc", "vanleer"):
        t.set_limiter(name)
        assert t.limiter == name
    P = bhfluid._Params()
    P.limiter = 3
    assert t._L.bh_tree_set_params(t._h, ctypes.byref(P)) < 0
    P.limiter = 0
    P.reserved[3] = 1
    assert t._L.bh_tree_set_params(t._h, ctypes.byref(P)) < 0
    assert t.limiter == "vanleer"               # rejected calls change nothing
