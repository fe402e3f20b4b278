"""ctypes bindings for the two CPU checkers under oracle/.

* ``Oracle``  -- oracle/liboracle.so, the CPU restatement of the reference's
  fluid step (test infrastructure; the product never loads it).
* ``RefTree`` -- oracle/_ref/libbhref.so, the reference's own fluid sources
  compiled by oracle/Makefile plus a C-ABI harness.  Only present where it
  was built; tests that need it skip otherwise.

Both exchange state in the product's SoA layout (U, W: [nb][8][8000];
Bf: [nb][3][8400]) and the block table ``meta[nb][14]`` / ``edges[nb][3][17]``.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libbhref.so"

NCELL = 8000
NFACE = 8400
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_d = ctypes.c_double
_i = ctypes.c_int
_v = ctypes.c_void_p
_ll = ctypes.c_longlong


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _pi(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            raise FileNotFoundError(f"{ORACLE_SO} not built (make -C oracle)")
        L = ctypes.CDLL(str(ORACLE_SO))
        _sig(L, "or_tree_new", _v, [_i, _ip, _dp, _d, _d, _d, _d, _i])
        _sig(L, "or_tree_free", None, [_v])
        _sig(L, "or_init_torus", _i, [_v, _i, _dp])
        _sig(L, "or_set_state", None, [_v, _dp, _dp, _dp, _dp])
        _sig(L, "or_get_state", None, [_v, _dp, _dp, _dp, _dp])
        _sig(L, "or_step", _ll, [_v, _d, _dp])
        _sig(L, "or_dtcomp", _ll, [_v, _d, _dp])
        _sig(L, "or_rkstg", _ll, [_v, _d, _i])
        _sig(L, "or_floors", _ll, [_v])
        _sig(L, "or_ghosts_all", _ll, [_v])
        _sig(L, "or_c2p_leaves", _ll, [_v])
        _sig(L, "or_flux_ct_patch", _ll, [_v, _i, _d, _dp, _dp])
        _sig(L, "or_update_terms", _ll, [_v, _i, _dp, _dp])
        _sig(L, "or_maxdivB", _d, [_v, _i])
        _sig(L, "or_cell_geom", None, [_v, _i, _dp])
        _sig(L, "or_diagnostics", None, [_v, _dp, _dp])
        _sig(L, "or_set_threads", None, [_i])
        _sig(L, "or_set_limiter", None, [_v, _i])
        _sig(L, "or_metric", None, [_d, _d, _d, _d, _d, _dp])
        _sig(L, "or_ptoc", _i, [_d, _d, _d, _d, _dp, _d, _d, _dp, _dp])
        _sig(L, "or_ctop", _i, [_d, _d, _d, _d, _dp, _d, _d, _dp, _dp, _i])
        _sig(L, "or_hlld", _i, [_d, _d, _d, _d, _dp, _dp, _dp, _dp, _d, _d, _i, _dp, _dp])
        _sig(L, "or_reconfp", None, [_v, _i, _i, _i, _i, _i, _i, _dp, _dp])
        _sig(L, "or_log", _d, [_d, _i])
        _sig(L, "or_exp", _d, [_d, _i])
        _sig(L, "or_ic_fn", _d, [_i, _d, _d, _i])
        _oracle = L
    return _oracle


def ref_available() -> bool:
    return REF_SO.exists()


def ref_one_thread(L, request):
    """Run the compiled reference on one OpenMP thread until the calling test
    ends.  omp_set_num_threads sets the count of the one OpenMP runtime that
    the compiled reference, the oracle and the product share in this process,
    so the previous count is put back at the end of the test; left at 1, every
    later oracle step in the process would run on one thread."""
    get = L.omp_get_max_threads
    get.restype, get.argtypes = ctypes.c_int, []
    n = get()
    L.ref_set_threads(1)
    request.addfinalizer(lambda: L.ref_set_threads(n))


ADAPTER_SO = REF_SO.parent / "libbhadapter.so"
_adapter = None


def adapter_lib():
    """oracle/_ref/libbhadapter.so: the mode-B drop-in (oracle/gpu_rk2_adapter.hpp)
    compiled against the reference's headers, stepping reference trees on the GPU."""
    global _adapter
    if _adapter is None:
        if not ADAPTER_SO.exists():
            raise FileNotFoundError(f"{ADAPTER_SO} not built (make -C oracle ref)")
        L = ctypes.CDLL(str(ADAPTER_SO))
        _sig(L, "adapter_new", _v, [_d, _dp, _dp, _i])
        _sig(L, "adapter_free", None, [_v])
        _sig(L, "adapter_error", ctypes.c_char_p, [])
        _sig(L, "adapter_step", _i, [_v, _v, _dp])
        _adapter = L
    return _adapter


def ref_lib():
    global _ref
    if _ref is None:
        if not REF_SO.exists():
            raise FileNotFoundError(f"{REF_SO} not built (make -C oracle ref)")
        L = ctypes.CDLL(str(REF_SO))
        _sig(L, "ref_tree_new", _v, [_dp, _dp, _i, _d, _d, _d, _d])
        _sig(L, "ref_tree_free", None, [_v])
        _sig(L, "ref_tree_error", ctypes.c_char_p, [_v])
        _sig(L, "ref_nblocks", _i, [_v])
        _sig(L, "ref_get_blocks", None, [_v, _ip, _dp])
        _sig(L, "ref_get_state", None, [_v, _dp, _dp, _dp, _dp])
        _sig(L, "ref_set_state", None, [_v, _dp, _dp, _dp, _dp])
        _sig(L, "ref_fm_init", _i, [_v])
        _sig(L, "ref_rho_max", _d, [_v])
        _sig(L, "ref_B_pot_init_all", None, [_v, _d])
        _sig(L, "ref_ptoc_interior", _i, [_v])
        _sig(L, "ref_regrid", _i, [_v])
        _sig(L, "ref_set_threads", None, [_i])
        _sig(L, "ref_step", _i, [_v, _d, _dp])
        _sig(L, "ref_dtcomp", _i, [_v, _d, _dp])
        _sig(L, "ref_rkstg", _i, [_v, _d, _i])
        _sig(L, "ref_floors", _i, [_v])
        _sig(L, "ref_ghosts_all", _i, [_v])
        _sig(L, "ref_flux_ct_patch", _i, [_v, _i, _d, _dp, _dp])
        _sig(L, "ref_update_terms", _i, [_v, _i, _dp, _dp])
        _sig(L, "ref_maxdivB", _d, [_v, _i])
        _sig(L, "ref_cell_geom", None, [_v, _i, _dp])
        _sig(L, "ref_diagnostics", None, [_v, _dp])
        _sig(L, "ref_metric", None, [_d, _d, _d, _d, _d, _dp])
        _sig(L, "ref_ptoc", _i, [_d, _d, _d, _d, _dp, _d, _d, _dp, _dp])
        _sig(L, "ref_ctop", _i, [_d, _d, _d, _d, _dp, _d, _d, _dp, _dp])
        _sig(L, "ref_hlld", _i, [_d, _d, _d, _d, _dp, _dp, _dp, _dp, _d, _d, _i, _dp, _dp])
        _sig(L, "ref_reconfp", None, [_v, _i, _i, _i, _i, _i, _i, _dp, _dp])
        _sig(L, "ref_B_init", None, [_v, _i])
        _ref = L
    return _ref


class State:
    """A tree snapshot: block table plus SoA arrays."""

    def __init__(self, meta, edges, U, W, Bf, dU=None):
        self.meta = np.ascontiguousarray(meta, dtype=np.int32)
        self.edges = np.ascontiguousarray(edges, dtype=np.float64)
        self.U = np.ascontiguousarray(U, dtype=np.float64)
        self.W = np.ascontiguousarray(W, dtype=np.float64)
        self.Bf = np.ascontiguousarray(Bf, dtype=np.float64)
        self.dU = None if dU is None else np.ascontiguousarray(dU, dtype=np.float64)

    @property
    def nb(self):
        return self.meta.shape[0]

    def copy(self):
        return State(self.meta.copy(), self.edges.copy(), self.U.copy(), self.W.copy(), self.Bf.copy(),
                     None if self.dU is None else self.dU.copy())


class RefTree:
    """The compiled reference's ``grid::amrtree`` (grid.cpp:177)."""

    def __init__(self, lo, hi, nqlt, M=1.0, a=0.9, Q=0.1, gm=5.0 / 3.0):
        self.L = ref_lib()
        lo = np.asarray(lo, np.float64)
        hi = np.asarray(hi, np.float64)
        self.h = self.L.ref_tree_new(_p(lo), _p(hi), int(nqlt), M, a, Q, gm)
        self.params = (M, a, Q, gm)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_tree_free(self.h)
            self.h = None

    def error(self):
        return self.L.ref_tree_error(self.h).decode()

    def torus_init(self, regrid=False):
        """main.cpp:98-159 (+ optional regrid at main.cpp:168)."""
        L = self.L
        if L.ref_fm_init(self.h):
            raise RuntimeError(self.error())
        rm = L.ref_rho_max(self.h)
        L.ref_B_pot_init_all(self.h, rm)
        if L.ref_ptoc_interior(self.h):
            raise RuntimeError(self.error())
        if regrid and L.ref_regrid(self.h):
            raise RuntimeError(self.error())
        return rm

    @property
    def nb(self):
        return self.L.ref_nblocks(self.h)

    def blocks(self):
        nb = self.nb
        meta = np.zeros((nb, 14), np.int32)
        edges = np.zeros((nb, 3, 17), np.float64)
        self.L.ref_get_blocks(self.h, _pi(meta), _p(edges))
        return meta, edges

    def state(self, with_dU=False) -> State:
        nb = self.nb
        meta, edges = self.blocks()
        U = np.zeros((nb, 8, NCELL))
        W = np.zeros((nb, 8, NCELL))
        Bf = np.zeros((nb, 3, NFACE))
        dU = np.zeros((nb, 8, NCELL)) if with_dU else None
        self.L.ref_get_state(self.h, _p(U), _p(W), _p(Bf), _p(dU))
        return State(meta, edges, U, W, Bf, dU)

    def set_state(self, s: State):
        self.L.ref_set_state(self.h, _p(s.U), _p(s.W), _p(s.Bf), _p(s.dU))

    def step(self, cfl=0.4):
        dt = _d(0.0)
        rc = self.L.ref_step(self.h, cfl, ctypes.byref(dt))
        return dt.value, rc

    def dtcomp(self, cfl=0.4):
        dt = _d(0.0)
        rc = self.L.ref_dtcomp(self.h, cfl, ctypes.byref(dt))
        return dt.value, rc

    def rkstg(self, dt, stage):
        return self.L.ref_rkstg(self.h, dt, stage)

    def floors(self):
        return self.L.ref_floors(self.h)

    def ghosts_all(self):
        return self.L.ref_ghosts_all(self.h)

    def diagnostics(self):
        """mass, energy, L_pol, Edot, Phi_B, thermE, Hscal (quantitatives.cpp)."""
        out = np.zeros(7)
        self.L.ref_diagnostics(self.h, _p(out))
        return out


class OracleTree:
    """The oracle restatement, loaded with a block table and a state."""

    def __init__(self, meta, edges, M=1.0, a=0.9, Q=0.1, gm=5.0 / 3.0, libm=0):
        self.L = oracle_lib()
        self.meta = np.ascontiguousarray(meta, np.int32)
        self.edges = np.ascontiguousarray(edges, np.float64)
        self.nb = self.meta.shape[0]
        self.h = self.L.or_tree_new(self.nb, _pi(self.meta), _p(self.edges), M, a, Q, gm, int(libm))

    @classmethod
    def from_state(cls, s: State, params, libm=0):
        t = cls(s.meta, s.edges, *params, libm=libm)
        t.set_state(s)
        return t

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_tree_free(self.h)
            self.h = None

    def set_state(self, s: State):
        self.L.or_set_state(self.h, _p(s.U), _p(s.W), _p(s.Bf), _p(s.dU))

    def set_limiter(self, lim: int):
        """reconfp's limiter in the step: 0 minmod, 1 van Leer (default), 2 MC."""
        self.L.or_set_limiter(self.h, int(lim))

    def state(self, with_dU=False) -> State:
        nb = self.nb
        U = np.zeros((nb, 8, NCELL))
        W = np.zeros((nb, 8, NCELL))
        Bf = np.zeros((nb, 3, NFACE))
        dU = np.zeros((nb, 8, NCELL)) if with_dU else None
        self.L.or_get_state(self.h, _p(U), _p(W), _p(Bf), _p(dU))
        return State(self.meta, self.edges, U, W, Bf, dU)

    def init_torus(self, libm=0):
        """The initial data of main.cpp:94-159 on this (fresh) block table;
        libm 0 = C library, 1 = correctly rounded.  Returns (code, rho_max)
        with code 0, 1 (L0_comp throws), 2 (no torus), 3 (rho_max < 1e-14)."""
        rm = _d(0.0)
        rc = self.L.or_init_torus(self.h, int(libm), ctypes.byref(rm))
        return rc, rm.value

    def step(self, cfl=0.4):
        dt = _d(0.0)
        faults = self.L.or_step(self.h, cfl, ctypes.byref(dt))
        return dt.value, faults

    def dtcomp(self, cfl=0.4):
        dt = _d(0.0)
        faults = self.L.or_dtcomp(self.h, cfl, ctypes.byref(dt))
        return dt.value, faults

    def rkstg(self, dt, stage):
        return self.L.or_rkstg(self.h, dt, stage)

    def floors(self):
        return self.L.or_floors(self.h)

    def ghosts_all(self):
        return self.L.or_ghosts_all(self.h)

    def diagnostics(self):
        """(values[7], sum of |terms|[8]) as or_diagnostics."""
        out = np.zeros(7)
        ab = np.zeros(8)
        self.L.or_diagnostics(self.h, _p(out), _p(ab))
        return out, ab


DIAG_NAMES = ("mass", "energy", "L_pol", "Edot", "Phi_B", "thermE", "Hscal")


def bits_equal(a, b):
    """Bitwise equality that treats every NaN alike."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(np.all((a.view(np.int64) == b.view(np.int64)) | both_nan))


def first_diff(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    bad = ~((a.view(np.int64) == b.view(np.int64)) | (np.isnan(a) & np.isnan(b)))
    idx = np.flatnonzero(bad)
    return (len(idx), None) if len(idx) == 0 else (len(idx), (int(idx[0]), a[idx[0]], b[idx[0]]))
