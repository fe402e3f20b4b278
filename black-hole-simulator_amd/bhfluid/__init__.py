"""bhfluid: MI355X-native GRMHD fluid step, a drop-in for the fluid path of
Fjoerde/black-hole-simulator (src/fluid).

The reference exposes this path as C++ classes; this package mirrors them:

=========================  ==========================================================
bhfluid                    reference
=========================  ==========================================================
``AmrTree``                ``grid::amrtree`` (src/fluid/grid.hpp:109-144)
``init.fm_init``           ``torus::init::fm_init`` (solver/initconds.cpp:228-591)
``init.B_pot_init``        ``torus::init::B_pot_init`` over every patch (main.cpp:142-145)
``RK2Integrator``          ``integ::rk2integrator`` (src/fluid/rk2.hpp:19-30)
``RK2Integrator.step``     ``rk2integrator::step`` (solver/integrator.cpp:262-275)
``DomainError``            ``std::domain_error`` thrown by the EOS (physics/thermo.cpp:11)
=========================  ==========================================================

The step runs on the GPU through the C ABI in include/bhfluid.h
(``libbhfluid.so``, built in-tree by ``make -C black-hole-simulator_amd``).
There is no CPU fallback: without the library, or without a GPU, the step
raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Optional, Sequence, Tuple

import numpy as np

__all__ = ["AmrTree", "RK2Integrator", "DomainError", "BhError", "init", "torus_setup", "library_path",
           "NCELL", "NFACE", "BLOCK", "GHOST"]

BLOCK = 16
GHOST = 2
CPN = BLOCK + 2 * GHOST
NCELL = CPN ** 3
NFACE = (CPN + 1) * CPN * CPN

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libbhfluid.so"
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class _Params(ctypes.Structure):
    """bh_params (include/bhfluid.h)."""
    _fields_ = [("limiter", ctypes.c_int), ("reserved", ctypes.c_int * 7)]
_d = ctypes.c_double
_i = ctypes.c_int
_v = ctypes.c_void_p


class BhError(RuntimeError):
    """A call into libbhfluid failed (bad arguments, device or state error)."""


class DomainError(ValueError):
    """The equation of state saw rho <= 0.

    The reference throws ``std::domain_error`` at that point
    (src/fluid/physics/thermo.cpp:11) and the process terminates.  Here the
    step completes, the fault is counted, and this exception is raised
    afterwards; ``faults`` holds the count.
    """

    def __init__(self, msg, faults):
        super().__init__(msg)
        self.faults = faults


def library_path() -> Path:
    return _LIB_PATH


def _sig(L, name, res, args):
    f = getattr(L, name)
    f.restype = res
    f.argtypes = args


def lib():
    """Load libbhfluid.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise BhError(f"{_LIB_PATH} is missing: build it with `make -C black-hole-simulator_amd` "
                          "(or __graft_entry__.build())")
        L = ctypes.CDLL(str(_LIB_PATH))
        _sig(L, "bh_tree_create", _i, [_dp, _dp, _i, _d, _d, _d, _d, ctypes.POINTER(_v)])
        _sig(L, "bh_tree_destroy", None, [_v])
        _sig(L, "bh_last_error", ctypes.c_char_p, [_v])
        _sig(L, "bh_tree_fm_init", _i, [_v])
        _sig(L, "bh_tree_rho_max", _i, [_v, _dp])
        _sig(L, "bh_tree_B_pot_init", _i, [_v, _d])
        _sig(L, "bh_tree_ptoc_interior", _i, [_v])
        _sig(L, "bh_tree_regrid", _i, [_v])
        _sig(L, "bh_tree_set_blocks", _i, [_v, _i, _ip])
        _sig(L, "bh_tree_set_params", _i, [_v, ctypes.POINTER(_Params)])
        _sig(L, "bh_tree_get_params", _i, [_v, ctypes.POINTER(_Params)])
        _sig(L, "bh_tree_nblocks", _i, [_v, _ip])
        _sig(L, "bh_tree_get_blocks", _i, [_v, _ip, _dp])
        _sig(L, "bh_tree_get_state", _i, [_v, _dp, _dp, _dp])
        _sig(L, "bh_tree_set_state", _i, [_v, _dp, _dp, _dp])
        _sig(L, "bh_rk2_step", _i, [_v, _d, _dp])
        _sig(L, "bh_rk2_dtcomp", _i, [_v, _d, _dp])
        _sig(L, "bh_rk2_rkstg", _i, [_v, _d, _i])
        _sig(L, "bh_tree_floors", _i, [_v])
        _sig(L, "bh_rk2_advance", _i, [_v, _d, _i, _d, _dp, _ip])
        _sig(L, "bh_tree_take_faults", _i, [_v, ctypes.POINTER(ctypes.c_longlong)])
        _sig(L, "bh_rk2_steps", _i, [_v, _d, _i, _dp])
        _sig(L, "bh_tree_prepare", _i, [_v])
        _sig(L, "bh_profile", _i, [_v, _i])
        _sig(L, "bh_profile_report", _i, [_v, ctypes.c_char_p, _i])
        _sig(L, "bh_device_sync", _i, [])
        _sig(L, "bh_tree_maxdivB", _i, [_v, _i, _dp])
        _sig(L, "bh_ptoc_batch", _i, [_d, _d, _d, _d, _i, _dp, _dp, _dp, _dp])
        _sig(L, "bh_ctop_batch", _i, [_d, _d, _d, _d, _i, _dp, _dp, _dp, _dp, _ip])
        _sig(L, "bh_hlld_batch", _i, [_d, _d, _d, _d, _i, _dp, _dp, _i, _dp, _dp, _dp, _dp, _dp, _dp])
        _sig(L, "bh_math_batch", _i, [_i, _i, _dp, _dp, _i])
        _sig(L, "bh_c2p_sign_batch", _i, [_d, _i, _dp, _ip, _dp, _i])
        _sig(L, "bh_tree_diagnostics", _i, [_v, _dp])
        _sig(L, "bh_tree_regrid_device", _i, [_v, _ip])
        _sig(L, "bh_tree_init_torus_device", _i, [_v, _dp])
        _sig(L, "bh_math2_batch", _i, [_i, _i, _dp, _dp, _dp, _i])
        _sig(L, "bh_tree_use_graph", _i, [_v, _i])
        _sig(L, "bh_tree_set_exact_copies", _i, [_v, _i])
        _sig(L, "bh_tree_maxdivB_device", _i, [_v, _dp, _dp])
        _sig(L, "bh_profile_only", _i, [_v, ctypes.c_char_p])
        _sig(L, "bh_tree_write_dump", _i, [_v, ctypes.c_char_p, _d, ctypes.c_longlong])
        _sig(L, "bh_device_info", _i, [ctypes.c_char_p, _i])
        _sig(L, "bh_copy_probe", _i, [ctypes.c_longlong, _i, _dp])
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def _pi(a):
    if a.dtype != np.int32 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous int32 array")
    return a.ctypes.data_as(_ip)


class Metric:
    """Black-hole parameters (``metric``, src/fluid/metric.hpp:23-31)."""

    def __init__(self, M=1.0, a=0.0, Q=0.0):
        self.M, self.a, self.Q = float(M), float(a), float(Q)

    def r_horizon(self):
        return self.M + np.sqrt(self.M * self.M - self.a * self.a - self.Q * self.Q)


class State:
    """Ideal-gas EOS parameter (``state``, src/fluid/state.hpp:9-20)."""

    def __init__(self, gm=5.0 / 3.0):
        self.gamma = float(gm)


class AmrTree:
    """Block-structured tree of 16^3 patches with two ghost layers.

    Mirrors ``grid::amrtree(dom_l, dom_h, nqlt, M, a, Q, gm)``
    (src/fluid/grid/grid.cpp:177-206): ``nqlt**3`` root patches tiling the box,
    every cell at the floor state.  The state lives on the GPU while it is
    being stepped and is copied back on demand (``state()``).
    """

    def __init__(self, dom_l: Sequence[float], dom_h: Sequence[float], nqlt: int, M: float = 1.0, a: float = 0.0,
                 Q: float = 0.0, gm: float = 5.0 / 3.0):
        L = lib()
        lo = np.ascontiguousarray(dom_l, dtype=np.float64)
        hi = np.ascontiguousarray(dom_h, dtype=np.float64)
        if lo.shape != (3,) or hi.shape != (3,) or int(nqlt) < 1:
            raise ValueError("dom_l and dom_h need three coordinates and nqlt >= 1")
        h = _v()
        rc = L.bh_tree_create(_p(lo), _p(hi), int(nqlt), float(M), float(a), float(Q), float(gm), ctypes.byref(h))
        if rc != 0:
            raise BhError(L.bh_last_error(None).decode())
        self._h = h
        self._L = L
        self.mtr = Metric(M, a, Q)
        self.stt = State(gm)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.bh_tree_destroy(h)
            self._h = None

    # -- error handling -------------------------------------------------
    def _check(self, rc, what):
        if rc == 0:
            return
        if rc == 1:
            n = ctypes.c_longlong(0)
            self._L.bh_tree_take_faults(self._h, ctypes.byref(n))
            raise DomainError(f"{what}: Thermodynamic calculations threw back an error: the density is not "
                              f"positive! ({n.value} equation-of-state calls)", n.value)
        raise BhError(f"{what}: {self._L.bh_last_error(self._h).decode()} (code {rc})")

    # -- structure --------------------------------------------------------
    @property
    def nblocks(self) -> int:
        n = ctypes.c_int(0)
        self._L.bh_tree_nblocks(self._h, ctypes.byref(n))
        return n.value

    def blocks(self) -> Tuple[np.ndarray, np.ndarray]:
        """Block table: meta[nb, 14] (lvl, i, j, k, leaf, parent, children[8])
        and edges[nb, 3, 17] (cell-edge coordinates), in quilt order."""
        nb = self.nblocks
        meta = np.zeros((nb, 14), np.int32)
        edges = np.zeros((nb, 3, 17), np.float64)
        self._L.bh_tree_get_blocks(self._h, _pi(meta), _p(edges))
        return meta, edges

    @property
    def leaves(self) -> np.ndarray:
        meta, _ = self.blocks()
        return np.flatnonzero(meta[:, 4])

    def state(self):
        """Copy of (U, W, Bf): U, W [nb, 8, 8000], Bf [nb, 3, 8400]."""
        nb = self.nblocks
        U = np.zeros((nb, 8, NCELL))
        W = np.zeros((nb, 8, NCELL))
        Bf = np.zeros((nb, 3, NFACE))
        self._check(self._L.bh_tree_get_state(self._h, _p(U), _p(W), _p(Bf)), "get_state")
        return U, W, Bf

    def set_state(self, U=None, W=None, Bf=None):
        nb = self.nblocks
        args = []
        for a, shp in ((U, (nb, 8, NCELL)), (W, (nb, 8, NCELL)), (Bf, (nb, 3, NFACE))):
            if a is None:
                args.append(None)
                continue
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.shape != shp:
                raise ValueError(f"expected shape {shp}, got {a.shape}")
            args.append(a)
        self._check(self._L.bh_tree_set_state(self._h, *[_p(a) for a in args]), "set_state")

    # -- host-side tree operations ---------------------------------------
    def rho_max(self) -> float:
        """Maximum rho over active zones (src/fluid/main.cpp:101-111)."""
        out = ctypes.c_double(0.0)
        self._check(self._L.bh_tree_rho_max(self._h, ctypes.byref(out)), "rho_max")
        return out.value

    def ptoc_interior(self):
        """U = ptoc(W) on active zones (src/fluid/main.cpp:148-159)."""
        self._check(self._L.bh_tree_ptoc_interior(self._h), "ptoc_interior")

    def regrid_device(self) -> int:
        """``amrtree::regrid`` with the state kept on the GPU (flags and data
        work on the device, block-table changes on the host); identical
        result to ``regrid``.  Returns blocks refined plus parents
        flattened."""
        n = ctypes.c_int(0)
        self._check(self._L.bh_tree_regrid_device(self._h, ctypes.byref(n)), "regrid_device")
        return n.value

    def init_torus_device(self) -> float:
        """The whole initial data of src/fluid/main.cpp:94-159 computed on the
        GPU (constructor floor state, ``fm_init``, global rho_max,
        ``B_pot_init``, ``ptoc`` of the active zones), with correctly rounded
        transcendental functions; the state stays on the device.  Returns the
        global rho_max.  A torus that cannot be placed leaves the atmosphere,
        as in the reference (``self.torus_placed`` is then False)."""
        out = ctypes.c_double(0.0)
        rc = self._L.bh_tree_init_torus_device(self._h, ctypes.byref(out))
        self.torus_placed = rc != 2
        if rc != 2:
            self._check(rc, "init_torus_device")
        return out.value

    def set_exact_copies(self, enable: bool = True):
        """Testing switch: with False every copied ghost runs its own con2prim
        (no reuse of the source zone's outcome); results are identical."""
        self._check(self._L.bh_tree_set_exact_copies(self._h, int(bool(enable))), "set_exact_copies")

    def use_graph(self, enable: bool = True):
        """Replay the step from a captured HIP graph (default) or launch its
        kernels one by one; identical results."""
        self._check(self._L.bh_tree_use_graph(self._h, int(bool(enable))), "use_graph")

    def regrid(self):
        """``amrtree::regrid`` (src/fluid/grid/grid.cpp:739-835)."""
        self._check(self._L.bh_tree_regrid(self._h), "regrid")

    LIMITERS = {"minmod": 0, "vanleer": 1, "mc": 2}

    def set_limiter(self, name: str):
        """reconfp's slope limiter in the step: "minmod", "vanleer" (the
        reference's, the default) or "mc" (src/fluid/recon.hpp:16-34)."""
        p = _Params()
        self._check(self._L.bh_tree_get_params(self._h, ctypes.byref(p)), "get_params")
        p.limiter = self.LIMITERS[name]
        self._check(self._L.bh_tree_set_params(self._h, ctypes.byref(p)), "set_params")

    @property
    def limiter(self) -> str:
        p = _Params()
        self._check(self._L.bh_tree_get_params(self._h, ctypes.byref(p)), "get_params")
        return {v: k for k, v in self.LIMITERS.items()}[p.limiter]

    def set_blocks(self, meta: np.ndarray):
        """Load a caller-owned block table (``meta[nb][14]`` as ``blocks()``
        returns it, quilt order); every zone is reset to the floor state."""
        meta = np.ascontiguousarray(meta, dtype=np.int32)
        if meta.ndim != 2 or meta.shape[1] != 14:
            raise ValueError("meta must be [nb, 14]")
        self._check(self._L.bh_tree_set_blocks(self._h, int(meta.shape[0]), _pi(meta)), "set_blocks")

    def floors(self):
        """``patch::floors`` on every patch (grid.cpp:331-358)."""
        self._check(self._L.bh_tree_floors(self._h), "floors")

    def prepare(self):
        """Build the device copy (geometry tables, ghost plan, state upload)."""
        self._check(self._L.bh_tree_prepare(self._h), "prepare")

    def profile(self, enable: bool = True, only: str = ""):
        """Per-kernel HIP-event timing of subsequent launches (of the kernels
        named ``only``, or of all)."""
        self._check(self._L.bh_profile_only(self._h, only.encode()), "profile")
        self._check(self._L.bh_profile(self._h, 1 if enable else 0), "profile")

    def profile_report(self):
        """{kernel: (launches, total_ms)} accumulated since profiling was enabled,
        plus ``queue:<pass>: (cells, 0.0)``, the cells each con2prim pass queued
        for the Brent fallback over the same period."""
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(self._L.bh_profile_report(self._h, buf, 1 << 16), "profile_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split()
            out[name] = (int(n), float(ms))
        return out

    DIAGNOSTICS = ("mass", "energy", "L_pol", "Edot", "Phi_B", "thermE", "Hscal")

    def diagnostics(self) -> dict:
        """The ``analysis::analyser`` quantities that read only the evolved
        state (src/fluid/data/quantitatives.cpp), reduced on the GPU over the
        leaf active zones: mass, energy, L_pol, Edot and Phi_B at the
        horizon, thermE, Hscal at 10 M."""
        out = np.zeros(7)
        self._check(self._L.bh_tree_diagnostics(self._h, _p(out)), "diagnostics")
        return dict(zip(self.DIAGNOSTICS, out.tolist()))

    def write_dump(self, path, time: float = 0.0, step: int = 0):
        """Write W of every leaf's active zones to a dump file (format in
        include/bhfluid.h; read it back with ``read_dump``)."""
        self._check(self._L.bh_tree_write_dump(self._h, str(path).encode(), float(time), int(step)), "write_dump")

    def maxdivB_device(self):
        """constrans::maxdivB of every block on the device: (per-block array,
        max over the leaves)."""
        per = np.zeros(self.nblocks)
        mx = ctypes.c_double(0.0)
        self._check(self._L.bh_tree_maxdivB_device(self._h, per.ctypes.data_as(_dp), ctypes.byref(mx)),
                    "maxdivB_device")
        return per, mx.value

    def maxdivB(self, b: int) -> float:
        """``constrans::maxdivB`` of block b (grid/transport.cpp:77-89)."""
        out = ctypes.c_double(0.0)
        self._check(self._L.bh_tree_maxdivB(self._h, int(b), ctypes.byref(out)), "maxdivB")
        return out.value


class init:
    """``torus::init`` (src/fluid/initial.hpp:17-22)."""

    @staticmethod
    def fm_init(tree: AmrTree) -> bool:
        """Fishbone-Moncrief torus on active zones.  Returns False where the
        reference prints an error and returns without a torus."""
        rc = tree._L.bh_tree_fm_init(tree._h)
        if rc == 2:
            return False
        tree._check(rc, "fm_init")
        return True

    @staticmethod
    def B_pot_init(tree: AmrTree, rho_max: float):
        """Face fields from the curl of A_phi on every patch, in quilt order."""
        tree._check(tree._L.bh_tree_B_pot_init(tree._h, float(rho_max)), "B_pot_init")


def torus_setup(tree: AmrTree, regrid: bool = False, device: bool = False) -> float:
    """The initialisation sequence of src/fluid/main.cpp:98-168: torus, global
    rho_max, magnetic field, conserved variables, optional regrid.  Returns
    rho_max.  ``device=True`` runs all of it on the GPU
    (``AmrTree.init_torus_device`` and ``regrid_device``) for a fresh tree."""
    if device:
        rho_max = tree.init_torus_device()
        if regrid:
            tree.regrid_device()
        return rho_max
    init.fm_init(tree)  # main.cpp:98 ignores a torus that could not be placed
    rho_max = tree.rho_max()
    if rho_max < 1e-14:
        raise BhError("rho_max is zero after the torus initialisation")
    init.B_pot_init(tree, rho_max)
    tree.ptoc_interior()
    if regrid:
        tree.regrid()
    return rho_max


class RK2Integrator:
    """``integ::rk2integrator`` (src/fluid/rk2.hpp:19-30)."""

    def __init__(self, cfl: float = 0.4):
        self.cfl = float(cfl)

    def step(self, tree: AmrTree) -> float:
        """One RK2 step on the GPU: dtcomp, two stages, floors.  Returns dt."""
        dt = ctypes.c_double(0.0)
        tree._check(tree._L.bh_rk2_step(tree._h, self.cfl, ctypes.byref(dt)), "step")
        return dt.value

    @staticmethod
    def dtcomp(tree: AmrTree, cfl: float) -> float:
        dt = ctypes.c_double(0.0)
        tree._check(tree._L.bh_rk2_dtcomp(tree._h, float(cfl), ctypes.byref(dt)), "dtcomp")
        return dt.value

    def rkstg(self, tree: AmrTree, dt: float, stg: int):
        tree._check(tree._L.bh_rk2_rkstg(tree._h, float(dt), int(stg)), "rkstg")

    def advance(self, tree: AmrTree, nsteps: Optional[int] = None, t_end: Optional[float] = None,
                t: float = 0.0) -> Tuple[float, int]:
        """The time loop of src/fluid/main.cpp:169-173: step while t < t_end
        and fewer than nsteps steps were taken.  Returns (t, steps)."""
        if nsteps is None and t_end is None:
            raise ValueError("give nsteps or t_end")
        tt = ctypes.c_double(float(t))
        n = ctypes.c_int(0)
        rc = tree._L.bh_rk2_advance(tree._h, self.cfl, int(nsteps if nsteps is not None else 2 ** 31 - 1),
                                    float(t_end if t_end is not None else np.inf), ctypes.byref(tt), ctypes.byref(n))
        try:
            tree._check(rc, "advance")
        except DomainError as e:
            # the loop stopped after the step that faulted (where the
            # reference's throws); that step is counted
            e.t, e.steps = tt.value, n.value
            raise
        return tt.value, n.value


def run_steps(tree: AmrTree, nsteps: int, cfl: float = 0.4) -> np.ndarray:
    """nsteps RK2 steps enqueued back to back, one synchronisation at the end;
    returns each step's dt.  Raises DomainError after the run if any step saw
    rho <= 0."""
    dts = np.zeros(max(int(nsteps), 1))
    tree._check(tree._L.bh_rk2_steps(tree._h, float(cfl), int(nsteps), _p(dts)), "steps")
    return dts[:nsteps]


DUMP_MAGIC = b"BHFLUIDW"


def read_dump(path) -> dict:
    """Read a dump written by ``AmrTree.write_dump``: header fields, the leaf
    table (level, ijk, lo, hi) and W as a read-only array [leaf][8][16][16][16]
    (rho, eps, v0, v1, v2, B0, B1, B2; indices i, j, k)."""
    raw = np.memmap(path, dtype=np.uint8, mode="r")
    if bytes(raw[:8]) != DUMP_MAGIC:
        raise BhError(f"{path}: not a bhfluid dump")
    version, hbytes = np.frombuffer(raw[8:16], "<u4")
    if version != 1:
        raise BhError(f"{path}: dump version {version} not supported")
    nl, nvar, nz, _ = np.frombuffer(raw[16:32], "<i4")
    M, a, Q, gm, time = np.frombuffer(raw[32:72], "<f8")
    step = int(np.frombuffer(raw[72:80], "<i8")[0])
    tab = raw[hbytes:hbytes + 64 * nl].reshape(nl, 64)
    ints = np.frombuffer(tab[:, :16].tobytes(), "<i4").reshape(nl, 4)
    box = np.frombuffer(tab[:, 16:].tobytes(), "<f8").reshape(nl, 6)
    off = int(hbytes) + 64 * int(nl)
    W = np.frombuffer(raw, dtype="<f8", count=int(nl) * int(nvar) * int(nz), offset=off)
    return {"version": int(version), "M": float(M), "a": float(a), "Q": float(Q), "gamma": float(gm),
            "time": float(time), "step": step, "level": ints[:, 0], "ijk": ints[:, 1:4],
            "lo": box[:, :3], "hi": box[:, 3:], "W": W.reshape(int(nl), int(nvar), BLOCK, BLOCK, BLOCK)}


def device_sync():
    if lib().bh_device_sync() != 0:
        raise BhError("hipDeviceSynchronize failed")


def copy_probe(nbytes: int, reps: int = 5) -> float:
    """Milliseconds per plain device copy of nbytes (HBM read + write)."""
    ms = ctypes.c_double(0.0)
    if lib().bh_copy_probe(int(nbytes), int(reps), ctypes.byref(ms)) != 0:
        raise BhError(lib().bh_last_error(None).decode())
    return ms.value


def device_info() -> str:
    buf = ctypes.create_string_buffer(256)
    lib().bh_device_info(buf, 256)
    return buf.value.decode()
