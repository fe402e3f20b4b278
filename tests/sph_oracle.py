"""ctypes binding of the spherical-engine oracle (oracle/sphoracle.cpp in
liboracle.so): test infrastructure, the checker of the spherical Kerr-Schild
engine.  Also the block-table builder the tests and the benchmark share with
the product's Python layer (bhfluid.sph)."""
import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u64p = ctypes.POINTER(ctypes.c_ulonglong)
_llp = ctypes.POINTER(ctypes.c_longlong)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(str(ROOT / "oracle" / "liboracle.so"))
        L.os_run_new.restype = ctypes.c_void_p
        L.os_run_new.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip]
        L.os_run_free.argtypes = [ctypes.c_void_p]
        L.os_tables.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.os_init_torus.restype = ctypes.c_int
        L.os_init_torus.argtypes = [ctypes.c_void_p, _dp, _u64p, _dp]
        L.os_get_state.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.os_set_state.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.os_step.restype = ctypes.c_double
        L.os_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.os_c2p.restype = ctypes.c_double
        L.os_c2p.argtypes = [ctypes.c_void_p]
        L.os_counters.argtypes = [ctypes.c_void_p, _llp]
        L.os_diag.argtypes = [ctypes.c_void_p, _dp]
        L.os_analysis.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        L.os_analysis_point.argtypes = [_dp, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_int, _dp]
        L.os_ptoc.argtypes = [_dp, ctypes.c_double, _dp, _dp]
        L.os_c2p_point.restype = ctypes.c_int
        L.os_c2p_point.argtypes = [_dp, ctypes.c_double, _dp, _dp]
        L.os_flux_point.argtypes = [_dp, ctypes.c_double, _dp, ctypes.c_int, _dp, _dp, _dp]
        L.os_hll_point.argtypes = [_dp, ctypes.c_double, _dp, _dp, ctypes.c_int, _dp]
        L.os_hlld_point.restype = ctypes.c_int
        L.os_hlld_point.argtypes = [_dp, ctypes.c_double, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.os_source_point.argtypes = [_dp, ctypes.c_double, _dp, _dp]
        L.os_metric_point.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.os_pcg64.restype = ctypes.c_double
        L.os_pcg64.argtypes = [ctypes.c_ulonglong] * 5
        L.os_torus_point.argtypes = [ctypes.c_double] * 7 + [_dp]
        _lib = L
    return _lib


def P(a):
    return a.ctypes.data_as(_dp)


NGC = 17 + 38
NMET = 17


def lay(bz):
    sz = bz + 4
    ncell = 20 * 20 * sz
    nint = 16 * 16 * bz
    nf = max(21 * 20 * sz, 20 * 20 * (sz + 1))
    return dict(bz=bz, sz=sz, ncell=ncell, nint=nint, nf=nf)


def pcg_words(seed):
    """numpy PCG64(seed)'s 128-bit state and increment as (hi, lo, hi, lo)."""
    st = np.random.PCG64(seed).state["state"]
    m = (1 << 64) - 1
    s, i = st["state"], st["inc"]
    return np.array([s >> 64, s & m, i >> 64, i & m], dtype=np.uint64)


class OracleRun:
    """One run of the spherical engine restated on the CPU."""

    def __init__(self, cfg, blocks):
        self.cfg = cfg
        self.L = lay(cfg["bz"])
        self.blocks = np.ascontiguousarray(blocks, dtype=np.int32)
        # phys[8]: the Riemann solver, 0 HLL, 1 HLLD (the default)
        phys = np.array([cfg["a"], cfg.get("Q", 0.0), cfg["gamma"], cfg["cfl"], cfg["rho_fl"], cfg["eps_fl"],
                         cfg["sigma_max"], cfg["w_max"], float(cfg.get("riemann", 1))], dtype=np.float64)
        self.nb = len(self.blocks)
        self._h = lib().os_run_new(cfg["bz"], P(phys), cfg["nr"], cfg["nth"], cfg["nph"], cfg["r_in"], cfg["r_out"],
                                   cfg["nlev"], self.nb, self.blocks.ctypes.data_as(_ip))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().os_run_free(self._h)
            self._h = None

    def init_torus(self, tor, seed):
        t = np.array([tor["r_in"], tor["r_max"], tor["kappa"], tor["beta"], tor["acut"], tor["pert"]], np.float64)
        w = pcg_words(seed)
        info = np.zeros(4)
        rc = lib().os_init_torus(self._h, P(t), w.ctypes.data_as(_u64p), P(info))
        if rc:
            raise RuntimeError(f"os_init_torus rc={rc}")
        return info

    def state(self):
        L = self.L
        U = np.zeros((self.nb, 8, L["ncell"]))
        W = np.zeros((self.nb, 8, L["ncell"]))
        Bf = np.zeros((self.nb, 3, L["nf"]))
        lib().os_get_state(self._h, P(U), P(W), P(Bf))
        return U, W, Bf

    def set_state(self, U, W, Bf):
        lib().os_set_state(self._h, P(np.ascontiguousarray(U)), P(np.ascontiguousarray(W)),
                           P(np.ascontiguousarray(Bf)))

    def step(self, mixed=False):
        return lib().os_step(self._h, int(mixed))

    def c2p(self):
        return lib().os_c2p(self._h)

    def counters(self):
        """con2prim failures, floored zones, HLLD faces solved, HLLD faces
        that fell back to HLL, and those by reason: no convergence,
        fast-wave states, fan ordering, non-finite flux."""
        c = np.zeros(8, np.int64)
        lib().os_counters(self._h, c.ctypes.data_as(_llp))
        return c

    def diag(self):
        d = np.zeros(6)
        lib().os_diag(self._h, P(d))
        return dict(mass=d[0], energy=d[1], divb_max=d[2], divb_scale=d[3], mdot=d[4], phi_b=d[5])

    def analysis(self, rho_min=0.0):
        d = np.zeros(10)
        lib().os_analysis(self._h, float(rho_min), P(d))
        return dict(thermE=d[0], magE=d[1], Ldot=d[2], L_BZ=d[3], eta_BZ=d[4], phi_BH=d[5], P_beta=d[6], alpha_SS=d[7], maxwell_reynolds=d[8], mdot=d[9])
