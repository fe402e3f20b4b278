"""Spherical Kerr-Schild engine: runs of the GRMHD fluid step on
logarithmic-r x theta x phi block grids, advanced on the GPU in shared
launches (C ABI ``include/bhsph.h`` in ``libbhfluid.so``).

=========================  ==========================================================
bhfluid.sph                reference
=========================  ==========================================================
``Engine``                 the storage of ``grid::amrtree`` (src/fluid/grid.hpp:109-144)
                           for any number of runs
``Engine.add_run``         ``amrtree`` construction + ``regrid`` (grid.cpp:177-206,
                           739-835), with a caller-owned block table
``Run.init_torus``         ``torus::init::fm_init`` + ``B_pot_init`` (initconds.cpp:121-668,
                           main.cpp:94-159)
``Engine.step``            ``rk2integrator::step`` (integrator.cpp:262-275) of every listed run
``Run.diagnostics``        conserved totals and ``constrans::maxdivB`` (transport.cpp:77-89)
=========================  ==========================================================

The physics is the reference's step made self-consistent (see
``csrc/sph_physics.hpp``); the configurations of the benchmark are built by
``c1_run`` .. ``amr_blocks`` below.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

from . import lib as _base_lib, BhError

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_d = ctypes.c_double
_i = ctypes.c_int
_v = ctypes.c_void_p
_ll = ctypes.c_longlong

BX = BY = 16
G = 2


_PHYS_WORDS = ("a", "Q", "gamma", "cfl", "rho_floor", "eps_floor", "sigma_max", "gamma_max")
# Riemann solvers (include/bhsph.h BHS_HLL, BHS_HLLD)
RIEMANN = {"hll": 0, "hlld": 1}


class Physics(ctypes.Structure):
    _fields_ = [(n, _d) for n in _PHYS_WORDS] + [("riemann", _i), ("reserved", _i)]


def riemann_code(physics: dict) -> int:
    """The solver code of a physics dict ("hlld" when absent)."""
    name = physics.get("riemann", "hlld")
    if name not in RIEMANN:
        raise ValueError(f"riemann must be one of {sorted(RIEMANN)}, got {name!r}")
    return RIEMANN[name]


class Grid(ctypes.Structure):
    _fields_ = [("nr", _i), ("nth", _i), ("nph", _i), ("nlevels", _i), ("r_in", _d), ("r_out", _d)]


class Torus(ctypes.Structure):
    _fields_ = [(n, _d) for n in ("r_in", "r_max", "kappa", "beta", "a_cut", "perturbation")] + \
               [(n, ctypes.c_ulonglong) for n in ("pcg_state_hi", "pcg_state_lo", "pcg_inc_hi", "pcg_inc_lo")]


_sigs_done = False


def lib():
    global _sigs_done
    L = _base_lib()
    if not _sigs_done:
        def sig(name, res, args):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        sig("bhs_create", _i, [_i, _ll, ctypes.POINTER(_v)])
        sig("bhs_destroy", None, [_v])
        sig("bhs_last_error", ctypes.c_char_p, [_v])
        sig("bhs_block_bytes", _ll, [_i])
        sig("bhs_run_add", _i, [_v, ctypes.POINTER(Physics), ctypes.POINTER(Grid), _i, _ip, _ip])
        sig("bhs_run_remove", _i, [_v, _i])
        sig("bhs_run_nblocks", _i, [_v, _i, _ip])
        sig("bhs_run_init_torus", _i, [_v, _i, ctypes.POINTER(Torus), _dp])
        sig("bhs_run_get_state", _i, [_v, _i, _dp, _dp, _dp])
        sig("bhs_run_set_state", _i, [_v, _i, _dp, _dp, _dp])
        sig("bhs_steps", _i, [_v, _i, _ip, _i, _dp])
        sig("bhs_c2p", _i, [_v, _i])
        sig("bhs_run_diagnostics", _i, [_v, _i, _dp])
        sig("bhs_run_analysis", _i, [_v, _i, _d, _dp])
        sig("bhs_set_precision", _i, [_v, _i])
        sig("bhs_set_schedule", _i, [_v, _i, _i])
        sig("bhs_use_graph", _i, [_v, _i])
        sig("bhs_sync", _i, [_v])
        sig("bhs_steps_timed", _i, [_v, _i, _ip, _i, _dp])
        sig("bhs_set_variants", _i, [_v, _i, _i, _i])
        sig("bhs_variant_info", _i, [_i, _ip, _ip])
        sig("bhs_set_xcd_remap", _i, [_v, _i])
        sig("bhs_set_face_variant", _i, [_v, _i])
        sig("bhs_set_barrier_spin_limit", _i, [_v, ctypes.c_uint])
        sig("bhs_run_diagnostics_static", _i, [_v, _i, _dp])
        sig("bhs_profile", _i, [_v, _i, ctypes.c_char_p])
        sig("bhs_profile_report", _i, [_v, ctypes.c_char_p, _i])
        _sigs_done = True
    return L


def layout(bz: int) -> Dict[str, int]:
    sz = bz + 2 * G
    ncell = 20 * 20 * sz
    return dict(bz=bz, sz=sz, ncell=ncell, nint=16 * 16 * bz, nf=max(21 * 20 * sz, 20 * 20 * (sz + 1)))


def pcg_words(seed: int):
    """numpy PCG64(seed)'s 128-bit state and increment as (hi, lo, hi, lo)."""
    st = np.random.PCG64(seed).state["state"]
    m = (1 << 64) - 1
    return (st["state"] >> 64, st["state"] & m, st["inc"] >> 64, st["inc"] & m)


def _p(a):
    return a.ctypes.data_as(_dp)


class Engine:
    """Device storage for up to ``max_blocks`` blocks of 16 x 16 x bz zones
    (bz = 16, or 1 for axisymmetric runs) and the runs that use them."""

    def __init__(self, bz: int, max_blocks: int):
        self._L = lib()
        self.bz = bz
        self.lay = layout(bz)
        h = _v()
        rc = self._L.bhs_create(bz, int(max_blocks), ctypes.byref(h))
        if rc:
            raise BhError(f"bhs_create: {self._L.bhs_last_error(None).decode()} (code {rc})")
        self._h = h
        self.runs: Dict[int, "Run"] = {}

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.bhs_destroy(self._h)
            self._h = None

    def _check(self, rc, what):
        if rc:
            raise BhError(f"{what}: {self._L.bhs_last_error(self._h).decode()} (code {rc})")

    @staticmethod
    def block_bytes(bz: int) -> int:
        return int(lib().bhs_block_bytes(bz))

    def add_run(self, physics: dict, grid: dict, blocks) -> "Run":
        ph = Physics(*[float(physics[n]) for n in _PHYS_WORDS], riemann_code(physics), 0)
        g = Grid(int(grid["nr"]), int(grid["nth"]), int(grid.get("nph", 1)), int(grid.get("nlevels", 1)),
                 float(grid["r_in"]), float(grid["r_out"]))
        b = np.ascontiguousarray(blocks, dtype=np.int32).reshape(-1, 4)
        rid = ctypes.c_int(-1)
        self._check(self._L.bhs_run_add(self._h, ctypes.byref(ph), ctypes.byref(g), len(b),
                                        b.ctypes.data_as(_ip), ctypes.byref(rid)), "run_add")
        r = Run(self, rid.value, physics, grid, b)
        self.runs[rid.value] = r
        return r

    def remove_run(self, run: "Run"):
        self._check(self._L.bhs_run_remove(self._h, run.id), "run_remove")
        self.runs.pop(run.id, None)
        run.id = -1

    def set_precision(self, mixed):
        """False/0: FP64; True/1: mixed (FP32 reconstruction, Riemann fluxes
        and EMFs, rounded as the CPU rounds them); 2: mixed with the FP32
        divisions and square roots of the hardware reciprocal and square-root
        instructions (faster; judged by its accuracy against FP64)."""
        mode = 2 if mixed == 2 else int(bool(mixed))
        self._check(self._L.bhs_set_precision(self._h, mode), "set_precision")

    def set_schedule(self, schedule: int, wg_per_cu: int = 0):
        """0 (or False): one kernel per sub-phase; 1 (or True): one
        persistent queue-driven kernel per stage, a block's phase B in one
        workgroup; 2: sub-phase kernels with the fluxes and the hydro update
        fused per block (slab); 3: schedule 0 with stage 1's con2prim inside
        stage 0's update; 4: one launch per stage whose phases are work
        queues over all blocks, separated by grid barriers."""
        self._check(self._L.bhs_set_schedule(self._h, int(schedule), int(wg_per_cu)), "set_schedule")

    def set_variants(self, flux: int = 0, update: int = 0, c2p: int = 0):
        self._check(self._L.bhs_set_variants(self._h, int(flux), int(update), int(c2p)), "set_variants")

    def set_face_variant(self, face: int = 2):
        self._check(self._L.bhs_set_face_variant(self._h, int(face)), "set_face_variant")

    def set_barrier_spin_limit(self, polls: int = 0):
        """Polls of a grid-barrier wait (schedules 1 and 4); 0 restores the
        default.  A small value makes the barriers time out (error path)."""
        self._check(self._L.bhs_set_barrier_spin_limit(self._h, int(polls)), "set_barrier_spin_limit")

    def set_xcd_remap(self, enable: bool = True):
        self._check(self._L.bhs_set_xcd_remap(self._h, int(bool(enable))), "set_xcd_remap")

    @staticmethod
    def variant_info(v: int):
        tb, mw = ctypes.c_int(), ctypes.c_int()
        if lib().bhs_variant_info(int(v), ctypes.byref(tb), ctypes.byref(mw)):
            raise BhError(f"no launch variant {v}")
        return tb.value, mw.value

    def use_graph(self, enable: bool = True):
        self._check(self._L.bhs_use_graph(self._h, int(bool(enable))), "use_graph")

    def step(self, runs: Sequence["Run"], nsteps: int = 1, want_dt: bool = True) -> Optional[np.ndarray]:
        """``nsteps`` RK2 steps of every listed run in shared launches; returns
        dt[nsteps][nruns] (each run's own CFL step) when ``want_dt``."""
        ids = np.array([r.id for r in runs], dtype=np.int32)
        dt = np.zeros((nsteps, len(ids))) if want_dt else None
        self._check(self._L.bhs_steps(self._h, len(ids), ids.ctypes.data_as(_ip), int(nsteps),
                                      _p(dt) if want_dt else None), "steps")
        return dt

    def sync(self):
        self._check(self._L.bhs_sync(self._h), "sync")

    def step_timed(self, runs: Sequence["Run"], nsteps: int) -> np.ndarray:
        """``nsteps`` steps (graph replays) with a HIP event between
        consecutive steps; returns each step's device time in ms."""
        ids = np.array([r.id for r in runs], dtype=np.int32)
        ms = np.zeros(nsteps)
        self._check(self._L.bhs_steps_timed(self._h, len(ids), ids.ctypes.data_as(_ip), int(nsteps), _p(ms)),
                    "steps_timed")
        return ms

    def profile(self, enable: bool = True, only: str = ""):
        self._check(self._L.bhs_profile(self._h, int(bool(enable)), only.encode()), "profile")

    def profile_report(self) -> Dict[str, dict]:
        buf = ctypes.create_string_buffer(1 << 16)
        n = self._L.bhs_profile_report(self._h, buf, len(buf))
        if n < 0:
            self._check(n, "profile_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, tot = line.split()
            out[name] = dict(launches=int(cnt), total_ms=float(tot), avg_ms=float(tot) / max(1, int(cnt)))
        return out


class Run:
    """One run (a torus on its own grid) inside an ``Engine``."""

    def __init__(self, eng: Engine, rid: int, physics: dict, grid: dict, blocks: np.ndarray):
        self.eng = eng
        self.id = rid
        self.physics = dict(physics)
        self.grid = dict(grid)
        self.blocks = blocks
        self.nb = len(blocks)

    def init_torus(self, torus: dict, seed: int = 42) -> np.ndarray:
        w = pcg_words(seed)
        t = Torus(float(torus["r_in"]), float(torus["r_max"]), float(torus["kappa"]), float(torus["beta"]),
                  float(torus.get("a_cut", 0.2)), float(torus.get("perturbation", 0.04)), *[int(x) for x in w])
        info = np.zeros(4)
        self.eng._check(self.eng._L.bhs_run_init_torus(self.eng._h, self.id, ctypes.byref(t), _p(info)),
                        "run_init_torus")
        return info

    def state(self):
        L = self.eng.lay
        U = np.zeros((self.nb, 8, L["ncell"]))
        W = np.zeros((self.nb, 8, L["ncell"]))
        Bf = np.zeros((self.nb, 3, L["nf"]))
        self.eng._check(self.eng._L.bhs_run_get_state(self.eng._h, self.id, _p(U), _p(W), _p(Bf)), "get_state")
        return U, W, Bf

    def set_state(self, U=None, W=None, Bf=None):
        arrs = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (U, W, Bf)]
        self.eng._check(self.eng._L.bhs_run_set_state(self.eng._h, self.id,
                                                      *[None if x is None else _p(x) for x in arrs]), "set_state")

    def c2p(self):
        self.eng._check(self.eng._L.bhs_c2p(self.eng._h, self.id), "c2p")

    def diagnostics(self) -> dict:
        out = np.zeros(8)
        self.eng._check(self.eng._L.bhs_run_diagnostics(self.eng._h, self.id, _p(out)), "diagnostics")
        return dict(mass=out[0], energy=out[1], divb_max=out[2], divb_scale=out[3], c2p_fail=int(out[4]),
                    floors=int(out[5]), mdot=out[6], phi_b=out[7])

    def analysis(self, rho_min: float = 0.0) -> dict:
        """The analyser quantities that read derived primitives
        (src/fluid/data/quantitatives.cpp:109-193) in covariant form; the
        stress ratios over zones with rho >= rho_min (include/bhsph.h)."""
        d = np.zeros(10)
        self.eng._check(self.eng._L.bhs_run_analysis(self.eng._h, self.id, float(rho_min), _p(d)), "analysis")
        return dict(thermE=d[0], magE=d[1], Ldot=d[2], L_BZ=d[3], eta_BZ=d[4], phi_BH=d[5], P_beta=d[6], alpha_SS=d[7], maxwell_reynolds=d[8], mdot=d[9])

    def active(self, W: np.ndarray) -> np.ndarray:
        """Active zones of a W (or U) array: [nb][8][bz][16][16] (phi, theta, r)."""
        bz = self.eng.bz
        return W.reshape(self.nb, W.shape[1], bz + 4, 20, 20)[:, :, 2:2 + bz, 2:18, 2:18]


# ----------------------------------------------------------------- presets

# Physics of the torus runs: a = 0.9375, Gamma = 13/9, CFL 0.4, the
# reference's floor constants (src/fluid/grid.hpp:140-143)
PHYSICS = dict(a=0.9375, Q=0.0, gamma=13.0 / 9.0, cfl=0.4, rho_floor=1e-7, eps_floor=1e-5, sigma_max=100.0,
               gamma_max=50.0, riemann="hlld")
TORUS = dict(r_in=6.0, r_max=12.0, kappa=1e-3, beta=100.0, a_cut=0.2, perturbation=0.04)
R_IN, R_OUT = 1.2, 40.0


def uniform_blocks(grid: dict, bz: int) -> np.ndarray:
    nbr, nbt = grid["nr"] // BX, grid["nth"] // BY
    nbp = 1 if bz == 1 else grid["nph"] // bz
    return np.array([(0, i, j, k) for k in range(nbp) for j in range(nbt) for i in range(nbr)], dtype=np.int32)


def fm_density(a, gamma, r_in, r_max, kappa, r, th):
    """Unnormalised Fishbone-Moncrief density at (r, theta) (numpy; used
    only to flag blocks for refinement)."""
    sr = math.sqrt(r_max)
    r32 = r_max * sr
    l = (r32 + a) * (r_max ** 2 - 2 * a * sr + a * a) / (r32 * (r32 - 3 * sr + 2 * a))

    def lnh_raw(r, s, c):
        Sig = r * r + a * a * c * c
        Del = r * r - 2 * r + a * a
        A = (r * r + a * a) ** 2 - Del * a * a * s * s
        q = np.sqrt(1 + 4 * l * l * Sig * Sig * Del / (A * A * s * s))
        return 0.5 * np.log((1 + q) * A / (Sig * Del)) - 0.5 * q - 2 * a * r * l / A

    with np.errstate(invalid="ignore", divide="ignore"):
        lnh = lnh_raw(r, np.sin(th), np.cos(th)) - lnh_raw(r_in, 1.0, 0.0)
        inside = (r >= r_in) & (lnh > 0)
        h = np.exp(np.where(inside, lnh, 0.0))
        return np.where(inside, ((h - 1) * (gamma - 1) / (kappa * gamma)) ** (1 / (gamma - 1)), 0.0)


def amr_blocks(grid: dict, nlevels: int, physics=PHYSICS, torus=TORUS, thresh: float = 0.1) -> np.ndarray:
    """Block table of a run refined where the torus density exceeds
    ``thresh`` x rho_max (the criterion evaluated at each level's zone
    centres), then 2:1 balanced over faces, edges and corners (with the
    axis mirror and periodic phi); blocks ordered by level, then phi,
    theta, r."""
    a, gam = physics["a"], physics["gamma"]
    rmax_rho = fm_density(a, gam, torus["r_in"], torus["r_max"], torus["kappa"], np.array([torus["r_max"]]),
                          np.array([math.pi / 2]))[0]
    lnr0, lnr1 = math.log(grid["r_in"]), math.log(grid["r_out"])

    def nb_at(l):
        f = 1 << l
        return grid["nr"] * f // BX, grid["nth"] * f // BY, grid["nph"] * f // BX

    def flag(l, bi, bj, bk):
        f = 1 << l
        nr, nth = grid["nr"] * f, grid["nth"] * f
        i = np.arange(bi * BX, bi * BX + BX) + 0.5
        j = np.arange(bj * BY, bj * BY + BY) + 0.5
        r = np.exp(lnr0 + i * (lnr1 - lnr0) / nr)
        th = j * math.pi / nth
        R, T = np.meshgrid(r, th)
        rho = fm_density(a, gam, torus["r_in"], torus["r_max"], torus["kappa"], R, T)
        return rho.max() > thresh * rmax_rho

    leaves = set()
    nbr, nbt, nbp = nb_at(0)
    for k in range(nbp):
        for j in range(nbt):
            for i in range(nbr):
                leaves.add((0, i, j, k))
    for l in range(nlevels - 1):
        for (ll, i, j, k) in sorted(leaves):
            if ll == l and flag(l, i, j, k):
                leaves.discard((ll, i, j, k))
                for c in range(8):
                    leaves.add((l + 1, 2 * i + (c & 1), 2 * j + ((c >> 1) & 1), 2 * k + (c >> 2)))

    def cover(l, i, j, k):
        """the leaf holding block position (l; i, j, k), searching coarser levels"""
        for m in range(l, -1, -1):
            s = l - m
            key = (m, i >> s, j >> s, k >> s)
            if key in leaves:
                return key
        return None

    changed = True
    while changed:
        changed = False
        for (l, i, j, k) in sorted(leaves):
            if (l, i, j, k) not in leaves:
                continue
            nr_, nt_, np_ = nb_at(l)
            for dk in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    for di in (-1, 0, 1):
                        ii, jj, kk = i + di, j + dj, (k + dk) % np_
                        if ii < 0 or ii >= nr_:
                            continue
                        if jj < 0:
                            jj = -1 - jj
                        if jj >= nt_:
                            jj = 2 * nt_ - 1 - jj
                        c = cover(l, ii, jj, kk)
                        if c is not None and c[0] < l - 1:
                            leaves.discard(c)
                            m, ci, cj, ck = c
                            for q in range(8):
                                leaves.add((m + 1, 2 * ci + (q & 1), 2 * cj + ((q >> 1) & 1), 2 * ck + (q >> 2)))
                            changed = True
    return np.array(sorted(leaves, key=lambda t: (t[0], t[3], t[2], t[1])), dtype=np.int32)


def c1_grid():
    """The 2D axisymmetric torus grid (C1): 128 (r, log 1.2-40) x 128 (theta)."""
    return dict(nr=128, nth=128, nph=1, nlevels=1, r_in=R_IN, r_out=R_OUT)


def c2_grid():
    """The uniform 3D torus grid (C2): 256 (r) x 128 (theta) x 128 (phi)."""
    return dict(nr=256, nth=128, nph=128, nlevels=1, r_in=R_IN, r_out=R_OUT)


def c3_grid():
    """The 3-level AMR torus grid (C3): 16^3 blocks refined where rho > 0.1
    rho_max; the level-0 grid is 128 x 64 x 64 zones (1976 blocks: 56, 384
    and 1536 per level, 8.09 M zones)."""
    return dict(nr=128, nth=64, nph=64, nlevels=3, r_in=R_IN, r_out=R_OUT)


def zone_volumes(grid: dict, blocks: np.ndarray, bz: int = 16) -> np.ndarray:
    """Coordinate volume dr dtheta dphi of every active zone, shaped like
    ``Run.active(W)[:, q]``: [nb][bz][16][16] (phi, theta, r)."""
    lnr0, lnr1 = math.log(grid["r_in"]), math.log(grid["r_out"])
    out = np.empty((len(blocks), bz, BY, BX))
    for b, (l, bi, bj, bk) in enumerate(blocks):
        f = 1 << int(l)
        nr, nth = grid["nr"] * f, grid["nth"] * f
        nph = 1 if bz == 1 else grid["nph"] * f
        i = np.arange(bi * BX, bi * BX + BX)
        dr = np.exp(lnr0 + (i + 1) * (lnr1 - lnr0) / nr) - np.exp(lnr0 + i * (lnr1 - lnr0) / nr)
        out[b] = dr[None, None, :] * (math.pi / nth) * (2 * math.pi / nph)
    return out


def zone_count(blocks: np.ndarray, bz: int) -> int:
    return int(len(blocks)) * BX * BY * bz
