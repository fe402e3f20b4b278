#!/usr/bin/env python3
"""Generate the golden fixtures from the compiled reference.

Runs only where oracle/_ref/libbhref.so was built (the container that holds
the reference sources); no GPU run executes it.  Writes:

* pointwise.npz -- inputs and outputs of the reference's metric::comp,
  conserved::ptoc / ctop and hlld on seeded random points;
* regrid_flatten.json -- for R2 and R3, digests of the block table and
  state after a second regrid that flattens (tests/regrid_cases.py);
* trees.json    -- for each parity configuration, the SHA-256 of the initial
  state (block table and U, W, Bf after src/fluid/main.cpp:94-168) and of the
  state after every phase of every step (dtcomp, rkstg 0, rkstg 1, floors),
  with dt, the reference's exception if one was thrown, and non-finite
  counts, and the reference analyser's state-only diagnostics (mass,
  energy, L_pol, Edot, Phi_B, thermE, Hscal; exact values as float.hex)
  initially and after every completed step.  NaNs are canonicalised before
  hashing.

Usage: python tests/golden/make_golden.py [--flatten-only]
"""
import ctypes
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

import configs  # noqa: E402
from oracle_libs import RefTree, ref_lib  # noqa: E402

SEED = 20261203
_dp = ctypes.POINTER(ctypes.c_double)


def P(a):
    return a.ctypes.data_as(_dp)


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        if a.dtype == np.float64:
            a = a.copy()
            a[np.isnan(a)] = np.nan  # one NaN bit pattern
        h.update(a.tobytes())
    return h.hexdigest()


def state_digest(s):
    return digest(s.U, s.W, s.Bf)


def random_prims(rng, n, r):
    w = np.empty((n, 8))
    w[:, 0] = np.exp(rng.uniform(np.log(1e-8), 0, n))
    w[:, 1] = np.exp(rng.uniform(np.log(1e-7), 0, n))
    w[:, 2] = rng.uniform(-0.3, 0.3, n)
    w[:, 3] = rng.uniform(-0.3, 0.3, n) / r
    w[:, 4] = rng.uniform(-0.3, 0.3, n) / r
    w[:, 5:8] = rng.uniform(-0.3, 0.3, (n, 3))
    return w


def pointwise():
    L = ref_lib()
    rng = np.random.default_rng(SEED)
    out = {}
    # metric::comp at (r, th) for several (a, Q)
    n = 256
    aq = np.array([(a, q) for a in (0.0, 0.9, 0.9375) for q in (0.0, 0.1)])
    sel = rng.integers(0, len(aq), n)
    r = np.exp(rng.uniform(np.log(1.2), np.log(60.0), n))
    th = rng.uniform(1e-3, np.pi - 1e-3, n)
    M = np.zeros((n, 120))
    for i in range(n):
        L.ref_metric(1.0, aq[sel[i], 0], aq[sel[i], 1], r[i], th[i], P(M[i]))
    out.update(metric_aq=aq[sel], metric_r=r, metric_th=th, metric_out=M)
    # ptoc / ctop
    n = 512
    a = rng.choice([0.0, 0.9], n)
    q = rng.choice([0.0, 0.1], n)
    r = np.exp(rng.uniform(np.log(1.5), np.log(60.0), n))
    th = rng.uniform(0.02, np.pi - 0.02, n)
    W = random_prims(rng, n, r)
    U = np.zeros((n, 8))
    W2 = np.zeros((n, 8))
    rc = np.zeros(n, np.int32)
    for i in range(n):
        L.ref_ptoc(1.0, a[i], q[i], configs.GAMMA, P(W[i]), r[i], th[i], P(U[i]), None)
        rc[i] = L.ref_ctop(1.0, a[i], q[i], configs.GAMMA, P(U[i]), r[i], th[i], P(W2[i]), None)
    out.update(pc_a=a, pc_q=q, pc_r=r, pc_th=th, pc_W=W, pc_U=U, ctop_W=W2, ctop_rc=rc)
    # hlld at faces, dim 0..2
    n = 384
    r = np.exp(rng.uniform(np.log(1.5), np.log(60.0), n))
    th = rng.uniform(0.02, np.pi - 0.02, n)
    dim = rng.integers(0, 3, n).astype(np.int32)
    WL, WR = random_prims(rng, n, r), random_prims(rng, n, r)
    UL, UR = np.zeros((n, 8)), np.zeros((n, 8))
    F = np.zeros((n, 8))
    ct = np.zeros(n)
    for i in range(n):
        L.ref_ptoc(1.0, 0.9, 0.1, configs.GAMMA, P(WL[i]), r[i], th[i], P(UL[i]), None)
        L.ref_ptoc(1.0, 0.9, 0.1, configs.GAMMA, P(WR[i]), r[i], th[i], P(UR[i]), None)
        c = ctypes.c_double(0.0)
        L.ref_hlld(1.0, 0.9, 0.1, configs.GAMMA, P(WL[i]), P(WR[i]), P(UL[i]), P(UR[i]), r[i], th[i], int(dim[i]),
                   P(F[i]), ctypes.byref(c))
        ct[i] = c.value
    out.update(hl_r=r, hl_th=th, hl_dim=dim, hl_WL=WL, hl_WR=WR, hl_UL=UL, hl_UR=UR, hl_F=F, hl_ct=ct)
    np.savez_compressed(HERE / "pointwise.npz", **out)


def tree_record(name, cfg, nsteps):
    lo, hi, nqlt, regrid, a, q = cfg
    t = RefTree(lo, hi, nqlt, 1.0, a, q, configs.GAMMA)
    t.L.ref_set_threads(1)  # an exception must not escape an OpenMP region
    t.torus_init(regrid=regrid)
    s = t.state()
    rec = {"config": {"dom_l": lo, "dom_h": hi, "nqlt": nqlt, "regrid": regrid, "a": a, "Q": q,
                      "gamma": configs.GAMMA, "cfl": configs.CFL},
           "nblocks": s.nb, "leaves": int(s.meta[:, 4].sum()),
           "blocks": digest(s.meta, s.edges), "initial": state_digest(s),
           "diag_initial": [float(x).hex() for x in t.diagnostics()], "steps": []}
    for step in range(nsteps):
        st = {}
        dt, rc = t.dtcomp(configs.CFL)
        st["dt"] = dt
        for phase, fn in (("rkstg0", lambda: t.rkstg(dt, 0)), ("rkstg1", lambda: t.rkstg(dt, 1)),
                          ("floors", t.floors)):
            rc = fn()
            if rc:
                st["threw"] = phase
                st["error"] = t.error()
                break
            s = t.state()
            st[phase] = state_digest(s)
        if "threw" not in st:
            U = s.U.reshape(s.nb, 8, 20, 20, 20)[:, :, 2:18, 2:18, 2:18]
            st["nonfinite_active"] = int((~np.isfinite(U)).any(axis=1).sum())
            st["diag"] = [float(x).hex() for x in t.diagnostics()]
        rec["steps"].append(st)
        print(name, step, {k: (v[:12] if isinstance(v, str) else v) for k, v in st.items()}, flush=True)
        if "threw" in st:
            break
    return rec


def flatten_record(name, cfg):
    """Digests of the block table and state after a second regrid that
    flattens (tests/regrid_cases.py), with the counts before and after."""
    from regrid_cases import flat_state
    lo, hi, nqlt, regrid, a, q = cfg
    t = RefTree(lo, hi, nqlt, 1.0, a, q, configs.GAMMA)
    t.L.ref_set_threads(1)
    t.torus_init(regrid=True)
    s = t.state()
    nb1, leaves1 = s.nb, int(s.meta[:, 4].sum())
    s.W[:] = flat_state(s.W.shape)
    t.set_state(s)
    if t.L.ref_ptoc_interior(t.h) or t.L.ref_regrid(t.h):
        raise RuntimeError(t.error())
    s2 = t.state()
    rec = {"config": name, "nblocks_before": nb1, "leaves_before": leaves1, "nblocks": s2.nb,
           "leaves": int(s2.meta[:, 4].sum()), "blocks": digest(s2.meta, s2.edges), "state": state_digest(s2)}
    print(name, rec, flush=True)
    return rec


def main():
    if "--flatten-only" in sys.argv:
        (HERE / "regrid_flatten.json").write_text(
            json.dumps({n: flatten_record(n, getattr(configs, n)) for n in ("R2", "R3")}, indent=1))
        return
    (HERE / "regrid_flatten.json").write_text(
        json.dumps({n: flatten_record(n, getattr(configs, n)) for n in ("R2", "R3")}, indent=1))
    pointwise()
    trees = {}
    for name, cfg, n in (("R0", configs.R0, 6), ("R1", configs.R1, 10), ("R2", configs.R2, 2),
                         ("R3", configs.R3, 2), ("R4", configs.R4, 3)):
        trees[name] = tree_record(name, cfg, n)
    (HERE / "trees.json").write_text(json.dumps(trees, indent=1))


if __name__ == "__main__":
    main()
