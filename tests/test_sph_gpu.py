"""GPU parity of the spherical Kerr-Schild engine against its oracle, through
the C ABI (include/bhsph.h).

The oracle (oracle/sphoracle.cpp) restates the engine's orchestration in
sequential CPU loops and shares only the pointwise physics headers, which
tests/test_sph_oracle.py pins against mathematics.  The bar is the north
star's: relative L-infinity (per variable, max |a - b| / max |b|) at most
1e-10 on primitives after 10 steps, FP64; the device and the oracle perform
the same floating-point operations, so the measured difference is 0.
"""
import numpy as np
import pytest

from bhfluid import sph
from sph_oracle import OracleRun

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _cfg(grid, bz):
    p = sph.PHYSICS
    return dict(bz=bz, a=p["a"], Q=p["Q"], gamma=p["gamma"], cfl=p["cfl"], rho_fl=p["rho_floor"],
                eps_fl=p["eps_floor"], sigma_max=p["sigma_max"], w_max=p["gamma_max"], nr=grid["nr"],
                nth=grid["nth"], nph=grid["nph"], r_in=grid["r_in"], r_out=grid["r_out"], nlev=grid["nlevels"],
                riemann=sph.riemann_code(p))


def _tor():
    t = sph.TORUS
    return dict(r_in=t["r_in"], r_max=t["r_max"], kappa=t["kappa"], beta=t["beta"], acut=t["a_cut"],
                pert=t["perturbation"])


def rel_linf(a, b):
    """max over variables of max|a - b| / max|b| over the zones where both
    are finite; the non-finite masks must agree."""
    assert (np.isfinite(a) == np.isfinite(b)).all(), "non-finite masks differ"
    worst = 0.0
    for v in range(a.shape[1]):
        x, y = a[:, v], b[:, v]
        m = np.isfinite(x) & np.isfinite(y)
        den = np.abs(y[m]).max() if m.any() else 0.0
        if den > 0:
            worst = max(worst, float(np.abs(x[m] - y[m]).max() / den))
    return worst


def make_pair(grid, bz, blocks, mixed=False, persistent=False, max_blocks=None, physics=None):
    physics = physics or sph.PHYSICS
    eng = sph.Engine(bz, max_blocks or len(blocks))
    eng.set_precision(mixed)
    eng.set_schedule(persistent)
    run = eng.add_run(physics, grid, blocks)
    info_g = run.init_torus(sph.TORUS, 42)
    orc = OracleRun(dict(_cfg(grid, bz), riemann=sph.riemann_code(physics)), blocks)
    info_o = orc.init_torus(_tor(), 42)
    return eng, run, orc, info_g, info_o


def compare_states(run, orc, what):
    Ug, Wg, Bg = run.state()
    Uo, Wo, Bo = orc.state()
    ru = rel_linf(run.active(Ug), run.active(Uo))
    rw = rel_linf(run.active(Wg), run.active(Wo))
    rb = rel_linf(Bg, Bo)
    nd = int((Ug != Uo).sum() + (Bg != Bo).sum())
    print(f"[{what}] rel Linf U {ru:.3e} W {rw:.3e} Bf {rb:.3e}; differing words {nd}", flush=True)
    return ru, rw, rb


def _steps(eng, run, orc, nsteps, mixed, name):
    for s in range(nsteps):
        dtg = eng.step([run])[0, 0]
        dto = orc.step(mixed)
        assert dtg == dto or abs(dtg - dto) <= TOL * abs(dto), (s, dtg, dto)
        ru, rw, rb = compare_states(run, orc, f"{name} step {s}")
        assert ru <= TOL and rb <= TOL, (s, ru, rb)
    run.c2p()
    orc.c2p()
    ru, rw, rb = compare_states(run, orc, f"{name} primitives after {nsteps} steps")
    assert rw <= TOL


def test_c1_initial_data_matches_oracle():
    g = sph.c1_grid()
    eng, run, orc, ig, io = make_pair(g, 1, sph.uniform_blocks(g, 1))
    print("init info gpu", ig, "oracle", io)
    assert np.array_equal(ig, io)
    ru, rw, rb = compare_states(run, orc, "C1 initial data")
    assert ru == 0.0 and rw == 0.0 and rb == 0.0


def test_c1_ten_steps_fp64():
    g = sph.c1_grid()
    eng, run, orc, _, _ = make_pair(g, 1, sph.uniform_blocks(g, 1))
    _steps(eng, run, orc, 10, False, "C1")
    d = run.diagnostics()
    assert d["divb_max"] <= 1e-13 * d["divb_scale"]


def test_c1_ten_steps_fp64_hll():
    """The two-wave solver (riemann = "hll"), C1, 10 steps."""
    g = sph.c1_grid()
    eng, run, orc, _, _ = make_pair(g, 1, sph.uniform_blocks(g, 1), physics=dict(sph.PHYSICS, riemann="hll"))
    _steps(eng, run, orc, 10, False, "C1 HLL")


def test_c1_ten_steps_mixed():
    g = sph.c1_grid()
    eng, run, orc, _, _ = make_pair(g, 1, sph.uniform_blocks(g, 1), mixed=True)
    _steps(eng, run, orc, 10, True, "C1 mixed")


def test_c2_cut_down_ten_steps_fp64():
    """C2's 3D torus on a 64 x 32 x 32 grid (16 blocks), 10 steps."""
    g = dict(sph.c2_grid(), nr=64, nth=32, nph=32)
    eng, run, orc, _, _ = make_pair(g, 16, sph.uniform_blocks(g, 16))
    _steps(eng, run, orc, 10, False, "C2/64")


def test_amr_three_levels_fp64():
    """Three levels (4, 40 and 192 blocks): prolongation, restriction and
    face-field transfer across levels."""
    g = dict(nr=48, nth=32, nph=32, nlevels=3, r_in=sph.R_IN, r_out=sph.R_OUT)
    blocks = sph.amr_blocks(g, 3)
    assert sorted(set(blocks[:, 0].tolist())) == [0, 1, 2]
    eng, run, orc, _, _ = make_pair(g, 16, blocks)
    _steps(eng, run, orc, 4, False, "AMR3")


@pytest.mark.parametrize("schedule,mixed,bz", [(1, False, 16), (2, False, 16), (2, True, 16), (2, False, 1),
                                               (3, False, 16), (3, True, 1), (4, False, 16), (4, True, 16),
                                               (4, False, 1), (3, 2, 16), (4, 2, 16), (1, 2, 16)])
def test_persistent_schedule_identical(schedule, mixed, bz):
    """The persistent (1), slab (2: fluxes and hydro update fused per block),
    fused-con2prim (3: stage 1's con2prim inside stage 0's update) and queue
    (4: one launch per stage, work-queue phases between grid barriers)
    schedules give the sub-phase schedule's state and counters bit for bit
    (3-level tree; bz = 1: the axisymmetric C1 grid)."""
    if bz == 16:
        g = dict(nr=48, nth=32, nph=32, nlevels=3, r_in=sph.R_IN, r_out=sph.R_OUT)
        blocks = sph.amr_blocks(g, 3)
    else:
        g = sph.c1_grid()
        blocks = sph.uniform_blocks(g, 1)
    a = sph.Engine(bz, len(blocks))
    b = sph.Engine(bz, len(blocks))
    a.set_precision(mixed)
    b.set_precision(mixed)
    b.set_schedule(schedule)
    ra = a.add_run(sph.PHYSICS, g, blocks)
    rb = b.add_run(sph.PHYSICS, g, blocks)
    ra.init_torus(sph.TORUS, 42)
    rb.init_torus(sph.TORUS, 42)
    for s in range(3):
        da = a.step([ra])
        db = b.step([rb])
        assert np.array_equal(da, db)
        for x, y in zip(ra.state(), rb.state()):
            assert np.array_equal(x, y), f"step {s}"
    xa, xb = ra.diagnostics(), rb.diagnostics()
    assert (xa["c2p_fail"], xa["floors"]) == (xb["c2p_fail"], xb["floors"])


@pytest.mark.parametrize("schedule", [1, 4])
def test_grid_barrier_timeout_is_reported(schedule):
    """The persistent schedules' grid barriers wait a bounded number of polls.
    With the bound forced down to one poll a barrier gives up: every
    workgroup still exits, the step reports the timeout (it is not lost when
    the next stage's launch clears the queue counters), and with the default
    bound restored the engine steps as an engine that never timed out."""
    from bhfluid import BhError
    g = dict(nr=48, nth=32, nph=32, nlevels=3, r_in=sph.R_IN, r_out=sph.R_OUT)
    blocks = sph.amr_blocks(g, 3)
    eng = sph.Engine(16, len(blocks))
    eng.set_schedule(schedule)
    run = eng.add_run(sph.PHYSICS, g, blocks)
    run.init_torus(sph.TORUS, 42)
    eng.set_barrier_spin_limit(1)
    with pytest.raises(BhError, match="grid barrier timed out"):
        eng.step([run], 1)
    eng.set_barrier_spin_limit(0)
    run.init_torus(sph.TORUS, 42)
    ref = sph.Engine(16, len(blocks))
    rr = ref.add_run(sph.PHYSICS, g, blocks)
    rr.init_torus(sph.TORUS, 42)
    assert np.array_equal(eng.step([run], 2), ref.step([rr], 2))
    for x, y in zip(run.state(), rr.state()):
        assert np.array_equal(x, y)


def test_mixed_fast_accuracy_against_fp64():
    """Mixed precision with the hardware FP32 reciprocal and square root
    (precision 2): not the oracle's bits, so it is held to the accuracy bar
    of the mixed precision instead -- after 100 steps of the 3-level tree,
    L1(rho) against the FP64 run (zone-volume weighted) at most 1e-5 and
    within a factor 4 of the exactly rounded mixed run's, mass and energy
    within 1e-8 of FP64, div B at round-off; replays are deterministic."""
    g = dict(nr=48, nth=32, nph=32, nlevels=3, r_in=sph.R_IN, r_out=sph.R_OUT)
    blocks = sph.amr_blocks(g, 3)
    w = sph.zone_volumes(g, blocks)
    out = {}
    for prec in (0, 1, 2, 2):
        eng = sph.Engine(16, len(blocks))
        eng.set_precision(prec)
        run = eng.add_run(sph.PHYSICS, g, blocks)
        run.init_torus(sph.TORUS, 42)
        eng.step([run], 100, want_dt=False)
        run.c2p()
        d = run.diagnostics()
        W = run.active(run.state()[1])
        assert np.isfinite(W).all()
        if prec in out:
            assert np.array_equal(W, out[prec][0]), "mixed-fast run not reproducible"
        out[prec] = (W.copy(), d)
    rho64 = out[0][0][:, 0]
    l1 = {p: float((np.abs(out[p][0][:, 0] - rho64) * w).sum() / (np.abs(rho64) * w).sum()) for p in (1, 2)}
    print("L1(rho) vs FP64 after 100 steps: mixed", l1[1], "mixed-fast", l1[2])
    assert l1[2] <= 1e-5 and l1[2] <= 4 * max(l1[1], 1e-12)
    f, q = out[0][1], out[2][1]
    assert abs(q["mass"] - f["mass"]) <= 1e-8 * abs(f["mass"])
    assert abs(q["energy"] - f["energy"]) <= 1e-8 * abs(f["energy"])
    assert q["divb_max"] <= 1e-13 * q["divb_scale"]


def test_graph_replays_identical_to_eager():
    g = dict(sph.c2_grid(), nr=64, nth=32, nph=32)
    blocks = sph.uniform_blocks(g, 16)
    a = sph.Engine(16, len(blocks))
    b = sph.Engine(16, len(blocks))
    b.use_graph(False)
    ra = a.add_run(sph.PHYSICS, g, blocks)
    rb = b.add_run(sph.PHYSICS, g, blocks)
    ra.init_torus(sph.TORUS, 42)
    rb.init_torus(sph.TORUS, 42)
    for s in range(3):
        a.step([ra])
        b.step([rb])
        ra.diagnostics()  # eager work between replays
        for x, y in zip(ra.state(), rb.state()):
            assert np.array_equal(x, y), f"replay {s}"


def test_c2_full_size_ten_steps_fp64():
    """C2 at its full size (256 x 128 x 128, 1024 blocks, 4.19 M zones): 10
    FP64 steps against the oracle (which takes a few seconds per step on the
    host's cores)."""
    g = sph.c2_grid()
    eng, run, orc, _, _ = make_pair(g, 16, sph.uniform_blocks(g, 16))
    _steps(eng, run, orc, 10, False, "C2 full")


@pytest.mark.parametrize("mixed", [False, True])
def test_c3_full_size_three_steps(mixed):
    """The benchmark workload itself: the 3-level AMR torus (1976 blocks of
    16^3, 8.09 M zones; 56, 384 and 1536 blocks per level), 3 steps in the
    launch configuration the engine uses by default, against the oracle
    (FP64, and the mixed-precision step against the oracle's mixed mode)."""
    g = sph.c3_grid()
    blocks = sph.amr_blocks(g, 3)
    assert len(blocks) == 1976
    eng, run, orc, _, _ = make_pair(g, 16, blocks, mixed=mixed)
    _steps(eng, run, orc, 3, mixed, "C3 mixed" if mixed else "C3")


def test_c2_full_size_stays_finite_div_b_round_off():
    """C2 at its full size (256 x 128 x 128, 1024 blocks), 10 steps FP64:
    finite, max |div B| at round-off of the stencil scale, mass and energy
    within the boundary outflow of their initial values."""
    g = sph.c2_grid()
    blocks = sph.uniform_blocks(g, 16)
    eng = sph.Engine(16, len(blocks))
    run = eng.add_run(sph.PHYSICS, g, blocks)
    run.init_torus(sph.TORUS, 42)
    d0 = run.diagnostics()
    eng.step([run], 10)
    run.c2p()
    d = run.diagnostics()
    U, W, Bf = run.state()
    assert np.isfinite(run.active(U)).all() and np.isfinite(run.active(W)).all()
    print("C2 diagnostics", d0, d)
    assert d["divb_max"] <= 1e-13 * d["divb_scale"]
    assert abs(d["mass"] - d0["mass"]) <= 1e-4 * d0["mass"]


@pytest.mark.parametrize("mixed", [False, True])
def test_launch_variants_identical(mixed):
    """Every launch variant (workgroup shapes, occupancy hints, the line
    schedule of the theta/phi fluxes, XCD numbering) gives the same bits."""
    g = dict(nr=48, nth=32, nph=32, nlevels=3, r_in=sph.R_IN, r_out=sph.R_OUT)
    blocks = sph.amr_blocks(g, 3)
    ref = None
    for flux, update, c2p, xcd in ((0, 0, 0, 1), (6, 2, 2, 1), (9, 5, 4, 0), (11, 1, 3, 1)):
        eng = sph.Engine(16, len(blocks))
        eng.set_precision(mixed)
        eng.set_variants(flux, update, c2p)
        eng.set_xcd_remap(bool(xcd))
        run = eng.add_run(sph.PHYSICS, g, blocks)
        run.init_torus(sph.TORUS, 42)
        dt = eng.step([run], 3)
        st = run.state()
        if ref is None:
            ref = (dt, st)
        else:
            assert np.array_equal(dt, ref[0]), (flux, update, c2p, xcd)
            for x, y in zip(st, ref[1]):
                assert np.array_equal(x, y), (flux, update, c2p, xcd)


@pytest.mark.parametrize("mixed,bz", [(False, 16), (True, 16), (False, 1), (2, 16)])
def test_face_variants_identical(mixed, bz):
    """The face-field update with each edge's EMF formed once in LDS (2, 4,
    8, 16 workgroups per block; bz = 1: one) gives the bits of the update
    that forms a face's four EMFs itself (variant 0), on a 3-level tree and
    on the axisymmetric C1 grid."""
    if bz == 16:
        g = dict(nr=48, nth=32, nph=32, nlevels=3, r_in=sph.R_IN, r_out=sph.R_OUT)
        blocks = sph.amr_blocks(g, 3)
    else:
        g = sph.c1_grid()
        blocks = sph.uniform_blocks(g, 1)
    ref = None
    for face in (0, 1, 2, 3, 4):
        eng = sph.Engine(bz, len(blocks))
        eng.set_precision(mixed)
        eng.set_face_variant(face)
        run = eng.add_run(sph.PHYSICS, g, blocks)
        run.init_torus(sph.TORUS, 42)
        dt = eng.step([run], 3)
        st = run.state()
        if ref is None:
            ref = (dt, st)
        else:
            assert np.array_equal(dt, ref[0]), face
            for x, y in zip(st, ref[1]):
                assert np.array_equal(x, y), face


@pytest.mark.parametrize("mixed", [True, 2])
def test_serving_results_same_alone_and_in_the_mix(mixed):
    """Continuous batching: requests of different grids, spins and fields
    arrive while others run, are admitted as block slots free up (a budget
    smaller than their sum), advance in shared launches, and leave when done.
    Each result is bit-identical to the same request run alone."""
    from bhfluid.serve import Request, Server
    reqs = []
    specs = [((64, 64, 64), 1, 0.3, 30.0), ((32, 32, 32), 2, 0.9, 100.0), ((96, 64, 64), 1, 0.6, 300.0),
             ((48, 32, 32), 2, 0.1, 100.0), ((64, 64, 64), 1, 0.95, 30.0)]
    for i, (sz, nl, a, beta) in enumerate(specs):
        g = dict(nr=sz[0], nth=sz[1], nph=sz[2], nlevels=nl, r_in=sph.R_IN, r_out=sph.R_OUT)
        reqs.append(Request(i, 0.0, g, dict(sph.PHYSICS, a=a), dict(sph.TORUS, beta=beta), seed=100 + i,
                            nsteps=4 + 2 * i))
    from bhfluid.serve import request_blocks
    nbs = [len(request_blocks(r)) for r in reqs]
    budget = max(nbs) + min(nbs)          # never all at once: arrivals wait for slots
    srv = Server(budget, mixed=mixed, keep_state=True)
    summary = srv.run(reqs, realtime=False)
    print("served", summary, "blocks", nbs)
    assert summary["max_batch_runs"] >= 2
    batch_sizes = {x[1] for x in srv.log}
    assert len(batch_sizes) >= 2            # the mix changed while running
    for r in reqs:
        alone = Server(nbs[r.rid], mixed=mixed, keep_state=True)
        solo = Request(r.rid, 0.0, r.grid, r.physics, r.torus, r.seed, r.nsteps)
        alone.run([solo], realtime=False)
        assert solo.result["t"] == r.result["t"]
        for x, y in zip(solo.result["state"], r.result["state"]):
            assert np.array_equal(x, y), f"request {r.rid} differs alone and in the mix"


@pytest.mark.parametrize("mixed", [False, True])
def test_runs_with_different_solvers_share_launches(mixed):
    """An HLL run and an HLLD run stepped together (the flux kernels that
    follow each run's solver) give the bits of each run alone (an HLLD run
    alone takes the kernels compiled with the HLLD solver only)."""
    g = dict(sph.c2_grid(), nr=64, nth=32, nph=32)
    blocks = sph.uniform_blocks(g, 16)
    phys = [dict(sph.PHYSICS, riemann="hll"), dict(sph.PHYSICS, riemann="hlld", a=0.5)]
    mix = sph.Engine(16, 2 * len(blocks))
    mix.set_precision(mixed)
    runs = [mix.add_run(p, g, blocks) for p in phys]
    for r in runs:
        r.init_torus(sph.TORUS, 42)
    dts = mix.step(runs, 3)
    for n, p in enumerate(phys):
        e = sph.Engine(16, len(blocks))
        e.set_precision(mixed)
        r = e.add_run(p, g, blocks)
        r.init_torus(sph.TORUS, 42)
        d = e.step([r], 3)
        assert np.array_equal(d[:, 0], dts[:, n]), p["riemann"]
        for x, y in zip(r.state(), runs[n].state()):
            assert np.array_equal(x, y), p["riemann"]


def test_autotuner_choice_is_applied_and_exact(tmp_path, monkeypatch):
    from bhfluid import tune
    monkeypatch.setenv("BHFLUID_TUNE_CACHE", str(tmp_path / "tune.json"))
    g = dict(sph.c2_grid(), nr=64, nth=32, nph=32)
    blocks = sph.uniform_blocks(g, 16)
    eng = sph.Engine(16, len(blocks))
    run = eng.add_run(sph.PHYSICS, g, blocks)
    run.init_torus(sph.TORUS, 42)
    t = tune.tune(eng, [run], False, steps=2)
    assert not t["cached"] and len(t["timings"]) >= 10
    t2 = tune.tune(eng, [run], False, steps=2)
    assert t2["cached"] and t2["config"] == t["config"]
    # the tuned configuration steps to the same bits as the default one
    ref = sph.Engine(16, len(blocks))
    rr = ref.add_run(sph.PHYSICS, g, blocks)
    rr.init_torus(sph.TORUS, 42)
    run.init_torus(sph.TORUS, 42)
    eng.step([run], 3)
    ref.step([rr], 3)
    for x, y in zip(run.state(), rr.state()):
        assert np.array_equal(x, y)


def test_diagnostics_against_oracle():
    """Device reductions (rest mass, Killing energy, max |div B|, Mdot and
    Phi_B on the horizon row) against the oracle's sequential sums after 5
    steps of the 3-level tree: the maxima bit for bit, the sums within the
    summation-order bound."""
    g = dict(nr=48, nth=32, nph=32, nlevels=3, r_in=sph.R_IN, r_out=sph.R_OUT)
    blocks = sph.amr_blocks(g, 3)
    eng, run, orc, _, _ = make_pair(g, 16, blocks)
    for _ in range(5):
        eng.step([run])
        orc.step(False)
    dg, do = run.diagnostics(), orc.diag()
    print("device", dg, "oracle", do)
    assert dg["divb_max"] == do["divb_max"] and dg["divb_scale"] == do["divb_scale"]
    n = len(blocks) * 4096
    for k in ("mass", "energy", "mdot", "phi_b"):
        assert abs(dg[k] - do[k]) <= n * 2.2e-16 * abs(do[k]) + 1e-300, k
    assert dg["mdot"] > 0 and dg["phi_b"] >= 0


def test_analysis_against_oracle():
    """The analyser quantities (bhs_run_analysis: thermal and magnetic
    energy, Ldot, L_BZ, eta_BZ, phi_BH, P_beta, alpha_SS, Maxwell/Reynolds)
    from the device's block sums against the oracle's sequential sums after
    5 steps of the 3-level tree, within the summation-order bound."""
    g = dict(nr=48, nth=32, nph=32, nlevels=3, r_in=sph.R_IN, r_out=sph.R_OUT)
    blocks = sph.amr_blocks(g, 3)
    eng, run, orc, _, _ = make_pair(g, 16, blocks)
    for _ in range(5):
        eng.step([run])
        orc.step(False)
    rho_min = 1e-3
    ag, ao = run.analysis(rho_min), orc.analysis(rho_min)
    print("device", ag, "oracle", ao)
    for k, vo in ao.items():
        vg = ag[k]
        assert np.isfinite(vg), k
        assert abs(vg - vo) <= 1e-8 * abs(vo) + 1e-300, (k, vg, vo)
    assert ag["thermE"] > 0 and ag["magE"] > 0 and ag["P_beta"] > 1.0
