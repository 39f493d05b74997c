"""Parity of refined (block-octree) plans on the GPU against the oracle's refined runs, through the C ABI.

Both sides build the same covered / ghost blocks and transfer stencils from the same caller levels, start
from the same per-level init_eq! state and run the same recursive steps (src/octree_q19/src/coupling.jl:
236-268).  The populations of every level's caller cells must agree bit for bit with the oracle's FAST
form in every storage form; the FP64 force sums differ only in summation order (relative 1e-11).
"""
import numpy as np
import pytest

from openaerolab_alpha_01_amd import solver
from openaerolab_alpha_01_amd.geometry import block_octree, sphere_octree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def octree():
    from oracle import octree
    return octree


def _slab():
    return block_octree((32, 16, 32), [{(1, 0, z) for z in range(2)}], periodic=(True, True, True))


def _sphere2():
    # non-periodic channel, inlet at x = 0; the sphere and its surroundings on the fine level
    return sphere_octree((48, 32, 32), 2, (40.0, 32.0, 32.0), 5.0, (8,))


def _three():
    return block_octree((64, 16, 64), [{(1, 0, 1)}, {(2, 0, 2)}], periodic=(True, True, True))


def _sphere4():
    # the block-AMR sphere case at reduced extent: 4 levels, inlet / outlet channel, the sphere (radius 6
    # finest cells) on the finest level, refinement by distance to its surface (2, 16, 64, 32 blocks)
    return sphere_octree((32, 16, 16), 4, (80.0, 64.0, 64.0), 6.0, (8, 6, 3))


CASES = {
    "slab_periodic": dict(grids=_slab, tau=0.56, u_inlet=0.0, u0=(0.03, 0.0, 0.0)),
    "sphere_channel": dict(grids=_sphere2, tau=0.54, u_inlet=0.04, u0=(0.04, 0.0, 0.0)),
    "three_levels": dict(grids=_three, tau=0.58, u_inlet=0.0, u0=(0.02, 0.01, 0.0)),
}
FOUR = {
    "acoustic": dict(grids=_sphere4, tau=0.52, u_inlet=0.04, u0=(0.04, 0.0, 0.0)),
    "legacy": dict(grids=_sphere4, tau=0.52, u_inlet=0.04, u0=(0.04, 0.0, 0.0), rule="legacy"),
}


def level_fields(g, seed, u0, amp=0.02):
    rng = np.random.default_rng(seed)
    x = g.cell_coords().astype(np.float64) + 0.5
    ph = 2 * np.pi * x / np.asarray(g.n, dtype=np.float64)
    u = np.stack([u0[0] + amp * np.sin(ph[:, 1]) * np.cos(ph[:, 2]),
                  u0[1] + 0.7 * amp * np.cos(ph[:, 0]) * np.sin(ph[:, 2]),
                  u0[2] + 0.5 * amp * np.sin(ph[:, 0] + ph[:, 1])])
    u += 1e-3 * rng.uniform(-1, 1, u.shape)
    rho = 1.0 + 0.01 * np.sin(ph[:, 0]) * np.cos(ph[:, 1]) + 1e-4 * rng.uniform(-1, 1, x.shape[0])
    if g.solid is not None:
        s = np.asarray(g.solid) != 0
        rho[s] = 1.0
        u[:, s] = 0.0
    return rho, u


def run_both(oracle, octree, case, storage, steps, rounding="stochastic", force=False, cfg=None):
    cfg = cfg or CASES[case]
    grids = cfg["grids"]()
    fields = [level_fields(g, 7 + l, cfg["u0"]) for l, g in enumerate(grids)]
    rule = cfg.get("rule")
    taus = solver.level_taus(cfg["tau"], len(grids), rule) if rule else solver.acoustic_taus(cfg["tau"], len(grids))
    rcode = {"stochastic": oracle.STOCHASTIC, "nearest": oracle.NEAREST}[rounding]
    wale = dict(cfg.get("wale", {}))
    tfs = cfg.get("tfs", 0.0)   # transverse freestream blend on every level (0: off)
    if wale.get("amr_scales") == "halving":      # the reference's earlier per-level rule, 2^-(L-1-l)
        wale["amr_scales"] = [0.5 ** (len(grids) - 1 - l) for l in range(len(grids))]
    run = octree.OctreeRun(grids, storage, taus, u_inlet=cfg["u_inlet"], rounding=rcode,
                           amr_scales=wale.get("amr_scales"), c_wale=wale.get("c_wale", 0.0),
                           nu_sgs_background=wale.get("nu_sgs_background", 0.0),
                           transverse_freestream=1 if tfs else 0, transverse_freestream_blend=tfs or 1.0)
    run.init([f[0] for f in fields], [f[1] for f in fields])
    run.step(steps, force=force)
    params = solver.StepParams(c_wale=wale.get("c_wale", 0.0), nu_sgs_background=wale.get("nu_sgs_background", 0.0),
                               inlet_turbulence=0.0, transverse_freestream=1 if tfs else 0,
                               transverse_freestream_blend=tfs or 1.0)
    plan = solver.octree_plan(grids, cfg["tau"], storage=storage, params=params, u_inlet=cfg["u_inlet"],
                              rounding=rounding, taus=taus, amr_sgs_scales=wale.get("amr_scales"))
    for l, (rho, u) in enumerate(fields):
        plan.init_equilibrium(rho, u, level=l)
    print(f"octree case {case} storage {storage} rounding {rounding} levels {len(grids)} "
          f"blocks {[g.n_blocks for g in grids]}", flush=True)
    plan.step(steps, force=force)
    plan.sync()
    return grids, run, plan


def compare_level(oracle, plan, run, storage, l):
    want_st = run.caller_f(l)
    if storage == "f32":
        got = plan.download_f(level=l, dtype=np.float32)
        want = want_st
    else:
        got = plan.download_f(level=l, dtype=np.float64)
        want = oracle.to_unshifted(storage, want_st) if storage in ("f16s", "bf16s") else want_st
    bad = np.argwhere(got != want)
    assert bad.size == 0, (
        f"level {l}: {len(bad)} mismatches, first {bad[:3].tolist()}, got {got[tuple(bad[0])]!r} "
        f"want {want[tuple(bad[0])]!r}, max abs diff {np.abs(got.astype(np.float64) - want).max():.3e}")


@pytest.mark.parametrize("storage", ["f64", "f32", "f16s", "bf16s"])
@pytest.mark.parametrize("case", list(CASES))
def test_octree_parity(oracle, octree, case, storage):
    grids, run, plan = run_both(oracle, octree, case, storage, steps=3)
    with plan:
        for l in range(len(grids)):
            compare_level(oracle, plan, run, storage, l)


@pytest.mark.parametrize("rule,storage", [("acoustic", "f32"), ("legacy", "f16s")])
def test_octree_four_levels_sphere(oracle, octree, rule, storage):
    """The block-AMR sphere case's depth: 4 levels with inlet, outlet and the sphere on the finest level,
    one coarse step (8 finest sub-steps), with the acoustic relaxation times and with the GPU solver's
    legacy per-level rule (src/solver/physics_scaling.jl:654-682): every level's caller cells bit for bit."""
    grids, run, plan = run_both(oracle, octree, rule, storage, steps=1, force=True, cfg=FOUR[rule])
    with plan:
        assert len(grids) == 4
        for l in range(4):
            compare_level(oracle, plan, run, storage, l)
        fine = 3
        np.testing.assert_allclose(plan.force(level=fine)[0], np.asarray(run.force[fine], dtype=np.float64),
                                   rtol=1e-11, atol=1e-14)


WALE = dict(c_wale=0.2, nu_sgs_background=1e-4, amr_scales="halving")


@pytest.mark.parametrize("storage", ["f64", "f32", "f16s", "bf16s"])
@pytest.mark.parametrize("case", ["sphere_channel", "three_levels"])
def test_octree_wale(oracle, octree, case, storage):
    """WALE on refined plans (kernel_cumulant.jl:496-595 on every level, each level's amr_sgs_scale on
    the eddy and background viscosity): the fill and the restriction record the velocity of each cell
    they write, which the receiving level's gradient gather reads; every level's caller cells bit for
    bit after 3 coarse steps."""
    grids, run, plan = run_both(oracle, octree, case, storage, steps=3, cfg=dict(CASES[case], wale=WALE))
    with plan:
        for l in range(len(grids)):
            compare_level(oracle, plan, run, storage, l)


def test_octree_wale_four_levels_sphere(oracle, octree):
    """WALE at the block-AMR sphere case's depth (4 levels, FP16, one coarse step): bit for bit."""
    grids, run, plan = run_both(oracle, octree, "acoustic", "f16s", steps=1, cfg=dict(FOUR["acoustic"], wale=WALE))
    with plan:
        for l in range(4):
            compare_level(oracle, plan, run, "f16s", l)


@pytest.mark.parametrize("storage", ["f32", "f16s"])
def test_octree_transverse_freestream(oracle, octree, storage):
    """The transverse freestream blend (kernel_cumulant.jl:323-339) on every level of the 2-level sphere
    channel, whose fine level reaches the y and z faces: every level bit for bit after 3 coarse steps."""
    grids, run, plan = run_both(oracle, octree, "sphere_channel", storage, steps=3,
                                cfg=dict(CASES["sphere_channel"], tfs=0.6))
    with plan:
        for l in range(len(grids)):
            compare_level(oracle, plan, run, storage, l)


def test_octree_parity_nearest_rounding(oracle, octree):
    grids, run, plan = run_both(oracle, octree, "slab_periodic", "f16s", steps=3, rounding="nearest")
    with plan:
        for l in range(len(grids)):
            compare_level(oracle, plan, run, "f16s", l)


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_octree_force(oracle, octree, storage):
    grids, run, plan = run_both(oracle, octree, "sphere_channel", storage, steps=3, force=True)
    with plan:
        fine = len(grids) - 1
        assert set(run.force) == set(range(len(grids)))      # every level has an obstacle array
        scale = np.abs(np.asarray(run.force[fine])).max()
        assert scale > 0
        for l, F in run.force.items():
            # the coarse level's sphere cells lie in covered blocks, which are not stepped: no force there
            np.testing.assert_allclose(plan.force(level=l)[0], np.asarray(F, dtype=np.float64), rtol=1e-11,
                                       atol=1e-11 * scale)


def test_octree_level_info(octree):
    grids = _three()
    aug = octree.build_levels(grids)
    plan = solver.octree_plan(grids, 0.58, storage="f32", params=solver.StepParams(c_wale=0.0, inlet_turbulence=0.0))
    with plan:
        for l, (g, A) in enumerate(zip(grids, aug)):
            info = plan.level_info(l)
            assert info["n_blocks"] == g.n_blocks
            assert info["n_cells"] == g.n_cells
            assert info["n_ghost_blocks"] == int((A.status == octree.GHOST).sum())
            assert info["n_covered_blocks"] == int((A.status == octree.COVERED).sum())


def _states(plan, n_levels):
    return [plan.download_f(level=l, dtype=np.float64) for l in range(n_levels)]


@pytest.mark.parametrize("case,storage,overlap,c_wale", [("sphere_channel", "f16s", 1, 0.0), ("three_levels", "bf16s", 0, 0.0),
                                                         ("sphere_channel", "f32", 0, 0.0), ("three_levels", "f32", 1, 0.2)])
def test_graph_replay_matches_eager(case, storage, overlap, c_wale):
    """oal_step replaying captured coarse-step graphs gives the eager launches' bits, call after call
    (with downloads and odd step counts between the replays, so both buffer arrangements of level 0
    are captured), and the force of a last step that collects it; with WALE the velocity buffers swap
    with the populations."""
    cfg = CASES[case]
    grids = cfg["grids"]()
    fields = [level_fields(g, 3 + l, cfg["u0"]) for l, g in enumerate(grids)]
    params = solver.StepParams(c_wale=c_wale, inlet_turbulence=0.0)
    plans = []
    for graph in (1, 0):
        p = solver.octree_plan(grids, cfg["tau"], storage=storage, params=params, u_inlet=cfg["u_inlet"])
        p.set_tuning(graph=graph, overlap=overlap)
        for l, (rho, u) in enumerate(fields):
            p.init_equilibrium(rho, u, level=l)
        plans.append(p)
    g, e = plans
    assert g.tuning["graph"] == 1 and e.tuning["graph"] == 0
    t = 0
    for n in (2, 3, 1, 2):
        print(f"graph replay {case} {storage} overlap {overlap}: steps {t}..{t + n - 1}", flush=True)
        for p in plans:
            p.step(n, first_time_step=t)
            p.sync()
        t += n
        for a, b in zip(_states(g, len(grids)), _states(e, len(grids))):
            np.testing.assert_array_equal(a, b)
    if grids[-1].solid is not None:
        for p in plans:
            p.step(2, first_time_step=t, force=True)
            p.sync()
        np.testing.assert_array_equal(g.force(level=len(grids) - 1), e.force(level=len(grids) - 1))
        for a, b in zip(_states(g, len(grids)), _states(e, len(grids))):
            np.testing.assert_array_equal(a, b)
    for p in plans:
        p.close()


def test_graph_replay_one_level():
    """a one-level plan with the graph switched on: same bits as the eager launches (FP16 stochastic
    rounding keys on the replayed step index)"""
    from openaerolab_alpha_01_amd.geometry import sphere_mask, uniform_grid
    grid = uniform_grid((80, 48, 48))
    grid.solid = sphere_mask(grid, (20.0, 24.0, 24.0), 5.0)
    rho, u = level_fields(grid, 11, (0.05, 0.0, 0.0))
    plans = []
    for graph in (1, 0):
        p = solver.uniform_plan(grid, 0.53, storage="f16s", u_inlet=0.05,
                                params=solver.StepParams(c_wale=0.0, inlet_turbulence=0.0))
        p.set_tuning(graph=graph)
        p.init_equilibrium(rho, u)
        plans.append(p)
    t = 0
    for n in (3, 2, 4):
        for p in plans:
            p.step(n, first_time_step=t)
            p.sync()
        t += n
        np.testing.assert_array_equal(plans[0].download_f(), plans[1].download_f())
    for p in plans:
        p.close()
