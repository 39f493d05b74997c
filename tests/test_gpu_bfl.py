"""Parity of the interpolated bounce-back (BFL) on the GPU against the oracle, through the C ABI.

A channel with a sphere that cuts the y = 0 face, so that links of every kind occur: q < 1/2 with a
fluid cell behind, q < 1/2 whose cell behind lies beyond the domain face (the fallback), and
q >= 1/2.  Both sides start from the same init_eq! state and run the same steps with the same
Float16 q-map; the populations must agree bit for bit with the oracle's FAST form in every storage
form and under every fallback, and the force taken at the gather to relative 1e-11 (summation order).
"""
import numpy as np
import pytest

from test_gpu_parity import compare, fields, gpu_plan
from test_oracle_bfl import bfl_case as _case

pytestmark = pytest.mark.gpu

FALLBACK = {0: "halfway_bb", 1: "wall_eq", 2: "bounce_back"}


def _run_both(oracle, storage, fallback, steps=4, force_steps=(1, 3)):
    grid, q = _case()
    okw = dict(tau=0.54, u_inlet=0.05)
    rho, u = fields(grid, seed=5, u0=(0.05, 0.0, 0.0))
    G = oracle.Grid(grid, qmap=q)
    f = oracle.init_eq(storage, G, rho, u)
    f_forces = []
    for t in range(steps):
        p = oracle.make_params(time_step=t, arith=oracle.FAST, bfl_fallback=fallback, **okw)
        f, F, _ = oracle.step(storage, G, p, f, want_force=t in force_steps)
        if t in force_steps:
            f_forces.append(np.asarray(F, dtype=np.float64))
    print(f"bfl storage {storage} fallback {FALLBACK[fallback]} cells {grid.n_cells} "
          f"links {int((q > 0).sum())}", flush=True)
    plan = gpu_plan(grid, storage, okw)
    plan.set_bouzidi(q, ff_fallback=FALLBACK[fallback])
    plan.init_equilibrium(rho, u)
    g_forces = []
    for t in range(steps):
        plan.level_step(0, t, force=t in force_steps)
        if t in force_steps:
            g_forces.append(plan.force()[0].copy())
    plan.sync()
    return grid, q, plan, f, f_forces, g_forces


@pytest.mark.parametrize("storage", ["f64", "f32", "f16s", "bf16s"])
def test_bfl_parity(oracle, storage):
    grid, q, plan, f, fo, fg = _run_both(oracle, storage, 0)
    with plan:
        compare(oracle, plan, storage, f, grid)
        for a, b in zip(fg, fo):
            assert np.abs(b).max() > 0
            np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-11 * np.abs(b).max())


@pytest.mark.parametrize("fallback", [1, 2])
def test_bfl_fallbacks(oracle, fallback):
    grid, q, plan, f, fo, fg = _run_both(oracle, "f32", fallback)
    with plan:
        compare(oracle, plan, "f32", f, grid)
        for a, b in zip(fg, fo):
            np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-11 * np.abs(b).max())


def test_bfl_off_again_is_halfway(oracle):
    """set_bouzidi(None) returns the level to halfway bounce-back (same bits as a plan without a q-map)"""
    grid, q = _case()
    okw = dict(tau=0.54, u_inlet=0.05)
    rho, u = fields(grid, seed=6, u0=(0.05, 0.0, 0.0))
    plans = [gpu_plan(grid, "f32", okw) for _ in range(2)]
    plans[0].set_bouzidi(q)
    plans[0].set_bouzidi(None)
    for p in plans:
        p.init_equilibrium(rho, u)
        p.step(3)
        p.sync()
    np.testing.assert_array_equal(plans[0].download_f(), plans[1].download_f())
    for p in plans:
        p.close()


@pytest.mark.parametrize("storage", ["f32", "f16s"])
def test_bfl_refined_level_half_q_is_halfway(storage):
    """on the finest level of a refined plan (caller blocks, then ghost blocks in the library's
    numbering), q = 1/2 on every solid link gives the halfway populations bit for bit on every level,
    and a sphere q-map runs with a finite state"""
    from openaerolab_alpha_01_amd import solver
    from openaerolab_alpha_01_amd.geometry import link_solid, sphere_octree, sphere_qmap
    centre, R = (40.0, 32.0, 32.0), 5.0
    grids = sphere_octree((48, 32, 32), 2, centre, R, (8,))
    fine = grids[-1]
    fl = (np.asarray(fine.solid) == 0) & fine.inside()
    q_half = np.where(link_solid(fine) & fl[None, :], 0.5, 0.0).astype(np.float16)
    params = solver.StepParams(c_wale=0.0, inlet_turbulence=0.0)
    plans = [solver.octree_plan(grids, 0.54, storage=storage, params=params, u_inlet=0.04) for _ in range(3)]
    plans[0].set_bouzidi(q_half, level=len(grids) - 1)
    plans[2].set_bouzidi(sphere_qmap(fine, centre, R), level=len(grids) - 1)
    for p in plans:
        for l, g in enumerate(grids):
            rho, u = fields(g, seed=30 + l, u0=(0.04, 0.0, 0.0))
            p.init_equilibrium(rho, u, level=l)
        print(f"bfl refined {storage}", flush=True)
        p.step(3)
        p.sync()
    for l in range(len(grids)):
        np.testing.assert_array_equal(plans[0].download_f(level=l, dtype=np.float64),
                                      plans[1].download_f(level=l, dtype=np.float64))
        assert plans[2].guard(0.2, 0.5, level=l) == 0
        assert not np.array_equal(plans[2].download_f(level=l, dtype=np.float64),
                                  plans[1].download_f(level=l, dtype=np.float64)) or l == 0
    for p in plans:
        p.close()
