"""Levels in the reference's own 8^3 blocks (src/geometry/blocks.jl:14, BLOCK_SIZE = 8) through the C ABI:
the library regroups them into its 16^3 blocks, and every per-cell host array stays in the caller's 8^3
layout.  Stepped against the oracle run on the same 8^3 grid, in the reference layout: bit for bit in
FP32 and FP64, and in the 16-bit forms with round-to-nearest storage (stochastic rounding draws its
bits by the library's cell numbering, which differs from an 8^3 grid's)."""
import numpy as np
import pytest

from openaerolab_alpha_01_amd import _lib, solver
from openaerolab_alpha_01_amd.geometry import sphere_mask, uniform_grid

from test_gpu_parity import compare, fields, gpu_plan, oracle_run

pytestmark = pytest.mark.gpu

CASES = {
    # 5 x 3 x 3 blocks of 8^3: the 16^3 blocks overhang the extent on every axis
    "channel_sphere": dict(n=(40, 24, 24), periodic=(False, False, False), solid=((14, 12, 11), 4.0),
                           okw=dict(tau=0.53, u_inlet=0.05, inlet_turbulence=0.01)),
    "periodic": dict(n=(32, 32, 16), periodic=(True, True, True), solid=None,
                     okw=dict(tau=0.52, body_force=(1e-5, 0.0, 0.0))),
    "interior": dict(n=(96, 64, 64), periodic=(False, False, False), solid=((22, 30, 33), 4.5),
                     okw=dict(tau=0.53, u_inlet=0.05, ymin_wall_bc=1, zwall_bc=2)),
}


def _grid(case):
    cfg = CASES[case]
    g = uniform_grid(cfg["n"], periodic=cfg["periodic"], bs=8)
    if cfg["solid"]:
        g.solid = sphere_mask(g, *cfg["solid"])
    return g, dict(cfg["okw"])


@pytest.mark.parametrize("storage,rounding", [("f32", 0), ("f64", 0), ("f16s", 1), ("bf16s", 1)])
@pytest.mark.parametrize("case", list(CASES))
def test_block8_parity(oracle, case, storage, rounding):
    grid, okw = _grid(case)
    okw["rounding"] = rounding
    rho, u = fields(grid, seed=11, u0=(okw.get("u_inlet", 0.02), 0, 0), amp=0.02)
    steps = 4
    f_or, F_or, _ = oracle_run(oracle, grid, storage, steps, okw, rho, u, force_steps=(3,) if grid.solid is not None else ())
    plan = gpu_plan(grid, storage, okw)
    assert plan.n_cells() == grid.n_cells and plan.level_info()["n_cells"] == grid.n_cells
    assert plan.level_info()["n_blocks"] == grid.n_blocks
    plan.init_equilibrium(rho, u)
    for t in range(steps):
        plan.level_step(0, t, force=grid.solid is not None and t == 3)
    compare(oracle, plan, storage, f_or, grid)
    if grid.solid is not None:
        np.testing.assert_allclose(plan.force()[0], F_or[0], rtol=1e-11, atol=1e-14)
    # macroscopic output in the caller's layout
    rho_g, u_g = plan.macro()
    _, rho_o, u_o = oracle.moments(storage, oracle.Grid(grid), f_or)
    ins = grid.inside() & (grid.solid == 0 if grid.solid is not None else True)
    np.testing.assert_allclose(rho_g[ins], rho_o[ins], rtol=0, atol=1e-6 if storage != "f64" else 1e-14)
    plan.close()


def test_block8_upload_download_and_raw_state(oracle, tmp_path):
    grid, okw = _grid("channel_sphere")
    rng = np.random.default_rng(3)
    n = grid.n_cells
    f = (np.asarray(oracle.lattice()[-1])[:, None] * (1.0 + 0.01 * rng.standard_normal((27, n)))).astype(np.float64)
    plan = gpu_plan(grid, "f64", okw)
    plan.upload_f(f)
    got = plan.download_f(dtype=np.float64)
    ins = grid.inside()
    np.testing.assert_array_equal(got[:, ins], f[:, ins])
    # the LBM1 file carries the caller's 8^3 blocks
    path = tmp_path / "level_1.bin"
    plan.write_raw_state(path, dx=0.5)
    head = np.fromfile(path, dtype=np.int32, count=7)
    assert head[2] == grid.n_blocks and head[3] == 8 and tuple(head[4:7]) == (5, 3, 3)
    rho0, u0 = plan.macro()
    plan.read_raw_state(path)
    rho1, u1 = plan.macro()
    fl = ins & (grid.solid == 0)
    np.testing.assert_allclose(rho1[fl], rho0[fl].astype(np.float32), rtol=0, atol=1e-6)
    np.testing.assert_allclose(u1[:, fl], u0[:, fl].astype(np.float32), rtol=0, atol=1e-7)
    plan.close()


def test_block8_rejects_holes_and_refined(gpu):
    # a 16^3 block with an octant missing inside the extent
    grid = uniform_grid((32, 32, 32), bs=8)
    keep = np.ones(grid.n_blocks, dtype=bool)
    keep[0] = False
    from openaerolab_alpha_01_amd.geometry import BlockGrid, build_neighbor_table
    coords = grid.coords[keep]
    holey = BlockGrid(n=grid.n, periodic=grid.periodic, coords=coords,
                      neighbor=build_neighbor_table(coords, (4, 4, 4)), bs=8)
    with pytest.raises(_lib.OalError) as e:
        gpu_plan(holey, "f32", dict(tau=0.6))
    assert e.value.code == _lib.OAL_ERR_UNSUPPORTED
    with pytest.raises(_lib.OalError) as e:
        solver.octree_plan([uniform_grid((32, 32, 32), bs=8), uniform_grid((64, 64, 64), bs=8)], 0.55)
    assert e.value.code == _lib.OAL_ERR_UNSUPPORTED
