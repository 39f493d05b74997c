"""The reference's LBM1 raw level state (src/io/raw_state_dump.jl:10-47) written and read through the C ABI.

The file is parsed here from the layout the reference states (little-endian Int32 header, Float64 dx,
Int32 coords[3, n_blocks], Float32 rho and vel, UInt8 obstacle) and compared with the plan's own
macroscopic fields; reading it back starts a plan at the equilibrium of the stored Float32 fields,
bit for bit the state that oal_init_equilibrium gives from the same fields.
"""
import numpy as np
import pytest

from test_gpu_parity import fields, gpu_plan

from openaerolab_alpha_01_amd import solver
from openaerolab_alpha_01_amd._lib import OalError
from openaerolab_alpha_01_amd.geometry import block_octree, sphere_mask, uniform_grid

pytestmark = pytest.mark.gpu


def parse_lbm1(path):
    b = open(path, "rb").read()
    head = np.frombuffer(b, "<i4", 7, 0)
    dx = float(np.frombuffer(b, "<f8", 1, 28)[0])
    nb, bs = int(head[2]), int(head[3])
    nc = nb * bs ** 3
    off = 36
    coords = np.frombuffer(b, "<i4", 3 * nb, off).reshape(nb, 3)
    off += 12 * nb
    rho = np.frombuffer(b, "<f4", nc, off)
    off += 4 * nc
    vel = np.frombuffer(b, "<f4", 3 * nc, off).reshape(3, nc)
    off += 12 * nc
    obs = np.frombuffer(b, "u1", nc, off)
    off += nc
    assert off == len(b)
    return head, dx, coords, rho, vel, obs


def _grid():
    grid = uniform_grid((48, 32, 40))
    grid.solid = sphere_mask(grid, (17.2, 15.9, 20.4), 5.5)
    return grid


def test_write_matches_macro_fields(tmp_path):
    grid = _grid()
    okw = dict(tau=0.55, u_inlet=0.05)
    rho0, u0 = fields(grid, seed=3, u0=(0.05, 0.0, 0.0))
    with gpu_plan(grid, "f32", okw) as plan:
        plan.init_equilibrium(rho0, u0)
        plan.step(5)
        plan.sync()
        path = tmp_path / "level_1.bin"
        print(f"raw state write {path}", flush=True)
        plan.write_raw_state(path, dx=0.125)
        rho, u = plan.macro()
    head, dx, coords, r, v, obs = parse_lbm1(path)
    assert head[0] == 0x4C424D31 and head[1] == 1 and head[2] == grid.n_blocks and head[3] == 16
    np.testing.assert_array_equal(head[4:7], np.asarray(grid.coords).max(0))
    assert dx == 0.125
    np.testing.assert_array_equal(coords, grid.coords)
    solid = np.asarray(grid.solid) != 0
    np.testing.assert_array_equal(obs, solid.astype(np.uint8))
    np.testing.assert_array_equal(r[~solid], rho[~solid].astype(np.float32))
    np.testing.assert_array_equal(v[:, ~solid], u[:, ~solid].astype(np.float32))
    assert (r[solid] == 1.0).all() and (v[:, solid] == 0.0).all()


@pytest.mark.parametrize("storage", ["f32", "f16s"])
def test_read_restarts_at_equilibrium(tmp_path, storage):
    grid = _grid()
    okw = dict(tau=0.55, u_inlet=0.05)
    rho0, u0 = fields(grid, seed=4, u0=(0.05, 0.0, 0.0))
    path = tmp_path / "level_1.bin"
    with gpu_plan(grid, "f32", okw) as plan:
        plan.init_equilibrium(rho0, u0)
        plan.step(4)
        plan.write_raw_state(path, level_id=3, dx=0.5)
    _, _, _, r, v, _ = parse_lbm1(path)
    with gpu_plan(grid, storage, okw) as a, gpu_plan(grid, storage, okw) as b:
        print(f"raw state read {storage}", flush=True)
        assert a.read_raw_state(path) == (3, 0.5)
        b.init_equilibrium(r.astype(np.float64), v.astype(np.float64))
        np.testing.assert_array_equal(a.download_f(dtype=np.float64), b.download_f(dtype=np.float64))
        a.step(2)
        b.step(2)
        np.testing.assert_array_equal(a.download_f(dtype=np.float64), b.download_f(dtype=np.float64))


def test_read_rejects_other_levels(tmp_path):
    grid = _grid()
    okw = dict(tau=0.55, u_inlet=0.05)
    path = tmp_path / "level_1.bin"
    with gpu_plan(grid, "f32", okw) as plan:
        plan.init_equilibrium(None, None, 1.0, (0.05, 0.0, 0.0))
        plan.write_raw_state(path)
    other = uniform_grid((48, 32, 40))
    other.solid = sphere_mask(other, (17.2, 15.9, 20.4), 4.5)          # same blocks, another obstacle
    smaller = uniform_grid((48, 32, 32))
    smaller.solid = sphere_mask(smaller, (17.2, 15.9, 20.4), 5.5)
    bad = tmp_path / "bad.bin"
    data = bytearray(open(path, "rb").read())
    data[0] ^= 0xFF
    bad.write_bytes(bytes(data))
    trunc = tmp_path / "trunc.bin"
    trunc.write_bytes(open(path, "rb").read()[:-10])
    for g, p in ((other, path), (smaller, path), (grid, bad), (grid, trunc)):
        with gpu_plan(g, "f32", okw) as plan:
            with pytest.raises(OalError):
                plan.read_raw_state(p)


def test_refined_level_round_trip(tmp_path):
    grids = block_octree((32, 16, 32), [{(1, 0, z) for z in range(2)}], periodic=(True, True, True))
    params = solver.StepParams(c_wale=0.0, inlet_turbulence=0.0)
    path = tmp_path / "level_2.bin"
    plans = [solver.octree_plan(grids, 0.56, storage="f32", params=params) for _ in range(2)]
    with plans[0] as a, plans[1] as b:
        for l, g in enumerate(grids):
            rho, u = fields(g, seed=10 + l, u0=(0.03, 0.0, 0.0))
            a.init_equilibrium(rho, u, level=l)
            b.init_equilibrium(rho, u, level=l)
        a.step(2)
        a.write_raw_state(path, level=1)
        head, _, coords, r, v, _ = parse_lbm1(path)
        assert head[1] == 2 and head[2] == grids[1].n_blocks
        np.testing.assert_array_equal(coords, grids[1].coords)
        assert b.read_raw_state(path, level=1) == (2, 1.0)
        rho, u = b.macro(level=1)
        np.testing.assert_allclose(rho, r, rtol=2e-7)
        np.testing.assert_allclose(u, v, rtol=0, atol=2e-7)
