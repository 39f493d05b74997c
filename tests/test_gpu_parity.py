"""Parity of liboal (HIP, gfx950) against the CPU oracle, through the C ABI.

Both sides start from the same init_eq! state (same FP64 fields, rounded to the working precision the
same way) and run the same steps.

  * Against the oracle's FAST form (the library's fused arithmetic restated on the CPU, operation for
    operation) the populations must agree bit for bit in every storage form; FP64 reductions (force,
    mass) differ only in summation order: relative 1e-11.
  * Against the LITERAL form (the reference arithmetic operation for operation) the plain FP32 path
    must agree within 1e-5 relative (max-norm of the populations and of the velocity field).
"""
import numpy as np
import pytest

from openaerolab_alpha_01_amd import solver
from openaerolab_alpha_01_amd.geometry import sphere_mask, uniform_grid

pytestmark = pytest.mark.gpu

NP = {"f32": np.float32, "f64": np.float64, "f16s": np.float32, "bf16s": np.float32}


def fields(grid, seed=0, u0=(0.0, 0.0, 0.0), amp=0.02):
    """Seeded smooth-plus-random initial rho and u (FP64) in block layout."""
    rng = np.random.default_rng(seed)
    x = grid.cell_coords().astype(np.float64) + 0.5
    n = np.asarray(grid.n, dtype=np.float64)
    ph = 2 * np.pi * x / n
    u = np.stack([u0[0] + amp * np.sin(ph[:, 1]) * np.cos(ph[:, 2]),
                  u0[1] + 0.7 * amp * np.cos(ph[:, 0]) * np.sin(ph[:, 2]),
                  u0[2] + 0.5 * amp * np.sin(ph[:, 0] + ph[:, 1])])
    u += 1e-3 * rng.uniform(-1, 1, u.shape)
    rho = 1.0 + 0.01 * np.sin(ph[:, 0]) * np.cos(ph[:, 1]) + 1e-4 * rng.uniform(-1, 1, x.shape[0])
    ins = grid.inside()
    rho[~ins] = 1.0
    u[:, ~ins] = 0.0
    return rho, u


def oracle_run(oracle, grid, storage, steps, okw, rho, u, force_steps=(), arith=1):
    G = oracle.Grid(grid)
    f = oracle.init_eq(storage, G, rho, u)
    vel = u.astype(NP[storage]) if okw.get("c_wale", 0) else None
    forces = []
    for t in range(steps):
        p = oracle.make_params(time_step=t, arith=arith, **okw)
        f, F, mac = oracle.step(storage, G, p, f, vel_in=vel, want_macro=vel is not None, want_force=t in force_steps)
        if vel is not None:
            vel = mac[1]
        if t in force_steps:
            forces.append(F)
    return f, forces, G


def gpu_plan(grid, storage, okw, stream="ab"):
    params = solver.StepParams(omega_bulk=okw.get("omega_bulk", 1.0), omega_3=okw.get("omega_3", 1.0),
                               omega_4=okw.get("omega_4", 1.0), lambda_=okw.get("lambda_", 1 / 6),
                               adaptive_omega4=okw.get("adaptive_omega4", 1), limiter=okw.get("limiter", 1),
                               compressible=okw.get("compressible", 0), c_wale=okw.get("c_wale", 0.0),
                               nu_sgs_background=okw.get("nu_sgs_background", 0.0),
                               inlet_turbulence=okw.get("inlet_turbulence", 0.0),
                               transverse_freestream=okw.get("transverse_freestream", 0),
                               transverse_freestream_blend=okw.get("transverse_freestream_blend", 1.0))
    dom = solver.Domain(extent=grid.n, tau=okw["tau"], periodic=grid.periodic, u_inlet=okw.get("u_inlet", 0.0),
                        body_force=tuple(okw.get("body_force", (0.0, 0.0, 0.0))),
                        ymin_wall_bc=okw.get("ymin_wall_bc", 0), zwall_bc=okw.get("zwall_bc", 0))
    bb = {1: "halfway", 0: "reference"}[okw.get("bb_mode", 1)]
    rounding = {0: "stochastic", 1: "nearest"}[okw.get("rounding", 0)]
    return solver.Plan(solver.Level(grid, [dom]), storage=storage, stream=stream, bb_mode=bb, params=params,
                       rounding=rounding)


def gpu_run(grid, storage, steps, okw, rho, u, stream="ab", force_steps=(), tuning=None):
    plan = gpu_plan(grid, storage, okw, stream)
    if tuning:
        plan.set_tuning(**tuning)
    plan.init_equilibrium(rho, u)
    forces = []
    for t in range(steps):
        plan.level_step(0, t, force=t in force_steps)
        if t in force_steps:
            forces.append(plan.force()[0].copy())
    return plan, forces


def compare(oracle, plan, storage, f_or, grid):
    """bit-exact comparison of all stored cells (populations in the working precision)"""
    if storage == "f32":
        got = plan.download_f(dtype=np.float32)
        want = f_or
    else:
        got = plan.download_f(dtype=np.float64)
        want = oracle.to_unshifted(storage, f_or)
    bad = np.argwhere(got != want)
    assert bad.size == 0, (
        f"{len(bad)} mismatches, first {bad[:3].tolist()}, got {got[tuple(bad[0])]!r} want {want[tuple(bad[0])]!r}, "
        f"max abs diff {np.abs(got.astype(np.float64) - want).max():.3e}")


CASES = {
    "periodic": dict(n=(32, 32, 32), periodic=(True, True, True), solid=None,
                     okw=dict(tau=0.52)),
    "channel_sphere": dict(n=(48, 32, 32), periodic=(False, False, False), solid=((16, 16, 16), 6.0),
                           okw=dict(tau=0.53, u_inlet=0.05)),
    "partial_blocks": dict(n=(40, 24, 20), periodic=(False, False, False), solid=((14, 12, 9), 4.0),
                           okw=dict(tau=0.56, u_inlet=0.04)),
    "interior_blocks": dict(n=(80, 48, 48), periodic=(False, False, False), solid=((18, 24, 24), 4.0),
                            okw=dict(tau=0.52, u_inlet=0.05)),
}


@pytest.mark.parametrize("storage", ["f32", "f64", "f16s", "bf16s"])
@pytest.mark.parametrize("case", list(CASES))
def test_step_parity(oracle, storage, case):
    cfg = CASES[case]
    grid = uniform_grid(cfg["n"], periodic=cfg["periodic"])
    if cfg["solid"]:
        grid.solid = sphere_mask(grid, *cfg["solid"])
    okw = dict(cfg["okw"])
    rho, u = fields(grid, seed=1, u0=(okw.get("u_inlet", 0.0), 0, 0))
    steps = 6
    fst = (2, 5) if grid.solid is not None else ()
    f_or, F_or, _ = oracle_run(oracle, grid, storage, steps, okw, rho, u, force_steps=fst)
    plan, F_gpu = gpu_run(grid, storage, steps, okw, rho, u, force_steps=fst)
    compare(oracle, plan, storage, f_or, grid)
    for a, b in zip(F_gpu, F_or):
        np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-14)
    plan.close()


VARIANTS = {
    "reference_reflection": dict(bb_mode=0),
    "no_limiter_fixed_omega4": dict(limiter=0, adaptive_omega4=0, omega_bulk=1.2, omega_3=0.9, omega_4=1.1),
    "positivity_limiter": dict(limiter=2, tau=0.5005, omega_3=1.9, omega_4=1.9),
    "compressible_inlet": dict(compressible=1, u_inlet=0.08),
    "inlet_noise": dict(inlet_turbulence=0.01),
    "walls": dict(ymin_wall_bc=1, zwall_bc=2),
    "wale": dict(c_wale=0.2, nu_sgs_background=1e-4),
    "body_force": dict(body_force=(1e-5, -2e-6, 3e-6)),
    "nearest_rounding": dict(rounding=1),
    "specular_zwall": dict(zwall_bc=1),
    "transverse_freestream": dict(transverse_freestream=1, transverse_freestream_blend=0.7),
}


@pytest.mark.parametrize("face_kernel", [0, 1])
@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("storage", ["f32", "f16s", "bf16s"])
def test_variant_parity(oracle, variant, storage, face_kernel):
    # 4 x 2 x 2 blocks: the sphere's blocks in the middle, the inlet and outlet blocks on the regular
    # path with its in-kernel face rules
    grid = uniform_grid((64, 32, 32))
    grid.solid = sphere_mask(grid, (31, 15, 17), 5.5)
    okw = dict(tau=0.53, u_inlet=0.05)
    okw.update(VARIANTS[variant])
    rho, u = fields(grid, seed=3, u0=(0.05, 0, 0), amp=0.03)
    steps = 5
    fst = (4,)
    f_or, F_or, _ = oracle_run(oracle, grid, storage, steps, okw, rho, u, force_steps=fst)
    plan, F_gpu = gpu_run(grid, storage, steps, okw, rho, u, force_steps=fst, tuning=dict(face_kernel=face_kernel))
    compare(oracle, plan, storage, f_or, grid)
    np.testing.assert_allclose(F_gpu[0], F_or[0], rtol=1e-11, atol=1e-14)
    plan.close()


@pytest.mark.parametrize("variant", ["specular_zwall", "transverse_freestream", "inlet_noise", "walls"])
def test_variant_parity_f64(oracle, variant):
    """The FP64 path of the optional face rules and the inlet noise (the same case as test_variant_parity)."""
    grid = uniform_grid((64, 32, 32))
    grid.solid = sphere_mask(grid, (31, 15, 17), 5.5)
    okw = dict(tau=0.53, u_inlet=0.05)
    okw.update(VARIANTS[variant])
    rho, u = fields(grid, seed=3, u0=(0.05, 0, 0), amp=0.03)
    f_or, F_or, _ = oracle_run(oracle, grid, "f64", 5, okw, rho, u, force_steps=(4,))
    plan, F_gpu = gpu_run(grid, "f64", 5, okw, rho, u, force_steps=(4,))
    compare(oracle, plan, "f64", f_or, grid)
    np.testing.assert_allclose(F_gpu[0], F_or[0], rtol=1e-11, atol=1e-14)
    plan.close()


PAIR_VARIANTS = {k: VARIANTS[k] for k in ("no_limiter_fixed_omega4", "positivity_limiter", "body_force",
                                           "nearest_rounding", "inlet_noise")}
PAIR_VARIANTS["unit_rates"] = {}


@pytest.mark.parametrize("pair", [1, 2])
@pytest.mark.parametrize("variant", list(PAIR_VARIANTS) + ["walls"])
@pytest.mark.parametrize("storage", ["f32", "f16s", "bf16s"])
def test_interior_blocks_rates_and_forcing(oracle, variant, storage, pair):
    """A box with 16 interior blocks on the regular grid (6 x 4 x 4 blocks), stepped by the pipelined
    kernels (the 16-bit forms two cells per lane) with the reference's unit rates and with others, a
    body force, the positivity limiter and round-to-nearest storage: the oracle's bits."""
    grid = uniform_grid((96, 64, 64))
    grid.solid = sphere_mask(grid, (22, 30, 33), 4.5)
    okw = dict(tau=0.53, u_inlet=0.05)
    okw.update(VARIANTS[variant] if variant == "walls" else PAIR_VARIANTS[variant])
    rho, u = fields(grid, seed=9, u0=(0.05, 0, 0), amp=0.03)
    steps = 3
    f_or, _, _ = oracle_run(oracle, grid, storage, steps, okw, rho, u)
    plan = gpu_plan(grid, storage, okw)
    assert plan.level_info()["n_regular"] >= 40
    # the blocks around the sphere: the pipelined kernel's solid-link form (steps without the force)
    assert plan.level_info()["n_solid_regular"] > 0
    plan.set_tuning(pipe_planes=4, overlap=1, face_kernel=0, graph=0, pair_kernel=pair)
    plan.init_equilibrium(rho, u)
    for t in range(steps):
        plan.level_step(0, t)
    compare(oracle, plan, storage, f_or, grid)
    plan.close()


def limiter_state(oracle, storage, grid, which):
    """Far-from-equilibrium populations (reference layout, exactly representable in `storage`) on which
    a limiter acts in some cells and not in others: random populations for the positivity limiter,
    random populations plus a strong random shear (Pi_xy, Pi_xz, Pi_yz) for the realizability clamp."""
    cx, cy, cz, w = (oracle.lattice()[i] for i in (0, 1, 2, 7))
    n = grid.n_blocks * 4096
    rng = np.random.default_rng(7)
    if which == "clamp":
        sh = 4.0 * rng.uniform(-1, 1, (3, n))
        f = w[:, None] * (1 + 0.3 * rng.uniform(-1, 1, (27, n)) + (cx * cy)[:, None] * sh[0]
                          + (cx * cz)[:, None] * sh[1] + (cy * cz)[:, None] * sh[2])
    else:
        f = w[:, None] * (1 + 0.4 * rng.uniform(-1, 1, (27, n)))
    return oracle.to_unshifted(storage, oracle.from_unshifted(storage, f))


LIMITER_OKW = {"clamp": dict(tau=0.52, limiter=1), "positivity": dict(tau=0.5005, limiter=2, omega_3=1.9, omega_4=1.9)}


@pytest.mark.parametrize("tuning", ["pair", "pipe", "one_plane"])
@pytest.mark.parametrize("which", list(LIMITER_OKW))
@pytest.mark.parametrize("storage", ["f32", "f64", "f16s", "bf16s"])
def test_limiters_far_from_equilibrium(oracle, storage, which, tuning):
    """The realizability clamp of the relaxed second-order cumulants (kernel_cumulant.jl:715-735, the
    square roots taken only by waves with a lane that needs one) and the positivity limiter (:849-884,
    its test as a minimum tree) on states where each acts in a few per cent of the cells (checked on
    the oracle against the same step without the limiter): the oracle's bits for two steps, in the two-
    cells-per-lane, one-cell pipelined and one-plane kernels."""
    # 4 x 4 x 4 periodic blocks: the inner 8 have their neighbours at lexicographic offsets (no wrap)
    # and take the pipelined kernels, the others the boundary kernel
    grid = uniform_grid((64, 64, 64), periodic=(True, True, True))
    okw = LIMITER_OKW[which]
    f0 = limiter_state(oracle, storage, grid, which)
    G = oracle.Grid(grid)
    fs = oracle.from_unshifted(storage, f0)
    plain = dict(okw, limiter=0)
    once = oracle.step(storage, G, oracle.make_params(time_step=0, arith=oracle.FAST, **okw), fs)[0]
    free = oracle.step(storage, G, oracle.make_params(time_step=0, arith=oracle.FAST, **plain), fs)[0]
    acted = np.mean(np.any(once != free, axis=0))
    assert 0.002 < acted < 0.5, acted
    f_or = fs
    for t in range(2):
        f_or = oracle.step(storage, G, oracle.make_params(time_step=t, arith=oracle.FAST, **okw), f_or)[0]
    plan = gpu_plan(grid, storage, okw)
    assert plan.level_info()["n_regular"] == 8
    shape = {"pair": dict(pipe_planes=2, pair_kernel=1), "pipe": dict(pipe_planes=4, pair_kernel=0),
             "one_plane": dict(pipe_planes=1, pair_kernel=0)}[tuning]
    plan.set_tuning(overlap=1, face_kernel=0, graph=0, **shape)
    plan.upload_f(f0 if storage == "f64" else f0.astype(np.float32))
    for t in range(2):
        plan.level_step(0, t)
    compare(oracle, plan, storage, f_or, grid)
    plan.close()


WALE_GRIDS = {"interior": ((96, 64, 64), (22, 30, 33), 4.5), "partial": ((72, 40, 36), (14, 14, 14), 3.0)}


@pytest.mark.parametrize("rates", ["unit", "fixed"])
@pytest.mark.parametrize("planes", [1, 4, 16])
@pytest.mark.parametrize("where", list(WALE_GRIDS))
@pytest.mark.parametrize("storage", ["f32", "f16s", "bf16s"])
def test_wale_pipelined(oracle, storage, where, planes, rates):
    """WALE (kernel_cumulant.jl:496-595) on the regular blocks' pipelined kernel: the previous step's
    velocity gathered per plane before the next plane's populations, the eddy viscosity into the rates,
    the velocity written; whole blocks on and off the domain faces and blocks that overhang the extent,
    the reference's unit rates and others, every plane count (16 runs as 8).  The oracle's bits."""
    n, c, r = WALE_GRIDS[where]
    grid = uniform_grid(n)
    grid.solid = sphere_mask(grid, c, r)
    okw = dict(tau=0.52, u_inlet=0.05, c_wale=0.2, nu_sgs_background=1e-4, inlet_turbulence=0.01)
    if rates == "fixed":
        okw.update(omega_bulk=1.2, omega_3=0.9, omega_4=1.1, adaptive_omega4=0)
    rho, u = fields(grid, seed=21, u0=(0.05, 0, 0), amp=0.03)
    steps = 3
    f_or, _, _ = oracle_run(oracle, grid, storage, steps, okw, rho, u)
    plan = gpu_plan(grid, storage, okw)
    info = plan.level_info()
    assert info["n_regular"] + info["n_partial"] >= 10
    plan.set_tuning(pipe_planes=planes, overlap=1, face_kernel=planes == 16, graph=0, pair_kernel=2)
    plan.init_equilibrium(rho, u)
    for t in range(steps):
        plan.level_step(0, t)
    compare(oracle, plan, storage, f_or, grid)
    plan.close()


@pytest.mark.parametrize("rates", ["unit", "fixed"])
@pytest.mark.parametrize("planes", [1, 2, 4, 8])
@pytest.mark.parametrize("storage", ["f32", "f16s", "bf16s"])
def test_aa_pipelined(oracle, storage, planes, rates):
    """The AA pattern's regular blocks (8 of 6 x 4 x 4, off every face and the sphere) on the pipelined
    even and odd half-step kernels: after an even number of steps the state is the pull step's, the
    oracle's bits."""
    grid = uniform_grid((96, 64, 64))
    grid.solid = sphere_mask(grid, (22, 30, 33), 4.5)
    okw = dict(tau=0.53, u_inlet=0.05, inlet_turbulence=0.01)
    if rates == "fixed":
        okw.update(omega_bulk=1.2, omega_3=0.9, omega_4=1.1, adaptive_omega4=0)
    rho, u = fields(grid, seed=17, u0=(0.05, 0, 0), amp=0.03)
    steps = 4
    f_or, _, _ = oracle_run(oracle, grid, storage, steps, okw, rho, u)
    plan = gpu_plan(grid, storage, okw, stream="aa")
    assert plan.level_info()["n_regular"] >= 8
    plan.set_tuning(pipe_planes=planes, overlap=1, face_kernel=0, graph=0, pair_kernel=2)
    plan.init_equilibrium(rho, u)
    for t in range(steps):
        plan.level_step(0, t)
    compare(oracle, plan, storage, f_or, grid)
    plan.close()


@pytest.mark.parametrize("storage", ["f32", "f16s", "bf16s", "f64"])
def test_aa_pattern_equals_ab(storage):
    """The in-place AA pattern reproduces the pull/AB step bit for bit (after every even step),
    boundary rules, halfway bounce-back and momentum exchange included."""
    grid = uniform_grid((48, 32, 40))
    grid.solid = sphere_mask(grid, (20, 16, 19), 6.0)
    okw = dict(tau=0.52, u_inlet=0.05, ymin_wall_bc=1, inlet_turbulence=0.01)
    rho, u = fields(grid, seed=5, u0=(0.05, 0, 0))
    ab = gpu_plan(grid, storage, okw, "ab")
    aa = gpu_plan(grid, storage, okw, "aa")
    ab.init_equilibrium(rho, u)
    aa.init_equilibrium(rho, u)
    for t in range(8):
        want_f = t in (3, 6, 7)
        ab.level_step(0, t, force=want_f)
        aa.level_step(0, t, force=want_f)
        if want_f:
            np.testing.assert_allclose(aa.force(), ab.force(), rtol=1e-11, atol=1e-14)
        if t % 2 == 1:
            dt = np.float32 if storage == "f32" else np.float64
            assert np.array_equal(aa.download_f(dtype=dt), ab.download_f(dtype=dt)), f"step {t}"
            np.testing.assert_allclose(aa.mass(), ab.mass(), rtol=1e-13)
    ab.close()
    aa.close()


def test_mass_macro_guard(oracle):
    grid = uniform_grid((32, 32, 48))
    grid.solid = sphere_mask(grid, (16, 16, 20), 5.0)
    okw = dict(tau=0.55, u_inlet=0.03)
    rho, u = fields(grid, seed=7, u0=(0.03, 0, 0))
    f_or, _, G = oracle_run(oracle, grid, "f32", 3, okw, rho, u)
    plan, _ = gpu_run(grid, "f32", 3, okw, rho, u)
    m_or, r_or, u_or = oracle.moments("f32", G, f_or)
    np.testing.assert_allclose(plan.mass()[0], m_or, rtol=1e-12, atol=1e-12)
    r_g, u_g = plan.macro()
    fl = (grid.solid == 0)
    np.testing.assert_allclose(r_g[fl], r_or[fl], rtol=1e-14)
    np.testing.assert_allclose(u_g[:, fl], u_or[:, fl], rtol=1e-12, atol=1e-16)
    assert plan.guard(0.2, 0.5) == 0
    # a NaN population uploaded from the host trips the guard
    f = plan.download_f(dtype=np.float64)
    f[5, 1000] = np.nan
    f[7, 2000] = 5.0
    plan.upload_f(f)
    from openaerolab_alpha_01_amd import _lib
    with pytest.raises(_lib.OalError) as e:
        plan.guard(0.2, 0.5)
    assert e.value.code == _lib.OAL_ERR_NONFINITE
    plan.close()


@pytest.mark.parametrize("storage", ["f32", "f64", "f16s", "bf16s"])
def test_upload_download_roundtrip(oracle, storage):
    grid = uniform_grid((32, 16, 16))
    rho, u = fields(grid, seed=11)
    G = oracle.Grid(grid)
    f_or = oracle.init_eq(storage, G, rho, u)
    want = oracle.to_unshifted(storage, f_or)
    for stream in ("ab", "aa"):
        plan = gpu_plan(grid, storage, dict(tau=0.6), stream)
        plan.upload_f(want)
        got = plan.download_f(dtype=np.float64)
        assert np.array_equal(got, want)
        plan.init_equilibrium(rho, u)
        assert np.array_equal(plan.download_f(dtype=np.float64), want)
        plan.close()


def test_many_domains_in_one_level(oracle):
    """Independent boxes in one block list and one launch equal their separate runs."""
    from openaerolab_alpha_01_amd.geometry import BlockGrid
    specs = [((48, 32, 32), 0.05, 0.53, (18, 16, 16), 5.0), ((40, 24, 20), 0.03, 0.56, (15, 12, 10), 4.0),
             ((32, 48, 16), 0.07, 0.52, (12, 20, 8), 3.5)]
    grids, doms, coords, nbrs, solids, bdom = [], [], [], [], [], []
    off = 0
    for i, (n, uin, tau, c, r) in enumerate(specs):
        g = uniform_grid(n)
        g.solid = sphere_mask(g, c, r)
        grids.append(g)
        doms.append(solver.Domain(extent=n, tau=tau, u_inlet=uin))
        coords.append(g.coords)
        nb = g.neighbor.copy()
        nb[nb > 0] += off
        nbrs.append(nb)
        solids.append(g.solid)
        bdom.append(np.full(g.n_blocks, i, dtype=np.int32))
        off += g.n_blocks
    big = BlockGrid(n=(0, 0, 0), periodic=(False,) * 3, coords=np.concatenate(coords), neighbor=np.concatenate(nbrs),
                    solid=np.concatenate(solids))
    params = solver.StepParams(c_wale=0.0, inlet_turbulence=0.0)
    plan = solver.Plan(solver.Level(big, doms, block_domain=np.concatenate(bdom)), storage="f16s", params=params)
    rhos, us = zip(*[fields(g, seed=20 + i, u0=(specs[i][1], 0, 0)) for i, g in enumerate(grids)])
    plan.init_equilibrium(np.concatenate(rhos), np.concatenate(us, axis=1))
    steps = 4
    for t in range(steps):
        plan.level_step(0, t, force=t == steps - 1)
    got = plan.download_f(dtype=np.float64)
    F = plan.force()
    o = 0
    for i, g in enumerate(grids):
        okw = dict(tau=specs[i][2], u_inlet=specs[i][1])
        f_or, F_or, _ = oracle_run(oracle, g, "f16s", steps, okw, rhos[i], us[i], force_steps=(steps - 1,))
        want = oracle.to_unshifted("f16s", f_or)
        assert np.array_equal(got[:, o:o + g.n_cells], want), f"domain {i}"
        np.testing.assert_allclose(F[i], F_or[0], rtol=1e-11, atol=1e-14)
        o += g.n_cells
    plan.close()


@pytest.mark.parametrize("case", ["channel_sphere", "periodic"])
def test_fp32_within_1e5_of_literal_reference(oracle, case):
    """The plain FP32 path against the literal reference arithmetic after 40 steps: every population
    within 1e-5 relative to its direction's weight w_k (the scale of the populations of direction k, so
    the 1/216-weight corners are held to the same relative bar as the rest population).  Velocity is a small difference of O(1) populations, so its FP32 rounding
    floor is itself ~1e-5 of max|u| here; the bar for it is accuracy: the library's velocity error
    against an FP64 run of the literal arithmetic is at most that of the literal FP32 arithmetic."""
    cfg = CASES[case]
    grid = uniform_grid(cfg["n"], periodic=cfg["periodic"])
    if cfg["solid"]:
        grid.solid = sphere_mask(grid, *cfg["solid"])
    okw = dict(cfg["okw"])
    rho, u = fields(grid, seed=2, u0=(okw.get("u_inlet", 0.0), 0, 0))
    steps = 40
    f_lit, _, G = oracle_run(oracle, grid, "f32", steps, okw, rho, u, arith=0)
    f_64, _, _ = oracle_run(oracle, grid, "f64", steps, okw, rho, u, arith=0)
    plan, _ = gpu_run(grid, "f32", steps, okw, rho, u)
    got = plan.download_f(dtype=np.float64)
    want = f_lit.astype(np.float64)
    w = oracle.lattice()[7]
    assert (np.abs(got - want) / w[:, None]).max() < 1e-5
    _, _, u_lit = oracle.moments("f32", G, f_lit)
    _, _, u_64 = oracle.moments("f64", G, f_64)
    _, u_gpu = plan.macro()
    fl = grid.inside() if grid.solid is None else (grid.inside() & (grid.solid == 0))
    scale = np.abs(u_64[:, fl]).max()
    err_lit = np.abs(u_lit[:, fl] - u_64[:, fl]).max() / scale
    err_gpu = np.abs(u_gpu[:, fl] - u_64[:, fl]).max() / scale
    assert err_gpu <= 1.25 * err_lit + 1e-7, (err_gpu, err_lit)
    assert np.abs(u_gpu[:, fl] - u_lit[:, fl]).max() / scale < 3e-5
    plan.close()


@pytest.mark.parametrize("storage,pair", [("f32", 0), ("f16s", 0), ("f16s", 1), ("f16s", 2), ("bf16s", 0), ("bf16s", 1),
                                          ("bf16s", 2)])
@pytest.mark.parametrize("face_kernel", [0, 1])
@pytest.mark.parametrize("planes", [1, 2, 4, 8, 16])
def test_interior_pipeline_shapes(oracle, storage, pair, planes, face_kernel):
    """Every launch shape of the pipelined interior kernel gives the same bits (regular interior
    blocks present: 5 x 3 x 3 blocks), the 16-bit forms with one and with two cells per lane."""
    grid = uniform_grid((80, 48, 48))
    grid.solid = sphere_mask(grid, (18, 24, 24), 4.0)
    okw = dict(tau=0.52, u_inlet=0.05)
    rho, u = fields(grid, seed=4, u0=(0.05, 0, 0))
    steps = 3
    f_or, _, _ = oracle_run(oracle, grid, storage, steps, okw, rho, u)
    plan = gpu_plan(grid, storage, okw)
    shape = {"pipe_planes": planes, "overlap": int(planes % 4 == 0), "face_kernel": face_kernel, "graph": 0,
             "pair_kernel": pair}
    plan.set_tuning(**shape)
    assert plan.tuning == shape
    plan.init_equilibrium(rho, u)
    for t in range(steps):
        plan.level_step(0, t)
    compare(oracle, plan, storage, f_or, grid)
    plan.close()


@pytest.mark.parametrize("rules", ["plain", "walls_noise"])
@pytest.mark.parametrize("storage", ["f32", "f16s", "bf16s"])
@pytest.mark.parametrize("planes", [1, 2, 4, 8, 16])
def test_partial_pipeline_shapes(oracle, storage, planes, rules):
    """Blocks that overhang the extent (72 x 40 x 36: partial in every axis) with no solid near take the
    pipelined kernel with the face rules at the extent's last cells; every launch shape gives the
    oracle's bits, the cells beyond the extent included (never written)."""
    grid = uniform_grid((72, 40, 36))
    grid.solid = sphere_mask(grid, (14, 14, 14), 3.0)
    okw = dict(tau=0.53, u_inlet=0.05)
    if rules == "walls_noise":
        okw.update(ymin_wall_bc=1, zwall_bc=2, inlet_turbulence=0.01)
    rho, u = fields(grid, seed=8, u0=(0.05, 0, 0))
    steps = 3
    f_or, _, _ = oracle_run(oracle, grid, storage, steps, okw, rho, u)
    plan = gpu_plan(grid, storage, okw)
    info = plan.level_info()
    # 5 x 3 x 3 blocks; those clear of the sphere's neighbourhood are regular, the ones that overhang
    # the extent in the partial list
    assert info["n_regular"] + info["n_partial"] >= 10 and info["n_partial"] > 0, info
    plan.set_tuning(pipe_planes=planes, overlap=1, face_kernel=0, graph=0)
    plan.init_equilibrium(rho, u)
    for t in range(steps):
        plan.level_step(0, t)
    compare(oracle, plan, storage, f_or, grid)
    plan.close()


@pytest.mark.timeout(300)
def test_partial_blocks_of_a_level_with_layers_beyond_2gib():
    """FP32 1320 x 1000 x 20: a layer of 83 x 63 blocks spans 2.3 GB, beyond the signed 32-bit byte
    offsets of the in-kernel face rules, and the z extent leaves partial blocks in the second layer.
    Such blocks take the one-plane boundary kernel (64-bit addresses): every launch shape of the
    pipelined regular kernel gives the bits of the one-plane kernels."""
    grid = uniform_grid((1320, 1000, 20))
    n = grid.n_cells
    rng = np.random.default_rng(12)
    rho = 1.0 + 0.01 * rng.standard_normal(n)
    u = np.stack([0.05 + 0.01 * rng.standard_normal(n), 0.01 * rng.standard_normal(n), 0.01 * rng.standard_normal(n)])
    okw = dict(tau=0.55, u_inlet=0.05)
    plan = gpu_plan(grid, "f32", okw)
    info = plan.level_info()
    assert info["n_partial"] == 0 and info["n_boundary_regular"] >= 83 * 63, info
    print(f"1320x1000x20 FP32: {n} cells, {info}", flush=True)
    results = []
    for planes in (4, 1):
        plan.set_tuning(pipe_planes=planes, overlap=1, face_kernel=0, graph=0)
        plan.init_equilibrium(rho, u)
        for t in range(3):
            plan.level_step(0, t)
        results.append(plan.download_f(dtype=np.float32))
    plan.close()
    assert np.array_equal(results[0], results[1])
    assert np.isfinite(results[0]).all()


def test_autotune_restores_state():
    grid = uniform_grid((80, 48, 48))
    rho, u = fields(grid, seed=6, u0=(0.05, 0, 0))
    plan = gpu_plan(grid, "f32", dict(tau=0.52, u_inlet=0.05))
    plan.init_equilibrium(rho, u)
    before = plan.download_f(dtype=np.float32)
    best, ms = plan.autotune(4)
    assert best["pipe_planes"] in (1, 2, 4, 8, 16) and best["overlap"] in (0, 1) and plan.tuning == best
    assert len(ms) == 20 and all(v > 0 for v in ms.values())
    assert ms[(best["pipe_planes"], best["overlap"], best["face_kernel"])] == min(ms.values())
    assert np.array_equal(plan.download_f(dtype=np.float32), before)
    plan.close()
