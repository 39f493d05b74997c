"""Pins of the oracle's interpolated bounce-back (linear BFL rule in pull form) and of the sphere q-map.

Expectations come from the mathematics or the physics, not from the oracle's formulas:
  * with q = 1/2 on every solid link the rule is halfway bounce-back (both branches reduce to f_k*):
    the same populations bit for bit, and the force taken at the gather equals the halfway force of
    the step before (the same links, summed in another order);
  * force-driven Poiseuille flow between flat walls placed off the half-link (q = 1/4 and 3/4: both
    branches of the rule): the profile follows the parabola of walls at the q positions, which
    halfway bounce-back misses by about 10 %; the momentum exchange balances the body force
    (bouzidi_kernel.jl:143-148 states the rule, tools/tests/q19_octree_poiseuille.jl:78-113 the 1 %
    friction gate);
  * a rest state stays at rest under every fallback of a q < 1/2 link with no fluid cell behind;
  * the sphere q-map: every kept q puts the wall on the sphere (distance found by bisection along
    the link), the segment before it lies outside the sphere, and q is kept on exactly the links
    from fluid to solid cells (found on the dense mask).
"""
import numpy as np
import pytest

from openaerolab_alpha_01_amd.geometry import (channel_walls, lattice_d3q27, link_solid, sphere_mask, sphere_qmap,
                                               uniform_grid)


def _steps(oracle, G, storage, f, steps, force_at=(), **kw):
    buf = f.copy()
    forces = {}
    for t in range(steps):
        p = oracle.make_params(time_step=t, **kw)
        _, F, _ = oracle.step(storage, G, p, f, out=buf, want_force=t in force_at)
        if t in force_at:
            forces[t] = np.asarray(F, dtype=np.float64)
        f, buf = buf, f
    return f, forces


def _sphere_case():
    grid = uniform_grid((48, 32, 32))
    grid.solid = sphere_mask(grid, (20.3, 16.1, 15.8), 5.2)
    return grid


@pytest.mark.parametrize("storage", ["f64", "f32", "f16s"])
def test_half_q_is_halfway(oracle, storage):
    grid = _sphere_case()
    fluid = (grid.solid == 0) & grid.inside()
    q = np.where(link_solid(grid) & fluid[None, :], 0.5, 0.0).astype(np.float16)
    rho = np.ones(grid.n_cells)
    u = np.zeros((3, grid.n_cells))
    u[0] = np.where(grid.solid != 0, 0.0, 0.05)
    kw = dict(tau=0.56, u_inlet=0.05, arith=oracle.FAST)
    G0, G1 = oracle.Grid(grid), oracle.Grid(grid, qmap=q)
    f0 = oracle.init_eq(storage, G0, rho, u)
    a, Fa = _steps(oracle, G0, storage, f0.copy(), 4, force_at=(2,), **kw)
    b, Fb = _steps(oracle, G1, storage, f0.copy(), 4, force_at=(3,), **kw)
    np.testing.assert_array_equal(a, b)
    assert np.abs(Fa[2]).max() > 0
    np.testing.assert_allclose(Fb[3], Fa[2], rtol=1e-12, atol=1e-12 * np.abs(Fa[2]).max())


def _poiseuille(oracle, delta, use_q, steps=1600, ny=12, tau=0.9, g=1e-6):
    grid = uniform_grid((16, ny, 16), periodic=(True, False, True))
    grid.solid = channel_walls(grid, axis=1)
    fl = (grid.solid == 0) & grid.inside()
    q = np.where(link_solid(grid) & fl[None, :], delta, 0.0).astype(np.float16) if use_q else None
    G = oracle.Grid(grid, qmap=q)
    f = oracle.init_eq("f64", G, np.ones(grid.n_cells), np.zeros((3, grid.n_cells)))
    f, F = _steps(oracle, G, "f64", f, steps, force_at=(steps - 1,), tau=tau, arith=oracle.FAST,
                  body_force=(g, 0.0, 0.0), adaptive_omega4=0)
    _, _, v = oracle.moments("f64", G, f)
    Y = grid.cell_coords()[:, 1] + 0.5
    nu = (np.float32(tau) - 0.5) / 3.0
    g_eff = np.float32(g) * (3.0 - 1.0 / np.float32(tau)) / 2.0   # what the forcing slot delivers per step
    lo, hi = 1.5 - delta, ny - 1.5 + delta                         # walls q cells beyond the outer fluid centres
    exact = g_eff / (2 * nu) * (Y - lo) * (hi - Y)
    err = np.abs(v[0][fl] - exact[fl]).max() / exact[fl].max()
    return err, F[steps - 1][0] / (g_eff * int(fl.sum())) - 1.0


@pytest.mark.parametrize("delta", [0.25, 0.75])
def test_poiseuille_walls_off_the_half_link(oracle, delta):
    err_bfl, fbal = _poiseuille(oracle, delta, True)
    err_hw, _ = _poiseuille(oracle, delta, False)
    # halfway bounce-back puts the walls at the half-link whatever q says
    assert err_hw > 0.08, err_hw
    # the interpolated walls sit where q puts them (the linear rule's viscosity-dependent slip
    # leaves a few per cent at q < 1/2 with tau = 0.9)
    assert err_bfl < 0.03 and err_bfl < 0.3 * err_hw, (err_bfl, err_hw)
    assert abs(fbal) < 1e-6, fbal


@pytest.mark.parametrize("fallback", [0, 1, 2])
def test_rest_state_under_every_fallback(oracle, fallback):
    """one-cell fluid gap between solid rows y = 0 and y = 2: the links of row 1 toward either wall have
    no fluid cell behind, so every q < 1/2 link takes the fallback"""
    grid = uniform_grid((16, 16, 16), periodic=(True, False, True))
    y = grid.cell_coords()[:, 1]
    grid.solid = ((y == 0) | (y == 2) | (y == 15)).astype(np.uint8)
    fl = (grid.solid == 0) & grid.inside()
    q = np.where(link_solid(grid) & fl[None, :], 0.3, 0.0).astype(np.float16)
    G = oracle.Grid(grid, qmap=q)
    f = oracle.init_eq("f64", G, np.ones(grid.n_cells), np.zeros((3, grid.n_cells)))
    f, _ = _steps(oracle, G, "f64", f, 10, tau=0.6, bfl_fallback=fallback)
    w = oracle.lattice()[7]
    np.testing.assert_allclose(f[:, fl], np.repeat(w[:, None], int(fl.sum()), 1), rtol=1e-14, atol=0)


def test_sphere_qmap():
    grid = uniform_grid((40, 36, 32))
    centre, R = np.array([19.7, 18.2, 15.6]), 6.3
    grid.solid = sphere_mask(grid, centre, R)
    q = sphere_qmap(grid, centre, R).astype(np.float64)
    c, _, _, _, _ = lattice_d3q27()
    # links from fluid to solid cells, on the dense mask
    dense = grid.to_dense(grid.solid.astype(np.int64))
    g = grid.cell_coords()
    ins = grid.inside()
    want = np.zeros((27, grid.n_cells), dtype=bool)
    for k in range(27):
        t = g + c[k]
        ok = ins & (t >= 0).all(1) & (t < np.asarray(grid.n)).all(1)
        tt = np.where(ok[:, None], t, 0)
        want[k] = ok & (grid.solid == 0) & (dense[tt[:, 0], tt[:, 1], tt[:, 2]] != 0)
    np.testing.assert_array_equal(link_solid(grid) & ((grid.solid == 0) & ins)[None, :], want)
    np.testing.assert_array_equal(q > 0, want)
    assert (q[want] <= 1.0).all()
    ks, cells = np.nonzero(want)
    assert len(ks) > 500
    p = g[cells] + 0.5
    d = c[ks] / np.linalg.norm(c[ks], axis=1, keepdims=True)
    step = np.linalg.norm(c[ks], axis=1)
    # bisection for the first crossing of the sphere along the link (the end point is inside)
    lo, hi = np.zeros(len(ks)), step.copy()
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        inside = np.linalg.norm(p + mid[:, None] * d - centre, axis=1) <= R
        hi = np.where(inside, mid, hi)
        lo = np.where(inside, lo, mid)
    q_true = hi / step
    # float16 keeps 11 significant bits
    np.testing.assert_allclose(q[ks, cells], q_true, rtol=2.0 ** -11, atol=0)
    for s in (0.25, 0.5, 0.75):
        assert (np.linalg.norm(p + (s * q_true * step)[:, None] * d - centre, axis=1) > R).all()


def bfl_case():
    """channel with a sphere that cuts the y = 0 face, and its q-map (the GPU parity case)"""
    grid = uniform_grid((64, 32, 40))
    centre, R = (22.3, 4.6, 19.7), 6.1
    grid.solid = sphere_mask(grid, centre, R)
    return grid, sphere_qmap(grid, centre, R)


def test_bfl_case_has_every_link_kind():
    """the GPU parity case holds links of both branches and links whose cell behind lies beyond a face"""
    grid, q = bfl_case()
    c, _, _, _, _ = lattice_d3q27()
    qq = q.astype(np.float64)
    assert ((qq > 0.001) & (qq < 0.5)).any() and (qq >= 0.5).any()
    ks, cells = np.nonzero((qq > 0.001) & (qq < 0.5))
    behind = grid.cell_coords()[cells] - c[ks]
    assert (behind[:, 1] < 0).any()
    assert (behind[:, 1] >= 0).any()
