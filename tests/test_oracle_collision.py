"""Pins of the oracle's chimera transforms and cumulant collision (one cell, FP64 and FP32).

Every expectation is computed independently by brute force: central moments as explicit sums
sum_k f_k (c_k - u)^a over the 27 directions, cumulants from those moments by their definition.
A wrong index, sign or relaxation rate in the oracle fails one of these.
"""
import itertools

import numpy as np
import pytest

C = np.array([(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)], dtype=float)
W = np.select([(C * C).sum(1) == d for d in range(4)], [8 / 27, 2 / 27, 1 / 54, 1 / 216])


def central(f, u):
    """normalised central moments m[a,b,c] = sum f (cx-ux)^a (cy-uy)^b (cz-uz)^c / rho, a,b,c in 0..2"""
    rho = f.sum()
    d = C - u
    m = np.zeros((3, 3, 3))
    for a, b, c in itertools.product(range(3), repeat=3):
        m[a, b, c] = (f * d[:, 0] ** a * d[:, 1] ** b * d[:, 2] ** c).sum() / rho
    return m


def cumulants4(m):
    """4th-order cumulants of the (2,2,0)-type and (2,1,1)-type from normalised central moments"""
    return {
        "220": m[2, 2, 0] - m[2, 0, 0] * m[0, 2, 0] - 2 * m[1, 1, 0] ** 2,
        "202": m[2, 0, 2] - m[2, 0, 0] * m[0, 0, 2] - 2 * m[1, 0, 1] ** 2,
        "022": m[0, 2, 2] - m[0, 2, 0] * m[0, 0, 2] - 2 * m[0, 1, 1] ** 2,
        "211": m[2, 1, 1] - m[2, 0, 0] * m[0, 1, 1] - 2 * m[1, 1, 0] * m[1, 0, 1],
        "121": m[1, 2, 1] - m[0, 2, 0] * m[1, 0, 1] - 2 * m[1, 1, 0] * m[0, 1, 1],
        "112": m[1, 1, 2] - m[0, 0, 2] * m[1, 1, 0] - 2 * m[1, 0, 1] * m[0, 1, 1],
    }


ARITH = pytest.mark.parametrize("arith", [0, 1], ids=["literal", "fast"])


def random_state(rng, amp=0.03, u=(0.04, -0.02, 0.03)):
    """a population vector near equilibrium at velocity u with a random non-equilibrium part"""
    u = np.asarray(u)
    cu = C @ u
    feq = W * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * u @ u)
    return feq * (1 + amp * rng.standard_normal(27))


def test_chimera_triple_is_central_moments(oracle):
    rng = np.random.default_rng(1)
    for _ in range(50):
        abc = rng.random(3)
        u = rng.uniform(-0.2, 0.2)
        k = oracle.chimera_f64(abc, u, +1)
        c = np.array([-1.0, 0.0, 1.0])
        ref = [abc.sum(), (abc * (c - u)).sum(), (abc * (c - u) ** 2).sum()]
        np.testing.assert_allclose(k, ref, rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(oracle.chimera_f64(k, u, -1), abc, rtol=1e-13, atol=1e-15)


@ARITH
def test_rest_state_is_fixed_point(oracle, arith):
    p = oracle.make_params(0.6, arith=arith)
    out, ru = oracle.collide_cell_f64(W, p)
    np.testing.assert_allclose(out, W, rtol=1e-15, atol=0)
    assert abs(ru[0] - 1) < 4e-16 and np.all(np.abs(ru[1:]) < 1e-16)


@pytest.mark.parametrize("limiter", [0, 1, 2])
@pytest.mark.parametrize("adaptive", [0, 1])
@ARITH
def test_conservation(oracle, limiter, adaptive, arith):
    rng = np.random.default_rng(2)
    for _ in range(20):
        f = random_state(rng)
        p = oracle.make_params(0.53, omega_bulk=1.3, omega_3=0.8, omega_4=1.1, limiter=limiter, adaptive_omega4=adaptive, arith=arith)
        out, ru = oracle.collide_cell_f64(f, p)
        assert abs(out.sum() - f.sum()) < 2e-15
        np.testing.assert_allclose(out @ C, f @ C, atol=2e-16)
        assert abs(ru[0] - f.sum()) < 1e-15
        np.testing.assert_allclose(ru[1:] * ru[0], f @ C, atol=1e-16)


@ARITH
def test_full_relaxation_gives_cumulant_equilibrium(oracle, arith):
    """omega1 = omega_bulk = omega_3 = omega_4 = 1: post-collision normalised central moments are those of
    the product equilibrium, m[a,b,c] = prod over axes of (1, 0, 1/3)[a]."""
    rng = np.random.default_rng(3)
    g1 = np.array([1.0, 0.0, 1.0 / 3.0])
    expect = g1[:, None, None] * g1[None, :, None] * g1[None, None, :]
    for _ in range(10):
        f = random_state(rng, amp=0.05)
        p = oracle.make_params(1.0, adaptive_omega4=0, limiter=0, arith=arith)   # tau = 1: nu = 1/6, omega1 = 1
        out, ru = oracle.collide_cell_f64(f, p)
        u = (f @ C) / f.sum()
        np.testing.assert_allclose(central(out, u), expect, atol=1e-14)


@pytest.mark.parametrize("tau,ob,o3,o4", [(0.52, 1.0, 1.0, 1.0), (0.7, 1.4, 0.6, 1.2), (0.51, 0.8, 1.5, 0.3)])
@ARITH
def test_relaxation_rates_by_moment(oracle, tau, ob, o3, o4, arith):
    """Each moment group relaxes at its own rate, measured by brute-force moments of input and output."""
    rng = np.random.default_rng(4)
    omega1 = 1.0 / (0.5 + 3.0 * ((np.float32(tau) - 0.5) / 3.0))
    for _ in range(10):
        f = random_state(rng, amp=0.04)
        p = oracle.make_params(tau, omega_bulk=ob, omega_3=o3, omega_4=o4, adaptive_omega4=0, limiter=0, arith=arith)
        out, _ = oracle.collide_cell_f64(f, p)
        u = (f @ C) / f.sum()
        mi, mo = central(f, u), central(out, u)
        ob32, o332, o432 = (float(np.float32(v)) for v in (ob, o3, o4))
        # trace toward 1 at omega_bulk, deviators at omega1
        tr_i = mi[2, 0, 0] + mi[0, 2, 0] + mi[0, 0, 2]
        tr_o = mo[2, 0, 0] + mo[0, 2, 0] + mo[0, 0, 2]
        assert abs(tr_o - ((1 - ob32) * tr_i + ob32)) < 1e-13
        assert abs((mo[2, 0, 0] - mo[0, 2, 0]) - (1 - omega1) * (mi[2, 0, 0] - mi[0, 2, 0])) < 1e-7
        assert abs((mo[2, 0, 0] - mo[0, 0, 2]) - (1 - omega1) * (mi[2, 0, 0] - mi[0, 0, 2])) < 1e-7
        for ab in [(1, 1, 0), (1, 0, 1), (0, 1, 1)]:
            assert abs(mo[ab] - (1 - omega1) * mi[ab]) < 1e-7
        # third order central moments (= cumulants) at omega_3; first moments vanish
        for ab in [(2, 1, 0), (1, 2, 0), (2, 0, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2), (1, 1, 1)]:
            assert abs(mo[ab] - (1 - o332) * mi[ab]) < 1e-13
        for ab in [(1, 0, 0), (0, 1, 0), (0, 0, 1)]:
            assert abs(mo[ab]) < 1e-15
        # fourth-order cumulants at omega_4
        ci, co = cumulants4(mi), cumulants4(mo)
        for key in ci:
            assert abs(co[key] - (1 - o432) * ci[key]) < 1e-12


@ARITH
def test_adaptive_omega4_diagonal_rate(oracle, arith):
    """The reference's adaptive rate -(omega1-2)^2 / (lambda omega1) is negative for every omega1 in (0, 2)
    and is clamped to 0.01: the diagonal 4th-order cumulants keep 99 % of their value."""
    rng = np.random.default_rng(5)
    f = random_state(rng, amp=0.04)
    p = oracle.make_params(0.55, adaptive_omega4=1, limiter=0, omega_4=1.0, arith=arith)
    out, _ = oracle.collide_cell_f64(f, p)
    u = (f @ C) / f.sum()
    ci, co = cumulants4(central(f, u)), cumulants4(central(out, u))
    for key in ("220", "202", "022"):
        assert abs(co[key] - (1 - float(np.float32(0.01))) * ci[key]) < 1e-12
    for key in ("211", "121", "112"):
        assert abs(co[key]) < 1e-14


@ARITH
def test_factored_limiter_clamps_offdiagonal(oracle, arith):
    """limiter 1: |c_ab| <= sqrt(c_aa c_bb) after relaxation (kernel_cumulant.jl:726-735)"""
    f = W.copy()
    # a strong xy shear: populations with cx*cy > 0 raised, < 0 lowered
    cxy = C[:, 0] * C[:, 1]
    f = f * (1 + 0.9 * cxy)
    p = oracle.make_params(0.5001, omega_bulk=1.0, adaptive_omega4=0, limiter=1, arith=arith)
    out, _ = oracle.collide_cell_f64(f, p)
    u = (f @ C) / f.sum()
    m = central(out, u)
    assert abs(m[1, 1, 0]) <= np.sqrt(m[2, 0, 0] * m[0, 2, 0]) * (1 + 1e-12)


@ARITH
def test_positivity_limiter(oracle, arith):
    """limiter 2 leaves no negative population when the equilibrium is positive"""
    rng = np.random.default_rng(6)
    hits = 0
    for _ in range(200):
        f = random_state(rng, amp=0.6, u=(0.1, 0.05, -0.08))
        f = np.abs(f)
        p0 = oracle.make_params(0.5005, limiter=0, adaptive_omega4=0, omega_3=1.9, omega_4=1.9, arith=arith)
        out0, _ = oracle.collide_cell_f64(f, p0)
        p2 = oracle.make_params(0.5005, limiter=2, adaptive_omega4=0, omega_3=1.9, omega_4=1.9, arith=arith)
        out2, _ = oracle.collide_cell_f64(f, p2)
        if out0.min() < 0:
            hits += 1
            assert out2.min() >= -1e-15
        assert abs(out2.sum() - f.sum()) < 1e-14
    assert hits > 0


@ARITH
def test_fp32_matches_fp64(oracle, arith):
    rng = np.random.default_rng(9)
    for _ in range(20):
        f = random_state(rng).astype(np.float32)
        p = oracle.make_params(0.55, arith=arith)
        o32, _ = oracle.collide_cell_f32(f, p)
        o64, _ = oracle.collide_cell_f64(f.astype(np.float64), p)
        np.testing.assert_allclose(o32, o64, rtol=0, atol=3e-7)


def test_fast_form_equals_literal_to_rounding(oracle):
    """The library's fused arithmetic and the literal reference arithmetic differ by rounding only:
    FP64 results agree to ~1e-16, FP32 ones to a few ulp of the populations."""
    rng = np.random.default_rng(10)
    for lim in (0, 1, 2):
        for _ in range(50):
            f = random_state(rng, amp=0.05)
            kw = dict(limiter=lim, omega_bulk=1.1, omega_3=0.9, omega_4=1.05, adaptive_omega4=1)
            lit, _ = oracle.collide_cell_f64(f, oracle.make_params(0.53, arith=0, **kw))
            fst, _ = oracle.collide_cell_f64(f, oracle.make_params(0.53, arith=1, **kw))
            np.testing.assert_allclose(fst, lit, rtol=0, atol=5e-16)
            l32, _ = oracle.collide_cell_f32(f.astype(np.float32), oracle.make_params(0.53, arith=0, **kw))
            f32_, _ = oracle.collide_cell_f32(f.astype(np.float32), oracle.make_params(0.53, arith=1, **kw))
            np.testing.assert_allclose(f32_, l32, rtol=0, atol=2e-7)
