"""Pins of the oracle's WALE eddy viscosity (kernel_cumulant.jl:500-535) against the model's published
definition (Nicoud & Ducros 1999), with filter width 1 (lattice units):

    S_ij = (g_ij + g_ji) / 2,   S^d_ij = ((g^2)_ij + (g^2)_ji) / 2 - delta_ij (g^2)_kk / 3,
    nu_t = C_w^2 (S^d:S^d)^{3/2} / ((S:S)^{5/2} + (S^d:S^d)^{5/4}),   g_ij = d u_i / d x_j.

The tensor form is evaluated here with einsum, independently of the oracle's written-out components,
so a transposed operand or a wrong index fails; the model's own properties are checked too: no eddy
viscosity in pure shear, degree-1 homogeneity, invariance under a rotation of the frame.
"""
import ctypes as C

import numpy as np
import pytest


def wale(oracle, g, c_wale=0.2):
    lib = oracle.load()
    g = np.ascontiguousarray(g, dtype=np.float64).reshape(9)
    return lib.oracle_wale_nu_f64(g.ctypes.data_as(C.c_void_p), c_wale)


def wale_reference(g, c_wale=0.2):
    g2 = np.einsum("ik,kj->ij", g, g)
    Sd = 0.5 * (g2 + g2.T) - np.eye(3) * np.trace(g2) / 3.0
    S = 0.5 * (g + g.T)
    op1 = np.einsum("ij,ij->", Sd, Sd)
    op2 = np.einsum("ij,ij->", S, S)
    return c_wale ** 2 * op1 ** 1.5 / (op2 ** 2.5 + op1 ** 1.25)


def test_wale_matches_tensor_definition(oracle):
    rng = np.random.default_rng(5)
    for _ in range(200):
        g = rng.normal(0.0, 1e-2, (3, 3))
        want = wale_reference(g)
        assert abs(wale(oracle, g) - want) <= 1e-12 * want


@pytest.mark.parametrize("i,j", [(0, 1), (0, 2), (1, 2), (2, 0)])
def test_wale_vanishes_in_pure_shear(oracle, i, j):
    g = np.zeros((3, 3))
    g[i, j] = 3e-2
    assert wale(oracle, g) == 0.0


def test_wale_homogeneity_and_frame_invariance(oracle):
    rng = np.random.default_rng(6)
    g = rng.normal(0.0, 1e-2, (3, 3))
    nu = wale(oracle, g)
    assert nu > 0
    for a in (0.5, 2.0, 7.0):
        assert abs(wale(oracle, a * g) / nu - a) < 1e-12
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    assert abs(wale(oracle, q @ g @ q.T) / nu - 1) < 1e-12
