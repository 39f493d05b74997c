"""The oracle and the library against the committed golden fixtures (tests/golden/, written by
make_golden.py from the oracle in FP64).  CPU: the oracle reproduces every fixture bit for bit, so a
change of the oracle's arithmetic shows here first.  GPU: the FP64 library reproduces the periodic
box's populations bit for bit and the sphere channel's force to 1e-11 relative."""
import os
import sys

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import make_golden as mg  # noqa: E402


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def test_collision_cells(oracle):
    z = load("collision_cells.npz")
    lat = oracle.lattice()
    w = lat[7]
    p_lit = oracle.make_params(float(z["tau"]), arith=oracle.LITERAL)
    p_fast = oracle.make_params(float(z["tau"]), arith=oracle.FAST)
    for i, (rho, u, neq) in enumerate(mg.collision_inputs()):
        cu = lat[0] * u[0] + lat[1] * u[1] + lat[2] * u[2]
        f = w * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * float(u @ u)) + neq * w
        np.testing.assert_array_equal(f, z["f_in"][i])
        np.testing.assert_array_equal(oracle.collide_cell_f64(f, p_lit)[0], z["literal"][i])
        np.testing.assert_array_equal(oracle.collide_cell_f64(f, p_fast)[0], z["fast"][i])
        # the two forms differ by rounding only
        np.testing.assert_allclose(z["literal"][i], z["fast"][i], rtol=0, atol=1e-15)


def test_periodic16(oracle):
    z = load("periodic16.npz")
    g, rho, u, okw = mg.periodic16()
    f, _ = mg.run_steps(oracle, g, rho, u, okw, int(z["steps"]))
    np.testing.assert_array_equal(f[:, z["cells"]], z["sampled"])
    np.testing.assert_array_equal(f.sum(1), z["sums"])
    np.testing.assert_array_equal((f * f).sum(1), z["sq"])


def test_sphere32(oracle):
    z = load("sphere32.npz")
    g, rho, u, okw = mg.sphere32()
    _, F = mg.run_steps(oracle, g, rho, u, okw, int(z["steps"]), force_step=int(z["steps"]) - 1)
    np.testing.assert_array_equal(F, z["force"])


def _gpu_plan(g, okw, storage="f64"):
    from openaerolab_alpha_01_amd import solver
    return solver.uniform_plan(g, okw["tau"], storage=storage, u_inlet=okw.get("u_inlet", 0.0),
                               params=solver.StepParams(c_wale=0.0, inlet_turbulence=0.0))


@pytest.mark.gpu
def test_gpu_periodic16():
    z = load("periodic16.npz")
    g, rho, u, okw = mg.periodic16()
    with _gpu_plan(g, okw) as plan:
        plan.init_equilibrium(rho, u)
        plan.step(int(z["steps"]))
        f = plan.download_f(dtype=np.float64)
    np.testing.assert_array_equal(f[:, z["cells"]], z["sampled"])
    np.testing.assert_array_equal(f.sum(1), z["sums"])
    np.testing.assert_array_equal((f * f).sum(1), z["sq"])


@pytest.mark.gpu
def test_gpu_sphere32():
    z = load("sphere32.npz")
    g, rho, u, okw = mg.sphere32()
    with _gpu_plan(g, okw) as plan:
        plan.init_equilibrium(rho, u)
        plan.step(int(z["steps"]), force=True)
        F = plan.force()[0]
    np.testing.assert_allclose(F, z["force"], rtol=1e-11, atol=1e-11 * np.abs(z["force"]).max())
