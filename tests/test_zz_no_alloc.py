"""oal_mass_noalloc, the entry point that allocates nothing, against oal_mass through the C ABI.

Two calls in a row on one GPU, the second on a smaller level (fewer blocks, another storage form,
several domains), each compared bit for bit with oal_mass of the same plan: the no-allocation variant
runs the same kernels with its partials in the library's fixed scratch array, which keeps the first
call's contents beyond the second level's blocks.
"""
import ctypes as C

import numpy as np
import pytest

from openaerolab_alpha_01_amd import _lib, solver
from openaerolab_alpha_01_amd.geometry import sphere_mask, uniform_grid

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """test-side device memory for the entry point's output (the caller owns its outputs)"""

    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so")
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def read(self, n):
        out = np.zeros(n, dtype=np.float64)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(n * 8), 2) == 0  # D2H
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _fields(grid, seed):
    rng = np.random.default_rng(seed)
    n = grid.n_cells
    u = np.stack([0.04 + 0.01 * rng.standard_normal(n), 0.01 * rng.standard_normal(n), 0.01 * rng.standard_normal(n)])
    return 1.0 + 0.01 * rng.standard_normal(n), u


def _check(gpu, plan, out):
    lib = _lib.load()
    nd = len(plan.levels[0].domains)
    want = plan.mass(0)
    print(f"oal_mass_noalloc: {plan.n_cells(0)} cells, {nd} domain(s)", flush=True)
    _lib.check(lib.oal_mass_noalloc(plan.h, 0, out.p, None), "oal_mass_noalloc")
    plan.sync()
    got = out.read(4 * nd).reshape(nd, 4)
    assert np.array_equal(got, want), (got, want)
    return got


def test_mass_noalloc_matches_mass(gpu):
    out = DeviceBuffer(4096)
    try:
        # first call: 64 x 48 x 48 channel with a sphere, FP32, after 3 steps
        g1 = uniform_grid((64, 48, 48))
        g1.solid = sphere_mask(g1, (20.0, 24.0, 24.0), 7.0)
        p1 = solver.uniform_plan(g1, 0.55, storage="f32", params=solver.StepParams(c_wale=0.0, inlet_turbulence=0.0),
                                 u_inlet=0.04)
        p1.init_equilibrium(*_fields(g1, 1))
        p1.step(3)
        m1 = _check(gpu, p1, out)
        assert abs(m1[0, 0] - p1.level_info(0)["n_fluid_cells"]) < 0.05 * m1[0, 0]
        p1.close()
        # second call, smaller: three 32 x 16 x 16 boxes in one level, FP16 storage, after 2 steps
        grids = [uniform_grid((32, 16, 16)) for _ in range(3)]
        p2 = solver.batched_plan(grids, [0.56, 0.6, 0.52], [0.03, 0.05, 0.02], storage="f16s",
                                 params=solver.StepParams(c_wale=0.0, inlet_turbulence=0.0))
        n2 = p2.n_cells(0)
        p2.init_equilibrium(*_fields(type("G", (), {"n_cells": n2})(), 2))
        p2.step(2)
        _check(gpu, p2, out)
        # and again on the same plan: the result does not depend on what the scratch held
        _check(gpu, p2, out)
        p2.close()
    finally:
        out.free()
