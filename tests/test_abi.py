"""The C ABI library loads on the CPU and exports every entry point include/oal.h declares (no compute)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "oal.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(oal_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_binding_list():
    from openaerolab_alpha_01_amd import _lib
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib_path = ROOT / "openaerolab_alpha_01_amd" / "liboal.so"
    if not lib_path.exists():
        pytest.fail("liboal.so is not built: run `python -m openaerolab_alpha_01_amd.build`")
    lib = ctypes.CDLL(str(lib_path))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_abi_version_and_strerror_without_gpu():
    from openaerolab_alpha_01_amd import _lib
    lib = _lib.load()
    assert lib.oal_abi_version() == _lib.OAL_ABI_VERSION
    assert lib.oal_strerror(_lib.OAL_ERR_UNSUPPORTED) == b"unsupported feature"


def test_descriptor_validation_without_gpu():
    """A bad descriptor is rejected before any device call."""
    import numpy as np
    from openaerolab_alpha_01_amd import _lib, solver
    from openaerolab_alpha_01_amd.geometry import uniform_grid
    g = uniform_grid((32, 32, 32))
    dom = solver.Domain(extent=g.n, tau=0.4)     # tau <= 1/2 is invalid
    with pytest.raises(_lib.OalError) as e:
        solver.Plan(solver.Level(g, [dom]))
    assert e.value.code == _lib.OAL_ERR_ARG
    # the exact specular z wall reads the image cell at the gather: the AB pattern only
    dom = solver.Domain(extent=g.n, tau=0.6, zwall_bc=1)
    with pytest.raises(_lib.OalError) as e:
        solver.Plan(solver.Level(g, [dom]), stream="aa")
    assert e.value.code == _lib.OAL_ERR_UNSUPPORTED
    with pytest.raises(_lib.OalError) as e:
        solver.Plan(solver.Level(g, [solver.Domain(extent=g.n, tau=0.6, zwall_bc=3)]))
    assert e.value.code == _lib.OAL_ERR_ARG
    # z walls and the transverse freestream act on the same faces; the blend lies in [0, 1]
    tfs = solver.StepParams(transverse_freestream=1)
    with pytest.raises(_lib.OalError) as e:
        solver.Plan(solver.Level(g, [solver.Domain(extent=g.n, tau=0.6, zwall_bc=2)]), params=tfs)
    assert e.value.code == _lib.OAL_ERR_ARG
    with pytest.raises(_lib.OalError) as e:
        solver.Plan(solver.Level(g, [solver.Domain(extent=g.n, tau=0.6)]),
                    params=solver.StepParams(transverse_freestream=1, transverse_freestream_blend=1.5))
    assert e.value.code == _lib.OAL_ERR_ARG
    with pytest.raises(_lib.OalError) as e:
        solver.Plan(solver.Level(g, [solver.Domain(extent=g.n, tau=0.6)]), stream="aa", bb_mode="reference")
    assert e.value.code == _lib.OAL_ERR_UNSUPPORTED


def test_operator_registry():
    from openaerolab_alpha_01_amd.solver import get_collision_kernel
    assert get_collision_kernel(":cumulant") == "cumulant"
    with pytest.raises(ValueError):
        get_collision_kernel("bgk")
    with pytest.raises(NotImplementedError):
        get_collision_kernel("entropic")


def test_refined_plan_validation_without_gpu():
    """Refined-plan descriptors are checked before any device call: nesting, margin, pattern."""
    import numpy as np
    from openaerolab_alpha_01_amd import _lib, solver
    from openaerolab_alpha_01_amd.geometry import BlockGrid, block_octree, build_neighbor_table
    params = solver.StepParams(c_wale=0.0, inlet_turbulence=0.0)
    grids = block_octree((64, 16, 64), [{(1, 0, 1)}, {(2, 0, 2)}], periodic=(True, True, True))

    def level_from(g, keep):
        coords = np.asarray(g.coords)[keep]
        nblk = tuple(int(v) // 16 for v in g.n)
        return BlockGrid(n=g.n, periodic=g.periodic, coords=coords.astype(np.int32),
                         neighbor=build_neighbor_table(coords, nblk, g.periodic))

    def code_of(levels, **kw):
        with pytest.raises(_lib.OalError) as e:
            solver.octree_plan(levels, 0.56, params=kw.pop("params", params), **kw)
        return e.value.code

    # level 1 cut down to the 8 children around the finest block: no margin for level 2's ghosts
    c1 = (np.asarray(grids[1].coords) - 1).tolist()
    core = {(x, y, z) for x in (2, 3) for y in (0, 1) for z in (2, 3)}
    assert code_of([grids[0], level_from(grids[1], [i for i, c in enumerate(c1) if tuple(c) in core]), grids[2]]) \
        == _lib.OAL_ERR_ARG
    # a child without its 7 siblings
    c2 = (np.asarray(grids[2].coords) - 1).tolist()
    assert code_of([grids[0], grids[1], level_from(grids[2], list(range(1, len(c2))))]) == _lib.OAL_ERR_ARG
    # the AA pattern is not offered on refined plans
    levels = [solver.Level(g, [solver.Domain(extent=g.n, tau=t, periodic=g.periodic)])
              for g, t in zip(grids, solver.acoustic_taus(0.56, 3))]
    with pytest.raises(_lib.OalError) as e:
        solver.Plan(levels, stream="aa", params=params)
    assert e.value.code == _lib.OAL_ERR_UNSUPPORTED
