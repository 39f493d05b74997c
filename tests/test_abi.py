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
    dom = solver.Domain(extent=g.n, tau=0.6, zwall_bc=1)
    with pytest.raises(_lib.OalError) as e:
        solver.Plan(solver.Level(g, [dom]))
    assert e.value.code == _lib.OAL_ERR_UNSUPPORTED
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
