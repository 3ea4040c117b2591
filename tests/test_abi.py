"""CPU-side checks of the C ABI library: it loads, exports every entry point
include/pamhd_mi355x.h declares, and rejects invalid arguments before any
device call.  No kernel runs here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pamhd_mi355x.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^int\s+(pamhd_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from pamhd_amd import _lib
    L = _lib.load()
    names = declared()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.EXPORTS)


def test_abi_version():
    from pamhd_amd import _lib
    assert _lib.load().pamhd_abi_version() == 1


def _desc(cells=(4, 4, 4), periodic=(1, 1, 1), length=(0.25, 0.25, 0.25)):
    from pamhd_amd import _lib
    return _lib.GridDesc((ctypes.c_int32 * 3)(*cells), (ctypes.c_double * 3)(0, 0, 0),
                         (ctypes.c_double * 3)(*length), (ctypes.c_int32 * 3)(*periodic))


@pytest.mark.parametrize("kw,code", [
    (dict(cells=(0, 4, 4)), 16),
    (dict(length=(0.25, -1.0, 0.25)), 16),
    (dict(periodic=(1, 0, 1)), 17),
])
def test_create_rejects_invalid_grids(kw, code):
    from pamhd_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    phys = _lib.Physics(5 / 3, 1.0, 1.0)
    assert L.pamhd_create(ctypes.byref(_desc(**kw)), ctypes.byref(phys), 0, ctypes.byref(h)) == code
    assert h.value is None
    assert _lib.last_error()


@pytest.mark.parametrize("gamma,mu0,prec", [(1.0, 1.0, 0), (5 / 3, 0.0, 0), (5 / 3, 1.0, 7)])
def test_create_rejects_invalid_physics(gamma, mu0, prec):
    from pamhd_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    phys = _lib.Physics(gamma, mu0, 1.0)
    assert L.pamhd_create(ctypes.byref(_desc()), ctypes.byref(phys), prec, ctypes.byref(h)) == 16


def test_batch_argument_checks_without_device():
    from pamhd_amd import _lib
    L = _lib.load()
    # zero faces is a no-op; null buffers are rejected before any launch
    assert L.pamhd_face_flux_batch(3, 0, 0, None, None, None, 5 / 3, 1.0, None, None, None, None) == 0
    assert L.pamhd_face_flux_batch(3, 0, 8, None, None, None, 5 / 3, 1.0, None, None, None, None) == 16
    assert L.pamhd_face_flux_batch(3, 5, 8, 1, 1, None, 5 / 3, 1.0, 1, 1, None, None) == 16


def test_solver_names_follow_the_reference():
    from pamhd_amd import solver
    assert solver.solver_id("rusanov") == 0
    assert solver.solver_id("hll-athena") == 2
    assert solver.solver_id("hlld-athena") == 3
    assert solver.solver_id("roe-athena") == 4
    with pytest.raises(ValueError):
        solver.solver_id("roe_athena")   # the reference's unusable default string


def test_product_path_does_not_use_the_oracle():
    """The package never imports or loads anything under oracle/."""
    pkg = os.path.join(ROOT, "pamhd_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.replace("oracle's", ""), f


def test_plain1d_argument_checks_without_device():
    """pamhd_plain1d_run rejects a null state, fewer than two cells, bad
    constants and unknown solvers before allocating anything"""
    from pamhd_amd import _lib
    L = _lib.load()
    s = (ctypes.c_double * 16)()
    n = ctypes.c_long(-1)
    assert L.pamhd_plain1d_run(2, 0.5, 2.0, 1.0, 0.5, 2, 0.1, 10, None, None, None) == 16
    assert L.pamhd_plain1d_run(1, 0.5, 2.0, 1.0, 0.5, 2, 0.1, 10, s, ctypes.byref(n), None) == 16
    assert L.pamhd_plain1d_run(2, -0.5, 2.0, 1.0, 0.5, 2, 0.1, 10, s, None, None) == 16
    assert L.pamhd_plain1d_run(2, 0.5, 1.0, 1.0, 0.5, 2, 0.1, 10, s, None, None) == 16
    assert L.pamhd_plain1d_run(2, 0.5, 2.0, 1.0, 0.5, 1, 0.1, 10, s, None, None) == 8
    assert L.pamhd_plain1d_run(2, 0.5, 2.0, 1.0, 0.5, 9, 0.1, 10, s, None, None) == 8
    assert n.value == -1
    assert _lib.last_error()
