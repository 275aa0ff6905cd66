"""The C ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

from conftest import ROOT

HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b((?:shesolve|shefill)_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_entry_points():
    names = declared_functions()
    for required in ("shesolve_create", "shesolve_set_matrix", "shesolve_solve", "shesolve_destroy",
                     "shefill_build", "shefill_refill", "shefill_device_system"):
        assert required in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    import shesolve
    lib = shesolve.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding lists exactly what the headers declare
    assert set(shesolve.EXPORTS) | set(shesolve.FILL_EXPORTS) == set(declared_functions())


def test_default_params_and_argument_errors_without_device():
    import shesolve
    p = shesolve.default_params()
    assert p.tol == 1e-10 and p.max_iters == 1000 and p.precond == shesolve.PRECOND_ILU0
    assert p.precision == shesolve.FP64 and p.inner_tol == 1e-2
    lib = shesolve.lib()
    # NULL handle pointer is argument 1
    assert lib.shesolve_create(None, 0, None) == 1
    assert lib.shesolve_destroy(None) == 1
    import numpy as np
    z = np.zeros(1)
    assert lib.shesolve_solve(None, z, z, 0, None) == 1


def test_create_reports_missing_device_cleanly():
    import shesolve
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a HIP device is present")
    except Exception:
        pass
    h = ctypes.c_void_p()
    rc = shesolve.lib().shesolve_create(ctypes.byref(h), 0, None)
    assert rc < 0
    assert b"device" in shesolve.lib().shesolve_last_error()
