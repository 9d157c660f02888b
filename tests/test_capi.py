"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and rejects unsupported descriptors before touching HIP."""
import os
import re

import numpy as np
import pytest

import pfdd_gl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "pfdd_gl.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(pfdd_gl_\w+)\s*\(", txt, re.M)))


def test_library_exports_header_symbols():
    L = pfdd_gl.load_library()
    names = header_functions()
    assert len(names) >= 16
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(pfdd_gl.EXPORTS)
    assert L.pfdd_gl_version() >= 1


@pytest.mark.parametrize("shape,ns,code", [((10, 16, 16), 3, 2), ((16, 16, 16), 5, 2),
                                          ((16, 16, 16), 13, 2)])
def test_unsupported_descriptors_rejected(shape, ns, code):
    xn = np.tile([0.0, 0.0, 1.0], (ns, 1))
    xb = np.tile([1.0, 0.0, 0.0], (ns, 1))
    with pytest.raises(pfdd_gl.PfddError) as ei:
        pfdd_gl.GLSolver(shape, xn, xb, np.eye(3), 48e9, 1e11, 2e11, 0.8165, 0.01)
    assert ei.value.code == code
    assert "UNSUPPORTED" in str(ei.value)


def test_missing_coefficients_rejected():
    xn = np.tile([0.0, 0.0, 1.0], (3, 1))
    xb = np.tile([1.0, 0.0, 0.0], (3, 1))
    with pytest.raises(pfdd_gl.PfddError) as ei:
        pfdd_gl.GLSolver((16, 16, 16), xn, xb, np.eye(3), 48e9, 1e11, 2e11, 0.8165, 0.01,
                         core_model=pfdd_gl.CORE_SINE)
    assert ei.value.code == 1


def test_missing_library_fails_loudly(tmp_path):
    import importlib
    saved = pfdd_gl._lib
    try:
        pfdd_gl._lib = None
        with pytest.raises(RuntimeError, match="not found"):
            pfdd_gl.load_library(str(tmp_path / "nope.so"))
    finally:
        pfdd_gl._lib = saved
