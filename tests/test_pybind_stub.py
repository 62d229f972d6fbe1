"""The reference-side pybind11 binding (diff_fem_amd/pybind/pdm_pybind.cc)
on the CPU: it is built in-tree against include/pdm.h and libpdm.so, exports
the reference's forward_all / backward_all / Environment, INTEGRATION.md shows
the compiled file verbatim, and with no HIP device it fails loudly."""
import ctypes as C
import importlib.util
import os
import re

import numpy as np
import pytest

from diff_fem_amd import _lib, build, scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load():
    path = build.pybind_module_path()
    if not os.path.exists(path):
        build.build_pybind()
    _lib.load()  # torch's HIP runtime first, then libpdm (as for the ctypes binding)
    spec = importlib.util.spec_from_file_location("pdm_pybind", path)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_integration_page_shows_the_compiled_binding():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    src = open(os.path.join(ROOT, "diff_fem_amd", "pybind", "pdm_pybind.cc")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", doc, re.S)
    assert src in blocks


def test_module_exports_the_reference_names():
    m = _load()
    for name in ("Environment", "forward_all", "backward_all", "__version__"):
        assert hasattr(m, name)


def test_environment_without_device_fails_loudly():
    n = C.c_int(-1)
    _lib.load().pdm_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a HIP device is present")
    m = _load()
    d = scene.config1()
    with pytest.raises(RuntimeError, match="no HIP device"):
        m.Environment(5, d.x_rest.ravel().tolist(), d.tets.ravel().tolist(), 24.0, 0.01, 10, [0, 0, -9.81],
                      d.attach_idx.tolist(), 1e20, 0.9999)
