"""CPU checks of the C-ABI library: it loads, exports every symbol
include/pdm.h declares, and its host-side setup (nested dissection, supernodal
symbolic + numeric factorisation, muscle selection) agrees with the oracle."""
import os
import re

import numpy as np
import pytest

from diff_fem_amd import _lib, engine, scene

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "pdm.h")


def test_library_exports_header_symbols():
    L = _lib.load()
    decl = set(re.findall(r"^\s*(?:int|void|const char\*)\s+(pdm_\w+)\s*\(", open(HDR).read(), re.M))
    assert decl and decl == set(_lib.SYMBOLS)
    for name in decl:
        assert hasattr(L, name)
    m = re.search(r"#define PDM_NOALLOC_SCRATCH_BYTES \((\d+)ull << (\d+)\)", open(HDR).read())
    assert m and int(m.group(1)) << int(m.group(2)) == _lib.NOALLOC_SCRATCH_BYTES


def test_create_without_device_fails_loudly():
    import ctypes as C
    n = C.c_int(-1)
    _lib.load().pdm_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(_lib.PdmError, match="no HIP device"):
        engine.Scene(scene.config1())


def test_invalid_mesh_is_rejected():
    d = scene.config1()
    d.tets = d.tets.copy()
    d.tets[0, 0] = 10 ** 6
    with pytest.raises(_lib.PdmError, match="out of range"):
        engine.host_analyse(d)


def test_host_factor_solve_matches_oracle(oracle_mod):
    scenes_root = os.path.join(os.path.dirname(__file__), "golden", "scenes")
    for d in (scene.config1(), scene.load_json_scene("/data/bridgesus/bridgesus.json", scenes_root)):
        o = oracle_mod.OracleScene(d)
        b = np.random.default_rng(0).standard_normal((2, 3 * d.N))
        x = engine.host_solve(d, b)
        for r in range(2):
            xo = o.solve3(b[r])
            assert np.abs(x[r] - xo).max() <= 1e-11 * np.abs(xo).max()


def test_host_analyse_muscle_selection_matches_oracle(oracle_mod):
    x, t = scene.box_grid(8, 3, 3)
    s0, s1, g, ng = scene.straight_fibres(2, 2, x.min(0), x.max(0), nseg=12)
    d = scene.SceneDesc(x_rest=x, tets=t, seg_start=s0, seg_end=s1, seg_group=g, n_groups=ng)
    info = engine.host_analyse(d)
    o = oracle_mod.OracleScene(d, factor=False)
    assert info["n_muscles"] == o.muscles[0].shape[0] > 0
    assert info["n_levels"] >= 2 and info["nnz_factor"] > 0


def test_nested_dissection_factor_is_sparser_than_band(oracle_mod):
    d = scene.config1()
    d.contact = False
    info = engine.host_analyse(d)
    # band envelope of the oracle's RCM order for comparison
    K = oracle_mod.OracleScene(d).dense_K()
    assert info["nnz_factor"] < 0.5 * d.N * d.N
    assert np.allclose(K, K.T)
