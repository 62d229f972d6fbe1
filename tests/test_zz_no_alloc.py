"""pdm_solve_noalloc (include/pdm.h): the solve of pdm_solve_host on device
buffers with its working memory in the library's fixed per-GPU array.

Checked through the C ABI against pdm_solve_host on the same right-hand
sides (FP64: 1e-12 relative max-norm), on two calls in a row where the second
has fewer rollouts than the first, on a one-level scene (config 1) and a
multi-level one (config 2), with b and x aliased, and the size limit's error.
"""
import ctypes as C

import numpy as np
import pytest

from diff_fem_amd import scene

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _noalloc(L, s, b_dev, x_dev, R):
    import torch
    from diff_fem_amd._lib import check
    check(L.pdm_solve_noalloc(s.h, int(R), C.c_void_p(b_dev.data_ptr()), C.c_void_p(x_dev.data_ptr())))
    torch.cuda.synchronize()
    return x_dev[:R].cpu().numpy()


@pytest.mark.parametrize("cfg,R0,R1", [("config1", 6, 2), ("config2", 8, 3)])
def test_noalloc_solve_matches_solve_host(cfg, R0, R1):
    import torch
    from diff_fem_amd import _lib, engine
    L = _lib.load()
    d = getattr(scene, cfg)()
    s = engine.Scene(d, R=R0)
    print("noalloc solve:", cfg, "R", R0, "then", R1, flush=True)
    rng = np.random.default_rng(3)
    b0 = rng.standard_normal((R0, 3 * d.N))
    b1 = rng.standard_normal((R0, 3 * d.N))
    ref0 = s.solve_host(b0)
    ref1 = s.solve_host(b1)
    dev = torch.device("cuda:0")
    bd = torch.from_numpy(b0).to(dev)
    xd = torch.full((R0, 3 * d.N), np.nan, dtype=torch.float64, device=dev)
    x0 = _noalloc(L, s, bd, xd, R0)
    assert rel(x0, ref0) < 1e-12
    # second call, fewer rollouts, different data: the scratch keeps the first
    # call's contents, which must not leak into the result
    bd = torch.from_numpy(b1[:R1].copy()).to(dev)
    xd = torch.full((R1, 3 * d.N), np.nan, dtype=torch.float64, device=dev)
    x1 = _noalloc(L, s, bd, xd, R1)
    assert rel(x1, ref1[:R1]) < 1e-12
    # b and x may alias
    bd = torch.from_numpy(b0[:R1].copy()).to(dev)
    x2 = _noalloc(L, s, bd, bd, R1)
    assert rel(x2, ref0[:R1]) < 1e-12
    s.close()


def test_noalloc_rejects_sizes_beyond_scratch():
    import torch
    from diff_fem_amd import _lib, engine
    L = _lib.load()
    d = scene.config1()
    s = engine.Scene(d, R=1)
    bd = torch.zeros(3 * d.N, dtype=torch.float64, device="cuda:0")
    R = _lib.NOALLOC_SCRATCH_BYTES // (8 * 3 * d.N) + 1  # the right-hand sides alone exceed it
    rc = L.pdm_solve_noalloc(s.h, int(R), C.c_void_p(bd.data_ptr()), C.c_void_p(bd.data_ptr()))
    assert rc != 0
    assert b"PDM_NOALLOC_SCRATCH_BYTES" in L.pdm_last_error()
    s.close()
