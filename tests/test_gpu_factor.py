"""Device re-factorisation (pdm_refactor(s, 0): what scene creation and
pdm_set_material run) against the same factorisation in host code
(pdm_refactor(s, 1)) on the same scene and material: the W blocks, solves,
the regularisation ladder's rung, and a material update.

Bar: W and solves within 1e-12 relative (max-norm). The device kernels fuse
multiply-adds where the host code rounds products and sums separately; every
sum runs over the same terms in the same order. The assembled matrix values
are checked bit for bit on the host when the device plan is built.
"""
import time

import numpy as np
import pytest

from diff_fem_amd import scene

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def eng():
    from diff_fem_amd import engine
    return engine


def muscle_scene():
    x, t = scene.box_grid(6, 3, 3)
    s0, s1, g, ng = scene.straight_fibres(2, 2, x.min(0), x.max(0), nseg=10)
    rng = np.random.default_rng(11)
    return scene.SceneDesc(x_rest=x, tets=t, total_mass=8.0, E=rng.uniform(5e4, 2e5, t.shape[0]), nu_fwd=0.4,
                           nu_ctor=np.full(t.shape[0], 0.4), attach_idx=np.arange(16, dtype=np.int32),
                           attach_w=1e20, seg_start=s0, seg_end=s1, seg_group=g, n_groups=ng, k_muscle=5e5,
                           h=0.01, max_iter=10, gravity=(0, 0, -9.81))


SCENES = {"config1": scene.config1, "config2": scene.config2, "muscle": muscle_scene, "config4": scene.config4}


def _compare(s, rng, R):
    Wd = s.factor_w()
    reg_d = s.info()["reg_used"]
    b = rng.standard_normal((R, 3 * s.N))
    xd = s.solve_host(b)
    s.refactor(on_host=True)
    Wh = s.factor_w()
    reg_h = s.info()["reg_used"]
    xh = s.solve_host(b)
    s.refactor()
    assert reg_d == reg_h
    assert rel(Wd, Wh) < 1e-12
    assert rel(xd, xh) < 1e-12
    assert np.array_equal(s.factor_w(), Wd)  # the device path is deterministic


@pytest.mark.parametrize("name", list(SCENES))
def test_device_factor_matches_host(eng, name):
    d = SCENES[name]()
    R = 3
    s = eng.Scene(d, R=R)
    rng = np.random.default_rng(31)
    _compare(s, rng, R)
    # a material update re-factorises on the device
    s.set_material(rng.uniform(3e4, 3e5, d.T), 0.3)
    _compare(s, rng, R)
    s.close()


def test_device_factor_regularisation_ladder(eng):
    """A negative modulus makes K_s indefinite: both paths climb the same
    rungs of the ladder and land on the same factor."""
    d = scene.config1()
    s = eng.Scene(d, R=1)
    s.set_material(np.full(d.T, -2e7), 0.4)
    assert s.info()["reg_used"] > 0
    _compare(s, np.random.default_rng(32), 1)
    s.close()


def test_device_factor_config5(eng):
    """Config 5's 64-mesh world (1.13M tets, 7253 supernodes): device against
    host factor, and the time of each."""
    d = scene.config5()
    s = eng.Scene(d, R=1)
    t0 = time.perf_counter()
    s.refactor()
    t_dev = time.perf_counter() - t0
    Wd = s.factor_w()
    t0 = time.perf_counter()
    s.refactor(on_host=True)
    t_host = time.perf_counter() - t0
    Wh = s.factor_w()
    print(f"config5 re-factorisation: device {1e3 * t_dev:.1f} ms, host {1e3 * t_host:.1f} ms", flush=True)
    assert rel(Wd, Wh) < 1e-12
    s.close()
