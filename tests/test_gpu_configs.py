"""The stated configurations at their stated sizes on the GPU, checked through
properties that hold at any size (the oracle is too slow for the full
horizons): the mixed-adjoint form against the FP64 path on the same inputs,
graph replay against eager launches, and the oracle on a prefix.

Config 2: 30x10x10 muscle body (18,000 tets, 64 fibres, per-tet E), dropped
onto a mu = 0.5 floor, uniform activations, 200 steps x 10 iterations.
Config 5: 64 packed meshes of 2k-60k tets (1.13M tets in one world), ground
and a sphere each, 150 steps.
Runtime caps: a vertex meeting more than 8 distinct obstacles in one step
raises instead of dropping one.
"""
import numpy as np
import pytest

from diff_fem_amd import scene

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _run(eng, s, x0, v0, steps, act, gx, gv):
    xT, vT = s.forward_host(x0, v0, steps, act=act, record=True)
    its = s.iterations().copy()
    cs = [c.copy() for c in s.contact_stats()]
    g = s.backward_host(gx, gv)
    return xT, vT, its, cs, g


def test_config2_mixed_adjoint_200_steps(eng, oracle_mod):
    """Config 2 at 200 steps with contact: the mixed-adjoint forward equals
    the FP64 forward bitwise (positions, PD and contact iteration counts) and
    every rollout's gradient (x0, v0, activations; per-tet E summed) is
    within 1e-4 of the FP64 gradient; the FP64 path holds the oracle's FP64
    bar over the first 40 steps (the drop and the first contacts)."""
    d = scene.config2()
    R, steps = 4, 200
    act = scene.activations_uniform(2, steps, d.n_groups, R)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    rng = np.random.default_rng(22)
    v0 = np.tile(rng.uniform(-0.2, 0.2, (R, 1, 3)), (1, d.N, 1)).reshape(R, -1)
    gx = rng.standard_normal((R, 3 * d.N)) / d.N
    gv = 0.01 * rng.standard_normal((R, 3 * d.N)) / d.N
    s = eng.Scene(d, R=R, precision=eng.PDM_FP64, flags=eng.PDM_HOST_EARLY_EXIT | eng.PDM_USE_GRAPH,
                  checkpoint_every=50)
    x64, v64, i64, c64, g64 = _run(eng, s, x0, v0, steps, act, gx, gv)
    s.set_batch(R, eng.PDM_MIXED_ADJOINT, eng.PDM_HOST_EARLY_EXIT | eng.PDM_USE_GRAPH, 50)
    xma, vma, ima, cma, gma = _run(eng, s, x0, v0, steps, act, gx, gv)
    assert np.array_equal(x64, xma) and np.array_equal(v64, vma)
    assert np.array_equal(i64, ima) and np.array_equal(c64[0], cma[0]) and np.array_equal(c64[1], cma[1])
    assert c64[1].sum() > 0  # the contact loop ran
    for r in range(R):
        assert max(rel(gma[k][r], g64[k][r]) for k in ("x0", "v0", "act")) < 1e-4, r
    assert rel(gma["E"], g64["E"]) < 1e-4
    # FP64 against the oracle over the first 40 steps
    s.set_batch(R, eng.PDM_FP64, eng.PDM_HOST_EARLY_EXIT, 40)
    xs, vs = s.forward_host(x0, v0, 40, act=act[:, :40], record=False)
    o = oracle_mod.OracleScene(d)
    for r in (0, R - 1):
        xo, vo, io = o.rollout(x0[r], v0[r], 40, act=act[r, :40])
        assert np.array_equal(s.iterations()[r], io[:, 0])
        assert rel(xs[r], xo) < 1e-9 and rel(vs[r], vo) < 1e-9
        assert io[:, 1].sum() > 0


def test_config5_stated_size_150_steps(eng):
    """Config 5 as stated: 64 meshes (log-uniform 2k-60k tets, 1.13M tets,
    half with muscles), ground and spheres, 150 steps. Graph replay equals
    eager launches bitwise (FP64); the mixed-adjoint forward equals the FP64
    forward bitwise and its gradients are within 1e-4 of FP64's."""
    d = scene.config5()
    assert d.T > 1_000_000 and len(d.dfobjs or []) in (0, 64) and d.spheres.shape[0] == 64
    steps = 150
    rng = np.random.default_rng(55)
    act = rng.uniform(0, 1, (1, steps, d.n_groups))
    x0 = d.x_rest.ravel()[None].copy()
    v0 = np.zeros_like(x0)
    gx = rng.standard_normal(x0.shape) / d.N
    gv = np.zeros_like(gx)
    s = eng.Scene(d, R=1, precision=eng.PDM_FP64, flags=eng.PDM_HOST_EARLY_EXIT, checkpoint_every=steps)
    xe, ve, ie, ce, ge = _run(eng, s, x0, v0, steps, act, gx, gv)
    assert ce[1].sum() > 0
    s.set_batch(1, eng.PDM_FP64, eng.PDM_HOST_EARLY_EXIT | eng.PDM_USE_GRAPH, steps)
    xg, vg, ig, cg, gg = _run(eng, s, x0, v0, steps, act, gx, gv)
    assert np.array_equal(xe, xg) and np.array_equal(ve, vg) and np.array_equal(ie, ig)
    for k in ("x0", "v0", "E", "act"):
        assert np.array_equal(ge[k], gg[k]), k
    s.set_batch(1, eng.PDM_MIXED_ADJOINT, eng.PDM_HOST_EARLY_EXIT | eng.PDM_USE_GRAPH, steps)
    xm, vm, im, cm, gm = _run(eng, s, x0, v0, steps, act, gx, gv)
    assert np.array_equal(xe, xm) and np.array_equal(ie, im)
    for k in ("x0", "v0", "act", "E"):
        assert rel(gm[k], ge[k]) < 1e-4, k
    s.close()


def test_obstacle_cap_raises(eng):
    """Nine copies of the floor are nine distinct obstacles: a vertex that
    crosses them all in one step exceeds the 8 kept per (vertex, rollout) and
    the forward raises instead of dropping one."""
    x, t = scene.box_grid(2, 1, 1, origin=(0.0, 0.0, 0.001))
    d = scene.SceneDesc(x_rest=x, tets=t, total_mass=1.0, E=np.full(t.shape[0], 1e5), h=0.01, max_iter=10,
                        gravity=(0, 0, -9.81), contact=True, max_iter1=10, mu=0.5,
                        planes=np.tile([[0, 0, 0, 0, 0, 1.0]], (9, 1)))
    s = eng.Scene(d, R=1)
    v0 = np.tile([0.0, 0.0, -1.0], d.N)[None]
    with pytest.raises(eng.PdmError if hasattr(eng, "PdmError") else RuntimeError, match="more than 8"):
        s.forward_host(d.x_rest.ravel()[None], v0, 2)
    # eight copies stay within the cap
    d.planes = np.tile([[0, 0, 0, 0, 0, 1.0]], (8, 1))
    s2 = eng.Scene(d, R=1)
    s2.forward_host(d.x_rest.ravel()[None], v0, 2)


@pytest.fixture(scope="module")
def eng():
    from diff_fem_amd import _lib, engine
    engine.PdmError = _lib.PdmError
    return engine
