"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle.

Error measure: relative max-norm, ||a - b||_inf / ||b||_inf.
Bars: 1e-9 for the plain FP64 path (positions and gradients), 1e-5 on
positions and 1e-4 on gradients for the mixed path (FP32 local step).
"""
import copy

import numpy as np
import pytest

from diff_fem_amd import scene

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


@pytest.fixture(scope="module")
def eng():
    from diff_fem_amd import engine
    return engine


def small_muscle_scene():
    x, t = scene.box_grid(6, 3, 3)
    s0, s1, g, ng = scene.straight_fibres(2, 2, x.min(0), x.max(0), nseg=10)
    rng = np.random.default_rng(11)
    return scene.SceneDesc(x_rest=x, tets=t, total_mass=8.0, E=rng.uniform(5e4, 2e5, t.shape[0]), nu_fwd=0.4,
                           nu_ctor=np.full(t.shape[0], 0.4), attach_idx=np.arange(16, dtype=np.int32),
                           attach_w=1e20, seg_start=s0, seg_end=s1, seg_group=g, n_groups=ng, k_muscle=5e5,
                           h=0.01, max_iter=10, gravity=(0, 0, -9.81))


def test_solve_matches_oracle(eng, oracle_mod):
    d = scene.config1()
    s = eng.Scene(d, R=2)
    o = oracle_mod.OracleScene(d)
    rng = np.random.default_rng(0)
    b = rng.standard_normal((2, 3 * d.N))
    x = s.solve_host(b)
    for r in range(2):
        assert rel(x[r], o.solve3(b[r])) < 1e-11


def test_solve_shapes_do_not_change_results(eng):
    """The autotuner only re-partitions columns and work between blocks:
    the tuned and the default launch shapes give bitwise equal solves."""
    d = scene.config2()
    rng = np.random.default_rng(5)
    b = rng.standard_normal((24, 3 * d.N))
    s1 = eng.Scene(d, R=24)
    s2 = eng.Scene(d, R=24, flags=eng.PDM_HOST_EARLY_EXIT | eng.PDM_NO_AUTOTUNE)
    x1 = s1.solve_host(b)
    x2 = s2.solve_host(b)
    assert np.array_equal(x1, x2)
    assert len(s1.solve_shapes()) == s1.info()["n_levels"]
    # every plan: subtree cut levels and column chains
    for cut in range(0, min(4, s1.info()["n_levels"])):
        for chains in (1, 2, 3):
            s1.set_solve_plan(cut, chains)
            assert np.array_equal(s1.solve_host(b), x2), (cut, chains)


@pytest.mark.parametrize("which", ["config2", "packed"])
def test_dense_blocks_match_oracle(eng, oracle_mod, which):
    """The top-separator blocks solved as one dense product per elimination
    tree (one tree; four trees of a packed world) and the dense leaf subtrees
    below them, against the level sweeps throughout (max_rows 0) and the
    oracle's LDL^T solve, with several block sizes (leaves from none to whole
    trees), and again after a material change re-forms the dense operators
    on the device."""
    d = scene.config2() if which == "config2" else packed_scene()
    R = 3
    s = eng.Scene(d, R=R)
    o = oracle_mod.OracleScene(d)
    rng = np.random.default_rng(7)
    b = rng.standard_normal((R, 3 * d.N))
    info = s.info()
    assert info["top_rows"] > 0 and info["top_blocks"] == (1 if which == "config2" else 4)
    assert info["top_rows"] % 32 == 0
    x = s.solve_host(b)
    for r in range(R):
        assert rel(x[r], o.solve3(b[r])) < 1e-11
    s.set_dense_top(0)
    assert s.info()["top_rows"] == 0
    x0 = s.solve_host(b)
    assert rel(x0, x) < 1e-12
    for cap in (200, 700, 100000):
        s.set_dense_top(cap)
        assert rel(s.solve_host(b), x) < 1e-12, cap
    # dense leaf subtrees below (small) top blocks: off, small, default, whole trees
    s.set_dense_top(200)
    for leaves in (0, 128, 512, 100000):
        s.set_dense_leaves(leaves)
        assert (s.info()["leaf_blocks"] > 0) == (leaves > 0)
        assert rel(s.solve_host(b), x) < 1e-12, leaves
    s.set_dense_top(1536)
    s.set_dense_leaves(512)
    # new material: the device factorisation forms Z again
    s.set_dense_top(1536)
    E2 = d.E * rng.uniform(0.5, 2.0, d.E.shape)
    s.set_material(E2)
    d2 = copy.copy(d)
    d2.E = E2
    o2 = oracle_mod.OracleScene(d2)
    x2 = s.solve_host(b)
    for r in range(R):
        assert rel(x2[r], o2.solve3(b[r])) < 1e-11
    s.close()


def test_force_matches_oracle(eng, oracle_mod):
    d = small_muscle_scene()
    s = eng.Scene(d, R=2)
    o = oracle_mod.OracleScene(d)
    assert s.info()["n_muscles"] == o.muscles[0].shape[0]
    rng = np.random.default_rng(1)
    x = d.x_rest.ravel()[None, :] + 0.003 * rng.standard_normal((2, 3 * d.N))
    act = rng.uniform(0, 1, (2, d.n_groups))
    Jd = s.force_host(x, act)
    for r in range(2):
        Jo, _ = o.force_terms(x[r], act[r])
        assert rel(Jd[r], Jo) < 1e-12


def test_forward_config1_fp64_host_exit(eng, oracle_mod):
    d = scene.config1()
    s = eng.Scene(d, R=1)
    o = oracle_mod.OracleScene(d)
    x0 = d.x_rest.ravel().copy()
    v0 = np.zeros_like(x0)
    steps = 100
    xT, vT = s.forward_host(x0, v0, steps)
    xo, vo, its = o.rollout(x0, v0, steps)
    assert np.array_equal(s.iterations()[0], its[:, 0])
    assert rel(xT[0], xo) < 1e-9
    assert rel(vT[0], vo) < 1e-9


def test_forward_batched_device_masking(eng, oracle_mod):
    d = scene.config1()
    R = 3
    s = eng.Scene(d, R=R, flags=0)
    o = oracle_mod.OracleScene(d)
    rng = np.random.default_rng(2)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    v0 = 0.05 * rng.standard_normal((R, 3 * d.N))
    v0[:, :] *= np.repeat((np.abs(d.x_rest[:, 0]) > 1e-9)[None, :, None], 3, axis=2).reshape(1, -1)
    steps = 40
    xT, vT = s.forward_host(x0, v0, steps)
    its = s.iterations()
    for r in range(R):
        xo, vo, io = o.rollout(x0[r], v0[r], steps)
        assert np.array_equal(its[r], io[:, 0])
        assert rel(xT[r], xo) < 1e-9 and rel(vT[r], vo) < 1e-9


def test_forward_mixed_precision(eng, oracle_mod):
    d = small_muscle_scene()
    R = 2
    s = eng.Scene(d, R=R, precision=eng.PDM_MIXED, flags=0)
    o = oracle_mod.OracleScene(d)
    steps = 30
    act = np.random.default_rng(3).uniform(0, 1, (R, steps, d.n_groups))
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    v0 = np.zeros_like(x0)
    xT, vT = s.forward_host(x0, v0, steps, act=act)
    for r in range(R):
        xo, vo, _ = o.rollout(x0[r], v0[r], steps, act=act[r])
        assert rel(xT[r], xo) < 1e-5


def _grad_case(eng, oracle_mod, d, precision, steps, R, tol_x, tol_g, ckpt=4, x_jitter=0.0):
    s = eng.Scene(d, R=R, precision=precision, flags=0, checkpoint_every=ckpt)
    o = oracle_mod.OracleScene(d)
    rng = np.random.default_rng(4)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    if x_jitter:
        x0 = x0 + x_jitter * rng.standard_normal(x0.shape)
    v0 = 0.01 * rng.standard_normal((R, 3 * d.N))
    act = rng.uniform(0, 1, (R, steps, d.n_groups)) if d.n_groups else None
    xT, vT = s.forward_host(x0, v0, steps, act=act, record=True)
    gx = rng.standard_normal((R, 3 * d.N))
    gv = 0.01 * rng.standard_normal((R, 3 * d.N))
    g = s.backward_host(gx, gv)
    gE_sum = np.zeros(d.T)
    nu_sum = 0.0
    for r in range(R):
        xo, vo, _ = o.rollout(x0[r], v0[r], steps, act=None if act is None else act[r], record=True)
        assert rel(xT[r], xo) < tol_x
        go, rc = o.backward(gx[r], gv[r])
        assert rc == 0
        assert rel(g["x0"][r], go["x0"]) < tol_g
        assert rel(g["v0"][r], go["v0"]) < tol_g
        if act is not None:
            assert rel(g["act"][r], go["act"]) < tol_g
        gE_sum += go["E"]
        nu_sum += go["nu"]
    assert rel(g["E"], gE_sum) < tol_g
    assert abs(g["nu"] - nu_sum) <= tol_g * abs(nu_sum)


def test_gradient_fp64_config1(eng, oracle_mod):
    _grad_case(eng, oracle_mod, scene.config1(), eng.PDM_FP64, steps=12, R=2, tol_x=1e-9, tol_g=1e-9)


def test_gradient_fp64_muscles(eng, oracle_mod):
    _grad_case(eng, oracle_mod, small_muscle_scene(), eng.PDM_FP64, steps=9, R=2, tol_x=1e-9, tol_g=1e-9)


def test_gradient_mixed_muscles(eng, oracle_mod):
    _grad_case(eng, oracle_mod, small_muscle_scene(), eng.PDM_MIXED, steps=9, R=2, tol_x=1e-5, tol_g=1e-4)


def contact_scene(mu=0.5, sphere=False):
    x, t = scene.box_grid(5, 2, 2, origin=(0, 0, 0.015))
    sp = np.array([[0.12, 0.05, -0.19, 0.2]]) if sphere else None
    return scene.SceneDesc(x_rest=x, tets=t, total_mass=2.0, E=np.full(t.shape[0], 8e4), h=0.01, max_iter=10,
                           gravity=(0, 0, -9.81), contact=True, max_iter1=10, mu=mu,
                           planes=np.array([[0, 0, 0, 0, 0, 1.0]]), spheres=sp)


@pytest.mark.parametrize("sphere", [False, True])
def test_contact_forward_matches_oracle(eng, oracle_mod, sphere):
    d = contact_scene(sphere=sphere)
    R = 2
    s = eng.Scene(d, R=R, flags=0)
    o = oracle_mod.OracleScene(d)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    v0 = np.zeros_like(x0)
    v0[1, 0::3] = 0.3  # second rollout slides along x
    steps = 40
    xT, vT = s.forward_host(x0, v0, steps)
    passes, jit = s.contact_stats()
    assert passes.sum() > 0
    for r in range(R):
        xo, vo, io = o.rollout(x0[r], v0[r], steps)
        assert np.array_equal(s.iterations()[r], io[:, 0])
        assert np.array_equal(passes[r], io[:, 2]) and np.array_equal(jit[r], io[:, 1])
        assert rel(xT[r], xo) < 1e-9 and rel(vT[r], vo) < 1e-9


@pytest.mark.parametrize("sphere", [False, True])
def test_contact_gradient_matches_oracle(eng, oracle_mod, sphere):
    """Adjoint through the contact phase (branches frozen, sphere normals
    differentiated, as in the oracle's reverse sweep) at the FP64 bar."""
    d = contact_scene(sphere=sphere)
    R = 2
    s = eng.Scene(d, R=R, flags=0, checkpoint_every=5)
    o = oracle_mod.OracleScene(d)
    rng = np.random.default_rng(13)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    v0 = np.zeros_like(x0)
    v0[1, 0::3] = 0.3
    steps = 14
    xT, vT = s.forward_host(x0, v0, steps, record=True)
    gx = rng.standard_normal((R, 3 * d.N))
    gv = 0.01 * rng.standard_normal((R, 3 * d.N))
    g = s.backward_host(gx, gv)
    gE = np.zeros(d.T)
    for r in range(R):
        xo, vo, io = o.rollout(x0[r], v0[r], steps, record=True)
        assert io[:, 1].sum() > 0
        go, _ = o.backward(gx[r], gv[r])
        assert rel(g["x0"][r], go["x0"]) < 1e-9
        assert rel(g["v0"][r], go["v0"]) < 1e-9
        gE += go["E"]
    assert rel(g["E"], gE) < 1e-9


@pytest.mark.parametrize("max_iter1", [10, 3])
def test_sphere_normal_gradient_matches_oracle(eng, oracle_mod, max_iter1):
    """Sphere contact normals move with the positions their detection read
    (start-of-step positions and a contact iterate, through the crossing
    time); the oracle's reverse sweep, pinned against central differences,
    carries that derivative. A tumbling bar on a small sphere, two rollouts
    whose contact passes end differently (out of iterations or converged):
    FP64 bar on positions, iteration counts and gradients."""
    from contact_scenes import sphere_tumble_scene
    d, v = sphere_tumble_scene(max_iter1=max_iter1)
    R = 2
    s = eng.Scene(d, R=R, flags=0, checkpoint_every=6)
    o = oracle_mod.OracleScene(d)
    rng = np.random.default_rng(21)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    v0 = np.stack([v, 0.8 * v])
    steps = 12
    xT, vT = s.forward_host(x0, v0, steps, record=True)
    passes, jit = s.contact_stats()
    gx = rng.standard_normal((R, 3 * d.N))
    gv = 0.01 * rng.standard_normal((R, 3 * d.N))
    g = s.backward_host(gx, gv)
    gE = np.zeros(d.T)
    for r in range(R):
        xo, vo, io = o.rollout(x0[r], v0[r], steps, record=True)
        assert np.array_equal(passes[r], io[:, 2]) and np.array_equal(jit[r], io[:, 1])
        assert rel(xT[r], xo) < 1e-9
        go, _ = o.backward(gx[r], gv[r])
        assert rel(g["x0"][r], go["x0"]) < 1e-9, r
        assert rel(g["v0"][r], go["v0"]) < 1e-9, r
        gE += go["E"]
    assert passes.max() >= 2
    assert rel(g["E"], gE) < 1e-9


@pytest.mark.parametrize("which,ckpt,prec", [("muscle", 20, 0), ("contact", 20, 0), ("contact", 25, 0),
                                              ("contact", 25, 1)])
def test_graph_replay_matches_eager(eng, which, ckpt, prec):
    """PDM_USE_GRAPH replays one captured step block per step (device-side
    early exit): untaped steps one graph, taped steps another whose tape slot
    comes from a device counter. Every replay must leave the same state as the
    eager path: positions, velocities and per-step iteration counts are
    compared after two forward calls (the second re-captures), and the adjoint
    that follows a graph forward (which replays the taped graph for the
    recomputed segments) must equal the eager one. ckpt = steps tapes the
    whole rollout in the forward pass, as the benchmark does."""
    d = small_muscle_scene() if which == "muscle" else contact_scene()
    R, steps = 3, 25
    rng = np.random.default_rng(17)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    v0 = 0.02 * rng.standard_normal((R, 3 * d.N))
    act = rng.uniform(0, 1, (R, steps, d.n_groups)) if d.n_groups else None
    eager = eng.Scene(d, R=R, precision=prec, flags=eng.PDM_HOST_EARLY_EXIT, checkpoint_every=ckpt)
    graph = eng.Scene(d, R=R, precision=prec, flags=eng.PDM_USE_GRAPH | eng.PDM_HOST_EARLY_EXIT,
                      checkpoint_every=ckpt)
    xa, va = eager.forward_host(x0, v0, steps, act=act, record=True)
    for call in range(2):
        xb, vb = graph.forward_host(x0, v0, steps, act=act, record=True)
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), call
        assert np.array_equal(eager.iterations(), graph.iterations()), call
    assert graph.graph_launches() == 2 * steps and eager.graph_launches() == 0
    gx = rng.standard_normal((R, 3 * d.N))
    gv = 0.01 * rng.standard_normal((R, 3 * d.N))
    ga = eager.backward_host(gx, gv)
    gb = graph.backward_host(gx, gv)
    for k in ("x0", "v0", "E"):
        assert np.array_equal(ga[k], gb[k]), k
    # the reverse sweep recomputes every segment but the last from its checkpoint
    assert graph.graph_launches() == 2 * steps + (steps - 1) // ckpt * ckpt


def packed_scene():
    return scene.config5(seed=9, n_meshes=4, tmin=300, tmax=1500)


def test_packed_forward_matches_oracle(eng, oracle_mod):
    """Four uneven meshes as the bodies of one world (per-body mass and nu,
    block-diagonal factor), with muscles, ground and spheres: FP64 bar."""
    d = packed_scene()
    R, steps = 2, 30
    s = eng.Scene(d, R=R, flags=0)
    o = oracle_mod.OracleScene(d)
    rng = np.random.default_rng(21)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    v0 = 0.05 * rng.standard_normal((R, 3 * d.N))
    act = rng.uniform(0, 1, (R, steps, d.n_groups))
    xT, vT = s.forward_host(x0, v0, steps, act=act)
    passes, jit = s.contact_stats()
    for r in range(R):
        xo, vo, io = o.rollout(x0[r], v0[r], steps, act=act[r])
        assert np.array_equal(s.iterations()[r], io[:, 0])
        assert np.array_equal(passes[r], io[:, 2]) and np.array_equal(jit[r], io[:, 1])
        assert rel(xT[r], xo) < 1e-9 and rel(vT[r], vo) < 1e-9


@pytest.mark.parametrize("x_jitter", [0.0, 2e-3])
def test_packed_gradient_matches_oracle(eng, oracle_mod, x_jitter):
    """FP64 bar on the packed world, from exact rest and from a state jittered
    by 2 mm. Near rest F has nearly equal singular values; the local-step VJP
    carries the divided differences of the volume projection exactly, so its
    result does not depend on how either SVD splits a near-degenerate
    singular subspace."""
    d = packed_scene()
    _grad_case(eng, oracle_mod, d, eng.PDM_FP64, steps=6, R=2, tol_x=1e-9, tol_g=1e-9, ckpt=4, x_jitter=x_jitter)


@pytest.mark.parametrize("contact,steps,eps_a,eps_e,tol_e", [(False, 200, 1e-4, 1e-3, 1e-3), (True, 50, 1e-6, 1e-6, 1e-3)])
def test_config3_gradient_vs_central_differences(eng, contact, steps, eps_a, eps_e, tol_e):
    """Config 3: the 30x10x10 muscle box (18,000 tets, 64 fibres), parameters
    per-tet E and activations a[steps x 64], loss = |centroid(x_T) - target|^2
    with a seeded target, FP64 adjoint with checkpoints every 20 steps.
    Directional derivatives along seeded directions agree with central
    differences, and the perturbed rollouts execute the same iterations as
    the base rollout (the adjoint differentiates the executed control flow, as
    the reference's CppAD tape does).

    The difference quotients have a noise floor: each rollout stops its PD
    iterations at the 5e-7 tolerance, so the loss carries a small
    parameter-dependent truncation error (about 1e-10 of the loss here). E
    directions therefore use a relative step of 1e-3 (1e-6 with contact), a 1e-3
    bar and directions without sign cancellation; activation directions agree
    to 1e-4.

    With ground contact the loss is only piecewise smooth: stick, slide and
    take-off switch per vertex and per iteration. A central difference that
    spans a switch measures the jump, so the contact case uses 50 steps
    (36 of them in contact) and small steps. Over 200 steps of sliding,
    differences of any size cross switches; there the GPU gradient is pinned
    to the oracle's instead (contact gradient tests)."""
    d = scene.config2()
    d.contact = contact
    rng = np.random.default_rng(30)
    x0 = d.x_rest.ravel().copy()
    v0 = np.zeros_like(x0)
    act = scene.activations_uniform(31, steps, d.n_groups, 1)
    target = d.x_rest.mean(0) + np.array([0.02, -0.01, 0.01])

    def loss(xT):
        c = xT.reshape(-1, 3).mean(0)
        return float(((c - target) ** 2).sum())

    s = eng.Scene(d, R=1, flags=eng.PDM_HOST_EARLY_EXIT, checkpoint_every=20)
    xT, _ = s.forward_host(x0[None], v0[None], steps, act=act, record=True)
    its0 = s.iterations().copy()
    c = xT[0].reshape(-1, 3).mean(0)
    gx = np.tile(2.0 * (c - target) / d.N, d.N)[None]
    g = s.backward_host(gx, np.zeros_like(gx))
    # activations: two directions, four rollouts in one batch
    da = [rng.uniform(-1, 1, act.shape) for _ in range(2)]
    sa = eng.Scene(d, R=4, flags=eng.PDM_HOST_EARLY_EXIT)
    acts = np.concatenate([act + eps_a * da[0], act - eps_a * da[0], act + eps_a * da[1], act - eps_a * da[1]])
    xa, _ = sa.forward_host(np.tile(x0, (4, 1)), np.tile(v0, (4, 1)), steps, act=acts)
    for r in range(4):
        assert np.array_equal(sa.iterations()[r], its0[0]), r
    for k in range(2):
        fd = (loss(xa[2 * k]) - loss(xa[2 * k + 1])) / (2 * eps_a)
        an = float((g["act"][0] * da[k][0]).sum())
        assert abs(fd - an) <= 1e-4 * abs(an), (k, fd, an)
    # per-tet E: a uniform scaling and (without contact) a random relative
    # direction whose signs follow the gradient's (no cancellation, so the
    # directional derivative stands well above the difference quotient's noise
    # floor). With contact, per-tet perturbations move single vertices across
    # stick/slide boundaries; the per-tet split of the contact E-gradient is
    # pinned on the oracle instead (test_contact_per_tet_E_gradient_vs_fd).
    dirs = [d.E.copy()] if contact else [d.E.copy(), rng.uniform(0.5, 1, d.T) * d.E * np.sign(g["E"])]
    for k, dE in enumerate(dirs):
        vals = []
        for sgn in (1, -1):
            se = eng.Scene(d, R=1, flags=eng.PDM_HOST_EARLY_EXIT)
            se.set_material(d.E + sgn * eps_e * dE)
            xe, _ = se.forward_host(x0[None], v0[None], steps, act=act)
            assert np.array_equal(se.iterations()[0], its0[0]), (k, sgn)
            vals.append(loss(xe[0]))
        fd = (vals[0] - vals[1]) / (2 * eps_e)
        an = float((g["E"] * dE).sum())
        assert abs(fd - an) <= tol_e * abs(an), (k, fd, an)


def _assert_kinks(o, kinked, x0, v0, act, gx, gv, steps):
    """Each (rollout, error, oracle gradients) in kinked: the oracle's own
    FP64 gradient must jump by a comparable amount when the rollout's start
    velocity changes by 1e-4 relative (a contact-branch kink next to the
    trajectory, within FP32 reach)."""
    prng = np.random.default_rng(7)
    for r, err, go in kinked:
        jump = 0.0
        for _ in range(8):
            vp = v0[r] * (1.0 + 1e-4 * prng.standard_normal(v0[r].shape))
            o.rollout(x0[r], vp, steps, act=act[r], record=True)
            gp, _ = o.backward(gx[r], gv[r])
            jump = max(jump, rel(gp["x0"], go["x0"]), rel(gp["v0"], go["v0"]))
        assert jump > 0.3 * err, (r, err, jump)


def _drop_batch_scene():
    x, t = scene.box_grid(4, 2, 2)
    x = x + np.array([0.0, 0.0, 0.002])
    s0, s1, g, ng = scene.straight_fibres(1, 2, x.min(0), x.max(0), nseg=10)
    rng = np.random.default_rng(64)
    return scene.SceneDesc(x_rest=x, tets=t, total_mass=1.0, E=rng.uniform(3e4, 8e4, t.shape[0]), nu_fwd=0.4,
                           nu_ctor=np.full(t.shape[0], 0.4), seg_start=s0, seg_end=s1, seg_group=g, n_groups=ng,
                           k_muscle=2e4, h=0.01, max_iter=10, gravity=(0, 0, -9.81), contact=True, mu=0.5,
                           max_iter1=10, planes=np.array([[0, 0, 0, 0, 0, 1.0]]))


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_adjoint"])
def test_batch64_drop_matches_oracle(eng, oracle_mod, precision):
    """R = 64 (the batched layout of the benchmark: one wave per tet for 64
    listed rollouts, wave-uniform vertex rows) with the benchmark's flags
    (hipGraph replay, device-side early exit) and a whole-rollout tape: a muscle
    box dropped onto a friction floor at 64 different velocities, so that the
    rollouts still iterating drop below 24 (the dense element kernels) as well
    as between 24 and 63. Forward and gradients per rollout against the oracle.

    Mixed precision: the gradient through friction contact is piecewise
    smooth (a contact vertex sticks or slides, the branch frozen at the one
    taken). Where the FP64 trajectory passes within FP32 reach of a branch
    boundary, the FP32 local step can take the other branch and the gradient
    then belongs to the neighbouring piece. Such a rollout may miss the 1e-4
    bar only if the oracle's own FP64 gradient jumps by a comparable amount
    when the start velocity changes by 1e-4 relative (a kink next to the
    trajectory); at most two of the 64 may, each within 1e-2.

    Mixed adjoint (FP64 forward, FP32 local-step VJP): the forward is the FP64
    path's, so positions and iteration counts hold the FP64 bar and every
    rollout's gradient the 1e-4 bar, without exception."""
    d = _drop_batch_scene()
    R, steps = 64, 12
    prec = {"fp64": eng.PDM_FP64, "mixed": eng.PDM_MIXED, "mixed_adjoint": eng.PDM_MIXED_ADJOINT}[precision]
    tol_x, tol_g = {"fp64": (1e-9, 1e-9), "mixed": (1e-5, 1e-4), "mixed_adjoint": (1e-9, 1e-4)}[precision]
    s = eng.Scene(d, R=R, precision=prec, flags=eng.PDM_HOST_EARLY_EXIT | eng.PDM_USE_GRAPH,
                  checkpoint_every=steps)
    o = oracle_mod.OracleScene(d)
    rng = np.random.default_rng(65)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    v0 = np.zeros((R, 3 * d.N))
    for r in range(R):
        u = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(-1.2, 1.0)])
        v0[r] = np.tile(u, d.N) + 0.02 * rng.standard_normal(3 * d.N)
    act = rng.uniform(0, 1, (R, steps, d.n_groups))
    xT, vT = s.forward_host(x0, v0, steps, act=act, record=True)
    its = s.iterations()
    passes, jac = s.contact_stats()
    # rollouts still iterating at PD iteration k of each step
    live = np.array([[(its[:, n] > k).sum() for k in range(d.max_iter)] for n in range(steps)])
    assert ((live > 0) & (live < 24)).any() and ((live >= 24) & (live < 64)).any()
    assert jac.sum() > 0 and (jac.sum(1) == 0).any()
    gx = rng.standard_normal((R, 3 * d.N))
    gv = 0.01 * rng.standard_normal((R, 3 * d.N))
    g = s.backward_host(gx, gv)
    gE_sum = np.zeros(d.T)
    kinked = []
    for r in range(R):
        xo, vo, io = o.rollout(x0[r], v0[r], steps, act=act[r], record=True)
        if precision != "mixed":
            assert np.array_equal(its[r], io[:, 0]) and np.array_equal(jac[r], io[:, 1]), r
        assert rel(xT[r], xo) < tol_x, r
        go, rc = o.backward(gx[r], gv[r])
        assert rc == 0
        err = max(rel(g["x0"][r], go["x0"]), rel(g["v0"][r], go["v0"]), rel(g["act"][r], go["act"]))
        if err >= tol_g:
            assert precision == "mixed" and err < 1e-2, (r, err)
            kinked.append((r, err, go))
        gE_sum += go["E"]
    assert len(kinked) <= 2, [k[:2] for k in kinked]
    _assert_kinks(o, kinked, x0, v0, act, gx, gv, steps)
    # the per-tet E gradient sums over rollouts, kinked ones included
    assert rel(g["E"], gE_sum) < (tol_g if not kinked else 1e-2)


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_adjoint"])
def test_element_variants_agree_bitwise(eng, precision):
    """A rollout's forward does not depend on which element kernels ran for
    it. With 24 or more rollouts iterating, the strip kernel adds the corner
    forces into per-vertex strip sums held in registers; with fewer, the dense
    kernel writes corner slots and k_strip_fold sums them in the same order.
    In the 64-rollout drop each rollout passes through both; 64 copies of one
    rollout keep all 64 iterating together (strip kernel only). The reverse
    sweep's wave-uniform and dense VJP kernels must agree in the same way.
    States, PD and contact iteration counts and the rollout's gradients (x0,
    v0, activations) must be bitwise equal."""
    d = _drop_batch_scene()
    R, steps = 64, 12
    prec = {"fp64": eng.PDM_FP64, "mixed": eng.PDM_MIXED, "mixed_adjoint": eng.PDM_MIXED_ADJOINT}[precision]
    flags = eng.PDM_HOST_EARLY_EXIT | eng.PDM_USE_GRAPH
    rng = np.random.default_rng(65)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    v0 = np.zeros((R, 3 * d.N))
    for r in range(R):
        u = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(-1.2, 1.0)])
        v0[r] = np.tile(u, d.N) + 0.02 * rng.standard_normal(3 * d.N)
    act = rng.uniform(0, 1, (R, steps, d.n_groups))
    gx = rng.standard_normal((R, 3 * d.N))
    gv = 0.01 * rng.standard_normal((R, 3 * d.N))
    sa = eng.Scene(d, R=R, precision=prec, flags=flags, checkpoint_every=steps)
    xa, va = sa.forward_host(x0, v0, steps, act=act, record=True)
    its = sa.iterations()
    jac = sa.contact_stats()[1]
    ga = sa.backward_host(gx, gv)
    sa.close()
    live = np.array([[(its[:, n] > k).sum() for k in range(d.max_iter)] for n in range(steps)])
    dense = np.array([sum(int((live[n, :its[r, n]] < 24).sum()) for n in range(steps)) for r in range(R)])
    rs = int(np.argmax(dense))
    assert dense[rs] > 0
    sb = eng.Scene(d, R=R, precision=prec, flags=flags, checkpoint_every=steps)
    xb, vb = sb.forward_host(np.tile(x0[rs], (R, 1)), np.tile(v0[rs], (R, 1)), steps, act=np.tile(act[rs], (R, 1, 1)),
                             record=True)
    itb = sb.iterations()
    jb = sb.contact_stats()[1]
    gb = sb.backward_host(np.tile(gx[rs], (R, 1)), np.tile(gv[rs], (R, 1)))
    sb.close()
    assert np.array_equal(itb[0], its[rs]) and np.array_equal(jb[0], jac[rs])
    assert np.array_equal(xb[0], xa[rs]) and np.array_equal(vb[0], va[rs])
    assert np.array_equal(xb, np.tile(xb[0], (R, 1)))
    for k in ("x0", "v0", "act"):
        assert np.array_equal(gb[k][0], ga[k][rs]), k


def test_mixed_adjoint_full_horizon_all_rollouts(eng, oracle_mod):
    """60 steps of a muscle box bouncing on a friction floor, 64 rollouts:
    the mixed-adjoint forward equals the FP64 forward bitwise (positions and
    iteration counts), and every rollout's gradient (x0, v0, activations) is
    within 1e-4 of the FP64 gradient; two rollouts also against the oracle."""
    d = _drop_batch_scene()
    R, steps = 64, 60
    rng = np.random.default_rng(66)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    v0 = np.zeros((R, 3 * d.N))
    for r in range(R):
        u = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(-1.2, 1.0)])
        v0[r] = np.tile(u, d.N) + 0.02 * rng.standard_normal(3 * d.N)
    act = rng.uniform(0, 1, (R, steps, d.n_groups))
    gx = rng.standard_normal((R, 3 * d.N))
    gv = 0.01 * rng.standard_normal((R, 3 * d.N))
    res = {}
    for name, prec in (("fp64", eng.PDM_FP64), ("ma", eng.PDM_MIXED_ADJOINT)):
        s = eng.Scene(d, R=R, precision=prec, flags=eng.PDM_HOST_EARLY_EXIT | eng.PDM_USE_GRAPH,
                      checkpoint_every=20)
        xT, vT = s.forward_host(x0, v0, steps, act=act, record=True)
        res[name] = (xT, vT, s.iterations(), s.contact_stats(), s.backward_host(gx, gv))
        s.close()
    (x64, v64, i64, c64, g64), (xma, vma, ima, cma, gma) = res["fp64"], res["ma"]
    assert np.array_equal(x64, xma) and np.array_equal(v64, vma)
    assert np.array_equal(i64, ima) and np.array_equal(c64[1], cma[1])
    assert c64[1].sum() > 0
    for r in range(R):
        err = max(rel(gma[k][r], g64[k][r]) for k in ("x0", "v0", "act"))
        assert err < 1e-4, (r, err)
    for r in (0, R - 1):
        xo, _, _ = oracle_mod.OracleScene(d).rollout(x0[r], v0[r], steps, act=act[r], record=False)
        assert rel(xma[r], xo) < 1e-9
    o = oracle_mod.OracleScene(d)
    o.rollout(x0[0], v0[0], steps, act=act[0], record=True)
    go, rc = o.backward(gx[0], gv[0])
    assert rc == 0 and rel(gma["x0"][0], go["x0"]) < 1e-4 and rel(gma["act"][0], go["act"]) < 1e-4


def test_rollout_backward_refuses_a_replaced_tape(eng):
    """Two rollout nodes on one scene: the later node's backward runs first
    and matches a fresh single-node run; the earlier node's tape is gone, so
    its backward raises instead of returning the later rollout's gradients."""
    import torch
    from diff_fem_amd import api_diff
    d = small_muscle_scene()
    R, steps = 2, 6
    s = eng.Scene(d, R=R)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(5)
    x0 = torch.tensor(np.tile(d.x_rest.ravel(), (R, 1)), device=dev, requires_grad=True)
    v0 = torch.tensor(0.01 * rng.standard_normal((R, 3 * d.N)), device=dev, requires_grad=True)
    E = torch.tensor(d.E, dtype=torch.float64)
    act1 = torch.tensor(rng.uniform(0, 1, (R, steps, d.n_groups)), device=dev)
    act2 = torch.tensor(rng.uniform(0, 1, (R, steps, d.n_groups)), device=dev)
    xa, _ = api_diff.rollout(x0, v0, E, act1, s, steps)
    xb, _ = api_diff.rollout(x0, v0, E, act2, s, steps)
    with pytest.raises(RuntimeError, match="tape is gone"):
        (xa.sum() + xb.sum()).backward()
    # one node per forward works
    x0.grad = None
    xb, _ = api_diff.rollout(x0, v0, E, act2, s, steps)
    xb.sum().backward()
    assert torch.isfinite(x0.grad).all() and x0.grad.abs().max() > 0


def test_split_scene_matches_one_scene(eng):
    """engine.SplitScene (two sub-batches, each on its own stream and host
    thread) against one Scene on the same batch with contact, graphs and a
    whole-rollout tape: per-rollout states, iteration counts and gradients
    bitwise equal; the parameter gradients summed over the parts equal to
    rounding."""
    import torch
    d = _drop_batch_scene()
    R, steps = 8, 8
    flags = eng.PDM_HOST_EARLY_EXIT | eng.PDM_USE_GRAPH
    rng = np.random.default_rng(80)
    dev = torch.device("cuda:0")
    x0 = torch.from_numpy(np.tile(d.x_rest.ravel(), (R, 1))).to(dev)
    v0 = torch.from_numpy(np.tile([0.1, -0.05, -0.8], (R, d.N)) + 0.05 * rng.standard_normal((R, 3 * d.N))).to(dev)
    act = torch.from_numpy(rng.uniform(0, 1, (R, steps, d.n_groups))).to(dev)
    gx = torch.from_numpy(rng.standard_normal((R, 3 * d.N))).to(dev)
    gv = torch.from_numpy(0.01 * rng.standard_normal((R, 3 * d.N))).to(dev)
    out = []
    for sc in (eng.Scene(d, R=R, precision=eng.PDM_FP64, flags=flags, checkpoint_every=steps),
               eng.SplitScene(d, R=R, parts=2, precision=eng.PDM_FP64, flags=flags, checkpoint_every=steps)):
        xT, vT = sc.forward_device(x0, v0, steps, act=act, record=True)
        g = sc.backward_device(gx, gv)
        torch.cuda.synchronize()
        out.append((xT.cpu().numpy(), vT.cpu().numpy(), {k: v.cpu().numpy() for k, v in g.items()}, sc.iterations(),
                    sc.contact_stats()[1]))
        sc.close()
    (x1, v1, g1, i1, j1), (x2, v2, g2, i2, j2) = out
    assert j1.sum() > 0
    assert np.array_equal(i1, i2) and np.array_equal(j1, j2)
    assert np.array_equal(x1, x2) and np.array_equal(v1, v2)
    for k in ("x0", "v0", "act"):
        assert np.array_equal(g1[k], g2[k]), k
    assert rel(g2["E"], g1["E"]) < 1e-12 and abs(g2["nu"][0] - g1["nu"][0]) <= 1e-12 * abs(g1["nu"][0])


@pytest.mark.parametrize("parts", [1, 2])
def test_rollout_autograd_matches_oracle(eng, oracle_mod, parts):
    """api_diff.rollout (the torch.autograd.Function over the device API) on a
    Scene and on a two-part SplitScene: x_T, v_T and the gradients of a scalar
    loss with respect to x0, v0, the per-tet E and the activations, against the
    oracle's forward and reverse sweep per rollout (FP64 bar)."""
    import torch
    from diff_fem_amd import api_diff
    d = _drop_batch_scene()
    R, steps = 4, 8
    dev = torch.device("cuda:0")
    kw = dict(precision=eng.PDM_FP64, flags=eng.PDM_HOST_EARLY_EXIT, checkpoint_every=4)
    sc = eng.Scene(d, R=R, **kw) if parts == 1 else eng.SplitScene(d, R=R, parts=parts, **kw)
    rng = np.random.default_rng(90)
    x0n = np.tile(d.x_rest.ravel(), (R, 1))
    v0n = np.tile([0.05, 0.1, -0.7], (R, d.N)) + 0.05 * rng.standard_normal((R, 3 * d.N))
    actn = rng.uniform(0, 1, (R, steps, d.n_groups))
    wx, wv = rng.standard_normal((R, 3 * d.N)), 0.01 * rng.standard_normal((R, 3 * d.N))
    x0 = torch.tensor(x0n, device=dev, requires_grad=True)
    v0 = torch.tensor(v0n, device=dev, requires_grad=True)
    E = torch.tensor(np.asarray(d.E, np.float64), device=dev, requires_grad=True)
    act = torch.tensor(actn, device=dev, requires_grad=True)
    xT, vT = api_diff.rollout(x0, v0, E, act, sc, steps)
    loss = (xT * torch.tensor(wx, device=dev)).sum() + (vT * torch.tensor(wv, device=dev)).sum()
    loss.backward()
    o = oracle_mod.OracleScene(d)
    gE = np.zeros(d.T)
    for r in range(R):
        xo, vo, _ = o.rollout(x0n[r], v0n[r], steps, act=actn[r], record=True)
        go, rc = o.backward(wx[r], wv[r])
        assert rc == 0
        assert rel(xT[r].detach().cpu().numpy(), xo) < 1e-9 and rel(vT[r].detach().cpu().numpy(), vo) < 1e-9
        assert rel(x0.grad[r].cpu().numpy(), go["x0"]) < 1e-9
        assert rel(v0.grad[r].cpu().numpy(), go["v0"]) < 1e-9
        assert rel(act.grad[r].cpu().numpy(), go["act"]) < 1e-9
        gE += go["E"]
    assert rel(E.grad.cpu().numpy(), gE) < 1e-9
    sc.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config4_full_size_matches_oracle(eng, oracle_mod, precision):
    """Config 4 at full size in the benchmark's layout: the 40x12x12 muscle
    box (34,560 tets, 64 fibres), R = 256 as two concurrent parts of 128
    (wave-uniform kernels), hipGraph replay, device early exit, whole-rollout
    tape, the benchmark's sinusoidal activations, 12 steps dropping onto the
    friction floor. The first rollout of the first part and the last of the
    second against the oracle, forward and gradients; the other rollouts get zero adjoint seeds,
    so the per-tet E gradient is the checked rollouts' alone. Mixed gradients
    follow the kink rule of test_batch64_drop_matches_oracle."""
    import torch
    d = scene.config4(seed=4)
    R, steps = 256, 12
    prec = eng.PDM_FP64 if precision == "fp64" else eng.PDM_MIXED
    tol_x, tol_g = (1e-9, 1e-9) if precision == "fp64" else (1e-5, 1e-4)
    sc = eng.SplitScene(d, R=R, parts=2, precision=prec, flags=eng.PDM_HOST_EARLY_EXIT | eng.PDM_USE_GRAPH,
                        checkpoint_every=steps)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(44)
    x0 = np.tile(d.x_rest.ravel(), (R, 1))
    v0 = np.zeros((R, 3 * d.N))
    for r in range(R):
        v0[r] = np.tile([rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2), rng.uniform(-1.3, -0.8)], d.N)
    act = scene.activations_sinusoid(4, steps, d.n_groups, R, d.h)
    check = [0, 255]
    gx = np.zeros((R, 3 * d.N))
    gv = np.zeros((R, 3 * d.N))
    for r in check:
        gx[r] = rng.standard_normal(3 * d.N)
        gv[r] = 0.01 * rng.standard_normal(3 * d.N)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    xT, vT = sc.forward_device(t(x0), t(v0), steps, act=t(act), record=True)
    g = sc.backward_device(t(gx), t(gv))
    torch.cuda.synchronize()
    xT, g = xT.cpu().numpy(), {k: v.cpu().numpy() for k, v in g.items()}
    its = sc.iterations()
    _, jac = sc.contact_stats()
    sc.close()
    assert jac[check].sum() > 0
    o = oracle_mod.OracleScene(d)
    gE = np.zeros(d.T)
    kinked = []
    for r in check:
        xo, vo, io = o.rollout(x0[r], v0[r], steps, act=act[r], record=True)
        if precision != "mixed":
            assert np.array_equal(its[r], io[:, 0]) and np.array_equal(jac[r], io[:, 1]), r
        assert rel(xT[r], xo) < tol_x, r
        go, rc = o.backward(gx[r], gv[r])
        assert rc == 0
        err = max(rel(g["x0"][r], go["x0"]), rel(g["v0"][r], go["v0"]), rel(g["act"][r], go["act"]))
        if err >= tol_g:
            assert precision == "mixed" and err < 1e-2, (r, err)
            kinked.append((r, err, go))
        gE += go["E"]
    assert len(kinked) <= 1, [k[:2] for k in kinked]
    _assert_kinks(o, kinked, x0, v0, act, gx, gv, steps)
    assert rel(g["E"], gE) < (tol_g if not kinked else 1e-2)
