"""Pins of the CPU FP64 oracle against mathematics, independent library
routines and brute force (the reference ships no golden vectors for this path;
its compiled build needs Eigen/CGAL/Boost/CppAD, absent here)."""
import os

import numpy as np
import pytest

from diff_fem_amd import scene

from contact_scenes import sphere_tumble_scene

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rand_rot(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def polar_numpy(F):
    u, s, vt = np.linalg.svd(F)
    return u @ vt, u, s, vt


# ---------------------------------------------------------------- local step
def test_svd_matches_numpy(oracle_mod):
    rng = np.random.default_rng(0)
    for _ in range(200):
        F = rng.standard_normal((3, 3))
        U, S, V = oracle_mod.svd3(F)
        _, _, s_np, _ = polar_numpy(F)
        assert np.allclose(S, s_np, rtol=1e-13, atol=1e-14)
        assert np.allclose(U @ np.diag(S) @ V.T, F, atol=1e-13)
        assert np.allclose(U.T @ U, np.eye(3), atol=1e-13)
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-13)


def test_projection_closed_forms(oracle_mod):
    rng = np.random.default_rng(1)
    md, mdv = oracle_mod.corotate_project(np.eye(3))
    assert np.allclose(md, np.eye(3), atol=1e-15) and np.allclose(mdv, np.eye(3), atol=1e-15)
    for _ in range(20):
        Q = rand_rot(rng)
        md, mdv = oracle_mod.corotate_project(Q)
        assert np.allclose(md, Q, atol=1e-13) and np.allclose(mdv, Q, atol=1e-13)
        s = rng.uniform(0.6, 1.5, 3)
        F = Q @ np.diag(s) @ rand_rot(rng).T
        md, mdv = oracle_mod.corotate_project(F)
        R, u, sv, vt = polar_numpy(F)
        assert np.allclose(md, R, atol=1e-12)  # polar factor is unique: numpy agrees
        D = oracle_mod.volume_projection(sv)
        assert np.allclose(mdv, u @ np.diag(sv + D) @ vt, atol=1e-12)
        # the 5 linearised steps land on the volume-preserving manifold
        assert abs(np.prod(sv + D) - 1.0) < 1e-6


def test_volume_projection_properties(oracle_mod):
    rng = np.random.default_rng(2)
    for _ in range(50):
        S = rng.uniform(0.7, 1.4, 3)
        D = oracle_mod.volume_projection(S)
        g = np.array([S[1] * S[2], S[0] * S[2], S[0] * S[1]])
        # isotropic stretch: correction is isotropic and exact to 1e-9
        a = rng.uniform(0.8, 1.2)
        Da = oracle_mod.volume_projection(np.array([a, a, a]))
        assert np.allclose(Da, Da[0]) and abs((a + Da[0]) ** 3 - 1) < 1e-9
        # permutation equivariance
        p = rng.permutation(3)
        assert np.allclose(oracle_mod.volume_projection(S[p]), D[p], atol=1e-15)
        assert abs(np.prod(S + D) - 1.0) < 1e-6
        # identity -> no correction
    assert np.allclose(oracle_mod.volume_projection(np.ones(3)), 0)


def test_inverted_tet_flip(oracle_mod):
    rng = np.random.default_rng(3)
    for _ in range(20):
        F = rand_rot(rng) @ np.diag(rng.uniform(0.5, 1.5, 3)) @ rand_rot(rng).T
        F[:, 2] *= -1  # det < 0
        md, _ = oracle_mod.corotate_project(F)
        R, *_ = polar_numpy(F)
        assert np.linalg.det(R) < 0
        assert np.allclose(md, R @ np.diag([1, 1, -1]), atol=1e-12)
        assert np.linalg.det(md) > 0


def test_rotation_equivariance(oracle_mod):
    rng = np.random.default_rng(4)
    for _ in range(20):
        F = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        Q = rand_rot(rng)
        a, b = oracle_mod.corotate_project(F)
        qa, qb = oracle_mod.corotate_project(Q @ F)
        assert np.allclose(qa, Q @ a, atol=1e-12) and np.allclose(qb, Q @ b, atol=1e-12)


def test_vjp_matches_finite_differences(oracle_mod):
    rng = np.random.default_rng(5)
    for trial in range(30):
        F = np.eye(3) + 0.35 * rng.standard_normal((3, 3))
        if trial % 5 == 0:
            F[:, 2] *= -1
        if trial % 7 == 1:  # repeated singular values
            F = rand_rot(rng) @ np.diag([1.1, 1.1, 0.9]) @ rand_rot(rng).T
        wr, wv = 1.3, 0.7
        Z = rng.standard_normal((3, 3))
        dF = rng.standard_normal((3, 3))
        Fb = oracle_mod.corotate_vjp(F, wr, wv, Z)
        h = 1e-6

        def f(M):
            a, b = oracle_mod.corotate_project(M)
            return np.sum(Z * (wr * a + wv * b))

        fd = (f(F + h * dF) - f(F - h * dF)) / (2 * h)
        assert abs(fd - np.sum(Fb * dF)) < 1e-6 * max(1.0, abs(fd)), trial


def _volproj_exact(S):
    """The 5-step volume projection (CorotateFEMConstraint.h:366-374) in exact
    rational arithmetic: phi = S + D."""
    from fractions import Fraction as Fr
    S = [Fr(float(v)) for v in S]
    D = [Fr(0)] * 3
    for _ in range(5):
        a = [S[i] + D[i] for i in range(3)]
        c = a[0] * a[1] * a[2] - 1
        g = [a[1] * a[2], a[0] * a[2], a[0] * a[1]]
        coef = (sum(g[i] * D[i] for i in range(3)) - c) / sum(x * x for x in g)
        D = [coef * x for x in g]
    return [S[i] + D[i] for i in range(3)], S


def test_volume_projection_divided_differences_are_exact(oracle_mod):
    """The VJP's (phi_i - phi_j)/(S_i - S_j) is carried through the projection
    instead of formed by division: it matches the exact rational divided
    difference also for singular values 1e-12 apart, and equals the
    derivative limit J_ii - J_ij where two coincide."""
    rng = np.random.default_rng(12)
    pairs = [(0, 1), (0, 2), (1, 2)]
    for gap in (0.3, 1e-3, 1e-6, 1e-9, 1e-12):
        for _ in range(4):
            S = np.sort(rng.uniform(0.6, 1.4, 3))[::-1].copy()
            S[1] = S[0] - gap
            D, J, r = oracle_mod.volume_projection_jac(S)
            np.testing.assert_allclose(D, oracle_mod.volume_projection(S), rtol=0, atol=1e-15)
            phi, Sx = _volproj_exact(S)
            for k, (i, j) in enumerate(pairs):
                exact = float((phi[i] - phi[j]) / (Sx[i] - Sx[j]))
                assert abs(r[k] - exact) < 1e-13 * max(1.0, abs(exact)), (gap, k, r[k], exact)
    S = np.array([1.2, 1.2, 0.8])
    D, J, r = oracle_mod.volume_projection_jac(S)
    assert abs(r[0] - (J[0, 0] - J[0, 1])) < 1e-13


# ---------------------------------------------------------------- setup + assembly
def small_muscle_scene(contact=False):
    x, t = scene.box_grid(4, 2, 2)
    s0, s1, g, ng = scene.straight_fibres(1, 2, x.min(0), x.max(0), nseg=6)
    rng = np.random.default_rng(7)
    return scene.SceneDesc(x_rest=x, tets=t, total_mass=5.0, E=rng.uniform(5e4, 2e5, t.shape[0]), nu_fwd=0.4,
                           nu_ctor=np.full(t.shape[0], 0.4), attach_idx=np.array([0, 1, 2], np.int32),
                           attach_w=1e7, seg_start=s0, seg_end=s1, seg_group=g, n_groups=ng, k_muscle=5e5,
                           h=0.01, max_iter=10, gravity=(0, 0, -9.81), contact=contact, mu=0.5,
                           planes=np.array([[0, 0, -0.01, 0, 0, 1.0]]))


def test_masses_and_orientation(oracle_mod):
    x, t = scene.box_grid(1, 1, 1)
    d = scene.SceneDesc(x_rest=x, tets=t, total_mass=6.0)
    s = oracle_mod.OracleScene(d)
    m = s.masses()
    assert abs(m.sum() - 6.0) < 1e-14
    v, vol, G = s.tets()
    assert np.all(vol > 0) and abs(vol.sum() - 0.05 ** 3) < 1e-18
    # m_i proportional to incident volume / 4 (Kuhn split: diagonal corners in all 6 tets)
    inc = np.zeros(8)
    for tt in t:
        inc[tt] += 1
    assert np.allclose(m, 6.0 * inc / 24.0)


def test_kronecker_assembly_matches_bruteforce(oracle_mod):
    d = small_muscle_scene()
    s = oracle_mod.OracleScene(d)
    assert s.muscles[0].shape[0] > 0
    K = s.dense_K()
    H = s.dense_H2ML_bruteforce()
    assert np.allclose(np.kron(K, np.eye(3)), H, rtol=1e-13, atol=1e-9)


def test_solve_matches_dense(oracle_mod):
    d = scene.config1()
    s = oracle_mod.OracleScene(d)
    K = s.dense_K()
    rng = np.random.default_rng(8)
    b = rng.standard_normal((s.N, 3))
    x = s.solve3(b)
    xr = np.linalg.solve(K, b)
    assert np.allclose(x, xr, rtol=1e-10, atol=1e-30)
    assert np.linalg.norm(K @ x - b) < 1e-8 * np.linalg.norm(b) * 1e6


# ---------------------------------------------------------------- PD loop invariants
def test_rest_state_is_fixed_point(oracle_mod):
    d = scene.config1()
    d.gravity = (0.0, 0.0, 0.0)
    s = oracle_mod.OracleScene(d)
    x0 = d.x_rest.ravel().copy()
    xT, vT, its = s.rollout(x0, np.zeros_like(x0), 5)
    assert np.abs(xT - x0).max() < 1e-13 and np.abs(vT).max() < 1e-11


def test_free_fall_velocity_recursion(oracle_mod):
    d = scene.config1()
    d.attach_idx = np.zeros(0, np.int32)
    s = oracle_mod.OracleScene(d)
    x0 = d.x_rest.ravel().copy()
    v = np.zeros(3)
    for n in range(1, 6):
        xT, vT, _ = s.rollout(x0, np.zeros_like(x0), n)
    vz = 0.0
    for _ in range(5):
        vz = (vz - 9.81 * 0.01) * 0.9999
    assert np.allclose(vT.reshape(-1, 3)[:, 2], vz, rtol=1e-10)
    assert np.allclose(vT.reshape(-1, 3)[:, :2], 0, atol=1e-12)


def test_early_exit_counts(oracle_mod):
    d = scene.config1()
    s = oracle_mod.OracleScene(d)
    x0 = d.x_rest.ravel().copy()
    _, _, its = s.rollout(x0, np.zeros_like(x0), 20)
    assert np.all(its[:, 0] >= 1) and np.all(its[:, 0] <= 10)


# ---------------------------------------------------------------- gradient vs FD
def _fd_check(s, x0, v0, steps, act, gx, gv, param_eval, analytic, h):
    fd = (param_eval(+h) - param_eval(-h)) / (2 * h)
    assert abs(fd - analytic) <= 1e-5 * max(abs(fd), 1e-12), (fd, analytic)


def test_gradient_vs_central_fd_config1(oracle_mod):
    d = scene.config1()
    s = oracle_mod.OracleScene(d)
    x0 = d.x_rest.ravel().copy()
    v0 = np.zeros_like(x0)
    steps = 15
    rng = np.random.default_rng(9)
    gx = rng.standard_normal(x0.shape)
    gv = 0.01 * rng.standard_normal(x0.shape)
    s.rollout(x0, v0, steps, record=True)
    g, rc = s.backward(gx, gv)
    assert rc == 0

    def loss(E=1e5, nu=0.4, dx=None, dv=None):
        s.set_material(np.full(s.T, E), nu)
        xT, vT, _ = s.rollout(x0 + (0 if dx is None else dx), v0 + (0 if dv is None else dv), steps)
        return gx @ xT + gv @ vT

    hE = 1e5 * 1e-6
    fd = (loss(E=1e5 + hE) - loss(E=1e5 - hE)) / (2 * hE)
    assert abs(fd - g["E"].sum()) < 1e-5 * abs(fd)
    hn = 1e-6
    fd = (loss(nu=0.4 + hn) - loss(nu=0.4 - hn)) / (2 * hn)
    assert abs(fd - g["nu"]) < 1e-5 * abs(fd)
    dirx = rng.standard_normal(x0.shape) * 1e-7
    fd = (loss(dx=dirx) - loss(dx=-dirx)) / 2
    assert abs(fd - g["x0"] @ dirx) < 1e-5 * abs(fd)
    dirv = rng.standard_normal(x0.shape) * 1e-6
    fd = (loss(dv=dirv) - loss(dv=-dirv)) / 2
    assert abs(fd - g["v0"] @ dirv) < 1e-5 * abs(fd)
    s.set_material(np.full(s.T, 1e5), 0.4)


def test_gradient_per_tet_E_and_activations(oracle_mod):
    d = small_muscle_scene()
    s = oracle_mod.OracleScene(d)
    x0 = d.x_rest.ravel().copy()
    v0 = np.zeros_like(x0)
    steps = 8
    rng = np.random.default_rng(10)
    act = rng.uniform(0, 1, (steps, d.n_groups))
    gx = rng.standard_normal(x0.shape)
    gv = 0.01 * rng.standard_normal(x0.shape)
    s.rollout(x0, v0, steps, act=act, record=True)
    g, rc = s.backward(gx, gv)
    assert rc == 0
    E0 = d.E.copy()

    def loss(E=E0, a=act):
        s.set_material(E, d.nu_fwd)
        xT, vT, _ = s.rollout(x0, v0, steps, act=a)
        return gx @ xT + gv @ vT

    for t in [0, 5, 17]:
        e = np.zeros_like(E0)
        e[t] = E0[t] * 1e-5
        fd = (loss(E=E0 + e) - loss(E=E0 - e)) / (2 * e[t])
        assert abs(fd - g["E"][t]) < 1e-4 * abs(fd) + 1e-12, (t, fd, g["E"][t])
    for (n, k) in [(0, 0), (3, 1), (7, 0)]:
        da = np.zeros_like(act)
        da[n, k] = 1e-6
        fd = (loss(a=act + da) - loss(a=act - da)) / 2e-6
        assert abs(fd - g["act"][n, k]) < 1e-4 * abs(fd) + 1e-12, (n, k, fd, g["act"][n, k])
    s.set_material(E0, d.nu_fwd)


# ---------------------------------------------------------------- contact pins
def test_plane_detection_table(oracle_mod):
    p, n = np.zeros(3), np.array([0, 0, 1.0])
    cases = [
        ((0, 0, 0.1), (0, 0, -0.05), True, +1),    # crossing downwards: normal +z
        ((0, 0, 0.1), (0, 0, 0.05), False, 0),     # above, no crossing
        ((0, 0, 0.1), (0, 0, 5e-5), True, +1),     # ends inside thickness
        ((0, 0, -0.1), (0, 0, -0.05), False, 0),   # below, no crossing
        ((0, 0, -0.1), (0, 0, 0.05), True, -1),    # crossing upwards: normal opposes motion
        ((0, 0, -5e-5), (0, 0, 2e-5), True, -1),
    ]
    for x0, x1, hit, sg in cases:
        h, nn = oracle_mod.detect_plane(p, n, np.array(x0, float), np.array(x1, float))
        assert h == hit, (x0, x1)
        if hit:
            assert np.allclose(nn, sg * n)


def test_cone_projection_cases(oracle_mod):
    n = np.array([0, 0, 1.0])
    # take-off: f_n > 0 -> unchanged
    f = np.array([0.3, -0.2, 1.0])
    assert np.allclose(oracle_mod.project_contact(n, 0.5, f), f)
    # stick: |f_t| <= mu |f_n| -> reaction -f
    f = np.array([0.1, 0.05, -1.0])
    assert np.allclose(oracle_mod.project_contact(n, 0.5, f), -f)
    # slide: tangential clamped to the cone
    f = np.array([2.0, 0.0, -1.0])
    r = oracle_mod.project_contact(n, 0.5, f)
    assert np.allclose(r, [-0.5, 0.0, 1.0])
    r = oracle_mod.project_contact(n, 1.0, np.array([0.0, 3.0, -1.0]))
    assert np.allclose(r, [0.0, -1.0, 1.0])


def test_contact_rollout_keeps_body_above_ground(oracle_mod):
    x, t = scene.box_grid(4, 2, 2, origin=(0, 0, 0.02))
    d = scene.SceneDesc(x_rest=x, tets=t, total_mass=2.0, E=np.full(t.shape[0], 1e5), h=0.01, max_iter=10,
                        gravity=(0, 0, -9.81), contact=True, max_iter1=10, mu=0.5,
                        planes=np.array([[0, 0, 0, 0, 0, 1.0]]))
    s = oracle_mod.OracleScene(d)
    x0 = x.ravel().copy()
    xT, vT, its = s.rollout(x0, np.zeros_like(x0), 40)
    assert its[:, 2].max() >= 1  # contact phase ran
    assert xT.reshape(-1, 3)[:, 2].min() > -0.02


def test_bridgesus_fixture_runs(oracle_mod):
    d = scene.load_json_scene("/data/bridgesus/bridgesus.json", os.path.join(GOLD, "scenes"))
    assert d.N == 212 and d.T == 668
    s = oracle_mod.OracleScene(d)
    x0 = d.x_rest.ravel().copy()
    xT, vT, its = s.rollout(x0, np.zeros_like(x0), 3)
    assert np.all(np.isfinite(xT))
    assert np.abs(xT.reshape(-1, 3)[d.attach_idx] - d.x_rest[d.attach_idx]).max() < 1e-12


def test_contact_gradient_vs_central_fd_plane(oracle_mod):
    """Contact-phase reverse sweep (take-off / stick / slide branches frozen):
    exact for static planes (constant contact normal)."""
    x, t = scene.box_grid(5, 2, 2, origin=(0, 0, 0.015))
    for mu in (0.5, 0.2):
        d = scene.SceneDesc(x_rest=x, tets=t, total_mass=2.0, E=np.full(t.shape[0], 8e4), h=0.01, max_iter=10,
                            gravity=(0, 0, -9.81), contact=True, max_iter1=10, mu=mu,
                            planes=np.array([[0, 0, 0, 0, 0, 1.0]]))
        s = oracle_mod.OracleScene(d)
        x0 = x.ravel().copy()
        v0 = np.zeros_like(x0)
        v0[0::3] = 0.3
        steps = 14
        rng = np.random.default_rng(12)
        gx = rng.standard_normal(x0.shape)
        gv = 0.01 * rng.standard_normal(x0.shape)
        _, _, its = s.rollout(x0, v0, steps, record=True)
        assert its[:, 1].sum() > 0
        g, rc = s.backward(gx, gv)

        def loss(E=8e4, dx=0.0, dv=0.0):
            s.set_material(np.full(s.T, E), 0.4)
            xT, vT, _ = s.rollout(x0 + dx, v0 + dv, steps)
            return gx @ xT + gv @ vT

        hE = 8e4 * 1e-6
        fd = (loss(E=8e4 + hE) - loss(E=8e4 - hE)) / (2 * hE)
        assert abs(fd - g["E"].sum()) < 1e-5 * abs(fd)
        dx = 1e-7 * rng.standard_normal(x0.shape)
        fd = (loss(dx=dx) - loss(dx=-dx)) / 2
        assert abs(fd - g["x0"] @ dx) < 1e-5 * abs(fd)
        dv = 1e-6 * rng.standard_normal(x0.shape)
        fd = (loss(dv=dv) - loss(dv=-dv)) / 2
        assert abs(fd - g["v0"] @ dv) < 1e-5 * abs(fd)
        s.set_material(np.full(s.T, 8e4), 0.4)


@pytest.mark.parametrize("mu,sphere", [(0.5, (0.12, 0.05, -0.19, 0.2)), (0.2, (0.12, 0.05, -0.19, 0.2)),
                                       (0.5, (0.02, 0.03, -0.28, 0.3))])
def test_contact_gradient_vs_central_fd_sphere(oracle_mod, mu, sphere):
    """Sphere obstacles: the contact normal moves with the positions it was
    detected on (through the crossing time, or the end-of-step position), and
    the reverse sweep carries that derivative back to the start-of-step
    positions and to the contact iterate the detection read. Both normal
    branches occur in these drops; freezing the normal, dropping either branch,
    or sending the root branch's adjoint to the wrong position misses the
    central differences by 1e-4 to 9e-2."""
    x, t = scene.box_grid(5, 2, 2, origin=(0, 0, 0.015))
    d = scene.SceneDesc(x_rest=x, tets=t, total_mass=2.0, E=np.full(t.shape[0], 8e4), h=0.01, max_iter=10,
                        gravity=(0, 0, -9.81), contact=True, max_iter1=10, mu=mu,
                        planes=np.zeros((0, 6)), spheres=np.array([sphere]))
    s = oracle_mod.OracleScene(d)
    x0 = x.ravel().copy()
    v0 = np.zeros_like(x0)
    v0[0::3] = 0.3
    steps = 14
    rng = np.random.default_rng(12)
    gx = rng.standard_normal(x0.shape)
    gv = 0.01 * rng.standard_normal(x0.shape)
    _, _, its = s.rollout(x0, v0, steps, record=True)
    assert its[:, 1].sum() > 0 and its[:, 2].sum() > 0
    g, _ = s.backward(gx, gv)

    def loss(E=8e4, dx=0.0, dv=0.0):
        s.set_material(np.full(s.T, E), 0.4)
        xT, vT, its1 = s.rollout(x0 + dx, v0 + dv, steps)
        assert np.array_equal(its1, its)
        return gx @ xT + gv @ vT

    hE = 8e4 * 1e-6
    fd = (loss(E=8e4 + hE) - loss(E=8e4 - hE)) / (2 * hE)
    assert abs(fd - g["E"].sum()) < 1e-5 * abs(fd)
    dx = 1e-7 * rng.standard_normal(x0.shape)
    fd = (loss(dx=dx) - loss(dx=-dx)) / 2
    assert abs(fd - g["x0"] @ dx) < 1e-5 * abs(fd)
    dv = 1e-6 * rng.standard_normal(x0.shape)
    fd = (loss(dv=dv) - loss(dv=-dv)) / 2
    assert abs(fd - g["v0"] @ dv) < 1e-5 * abs(fd)


def test_contact_gradient_vs_central_fd_sphere_later_pass(oracle_mod):
    """Sphere normals detected in later contact passes: their adjoint goes to
    the position that detection read (here the input of the pass's last
    iteration). Reading the output instead, or sending every normal's adjoint
    to the PD result, misses the central differences by 4e-4 to 5e-1."""
    d, v0 = sphere_tumble_scene()
    s = oracle_mod.OracleScene(d)
    x0 = d.x_rest.ravel().copy()
    steps = 12
    rng = np.random.default_rng(12)
    gx = rng.standard_normal(x0.shape)
    gv = 0.01 * rng.standard_normal(x0.shape)
    _, _, its = s.rollout(x0, v0, steps, record=True)
    assert its[:, 2].max() >= 2  # a step with more than one detection pass
    g, _ = s.backward(gx, gv)

    def loss(E=8e4, dx=0.0, dv=0.0):
        s.set_material(np.full(s.T, E), 0.4)
        xT, vT, its1 = s.rollout(x0 + dx, v0 + dv, steps)
        assert np.array_equal(its1, its)
        return gx @ xT + gv @ vT

    hE = 8e4 * 1e-6
    fd = (loss(E=8e4 + hE) - loss(E=8e4 - hE)) / (2 * hE)
    assert abs(fd - g["E"].sum()) < 1e-5 * abs(fd)
    for _ in range(2):
        dx = 1e-7 * rng.standard_normal(x0.shape)
        fd = (loss(dx=dx) - loss(dx=-dx)) / 2
        assert abs(fd - g["x0"] @ dx) < 1e-5 * abs(fd)
        dv = 1e-6 * rng.standard_normal(x0.shape)
        fd = (loss(dv=dv) - loss(dv=-dv)) / 2
        assert abs(fd - g["v0"] @ dv) < 1e-5 * abs(fd)


def test_contact_per_tet_E_gradient_vs_fd(oracle_mod):
    """Per-tet split of the E-gradient through the contact phase (the D and C
    weight dependence of the Jacobi update lands on the right tets): single-tet
    and random per-tet directions against central differences, sliding box on
    a plane."""
    x, t = scene.box_grid(5, 2, 2, origin=(0, 0, 0.015))
    d = scene.SceneDesc(x_rest=x, tets=t, total_mass=2.0, E=np.full(t.shape[0], 8e4), h=0.01, max_iter=10,
                        gravity=(0, 0, -9.81), contact=True, max_iter1=10, mu=0.5,
                        planes=np.array([[0, 0, 0, 0, 0, 1.0]]))
    s = oracle_mod.OracleScene(d)
    x0 = x.ravel().copy()
    v0 = np.zeros_like(x0)
    v0[0::3] = 0.3
    steps = 14
    rng = np.random.default_rng(12)
    gx = rng.standard_normal(x0.shape)
    gv = 0.01 * rng.standard_normal(x0.shape)
    _, _, its0 = s.rollout(x0, v0, steps, record=True)
    assert its0[:, 1].sum() > 0
    g, rc = s.backward(gx, gv)
    E0 = d.E.copy()

    def loss(E):
        s.set_material(E, 0.4)
        xT, vT, its = s.rollout(x0, v0, steps)
        assert np.array_equal(its, its0)
        return gx @ xT + gv @ vT

    for dE in (np.eye(s.T)[3] * E0, np.eye(s.T)[40] * E0, rng.uniform(-1, 1, s.T) * E0):
        eps = 1e-5
        fd = (loss(E0 + eps * dE) - loss(E0 - eps * dE)) / (2 * eps)
        assert abs(fd - g["E"] @ dE) < 1e-5 * abs(fd), (fd, g["E"] @ dE)
    s.set_material(E0, 0.4)


def _packed_pair(seed=7, contact=False, max_iter=3):
    ms = scene.config5_meshes(seed=seed, n_meshes=3, tmin=200, tmax=800)
    for m in ms:
        m.contact = contact
        m.max_iter = max_iter
    return ms, scene.pack_scenes(ms)


def test_packed_body_masses_equal_single_body_masses(oracle_mod):
    """Each body of a packed scene keeps the lumped masses it has alone
    (its own "Mass" over its own volume, Dfobj.h:506-525)."""
    ms, p = _packed_pair()
    mp = oracle_mod.OracleScene(p, factor=False).masses()
    off = 0
    for m, mass in zip(ms, p.body_mass):
        mm = oracle_mod.OracleScene(m, factor=False).masses()
        assert np.allclose(mp[off:off + m.N], mm, rtol=1e-14, atol=0)
        assert abs(mm.sum() - mass) < 1e-12 * mass
        off += m.N


def test_packed_rollout_equals_separate_rollouts(oracle_mod):
    """Without contact, and with every body running all iterations, the bodies
    of one world evolve exactly as they do alone (block-diagonal system,
    per-body nu in mMu): translation of the packed layout aside, positions
    agree to rounding."""
    ms, p = _packed_pair()
    rng = np.random.default_rng(0)
    v0 = 0.5 * rng.standard_normal(3 * p.N)
    xT, vT, its = oracle_mod.OracleScene(p).rollout(p.x_rest.ravel().copy(), v0, 5)
    assert np.all(its[:, 0] == 3)
    off = 0
    for m in ms:
        n3 = 3 * m.N
        x, v, i2 = oracle_mod.OracleScene(m).rollout(m.x_rest.ravel().copy(), v0[off:off + n3], 5)
        assert np.all(i2[:, 0] == 3)
        shift = p.x_rest.ravel()[off:off + n3] - m.x_rest.ravel()
        assert np.abs(x - (xT[off:off + n3] - shift)).max() < 1e-12
        assert np.abs(v - vT[off:off + n3]).max() < 1e-10
        off += n3


def test_fwd_adj_batch_equals_single_rollouts(oracle_mod):
    """The OpenMP forward + reverse sweep over rollouts (the multi-core CPU
    baseline) gives each rollout's gradient of the single-rollout sweep, and
    the per-tet E gradient summed over rollouts."""
    x, t = scene.box_grid(4, 2, 2)
    x = x.copy()
    x[:, 2] += 0.01
    s0, s1, g, ng = scene.straight_fibres(2, 1, x.min(0), x.max(0), nseg=8)
    d = scene.SceneDesc(x_rest=x, tets=t, total_mass=2.0, E=np.full(t.shape[0], 1e5), nu_fwd=0.4,
                        nu_ctor=np.full(t.shape[0], 0.4), seg_start=s0, seg_end=s1, seg_group=g, n_groups=ng,
                        h=0.01, max_iter=10, gravity=(0, 0, -9.81), contact=True, max_iter1=30, mu=0.5,
                        planes=np.array([[0, 0, 0, 0, 0, 1.0]]))
    o = oracle_mod.OracleScene(d)
    R, steps = 3, 8
    rng = np.random.default_rng(4)
    x0 = np.tile(x.ravel(), (R, 1))
    v0 = 0.05 * rng.standard_normal((R, 3 * d.N))
    act = rng.uniform(0, 1, (R, steps, ng))
    gx = rng.standard_normal((R, 3 * d.N))
    gv = rng.standard_normal((R, 3 * d.N))
    gx0, gE, its = o.fwd_adj_batch(x0, v0, steps, gx, gv, act=act, nthreads=3)
    gE_sum = np.zeros(d.T)
    for r in range(R):
        _, _, it1 = o.rollout(x0[r], v0[r], steps, act=act[r], record=True)
        g1, rc = o.backward(gx[r], gv[r])
        assert rc == 0
        np.testing.assert_array_equal(its[r], it1[:, 0])
        assert np.abs(gx0[r] - g1["x0"]).max() <= 1e-12 * np.abs(g1["x0"]).max()
        gE_sum += g1["E"]
    assert np.abs(gE - gE_sum).max() <= 1e-12 * np.abs(gE_sum).max()
