"""Static triangle-mesh obstacles in the oracle (ArcSim's vertex-face test,
collision.cpp:127-278, with the face at rest), on the CPU.

Pins: closed-form cases of the test (interior crossing, a crossing outside
the face, a vertex coming to rest within the thickness, the t = 1 edge-line
test, the single-precision box filter, a degenerate face, normal orientation
and its unnormalised length); a floor of two triangles reproduces the
analytic plane floor step for step; the OBJ obstacle reader (fan
triangulation, 1-based and relative indices).
"""
import numpy as np
import pytest

from diff_fem_amd import scene

TRI = np.array([0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0])      # right angle at y0, normal +z


def test_crossing_the_face_hits_with_the_normal_against_the_motion(oracle_mod):
    hit, n = oracle_mod.detect_tri(TRI, [0.2, 0.3, 0.01], [0.2, 0.3, -0.01])
    assert hit and np.allclose(n, [0, 0, 1], atol=0, rtol=0)
    hit, n = oracle_mod.detect_tri(TRI, [0.2, 0.3, -0.01], [0.2, 0.3, 0.01])
    assert hit and np.array_equal(n, [0, 0, -1])
    # a face with a 60-degree corner at y0: the normal keeps |n| = sin 60
    t60 = np.array([0, 0, 0, 1, 0, 0, np.cos(np.pi / 3), np.sin(np.pi / 3), 0.0])
    hit, n = oracle_mod.detect_tri(t60, [0.4, 0.2, 0.01], [0.4, 0.2, -0.01])
    assert hit and abs(n[2] - np.sin(np.pi / 3)) < 1e-15 and n[0] == n[1] == 0


def test_crossing_outside_the_face_misses(oracle_mod):
    assert not oracle_mod.detect_tri(TRI, [0.7, 0.7, 0.01], [0.7, 0.7, -0.01])[0]
    assert not oracle_mod.detect_tri(TRI, [-0.2, 0.3, 0.01], [-0.2, 0.3, -0.01])[0]


def test_resting_within_the_thickness(oracle_mod):
    assert oracle_mod.detect_tri(TRI, [0.2, 0.3, 0.01], [0.2, 0.3, 5e-5])[0]
    assert not oracle_mod.detect_tri(TRI, [0.2, 0.3, 0.01], [0.2, 0.3, 2e-4])[0]


def test_edge_line_and_box_filter_at_rest(oracle_mod):
    """At t = 1 a vertex within the thickness of the plane and of an edge's
    line counts as inside, but the kDOP box of the (dilated) face must still
    reach it."""
    assert oracle_mod.detect_tri(TRI, [0.5, -5e-5, 0.01], [0.5, -5e-5, 5e-5])[0]     # just outside edge y0-y1
    assert not oracle_mod.detect_tri(TRI, [0.5, -3e-4, 0.01], [0.5, -3e-4, 5e-5])[0]  # beyond the thickness
    # on the line through y0-y1 but far beyond the face: the box test rejects it
    assert not oracle_mod.detect_tri(TRI, [3.0, 0.0, 0.01], [3.0, 0.0, 5e-5])[0]


def test_degenerate_face_never_hits(oracle_mod):
    flat = np.array([0, 0, 0, 1, 0, 0, 2, 0, 0.0])
    assert not oracle_mod.detect_tri(flat, [0.5, 0.0, 0.01], [0.5, 0.0, -0.01])[0]


def _drop(floor, oracle_mod, steps=40):
    x, t = scene.box_grid(4, 2, 2)
    x = x + np.array([0.3, 0.35, 0.03])
    kw = dict(planes=np.array([[0, 0, 0, 0, 0, 1.0]])) if floor == "plane" else dict(
        tris=np.array([[-1, -1, 0, 3, -1, 0, 3, 3, 0], [-1, -1, 0, 3, 3, 0, -1, 3, 0.0]]))
    d = scene.SceneDesc(x_rest=x, tets=t, total_mass=1.0, E=np.full(t.shape[0], 5e4), h=0.01, max_iter=10,
                        gravity=(0, 0, -9.81), contact=True, mu=0.5, **kw)
    o = oracle_mod.OracleScene(d)
    v0 = np.tile([0.3, 0.1, -0.5], d.N)
    return o.rollout(d.x_rest, v0, steps)


def test_two_triangle_floor_equals_the_plane_floor(oracle_mod):
    xp, vp, ip = _drop("plane", oracle_mod)
    xt, vt, it = _drop("tris", oracle_mod)
    assert np.array_equal(ip, it)
    assert ip[:, 1].sum() > 0                     # contact iterations happened
    np.testing.assert_allclose(xt, xp, rtol=0, atol=1e-13)
    np.testing.assert_allclose(vt, vp, rtol=0, atol=1e-11)


def test_obj_obstacle_reader(tmp_path):
    p = tmp_path / "o.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nf 1/1/1 2/2/1 3/3/1 4/4/1\nf -4 -2 -1\n")
    tr = scene.load_obj_triangles(str(p))
    np.testing.assert_array_equal(tr, [[0, 0, 0, 1, 0, 0, 1, 1, 0], [0, 0, 0, 1, 1, 0, 0, 1, 0],
                                       [0, 0, 0, 1, 1, 0, 0, 1, 0]])
    p.write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    with pytest.raises(ValueError):
        scene.load_obj_triangles(str(p))
