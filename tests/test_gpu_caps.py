"""Fixed per-step capacities of the contact phase: at the limit the HIP path
still matches the oracle; past it the call raises instead of dropping an
impact or a group member.

- per (contact vertex, rollout), the distinct obstacles seen in one step:
  at most 8 (a vertex falling through nine stacked planes in one step);
- per (surface face, rollout), the vertices of one self-contact group: at
  most 16 (a fine box landing flat on one large triangle of another body).
"""
import numpy as np
import pytest

from diff_fem_amd import scene
from diff_fem_amd._lib import PdmError

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def stacked_planes_scene(n_planes):
    x, t = scene.box_grid(2, 2, 1, origin=(0.0, 0.0, 0.002))
    planes = np.array([[0.0, 0.0, -0.001 * k, 0.0, 0.0, 1.0] for k in range(n_planes)])
    d = scene.SceneDesc(x_rest=x, tets=t, total_mass=1.0, E=np.full(t.shape[0], 1e5), h=0.01, max_iter=10,
                        gravity=(0.0, 0.0, -9.81), contact=True, max_iter1=10, mu=0.5, planes=planes)
    v0 = np.tile([0.0, 0.0, -6.0], x.shape[0])
    return d, v0


def test_obstacle_cap_at_limit_matches_oracle(oracle_mod):
    from diff_fem_amd import engine as eng
    d, v0 = stacked_planes_scene(8)
    s = eng.Scene(d, R=1, flags=0)
    o = oracle_mod.OracleScene(d)
    x0 = d.x_rest.ravel()
    steps = 3
    xT, vT = s.forward_host(x0, v0, steps)
    xo, vo, io = o.rollout(x0, v0, steps)
    passes, jac = s.contact_stats()
    assert passes.sum() > 0
    assert np.array_equal(s.iterations()[0], io[:, 0])
    assert rel(xT[0], xo) < 1e-9 and rel(vT[0], vo) < 1e-9
    s.close()


def test_obstacle_cap_exceeded_raises():
    from diff_fem_amd import engine as eng
    d, v0 = stacked_planes_scene(9)
    s = eng.Scene(d, R=1, flags=0)
    with pytest.raises(PdmError, match="more than 8 distinct obstacle"):
        s.forward_host(d.x_rest.ravel(), v0, 2)
    s.close()


def test_self_contact_group_cap_exceeded_raises():
    """One large cube (its top face two triangles) and a 6 x 6 x 1 box of
    small cubes 1 mm above it moving down at 0.5 m/s: its 49 bottom vertices
    cross the top face in the first step, about 24 per triangle."""
    from diff_fem_amd import engine as eng
    xa, ta = scene.box_grid(1, 1, 1, edge=0.2)
    xb, tb = scene.box_grid(6, 6, 1, edge=0.02, origin=(0.04, 0.04, 0.201))
    da = scene.SceneDesc(x_rest=xa, tets=ta, total_mass=4.0, E=np.full(ta.shape[0], 1e5),
                         attach_idx=np.arange(xa.shape[0], dtype=np.int32), attach_w=1e8)
    db = scene.SceneDesc(x_rest=xb, tets=tb, total_mass=0.5, E=np.full(tb.shape[0], 1e5))
    d = scene.pack_scenes([da, db], translate=False)
    d.h, d.max_iter, d.gravity, d.contact, d.max_iter1 = 0.01, 10, (0.0, 0.0, 0.0), True, 10
    d.self_collision, d.self_mu = True, 0.5
    na = xa.shape[0]
    s = eng.Scene(d, R=1, flags=0)
    assert s.info()["n_self_faces"] > 0
    v0 = np.zeros(3 * d.N)
    v0[3 * na + 2::3] = -0.5
    with pytest.raises(PdmError, match="self-contact list overflowed"):
        s.forward_host(d.x_rest.ravel(), v0, 2)
    s.close()
