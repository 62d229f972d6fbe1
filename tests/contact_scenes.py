"""Contact scenes shared by the oracle tests and the GPU parity tests."""
import numpy as np

from diff_fem_amd import scene


def sphere_tumble_scene(mu=0.5, max_iter1=10):
    """A tumbling 6x2x2 bar dropped onto a small sphere (no ground): some
    contact passes run out of Jacobi iterations, so the next detection reads
    the last iteration's input, and sphere normals detected in later passes
    act. Returns the scene and the initial velocity (3N)."""
    x, t = scene.box_grid(6, 2, 2, origin=(0, 0, 0.03))
    d = scene.SceneDesc(x_rest=x, tets=t, total_mass=2.0, E=np.full(t.shape[0], 8e4), h=0.01, max_iter=10,
                        gravity=(0, 0, -9.81), contact=True, max_iter1=max_iter1, mu=mu,
                        planes=np.zeros((0, 6)), spheres=np.array([[0.214, 0.069, -0.013, 0.031]]))
    X = x - x.mean(0)
    v0 = np.zeros(x.size)
    v0[0::3] = -0.056 - 19.72 * X[:, 2]
    v0[2::3] = -1.452 + 19.72 * X[:, 0]
    return d, v0


def side_push_scene(mu=0.5, self_mu=0.5, gap=0.004):
    """Two 2x2x1 boxes on a friction floor, the second sliding into the first's
    side: the face it hits has corners on the floor, so obstacle reactions and
    self-contact groups act on the same vertices. Returns the scene and the
    vertex count of the first box."""
    xa, ta = scene.box_grid(2, 2, 1, origin=(0, 0, 0.0015))
    xb, tb = scene.box_grid(2, 2, 1, origin=(0.1 + gap, 0.02, 0.0015))
    da = scene.SceneDesc(x_rest=xa, tets=ta, total_mass=1.0, E=np.full(ta.shape[0], 1e5))
    db = scene.SceneDesc(x_rest=xb, tets=tb, total_mass=0.5, E=np.full(tb.shape[0], 1e5))
    d = scene.pack_scenes([da, db], translate=False)
    d.h, d.max_iter, d.gravity, d.contact, d.max_iter1, d.mu = 0.01, 10, (0.0, 0.0, -9.81), True, 10, mu
    d.planes = np.array([[0, 0, 0, 0, 0, 1.0]])
    d.self_collision, d.self_mu = True, self_mu
    return d, xa.shape[0]
