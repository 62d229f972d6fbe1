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
