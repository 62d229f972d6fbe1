"""Drop-in mirror of the reference's ``pydifem`` module (python/pydifem.cc:329-389)
over libpdm. Same names, argument order, return shapes and error behaviour:

    env = init_env(total_step, config_path)
    q_T, qd_T = forward_all(q, qd, tau, env)
    dq0, dqd0, dtau = backward_all(q, qd, tau, dldq, dldqd, env)
    q, qd, jq, jqd = forward_act(q, qd, jq, jqd, tau, env, is_act)
    dq, dqd, djq, djqd, dtau = backward_act(q, qd, jq, jqd, tau, dldq, dldqd, dldjq, dldjqd, env, is_act)

Differences that are deliberate (each a documented behaviour of the
reference that this engine does not keep):
  * backward calls raise instead of ``exit(0)`` when no forward was recorded
    (the reference exits in its FORWARD_ONLY build, pydifem.cc:105-109, 291-295);
  * OBJ frames are not written per step (pydifem.cc:226-239); ``SaveObj`` writes
    one on request;
  * the config path is resolved against ``PDM_DATA_ROOT`` (or the path as
    given) instead of the compile-time SOFTCON_DIR (Environment.h:90).
Arithmetic runs in FP64 on the GPU; results are Python lists of floats, as the
reference's std::vector<double> returns.
"""
from __future__ import annotations

import numpy as np

from . import engine, scene

__version__ = "dev"


class World:
    """Subset of FEM::World exposed by the reference binding (pydifem.cc:357-365)."""

    def __init__(self, env):
        self._env = env
        self.mX = env._desc.x_rest.ravel().copy()
        self.mV = np.zeros_like(self.mX)
        self.mJointQ = np.zeros(0)
        self.mJointQd = np.zeros(0)
        self.mJointTorque = np.zeros(0)

    def GetTimeStep(self):
        return self._env._desc.h


class Dfobj:
    """Subset of Dfobj exposed by the reference binding (pydifem.cc:367-373)."""

    def __init__(self, env):
        self._env = env
        self.mNumUnknown = 3 * env._desc.N
        self.mActions = np.zeros(env._desc.n_groups)

    def GetActivationLevelsAggregate(self):
        return self._env._act.copy()

    def SetActivationLevelsAggregate(self, a):
        a = np.asarray(a, np.float64).ravel()
        if a.size != self._env._act.size:
            raise ValueError("expected %d activation levels" % self._env._act.size)
        self._env._act[:] = a


class Environment:
    def __init__(self, total_step: int, desc: scene.SceneDesc):
        self.mTotalStep = int(total_step)
        self._desc = desc
        self._scene = engine.Scene(desc, R=1, precision=engine.PDM_FP64, flags=engine.PDM_HOST_EARLY_EXIT,
                                   checkpoint_every=20)
        self._act = np.zeros(desc.n_groups)
        self._E = np.asarray(desc.E, np.float64).copy()
        self._nu = float(desc.nu_fwd)
        self._init_x = desc.x_rest.ravel().copy()
        self.mSoftWorld = World(self)
        self.mDfobjs = [Dfobj(self)]
        self.mPhase = 0
        self._recorded = None

    def SaveObj(self, filename: str):
        """Write the current positions as an OBJ of vertices and tets."""
        x = self.mSoftWorld.mX.reshape(-1, 3)
        with open(filename, "w") as f:
            for p in x:
                f.write("v %f %f %f\n" % tuple(p))
            for t in self._desc.tets:
                f.write("t %d %d %d %d\n" % tuple(t))

    def Reset(self):
        self.mSoftWorld.mX = self._init_x.copy()
        self.mSoftWorld.mV = np.zeros_like(self._init_x)
        self.mPhase = 0
        self._act[:] = 0.0

    def _set_material(self, tau):
        tau = np.asarray(tau, np.float64).ravel()
        if tau.size < 2:
            raise ValueError("tau must hold at least [E, nu]")
        E, nu = float(tau[0]), float(tau[1])
        self._scene.set_material(np.full(self._desc.T, E), nu)


def init_env(total_step: int, config_path: str, data_root: str | None = None) -> Environment:
    """Environment(total_step, config_path) (pydifem.cc:26-36; Environment.h:85-225)."""
    return Environment(total_step, scene.load_json_scene(config_path, data_root))


def init_env_desc(total_step: int, desc: scene.SceneDesc) -> Environment:
    """Environment from a programmatic scene description (synthetic meshes)."""
    return Environment(total_step, desc)


def _as_vec(a, n):
    a = np.asarray(a, np.float64).ravel()
    if a.size != n:
        raise ValueError("expected a vector of %d values, got %d" % (n, a.size))
    return a


def forward_all(q, qd, tau, env: Environment):
    """x=q, v=qd, mMu from tau[0:2], re-factor, mTotalStep steps
    (pydifem.cc:182-259). Returns [q_T, qd_T]."""
    n3 = 3 * env._desc.N
    x0, v0 = _as_vec(q, n3), _as_vec(qd, n3)
    env._set_material(tau)
    steps = env.mTotalStep
    act = np.tile(env._act, (steps, 1)) if env._desc.n_groups else None
    xT, vT = env._scene.forward_host(x0, v0, steps, act=act, record=True)
    env._recorded = ("all", np.asarray(tau, np.float64).ravel().size)
    env.mSoftWorld.mX = xT[0].copy()
    env.mSoftWorld.mV = vT[0].copy()
    return [xT[0].tolist(), vT[0].tolist()]


def backward_all(q, qd, tau, dldq, dldqd, env: Environment):
    """Gradient of <dldq, q_T> + <dldqd, qd_T> over the last forward_all
    (pydifem.cc:261-301): [dq0, dqd0, dtau] with dtau[0] = dE, dtau[1] = dnu,
    then the environment is reset."""
    if env._recorded is None or env._recorded[0] != "all":
        raise RuntimeError("backward_all needs a preceding forward_all on this environment")
    n3 = 3 * env._desc.N
    g = env._scene.backward_host(_as_vec(dldq, n3)[None], _as_vec(dldqd, n3)[None])
    ntau = env._recorded[1]
    dtau = np.zeros(ntau)
    dtau[0] = g["E"].sum()
    dtau[1] = g["nu"]
    env.Reset()
    env._recorded = None
    return [g["x0"][0].tolist(), g["v0"][0].tolist(), dtau.tolist()]


def forward_act(q, qd, jq, jqd, tau, env: Environment, is_act: bool):
    """One Environment::Step from (q, qd) (pydifem.cc:38-86). With is_act the
    activations are set from tau (one level per activation group) and the
    phase advances."""
    n3 = 3 * env._desc.N
    x0, v0 = _as_vec(q, n3), _as_vec(qd, n3)
    if is_act and env._desc.n_groups:
        env._act[:] = np.clip(np.asarray(tau, np.float64).ravel()[: env._desc.n_groups], 0.0, 1.0)
        env.mPhase += 1
    act = env._act[None, :].copy() if env._desc.n_groups else None
    xT, vT = env._scene.forward_host(x0, v0, 1, act=act, record=False)
    env.mSoftWorld.mX = xT[0].copy()
    env.mSoftWorld.mV = vT[0].copy()
    return [xT[0].tolist(), vT[0].tolist(), list(np.asarray(jq, float)), list(np.asarray(jqd, float))]


def backward_act(q, qd, jq, jqd, tau, dldq, dldqd, dldjq, dldjqd, env: Environment, is_act: bool):
    """Re-runs the step from (q, qd) and returns its gradient
    (pydifem.cc:90-180): [dq, dqd, djq, djqd, dtau]; dtau holds the
    activation gradient when is_act."""
    n3 = 3 * env._desc.N
    x0, v0 = _as_vec(q, n3), _as_vec(qd, n3)
    act = env._act[None, :].copy() if env._desc.n_groups else None
    if is_act and env._desc.n_groups:
        act = np.clip(np.asarray(tau, np.float64).ravel()[: env._desc.n_groups], 0.0, 1.0)[None, :]
    env._scene.forward_host(x0, v0, 1, act=act, record=True)
    g = env._scene.backward_host(_as_vec(dldq, n3)[None], _as_vec(dldqd, n3)[None])
    dtau = np.zeros(np.asarray(tau, np.float64).size)
    if is_act and env._desc.n_groups:
        k = min(dtau.size, env._desc.n_groups)
        dtau[:k] = g["act"][0, 0, :k]
    return [g["x0"][0].tolist(), g["v0"][0].tolist(), [0.0] * len(np.asarray(jq).ravel()),
            [0.0] * len(np.asarray(jqd).ravel()), dtau.tolist()]


def vector2double(ev):
    return [float(v) for v in np.asarray(ev, np.float64).ravel()]


def double2vector(v):
    return np.asarray(v, np.float64).ravel().copy()
