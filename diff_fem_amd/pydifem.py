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
    them on request, as the reference's Environment::SaveObj does;
  * the config path is resolved against ``PDM_DATA_ROOT`` (or the path as
    given) instead of the compile-time SOFTCON_DIR (Environment.h:90).
Arithmetic runs in FP64 on the GPU; results are Python lists of floats, as the
reference's std::vector<double> returns.
"""
from __future__ import annotations

import numpy as np

from . import actuation, engine, scene

__version__ = "dev"


class EigenVector(np.ndarray):
    """Eigen::Matrix<Scalar, Dynamic, 1> as the binding exposes it
    (pydifem.cc:354-356): ``rows()``, ``cols()``; FP64 numpy underneath."""

    def __new__(cls, a=()):
        return np.array(a, dtype=np.float64).ravel().view(cls)

    def rows(self):
        return int(self.size)

    def cols(self):
        return 1


class EigenMatrix(np.ndarray):
    """Eigen::Matrix<Scalar, Dynamic, Dynamic> (pydifem.cc:351-353)."""

    def __new__(cls, a=((),)):
        return np.array(a, dtype=np.float64, ndmin=2).view(cls)

    def rows(self):
        return int(self.shape[0])

    def cols(self):
        return int(self.shape[1])


class Scalar(float):
    """The scalar type of the build (double in the FORWARD_ONLY build,
    pydifem.cc:14)."""


class Utils:
    """DoubleUtils as the binding exposes it (pydifem.cc:381-385;
    sim/diff/tiny_double_utils.h)."""

    @staticmethod
    def getDouble(v):
        return float(v)

    @staticmethod
    def scalar_from_double(v):
        return float(v)

    @staticmethod
    def zero():
        return 0.0

    @staticmethod
    def fraction(num, denom):
        return float(int(num)) / float(int(denom))


def _ev_prop(name):
    def get(self):
        return getattr(self, "_" + name)

    def put(self, v):
        setattr(self, "_" + name, EigenVector(v))
    return property(get, put)


class World:
    """Subset of FEM::World exposed by the reference binding (pydifem.cc:357-365);
    its state vectors are EigenVectors (assignments are converted)."""
    mX = _ev_prop("mX")
    mV = _ev_prop("mV")
    mJointQ = _ev_prop("mJointQ")
    mJointQd = _ev_prop("mJointQd")
    mJointTorque = _ev_prop("mJointTorque")

    def __init__(self, env):
        self._env = env
        self.mX = env._desc.x_rest.ravel()
        self.mV = np.zeros(self.mX.size)
        self.mJointQ = ()
        self.mJointQd = ()
        self.mJointTorque = ()

    def GetTimeStep(self):
        return self._env._desc.h


class Dfobj:
    """One mesh of the scene: the subset of Dfobj exposed by the reference
    binding (pydifem.cc:367-373), plus its phase (Dfobj::mPhase) and actuator.
    ``mActions`` has 4 entries per muscle for actuator type 2
    (Dfobj.h:335-383), none otherwise."""

    mActions = _ev_prop("mActions")

    def __init__(self, env, info: scene.DfobjInfo):
        self._env = env
        self._info = info
        self.mNumUnknown = 3 * info.n_vertices
        self.mPhase = 0
        n_act = 4 * info.n_muscles if info.actuator_type == 2 else 0
        self.mActions = np.zeros(n_act)
        self._key = actuation.key_table(info.n_muscles) if n_act else None
        g0 = info.group_offset
        self._groups = slice(g0, g0 + info.n_sampling * info.n_muscles)

    def GetActivationLevelsAggregate(self):
        """The muscles' levels, 50 per muscle (Dfobj.h:689-707)."""
        return EigenVector(self._env._act_state[self._groups])

    def SetActivationLevelsAggregate(self, a):
        """Sets the muscles' levels (Dfobj.h:709-720); the levels the
        constraints use change at the next SetActions or backward_act."""
        a = np.asarray(a, np.float64).ravel()
        n = self._groups.stop - self._groups.start
        if a.size != n:
            raise ValueError("expected %d activation levels" % n)
        self._env._act_state[self._groups] = a

    def _contours(self):
        """This mesh's contour faces in its own vertex numbering (cached)."""
        if getattr(self, "_cont", None) is None:
            i, d = self._info, self._env._desc
            tets = d.tets[i.tet_offset:i.tet_offset + i.n_tets] - i.vertex_offset
            self._cont = scene.contour_faces(d.x_rest[i.vertex_offset:i.vertex_offset + i.n_vertices], tets)
        return self._cont

    def _set_actions(self, actions):
        """Dfobj::SetActions for a muscle actuator (Dfobj.h:316-331): new
        levels from the normalised actions at this Dfobj's phase; they become
        both the muscles' levels and the constraints' activations. Returns the
        tape of the gradient."""
        env = self._env
        lv, tape = actuation.set_actions(env._act_state[self._groups], self._key, self.mPhase, actions)
        env._act_state[self._groups] = lv
        env._act_seg[self._groups] = lv
        return tape


class Environment:
    """Environment (Environment.h:85-225): one World, one Dfobj per mesh, the
    phase counter ``mPhase`` (readable and writable, as the binding exposes
    it) and the activation state of every muscle segment. Two activation
    vectors are kept, as the reference keeps two: the muscles' levels
    (Muscle::mActivationLevels, read and written by the *ActivationLevels*
    calls) and the levels the muscle constraints use (set by SetActions,
    Reset and the non-actuated branch of backward_act)."""

    def __init__(self, total_step: int, desc: scene.SceneDesc):
        self.mTotalStep = int(total_step)
        self._desc = desc
        self._scene = engine.Scene(desc, R=1, precision=engine.PDM_FP64, flags=engine.PDM_HOST_EARLY_EXIT,
                                   checkpoint_every=20)
        self._act_state = np.zeros(desc.n_groups)
        self._act_seg = np.zeros(desc.n_groups)
        self._E = np.asarray(desc.E, np.float64).copy()
        self._nu = float(desc.nu_fwd)
        self._init_x = desc.x_rest.ravel().copy()
        self.mSoftWorld = World(self)
        infos = desc.dfobjs
        if not infos:
            infos = [scene.DfobjInfo(vertex_offset=0, n_vertices=desc.N, tet_offset=0, n_tets=desc.T)]
        self.mDfobjs = [Dfobj(self, info) for info in infos]
        self.mPhase = 0
        self._recorded = None

    def SaveObj(self, filename: str, x):
        """Environment::SaveObj (Environment.h:296-326): for every Dfobj i,
        `<filename>_DfObj<i>.obj` with its vertices taken from consecutive
        segments of x (mNumUnknown values each) as `v` lines and its contour
        faces (1-based, outward) as `f` lines."""
        x = np.asarray(x, np.float64).ravel()
        need = sum(d.mNumUnknown for d in self.mDfobjs)
        if x.size < need:
            raise ValueError("SaveObj needs %d values of x, got %d" % (need, x.size))
        off = 0
        for i, d in enumerate(self.mDfobjs):
            seg = x[off:off + d.mNumUnknown].reshape(-1, 3)
            with open("%s_DfObj%d.obj" % (filename, i), "w") as f:
                f.writelines("v %06f %06f %06f\n" % (p[0], p[1], p[2]) for p in seg)
                f.writelines("f %d %d %d\n" % (t[0] + 1, t[1] + 1, t[2] + 1) for t in d._contours())
            off += d.mNumUnknown

    def Reset(self):
        """Environment::Reset (Environment.h:388-399): muscle levels and the
        constraints' activations to 0, every Dfobj's phase to 0, the world to
        its rest state. ``mPhase`` itself is left as it is."""
        self.mSoftWorld.mX = self._init_x.copy()
        self.mSoftWorld.mV = np.zeros_like(self._init_x)
        for d in self.mDfobjs:
            d.mPhase = 0
        self._act_state[:] = 0.0
        self._act_seg[:] = 0.0

    def IncPhase(self):
        """Environment::IncPhase (Environment.h:372-378)."""
        self.mPhase += 1
        for d in self.mDfobjs:
            d.mPhase = self.mPhase

    def SetPhase(self, phase: int):
        """Environment::SetPhase (Environment.h:380-386): sets the Dfobjs'
        phase only."""
        for d in self.mDfobjs:
            d.mPhase = int(phase)

    def GetPhase(self):
        return self.mPhase

    def SetActions(self, actions):
        """Environment::SetActions (Environment.h:476-486): consecutive
        slices of ``actions`` go to the Dfobjs, each as long as its
        ``mActions``. Returns [(dfobj, slice, tape)] for the gradient."""
        actions = np.asarray(actions, np.float64).ravel()
        need = sum(d.mActions.size for d in self.mDfobjs)
        if actions.size < need:
            raise ValueError("expected at least %d actions, got %d" % (need, actions.size))
        out, k = [], 0
        for d in self.mDfobjs:
            n = d.mActions.size
            if n:
                out.append((slice(k, k + n), d._set_actions(actions[k:k + n])))
            k += n
        return out

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


def _seg_act(env, steps=1):
    return np.tile(env._act_seg, (steps, 1)) if env._desc.n_groups else None


def forward_all(q, qd, tau, env: Environment):
    """x=q, v=qd, mMu = tau[0]/(1+tau[1]) on every corotated constraint,
    re-factor, mTotalStep steps at the constraints' current activations
    (pydifem.cc:182-259). Returns [q_T, qd_T]."""
    n3 = 3 * env._desc.N
    x0, v0 = _as_vec(q, n3), _as_vec(qd, n3)
    env._set_material(tau)
    steps = env.mTotalStep
    xT, vT = env._scene.forward_host(x0, v0, steps, act=_seg_act(env, steps), record=True)
    env._recorded = ("all", np.asarray(tau, np.float64).ravel().size)
    env.mSoftWorld.mX = xT[0].copy()
    env.mSoftWorld.mV = vT[0].copy()
    return [xT[0].tolist(), vT[0].tolist()]


def backward_all(q, qd, tau, dldq, dldqd, env: Environment):
    """Gradient of <dldq, q_T> + <dldqd, qd_T> over the last forward_all
    (pydifem.cc:261-301): [dq0, dqd0, dtau] with dtau[0] = dE, dtau[1] = dnu,
    then Environment::Reset."""
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
    """One Environment::Step from (q, qd) (pydifem.cc:38-86). With is_act,
    SetActions(tau) at the Dfobjs' phase, then IncPhase."""
    n3 = 3 * env._desc.N
    x0, v0 = _as_vec(q, n3), _as_vec(qd, n3)
    if is_act:
        env.SetActions(tau)
        env.IncPhase()
    xT, vT = env._scene.forward_host(x0, v0, 1, act=_seg_act(env), record=False)
    env.mSoftWorld.mX = xT[0].copy()
    env.mSoftWorld.mV = vT[0].copy()
    return [xT[0].tolist(), vT[0].tolist(), list(np.asarray(jq, float)), list(np.asarray(jqd, float))]


def backward_act(q, qd, jq, jqd, tau, dldq, dldqd, dldjq, dldjqd, env: Environment, is_act: bool):
    """Gradient of one step from (q, qd) (pydifem.cc:90-180):
    [dq, dqd, djq, djqd, dtau]. With is_act, SetActions(tau) at the Dfobjs'
    phase, which then becomes mPhase + 1 (mPhase itself is not advanced);
    without, the constraints take the muscles' levels. dtau is the gradient
    through the muscle actuation (zero without is_act)."""
    n3 = 3 * env._desc.N
    x0, v0 = _as_vec(q, n3), _as_vec(qd, n3)
    tau = np.asarray(tau, np.float64).ravel()
    tapes = []
    if is_act:
        cur = env.mPhase
        tapes = env.SetActions(tau)
        env.SetPhase(cur + 1)
    else:
        env._act_seg[:] = env._act_state
    env._scene.forward_host(x0, v0, 1, act=_seg_act(env), record=True)
    g = env._scene.backward_host(_as_vec(dldq, n3)[None], _as_vec(dldqd, n3)[None])
    dtau = np.zeros(tau.size)
    if tapes:
        gseg = g["act"][0, 0]
        for (sl, tape), d in zip(tapes, [d for d in env.mDfobjs if d.mActions.size]):
            dtau[sl] = actuation.set_actions_vjp(gseg[d._groups], tape)
    return [g["x0"][0].tolist(), g["v0"][0].tolist(), [0.0] * len(np.asarray(jq).ravel()),
            [0.0] * len(np.asarray(jqd).ravel()), dtau.tolist()]


def vector2double(ev):
    return [float(v) for v in np.asarray(ev, np.float64).ravel()]


def double2vector(v):
    return EigenVector(v)
