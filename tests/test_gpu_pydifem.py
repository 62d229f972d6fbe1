"""The pydifem surface on the GPU against the FP64 oracle, on the reference's
own scene files (tests/golden/scenes): forward_all / backward_all on
bridgesus (attachments, gravity -100 y, collision handling on without
obstacles) and forward_act / backward_act on the octopus (16 muscles,
actuator type 2; drag forces switched off).

The oracle gets the same scene description and the segment activations that
the actuation produced; the actuation itself (host code) is pinned on the CPU
(test_scene_actuation.py). Bar: 1e-9 relative max-norm (FP64 path).
"""
import os

import numpy as np
import pytest

from diff_fem_amd import actuation, scene

pytestmark = pytest.mark.gpu
ROOT = os.path.join(os.path.dirname(__file__), "golden", "scenes")


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_bridgesus_forward_backward_all(oracle_mod):
    from diff_fem_amd import pydifem as pd
    steps = 12
    env = pd.init_env(steps, "/data/bridgesus/bridgesus.json", ROOT)
    d = env._desc
    q0 = pd.vector2double(env.mSoftWorld.mX)
    qd0 = [0.0] * len(q0)
    tau = [1.5e5, 0.35]
    qT, qdT = pd.forward_all(q0, qd0, tau, env)
    o = oracle_mod.OracleScene(d)
    o.set_material(np.full(d.T, tau[0]), tau[1])
    xo, vo, _ = o.rollout(np.asarray(q0), np.asarray(qd0), steps, record=True)
    assert np.abs(np.asarray(qT) - np.asarray(q0)).max() > 1e-3      # it moved
    assert rel(qT, xo) < 1e-9 and rel(qdT, vo) < 1e-9
    rng = np.random.default_rng(4)
    gx, gv = rng.standard_normal(len(q0)), rng.standard_normal(len(q0))
    dq, dqd, dtau = pd.backward_all(q0, qd0, tau, gx, gv, env)
    go, _ = o.backward(gx, gv)
    assert rel(dq, go["x0"]) < 1e-9 and rel(dqd, go["v0"]) < 1e-9
    assert abs(dtau[0] - go["E"].sum()) < 1e-9 * abs(go["E"].sum())
    assert abs(dtau[1] - go["nu"]) < 1e-9 * abs(go["nu"])
    # backward_all resets the environment
    np.testing.assert_array_equal(env.mSoftWorld.mX, np.asarray(q0))
    with pytest.raises(RuntimeError):
        pd.backward_all(q0, qd0, tau, gx, gv, env)


def test_octopus_act_steps_and_gradient(oracle_mod):
    from diff_fem_amd import pydifem as pd
    env = pd.init_env(1, "/data/octopus/octopus_dry.json", ROOT)
    d = env._desc
    o = oracle_mod.OracleScene(d)
    key = actuation.key_table(16)
    rng = np.random.default_rng(8)
    q = np.asarray(pd.vector2double(env.mSoftWorld.mX))
    qd = np.zeros_like(q)
    z = np.zeros(0)
    levels = np.zeros(800)
    phase = 0
    for k, is_act in enumerate([True, False, True, True]):
        tau = rng.uniform(-4.5, 4.5, 64)
        if is_act:
            levels = actuation.set_actions(levels, key, phase, tau)[0]
            phase += 1
        q1, qd1, _, _ = pd.forward_act(q, qd, z, z, tau, env, is_act)
        xo, vo, _ = o.rollout(q, qd, 1, act=levels[None])
        assert np.abs(np.asarray(q1) - q).max() > 1e-6 if k else True
        assert rel(q1, xo) < 1e-9 and rel(qd1, vo) < 1e-9, k
        q, qd = np.asarray(q1), np.asarray(qd1)
    assert env.mPhase == 3
    # gradient of the next actuated step, the way SimAct calls it
    tau = rng.uniform(-4.5, 4.5, 64)
    saved = env.mDfobjs[0].GetActivationLevelsAggregate()
    env.mPhase = 3
    env.mDfobjs[0].SetActivationLevelsAggregate(saved)
    gx, gv = rng.standard_normal(q.size), rng.standard_normal(q.size)
    dq, dqd, _, _, dtau = pd.backward_act(q, qd, z, z, tau, gx, gv, z, z, env, True)
    lv, tape = actuation.set_actions(saved, key, 3, tau)
    o.rollout(q, qd, 1, act=lv[None], record=True)
    go, _ = o.backward(gx, gv)
    assert rel(dq, go["x0"]) < 1e-9 and rel(dqd, go["v0"]) < 1e-9
    ref_tau = actuation.set_actions_vjp(go["act"][0], tape)
    assert np.abs(ref_tau).max() > 0
    assert rel(dtau, ref_tau) < 1e-9
