"""torch.autograd wrappers (the reference's python/api_diff.py, FP64).

``sim_all`` / ``sim_act`` keep the reference's call signatures and its
calls into the module (api_diff.py:4-70: SimAct saves the phase and the first
mesh's activation levels and restores both before backward_act) but return
float64 results (the reference casts them to float32, api_diff.py:16-20,
56-57); gradients come back in each input's dtype, so the reference's float32
callers (python/demo_octopus.py) run unchanged.

``rollout`` is the batched device path: R rollouts of one scene as one
autograd node whose backward is the hand-written adjoint (checkpointed
recompute, same Cholesky factor), with gradients for x0, v0, per-tet E and
per-step activations. A scene holds the tape of its latest forward only: the
backward of a node whose scene has run another forward since raises.
"""
from __future__ import annotations

import numpy as np
import torch

from . import pydifem as pd


class SimAll(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, qd, tau, env):
        ctx.env = env
        q_T, qd_T = pd.forward_all(q.detach().cpu().numpy(), qd.detach().cpu().numpy(),
                                   tau.detach().cpu().numpy(), env)
        ctx.save_for_backward(q, qd, tau)
        return torch.tensor(q_T, dtype=torch.float64), torch.tensor(qd_T, dtype=torch.float64)

    @staticmethod
    def backward(ctx, dldq, dldqd):
        q, qd, tau = ctx.saved_tensors
        dq, dqd, dtau = pd.backward_all(q.detach().cpu().numpy(), qd.detach().cpu().numpy(),
                                        tau.detach().cpu().numpy(), dldq.detach().cpu().numpy(),
                                        dldqd.detach().cpu().numpy(), ctx.env)
        # gradients in each input's dtype (the reference's callers pass float32)
        return (torch.tensor(dq, dtype=q.dtype), torch.tensor(dqd, dtype=qd.dtype), torch.tensor(dtau, dtype=tau.dtype),
                None)


sim_all = SimAll.apply


class SimAct(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, qd, jq, jqd, tau, env, is_act):
        ctx.env = env
        ctx.is_act = is_act
        ctx.act_level = pd.vector2double(env.mDfobjs[0].GetActivationLevelsAggregate())
        ctx.phase = env.mPhase
        ctx.save_for_backward(q, qd, jq, jqd, tau)
        out = pd.forward_act(q.detach().cpu().numpy(), qd.detach().cpu().numpy(), jq.detach().cpu().numpy(),
                             jqd.detach().cpu().numpy(), tau.detach().cpu().numpy(), env, is_act)
        return tuple(torch.tensor(o, dtype=torch.float64) for o in out)

    @staticmethod
    def backward(ctx, dldq, dldqd, dldjq, dldjqd):
        q, qd, jq, jqd, tau = ctx.saved_tensors
        env = ctx.env
        env.mPhase = ctx.phase
        env.mDfobjs[0].SetActivationLevelsAggregate(pd.double2vector(ctx.act_level))
        n = lambda t: t.detach().cpu().numpy()  # noqa: E731
        out = pd.backward_act(n(q), n(qd), n(jq), n(jqd), n(tau), n(dldq), n(dldqd), n(dldjq), n(dldjqd), env,
                              ctx.is_act)
        # gradients in each input's dtype (the reference's callers pass float32)
        return tuple(torch.tensor(o, dtype=i.dtype) for o, i in zip(out, (q, qd, jq, jqd, tau))) + (None, None)


sim_act = SimAct.apply


class Rollout(torch.autograd.Function):
    """x_T, v_T = rollout(x0, v0, E, act, scene, steps).

    x0, v0: (R, 3N) float64 CUDA tensors; E: (T,) float64 (shared by all
    rollouts; a change re-factorises); act: (R, steps, G) float64 or an empty
    tensor for scenes without muscles."""

    @staticmethod
    def forward(ctx, x0, v0, E, act, sc, steps):
        E_h = E.detach().cpu().numpy()
        if getattr(sc, "_E_cached", None) is None or not np.array_equal(sc._E_cached, E_h):
            sc.set_material(E_h)
            sc._E_cached = E_h.copy()
        a = act.contiguous() if act.numel() else None
        xT, vT = sc.forward_device(x0.contiguous(), v0.contiguous(), steps, act=a, record=True)
        ctx.sc = sc
        ctx.tape_id = sc.tape_id
        ctx.has_act = act.numel() > 0
        ctx.act_shape = act.shape
        return xT, vT

    @staticmethod
    def backward(ctx, gx, gv):
        # The scene holds one tape, that of its latest forward: a backward of
        # an earlier node would silently differentiate the later rollout.
        if ctx.sc.tape_id != ctx.tape_id:
            raise RuntimeError("rollout backward: the scene ran another forward (or changed its material or "
                               "batch) after this node's forward, so its tape is gone; run one rollout per scene "
                               "between backward passes, or use one scene per rollout node")
        g = ctx.sc.backward_device(gx.contiguous(), gv.contiguous())
        gact = g["act"].contiguous() if ctx.has_act else torch.zeros(ctx.act_shape, dtype=torch.float64,
                                                                     device=gx.device)
        return g["x0"], g["v0"], g["E"], gact, None, None


rollout = Rollout.apply
