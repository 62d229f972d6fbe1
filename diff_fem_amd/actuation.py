"""Muscle actuation of the reference's actuator type 2, on the host.

One action vector of 4 values per muscle, normalised to [-5, 5], drives the
50 segment activations of that muscle (Dfobj::SetActions, Dfobj.h:316-331):

  * ``norm_to_real`` maps the normalised actions to the real ranges
    delta in [-0.25, 0.25], alpha in [0, 1], (unused) in [0, 1],
    speed in [0, 10], clamping to the range (Normalizer::NormToReal,
    Normalizer.h:93-114; bounds Muscle.h:63-78; normalised bounds
    Dfobj.h:364-371);
  * ``propagate`` runs Muscle::SetActivationLevels (Muscle.h:224-258): the
    first segment takes ``key + delta``, and each pass moves every other
    segment towards its predecessor by ``alpha``; ``ceil(speed)`` passes
    start from the muscle's current levels, and the result is clamped to
    [0, 1]; with no pass all levels become 0;
  * ``key_table`` is Dfobj::SetKey (Dfobj.h:628-665): a 50-phase key per
    muscle, the running sum of a fixed increment table clamped to [0, 1] at
    every phase. The table is written for 16 muscles (the octopus).

Each forward function has a reverse (``*_vjp``) that the gradient of
``backward_act`` chains through. A clamp passes the gradient where its input
lies inside the closed range, the branch that the reference's taped
``std::min``/``std::max`` take (they return their first argument on ties).
The per-pass recurrence is linear in the levels and bilinear in alpha; the
pass count depends on ``speed`` only through a comparison, so ``speed`` has
no gradient, as in the reference's tape.
"""
from __future__ import annotations

import numpy as np

N_SAMPLING = 50
N_KEY_MUSCLES = 16
REAL_LO = np.array([-0.25, 0.0, 0.0, 0.0])
REAL_HI = np.array([0.25, 1.0, 1.0, 10.0])
NORM_LO = -5.0
NORM_HI = 5.0


def key_table(n_muscles: int) -> np.ndarray:
    """(50, n_muscles) activation key per phase (Dfobj.h:628-665). The
    reference fills the increment table with exactly 16 values per row, so it
    is defined for 16 muscles only; other counts raise."""
    if n_muscles != N_KEY_MUSCLES:
        raise NotImplementedError("the reference's activation key table is defined for %d muscles, not %d"
                                  % (N_KEY_MUSCLES, n_muscles))
    inc = np.zeros((N_SAMPLING, n_muscles))
    even = np.arange(n_muscles) % 2 == 0
    inc[:20] = np.where(even, 0.025, -0.0625)
    inc[20:30] = np.where(even, -0.125, 0.03125)
    inc[30:50] = np.where(even, 0.0, -0.125)
    key = np.zeros_like(inc)
    for j in range(n_muscles):
        acc = 0.0
        for i in range(N_SAMPLING):
            acc = min(max(acc + inc[i, j], 0.0), 1.0)
            key[i, j] = acc
    return key


def norm_to_real(a: np.ndarray):
    """Normalised actions (4 per muscle, flat) -> (real actions, mask of the
    entries whose clamp passed the gradient)."""
    a = np.asarray(a, np.float64).reshape(-1, 4)
    inv = 1.0 / (NORM_HI - NORM_LO)
    v = (a - NORM_LO) * inv
    mask = (v >= 0.0) & (v <= 1.0)
    v = np.minimum(np.maximum(v, 0.0), 1.0)
    real = REAL_LO + (REAL_HI - REAL_LO) * v
    return real, mask


def norm_to_real_vjp(g_real: np.ndarray, mask: np.ndarray) -> np.ndarray:
    inv = 1.0 / (NORM_HI - NORM_LO)
    return (np.asarray(g_real).reshape(-1, 4) * (REAL_HI - REAL_LO) * inv * mask).ravel()


def propagate(levels: np.ndarray, key: float, action: np.ndarray):
    """One muscle: current levels (50,), its key at this phase and its real
    action (4,) -> (new levels (50,), tape for ``propagate_vjp``)."""
    delta, alpha, speed = float(action[0]), float(action[1]), float(action[3])
    beta = 1.0 - alpha
    prev = np.asarray(levels, np.float64).copy()
    new = np.zeros(N_SAMPLING)
    passes = []
    i = 0
    while i < speed:
        new = np.empty(N_SAMPLING)
        new[0] = key + delta
        new[1:] = alpha * prev[:-1] + beta * prev[1:]
        passes.append(prev)
        prev = new
        i += 1
    mask = (new >= 0.0) & (new <= 1.0)
    out = np.minimum(np.maximum(new, 0.0), 1.0)
    return out, (passes, mask, alpha)


def propagate_vjp(g_out: np.ndarray, tape) -> np.ndarray:
    """Gradient of the real action (4,) from that of the new levels."""
    passes, mask, alpha = tape
    ga = np.zeros(4)
    g = np.asarray(g_out, np.float64) * mask
    for prev in reversed(passes):
        ga[0] += g[0]
        ga[1] += float(np.dot(g[1:], prev[:-1] - prev[1:]))
        gp = np.zeros(N_SAMPLING)
        gp[:-1] += alpha * g[1:]
        gp[1:] += (1.0 - alpha) * g[1:]
        g = gp
    return ga


def set_actions(levels: np.ndarray, key: np.ndarray, phase: int, actions: np.ndarray):
    """Dfobj::SetActions for actuator type 2: ``levels`` (n_muscles*50,) are
    the muscles' current levels, ``key`` the (50, n_muscles) table, ``actions``
    the normalised (4*n_muscles,) vector. Returns the new levels and the tape
    of ``set_actions_vjp``."""
    nm = key.shape[1]
    real, nmask = norm_to_real(actions)
    out = np.empty(nm * N_SAMPLING)
    tapes = []
    krow = key[phase % key.shape[0]]
    for m in range(nm):
        o, tp = propagate(levels[m * N_SAMPLING:(m + 1) * N_SAMPLING], krow[m], real[m])
        out[m * N_SAMPLING:(m + 1) * N_SAMPLING] = o
        tapes.append(tp)
    return out, (tapes, nmask)


def set_actions_vjp(g_levels: np.ndarray, tape) -> np.ndarray:
    """Gradient of the normalised actions from that of the new levels."""
    tapes, nmask = tape
    g_real = np.zeros((len(tapes), 4))
    for m, tp in enumerate(tapes):
        g_real[m] = propagate_vjp(g_levels[m * N_SAMPLING:(m + 1) * N_SAMPLING], tp)
    return norm_to_real_vjp(g_real, nmask)
