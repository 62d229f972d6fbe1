"""ctypes front-end of the CPU FP64 oracle (oracle/pd_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py``, as the checker. Nothing in the product
package imports it.

``OracleScene(desc)`` builds the reference-semantics scene from a scene
description (any object with the ``SceneDesc`` fields) and exposes rollout,
gradient and the unit-level pieces used by the parity pins.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB names another build of pd_oracle.cpp (a copy with a deliberate
# mistake planted, to check on the CPU that the pin tests catch it)
_LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_scene_new.restype = C.c_void_p
        L.orc_scene_new.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_int, C.c_double]
        L.orc_scene_free.argtypes = [C.c_void_p]
        L.orc_num_tets.argtypes = [C.c_void_p]
        L.orc_get_masses.argtypes = [C.c_void_p, _dp]
        L.orc_get_tets.argtypes = [C.c_void_p, _ip, _dp, _dp]
        L.orc_num_contact_vertices.argtypes = [C.c_void_p]
        L.orc_get_contact_vertices.argtypes = [C.c_void_p, _ip]
        L.orc_set_material.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.orc_set_attachments.argtypes = [C.c_void_p, C.c_int, _ip, C.c_double]
        L.orc_set_bodies.argtypes = [C.c_void_p, C.c_int, _ip, _dp, C.c_int]
        L.orc_set_nu_fwd_tet.argtypes = [C.c_void_p, _dp]
        L.orc_set_muscles.argtypes = [C.c_void_p, C.c_int, _ip, _dp, _dp, C.c_double, _ip, C.c_int]
        L.orc_muscle_segment_select.argtypes = [C.c_void_p, _ip, C.c_int, _dp, _dp, C.c_int, _ip, _dp]
        L.orc_set_sim.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, _dp, C.c_int, C.c_int, C.c_int,
                                  C.c_double, C.c_double]
        L.orc_set_obstacles.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _dp]
        L.orc_factor.argtypes = [C.c_void_p]
        L.orc_reg_used.argtypes = [C.c_void_p]
        L.orc_reg_used.restype = C.c_double
        L.orc_dense_K.argtypes = [C.c_void_p, _dp]
        L.orc_dense_H2ML_bruteforce.argtypes = [C.c_void_p, _dp]
        L.orc_solve3.argtypes = [C.c_void_p, _dp]
        L.orc_force_terms.argtypes = [C.c_void_p, _dp, C.c_void_p, _dp, _dp]
        L.orc_rollout.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_void_p, _dp, _dp, C.c_void_p, C.c_int]
        L.orc_rollout_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, C.c_void_p, _dp, _dp, C.c_int]
        L.orc_backward.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, C.c_void_p]
        L.orc_fwd_adj_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, C.c_void_p, _dp, _dp, _dp, _dp,
                                        C.c_void_p, C.c_int]
        L.orc_corotate_project.argtypes = [_dp, _dp, _dp]
        L.orc_corotate_vjp.argtypes = [_dp, C.c_double, C.c_double, _dp, _dp]
        L.orc_svd3.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_volume_projection.argtypes = [_dp, _dp]
        L.orc_volume_projection_jac.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_set_underwater.argtypes = [C.c_void_p, C.c_int]
        L.orc_self_normals.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_tris.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_set_self_collision.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_cubic_roots01.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_vf_moving.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_self_group.argtypes = [_dp, _dp, _dp, C.c_int, _dp, _dp, C.c_double, _dp, _dp]
        L.orc_detect_tri.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_num_drag_faces.argtypes = [C.c_void_p]
        L.orc_get_drag_faces.argtypes = [C.c_void_p, _ip]
        L.orc_drag_face.argtypes = [_dp, _dp, _dp, C.c_void_p]
        L.orc_drag_forces.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_detect_plane.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.orc_project_contact.argtypes = [_dp, C.c_double, _dp, _dp]
        L.orc_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def muscle_constraints(handle, desc):
    """Per-segment tet selection (Muscle.h:109-215) over every segment; returns
    (idx4 (M,4), fibre dirs (M,3), weights (M,), groups (M,))."""
    L = lib()
    tets = np.ascontiguousarray(desc.tets, dtype=np.int32)
    T = tets.shape[0]
    idx, fib, wts, grp = [], [], [], []
    cap = 4096
    out_t = np.zeros(cap, dtype=np.int32)
    out_w = np.zeros(cap)
    for s in range(desc.seg_start.shape[0]):
        st = np.ascontiguousarray(desc.seg_start[s], dtype=np.float64)
        en = np.ascontiguousarray(desc.seg_end[s], dtype=np.float64)
        n = L.orc_muscle_segment_select(handle, tets, T, st, en, cap, out_t, out_w)
        if n > cap:
            raise RuntimeError("segment selects more than %d constraints" % cap)
        d = (en - st) / np.linalg.norm(en - st)
        for k in range(n):
            idx.append(tets[out_t[k]])
            fib.append(d)
            wts.append(out_w[k])
            grp.append(desc.seg_group[s])
    if not idx:
        return (np.zeros((0, 4), np.int32), np.zeros((0, 3)), np.zeros(0), np.zeros(0, np.int32))
    return (np.ascontiguousarray(idx, dtype=np.int32), np.ascontiguousarray(fib, dtype=np.float64),
            np.ascontiguousarray(wts, dtype=np.float64), np.ascontiguousarray(grp, dtype=np.int32))


class OracleScene:
    def __init__(self, desc, factor=True):
        self.desc = desc
        L = lib()
        x = np.ascontiguousarray(desc.x_rest, dtype=np.float64).ravel()
        t = np.ascontiguousarray(desc.tets, dtype=np.int32).ravel()
        self.N = desc.x_rest.shape[0]
        self.T = desc.tets.shape[0]
        self.h = L.orc_scene_new(self.N, x, self.T, t, int(desc.mass_same), float(desc.total_mass))
        L.orc_set_material(self.h, np.ascontiguousarray(desc.E, dtype=np.float64), float(desc.nu_fwd),
                           np.ascontiguousarray(desc.nu_ctor, dtype=np.float64))
        if getattr(desc, "nu_fwd_tet", None) is not None:
            L.orc_set_nu_fwd_tet(self.h, np.ascontiguousarray(desc.nu_fwd_tet, dtype=np.float64))
        if getattr(desc, "body_of_tet", None) is not None:
            rc = L.orc_set_bodies(self.h, int(desc.body_mass.size), np.ascontiguousarray(desc.body_of_tet, np.int32),
                                  np.ascontiguousarray(desc.body_mass, np.float64), int(desc.mass_same))
            if rc:
                raise RuntimeError(L.orc_last_error().decode())
        if len(desc.attach_idx):
            L.orc_set_attachments(self.h, len(desc.attach_idx), np.ascontiguousarray(desc.attach_idx, np.int32),
                                  float(desc.attach_w))
        self.n_groups = int(desc.n_groups)
        self.muscles = muscle_constraints(self.h, desc)
        if self.muscles[0].shape[0]:
            idx, fib, wts, grp = self.muscles
            L.orc_set_muscles(self.h, idx.shape[0], idx.ravel().copy(), fib.ravel().copy(), wts, float(desc.k_muscle),
                              grp, self.n_groups)
        g = np.ascontiguousarray(desc.gravity, dtype=np.float64)
        L.orc_set_sim(self.h, float(desc.h), int(desc.max_iter), float(desc.damping), g, int(desc.gravity_on),
                      int(desc.contact), int(desc.max_iter1), float(desc.alpha), float(desc.mu))
        L.orc_set_obstacles(self.h, desc.planes.shape[0], np.ascontiguousarray(desc.planes).ravel().copy(),
                            desc.spheres.shape[0], np.ascontiguousarray(desc.spheres).ravel().copy())
        if getattr(desc, "underwater", False):
            L.orc_set_underwater(self.h, 1)
        if getattr(desc, "self_collision", False):
            L.orc_set_self_collision(self.h, 1, float(desc.self_mu))
        tris = getattr(desc, "tris", None)
        if tris is not None and len(tris):
            L.orc_set_tris(self.h, len(tris), np.ascontiguousarray(tris, np.float64).ravel().copy())
        if factor:
            self.factor()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_scene_free(self.h)
            self.h = None

    def factor(self):
        rc = lib().orc_factor(self.h)
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())

    def set_material(self, E, nu_fwd, nu_ctor=None):
        E = np.ascontiguousarray(np.broadcast_to(np.asarray(E, np.float64), (self.T,)))
        nc = self.desc.nu_ctor if nu_ctor is None else nu_ctor
        lib().orc_set_material(self.h, E, float(nu_fwd),
                               np.ascontiguousarray(np.broadcast_to(np.asarray(nc, np.float64), (self.T,))))
        self.factor()

    def drag_faces(self):
        out = np.zeros((lib().orc_num_drag_faces(self.h), 3), np.int32)
        lib().orc_get_drag_faces(self.h, out)
        return out

    def drag_forces(self, x, v):
        f = np.zeros(3 * self.N)
        lib().orc_drag_forces(self.h, np.ascontiguousarray(x, np.float64).ravel(),
                              np.ascontiguousarray(v, np.float64).ravel(), f)
        return f

    def masses(self):
        m = np.zeros(self.N)
        lib().orc_get_masses(self.h, m)
        return m

    def tets(self):
        v = np.zeros((self.T, 4), np.int32)
        vol = np.zeros(self.T)
        G = np.zeros((self.T, 4, 3))
        lib().orc_get_tets(self.h, v, vol, G)
        return v, vol, G

    def contact_vertices(self):
        n = lib().orc_num_contact_vertices(self.h)
        out = np.zeros(n, np.int32)
        lib().orc_get_contact_vertices(self.h, out)
        return out

    def dense_K(self):
        out = np.zeros((self.N, self.N))
        lib().orc_dense_K(self.h, out)
        return out

    def dense_H2ML_bruteforce(self):
        out = np.zeros((3 * self.N, 3 * self.N))
        lib().orc_dense_H2ML_bruteforce(self.h, out)
        return out

    def solve3(self, b):
        x = np.ascontiguousarray(b, dtype=np.float64).reshape(-1).copy()
        lib().orc_solve3(self.h, x)
        return x.reshape(np.shape(b))

    def force_terms(self, x, act=None):
        Jd = np.zeros(3 * self.N)
        Lx = np.zeros(3 * self.N)
        a = None if act is None else np.ascontiguousarray(act, dtype=np.float64)
        lib().orc_force_terms(self.h, np.ascontiguousarray(x, np.float64).ravel(), _ptr(a), Jd, Lx)
        return Jd, Lx

    def self_normals(self, mode):
        """Self-contact impact normals: 0 computed, 1 recorded by the next
        rollout, 2 replayed from that recording (central differences of the
        derivative with the normals held, as the reference's tape holds them)."""
        if lib().orc_self_normals(self.h, int(mode)):
            raise RuntimeError(lib().orc_last_error().decode())

    def rollout(self, x0, v0, steps, act=None, record=False):
        """x0, v0: (N,3) or (3N,). act: (steps, n_groups) or None.
        Returns xT, vT (3N,) and iteration counts (steps, 3)."""
        x0 = np.ascontiguousarray(x0, np.float64).ravel()
        v0 = np.ascontiguousarray(v0, np.float64).ravel()
        xT = np.zeros_like(x0)
        vT = np.zeros_like(v0)
        its = np.zeros((steps, 3), np.int32)
        a = None
        if act is not None:
            a = np.ascontiguousarray(act, np.float64).reshape(steps, self.n_groups)
        rc = lib().orc_rollout(self.h, steps, x0, v0, _ptr(a), xT, vT, _ptr(its), int(record))
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        self._steps = steps
        return xT, vT, its

    def rollout_batch(self, x0, v0, steps, act=None, nthreads=1):
        """x0, v0: (R, 3N). act: (R, steps, n_groups) or None."""
        x0 = np.ascontiguousarray(x0, np.float64)
        v0 = np.ascontiguousarray(v0, np.float64)
        R = x0.shape[0]
        xT = np.zeros_like(x0)
        vT = np.zeros_like(v0)
        a = None if act is None else np.ascontiguousarray(act, np.float64)
        rc = lib().orc_rollout_batch(self.h, R, steps, x0, v0, _ptr(a), xT, vT, int(nthreads))
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        return xT, vT

    def fwd_adj_batch(self, x0, v0, steps, gx, gv, act=None, nthreads=1):
        """R rollouts forward + reverse sweep, OpenMP over rollouts.
        Returns (gx0 (R, 3N), gE (T,) summed over rollouts, iters (R, steps))."""
        x0 = np.ascontiguousarray(x0, np.float64)
        R = x0.shape[0]
        v0 = np.ascontiguousarray(v0, np.float64)
        gx = np.ascontiguousarray(gx, np.float64)
        gv = np.ascontiguousarray(gv, np.float64)
        a = None if act is None else np.ascontiguousarray(act, np.float64).reshape(R, steps, self.n_groups)
        gx0 = np.zeros_like(x0)
        gE = np.zeros(self.T)
        its = np.zeros((R, steps), np.int32)
        rc = lib().orc_fwd_adj_batch(self.h, R, steps, x0, v0, _ptr(a), gx, gv, gx0, gE, _ptr(its), int(nthreads))
        if rc:
            raise RuntimeError(lib().orc_last_error().decode())
        return gx0, gE, its

    def backward(self, gx, gv):
        """Gradient of <gx,x_T>+<gv,v_T> over the last recorded rollout.
        Returns dict(x0, v0, E (per tet), nu, act (steps, n_groups)), flag."""
        gx = np.ascontiguousarray(gx, np.float64).ravel()
        gv = np.ascontiguousarray(gv, np.float64).ravel()
        gx0 = np.zeros_like(gx)
        gv0 = np.zeros_like(gv)
        gE = np.zeros(self.T)
        gnu = np.zeros(1)
        gact = np.zeros((self._steps, max(self.n_groups, 1)))
        rc = lib().orc_backward(self.h, gx, gv, gx0, gv0, gE, gnu, _ptr(gact))
        return dict(x0=gx0, v0=gv0, E=gE, nu=float(gnu[0]), act=gact[:, : self.n_groups]), rc


def corotate_project(F):
    F = np.ascontiguousarray(F, np.float64).reshape(3, 3)
    md = np.zeros((3, 3))
    mdv = np.zeros((3, 3))
    lib().orc_corotate_project(F, md, mdv)
    return md, mdv


def corotate_vjp(F, wr, wv, Z):
    Fb = np.zeros((3, 3))
    lib().orc_corotate_vjp(np.ascontiguousarray(F, np.float64).reshape(3, 3), float(wr), float(wv),
                           np.ascontiguousarray(Z, np.float64).reshape(3, 3), Fb)
    return Fb


def svd3(F):
    U = np.zeros((3, 3))
    S = np.zeros(3)
    V = np.zeros((3, 3))
    lib().orc_svd3(np.ascontiguousarray(F, np.float64).reshape(3, 3), U, S, V)
    return U, S, V


def volume_projection(S):
    D = np.zeros(3)
    lib().orc_volume_projection(np.ascontiguousarray(S, np.float64), D)
    return D


def drag_face(n, w, jac=False):
    """Drag + lift per vertex of one face (force, and its 3x6 Jacobian in
    (n, w) by the complex step)."""
    F = np.zeros(3)
    J = np.zeros((3, 6)) if jac else None
    lib().orc_drag_face(np.ascontiguousarray(n, np.float64), np.ascontiguousarray(w, np.float64), F, _ptr(J))
    return (F, J) if jac else F


def volume_projection_jac(S):
    """(D, J = d(S + D)/dS, rdd = divided differences of S + D for the
    pairs (0,1), (0,2), (1,2))."""
    D, J, r = np.zeros(3), np.zeros((3, 3)), np.zeros(3)
    lib().orc_volume_projection_jac(np.ascontiguousarray(S, np.float64), D, J, r)
    return D, J, r


def cubic_roots01(a, b, c, d):
    """Real roots in [0, 1] of a t^3 + b t^2 + c t + d, ascending (to 1e-12)."""
    x = np.zeros(3)
    n = lib().orc_cubic_roots01(float(a), float(b), float(c), float(d), x)
    return x[:n]


def vf_moving(p0, p1, y0, y1):
    """(hit, t, d, n) of a vertex p0 -> p1 against a face y0 -> y1 (3x3 each)."""
    t, d, n = np.zeros(1), np.zeros(1), np.zeros(3)
    f = lambda a: np.ascontiguousarray(a, np.float64).ravel()  # noqa: E731
    hit = lib().orc_vf_moving(f(p0), f(p1), f(y0), f(y1), t, d, n)
    return bool(hit), float(t[0]), float(d[0]), n


def self_group(n, f3, m3, fv, mv, mu):
    """Self-contact forces of one group: (face corners 3x3, members kx3)."""
    f = lambda a: np.ascontiguousarray(a, np.float64).ravel()  # noqa: E731
    fv = np.asarray(fv, np.float64).reshape(-1, 3)
    o3, ov = np.zeros(9), np.zeros(fv.size)
    lib().orc_self_group(f(n), f(f3), f(m3), fv.shape[0], f(fv), f(mv), float(mu), o3, ov)
    return o3.reshape(3, 3), ov.reshape(-1, 3)


def detect_tri(tri, x0, x1):
    """(hit, normal) of a vertex moving x0 -> x1 against a static triangle
    (9 values)."""
    out = np.zeros(3)
    hit = lib().orc_detect_tri(np.ascontiguousarray(tri, np.float64).ravel(), np.ascontiguousarray(x0, np.float64),
                               np.ascontiguousarray(x1, np.float64), out)
    return bool(hit), out


def detect_plane(p, n, x0, x1):
    out = np.zeros(3)
    hit = lib().orc_detect_plane(*(np.ascontiguousarray(a, np.float64) for a in (p, n, x0, x1)), out)
    return bool(hit), out


def project_contact(n, mu, f):
    out = np.zeros(3)
    lib().orc_project_contact(np.ascontiguousarray(n, np.float64), float(mu), np.ascontiguousarray(f, np.float64), out)
    return out
