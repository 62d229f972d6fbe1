"""Python handle over one libpdm scene (C ABI in include/pdm.h).

``Scene`` owns a ``pdm_scene``: build it from a ``SceneDesc``, choose the
rollout batch and precision, then run forward rollouts and reverse sweeps on
host numpy arrays or on device tensors (torch, FP64, caller AoS layout).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (PDM_FP64, PDM_HOST_EARLY_EXIT, PDM_MIXED, PDM_MIXED_ADJOINT, PDM_NO_AUTOTUNE,  # noqa: F401
                   PDM_TIME_SOLVES, PDM_USE_GRAPH, check)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _vp(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


def make_desc(desc):
    """SceneDesc -> (pdm_scene_desc, arrays kept alive for its pointers)."""
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    x = arr(desc.x_rest.ravel(), np.float64)
    t = arr(desc.tets.ravel(), np.int32)
    E = arr(desc.E, np.float64)
    nuc = arr(desc.nu_ctor, np.float64)
    att = arr(desc.attach_idx, np.int32)
    s0 = arr(desc.seg_start.ravel(), np.float64)
    s1 = arr(desc.seg_end.ravel(), np.float64)
    sg = arr(desc.seg_group, np.int32)
    pl = arr(desc.planes.ravel(), np.float64)
    sp = arr(desc.spheres.ravel(), np.float64)
    d = _lib.SceneDescC()
    if desc.nu_fwd_tet is not None:
        d.nu_fwd_tet = _dptr(arr(desc.nu_fwd_tet, np.float64))
    if desc.body_of_tet is not None:
        d.n_bodies = int(desc.body_mass.size)
        d.body_of_tet = _iptr(arr(desc.body_of_tet, np.int32))
        d.body_mass = _dptr(arr(desc.body_mass, np.float64))
    d.n_vertices = desc.N
    d.x_rest = _dptr(x)
    d.n_tets = desc.T
    d.tets = _iptr(t)
    d.mass_same = int(desc.mass_same)
    d.total_mass = float(desc.total_mass)
    d.E = _dptr(E)
    d.nu_fwd = float(desc.nu_fwd)
    d.nu_ctor = _dptr(nuc)
    d.n_attach = int(att.size)
    d.attach_idx = _iptr(att)
    d.attach_w = float(desc.attach_w)
    d.n_segments = int(sg.size)
    d.seg_start = _dptr(s0)
    d.seg_end = _dptr(s1)
    d.seg_group = _iptr(sg)
    d.n_groups = int(desc.n_groups)
    d.k_muscle = float(desc.k_muscle)
    d.h = float(desc.h)
    d.max_iter = int(desc.max_iter)
    d.damping = float(desc.damping)
    for i in range(3):
        d.gravity[i] = float(desc.gravity[i])
    d.gravity_on = int(desc.gravity_on)
    d.underwater = int(desc.underwater)
    d.self_collision = int(desc.self_collision)
    d.self_mu = float(desc.self_mu)
    d.contact = int(desc.contact)
    d.max_iter1 = int(desc.max_iter1)
    d.alpha = float(desc.alpha)
    d.mu = float(desc.mu)
    d.n_planes = int(desc.planes.shape[0])
    d.planes = _dptr(pl)
    d.n_spheres = int(desc.spheres.shape[0])
    d.spheres = _dptr(sp)
    if desc.tris.shape[0]:
        d.n_tris = int(desc.tris.shape[0])
        d.tris = _dptr(arr(desc.tris.ravel(), np.float64))
    return d, keep


class Scene:
    def __init__(self, desc, R: int = 1, precision: int = PDM_FP64, flags: int = PDM_HOST_EARLY_EXIT,
                 checkpoint_every: int = 20):
        L = _lib.load()
        self.desc = desc
        self.N = int(desc.N)
        self.T = int(desc.T)
        self.n_groups = int(desc.n_groups)
        d, self._keep = make_desc(desc)
        h = C.c_void_p()
        check(L.pdm_scene_create(C.byref(d), C.byref(h)))
        self.h = h
        self._steps = 0
        # Every forward, material update or batch change replaces (or drops)
        # the library's one tape; tape_id names the tape a backward would use.
        self.tape_id = 0
        self.set_batch(R, precision, flags, checkpoint_every)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _lib.load().pdm_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- config
    def set_batch(self, R: int, precision: int = PDM_FP64, flags: int = PDM_HOST_EARLY_EXIT,
                  checkpoint_every: int = 20):
        check(_lib.load().pdm_set_batch(self.h, int(R), int(precision), int(flags), int(checkpoint_every)))
        self.tape_id += 1
        self.R = int(R)
        self.precision = int(precision)
        self.flags = int(flags)

    def set_material(self, E=None, nu_fwd=None):
        L = _lib.load()
        self.tape_id += 1
        Ea = None
        if E is not None:
            Ea = np.ascontiguousarray(np.broadcast_to(np.asarray(E, np.float64), (self.T,)))
        # nu_fwd None keeps the per-tet nu of the scene (NaN in the C call)
        nu = float("nan") if nu_fwd is None else float(nu_fwd)
        check(L.pdm_set_material(self.h, _vp(Ea), nu))

    def refactor(self, on_host: bool = False):
        """Re-factorise for the current material: on the device (what
        set_material does) or, with on_host, in host code (a cross-check)."""
        self.tape_id += 1
        check(_lib.load().pdm_refactor(self.h, 1 if on_host else 0))

    def factor_w(self) -> np.ndarray:
        """W = [inv(L_ss); -L_below inv(L_ss)] of every supernode, column-major
        blocks in factor order (nnz_factor values)."""
        n = self.info()["nnz_factor"]
        W = np.empty(n)
        check(_lib.load().pdm_get_factor_w(self.h, _vp(W), n))
        return W

    def set_stream(self, stream_ptr):
        check(_lib.load().pdm_set_stream(self.h, C.c_void_p(stream_ptr) if stream_ptr else None))

    def info(self) -> dict:
        i = _lib.InfoC()
        check(_lib.load().pdm_scene_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in _lib.InfoC._fields_}

    # ------------------------------------------------------------- host API
    def forward_host(self, x0, v0, steps: int, act=None, record: bool = False):
        R, n3 = self.R, 3 * self.N
        x0 = np.ascontiguousarray(np.broadcast_to(np.asarray(x0, np.float64).reshape(-1, n3), (R, n3)))
        v0 = np.ascontiguousarray(np.broadcast_to(np.asarray(v0, np.float64).reshape(-1, n3), (R, n3)))
        a = None
        if act is not None and self.n_groups > 0:
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(act, np.float64).reshape(-1, steps, self.n_groups),
                                                     (R, steps, self.n_groups)))
        xT = np.zeros((R, n3))
        vT = np.zeros((R, n3))
        self.tape_id += 1
        check(_lib.load().pdm_forward_host(self.h, int(steps), _vp(x0), _vp(v0), _vp(a), _vp(xT), _vp(vT),
                                           int(record)))
        self._steps = steps
        return xT, vT

    def backward_host(self, gx, gv):
        R, n3 = self.R, 3 * self.N
        gx = np.ascontiguousarray(np.asarray(gx, np.float64).reshape(R, n3))
        gv = np.ascontiguousarray(np.asarray(gv, np.float64).reshape(R, n3))
        gx0 = np.zeros((R, n3))
        gv0 = np.zeros((R, n3))
        gE = np.zeros(self.T)
        gnu = np.zeros(1)
        gact = np.zeros((R, max(self._steps, 1), max(self.n_groups, 1)))
        check(_lib.load().pdm_backward_host(self.h, _vp(gx), _vp(gv), _vp(gx0), _vp(gv0), _vp(gE), _vp(gnu),
                                            _vp(gact) if self.n_groups else None))
        return dict(x0=gx0, v0=gv0, E=gE, nu=float(gnu[0]), act=gact[:, :self._steps, :self.n_groups])

    def iterations(self):
        it = np.zeros(self.R * self._steps, np.int32)
        if it.size:
            check(_lib.load().pdm_get_iterations(self.h, _iptr(it)))
        return it.reshape(self.R, self._steps)

    def contact_stats(self):
        """(passes, jacobi_iters), each (R, steps), of the last forward."""
        n = self.R * self._steps
        p = np.zeros(n, np.int32)
        j = np.zeros(n, np.int32)
        if n:
            check(_lib.load().pdm_get_contact_stats(self.h, _iptr(p), _iptr(j)))
        return p.reshape(self.R, self._steps), j.reshape(self.R, self._steps)

    def solve_host(self, b):
        b = np.ascontiguousarray(np.asarray(b, np.float64).reshape(self.R, 3 * self.N))
        x = np.zeros_like(b)
        check(_lib.load().pdm_solve_host(self.h, _vp(b), _vp(x)))
        return x

    def force_host(self, x, act_step=None):
        x = np.ascontiguousarray(np.asarray(x, np.float64).reshape(self.R, 3 * self.N))
        a = None
        if act_step is not None and self.n_groups:
            a = np.ascontiguousarray(np.asarray(act_step, np.float64).reshape(self.R, self.n_groups))
        out = np.zeros_like(x)
        check(_lib.load().pdm_force_host(self.h, _vp(x), _vp(a), _vp(out)))
        return out

    def time_solve(self, reps: int = 20) -> float:
        """Mean ms of one forward+backward triangular solve over all 3R
        right-hand sides (HIP events on the library stream)."""
        ms = C.c_double()
        check(_lib.load().pdm_time_solve(self.h, int(reps), C.byref(ms)))
        return ms.value

    def solve_shapes(self):
        """Per-level launch shapes of the triangular solve: a list of
        ((Ct, threads) forward, (Ct, threads) backward), level 0 = leaves."""
        cap = 64
        a = [(C.c_int * cap)() for _ in range(4)]
        n, chains, cut = C.c_int(), C.c_int(), C.c_int()
        check(_lib.load().pdm_get_solve_shapes(self.h, cap, C.byref(n), *a, C.byref(chains), C.byref(cut)))
        self.solve_chains = chains.value
        self.solve_cut = cut.value
        return [((a[0][l], a[1][l]), (a[2][l], a[3][l])) for l in range(min(n.value, cap))]

    def set_solve_plan(self, subtree_cut: int, n_chains: int):
        """Override the autotuned whole-solve plan (subtree cut level, column
        chains); results do not depend on it."""
        check(_lib.load().pdm_set_solve_plan(self.h, int(subtree_cut), int(n_chains)))

    def set_dense_top(self, max_rows: int):
        """Top-separator blocks solved as one dense product (rows per
        elimination tree at most max_rows; 0: off, level sweeps throughout)."""
        check(_lib.load().pdm_set_dense_top(self.h, int(max_rows)))

    def set_dense_leaves(self, max_rows: int):
        """Leaf subtrees of at most max_rows rows solved as dense products
        (0: off)."""
        check(_lib.load().pdm_set_dense_leaves(self.h, int(max_rows)))

    def solve_timing(self):
        """(number of solves, summed ms) since the last call; needs the
        PDM_TIME_SOLVES flag (HIP events around each solve)."""
        n, ms = C.c_longlong(), C.c_double()
        check(_lib.load().pdm_get_solve_timing(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def graph_launches(self) -> int:
        """Step-block hipGraph replays since the scene was made."""
        n = C.c_longlong()
        check(_lib.load().pdm_get_graph_launches(self.h, C.byref(n)))
        return n.value

    def timing(self):
        f, b, fa = C.c_double(), C.c_double(), C.c_double()
        check(_lib.load().pdm_get_timing(self.h, C.byref(f), C.byref(b), C.byref(fa)))
        return dict(forward_ms=f.value, backward_ms=b.value, factor_ms=fa.value)

    # ------------------------------------------------------------- device API
    def forward_device(self, x0, v0, steps: int, act=None, xT=None, vT=None, record: bool = False):
        """x0, v0: FP64 device tensors (R, 3N); act: (R, steps, G) or None.
        Inputs stay in HBM; outputs are written into xT/vT (allocated if None)."""
        import torch
        R, n3 = self.R, 3 * self.N
        for t in (x0, v0):
            assert t.dtype == torch.float64 and t.is_cuda and t.is_contiguous() and t.numel() == R * n3
        if act is not None:
            assert act.dtype == torch.float64 and act.is_contiguous() and act.numel() == R * steps * self.n_groups
        if xT is None:
            xT = torch.empty_like(x0)
        if vT is None:
            vT = torch.empty_like(v0)
        self.tape_id += 1
        check(_lib.load().pdm_forward(self.h, int(steps), _vp(x0), _vp(v0), _vp(act), _vp(xT), _vp(vT),
                                      int(record)))
        self._steps = steps
        return xT, vT

    def backward_device(self, gx, gv):
        import torch
        R, n3 = self.R, 3 * self.N
        gx0 = torch.empty((R, n3), dtype=torch.float64, device=gx.device)
        gv0 = torch.empty_like(gx0)
        gE = torch.empty(self.T, dtype=torch.float64, device=gx.device)
        gnu = torch.empty(1, dtype=torch.float64, device=gx.device)
        # filled in full by the library when the scene has muscle groups; torch.empty
        # launches nothing on torch's stream (a fill there would wait for other
        # scenes' streams)
        alloc = torch.empty if self.n_groups else torch.zeros
        gact = alloc((R, max(self._steps, 1), max(self.n_groups, 1)), dtype=torch.float64, device=gx.device)
        check(_lib.load().pdm_backward(self.h, _vp(gx.contiguous()), _vp(gv.contiguous()), _vp(gx0), _vp(gv0),
                                       _vp(gE), _vp(gnu), _vp(gact) if self.n_groups else None))
        return dict(x0=gx0, v0=gv0, E=gE, nu=gnu, act=gact[:, :self._steps, :self.n_groups])


class SplitScene:
    """R rollouts run as `parts` independent sub-batches, each a Scene with
    its own HIP stream and its own tape, driven concurrently from one host
    thread per part. Rollouts never interact, so each sub-batch computes
    exactly what the whole batch would; what changes is the overlap: the
    latency-bound upper levels of one part's triangular solve leave compute
    units idle that the other part's element kernels fill. Same device API as
    Scene (forward_device / backward_device on (R, 3N) tensors); the shared
    parameters' gradients (E, nu) are summed over the parts in part order."""

    def __init__(self, desc, R: int, parts: int = 2, precision: int = PDM_FP64, flags: int = PDM_HOST_EARLY_EXIT,
                 checkpoint_every: int = 20):
        if parts < 1 or parts > R:
            raise ValueError("parts must be between 1 and R")
        sizes = [R // parts + (1 if k < R % parts else 0) for k in range(parts)]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
        self.parts = [Scene(desc, R=r, precision=precision, flags=flags, checkpoint_every=checkpoint_every)
                      for r in sizes]
        # Batch buffers and the solve autotuning now, one part after the other:
        # their set-up copies and fills go through the null stream, which must
        # not happen in one part's thread while another part's blocking stream
        # captures its step graph.
        for p in self.parts:
            p.solve_shapes()
        self.R, self.N, self.T, self.n_groups = int(R), int(desc.N), int(desc.T), int(desc.n_groups)
        self._steps = 0

    def _rows(self, k):
        return slice(int(self.offsets[k]), int(self.offsets[k + 1]))

    def _run(self, fn):
        """fn(k) for every part, part 0 on the calling thread; the first
        exception is raised after all parts have returned."""
        import threading
        errs = [None] * len(self.parts)

        def call(k):
            try:
                fn(k)
            except BaseException as e:  # re-raised below
                errs[k] = e
        th = [threading.Thread(target=call, args=(k,)) for k in range(1, len(self.parts))]
        for t in th:
            t.start()
        call(0)
        for t in th:
            t.join()
        for e in errs:
            if e is not None:
                raise e

    def close(self):
        for p in self.parts:
            p.close()

    @property
    def tape_id(self):
        return tuple(p.tape_id for p in self.parts)

    def set_material(self, E=None, nu_fwd=None):
        for p in self.parts:
            p.set_material(E, nu_fwd)
            p.solve_shapes()  # any re-planning here, not in the parts' threads

    def set_dense_top(self, max_rows: int):
        for p in self.parts:
            p.set_dense_top(max_rows)
            p.solve_shapes()

    def set_dense_leaves(self, max_rows: int):
        for p in self.parts:
            p.set_dense_leaves(max_rows)
            p.solve_shapes()

    def forward_device(self, x0, v0, steps: int, act=None, xT=None, vT=None, record: bool = False):
        import torch
        if xT is None:
            xT = torch.empty_like(x0)
        if vT is None:
            vT = torch.empty_like(v0)
        rows = [self._rows(k) for k in range(len(self.parts))]
        self._run(lambda k: self.parts[k].forward_device(
            x0[rows[k]], v0[rows[k]], steps, act=None if act is None else act[rows[k]], xT=xT[rows[k]],
            vT=vT[rows[k]], record=record))
        self._steps = steps
        return xT, vT

    def backward_device(self, gx, gv):
        import torch
        out = [None] * len(self.parts)

        def bwd(k):
            r = self._rows(k)
            out[k] = self.parts[k].backward_device(gx[r].contiguous(), gv[r].contiguous())
        self._run(bwd)
        g = dict(x0=torch.cat([o["x0"] for o in out]), v0=torch.cat([o["v0"] for o in out]),
                 act=torch.cat([o["act"] for o in out]))
        g["E"] = out[0]["E"].clone()
        g["nu"] = out[0]["nu"].clone()
        for o in out[1:]:
            g["E"] += o["E"]
            g["nu"] += o["nu"]
        return g

    def iterations(self):
        return np.concatenate([p.iterations() for p in self.parts])

    def contact_stats(self):
        st = [p.contact_stats() for p in self.parts]
        return np.concatenate([a for a, _ in st]), np.concatenate([b for _, b in st])

    def graph_launches(self) -> int:
        return sum(p.graph_launches() for p in self.parts)

    def timing(self):
        t = [p.timing() for p in self.parts]
        return {k: max(x[k] for x in t) for k in t[0]}

    def solve_timing(self):
        """(number of solves, summed ms) over all parts since the last call;
        each solve covers one part's 3R/parts right-hand sides."""
        st = [p.solve_timing() for p in self.parts]
        return sum(n for n, _ in st), sum(ms for _, ms in st)

    def info(self) -> dict:
        i = dict(self.parts[0].info())
        i["batch"] = self.R
        i["parts"] = len(self.parts)
        return i


def host_analyse(desc) -> dict:
    """Setup + ordering + symbolic + numeric factorisation on the CPU only."""
    d, keep = make_desc(desc)
    i = _lib.InfoC()
    check(_lib.load().pdm_host_analyse(C.byref(d), C.byref(i)))
    return {k: getattr(i, k) for k, _ in _lib.InfoC._fields_}


def host_solve(desc, b):
    """Solve (M/h^2 + L) x = b on the CPU with the engine's supernodal factor."""
    d, keep = make_desc(desc)
    b = np.ascontiguousarray(np.asarray(b, np.float64).reshape(-1, 3 * desc.N))
    x = np.zeros_like(b)
    check(_lib.load().pdm_host_solve(C.byref(d), b.shape[0], _vp(b), _vp(x)))
    return x
