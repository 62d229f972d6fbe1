"""Scene descriptions for the projective-dynamics engine.

A ``SceneDesc`` holds everything ``Dfobj::SetDfobj`` + ``Environment``'s JSON
parameters carry for the hot path (sim/Environment.h:109-210,
sim/Dfobj.h:142-314): rest vertices, tets (unoriented; orientation is fixed at
setup as Dfobj.h:474-487 does), total mass, material, attachments, muscle
segments, contact obstacles and the stepping parameters.

Two sources build one:
  * ``box_grid`` / ``make_config`` — seeded Kuhn-split box grids (edge 0.05 m)
    for the five benchmark workloads;
  * ``load_json_scene`` — the reference's JSON scene + ``v``/``t`` OBJ tet mesh
    format (sim/fem/Mesh/OBJMesh.h:25-97), missing keys read as the reference's
    defaults.
"""
from __future__ import annotations

import dataclasses
import json
import math
import os
from typing import List, Optional

import numpy as np

EDGE = 0.05
MUSCLE_SEGMENTS = 50   # Dfobj::MakeMuscles' num_sampling (Dfobj.h:767)


@dataclasses.dataclass
class SceneDesc:
    x_rest: np.ndarray                 # (N,3) float64
    tets: np.ndarray                   # (T,4) int32, as read (unoriented)
    total_mass: float = 10.0           # Dfobj.h:114 default
    mass_same: bool = True             # "Is Mass Same"
    E: np.ndarray = None               # (T,) Young's modulus per tet
    nu_fwd: float = 0.4                # value entering mMu = E/(1+nu)
    nu_ctor: np.ndarray = None         # (T,) nu of the volume-term weight
    attach_idx: np.ndarray = None      # (A,) int32
    attach_w: float = 1e20
    # muscles: per segment start/end points and its activation group
    seg_start: np.ndarray = None       # (S,3)
    seg_end: np.ndarray = None         # (S,3)
    seg_group: np.ndarray = None       # (S,) int32
    n_groups: int = 0
    k_muscle: float = 5e5              # Environment.h:185 Dfobj(5E5, ...)
    # stepping
    h: float = 1.0 / 200
    max_iter: int = 30
    damping: float = 0.9999
    gravity: tuple = (0.0, 0.0, 0.0)
    gravity_on: bool = True
    underwater: bool = False           # "Is underwater": surface drag, no gravity
    self_collision: bool = False       # "Self-collision": surface vertices against surface faces
    self_mu: float = 1.0               # "World Friction": friction of self contact
    contact: bool = False
    max_iter1: int = 30
    alpha: float = 1.0
    mu: float = 1.0
    planes: np.ndarray = None          # (P,6): point, normal
    spheres: np.ndarray = None         # (Q,4): centre, radius
    tris: np.ndarray = None            # (F,9): static obstacle triangles, three corners each
    # several bodies in one world (one Dfobj per mesh): per-tet body id, each
    # body's "Mass", and the nu entering mMu per tet (bodies built with their
    # own Poisson ratio); None = one body / nu_fwd everywhere
    body_of_tet: np.ndarray = None     # (T,) int32
    body_mass: np.ndarray = None       # (B,)
    nu_fwd_tet: np.ndarray = None      # (T,)
    # one entry per reference Dfobj (mesh) of a loaded scene, in load order:
    # DfobjInfo records; None = one body without actuator
    dfobjs: list = None
    name: str = "scene"

    def __post_init__(self):
        self.x_rest = np.ascontiguousarray(self.x_rest, dtype=np.float64).reshape(-1, 3)
        self.tets = np.ascontiguousarray(self.tets, dtype=np.int32).reshape(-1, 4)
        T = self.tets.shape[0]
        if self.E is None:
            self.E = np.full(T, 2e5)
        self.E = np.ascontiguousarray(np.broadcast_to(np.asarray(self.E, dtype=np.float64), (T,)))
        if self.nu_ctor is None:
            self.nu_ctor = np.full(T, 0.4)
        self.nu_ctor = np.ascontiguousarray(np.broadcast_to(np.asarray(self.nu_ctor, dtype=np.float64), (T,)))
        if self.body_of_tet is not None:
            self.body_of_tet = np.ascontiguousarray(self.body_of_tet, dtype=np.int32).reshape(T)
            self.body_mass = np.ascontiguousarray(self.body_mass, dtype=np.float64).ravel()
        if self.nu_fwd_tet is not None:
            self.nu_fwd_tet = np.ascontiguousarray(np.broadcast_to(np.asarray(self.nu_fwd_tet, np.float64), (T,)))
        if self.attach_idx is None:
            self.attach_idx = np.zeros(0, dtype=np.int32)
        self.attach_idx = np.ascontiguousarray(self.attach_idx, dtype=np.int32)
        if self.seg_start is None:
            self.seg_start = np.zeros((0, 3))
            self.seg_end = np.zeros((0, 3))
            self.seg_group = np.zeros(0, dtype=np.int32)
        self.seg_start = np.ascontiguousarray(self.seg_start, dtype=np.float64).reshape(-1, 3)
        self.seg_end = np.ascontiguousarray(self.seg_end, dtype=np.float64).reshape(-1, 3)
        self.seg_group = np.ascontiguousarray(self.seg_group, dtype=np.int32)
        if self.planes is None:
            self.planes = np.zeros((0, 6))
        if self.spheres is None:
            self.spheres = np.zeros((0, 4))
        self.planes = np.ascontiguousarray(self.planes, dtype=np.float64).reshape(-1, 6)
        self.spheres = np.ascontiguousarray(self.spheres, dtype=np.float64).reshape(-1, 4)
        if self.tris is None:
            self.tris = np.zeros((0, 9))
        self.tris = np.ascontiguousarray(self.tris, dtype=np.float64).reshape(-1, 9)

    @property
    def N(self) -> int:
        return self.x_rest.shape[0]

    @property
    def T(self) -> int:
        return self.tets.shape[0]


@dataclasses.dataclass
class DfobjInfo:
    """What the host side keeps of one reference Dfobj: its vertex and tet
    ranges in the world, its actuator type (Dfobj.h:150-160; 0 none, 2 muscle)
    and its muscles, each with ``n_sampling`` segments whose activation groups
    are ``group_offset + n_sampling * muscle + segment``."""
    vertex_offset: int
    n_vertices: int
    tet_offset: int
    n_tets: int
    actuator_type: int = 0
    n_muscles: int = 0
    n_sampling: int = MUSCLE_SEGMENTS
    group_offset: int = 0


def box_grid(nx: int, ny: int, nz: int, edge: float = EDGE, origin=(0.0, 0.0, 0.0)):
    """Axis-aligned grid of nx*ny*nz cubes, each split into 6 tets around the
    (0,0,0)-(1,1,1) diagonal (Kuhn split). Vertex (i,j,k) has index
    (i*(ny+1)+j)*(nz+1)+k."""
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    x = np.stack([ii, jj, kk], -1).reshape(-1, 3).astype(np.float64) * edge + np.asarray(origin, dtype=np.float64)

    def vid(i, j, k):
        return (i * (ny + 1) + j) * (nz + 1) + k

    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    tets = []
    for (a, b, _c) in perms:
        e = np.zeros((4, 3), dtype=np.int64)
        e[1, a] = 1
        e[2, a] = 1
        e[2, b] = 1
        e[3, :] = 1
        tets.append(np.stack([vid(ci + e[q, 0], cj + e[q, 1], ck + e[q, 2]) for q in range(4)], -1))
    tets = np.stack(tets, 1).reshape(-1, 4).astype(np.int32)
    return x, tets


def straight_fibres(nfib_y: int, nfib_z: int, lo, hi, nseg: int = 50, jitter: float = 1.3e-3):
    """Straight fibres along +x on an nfib_y x nfib_z lattice of the box
    cross-section [lo, hi], each split into ``nseg`` equal segments (the 50
    segments of Dfobj::MakeMuscles, Dfobj.h:767). A small offset keeps the
    fibres off mesh faces and edges."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    starts, ends, groups = [], [], []
    g = 0
    x0 = lo[0] + 0.25 * EDGE
    x1 = hi[0] - 0.25 * EDGE
    for a in range(nfib_y):
        for b in range(nfib_z):
            y = lo[1] + (a + 0.5) / nfib_y * (hi[1] - lo[1]) + jitter
            z = lo[2] + (b + 0.5) / nfib_z * (hi[2] - lo[2]) + 0.7 * jitter
            xs = np.linspace(x0, x1, nseg + 1)
            for s in range(nseg):
                starts.append((xs[s], y, z))
                ends.append((xs[s + 1], y, z))
                groups.append(g)
            g += 1
    return np.array(starts), np.array(ends), np.array(groups, dtype=np.int32), g


def config1() -> SceneDesc:
    """12x4x4 cantilever bar, E=1e5, nu=0.4, rho=1000 (Mass = 24 kg), x=0 face
    pinned, gravity -z, h=0.01, 10 iterations, no contact."""
    nx, ny, nz = 12, 4, 4
    x, t = box_grid(nx, ny, nz)
    pinned = np.nonzero(np.isclose(x[:, 0], 0.0))[0].astype(np.int32)
    vol = nx * ny * nz * EDGE ** 3
    return SceneDesc(x_rest=x, tets=t, total_mass=1000.0 * vol, E=np.full(t.shape[0], 1e5), nu_fwd=0.4,
                     nu_ctor=np.full(t.shape[0], 0.4), attach_idx=pinned, attach_w=1e20, h=0.01, max_iter=10,
                     damping=0.9999, gravity=(0.0, 0.0, -9.81), contact=False, name="c1_bar_12x4x4")


def _muscle_box(nx, ny, nz, seed, drop=0.1, mu=0.5, fibres=(8, 8), name="box"):
    rng = np.random.default_rng(seed)
    x, t = box_grid(nx, ny, nz, origin=(0.0, 0.0, drop))
    T = t.shape[0]
    E = rng.uniform(5e4, 2e5, size=T)
    lo = np.array([0.0, 0.0, drop])
    hi = np.array([nx * EDGE, ny * EDGE, nz * EDGE + drop])
    s0, s1, grp, ng = straight_fibres(fibres[0], fibres[1], lo, hi)
    vol = nx * ny * nz * EDGE ** 3
    return SceneDesc(x_rest=x, tets=t, total_mass=1000.0 * vol, E=E, nu_fwd=0.4, nu_ctor=np.full(T, 0.4),
                     seg_start=s0, seg_end=s1, seg_group=grp, n_groups=ng, k_muscle=5e5, h=0.01, max_iter=10,
                     damping=0.9999, gravity=(0.0, 0.0, -9.81), contact=True, max_iter1=10, alpha=1.0, mu=mu,
                     planes=np.array([[0.0, 0.0, 0.0, 0.0, 0.0, 1.0]]), name=name)


def config2(seed: int = 2) -> SceneDesc:
    """30x10x10 muscle body (18,000 tets) with per-tet E ~ U[5e4,2e5], 64
    straight fibres along x, ground z=0 with mu=0.5, dropped from 0.1 m."""
    return _muscle_box(30, 10, 10, seed, name="c2_muscle_30x10x10")


def config4(seed: int = 4) -> SceneDesc:
    """40x12x12 muscle body (34,560 tets), 64 fibres, ground mu=0.5."""
    return _muscle_box(40, 12, 12, seed, name="c4_muscle_40x12x12")


def activations_uniform(seed: int, steps: int, n_groups: int, R: int = 1) -> np.ndarray:
    """Config-2 activations: uniform [0,1] per (rollout, step, fibre)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, size=(R, steps, n_groups))


def activations_sinusoid(seed: int, steps: int, n_groups: int, R: int, h: float) -> np.ndarray:
    """Config-4 activations: a(t) = (1 + sin(2 pi f t + phi))/2 with
    f ~ U[0.5, 4] Hz and phi ~ U[0, 2 pi) per (rollout, fibre)."""
    rng = np.random.default_rng(seed)
    f = rng.uniform(0.5, 4.0, size=(R, 1, n_groups))
    ph = rng.uniform(0.0, 2 * math.pi, size=(R, 1, n_groups))
    t = (np.arange(steps, dtype=np.float64) * h)[None, :, None]
    return 0.5 * (1.0 + np.sin(2 * math.pi * f * t + ph))


def contour_faces(x_rest: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """Surface triangles of a tet mesh as Dfobj::Initialize builds mContours
    (Dfobj.h:497-585): each tet contributes the faces (s0,s1,s2), (s0,s1,s3),
    (s0,s2,s3), (s1,s2,s3) of its sorted vertex ids, in tet order; a face no
    other tet shares is kept and turned outward (q0 and q1 swap when
    n = (q1 - q0) x (q2 - q1) points towards its tet's centre). Unshared faces
    are found by sorting the 4T face keys instead of the all-pairs scan."""
    x = np.asarray(x_rest, np.float64).reshape(-1, 3)
    t = np.asarray(tets, np.int64).reshape(-1, 4)
    st = np.sort(t, axis=1)
    tri = st[:, [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]].reshape(-1, 3)
    _, inv, cnt = np.unique(tri, axis=0, return_inverse=True, return_counts=True)
    keep = np.flatnonzero(cnt[inv.ravel()] == 1)
    f = tri[keep].copy()
    centre = x[t[keep // 4]].mean(axis=1)
    q0, q1, q2 = x[f[:, 0]], x[f[:, 1]], x[f[:, 2]]
    n = np.cross(q1 - q0, q2 - q1)
    flip = np.einsum("ij,ij->i", centre - q0, n) > 0.0
    f[flip, 0], f[flip, 1] = f[flip, 1], f[flip, 0].copy()
    return f.astype(np.int32)


def load_obj_tet(path: str, scale: float = 0.05):
    """``v x y z`` / ``t i0 i1 i2 i3`` (0-based) tet OBJ, vertices scaled by
    ``scale`` (OBJMesh.h:40-97; Dfobj.h:163-190 default Scale 0.05). Lines of
    other kinds are skipped, as the reference's reader does."""
    if not os.path.exists(path):
        raise ValueError("Can't read file " + path)
    vs, ts = [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                vs.append([float(p[1]), float(p[2]), float(p[3])])
            elif p[0] == "t":
                ts.append([int(p[1]), int(p[2]), int(p[3]), int(p[4])])
    x = np.asarray(vs, dtype=np.float64).reshape(-1, 3) * scale
    t = np.asarray(ts, dtype=np.int32).reshape(-1, 4)
    if t.size and (t.min() < 0 or t.max() >= x.shape[0]):
        raise ValueError("%s: tet vertex index out of range" % path)
    return x, t


def load_obj_triangles(path: str) -> np.ndarray:
    """Static obstacle triangles of an OBJ surface mesh, (F, 9) float64: ``v``
    lines and ``f`` polygons whose vertex tokens are ``i``, ``i/t``, ``i//n``
    or ``i/t/n`` with 1-based (or negative, relative) vertex indices; a
    polygon with k corners gives the k-2 triangles of a fan from its first
    corner. The reference reads obstacles unscaled at offset 0
    (Dfobstacle.h:110-160) with the OBJ reader of OBJMesh.h:40-97."""
    if not os.path.exists(path):
        raise ValueError("Can't read file " + path)
    vs, faces = [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                vs.append([float(p[1]), float(p[2]), float(p[3])])
            elif p[0] == "f":
                faces.append([int(tok.split("/")[0]) for tok in p[1:]])
    x = np.asarray(vs, dtype=np.float64).reshape(-1, 3)
    out = []
    for face in faces:
        if len(face) < 3:
            raise ValueError("%s: face with fewer than 3 corners" % path)
        ids = [i - 1 if i > 0 else x.shape[0] + i for i in face]
        if min(ids) < 0 or max(ids) >= x.shape[0] or 0 in face:
            raise ValueError("%s: face vertex index out of range" % path)
        for k in range(1, len(ids) - 1):
            out.append(np.concatenate([x[ids[0]], x[ids[k]], x[ids[k + 1]]]))
    return np.asarray(out, dtype=np.float64).reshape(-1, 9)


def load_muscle_file(path: str, scale: float = 0.05) -> List[np.ndarray]:
    """Fibre polylines of a ``.muscle`` file (Dfobj::MakeMuscles,
    Dfobj.h:734-797): one point per line, ``fibre segment x y z``; a fibre
    ends where the fibre index changes; points are scaled like the mesh
    (``mScalingMatrix``). Blank lines are skipped. The reader starts at fibre
    index 0, so a file whose first fibre has another index would give an empty
    first fibre (undefined in the reference): that is an error here."""
    if not os.path.exists(path):
        raise ValueError("Muscle file doesn't exist: " + path)
    fibres: List[list] = []
    cur, pts = 0, []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if len(p) < 5:
                raise ValueError("%s: malformed muscle line %r" % (path, line))
            idx = int(p[0])
            if idx != cur:
                if not pts:
                    raise ValueError("%s: fibre %d has no points" % (path, cur))
                fibres.append(pts)
                cur, pts = idx, []
            pts.append([float(p[2]), float(p[3]), float(p[4])])
    if pts:
        fibres.append(pts)
    return [np.asarray(q, dtype=np.float64) * scale for q in fibres]


def resample_fibre(points: np.ndarray, nseg: int = MUSCLE_SEGMENTS):
    """Segments of one fibre polyline (Muscle::Muscle, Muscle.h:59-107): walk
    the polyline accumulating edge lengths; when the sum exceeds
    arc_length / nseg, close a segment from the last start point to the
    current point and restart the sum there. The last end point is replaced by
    the polyline's last point.

    The reference loop also adds the edge that starts at the last point, which
    reads one point past the end (Muscle.h:92-93). Here that step always
    closes a segment ending at the last point: it is the one reading under
    which a fibre such as the octopus's (99 evenly spaced points) has the
    ``nseg`` segments that Muscle::Initialize reads; any other reading leaves
    it one short, and Initialize then reads past the segment list. A fibre
    with fewer than ``nseg`` segments is rejected; extra segments are unused,
    as Initialize uses the first ``nseg``."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    n = pts.shape[0]
    if n < 2:
        raise ValueError("a muscle fibre needs at least two points")
    edge = np.linalg.norm(pts[1:] - pts[:-1], axis=1)
    arc = 0.0
    for e in edge:
        arc += e
    samp = arc / nseg
    starts, ends = [], []
    acc, si = 0.0, 0
    for i in range(n - 1):
        acc += edge[i]
        if acc > samp:
            starts.append(pts[si])
            ends.append(pts[i])
            acc, si = 0.0, i
    starts.append(pts[si])
    ends.append(pts[n - 1])
    if len(starts) < nseg:
        raise ValueError("muscle fibre resamples to %d segments, fewer than %d" % (len(starts), nseg))
    return np.asarray(starts[:nseg]), np.asarray(ends[:nseg])


def _unsupported(what: str):
    raise NotImplementedError("scene feature not supported by this engine: " + what)


def _json_len(obj: dict, key: str) -> int:
    """Size of an optional JSON array: a missing key reads as empty (the
    reference is built with NDEBUG, and rapidjson's null value then has size 0,
    document.h:1163-1177)."""
    v = obj.get(key)
    return len(v) if isinstance(v, list) else 0


def load_json_scene(config_path: str, data_root: Optional[str] = None) -> SceneDesc:
    """Reference JSON scene (Environment.h:85-225, Dfobj.h:142-314,
    Dfobj.h:442-611): the "Parameters" block and every "Meshes" entry, each
    mesh a body of the one world with its own mass, material, attachments and
    muscle file. Missing keys take the reference's defaults (Dfobj(5E5, 2E5,
    0.4), Mass 10, Scale 0.05, Simulation HZ 200, Max Iteration 30, damping
    0.9999, gravity 0); a missing "Attach Stiffness" next to "Attach Verts"
    reads 0, as a null rapidjson value does in the reference's NDEBUG build.

    ``config_path`` is resolved against ``data_root`` (the reference prefixes
    SOFTCON_DIR, Environment.h:90), mesh and muscle files against
    ``<data_root>/data`` (Dfobj.h:192-199, 724-731), falling back to the
    config's own folder. "Obstacles" are static triangle meshes (OBJ
    ``f`` polygons, fan-triangulated). Scene features outside the
    projective-dynamics path (cloth, rigid bones, pneumatic volume muscles,
    skeleton or pneumatic actuators, drag on several meshes) raise
    NotImplementedError. "Self-collision" is simulated forward; its reverse
    sweep is not available."""
    root = data_root if data_root is not None else os.environ.get("PDM_DATA_ROOT", "")
    path = os.path.join(root, config_path.lstrip("/")) if root else config_path
    if not os.path.exists(path):
        raise ValueError("Couldn't open file " + path)
    with open(path) as f:
        text = f.read()
    try:
        js = json.loads(text)
    except json.JSONDecodeError:
        js = None
    if not isinstance(js, dict):
        raise ValueError("JSON reader - Error : root must be an object")
    para = js["Parameters"][0]
    hz = int(para.get("Simulation HZ", 200))
    max_iter = int(para.get("Max Iteration", 30))
    damping = float(para.get("Damping coefficient", 0.9999))
    contact = bool(para.get("Handle collision", False))
    g = para.get("Gravity", [0.0, 0.0, 0.0])
    if len(g) < 3:
        raise ValueError("The Gravity vector should have the length of 3 but got length %d" % len(g))
    gravity = tuple(float(v) for v in g[:3])
    alpha = float(para.get("World Alpha", 1.0))
    self_collision = bool(para.get("Self-collision", False))
    self_mu = float(para.get("World Friction", 1.0))
    underwater = bool(int(para.get("Is underwater", 0)))
    if _json_len(js, "Clothes"):
        _unsupported("Clothes")
    meshes = js.get("Meshes") or []
    if not meshes:
        raise ValueError("scene has no meshes")
    if underwater and len(meshes) > 1:
        # Dfobj::ComputeDragForces indexes the world's state with the mesh's own
        # vertex ids and returns a world-sized force for a mesh-sized segment
        # (Dfobj.h:874-881, Environment.h:247): defined for one mesh only
        _unsupported("Is underwater with several meshes")
    cfg_dir = os.path.dirname(path)
    data_dir = os.path.join(root, "data") if root else cfg_dir

    def resolve(name):
        p = os.path.join(data_dir, name)
        return p if os.path.exists(p) else os.path.join(cfg_dir, os.path.basename(name))

    obstacles = js.get("Obstacles") or []
    tris = [load_obj_triangles(resolve(o["Obj filename"])) for o in obstacles if "Obj filename" in o]
    if len(tris) != len(obstacles):
        raise ValueError("an obstacle has no \"Obj filename\"")
    tris = np.concatenate(tris) if tris else None

    bodies, infos = [], []
    nv = nt = ng = 0
    for k, m in enumerate(meshes):
        if "Obj filename" not in m:
            raise ValueError("mesh %d has no \"Obj filename\"" % k)
        act_type = int(m.get("Actuator type", 0))
        if act_type not in (0, 2):
            _unsupported("Actuator type %d" % act_type)
        if _json_len(m, "Bones"):
            _unsupported("Bones (rigid links)")
        if _json_len(m, "Volume Verts"):
            _unsupported("Volume Verts (pneumatic constraints)")
        scale = float(m.get("Scale", 0.05))
        if not scale > 0:
            raise ValueError("mesh %d: Scale must be positive" % k)
        x, t = load_obj_tet(resolve(m["Obj filename"]), scale)
        T = t.shape[0]
        E = float(m.get("Youngs Modulus", 2e5))
        nu = float(m.get("Poisson Ratio", 0.4))
        kw = {}
        n_mus = 0
        if "Muscle filename" in m:
            fibres = load_muscle_file(resolve(m["Muscle filename"]), scale)
            s0, s1, grp = [], [], []
            for i, fpts in enumerate(fibres):
                a, b = resample_fibre(fpts)
                s0.append(a)
                s1.append(b)
                grp.append(np.arange(MUSCLE_SEGMENTS, dtype=np.int32) + MUSCLE_SEGMENTS * i)
            n_mus = len(fibres)
            if n_mus:
                kw = dict(seg_start=np.concatenate(s0), seg_end=np.concatenate(s1), seg_group=np.concatenate(grp),
                          n_groups=MUSCLE_SEGMENTS * n_mus)
        attach = np.zeros(0, dtype=np.int32)
        attach_w = 0.0
        if "Attach Verts" in m:
            attach = np.asarray(sorted(set(int(a) for a in m["Attach Verts"])), dtype=np.int32)
            if attach.size and (attach.min() < 0 or attach.max() >= x.shape[0]):
                raise ValueError("mesh %d: attach vertex out of range" % k)
            attach_w = float(m.get("Attach Stiffness", 0.0) or 0.0)
        bodies.append(SceneDesc(x_rest=x, tets=t, total_mass=float(m.get("Mass", 10.0)),
                                mass_same=bool(m.get("Is Mass Same", True)), E=np.full(T, E), nu_fwd=nu,
                                nu_ctor=np.full(T, nu), attach_idx=attach, attach_w=attach_w,
                                k_muscle=float(m.get("Muscle Stiffness", 5e5)), h=1.0 / hz, max_iter=max_iter,
                                damping=damping, gravity=gravity, gravity_on=bool(int(m.get("Enable Gravity", 1))),
                                underwater=underwater, contact=contact, max_iter1=max_iter, alpha=alpha, mu=1.0,
                                tris=tris if k == 0 else None, self_collision=self_collision, self_mu=self_mu, **kw))
        infos.append(DfobjInfo(vertex_offset=nv, n_vertices=x.shape[0], tet_offset=nt, n_tets=T,
                               actuator_type=act_type, n_muscles=n_mus, group_offset=ng))
        nv += x.shape[0]
        nt += T
        ng += MUSCLE_SEGMENTS * n_mus
    name = os.path.splitext(os.path.basename(config_path))[0]
    if len(bodies) == 1:
        d = bodies[0]
        d.name = name
    else:
        d = pack_scenes(bodies, name=name, translate=False)
    d.dfobjs = infos
    return d


def config5_meshes(seed: int = 5, n_meshes: int = 64, tmin: int = 2000, tmax: int = 60000) -> List[SceneDesc]:
    """64 seeded box meshes with T ~ logU[2k,60k], E ~ logU[5e4,2e5],
    nu ~ U[0.3,0.45]; even-indexed meshes carry 16 fibres, odd ones are pure
    elastic; ground plane plus a rigid sphere r=0.2 m."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_meshes):
        target = math.exp(rng.uniform(math.log(tmin), math.log(tmax)))
        cubes = max(1, int(round(target / 6)))
        # aspect 3:1:1
        ny = max(2, int(round((cubes / 3.0) ** (1.0 / 3.0))))
        nz = ny
        nx = max(2, int(round(cubes / (ny * nz))))
        x, t = box_grid(nx, ny, nz, origin=(-0.5 * nx * EDGE, -0.5 * ny * EDGE, 0.35))
        T = t.shape[0]
        E = math.exp(rng.uniform(math.log(5e4), math.log(2e5)))
        nu = rng.uniform(0.3, 0.45)
        sph = np.array([[rng.uniform(-0.2, 0.2), rng.uniform(-0.1, 0.1), 0.1, 0.2]])
        kw = {}
        if i % 2 == 0:
            lo = x.min(0)
            hi = x.max(0)
            s0, s1, grp, ng = straight_fibres(4, 4, lo, hi)
            kw = dict(seg_start=s0, seg_end=s1, seg_group=grp, n_groups=ng)
        vol = nx * ny * nz * EDGE ** 3
        out.append(SceneDesc(x_rest=x, tets=t, total_mass=1000.0 * vol, E=np.full(T, E), nu_fwd=nu,
                             nu_ctor=np.full(T, nu), h=0.01, max_iter=10, damping=0.9999,
                             gravity=(0.0, 0.0, -9.81), contact=True, max_iter1=10, mu=0.5,
                             planes=np.array([[0.0, 0.0, 0.0, 0.0, 0.0, 1.0]]), spheres=sph,
                             name="c5_mesh%02d_%dx%dx%d" % (i, nx, ny, nz), **kw))
    return out


def pack_scenes(descs: List[SceneDesc], gap: float = 0.4, name: str = "packed", translate: bool = True) -> SceneDesc:
    """Several meshes as the bodies of one world (one Dfobj per mesh, as the
    reference's "Meshes" list loads them, Environment.h:109-210): with
    ``translate`` the meshes are moved onto a square grid of cells in the x-y
    plane, ``gap`` apart (without it they stay where they are, as the
    reference places every mesh at offset 0); each keeps its own mass, E, nu
    and muscle fibres (activation groups are renumbered consecutively);
    obstacles are the union of the meshes' (moved with their mesh). Step
    parameters, the attachment stiffness of bodies with attachments and the
    muscle stiffness of bodies with muscles must agree across meshes, as the
    engine holds one value of each.

    A fibre selects the tets it crosses or passes within about 8 mm of
    (Muscle.h:109-215), and the reference only searches its own mesh. Here the
    search covers the world, so untranslated bodies with muscles must stay
    1 cm apart (bounding boxes), which makes the two searches agree."""
    if not descs:
        raise ValueError("no meshes to pack")
    d0 = descs[0]
    for d in descs[1:]:
        for k in ("h", "max_iter", "damping", "gravity_on", "underwater", "contact", "max_iter1", "alpha", "mu",
                  "mass_same", "self_collision", "self_mu"):
            if getattr(d, k) != getattr(d0, k):
                raise ValueError("packed meshes disagree on %s" % k)
        if tuple(d.gravity) != tuple(d0.gravity):
            raise ValueError("packed meshes disagree on gravity")
    att_w = {d.attach_w for d in descs if d.attach_idx.size}
    k_mus = {d.k_muscle for d in descs if d.seg_group.size}
    if len(att_w) > 1:
        raise ValueError("packed meshes disagree on the attachment stiffness")
    if len(k_mus) > 1:
        raise ValueError("packed meshes disagree on the muscle stiffness")
    if not translate:
        mus = [d for d in descs if d.seg_group.size]
        for d in mus:
            for e in descs:
                if e is d:
                    continue
                lo = np.maximum(d.x_rest.min(0), e.x_rest.min(0))
                hi = np.minimum(d.x_rest.max(0), e.x_rest.max(0))
                if np.all(lo - hi < 0.01):
                    raise NotImplementedError("a body with muscles lies within 1 cm of another body")
    ext = np.array([d.x_rest.max(0) - d.x_rest.min(0) for d in descs])
    cell = ext[:, :2].max(0) + gap
    ncol = int(math.ceil(math.sqrt(len(descs))))
    xs, ts, bodies, E, nuc, nuf, mass = [], [], [], [], [], [], []
    offs = []
    s0, s1, sg, att, spheres = [], [], [], [], []
    nv = 0
    ng = 0
    for b, d in enumerate(descs):
        off = np.zeros(3)
        if translate:
            off = np.array([(b % ncol) * cell[0], (b // ncol) * cell[1], 0.0]) - np.array([*d.x_rest.min(0)[:2], 0.0])
        xs.append(d.x_rest + off)
        ts.append(d.tets + nv)
        bodies.append(np.full(d.T, b, dtype=np.int32))
        E.append(d.E)
        nuc.append(d.nu_ctor)
        nuf.append(d.nu_fwd_tet if d.nu_fwd_tet is not None else np.full(d.T, d.nu_fwd))
        mass.append(d.total_mass)
        att.append(d.attach_idx + nv)
        if d.seg_group.size:
            s0.append(d.seg_start + off)
            s1.append(d.seg_end + off)
            sg.append(d.seg_group + ng)
            ng += d.n_groups
        for sp in d.spheres:
            spheres.append(np.concatenate([sp[:3] + off, sp[3:]]))
        offs.append(off)
        nv += d.N
    planes = np.unique(np.concatenate([d.planes for d in descs]), axis=0) if any(d.planes.size for d in descs) else None
    tris = [d.tris + np.tile(offs[b], 3) for b, d in enumerate(descs) if d.tris.size]
    kw = {}
    if sg:
        kw = dict(seg_start=np.concatenate(s0), seg_end=np.concatenate(s1), seg_group=np.concatenate(sg), n_groups=ng)
    return SceneDesc(x_rest=np.concatenate(xs), tets=np.concatenate(ts), total_mass=float(sum(mass)),
                     mass_same=d0.mass_same, E=np.concatenate(E), nu_fwd=d0.nu_fwd, nu_ctor=np.concatenate(nuc),
                     attach_idx=np.concatenate(att), attach_w=att_w.pop() if att_w else d0.attach_w,
                     k_muscle=k_mus.pop() if k_mus else d0.k_muscle, h=d0.h,
                     max_iter=d0.max_iter, damping=d0.damping, gravity=d0.gravity, gravity_on=d0.gravity_on,
                     underwater=d0.underwater, self_collision=d0.self_collision, self_mu=d0.self_mu,
                     contact=d0.contact, max_iter1=d0.max_iter1, alpha=d0.alpha, mu=d0.mu, planes=planes,
                     spheres=np.array(spheres) if spheres else None, tris=np.concatenate(tris) if tris else None,
                     body_of_tet=np.concatenate(bodies),
                     body_mass=np.array(mass), nu_fwd_tet=np.concatenate(nuf), name=name, **kw)


def config5(seed: int = 5, n_meshes: int = 64, tmin: int = 2000, tmax: int = 60000) -> SceneDesc:
    """The packed uneven-mesh workload: ``config5_meshes`` as the bodies of
    one world (block-diagonal factor, one supernodal forest)."""
    return pack_scenes(config5_meshes(seed, n_meshes, tmin, tmax), name="c5_packed_%d" % n_meshes)
