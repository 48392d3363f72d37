"""Synthetic workloads of the implicit-step benchmark (seeded, reproducible).

  box_small       20x4x4 cubes, 1.0x0.2x0.2 m, corotational E=1e5 nu=0.45, dt=1/60, 100 steps
  box_mid_neo     80x16x16 cubes, 2.0x0.4x0.4 m, neo-Hookean E=5e4 nu=0.40, tip impulse, dt=1/120, 200 steps
  box_1m_dof      200x40x40 cubes, 5x1x1 m, corotational per-tet E log-uniform [1e4,1e6], dt=1/60, 50 steps
  batch_256       256 boxes, tets log-uniform [500, 5e4], aspect U(1,6), mixed materials, dt=1/60, 120 steps
  served_pool     pool of boxes, tets log-uniform [lo, hi], mixed materials (request stream in serving.py)

All use rho = 1000, the x = 0 face fixed, gravity -9.81 in z.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import mesh as M

GRAVITY = np.array([0.0, 0.0, -9.81])


@dataclass
class Workload:
    name: str
    X: np.ndarray
    tets: np.ndarray
    facets: np.ndarray
    fixed_nodes: np.ndarray
    kind: np.ndarray
    E: np.ndarray
    nu: np.ndarray
    rho: np.ndarray
    dt: float
    steps: int
    inst_node_start: np.ndarray | None = None
    inst_tet_start: np.ndarray | None = None
    impulse_node_forces: np.ndarray | None = None   # (nNodes, 3) applied at step 0 only
    shapes: list = field(default_factory=list)
    meta: dict = field(default_factory=dict)
    # axes held on the fixed nodes: all three, or one or two for the reference's partial
    # constraint types (Constrain_X .. Constrain_XZ, core/constrainer.h:16-22)
    constrain_axes: tuple = (0, 1, 2)

    @property
    def n_nodes(self):
        return self.X.shape[0]

    @property
    def n_tets(self):
        return self.tets.shape[0]

    @property
    def fixed_dofs(self):
        if tuple(self.constrain_axes) == (0, 1, 2):
            return M.fixed_dofs_from_nodes(self.fixed_nodes)
        n = np.asarray(self.fixed_nodes, np.int64)
        return np.sort((3 * n[:, None] + np.asarray(self.constrain_axes)[None, :]).ravel()).astype(np.int32)

    @property
    def body(self):
        """Per-tet body force rho * g (Forcer::addBodyForce input)."""
        return self.rho[:, None] * GRAVITY[None, :]

    def node_to_free(self, node_vec):
        """Node-space (nNodes,3) vector -> free-DOF vector (whole-node constraints)."""
        free = np.ones(self.n_nodes, bool)
        free[self.fixed_nodes] = False
        return np.ascontiguousarray(np.asarray(node_vec, np.float64).reshape(-1, 3)[free].ravel())

    def free_to_node(self, d):
        out = np.zeros((self.n_nodes, 3))
        free = np.ones(self.n_nodes, bool)
        free[self.fixed_nodes] = False
        out[free] = np.asarray(d).reshape(-1, 3)
        return out

    @property
    def impulse(self):
        if self.impulse_node_forces is None:
            return None
        return self.node_to_free(self.impulse_node_forces)


def _color_perm(m: M.TetMesh, ordering: str) -> np.ndarray:
    """Node permutation of a multicolour ordering: color8 / color5 / color4, or color4t (4
    colours, each colour's chunks compact tiles; the x = 0 face, fixed by the workloads, last)."""
    if ordering == "color4t":
        return M.tiled_color4_permutation(m, M.plane_fixed_nodes(m, axis=0, value=0.0))
    return M.multicolor_permutation(m, int(ordering[len("color"):]))


def _single(name, m: M.TetMesh, kind, E, nu, dt, steps, ordering="natural", impulse=None, meta=None):
    if ordering in ("color8", "color5", "color4", "color4t"):
        perm = _color_perm(m, ordering)
        m = M.relabel(m, perm)
        if impulse is not None:
            imp = np.zeros_like(impulse)
            imp[perm] = impulse
            impulse = imp
    elif ordering == "gps":
        m = M.gps_relabel(m)
        if impulse is not None:
            imp = np.zeros_like(impulse)
            imp[m.extra["perm"]] = impulse
            impulse = imp
    elif ordering != "natural":
        raise ValueError(f"unknown ordering {ordering}")
    fixed = M.plane_fixed_nodes(m, axis=0, value=0.0)
    n = m.n_tets
    return Workload(name=name, X=m.X, tets=m.tets, facets=m.facets, fixed_nodes=fixed,
                    kind=np.broadcast_to(np.int32(kind), (n,)).copy() if np.ndim(kind) == 0 else np.asarray(kind, np.int32),
                    E=np.broadcast_to(np.float64(E), (n,)).copy() if np.ndim(E) == 0 else np.asarray(E, np.float64),
                    nu=np.full(n, float(nu)), rho=np.full(n, 1000.0), dt=dt, steps=steps,
                    impulse_node_forces=impulse, shapes=[m.shape], meta=dict(meta or {}, ordering=ordering))


CONSTRAIN_AXES = {"X": (0,), "Y": (1,), "Z": (2,), "XY": (0, 1), "YZ": (1, 2), "XZ": (0, 2), "XYZ": (0, 1, 2)}


def box_small(ordering="natural", seed=1, constrain="XYZ"):
    """constrain: the axes held on the x = 0 face, a ConstrainType name (X .. XYZ)."""
    m = M.box_mesh(20, 4, 4, 1.0, 0.2, 0.2, seed=seed)
    w = _single("box_small_20x4x4_corot", m, 0, 1e5, 0.45, 1.0 / 60, 100, ordering)
    w.constrain_axes = CONSTRAIN_AXES[constrain]
    if constrain != "XYZ":
        w.name += f"_constrain{constrain}"
    return w


def tip_impulse(m: M.TetMesh, seed, max_newton=50.0):
    u = M.uniform01(seed, M.STREAM_IMPULSE, np.arange(4, dtype=np.uint64))
    mag = max_newton * u[0]
    z = 2.0 * u[1] - 1.0
    phi = 2.0 * math.pi * u[2]
    r = math.sqrt(max(0.0, 1.0 - z * z))
    direction = np.array([r * math.cos(phi), r * math.sin(phi), z])
    face = M.plane_fixed_nodes(m, axis=0, value=m.size[0])
    f = np.zeros((m.n_nodes, 3))
    f[face] = direction * (mag / len(face))
    return f


def box_mid_neo(ordering="natural", seed=2):
    m = M.box_mesh(80, 16, 16, 2.0, 0.4, 0.4, seed=seed)
    imp = tip_impulse(m, seed)
    return _single("box_mid_80x16x16_neohookean", m, 1, 5e4, 0.40, 1.0 / 120, 200, ordering, impulse=imp)


def box_1m_dof(ordering="natural", seed=3, scale=1):
    """scale < 1 shrinks the grid (for tests); scale = 1 is the 200x40x40 workload."""
    nx, ny, nz = max(2, int(200 * scale)), max(2, int(40 * scale)), max(2, int(40 * scale))
    m = M.box_mesh(nx, ny, nz, 5.0, 1.0, 1.0, seed=seed)
    E = M.log_uniform(seed, M.STREAM_YOUNG, np.arange(m.n_tets, dtype=np.uint64), 1e4, 1e6)
    return _single(f"box_{nx}x{ny}x{nz}_corot_perTetE", m, 0, E, 0.45, 1.0 / 60, 50, ordering)


def instance_shape(seed, tet_lo, tet_hi):
    u = M.uniform01(seed, M.STREAM_INSTANCE, np.arange(8, dtype=np.uint64))
    T = math.exp(math.log(tet_lo) + u[0] * (math.log(tet_hi) - math.log(tet_lo)))
    a = 1.0 + 5.0 * u[1]
    ny = max(1, int(round((T / (6.0 * a)) ** (1.0 / 3.0))))
    nx = max(1, int(round(a * ny)))
    kind = 0 if u[2] < 0.5 else 1
    E = math.exp(math.log(1e4) + u[3] * (math.log(1e6) - math.log(1e4)))
    nu = 0.3 + 0.18 * u[4]
    return nx, ny, ny, kind, E, nu


def batch(n_inst=256, seed=4, tet_lo=500, tet_hi=50000, cell=0.05, steps=120, ordering="natural", name=None):
    """Concatenation of n_inst independent boxes (one face fixed each)."""
    Xs, Ts, Fs, fixed, kinds, Es, nus, shapes = [], [], [], [], [], [], [], []
    node_start, tet_start = [0], [0]
    for i in range(n_inst):
        s = seed * 100003 + i
        nx, ny, nz, kind, E, nu = instance_shape(s, tet_lo, tet_hi)
        m = M.box_mesh(nx, ny, nz, nx * cell, ny * cell, nz * cell, seed=s)
        if ordering in ("color8", "color5", "color4", "color4t"):
            m = M.relabel(m, _color_perm(m, ordering))
        off = node_start[-1]
        Xs.append(m.X); Ts.append(m.tets + off); Fs.append(m.facets + off)
        fixed.append(M.plane_fixed_nodes(m, axis=0, value=0.0) + off)
        kinds.append(np.full(m.n_tets, kind, np.int32)); Es.append(np.full(m.n_tets, E)); nus.append(np.full(m.n_tets, nu))
        node_start.append(off + m.n_nodes); tet_start.append(tet_start[-1] + m.n_tets)
        shapes.append((nx, ny, nz))
    n = tet_start[-1]
    return Workload(name=name or f"batch_{n_inst}", X=np.concatenate(Xs), tets=np.concatenate(Ts).astype(np.int32),
                    facets=np.concatenate(Fs).astype(np.int32), fixed_nodes=np.concatenate(fixed).astype(np.int32),
                    kind=np.concatenate(kinds), E=np.concatenate(Es), nu=np.concatenate(nus), rho=np.full(n, 1000.0),
                    dt=1.0 / 60, steps=steps, inst_node_start=np.array(node_start, np.int32),
                    inst_tet_start=np.array(tet_start, np.int32), shapes=shapes, meta={"ordering": ordering})


def subset(w: Workload, i: int) -> Workload:
    """Instance i of a batch as a stand-alone workload (node/tet ids rebased)."""
    n0, n1 = int(w.inst_node_start[i]), int(w.inst_node_start[i + 1])
    t0, t1 = int(w.inst_tet_start[i]), int(w.inst_tet_start[i + 1])
    fixed = w.fixed_nodes[(w.fixed_nodes >= n0) & (w.fixed_nodes < n1)] - n0
    return Workload(name=f"{w.name}[{i}]", X=w.X[n0:n1], tets=(w.tets[t0:t1] - n0).astype(np.int32), facets=w.facets[:0],
                    fixed_nodes=fixed.astype(np.int32), kind=w.kind[t0:t1], E=w.E[t0:t1], nu=w.nu[t0:t1],
                    rho=w.rho[t0:t1], dt=w.dt, steps=w.steps, shapes=[w.shapes[i]], meta=dict(w.meta))


def served_pool(n_obj=1024, seed=5, tet_lo=1000, tet_hi=500000, cell=0.05, ordering="color4t"):
    """Pool of independent boxes for the serving stream: tets log-uniform in [tet_lo, tet_hi],
    aspect U(1, 6), corotational or neo-Hookean, E log-uniform [1e4, 1e6], nu U(0.3, 0.48),
    the x = 0 face fixed, gravity.  Nodes are numbered by the tiled 4-colouring by default (4
    IC(0) dependency levels per object, 64-row chunks that group by tile)."""
    objs = []
    for i in range(n_obj):
        s = seed * 100003 + i
        nx, ny, nz, kind, E, nu = instance_shape(s, tet_lo, tet_hi)
        m = M.box_mesh(nx, ny, nz, nx * cell, ny * cell, nz * cell, seed=s)
        if ordering in ("color8", "color5", "color4", "color4t"):
            m = M.relabel(m, _color_perm(m, ordering))
        elif ordering != "natural":
            raise ValueError(f"unknown ordering {ordering}")
        fixed = M.plane_fixed_nodes(m, axis=0, value=0.0)
        n = m.n_tets
        objs.append(Workload(name=f"pool{seed}[{i}]", X=m.X, tets=m.tets, facets=m.facets, fixed_nodes=fixed,
                             kind=np.full(n, kind, np.int32), E=np.full(n, E), nu=np.full(n, nu),
                             rho=np.full(n, 1000.0), dt=1.0 / 60, steps=0, shapes=[m.shape]))
    return objs


def served_requests(pool, rate, duration, seed=6, max_steps=8, max_force=100.0, max_nodes=32):
    """Poisson arrivals at `rate` requests/s over `duration` s.  Each request advances a
    uniformly chosen object 1..max_steps steps (uniform) with an impulse of U(0, max_force) N
    split over 1..max_nodes surface nodes (not fixed), direction uniform on the sphere.
    Returns (obj, steps, arrival, impulses) with impulses[k] = [(node, (fx, fy, fz)), ...]."""
    rng = np.random.default_rng(seed)
    t, arrival = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        arrival.append(t)
    k = len(arrival)
    obj = rng.integers(0, len(pool), k).astype(np.int32)
    steps = rng.integers(1, max_steps + 1, k).astype(np.int32)
    surface = []
    for w in pool:
        s = np.setdiff1d(np.unique(w.facets.ravel()), w.fixed_nodes)
        surface.append(s)
    impulses = []
    for q in range(k):
        cand = surface[obj[q]]
        nn = int(rng.integers(1, max_nodes + 1))
        nodes = rng.choice(cand, size=min(nn, cand.size), replace=False)
        mag = rng.uniform(0.0, max_force)
        d = rng.standard_normal(3); d /= np.linalg.norm(d)
        f = d * mag / nodes.size
        impulses.append([(int(n), tuple(f)) for n in nodes])
    return obj, steps, np.asarray(arrival), impulses


def plane_contacts(X, tets, normal, constant, eps=1e-6):
    """(plane, tet) collision pairs of a static plane n.x + c = 0: every tet with a vertex at
    n.x + c < eps, in tet order, as a MeshPlaneCollisionDampingForcer caller would set them
    (collision/penaltyDampingForcer.h:25-38).  Returns (planes (k, 4), tets (k,))."""
    n = np.asarray(normal, np.float64)
    below = (np.asarray(X, np.float64) @ n + constant) < eps
    hit = np.nonzero(below[np.asarray(tets)].any(axis=1))[0].astype(np.int32)
    planes = np.tile(np.r_[n, constant], (hit.size, 1))
    return planes, hit
