"""ctypes front-end of the CPU FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product path never does.  The C++ restatement it loads is
oracle/odoracle.cpp (built by oracle/Makefile into oracle/liboracle.so).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PRE_NONE, PRE_JACOBI, PRE_ICT, PRE_IC0, PRE_LDLT = 0, 1, 2, 3, 4


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, ip, dp, i, d = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int, C.c_double
        L.oro_system_create.restype = vp
        L.oro_system_create.argtypes = [i, dp, i, ip, i, ip]
        L.oro_system_destroy.argtypes = [vp]
        L.oro_set_threads.argtypes = [i]
        L.oro_set_threads.restype = i
        L.oro_system_sizes.argtypes = [vp, C.POINTER(C.c_longlong)]
        L.oro_pattern.argtypes = [vp, ip, ip, ip, ip]
        L.oro_global_index.argtypes = [vp, i, i]
        L.oro_constrain_plane.argtypes = [i, dp, i, ip, i, dp, dp, i, ip, i]
        L.oro_set_displacements.argtypes = [vp, dp]
        L.oro_set_free_displacements.argtypes = [vp, dp]
        L.oro_mass.argtypes = [vp, dp, dp]
        L.oro_body_force.argtypes = [vp, dp, dp]
        L.oro_nodal_force.argtypes = [vp, dp, i, ip, dp]
        L.oro_assemble.argtypes = [vp, ip, dp, dp, d, d, dp, dp]
        L.oro_set_plastic.argtypes = [vp, dp, dp, dp]
        L.oro_get_plastic_strain.argtypes = [vp, dp]
        L.oro_spmdv.argtypes = [vp, dp, dp, dp]
        L.oro_get_diagonal.argtypes = [vp, dp, dp]
        L.oro_cg.argtypes = [vp, dp, dp, dp, d, i, i, d, d, dp]
        L.oro_precond_apply.argtypes = [vp, dp, i, d, d, dp, dp]
        L.oro_ic0_factor.argtypes = [vp, dp, d, dp]
        L.oro_ict_dense.argtypes = [vp, dp, d, d, dp]
        L.oro_ldlt_dense.argtypes = [vp, dp, d, d, dp, dp]
        L.oro_polar.argtypes = [dp, dp, dp]
        L.oro_qr.argtypes = [dp, dp, dp]
        L.oro_psd3.argtypes = [dp]
        L.oro_psd2.argtypes = [dp, dp]
        L.oro_eigsym.argtypes = [dp, dp, dp]
        L.oro_corot_element.argtypes = [dp, dp, d, d, d, dp, dp, dp]
        L.oro_corot_k0.argtypes = [dp, d, d, dp]
        L.oro_neo_element.argtypes = [dp, dp, d, d, d, dp, dp]
        L.oro_neo_stress.argtypes = [dp, d, d, d, dp, dp]
        L.oro_newmark_create.restype = vp
        L.oro_newmark_create.argtypes = [vp, ip, dp, dp, d, d, dp, d, d, d, d, d, i, d, d, d, i, i]
        L.oro_newmark_destroy.argtypes = [vp]
        L.oro_newmark_set_external.argtypes = [vp, dp]
        L.oro_newmark_set_impulse.argtypes = [vp, dp]
        L.oro_newmark_set_state.argtypes = [vp, dp, dp, dp, i]
        L.oro_newmark_step.argtypes = [vp, i, ip, dp]
        L.oro_newmark_get_state.argtypes = [vp, dp, dp, dp]
        L.oro_newmark_mass.argtypes = [vp, dp]
        L.oro_be_create.restype = vp
        L.oro_be_create.argtypes = [vp, ip, dp, dp, d, d, dp, d, d, d, i, d, d, d, i]
        L.oro_be_destroy.argtypes = [vp]
        L.oro_be_set_external.argtypes = [vp, dp]
        L.oro_be_step.argtypes = [vp]
        L.oro_be_update_mesh.argtypes = [vp]
        L.oro_be_get_velocity.argtypes = [vp, dp]
        L.oro_get_node_displacements.argtypes = [vp, dp]
        L.oro_plane_damping.argtypes = [vp, i, dp, ip, d, dp, dp]
        L.oro_be_add_plane_damping.argtypes = [vp, i, dp, ip, d]
        L.oro_be_clear_damping.argtypes = [vp]
        L.oro_be_damping.argtypes = [vp, dp]
        L.oro_be_set_velocity.argtypes = [vp, dp]
        L.oro_plane_tet_intersection.argtypes = [dp, dp, dp, dp]
        L.oro_barycentric.argtypes = [dp, dp, dp]
        L.oro_gps.argtypes = [i, i, i, ip, ip]
        _LIB = L
    return _LIB


def set_threads(n: int) -> int:
    """Threads of the oracle's per-tet element evaluation (1 = the reference's single thread;
    the scatter stays sequential in tet order, so results are the same bits)."""
    return lib().oro_set_threads(int(n))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def constrain_plane(X, facets, origin, normal, ctype=6):
    X = f64(X); facets = i32(facets)
    cap = X.shape[0] * 3
    out = np.zeros(cap, np.int32)
    n = lib().oro_constrain_plane(X.shape[0], _d(X), facets.shape[0], _i(facets), facets.shape[1],
                                  _d(f64(origin)), _d(f64(normal)), ctype, _i(out), cap)
    return out[:n].copy()


class System:
    """Mesh + NodeIndexer + BlockedSymSparseMatrix pattern + 78-entry element map."""

    def __init__(self, X, tets, constrained_dofs):
        self.X = f64(X); self.tets = i32(tets); self.con = i32(constrained_dofs)
        self.n_nodes = self.X.shape[0]; self.n_tets = self.tets.shape[0]
        self.h = lib().oro_system_create(self.n_nodes, _d(self.X), self.n_tets, _i(self.tets),
                                         self.con.shape[0], _i(self.con))
        s = np.zeros(5, np.int64)
        lib().oro_system_sizes(self.h, s.ctypes.data_as(C.POINTER(C.c_longlong)))
        self.ndof, self.n_block_col, self.n_remained, self.n_blocks, self.n_values = (int(v) for v in s)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oro_system_destroy(self.h)
            self.h = None

    def pattern(self, with_map=True):
        nc = self.n_block_col + self.n_remained + 1
        pcol = np.zeros(nc, np.int32); oris = np.zeros(nc, np.int32); rows = np.zeros(self.n_blocks, np.int32)
        emap = np.zeros(78 * self.n_tets, np.int32) if with_map else None
        lib().oro_pattern(self.h, _i(pcol), _i(oris), _i(rows), _i(emap) if with_map else None)
        return pcol, oris, rows, (emap.reshape(self.n_tets, 78) if with_map else None)

    def global_index(self, node, axis):
        return lib().oro_global_index(self.h, int(node), int(axis))

    def mass(self, rho):
        rho = f64(np.broadcast_to(rho, (self.n_tets,)))
        M = np.zeros(self.n_values)
        lib().oro_mass(self.h, _d(rho), _d(M))
        return M

    def body_force(self, body):
        body = f64(np.broadcast_to(np.asarray(body, np.float64).reshape(-1, 3), (self.n_tets, 3)))
        f = np.zeros(self.ndof)
        lib().oro_body_force(self.h, _d(body), _d(f))
        return f

    def nodal_force(self, f3, nodes, out=None):
        out = np.zeros(self.ndof) if out is None else out
        nodes = i32(nodes)
        lib().oro_nodal_force(self.h, _d(f64(f3)), nodes.shape[0], _i(nodes), _d(out))
        return out

    def set_node_displacements(self, u):
        lib().oro_set_displacements(self.h, _d(f64(u).reshape(-1)))

    def set_free_displacements(self, d):
        lib().oro_set_free_displacements(self.h, _d(f64(d)))

    def node_displacements(self):
        u = np.zeros(3 * self.n_nodes)
        lib().oro_get_node_displacements(self.h, _d(u))
        return u.reshape(-1, 3)

    def set_plastic(self, yield_criterion, creep, max_strain):
        """Corotational plastic parameters per tet (kind 2); resets the plastic strain."""
        n = self.n_tets
        y = f64(np.broadcast_to(yield_criterion, (n,))); c = f64(np.broadcast_to(creep, (n,)))
        mx = f64(np.broadcast_to(max_strain, (n,)))
        lib().oro_set_plastic(self.h, _d(y), _d(c), _d(mx))

    def plastic_strain(self):
        out = np.zeros(9 * self.n_tets)
        lib().oro_get_plastic_strain(self.h, _d(out))
        return out.reshape(-1, 3, 3)

    def plane_damping(self, planes, tets, coeff, vel_nodes, values=None):
        """Adds MeshPlaneCollisionDampingForcer::addDampingMatrix of each (plane, tet) pair, in
        order, into values (positions: rest + current node displacements)."""
        values = np.zeros(self.n_values) if values is None else values
        planes = f64(np.asarray(planes, np.float64).reshape(-1, 4)); tets = i32(tets)
        lib().oro_plane_damping(self.h, len(tets), _d(planes), _i(tets), coeff, _d(f64(np.asarray(vel_nodes).reshape(-1))),
                                _d(values))
        return values

    def assemble(self, kind, E, nu, thr_corot=0.0, thr_neo=0.1):
        kind = i32(np.broadcast_to(kind, (self.n_tets,)))
        E = f64(np.broadcast_to(E, (self.n_tets,))); nu = f64(np.broadcast_to(nu, (self.n_tets,)))
        K = np.zeros(self.n_values); f = np.zeros(self.ndof)
        lib().oro_assemble(self.h, _i(kind), _d(E), _d(nu), thr_corot, thr_neo, _d(K), _d(f))
        return K, f

    def spmdv(self, values, x, y=None):
        y = np.zeros(self.ndof) if y is None else y
        lib().oro_spmdv(self.h, _d(f64(values)), _d(f64(x)), _d(y))
        return y

    def diagonal(self, values):
        d = np.zeros(self.ndof)
        lib().oro_get_diagonal(self.h, _d(f64(values)), _d(d))
        return d

    def cg(self, values, rhs, x0=None, tol=1e-6, max_iter=10000, pre=PRE_JACOBI, p1=0.0, p2=0.0):
        x = np.zeros(self.ndof) if x0 is None else f64(x0).copy()
        rr = C.c_double(0)
        it = lib().oro_cg(self.h, _d(f64(values)), _d(f64(rhs)), _d(x), tol, max_iter, pre, p1, p2, C.byref(rr))
        return x, it, rr.value

    def precond_apply(self, values, rhs, pre, p1=0.0, p2=0.0):
        out = np.zeros(self.ndof)
        lib().oro_precond_apply(self.h, _d(f64(values)), pre, p1, p2, _d(f64(rhs)), _d(out))
        return out

    def ic0(self, values, pivot_eps=1e-3):
        L = np.zeros(self.n_values)
        lib().oro_ic0_factor(self.h, _d(f64(values)), pivot_eps, _d(L))
        return L

    def ict_dense(self, values, eps, relax):
        L = np.zeros((self.ndof, self.ndof))
        lib().oro_ict_dense(self.h, _d(f64(values)), eps, relax, _d(L))
        return L

    def ldlt_dense(self, values, se, le):
        L = np.zeros((self.ndof, self.ndof)); D = np.zeros(self.ndof)
        lib().oro_ldlt_dense(self.h, _d(f64(values)), se, le, _d(L), _d(D))
        return L, D

    def dense(self, values):
        """Full symmetric dense matrix from canonical values (via the element-pair index map)."""
        n = self.ndof
        A = np.zeros((n, n))
        I = np.eye(n)
        for j in range(n):
            A[:, j] = self.spmdv(values, I[:, j])
        return A


class Newmark:
    def __init__(self, system: System, kind, E, nu, rho, dt, beta=0.25, gamma=0.5, alpha_m=0.0, alpha_k=0.0,
                 pre=PRE_IC0, p1=1e-3, p2=0.0, tol=1e-6, max_iter=10000, thr_corot=0.0, thr_neo=0.1, init_accel=True):
        self.sys = system
        n = system.n_tets
        self.kind = i32(np.broadcast_to(kind, (n,))); self.E = f64(np.broadcast_to(E, (n,)))
        self.nu = f64(np.broadcast_to(nu, (n,))); self.rho = f64(np.broadcast_to(rho, (n,)))
        self.h = lib().oro_newmark_create(system.h, _i(self.kind), _d(self.E), _d(self.nu), thr_corot, thr_neo,
                                          _d(self.rho), beta, gamma, dt, alpha_m, alpha_k, pre, p1, p2, tol, max_iter,
                                          1 if init_accel else 0)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oro_newmark_destroy(self.h)
            self.h = None

    def set_external(self, f):
        lib().oro_newmark_set_external(self.h, _d(f64(f)))

    def set_impulse(self, f):
        lib().oro_newmark_set_impulse(self.h, _d(f64(f)))

    def set_state(self, d, v, a, steps=1):
        lib().oro_newmark_set_state(self.h, _d(f64(d)), _d(f64(v)), _d(f64(a)), steps)

    def step(self, n=1):
        it = np.zeros(n, np.int32); rr = np.zeros(n)
        init_it = lib().oro_newmark_step(self.h, n, _i(it), _d(rr))
        return it, rr, init_it

    def state(self):
        n = self.sys.ndof
        d, v, a = np.zeros(n), np.zeros(n), np.zeros(n)
        lib().oro_newmark_get_state(self.h, _d(d), _d(v), _d(a))
        return d, v, a

    def mass(self):
        M = np.zeros(self.sys.n_values)
        lib().oro_newmark_mass(self.h, _d(M))
        return M


class BackwardEuler:
    def __init__(self, system: System, kind, E, nu, rho, dt, alpha_m=0.0, alpha_k=0.0, pre=PRE_JACOBI, p1=0.0, p2=0.0,
                 tol=1e-6, max_iter=10000, thr_corot=0.0, thr_neo=0.1):
        self.sys = system
        n = system.n_tets
        self.kind = i32(np.broadcast_to(kind, (n,))); self.E = f64(np.broadcast_to(E, (n,)))
        self.nu = f64(np.broadcast_to(nu, (n,))); self.rho = f64(np.broadcast_to(rho, (n,)))
        self.h = lib().oro_be_create(system.h, _i(self.kind), _d(self.E), _d(self.nu), thr_corot, thr_neo, _d(self.rho),
                                     dt, alpha_m, alpha_k, pre, p1, p2, tol, max_iter)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oro_be_destroy(self.h)
            self.h = None

    def set_external(self, f):
        lib().oro_be_set_external(self.h, _d(f64(f)))

    def step(self):
        return lib().oro_be_step(self.h)

    def update_mesh(self):
        lib().oro_be_update_mesh(self.h)

    def velocity(self):
        v = np.zeros(self.sys.ndof)
        lib().oro_be_get_velocity(self.h, _d(v))
        return v

    def set_velocity(self, v):
        lib().oro_be_set_velocity(self.h, _d(f64(v)))

    def add_plane_damping(self, planes, tets, coeff):
        """addExternalDampingForceMatrix of a MeshPlaneCollisionDampingForcer per (plane, tet)
        pair, reading this integrator's velocities and the mesh displacements."""
        planes = f64(np.asarray(planes, np.float64).reshape(-1, 4)); tets = i32(tets)
        lib().oro_be_add_plane_damping(self.h, len(tets), _d(planes), _i(tets), coeff)

    def clear_damping(self):
        lib().oro_be_clear_damping(self.h)

    def damping(self):
        D = np.zeros(self.sys.n_values)
        lib().oro_be_damping(self.h, _d(D))
        return D


def plane_tet_intersection(plane, pts):
    """(centroid, dir) of the part of tet pts (4x3) below plane (n, c) (intersection.cpp:20-117)."""
    c, d = np.zeros(3), np.zeros(3)
    lib().oro_plane_tet_intersection(_d(f64(plane)), _d(f64(np.asarray(pts).reshape(12))), _d(c), _d(d))
    return c, d


def barycentric(pts, q):
    out = np.zeros(4)
    lib().oro_barycentric(_d(f64(np.asarray(pts).reshape(12))), _d(f64(q)), _d(out))
    return out


def gps(n_nodes, elements):
    """MeshRelabeler::getNewLables (core/meshRelabler.cpp:468-498): (new label per node,
    elements relabelled and sorted by label sum)."""
    el = i32(np.array(elements, np.int32, copy=True))
    lab = np.zeros(n_nodes, np.int32)
    rc = lib().oro_gps(n_nodes, el.shape[0], el.shape[1], _i(el), _i(lab))
    if rc != 0:
        raise ValueError("mesh graph is not connected")
    return lab, el
