"""Python host mirror of the reference's interfaces over the C-ABI.

Class names follow the reference (OpenDeformer/): the Newmark and
backward-Euler integrators stand behind the `Intergrator` interface
(core/intergrator.h:13-35), `CGSolver` behind `LinearSolver`
(core/linearSolver.h:12-20, solver/cgSolver.cpp:13-57), the material call
behind `FullOrderNonlinearMaterial::generateMatrixAndVirtualWorks`
(material/fullOrderNonlinear.h:36-44), and `spmdv` is `SpMDV`
(core/numerMethod.h:134-262).  Compute happens in libodhip.so only.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check, dptr, f64, i32, lib


class Context:
    """One HIP device and stream."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        check(lib().od_ctx_create(device, C.byref(self.h)))

    def sync(self):
        check(lib().od_ctx_sync(self.h))

    def profile(self, enable: bool):
        """HIP-event timing of every PCG SpMV launch on this context's stream."""
        check(lib().od_ctx_profile(self.h, 1 if enable else 0))

    def profile_read(self):
        ms = C.c_double(0); n = C.c_longlong(0)
        check(lib().od_ctx_profile_read(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if self.h:
            lib().od_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device buffer owned by the context (float64/float32/int32)."""

    def __init__(self, ctx: Context, n: int, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.ptr = C.c_void_p()
        check(lib().od_dev_alloc(ctx.h, self.n * self.dtype.itemsize, C.byref(self.ptr)))

    @classmethod
    def from_host(cls, ctx, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(ctx, a.size, dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(lib().od_copy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def download(self):
        out = np.empty(self.n, self.dtype)
        check(lib().od_copy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def zero(self):
        check(lib().od_memset_d(self.ctx.h, self.ptr, 0, self.n * self.dtype.itemsize))

    def __del__(self):
        try:
            if self.ptr:
                lib().od_dev_free(self.ctx.h, self.ptr)
                self.ptr = C.c_void_p()
        except Exception:
            pass


class Mesh:
    """Device-resident mesh + NodeIndexer + canonical pattern (+ instances)."""

    def __init__(self, ctx: Context, X, tets, constrained_dofs, inst_node_start=None, inst_tet_start=None):
        self.ctx = ctx
        X = f64(X).reshape(-1, 3)
        tets = i32(tets).reshape(-1, 4)
        con = i32(constrained_dofs)
        self.X, self.tets = X, tets
        n_inst = 1 if inst_node_start is None else len(inst_node_start) - 1
        ins = i32(inst_node_start) if n_inst > 1 else None
        its = i32(inst_tet_start) if n_inst > 1 else None
        self.h = C.c_void_p()
        check(lib().od_mesh_create(ctx.h, X.shape[0], dptr(X), tets.shape[0], capi.iptr(tets), con.size,
                                   capi.iptr(con) if con.size else None, n_inst,
                                   capi.iptr(ins) if ins is not None else None,
                                   capi.iptr(its) if its is not None else None, C.byref(self.h)))
        info = np.zeros(16, np.int64)
        check(lib().od_mesh_info(self.h, info.ctypes.data_as(C.POINTER(C.c_longlong))))
        (self.ndof, self.n_rows, self.n_blocks, self.n_values, self.n_levels, self.n_inst, self.n_tets,
         self.n_nodes, self.n_chunks, self.n_pair_contrib, self.n_node_contrib, self.device_bytes) = (int(v) for v in info[:12])
        # some node partially constrained: the block rows are DOF triples (n_rows, n_blocks count
        # the kernels' padded blocks; ndof, n_values and pattern() are the reference's layout)
        self.partial = bool(info[15])

    def info_raw(self):
        """The 16 values of od_mesh_info (read now: the sweep lanes change when a solver tunes them)."""
        info = np.zeros(16, np.int64)
        check(lib().od_mesh_info(self.h, info.ctypes.data_as(C.POINTER(C.c_longlong))))
        return info

    def close(self):
        if self.h:
            lib().od_mesh_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pattern(self, with_map=True):
        ncols = self.ndof // 3 + self.ndof % 3 if self.partial else self.n_rows
        pcol = np.zeros(ncols + 1, np.int32); oris = np.zeros(ncols + 1, np.int32)
        if self.partial:
            check(lib().od_pattern_export(self.h, capi.iptr(pcol), None, None, None))
        rows = np.zeros(int(pcol[-1]) if self.partial else self.n_blocks, np.int32)
        emap = np.zeros(78 * self.n_tets, np.int32) if with_map else None
        check(lib().od_pattern_export(self.h, capi.iptr(pcol), capi.iptr(oris), capi.iptr(rows),
                                      capi.iptr(emap) if with_map else None))
        return pcol, oris, rows, (emap.reshape(-1, 78) if with_map else None)

    def rows(self):
        irs = np.zeros(self.n_inst + 1, np.int32); lv = np.zeros(self.n_rows, np.int32)
        check(lib().od_mesh_rows(self.h, capi.iptr(irs), capi.iptr(lv)))
        return irs, lv

    def set_material(self, kind, E, nu, rho, corot_threshold=0.0, neo_threshold=0.1):
        n = self.n_tets
        kind = i32(np.broadcast_to(kind, (n,))); E = f64(np.broadcast_to(E, (n,)))
        nu = f64(np.broadcast_to(nu, (n,))); rho = f64(np.broadcast_to(rho, (n,)))
        check(lib().od_mesh_set_material(self.h, capi.iptr(kind), dptr(E), dptr(nu), dptr(rho), corot_threshold,
                                         neo_threshold))

    def set_plastic(self, yield_criterion, creep, max_strain):
        """Corotational plastic parameters per tet (kind OD_MAT_COROT_PLASTIC); resets the state."""
        n = self.n_tets
        y = f64(np.broadcast_to(yield_criterion, (n,))); c = f64(np.broadcast_to(creep, (n,)))
        mx = f64(np.broadcast_to(max_strain, (n,)))
        check(lib().od_mesh_set_plastic(self.h, dptr(y), dptr(c), dptr(mx)))

    def plastic_strain(self):
        out = np.zeros(9 * self.n_tets)
        check(lib().od_mesh_plastic_strain(self.h, dptr(out), None))
        return out.reshape(-1, 3, 3)

    def body_force(self, body):
        body = f64(np.broadcast_to(np.asarray(body, np.float64).reshape(-1, 3), (self.n_tets, 3)))
        f = DeviceArray(self.ctx, self.ndof)
        check(lib().od_body_force(self.h, dptr(body), f.ptr))
        return f.download()


def mesh_autotune(mesh: "Mesh"):
    """Reruns the autotuner: ({(precision, lanes, R): ms}, (R, lanes) for FP64, (R, lanes) for FP32).
    Lanes 0 is the FP32 slot-value form, whose R is the slots per thread; lanes -1 the grouped
    form, whose R is the slots per thread (4 or 8), listed only where the mesh has it."""
    ms = np.zeros(25); ch = np.zeros(4, np.int32)
    check(lib().od_mesh_autotune(mesh.h, dptr(ms), capi.iptr(ch)))
    table = {(p, ls, r): float(ms[9 * i + 3 * k + j]) for i, p in enumerate(("fp64", "fp32"))
             for k, ls in enumerate((16, 8, 4)) for j, r in enumerate((1, 2, 4))}
    table.update({("fp32", 0, r): float(ms[18 + j]) for j, r in enumerate((1, 2, 4))})
    for i, p in enumerate(("fp64", "fp32")):
        for j, k in enumerate((4, 8)):
            if ms[21 + 2 * i + j] >= 0:
                table[(p, -1, k)] = float(ms[21 + 2 * i + j])
    return table, (int(ch[0]), int(ch[1])), (int(ch[2]), int(ch[3]))


def mesh_spmv_variant(mesh: "Mesh", v64, v32):
    """Sets the SpMV launch shapes: v64, v32 = (chunks per workgroup R, lanes per row); v32 may
    be (slots per thread, 0), the FP32 slot-value form; either may be (4 or 8, -1), the grouped
    form with that many slots per thread (where the mesh has it)."""
    check(lib().od_mesh_spmv_variant(mesh.h, int(v64[0]), int(v64[1]), int(v32[0]), int(v32[1])))


def mesh_sweep_form(mesh: "Mesh", lanes, slots=0):
    """Sets every IC(0) sweep level's form: lanes per row (4, 8, 16) of the lane-per-row
    kernels, or with slots (2, 4, 8, 16) the FP32 slot kernel."""
    check(lib().od_mesh_sweep_form(mesh.h, int(lanes), int(slots)))


class Matrix:
    """Values of a BlockedSymSparseMatrix<3,3> in the canonical layout, on the device."""

    def __init__(self, mesh: Mesh):
        self.mesh = mesh
        self.h = C.c_void_p()
        check(lib().od_matrix_create(mesh.h, C.byref(self.h)))
        p = C.c_void_p()
        check(lib().od_matrix_values(self.h, C.byref(p)))
        self.values_ptr = p

    def __del__(self):
        try:
            if self.h:
                lib().od_matrix_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass

    def upload(self, values):
        v = f64(values)
        assert v.size == self.mesh.n_values
        check(lib().od_matrix_upload(self.h, dptr(v)))

    def download(self):
        v = np.zeros(self.mesh.n_values)
        check(lib().od_matrix_download(self.h, dptr(v)))
        return v

    def add_plane_damping(self, planes, tets, coeff, disp_dev=None, vel_dev=None):
        """MeshPlaneCollisionDampingForcer::addDampingMatrix for each (plane, tet) pair, in order
        (collision/penaltyDampingForcer.cpp:8-61); planes: (k, 4) normal and constant."""
        planes = f64(np.asarray(planes, np.float64).reshape(-1, 4)); tets = np.ascontiguousarray(tets, np.int32)
        check(lib().od_plane_damping(self.h, int(tets.size), dptr(planes), capi.iptr(tets), float(coeff),
                                     disp_dev, vel_dev))

    def set_zeros(self):
        check(lib().od_matrix_zero(self.h))

    def scale_add(self, a, b=0.0, other=None):
        check(lib().od_matrix_axpby(self.h, a, b, other.h if other is not None else None))

    def add_mass(self):
        check(lib().od_mass_matrix(self.mesh.h, self.h))

    def sync_f32(self):
        check(lib().od_matrix_sync_f32(self.h))


def generate_matrix_and_virtual_works(mesh: Mesh, u_free, K: Matrix, fint: DeviceArray):
    """FullOrderNonlinearMaterial::generateMatrixAndVirtualWorks: K += K(u), fint += f(u)."""
    u = u_free if isinstance(u_free, DeviceArray) else DeviceArray.from_host(mesh.ctx, u_free)
    check(lib().od_assemble(mesh.h, u.ptr, K.h, fint.ptr if fint is not None else None))


def spmdv(A: Matrix, x: DeviceArray, y: DeviceArray, accumulate=True):
    """SpMDV: y += A x (accumulate) or y = A x."""
    check(lib().od_spmv(A.h, x.ptr, y.ptr, 1 if accumulate else 0))


def spmdv_f32(A: Matrix, x: DeviceArray, y: DeviceArray, accumulate=True):
    check(lib().od_spmv_f32(A.h, x.ptr, y.ptr, 1 if accumulate else 0))


class CGSolver:
    """PCG with a per-solve preconditioner rebuild (cgSolver.cpp:13-57)."""

    def __init__(self, mesh: Mesh, tol=1e-6, max_iter=10000, precond=capi.OD_PRE_JACOBI, precond_param=1e-3,
                 precision=capi.OD_FP64):
        self.mesh, self.tol, self.max_iter = mesh, tol, max_iter
        self.h = C.c_void_p()
        check(lib().od_solver_create(mesh.h, precond, precond_param, precision, C.byref(self.h)))

    def __del__(self):
        try:
            if self.h:
                lib().od_solver_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass

    def solve(self, A: Matrix, b: DeviceArray, x: DeviceArray):
        it = np.zeros(self.mesh.n_inst, np.int32); rr = np.zeros(self.mesh.n_inst)
        check(lib().od_pcg_solve(self.h, A.h, b.ptr, x.ptr, self.tol, self.max_iter, capi.iptr(it), dptr(rr)))
        return it, rr

    def precond_apply(self, A: Matrix, r: DeviceArray, z: DeviceArray):
        check(lib().od_precond_apply(self.h, A.h, r.ptr, z.ptr))

    def set_refinement(self, inner_tol, max_outer=12):
        check(lib().od_solver_set_refinement(self.h, inner_tol, max_outer))

    def set_graph(self, enable=True):
        check(lib().od_solver_set_graph(self.h, 1 if enable else 0))

    def set_precond_params(self, p1, p2):
        """InCholeskyPreconditioner(epsilon, relaxedScale) / InLDLTPreconditioner(sainvEpsilon,
        ldltEpsilon) parameters; IC(0): the pivot epsilon."""
        check(lib().od_solver_set_precond_params(self.h, float(p1), float(p2)))

    def info(self):
        """Graph captures, device bytes, refinement sweeps and largest iteration count of the
        last solve."""
        return _solver_info(self.h)

    def ic0_factor(self, A: Matrix):
        L = np.zeros(self.mesh.n_values)
        check(lib().od_ic0_factor(self.h, A.h, dptr(L)))
        return L


def _solver_info(h):
    info = np.zeros(4, np.int64)
    check(lib().od_solver_info(h, info.ctypes.data_as(C.POINTER(C.c_longlong))))
    return {"graph_captures": int(info[0]), "device_bytes": int(info[1]), "outer": int(info[2]),
            "iters_max": int(info[3])}


def pcg_solve_noalloc(A: Matrix, b: DeviceArray, x: DeviceArray, precond=capi.OD_PRE_JACOBI, tol=1e-6,
                      max_iter=10000):
    """od_pcg_solve_noalloc: the FP64 PCG solve (Jacobi or no preconditioner) with every
    temporary in the library's fixed device scratch; x is the warm start and the result."""
    n = A.mesh.n_inst
    it = np.zeros(n, np.int32); rr = np.zeros(n)
    check(lib().od_pcg_solve_noalloc(A.h, int(precond), b.ptr, x.ptr, tol, int(max_iter), capi.iptr(it), dptr(rr)))
    return it, rr


class _Integrator:
    def __del__(self):
        try:
            if self.h:
                lib().od_integrator_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass

    def set_external_virtual_work(self, f):
        check(lib().od_integrator_set_external(self.h, dptr(f64(f))))

    def set_impulse(self, f):
        check(lib().od_integrator_set_impulse(self.h, dptr(f64(f))))

    def set_active(self, mask):
        if mask is None:
            check(lib().od_integrator_set_active(self.h, None))
        else:
            m = np.ascontiguousarray(mask, dtype=np.uint8)
            check(lib().od_integrator_set_active(self.h, m.ctypes.data_as(C.POINTER(C.c_ubyte))))

    def run_one_time_step(self):
        check(lib().od_integrator_step(self.h, 1))

    def run_time_steps(self, n):
        check(lib().od_integrator_step(self.h, int(n)))

    def state(self):
        n = self.mesh.ndof
        d, v, a = np.zeros(n), np.zeros(n), np.zeros(n)
        check(lib().od_integrator_get_state(self.h, dptr(d), dptr(v), dptr(a)))
        return d, v, a

    def set_state(self, d=None, v=None, a=None):
        args = [dptr(f64(x)) if x is not None else None for x in (d, v, a)]
        check(lib().od_integrator_set_state(self.h, *args))

    def device_bytes(self):
        """Device bytes of the mesh, the integrator and its solver."""
        b = C.c_longlong(0)
        check(lib().od_integrator_device_bytes(self.h, C.byref(b)))
        return b.value

    def set_refinement(self, inner_tol, max_outer=12):
        """Mixed precision: FP32 inner tolerance and the cap on FP64 corrections."""
        check(lib().od_integrator_set_refinement(self.h, inner_tol, max_outer))

    def set_graph(self, enable=True, whole_step=False):
        """Graph mode: PCG iterations replayed from captured HIP graphs; with whole_step, Newmark
        steps after the first run as one captured graph each."""
        check(lib().od_integrator_set_graph(self.h, (2 if whole_step else 1) if enable else 0))

    def graph_info(self):
        """Step graphs captured, steps run as a graph, graph steps completed eagerly, cached graphs."""
        info = np.zeros(4, np.int64)
        check(lib().od_integrator_graph_info(self.h, info.ctypes.data_as(C.POINTER(C.c_longlong))))
        return {"captures": int(info[0]), "graph_steps": int(info[1]), "fallbacks": int(info[2]),
                "cached": int(info[3])}

    def set_precond_params(self, p1, p2):
        """od_solver_set_precond_params on the integrator's solver."""
        sh = C.c_void_p()
        check(lib().od_integrator_solver(self.h, C.byref(sh)))
        check(lib().od_solver_set_precond_params(sh, float(p1), float(p2)))

    def solver_info(self):
        """od_solver_info of the integrator's solver (graph captures among others)."""
        s = C.c_void_p()
        check(lib().od_integrator_solver(self.h, C.byref(s)))
        return _solver_info(s)

    def stats(self):
        it = np.zeros(self.mesh.n_inst, np.int32); rr = np.zeros(self.mesh.n_inst)
        outer = C.c_int(0); init = C.c_int(0)
        check(lib().od_integrator_stats(self.h, capi.iptr(it), dptr(rr), C.byref(outer), C.byref(init)))
        return it, rr, outer.value, init.value

    def matrix_handle(self):
        A = C.c_void_p()
        check(lib().od_integrator_matrix(self.h, C.byref(A)))
        return A

    def vertex_positions(self):
        pos = np.zeros(3 * self.mesh.n_nodes, np.float32)
        check(lib().od_integrator_positions(self.h, pos.ctypes.data_as(C.POINTER(C.c_float))))
        return pos.reshape(-1, 3)


class NewmarkIntegrator(_Integrator):
    """Full-order linearised Newmark (linearNewmark.cpp:27-30, 57-79 at full order)."""

    def __init__(self, mesh: Mesh, dt, beta=0.25, gamma=0.5, alpha_m=0.0, alpha_k=0.0, precond=capi.OD_PRE_IC0,
                 precond_param=1e-3, tol=1e-6, max_iter=10000, precision=capi.OD_FP64, init_accel=True):
        self.mesh = mesh
        self.h = C.c_void_p()
        check(lib().od_newmark_create(mesh.h, beta, gamma, dt, alpha_m, alpha_k, precond, precond_param, tol, max_iter,
                                      precision, 1 if init_accel else 0, C.byref(self.h)))


class BackwardEulerIntegrator(_Integrator):
    """NonlinearImplicitBackwardEuler (intergrator/nonlinearImplicitBackwardEuler.h:48-146)."""

    def __init__(self, mesh: Mesh, dt, alpha_m=0.0, alpha_k=0.0, precond=capi.OD_PRE_JACOBI, precond_param=1e-3,
                 tol=1e-6, max_iter=10000):
        self.mesh = mesh
        self.h = C.c_void_p()
        check(lib().od_backward_euler_create(mesh.h, dt, alpha_m, alpha_k, precond, precond_param, tol, max_iter,
                                             C.byref(self.h)))

    def update_mesh_vertices_displacements(self):
        check(lib().od_backward_euler_update_mesh(self.h))

    def add_external_damping_force_matrix(self, planes, tets, coeff):
        """addExternalDampingForceMatrix (nonlinearImplicitBackwardEuler.h:122-124) with a
        MeshPlaneCollisionDampingForcer per (plane, tet) pair reading this integrator's
        displacement and velocity."""
        planes = f64(np.asarray(planes, np.float64).reshape(-1, 4)); tets = np.ascontiguousarray(tets, np.int32)
        check(lib().od_integrator_add_plane_damping(self.h, int(tets.size), dptr(planes), capi.iptr(tets),
                                                    float(coeff)))

    def clear_external_damping_force_matrix(self):
        check(lib().od_integrator_clear_damping(self.h))

    def damping_matrix(self):
        D = np.zeros(self.mesh.n_values)
        check(lib().od_integrator_damping(self.h, dptr(D)))
        return D


class Server:
    """Serving loop (od_server_*): a resident pool of objects, HBM-budget admission and
    nSlots overlapping worker slots; requests advance one object by k Newmark steps."""

    def __init__(self, budget_bytes, slots=4, precision=capi.OD_MIXED, precond=capi.OD_PRE_IC0, tol=1e-6, device=0):
        self.h = C.c_void_p()
        check(lib().od_server_create(device, int(budget_bytes), slots, precision, precond, tol, C.byref(self.h)))
        self.ndof = []

    def __del__(self):
        try:
            if self.h:
                lib().od_server_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass

    def add_object(self, X, tets, constrained_dofs, kind, E, nu, rho, body_per_tet, dt):
        X = f64(X).reshape(-1, 3); tets = i32(tets).reshape(-1, 4); con = i32(constrained_dofs)
        n = tets.shape[0]
        kind = i32(np.broadcast_to(kind, (n,))); E = f64(np.broadcast_to(E, (n,)))
        nu = f64(np.broadcast_to(nu, (n,))); rho = f64(np.broadcast_to(rho, (n,)))
        body = f64(np.broadcast_to(body_per_tet, (n, 3)))
        oid = C.c_int()
        check(lib().od_server_add_object(self.h, X.shape[0], dptr(X), n, capi.iptr(tets), con.size,
                                         capi.iptr(con) if con.size else None, capi.iptr(kind), dptr(E), dptr(nu),
                                         dptr(rho), dptr(body), float(dt), C.byref(oid)))
        self.ndof.append(3 * X.shape[0] - con.size)
        return oid.value

    def run(self, obj, n_steps, arrival, impulses=None):
        """impulses: per request a list of (node, (fx, fy, fz)) or None.  Returns dict of
        start/finish seconds, PCG iterations per request and run stats."""
        obj = i32(obj); n_steps = i32(n_steps); arrival = f64(arrival)
        k = obj.size
        start = np.zeros(k); finish = np.zeros(k); iters = np.zeros(k, np.int32); stats = np.zeros(4, np.int64)
        if impulses is not None:
            starts = [0]; nodes = []; forces = []
            for imp in impulses:
                for node, fv in (imp or []):
                    nodes.append(int(node)); forces.extend(float(x) for x in fv)
                starts.append(len(nodes))
            starts = i32(starts); nodes = i32(nodes if nodes else [0]); forces = f64(forces if forces else [0.0] * 3)
            imp_args = (capi.iptr(starts), capi.iptr(nodes), dptr(forces))
        else:
            imp_args = (None, None, None)
        check(lib().od_server_run(self.h, k, capi.iptr(obj), capi.iptr(n_steps), dptr(arrival), *imp_args,
                                  dptr(start), dptr(finish), capi.iptr(iters),
                                  stats.ctypes.data_as(C.POINTER(C.c_longlong))))
        return {"start": start, "finish": finish, "iters": iters,
                "stats": {"peak_resident_bytes": int(stats[0]), "evictions": int(stats[1]), "loads": int(stats[2]),
                          "resident_bytes": int(stats[3])}}

    def object_state(self, oid):
        n = self.ndof[oid]
        d, v, a = np.zeros(n), np.zeros(n), np.zeros(n)
        check(lib().od_server_object_state(self.h, oid, dptr(d), dptr(v), dptr(a)))
        return d, v, a

    def object_info(self, oid):
        info = np.zeros(5, np.int64)
        check(lib().od_server_object_info(self.h, oid, info.ctypes.data_as(C.POINTER(C.c_longlong))))
        return {"ndof": int(info[0]), "resident": bool(info[1]), "bytes": int(info[2]), "steps": int(info[3]),
                "tets": int(info[4])}
