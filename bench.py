#!/usr/bin/env python3
"""Implicit Newmark step throughput at 1M DOF on one MI355X.

Workload (default): 200x40x40 Kuhn box, 5x1x1 m, 1,920,000 tets, 1,008,600
free DOF, corotational material with per-tet E log-uniform in [1e4, 1e6],
nu = 0.45, rho = 1000, x = 0 face fixed, gravity, Newmark beta = 1/4,
gamma = 1/2, dt = 1/60, IC(0)-PCG to 1e-6 relative residual, tiled 4-colour node
numbering.  A "step" is one full Newmark step on the device: element pass,
assembly, matrix combine, IC(0) factorisation, PCG to tolerance, corrector.

`value` is the plain FP64 path (the reference's arithmetic).  The precision
form (FP32 inner PCG with FP64 residual corrections to the same 1e-6 relative
residual) runs on the same mesh in the same process and is reported under
`mixed_form`, with its speed-up and its displacement difference to the FP64
path.

Prints one JSON line.  Extra fields: the SpMV roofline of each form from HIP
events on the library stream, the step roofline of each form, the benchmark's
other legs, and the CPU FP64 oracle timed on the full workload (single thread,
on a host thread while the GPU legs run).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Newmark steps/s + SpMV GB/s at 1M DOF, rel.res 1e-6; p50/p99 served step latency"


def spmv_kernel_f32(v32):
    """Name of the FP32 PCG product kernel for launch shape v32 = (R, lanes); lanes 0 is the
    slot-value form, R its slots per thread."""
    if v32[1] == -1:
        return f"k_spmv_grp<float, true, {v32[0]}>"
    return f"k_spmv_sv<{v32[0]}>" if v32[1] == 0 else f"k_spmv_dot<float, {v32[0]}, true, {v32[1]}>"


def spmv_kernel_f64(v64):
    return f"k_spmv_grp<double, false, {v64[0]}>" if v64[1] == -1 else f"k_spmv_dot<double, {v64[0]}, false, {v64[1]}>"


def shape_f32(v32):
    if v32[1] == -1:
        return {"form": "grouped", "slots_per_thread": v32[0]}
    if v32[1] == 0:
        return {"form": "slot values", "slots_per_thread": v32[0]}
    return {"chunks_per_workgroup": v32[0], "lanes_per_row": v32[1]}


def spmv_bytes(mesh, value_bytes=8, vec_bytes=8, fused=False):
    """Algorithmic bytes of one SpMV over the upper-symmetric canonical layout:
    values + block column indices + block row pointers + x read + y write; the PCG product
    (fused) also reads z and writes the new direction p = z + beta p_old."""
    b = value_bytes * mesh.n_values + 4 * mesh.n_blocks + 4 * (mesh.n_rows + 1) + 2 * vec_bytes * mesh.ndof
    return b + (2 * vec_bytes * mesh.ndof if fused else 0)


def step_bytes(mesh, precision, iters, outer=0.0):
    """Algorithmic HBM bytes of one Newmark step (per-unit figures times the mesh's counts): fixed work = element pass (130.7 B per tet: connectivity, material, node positions
    and state, K0 recomputed), the M-combine read (8 B per stored value), two right-hand-side
    SpMVs and the IC(0) factorisation (read A, write the factor); per PCG iteration = one SpMV,
    one IC(0) apply (two sweeps over a factor with A's pattern: twice the SpMV's bytes) and
    eight vector passes.  The mixed form's iterations use FP32 values and vectors, and each
    refinement sweep adds an FP64 residual product and four FP64 vector passes."""
    n = mesh.ndof
    spmv64 = spmv_bytes(mesh, 8, 8)
    fixed = 130.7 * mesh.n_tets + 8 * mesh.n_values + 2 * spmv64 + 16 * mesh.n_values
    if precision == "fp64":
        return fixed + iters * (3 * spmv64 + 8 * 8 * n)
    spmv32 = spmv_bytes(mesh, 4, 4)
    return fixed + iters * (3 * spmv32 + 8 * 4 * n) + outer * (spmv64 + 4 * 8 * n)


def traffic_from_profile(workload, args, precision=None):
    """HBM bytes per SpMV launch measured with FETCH_SIZE/WRITE_SIZE counters (scaled by a
    copy kernel of known size), read from the committed counter summary of this workload
    and configuration; None when there is none."""
    path = os.path.join(ROOT, "profiles", f"pmc_spmv_{workload}_{args.precond}_{args.ordering}_"
                                          f"{precision or args.precision}.md")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("traffic_bytes_per_launch:"):
            return int(float(ln.split(":", 1)[1]))
    return None


def copy_calibration(mesh):
    """Three launches of the library's streaming matrix update y = y + 0.5 x over two
    matrices of this mesh (k_axpby): 16 bytes read and 8 written per stored value, the
    known-size reference that scales the HBM byte counters."""
    import opendeformer_amd as od
    a = od.Matrix(mesh); b = od.Matrix(mesh)
    a.set_zeros(); b.set_zeros()
    for _ in range(3):
        a.scale_add(1.0, 0.5, b)
    mesh.ctx.sync()
    print(f"copy calibration: 3 x k_axpby over {mesh.n_values} values: {16 * mesh.n_values} bytes read, "
          f"{8 * mesh.n_values} bytes written per launch", flush=True)


def mid_neo_leg(args, steps=200):
    """The mid-size precision-form workload: 80x16x16 box, 2.0x0.4x0.4 m, 122,880 tets,
    compressible neo-Hookean E = 5e4, nu = 0.40, x = 0 face fixed, gravity plus a seeded tip
    impulse of U(0, 50) N at step 0, dt = 1/120, 200 steps, IC(0)-PCG to 1e-6 in FP64 and in
    the mixed form from the same start; the mixed form's per-step displacement error against
    the FP64 path is reported."""
    import opendeformer_amd as od
    from opendeformer_amd import capi
    from opendeformer_amd import workloads as W
    w = W.box_mid_neo("color4t", seed=2)
    ctx = od.Context(0)
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    f = mesh.body_force(w.body)
    out, traj = {}, {}
    for name, prec in (("fp64", capi.OD_FP64), ("mixed", capi.OD_MIXED)):
        nm = od.NewmarkIntegrator(mesh, w.dt, precond=capi.OD_PRE_IC0, precision=prec, tol=1e-6)
        nm.set_external_virtual_work(f)
        nm.set_impulse(w.impulse)
        ds, its = [], []
        ctx.sync()
        t0 = time.perf_counter()
        for k in range(steps):
            nm.run_one_time_step()
            its.append(int(nm.stats()[0][0]))
            if k % 20 == 19:
                ds.append(nm.state()[0])
        ctx.sync()
        dt = time.perf_counter() - t0
        out[name] = {"steps_per_s": round(steps / dt, 2), "pcg_iters_per_step": round(float(np.mean(its)), 2)}
        traj[name] = ds
        del nm
    errs = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(traj["mixed"], traj["fp64"])]
    return {"workload": f"{w.name}: {mesh.n_tets} tets, {mesh.ndof} DOF, tip impulse, dt=1/120, {steps} steps "
                        f"(tiled 4-colour ordering; the timing includes a state read-back every 20 steps)",
            "fp64": out["fp64"], "mixed": out["mixed"],
            "mixed_rel_err_vs_fp64_max_over_sampled_steps": max(errs)}


def batched_leg(args, fp64_dof_steps_per_s=None, n_inst=256, warmup=2, steps=120):
    """256 uneven boxes (tets log-uniform [500, 5e4], aspect U(1, 6), corotational or
    neo-Hookean, E log-uniform [1e4, 1e6], nu U(0.3, 0.48)) concatenated into one system and
    stepped together: mixed-precision IC(0)-PCG to 1e-6 with per-instance convergence, PCG
    iterations replayed from a captured HIP graph, SpMV launch shape from the autotuner."""
    import opendeformer_amd as od
    from opendeformer_amd import capi
    from opendeformer_amd import workloads as W
    w = W.batch(n_inst, seed=4, ordering="color4t")
    ctx = od.Context(0)
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs, w.inst_node_start, w.inst_tet_start)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    _, v64, v32 = od.mesh_autotune(mesh)
    fb = mesh.body_force(w.body)

    def timed(whole_step, n_steps):
        nm = od.NewmarkIntegrator(mesh, w.dt, precond=capi.OD_PRE_IC0, precision=capi.OD_MIXED, tol=1e-6)
        nm.set_graph(True, whole_step=whole_step)
        nm.set_external_virtual_work(fb)
        nm.run_time_steps(warmup)
        ctx.sync()
        its = []
        t0 = time.perf_counter()
        for _ in range(n_steps):
            nm.run_one_time_step()
            its.append(nm.stats()[0].copy())
        ctx.sync()
        dt = (time.perf_counter() - t0) / n_steps
        return dt, np.array(its), nm.graph_info()

    # PCG iteration graphs (the faster mode here), then whole-step graphs on the same steps
    dt, its, _ = timed(False, steps)
    dt_step, its_step, ginfo = timed(True, steps)
    assert np.array_equal(its, its_step), "whole-step graphs must give the iteration counts of the eager path"
    # the plain FP64 drop-in path on the same instances, one at a time (every 16th instance,
    # 10 steps each after 2 warm-up steps): instance-steps/s of stepping them alone
    alone_t, alone_n = 0.0, 0
    for i in range(0, n_inst, 16):
        wi = W.subset(w, i)
        mi = od.Mesh(ctx, wi.X, wi.tets, wi.fixed_dofs)
        mi.set_material(wi.kind, wi.E, wi.nu, wi.rho)
        od.mesh_spmv_variant(mi, v64, v32)
        ni = od.NewmarkIntegrator(mi, wi.dt, precond=capi.OD_PRE_IC0, precision=capi.OD_FP64, tol=1e-6)
        ni.set_external_virtual_work(mi.body_force(wi.body))
        ni.run_time_steps(2)
        ctx.sync()
        t0 = time.perf_counter()
        ni.run_time_steps(10)
        ctx.sync()
        alone_t += time.perf_counter() - t0
        alone_n += 10
        del ni, mi
    alone_rate = alone_n / alone_t
    out = {"workload": f"{n_inst} boxes, {mesh.n_tets} tets, {mesh.ndof} DOF in one concatenated system "
                       f"(tiled 4-colour ordering per box), mixed precision, graph replay",
           "ms_per_step": round(dt * 1e3, 3), "batch_steps_per_s": round(1.0 / dt, 3),
           "instance_steps_per_s": round(n_inst / dt, 1), "dof_steps_per_s": round(mesh.ndof / dt, 1),
           "steps": steps, "warmup": warmup,
           "pcg_iters_per_instance_step": {"median": float(np.median(its)), "max": int(its.max()),
                                           "min": int(its.min())},
           "plain_fp64_alone_instance_steps_per_s": round(alone_rate, 1),
           "plain_fp64_alone_sample": f"instances 0, 16, ..., {n_inst - 16}: 10 steps each, one at a time",
           "speedup_vs_plain_fp64_alone": round((n_inst / dt) / alone_rate, 2),
           "spmv_shape": {"fp64": {"chunks_per_workgroup": v64[0], "lanes_per_row": v64[1]},
                          "fp32": shape_f32(v32)},
           "whole_step_graph": {"ms_per_step": round(dt_step * 1e3, 3), "dof_steps_per_s": round(mesh.ndof / dt_step, 1),
                                "vs_iteration_graphs": round(dt / dt_step, 3), **ginfo,
                                "note": "each step one captured graph (element, gather, factor, the plan's PCG "
                                        "iterations, corrector), plan = previous count + 4 rounded up to 4; spare "
                                        "iterations exit at once but still launch over the whole batch"}}
    if fp64_dof_steps_per_s:
        # the like-for-like comparison: DOF advanced per second against the 1M-DOF box on the plain
        # FP64 path (large single meshes are not launch-bound, small lone instances are)
        out["speedup_vs_fp64_path_dof_steps"] = round(out["dof_steps_per_s"] / fp64_dof_steps_per_s, 3)
    return out


def serving_leg(args, n_obj=64, tet_lo=1000, tet_hi=50000, rates=(50.0, 200.0), duration=4.0):
    """Served stream on a pool scaled down from 1,024 boxes of [1k, 500k] tets (about 82M
    tets, minutes of host-side mesh building) to n_obj boxes of [tet_lo, tet_hi] tets, with the
    200 GB HBM budget scaled by the same tet ratio; Poisson arrivals for `duration` s per rate."""
    from opendeformer_amd import serving as SV
    from opendeformer_amd import workloads as W
    pool = W.served_pool(n_obj, seed=5, tet_lo=tet_lo, tet_hi=tet_hi)
    pool_tets = sum(w.n_tets for w in pool)
    budget = int(200 * 2 ** 30 * pool_tets / 82.0e6)
    srv = SV.build_server(pool, budget, slots=4)
    legs = [SV.run_stream(srv, pool, r, duration, seed=6 + i) for i, r in enumerate(rates)]
    return {"pool": {"objects": n_obj, "tets_range": [tet_lo, tet_hi], "tets": int(pool_tets),
                     "hbm_budget_bytes": budget, "slots": 4, "precision": "mixed", "tol": 1e-6},
            "streams": legs}


def damping_leg(ctx, mesh, w, calls=20):
    """Plane collision damping on the main mesh: a floor tilted by (0.01, 0.005) under the box
    cuts its bottom cells; every tet with a vertex below it is a (plane, tet) pair, and all
    nodes move down at 1 m/s.  Times od_plane_damping (pair upload, the four damping kernels,
    the synchronisation) into a fresh matrix, median of `calls` calls after one warm-up."""
    import opendeformer_amd as od
    from opendeformer_amd import workloads as W
    n = np.array([0.01, 0.005, 1.0]); n /= np.linalg.norm(n)
    planes, tets = W.plane_contacts(w.X, w.tets, n, -0.06)
    vel = od.DeviceArray.from_host(ctx, np.tile([0.0, 0.0, -1.0], mesh.ndof // 3))
    D = od.Matrix(mesh)
    D.add_plane_damping(planes, tets, 1e5, None, vel.ptr)
    ctx.sync()
    ts = []
    for _ in range(calls):
        t0 = time.perf_counter()
        D.add_plane_damping(planes, tets, 1e5, None, vel.ptr)
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    return {"workload": f"floor under {w.name}: {len(tets)} (plane, tet) pairs, nodes moving into it",
            "pairs": int(len(tets)), "ms_per_call_median": round(med * 1e3, 3),
            "pairs_per_s": round(len(tets) / med, 1), "calls": calls,
            "note": "per call: upload of 36 B per pair, k_damp_eval, k_damp_offsets, k_damp_fill, "
                    "k_damp_apply, stream synchronisation"}


def log(msg):
    """Progress on stderr (stdout carries only the result line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def ordering_leg(args, scale=0.5, max_iter=2000,
                 plan=(("natural", 1, 2), ("color4", 1, 2), ("color4t", 1, 2), ("color5", 1, 2), ("gps", 0, 1))):
    """The box workload at half size (100x20x20 cells, 132,300 DOF) under three node numberings:
    natural (x-slowest grid order), the 4- and 5-colourings, and the reference's GPS relabelling
    (od_relabel_gps, host code).  FP64 IC(0)-PCG to 1e-6 with at most `max_iter` iterations;
    IC(0) dependency levels, PCG iterations and step time of each.  The GPS numbering gives
    about 1,400 dependency levels here and hundreds of iterations per step, so it is timed over
    one step (after the initial-acceleration solve)."""
    import opendeformer_amd as od
    from opendeformer_amd import capi
    from opendeformer_amd import workloads as W
    out = {}
    for ordering, warmup, steps in plan:
        t0 = time.perf_counter()
        w = W.box_1m_dof(ordering, scale=scale)
        t_host = time.perf_counter() - t0
        ctx = od.Context(0)
        mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs)
        mesh.set_material(w.kind, w.E, w.nu, w.rho)
        od.mesh_autotune(mesh)
        nm = od.NewmarkIntegrator(mesh, w.dt, precond=capi.OD_PRE_IC0, precision=capi.OD_FP64, tol=1e-6,
                                  max_iter=max_iter)
        nm.set_external_virtual_work(mesh.body_force(w.body))
        log(f"ordering {ordering}: {mesh.n_levels} IC(0) levels, stepping")
        for k in range(warmup):
            nm.run_one_time_step()
            log(f"ordering {ordering}: warm-up step {k} done, {int(nm.stats()[0][0])} iterations")
        ctx.sync()
        its, rrs = [], []
        t0 = time.perf_counter()
        for k in range(steps):
            nm.run_one_time_step()
            it, rr, _, _ = nm.stats()
            its.append(int(it[0])); rrs.append(float(rr[0]))
            log(f"ordering {ordering}: step {k} done, {its[-1]} iterations")
        ctx.sync()
        dt = (time.perf_counter() - t0) / steps
        out[ordering] = {"ic0_levels": mesh.n_levels, "pcg_iters_per_step": round(float(np.mean(its)), 2),
                         "relres_max": max(rrs), "ms_per_step": round(dt * 1e3, 2), "steps_per_s": round(1.0 / dt, 3),
                         "warmup": warmup, "steps": steps, "mesh_build_and_relabel_s": round(t_host, 2)}
        del nm, mesh, ctx
    return {"workload": f"box {int(200 * scale)}x{int(40 * scale)}x{int(40 * scale)} (the 1M-DOF box at half size), "
                        f"FP64 IC(0)-PCG to 1e-6, at most {max_iter} iterations", **out}


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


class CpuBaseline(threading.Thread):
    """The FP64 oracle (single thread, as the reference) on the benchmark's own workload: same
    mesh, seed, numbering and preconditioner; one warm-up step (it includes the initial
    acceleration solve), then `steps` timed steps.  Runs on a host thread (the oracle's calls
    release the GIL) while the GPU legs run."""

    def __init__(self, w, args, steps=5):
        super().__init__(daemon=True)
        self.w, self.args, self.steps = w, args, steps
        self.out, self.err = None, None

    def run(self):
        try:
            import oracle as O
            O.lib()
            O.set_threads(1)
            w = self.w
            t0 = time.perf_counter()
            s = O.System(w.X, w.tets, w.fixed_dofs)
            pre = {"jacobi": O.PRE_JACOBI, "ic0": O.PRE_IC0}[self.args.precond]
            nm = O.Newmark(s, w.kind, w.E, w.nu, w.rho, w.dt, pre=pre, p1=1e-3)
            nm.set_external(s.body_force(w.body))
            t_setup = time.perf_counter() - t0
            t0 = time.perf_counter()
            nm.step(1)
            t_warm = time.perf_counter() - t0
            ts, its = [], []
            for _ in range(self.steps):
                t0 = time.perf_counter()
                it, _, _ = nm.step(1)
                ts.append(time.perf_counter() - t0)
                its.append(int(it[0]))
            med = float(np.median(ts))
            self.out = {"value": 1.0 / med, "unit": "steps/s", "cores": 1, "kind": "port",
                        "sample": (f"FP64 oracle (the reference's CPU algorithm restated), 1 thread, on the full "
                                   f"workload ({w.name}, {s.ndof} DOF, {self.args.ordering} numbering, "
                                   f"{self.args.precond}-PCG to 1e-6): set-up {t_setup:.1f} s, 1 warm-up step "
                                   f"{t_warm:.1f} s (with the initial-acceleration solve), then {self.steps} timed "
                                   f"steps, median {med:.2f} s/step (min {min(ts):.2f}, max {max(ts):.2f}), "
                                   f"{np.mean(its):.1f} PCG iterations/step; host CPU {cpu_model()}, "
                                   f"{os.cpu_count()} logical CPUs visible; run on a host thread beside the GPU legs")}
        except BaseException as e:  # reported in the line, the GPU numbers stay valid
            self.err = e

    def result(self):
        self.join()
        if self.err is not None:
            return {"value": None, "error": str(self.err)}
        return self.out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precond", choices=["jacobi", "ic0"], default="ic0")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="the form whose steps/s is `value` (default: the plain FP64 path); the other form "
                         "runs beside it unless --no-second-form")
    ap.add_argument("--ordering", choices=["natural", "color8", "color5", "color4", "color4t", "gps"],
                    default="color4t",
                    help="node numbering: natural (x-slowest grid order), grouped by an 8-, 5- or 4-colouring "
                         "(4 colours: 4 IC(0) dependency levels; color4t: each colour's chunks are 4x8x8 tiles, "
                         "the same tile in every colour, for the grouped SpMV), or the GPS relabeller's")
    ap.add_argument("--scale", type=float, default=1.0, help="grid scale (1.0 = 200x40x40)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-second-form", action="store_true",
                    help="run only the --precision form (no speed-up or error report of the other one)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-instances leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--no-damping", action="store_true", help="skip the plane collision damping leg")
    ap.add_argument("--no-orderings", action="store_true",
                    help="skip the natural and GPS node-numbering comparison leg")
    ap.add_argument("--copy-calib", action="store_true",
                    help="after the run, launch a streaming update of known size three times (the reference "
                         "that scales HBM byte counters)")
    ap.add_argument("--fp32-spmv", choices=["auto", "slots1", "slots2", "slots4", "group4", "group8"], default="auto",
                    help="FP32 product shape: the autotuner's choice, the slot-value form with 1, 2 or 4 slots "
                         "per thread, or the grouped form with 4 or 8 (for counter runs, whose profiling skews "
                         "the autotuner's timings)")
    ap.add_argument("--max-iter", type=int, default=10000,
                    help="PCG iteration cap per solve (the reference's maxIteration; the mixed form's inner "
                         "iterations count against it together)")
    ap.add_argument("--fp64-spmv", default="auto",
                    help="FP64 product shape 'R,L' (chunks per workgroup, lanes per row), 'group4' or "
                         "'group8' (the grouped form, slots per thread), or auto (the autotuner's choice); "
                         "counter runs pin the shape the timed runs chose")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(the sharded multi-GPU path is designed, not implemented)", file=sys.stderr)
        sys.exit(2)

    import opendeformer_amd as od
    from opendeformer_amd import capi
    from opendeformer_amd import workloads as W

    w = W.box_1m_dof(args.ordering, scale=args.scale)
    cpu = None
    if not args.no_cpu:
        cpu = CpuBaseline(w, args)
        cpu.start()
    ctx = od.Context(0)
    mesh = od.Mesh(ctx, w.X, w.tets, w.fixed_dofs)
    mesh.set_material(w.kind, w.E, w.nu, w.rho)
    f = mesh.body_force(w.body)
    pre = {"jacobi": capi.OD_PRE_JACOBI, "ic0": capi.OD_PRE_IC0}[args.precond]
    tune_table, v64, v32 = od.mesh_autotune(mesh)
    if args.fp64_spmv != "auto":
        v64 = ((int(args.fp64_spmv[len("group"):]), -1) if args.fp64_spmv.startswith("group")
               else tuple(int(x) for x in args.fp64_spmv.split(",")))
    if args.fp32_spmv != "auto":
        v32 = ((int(args.fp32_spmv[len("group"):]), -1) if args.fp32_spmv.startswith("group")
               else (int(args.fp32_spmv[len("slots"):]), 0))
    if args.fp64_spmv != "auto" or args.fp32_spmv != "auto":
        od.mesh_spmv_variant(mesh, v64, v32)      # the sweep forms stay as tuned

    def run_form(precision):
        """Warm-up, then the timed steps; returns the measurements and the final displacement."""
        prec = {"fp64": capi.OD_FP64, "mixed": capi.OD_MIXED}[precision]
        nm = od.NewmarkIntegrator(mesh, w.dt, beta=0.25, gamma=0.5, precond=pre, precision=prec, tol=1e-6,
                                  max_iter=args.max_iter)
        nm.set_graph(True)
        nm.set_external_virtual_work(f)
        nm.run_time_steps(args.warmup)

        def timed_steps():
            iters, relres, outer = [], [], []
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                nm.run_one_time_step()
                it, rr, ou, _ = nm.stats()
                iters.append(int(it[0])); relres.append(float(rr[0])); outer.append(ou)
            ctx.sync()
            return time.perf_counter() - t0, iters, relres, outer

        # K steps as a user runs them (PCG iterations replayed from HIP graphs, no instrumentation)
        elapsed, iters, relres, outer = timed_steps()
        # the SpMV's launch durations: K more steps with a HIP event pair recorded from every
        # SpMV dispatch on the library's stream (eager launches: events cannot sit in a graph)
        ctx.profile(True)
        elapsed_ev, iters_ev, _, _ = timed_steps()
        spmv_ms, spmv_n = ctx.profile_read()
        ctx.profile(False)
        vb = 4 if precision == "mixed" else 8
        alg = spmv_bytes(mesh, vb, vb)                       # the SpMV's algorithmic bytes
        fused = spmv_bytes(mesh, vb, vb, fused=(vb == 4))    # what the FP32 product also moves
        # Launches of a batch that finds its instance converged exit at once; their (small) time
        # stays in the timed total, but only the iterations that ran count as SpMV work.
        real_spmv = int(np.sum(iters_ev))
        avg_spmv_s = (spmv_ms / max(real_spmv, 1)) / 1e3
        achieved = alg / avg_spmv_s / 1e9 if real_spmv else 0.0
        d = nm.state()[0]
        del nm
        t_step = elapsed / args.steps
        sb = step_bytes(mesh, precision, float(np.mean(iters)), float(np.mean(outer)))
        spmv = {"kernel": (spmv_kernel_f64(v64) if vb == 8 else spmv_kernel_f32(v32)),
                "launches": spmv_n, "iterations": real_spmv,
                "avg_us_per_launch": round(spmv_ms / max(spmv_n, 1) * 1e3, 2),
                "avg_us_per_iteration": round(avg_spmv_s * 1e6, 2),
                "algorithmic_MB": round(alg / 1e6, 2), "GBps": round(achieved, 1),
                "timed_over": (f"{args.steps} further steps with a HIP event pair on every SpMV dispatch "
                               f"(eager launches, {args.steps / elapsed_ev:.2f} steps/s); the headline "
                               f"steps are not instrumented")}
        if fused != alg:
            spmv["fused_MB"] = round(fused / 1e6, 2)
            spmv["fused_GBps"] = round(fused / avg_spmv_s / 1e9, 1) if real_spmv else 0.0
            spmv["fused_note"] = ("the FP32 product also reads z and writes the new direction p = z + beta p_old "
                                  "(the direction update fused into it)")
        return {"steps_s": args.steps / elapsed, "ms_step": t_step * 1e3,
                "iters": float(np.mean(iters)), "relres_max": float(np.max(relres)),
                "outer": float(np.mean(outer)), "spmv": spmv, "achieved": achieved,
                "step_roofline": {"bytes_per_step_GB": round(sb / 1e9, 3), "achieved_GBps": round(sb / t_step / 1e9, 1),
                                  "frac": round(sb / t_step / 1e9 / HBM_PEAK_GBS, 4)}}, d

    def roofline(r, precision):
        return {"bound": "hbm", "achieved": round(r["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(r["achieved"] / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profile(w.name, args, precision)}

    log(f"{w.name}: {mesh.n_tets} tets, {mesh.ndof} DOF; SpMV shapes {v64}/{v32}; timing the {args.precision} form")
    main_r, d_main = run_form(args.precision)
    log(f"{args.precision}: {main_r['steps_s']:.2f} steps/s")
    other = "mixed" if args.precision == "fp64" else "fp64"
    other_r = None
    if not args.no_second_form:
        other_r, d_other = run_form(other)
        log(f"{other}: {other_r['steps_s']:.2f} steps/s")
        d64, d32 = (d_main, d_other) if args.precision == "fp64" else (d_other, d_main)
        err = float(np.linalg.norm(d32 - d64) / np.linalg.norm(d64))
    line = {
        "metric": METRIC, "value": round(main_r["steps_s"], 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(main_r["ms_step"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-inner/f64-refinement",
        "data": "synthetic (seeded box mesh, seed 3)",
        "config": {"workload": f"{w.name}: {mesh.n_tets} tets, {mesh.ndof} DOF, Newmark beta=1/4 gamma=1/2 dt=1/60, "
                               f"PCG tol 1e-6 relative residual",
                   "precond": args.precond, "precision": args.precision, "ordering": args.ordering,
                   "ic0_levels": mesh.n_levels},
        "pcg_iters_per_step": round(main_r["iters"], 2),
        "relres_max": main_r["relres_max"],
        "converged": bool(main_r["relres_max"] <= 1e-6),
        "spmv": main_r["spmv"],
        "roofline": roofline(main_r, args.precision),
        "step_roofline": main_r["step_roofline"],
    }
    if args.precision == "mixed":
        line["refinement_sweeps_per_step"] = round(main_r["outer"], 2)
    if other_r is not None:
        key = "mixed_form" if other == "mixed" else "fp64_path"
        line[key] = {"value": round(other_r["steps_s"], 4), "ms_per_step": round(other_r["ms_step"], 3),
                     "dtype": "f32-inner/f64-refinement" if other == "mixed" else "f64",
                     "pcg_iters_per_step": round(other_r["iters"], 2), "relres_max": other_r["relres_max"],
                     "spmv": other_r["spmv"], "roofline": roofline(other_r, other),
                     "step_roofline": other_r["step_roofline"]}
        if other == "mixed":
            line[key]["refinement_sweeps_per_step"] = round(other_r["outer"], 2)
        mixed_r, fp64_r = (other_r, main_r) if other == "mixed" else (main_r, other_r)
        line["mixed_speedup_vs_fp64"] = round(mixed_r["steps_s"] / fp64_r["steps_s"], 3)
        line["mixed_rel_err_vs_fp64_path"] = {"value": err,
                                              "at": f"displacement after step {args.warmup + 2 * args.steps} "
                                                    f"of each form's run"}
    def tune_key(p, ls, r):
        return f"{p}_grouped_K{r}" if ls == -1 else (f"{p}_L{ls}_R{r}" if ls else f"{p}_slotvalues_K{r}")
    line["spmv_autotune"] = {"fp64": ({"form": "grouped", "slots_per_thread": v64[0]} if v64[1] == -1 else
                                      {"chunks_per_workgroup": v64[0], "lanes_per_row": v64[1]}),
                             "fp32": shape_f32(v32),
                             "cold_us": {tune_key(p, ls, r): round(v * 1e3, 2) for (p, ls, r), v in tune_table.items()}}
    info = mesh.info_raw()
    def forms(word):
        # per level: lanes per row of the lane-per-row kernel, or slots per thread of the slot kernel
        codes = [(int(word) >> (4 * l)) & 15 for l in range(min(8, mesh.n_levels))]
        return [f"slots{1 << (c - 8)}" if c >= 8 else f"lanes{1 << c}" for c in codes]
    line["ic0_sweeps"] = {"contiguous_levels": bool(info[12] & 1), "form_fwd": forms(info[13]),
                          "form_bwd": forms(info[14])}
    if not args.no_damping:
        line["plane_damping"] = damping_leg(ctx, mesh, w)
        log("plane damping leg done")
    if not args.no_orderings:
        line["orderings"] = ordering_leg(args)
    if not args.no_batch:
        line["mid_neo_precision"] = mid_neo_leg(args)
        log("mid-size neo-Hookean leg done")
        r64 = main_r if args.precision == "fp64" else other_r
        line["batched"] = batched_leg(args, mesh.ndof * r64["steps_s"] if r64 else None)
        log("batched leg done")
    if not args.no_serving:
        line["serving"] = serving_leg(args)
        log("serving leg done")
    if cpu is not None:
        log("waiting for the CPU baseline")
        line["cpu_baseline"] = cpu.result()
    print(json.dumps(line), flush=True)
    if args.copy_calib:
        copy_calibration(mesh)


if __name__ == "__main__":
    main()
