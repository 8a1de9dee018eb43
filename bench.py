#!/usr/bin/env python3
"""Tet10 elasticity DOF/s (assemble + PCG to 1e-10), mixed vs FP64, one GPU.

Workload (default "tet10-160x20x20"): cantilever box 0.4 x 0.05 x 0.05 m,
160 x 20 x 20 Kuhn cells -> 384,000 Tet10 elements, 539,601 nodes,
1,618,803 DOF; E = 210 GPa, nu = 0.3, clamped at x = 0, uniform traction
1 MPa in -z on x = 0.4.  Synthetic mesh, generated on the host before the
timed region and resident in HBM when it starts.

One step = the whole drop-in path on device-resident inputs: element
geometry, node-pair CSR pattern (radix sort), stiffness values, traction
load, Dirichlet-masked solve to ||b_f - A_ff x_f|| / ||b_f|| <= 1e-10 on the
true FP64 residual (solvers restart from the true residual when their
recursion drifts; where 1e-10 lies below the FP64 floor eps ||A_ff| |x_f||
/ ||b_f|| of the system they stop at that floor and say so).  Solver paths,
each timed after the same full CSR assembly:
  fp64       plain FP64 Jacobi PCG on the assembled matrix (the baseline)
  matfree    FP64 Jacobi PCG, matrix-free Tet10 operator
  fp64_2l    two-level flexible CG, assembled outer products, FP64
             matrix-free Chebyshev + rigid-body coarse correction inside
  mixed_2l   the same with the inner Chebyshev products in FP32
  mixed_mf2l as mixed_2l with matrix-free FP64 outer products (solves the
             matrix-free operator, equal to the assembled one up to rounding)
  mixed_2l_mfma  mixed_2l with the FP32 element forces on the matrix cores
             (mf_t10_elem_mfma, v_mfma_f32_16x16x4f32) instead of VALU
  (also selectable: mixed_mf, mixed, mixed16)
`value` is the fastest path that solves the assembled matrix and meets the
bar max(1e-10, 3 x FP64 floor); `config.solver` names it.  DOF counts all
nodal DOF (3N).  fp64_2l vs mixed_2l is the same algorithm in two
precisions: `speedup_precision_only`.

The roofline object describes the dominant kernel of that path (its FP32
Chebyshev product): average duration from HIP events on sampled launches of
the timed solves (cf_spmv_profile), bytes per launch stated per path in
`paths.*.op_bytes` (implementation's algorithmic bytes) and `op_min_bytes`
(connectivity, geometry and the vector streams only).

Extra keys: `configs` holds the 80x10x10 cantilever (fp64 and mixed_2l),
the batched transient heat workload (64 jittered P1 boxes, node-steps/s and
the batched SpMV roofline) and the adjoint gradient at 40^3 (wall time).
"""

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (nx, ny, nz)
    "tet10-160x20x20": (160, 20, 20),
    "tet10-80x10x10": (80, 10, 10),
    # 64 jittered P1 boxes, sides U{10..40}, 50 Crank-Nicolson steps (batched path)
    "heat-batch64": None,
}
LIMS = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
E_MOD, NU, TRACTION, TOL = 210e9, 0.3, -1e6, 1e-10
HBM_PEAK_GBS = 8000.0
CPU_SAMPLE = (56, 6, 6)   # ~9 s of single-core CPU work per run on the GPU host
CPU_DNF_CAP_S = 30.0
DEFAULT_PATHS = "fp64,matfree,fp64_2l,mixed_mf2l,mixed_2l,mixed_2l_mfma"
ASSEMBLED = {"fp64", "fp64_2l", "mixed_2l", "mixed_2l_mfma", "mixed", "mixed16"}   # solve A itself


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="tet10-160x20x20", choices=sorted(WORKLOADS))
    ap.add_argument("--precond64", default="jacobi", choices=["jacobi", "block3"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra-configs", action="store_true",
                    help="skip the 80x10x10, batched-heat and adjoint keys")
    ap.add_argument("--paths", default=DEFAULT_PATHS,
                    help="comma-separated solver paths to time (fp64 always runs first)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="before the workload, copy a 256 MiB device buffer 3 times "
                         "(reference kernel for HBM byte counters)")
    ap.add_argument("--cpu-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-dnf", action="store_true", help=argparse.SUPPRESS)
    return ap.parse_args()


# ----------------------------------------------------------------- CPU baseline

def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline():
    """Oracle (NumPy/SciPy restatement of the reference path: per-element
    Python assembly, coo->csr, the reduced system of apply_dirichlet_vec3 and
    SuperLU spsolve) on a bounded sample of the same cantilever, one thread,
    median of 3 runs with assemble / BC / solve columns."""
    import resource

    import numpy as np
    import scipy.sparse.linalg as spla
    from oracle import elastic as oel
    from oracle import meshes as om
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        m = om.box_mesh_tet10(*CPU_SAMPLE, **LIMS)
        t1 = time.perf_counter()
        A, rhs = oel.assemble_elasticity_tet10(m, E_MOD, NU,
                                               traction={"xmax": (0.0, 0.0, TRACTION)})
        t2 = time.perf_counter()
        Aff, bf, free = oel.reduced_system_vec3(A, rhs, m, {"xmin": (0.0, 0.0, 0.0)})[:3]
        t3 = time.perf_counter()
        xf = spla.spsolve(Aff.tocsc(), bf)
        t4 = time.perf_counter()
        assert np.isfinite(xf).all()
        runs.append((t2 - t1, t3 - t2, t4 - t3, t1 - t0))
    runs.sort(key=lambda r: r[0] + r[1] + r[2])
    asm, bc, sol, mesh_s = runs[1]
    n = 3 * m.n_nodes
    return {"value": n / (asm + bc + sol), "unit": "DOF/s", "cores": 1, "kind": "port",
            "sample": (f"oracle assemble + reduced system (apply_dirichlet_vec3) + SuperLU "
                       f"spsolve on the same cantilever at {CPU_SAMPLE[0]}x{CPU_SAMPLE[1]}x"
                       f"{CPU_SAMPLE[2]} cells ({n} DOF), median of 3 runs; SuperLU fill "
                       "grows superlinearly, so the full-size rate is lower"),
            "median_s": {"assemble": asm, "bc": bc, "solve": sol, "mesh_untimed": mesh_s},
            "runs_s": [[round(v, 3) for v in r[:3]] for r in runs],
            "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "OPENBLAS_NUM_THREADS": os.environ.get("OPENBLAS_NUM_THREADS"),
            "peak_rss_mb": resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1024.0}


def cpu_dnf_child():
    """The oracle path on the full 160x20x20 cantilever; prints its stage."""
    from oracle import elastic as oel
    from oracle import meshes as om
    print("stage mesh", flush=True)
    m = om.box_mesh_tet10(*WORKLOADS["tet10-160x20x20"], **LIMS)
    print("stage assembly", flush=True)
    A, rhs = oel.assemble_elasticity_tet10(m, E_MOD, NU, traction={"xmax": (0.0, 0.0, TRACTION)})
    print("stage solve", flush=True)
    oel.apply_dirichlet_vec3(A, rhs, m, {"xmin": (0.0, 0.0, 0.0)})
    print("stage done", flush=True)


def _single_thread_env():
    return dict(os.environ, OPENBLAS_NUM_THREADS="1", OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")


def run_cpu_baseline_subprocess():
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-only"],
                         env=_single_thread_env(), capture_output=True, text=True, timeout=900)
    if out.returncode != 0:
        return {"value": None, "error": out.stderr[-500:]}
    return json.loads(out.stdout.strip().splitlines()[-1])


def run_cpu_dnf_subprocess():
    """The oracle path at full size under a wall-clock cap: DNF record with
    elapsed time, peak RSS and the stage it reached."""
    import psutil
    proc = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-dnf"],
                            env=_single_thread_env(), stdout=subprocess.PIPE,
                            stderr=subprocess.DEVNULL, text=True)
    ps = psutil.Process(proc.pid)
    t0 = time.perf_counter()
    peak = 0
    while proc.poll() is None and time.perf_counter() - t0 < CPU_DNF_CAP_S:
        try:
            peak = max(peak, ps.memory_info().rss)
        except psutil.Error:
            break
        time.sleep(0.25)
    finished = proc.poll() is not None
    if not finished:
        proc.kill()
    out = proc.communicate()[0]
    stages = [ln.split()[1] for ln in out.splitlines() if ln.startswith("stage ")]
    return {"status": "done" if finished and stages[-1:] == ["done"] else "DNF",
            "cap_s": CPU_DNF_CAP_S, "elapsed_s": round(time.perf_counter() - t0, 1),
            "peak_rss_mb": round(peak / 2 ** 20, 1), "stage_reached": stages[-1] if stages else None,
            "what": "oracle path (per-element assembly, reduced system, SuperLU) on the full "
                    "160x20x20 cantilever, one thread"}


def scipy_cg_estimate(A_csr, rhs, mask, iterations_needed, n_it=10):
    """SciPy CG with Jacobi on the exported FP64 CSR (not the reference path,
    which uses SuperLU): seconds per iteration, timed over n_it iterations,
    and the time the FP64 GPU path's iteration count would take."""
    import numpy as np
    import scipy.sparse.linalg as spla
    free = np.nonzero(~mask)[0]
    Aff = A_csr[free][:, free].tocsr()
    bf = rhs[free]
    dinv = 1.0 / Aff.diagonal()
    M = spla.LinearOperator(Aff.shape, matvec=lambda v: dinv * v)
    t0 = time.perf_counter()
    spla.cg(Aff, bf, rtol=1e-30, maxiter=n_it, M=M)
    per_it = (time.perf_counter() - t0) / n_it
    return {"label": "not the reference path: SciPy CG + Jacobi on the exported CSR, "
                     f"{n_it} iterations timed, one thread",
            "s_per_iteration": per_it, "iterations_of_gpu_fp64": iterations_needed,
            "estimated_solve_s": per_it * iterations_needed}


# ----------------------------------------------------------------- cantilever

def op_bytes(path, n_elem, n_runs, n_nodes, n_dof, nnz_b):
    """(algorithmic bytes, strict-minimum bytes) per launch of the path's
    dominant operator."""
    sell64 = 76 * nnz_b + 4 * (n_nodes + 1) + 16 * n_dof
    if path == "fp64":
        return sell64, sell64
    if path == "matfree":
        # per element conn 40 + geometry 80 + ranks 20 B; per run its start
        # 2 B, an FP64 partial written and read 2 x 24 B, its node-list
        # entry 4 B; node offsets; per DOF x gathered, x read for p.Ap, y
        # written 3 x 8 B, fixed flag 1 B
        return (140 * n_elem + 54 * n_runs + 4 * (n_nodes + 1) + 25 * n_dof,
                120 * n_elem + 24 * n_dof)
    if path in ("mixed", "mixed16"):
        vb = 36 if path == "mixed" else 18
        return (vb + 4) * nnz_b + 4 * n_nodes + 24 * n_dof, (vb + 4) * nnz_b + 24 * n_dof
    # matrix-free Chebyshev step in T (4 or 8 B): per element conn 40 +
    # geometry 10 T + ranks 20 B; per run its start 2 B, a partial written and
    # read 2 x 3 T, its node-list entry 4 B; node offsets; per DOF z_k
    # gathered, S, w, z_k, z_k-1 read, z_k+1 written: 6 T
    t = 8 if path in ("fp64_2l", "fp64_mf2l") else 4
    return ((60 + 10 * t) * n_elem + (6 + 6 * t) * n_runs + 4 * (n_nodes + 1) + 6 * t * n_dof,
            (40 + 10 * t) * n_elem + 6 * t * n_dof)


OP_NAMES = {
    "fp64": "spmv_sell<double,double,double>",
    "matfree": "mf_t10_elem + mf_gather_runs",
    "mixed": "cheb_step<float>", "mixed16": "cheb_step<_Float16>",
    "mixed_mf": "mf_t10_elem_lp<float> + mf_gather_cheb<float>",
    "mixed_mf2l": "mf_t10_elem_lp<float> + mf_gather_cheb<float>",
    "mixed_2l": "mf_t10_elem_lp<float> + mf_gather_cheb<float>",
    "mixed_2l_mfma": "mf_t10_elem_mfma + mf_gather_cheb<float>",
    "fp64_2l": "mf_t10_elem_lp<double> + mf_gather_cheb<double>",
    "fp64_mf2l": "mf_t10_elem_lp<double> + mf_gather_cheb<double>",
}


def run_cantilever(workload, path_list, steps, warmup, precond64="jacobi", keep_csr=False):
    """Time each solver path on the cantilever; returns (results, context)."""
    import ctypes

    import numpy as np
    import torch

    from cloudfem_amd import _native as N
    from cloudfem_amd.coef import dirichlet_vec3
    from cloudfem_amd.device import DeviceMesh, build_incidence, to_dev, vector_csr
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.mesh.faces import boundary_faces10
    from cloudfem_amd.physics.elastic3d import lame
    from cloudfem_amd.solver import (DeviceMatrix, pcg, pcg_matfree_t10, pcg_matfree_t10_mixed,
                                     pcg_t10_twolevel)

    dev = torch.device("cuda", 0)
    lib = N.load()
    dims = WORKLOADS[workload]
    mesh = box_mesh_tet10(*dims, **LIMS)
    n_nodes, n_dof, n_elem = mesh.n_nodes, 3 * mesh.n_nodes, mesh.n_elements
    xyz = to_dev(mesh.nodes.T.copy(), torch.float64, dev)
    conn = to_dev(mesh.elements.astype(np.int32), torch.int32, dev)
    faces_np = boundary_faces10(mesh, "xmax")
    faces = to_dev(faces_np.astype(np.int32), torch.int32, dev)
    coef = torch.zeros((faces_np.shape[0], 3, 3), dtype=torch.float64, device=dev)
    coef[:, :, 2] = TRACTION
    mask, fv = dirichlet_vec3(mesh, {"xmin": (0.0, 0.0, 0.0)})
    fixed = to_dev(mask.astype(np.uint8), torch.uint8, dev)
    x0 = to_dev(np.where(mask, fv, 0.0), torch.float64, dev)
    lam, mu = lame(E_MOD, NU)
    torch.cuda.synchronize()
    maxit = 200000
    last_sys = {}

    def step(path):
        dm = DeviceMesh.from_device(xyz[0], xyz[1], xyz[2], conn)
        vals = dm.matrix_values(N.OP_T10_ELAST, lam=lam, mu=mu)
        rhs = torch.zeros(n_dof, dtype=torch.float64, device=dev)
        dm.add_face_rhs_device(N.FACE_T10_TRACTION, faces, coef, rhs,
                               build_incidence(faces.reshape(-1), n_nodes, dev))
        x = x0.clone()
        A = DeviceMatrix(N.FMT_NB3, dm.pattern(), vals)
        kw = dict(tol=TOL, maxit=maxit)
        if path == "matfree":
            stats = pcg_matfree_t10(dm, lam, mu, rhs, x, fixed, precond=precond64, **kw)
        elif path == "mixed_mf":
            stats = pcg_matfree_t10_mixed(dm, lam, mu, rhs, x, fixed, **kw)
        elif path in ("mixed_mf2l", "fp64_mf2l"):
            stats = pcg_t10_twolevel(dm, lam, mu, rhs, x, fixed, A=None,
                                     precision="mixed" if path == "mixed_mf2l" else "fp64", **kw)
        elif path in ("mixed_2l", "fp64_2l"):
            stats = pcg_t10_twolevel(dm, lam, mu, rhs, x, fixed, A=A,
                                     precision="mixed" if path == "mixed_2l" else "fp64", **kw)
        elif path == "mixed_2l_mfma":
            os.environ["CF_INNER_KERNEL"] = "mfma"
            try:
                stats = pcg_t10_twolevel(dm, lam, mu, rhs, x, fixed, A=A, precision="mixed",
                                         **kw)
            finally:
                os.environ.pop("CF_INNER_KERNEL", None)
        elif path == "fp64":
            stats = pcg(A, rhs, x, fixed, precision="fp64", precond=precond64, **kw)
        else:
            stats = pcg(A, rhs, x, fixed, precision=path, precond="jacobi", **kw)
        last_sys.update(A=A, rhs=rhs, dm=dm, vals=vals)
        return x, stats, dm.pattern().nnz

    def timed(path):
        for _ in range(warmup):
            step(path)
        torch.cuda.synchronize()
        lib.cf_spmv_profile(1, None, None, None, None)
        times, last = [], None
        for _ in range(steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            last = step(path)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        a64, a32 = ctypes.c_double(), ctypes.c_double()
        c64, c32 = ctypes.c_int64(), ctypes.c_int64()
        lib.cf_spmv_profile(0, ctypes.byref(a64), ctypes.byref(c64), ctypes.byref(a32),
                            ctypes.byref(c32))
        return 1e3 * float(np.mean(times)), last, (a64.value, c64.value, a32.value, c32.value)

    # runs of the matrix-free operators: distinct (tile of 256 elements, node)
    # pairs with the elements in first-corner order (DeviceMesh.matfree_arrays)
    els = mesh.elements[np.argsort(mesh.elements[:, 0], kind="stable")].astype(np.int64)
    n_runs = int(np.unique((np.arange(n_elem, dtype=np.int64)[:, None] // 256) * n_nodes
                           + els).size)
    results, sols = {}, {}
    order = ["fp64"] + [p for p in path_list if p != "fp64"]
    floor = None
    ctx = {"mesh": mesh, "mask": mask, "n_dof": n_dof, "n_runs": n_runs}
    for path in order:
        try:
            ms, (x, st, nnz_b), prof = timed(path)
        except (N.SolverNotConverged, N.SolverBreakdown, N.CFError) as exc:
            results[path] = {"converged": False, "error": str(exc)[-300:]}
            continue
        sols[path] = x
        if path == "fp64":
            # FP64 floor of the system at the FP64 solution, |A_ff| |x_f| on the device
            floor = last_sys["A"].fp64_floor(x, last_sys["rhs"], fixed)
            ctx["nnz_b"] = nnz_b
            if keep_csr:
                ctx["A_csr"] = vector_csr(last_sys["dm"].pattern(), last_sys["vals"],
                                          dense3=True)
                ctx["rhs"] = last_sys["rhs"].cpu().numpy()
        inner = path not in ("fp64", "matfree")
        b_alg, b_min = op_bytes(path, n_elem, n_runs, n_nodes, n_dof, nnz_b)
        avg = prof[2] if inner else prof[0]
        r = {"ms_per_step": ms, "value": n_dof / (ms / 1e3), "iterations": st["iterations"],
             "restarts": st["restarts"], "floor_limited": st["floor_limited"],
             "rel_residual": st["rel_residual"],
             "op": OP_NAMES[path], "op_avg_ms": avg,
             "op_launches_sampled": prof[3] if inner else prof[1],
             "op_bytes": b_alg, "op_min_bytes": b_min,
             "op_gbs": b_alg / (avg / 1e3) / 1e9 if avg > 0 else None,
             "solves_assembled_matrix": path in ASSEMBLED}
        if inner:
            r["inner_products"] = st["residual_updates"]
            r["outer_op_avg_ms"] = prof[0]
        if path != "fp64":
            r["rel_diff_vs_fp64"] = float(torch.linalg.vector_norm(x - sols["fp64"])
                                          / torch.linalg.vector_norm(sols["fp64"]))
        results[path] = r
    bar = max(TOL, 3.0 * floor)
    for r in results.values():
        if "rel_residual" in r:
            r["residual_over_floor"] = r["rel_residual"] / floor
            r["meets_bar"] = bool(r["rel_residual"] <= bar)
    u = sols["fp64"].cpu().numpy()
    ctx.update(floor=floor, bar=bar, tip=float(u[2::3][mesh.boundaries["xmax"]].mean()))
    return results, ctx


# ----------------------------------------------------------------- other configs

def bench_batch_heat(steps, warmup):
    """64 jittered P1 box meshes (sides U{10..40}, k ~ lognormal(0, 0.5) per
    element, rho c = 3.6e6, Robin h = 20 at 20 C on all faces, T0 = 100 C),
    50 Crank-Nicolson steps of 0.1 s, FP64 and mixed.  One step of the
    bench = the whole batched simulation (assembly of all meshes + 50 steps);
    the stepping alone (cf_batch_theta_run) is reported beside it, and the
    batched SpMV (12 B per stored entry + row pointers + x and y) against
    HBM peak at the tuned launch shape."""
    import numpy as np
    from cloudfem_amd.fem.batch_heat import jittered_box_batch, last_batch_stats
    from cloudfem_amd.fem.batch_heat import transient_heat_batch
    meshes, ks = jittered_box_batch()
    robin = {f: (20.0, 20.0) for f in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")}
    kw = dict(k=ks, rho_cp=3.6e6, robin=robin, t_end=5.0, dt=0.1, T0=100.0, method="cn",
              history=False)
    n_nodes = sum(m.n_nodes for m in meshes)
    paths, finals = {}, {}
    for prec in ("fp64", "mixed"):
        for _ in range(max(warmup, 1)):   # first call also tunes the launch shape
            transient_heat_batch(meshes, precision=prec, **kw)
        runs = []
        for _ in range(steps):
            t0 = time.perf_counter()
            out = transient_heat_batch(meshes, precision=prec, **kw)
            runs.append((time.perf_counter() - t0, dict(last_batch_stats)))
        finals[prec] = np.concatenate([o[1] for o in out])
        st = runs[-1][1]
        run_s = float(np.mean([r[1]["run_s"] for r in runs]))
        paths[prec] = {"ms_per_step": 1e3 * float(np.mean([r[0] for r in runs])),
                       "stepping_ms": 1e3 * run_s,
                       "setup_ms": 1e3 * float(np.mean([r[1]["setup_s"] for r in runs])),
                       "node_steps_per_s_stepping": n_nodes * 50 / run_s,
                       "iterations": st["iterations"], "sweeps": st["sweeps"],
                       "worst_step": st["worst_step"], "graph_replays": st["graph_replays"],
                       "max_rel_residual": st["max_rel_residual"],
                       "launch_shape": {"lanes": st["lanes"], "rows_per_tile": st["rows_per_tile"]}}
        tune = st["autotune"].get("timings_us") or {}
        if tune:
            rows, nnz = st["rows"], st["nnz"]
            b = 12 * nnz + 4 * (rows + 1) + 16 * rows
            us = min(tune.values())
            paths[prec]["spmv_roofline"] = {"bytes": b, "best_us": us,
                                            "achieved_gbs": b / (us * 1e-6) / 1e9,
                                            "frac": b / (us * 1e-6) / 1e9 / HBM_PEAK_GBS}
    best = min(paths, key=lambda k: paths[k]["ms_per_step"])
    return {"workload": "64 jittered P1 boxes [0,0.1]^3, sides U{10..40} (seed 0), k "
                        "lognormal(0,0.5), rho c 3.6e6, Robin h=20 T=20 on 6 faces, T0=100, "
                        "CN dt=0.1 t_end=5",
            "n_nodes": n_nodes, "n_elements": sum(m.n_elements for m in meshes),
            "node_steps_per_s": n_nodes * 50 / (paths[best]["ms_per_step"] / 1e3),
            "solver": best, "paths": paths,
            "mixed_vs_fp64_max_rel_diff": float(np.abs(finals["mixed"] - finals["fp64"]).max()
                                                / np.abs(finals["fp64"]).max())}


def bench_adjoint(steps, warmup):
    """dJ/dk on the 40^3 Kuhn box (384,000 tets): k_e ~ lognormal(0, 0.5)
    (seed 0), f = 1e4, T = 0 on x = 0, J = 1/2 ||T_S - T_target||^2 over 500
    sensor nodes off x = 0 (seed 1), T_target ~ U(0, 1) (seed 2).  One step =
    forward solve + adjoint solve + per-element gradient kernel."""
    import numpy as np
    import torch
    from cloudfem_amd.adjoint import steady_diffusion
    from cloudfem_amd.device import DeviceMesh
    from cloudfem_amd.mesh import box_mesh
    m = box_mesh(40, 40, 40)
    k = np.random.default_rng(0).lognormal(0.0, 0.5, m.n_elements)
    cand = np.setdiff1d(np.arange(m.n_nodes), m.boundaries["xmin"])
    sens = np.random.default_rng(1).choice(cand, 500, replace=False)
    target = np.random.default_rng(2).uniform(0.0, 1.0, 500)
    dm = DeviceMesh(m)
    s = torch.as_tensor(sens, device="cuda")
    tgt = torch.as_tensor(target, device="cuda")

    def once():
        kap = torch.tensor(k, dtype=torch.float64, device="cuda", requires_grad=True)
        T = steady_diffusion(m, kap, f=1e4, bc={"xmin": 0.0}, tol=1e-12, dm=dm)
        J = 0.5 * ((T[s] - tgt) ** 2).sum()
        J.backward()
        return kap.grad

    for _ in range(warmup):
        once()
    torch.cuda.synchronize()
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        g = once()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    return {"workload": "P1 Kuhn 40^3 (384,000 tets, 68,921 nodes), dJ/dk of 500 sensors",
            "gradient_ms": 1e3 * float(np.mean(times)),
            "grad_norm": float(torch.linalg.vector_norm(g))}


# ----------------------------------------------------------------- main

def main():
    args = parse()
    if args.cpu_only:
        print(json.dumps(cpu_baseline()))
        return 0
    if args.cpu_dnf:
        cpu_dnf_child()
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    from cloudfem_amd import _native as N
    N.load()
    if args.workload == "heat-batch64":
        line = {"metric": "batched transient heat node-steps/s (64 jittered P1 meshes, 50 CN "
                          "steps)", "unit": "node-steps/s", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "data": "synthetic"}
        r = bench_batch_heat(args.steps, args.warmup)
        line.update(value=r["node_steps_per_s"],
                    ms_per_step=r["paths"][r["solver"]]["ms_per_step"],
                    dtype="f32+f64" if r["solver"] == "mixed" else "f64",
                    config={"workload": "heat-batch64: " + r["workload"], "solver": r["solver"]},
                    paths=r["paths"])
        print(json.dumps(line))
        return 0

    if args.copy_probe:
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    want_cpu = not args.no_cpu_baseline and args.workload == "tet10-160x20x20"
    results, ctx = run_cantilever(args.workload, args.paths.split(","), args.steps, args.warmup,
                                  args.precond64, keep_csr=want_cpu)
    valid = [k for k, r in results.items()
             if r.get("meets_bar") and r.get("solves_assembled_matrix")]
    best = min(valid or [k for k in results if "ms_per_step" in results[k]],
               key=lambda k: results[k]["ms_per_step"])
    rb = results[best]
    n_dof = ctx["n_dof"]
    fp64_ms = results["fp64"]["ms_per_step"]
    speed = {k: fp64_ms / r["ms_per_step"] for k, r in results.items()
             if k != "fp64" and "ms_per_step" in r}
    precision_only = None
    if "ms_per_step" in results.get("fp64_2l", {}) and "ms_per_step" in results.get("mixed_2l", {}):
        a, b = results["fp64_2l"], results["mixed_2l"]
        precision_only = {"fp64_2l_ms": a["ms_per_step"], "mixed_2l_ms": b["ms_per_step"],
                          "speedup": a["ms_per_step"] / b["ms_per_step"],
                          "inner_product_ms": {"fp64": a["op_avg_ms"], "fp32": b["op_avg_ms"]},
                          "iterations": {"fp64": a["iterations"], "mixed": b["iterations"]}}
    mfma_vs_valu = None
    if "ms_per_step" in results.get("mixed_2l", {}) and \
            "ms_per_step" in results.get("mixed_2l_mfma", {}):
        a, b = results["mixed_2l_mfma"], results["mixed_2l"]
        mfma_vs_valu = {"mfma_ms": a["ms_per_step"], "valu_ms": b["ms_per_step"],
                        "inner_product_ms": {"mfma": a["op_avg_ms"], "valu": b["op_avg_ms"]},
                        "inner_products": {"mfma": a.get("inner_products"),
                                           "valu": b.get("inner_products")},
                        "solution_rel_diff": a.get("rel_diff_vs_fp64")}
    line = {
        "metric": "Tet10 elasticity DOF/s (assemble+PCG to 1e-10), mixed vs FP64 on 1 GPU",
        "value": n_dof / (rb["ms_per_step"] / 1e3),
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rb["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if best.startswith("mixed") else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} Tet10 cantilever 0.4x0.05x0.05 m, "
                               f"{ctx['mesh'].n_elements} elements, {n_dof} DOF, E=210GPa "
                               "nu=0.3, clamped x=0, 1 MPa -z traction at x=0.4",
                   "n_dof": n_dof, "nnz_blocks": ctx["nnz_b"], "tol": TOL, "solver": best},
        "valid": bool(valid) and best in valid,
        "residual_bar": {"tol": TOL, "fp64_floor": ctx["floor"], "bar": ctx["bar"],
                         "rule": "true FP64 relative residual <= max(tol, 3 x eps "
                                 "||A_ff| |x_f|| / ||b_f||)"},
        "paths": results,
        "speedup_vs_fp64": speed,
        "speedup_precision_only": precision_only,
        "element_forces_mfma_vs_valu": mfma_vs_valu,
        "tip_uz": ctx["tip"],
        "roofline": {"bound": "hbm", "achieved": rb["op_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s",
                     "frac": rb["op_gbs"] / HBM_PEAK_GBS if rb["op_gbs"] else None,
                     "traffic": None, "kernel": rb["op"], "avg_ms": rb["op_avg_ms"],
                     "launches_sampled": rb["op_launches_sampled"],
                     "algorithmic_bytes": rb["op_bytes"],
                     "strict_min_bytes": rb["op_min_bytes"],
                     "frac_strict_min": (rb["op_min_bytes"] / (rb["op_avg_ms"] / 1e3) / 1e9
                                         / HBM_PEAK_GBS) if rb["op_avg_ms"] else None},
        "spmv_fp64_roofline": {"achieved": results["fp64"]["op_gbs"], "peak": HBM_PEAK_GBS,
                               "frac": (results["fp64"]["op_gbs"] / HBM_PEAK_GBS
                                        if results["fp64"]["op_gbs"] else None)},
    }
    traffic_file = os.path.join(ROOT, "profiles", f"pmc_traffic_{args.workload}.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            line["roofline"]["traffic"] = json.load(fh).get("paths", {}).get(best)
    if not args.no_extra_configs and args.workload == "tet10-160x20x20":
        st2 = max(1, min(args.steps, 5))
        r80, c80 = run_cantilever("tet10-80x10x10", ["fp64", "mixed_2l"], st2, 1)
        line["configs"] = {
            "tet10-80x10x10": {
                "n_dof": c80["n_dof"], "fp64_floor": c80["floor"],
                "paths": {k: {q: r[q] for q in ("ms_per_step", "value", "iterations",
                                                "rel_residual", "meets_bar")
                              if q in r} for k, r in r80.items()},
                "speedup_mixed_2l_vs_fp64": (r80["fp64"]["ms_per_step"]
                                             / r80["mixed_2l"]["ms_per_step"]
                                             if "ms_per_step" in r80.get("mixed_2l", {})
                                             else None)},
            "heat-batch64": bench_batch_heat(max(1, min(args.steps, 3)), 1),
            "adjoint-40": bench_adjoint(max(1, min(args.steps, 5)), 1)}
    if want_cpu:
        cb = run_cpu_baseline_subprocess()
        cb["full_size_reference_path"] = run_cpu_dnf_subprocess()
        cb["scipy_cg_full_size"] = scipy_cg_estimate(ctx["A_csr"], ctx["rhs"], ctx["mask"],
                                                     results["fp64"]["iterations"])
        line["cpu_baseline"] = cb
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
