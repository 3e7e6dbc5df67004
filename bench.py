"""Throughput of the Poisson reconstruction path on one MI355X.

A step is one whole reconstruction (normalise, octree, field, per-depth CG,
iso-value, geometry, marching cubes with subdivision) of a 200,000-point
torus at octree depth 7, the points already resident in HBM.  By default both
forms are timed, the plain FP32 form (the reference's CG rule) and the
mixed-precision form (FP16 field storage, compensated FP32 CG, FP64
refinement to a 1e-5 true relative residual); the JSON line's value is the
mixed form's, with the FP32 figures beside it.  The dominant kernel is the
concurrent CG solve; its roofline uses the algorithmic bytes
sum_d k_d (8 nnz_d + 48 n_d) over its launch time, which the library
measures with hipEvents on its own stream.  The CPU baseline is
the oracle's FP32 restatement of the same pipeline, timed on a bounded
sample on the host cores.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_TAG = {"torus": "torus_d7", "sphere": "sphere_d6", "displaced": "displaced_d9"}
WORKLOADS = {"torus": "200,000-point torus (R 0.3, r 0.1), normals rotated <= 2 deg, depth 7",
             "sphere": "20,000-point sphere (r 0.4), jittered, depth 6",
             "displaced": "2,000,000-point displaced sphere, depth 9"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="torus", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="both", choices=["both", "fp32", "mixed"],
                    help="both: time the FP32 form and the mixed-precision form; the line's value is the latter")
    ap.add_argument("--cpu-sample", type=int, default=5, help="timed reconstructions in the CPU-baseline median (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="OpenMP threads of the CPU baseline (0: all usable, capped by OMP_NUM_THREADS)")
    ap.add_argument("--batch-reps", type=int, default=3,
                    help="timed repetitions of the 64-cloud batch, sequential and batched (0: skip)")
    ap.add_argument("--batch-forms", default="fp32,mixed",
                    help="precision forms of the 64-cloud batch leg, comma-separated (fp32, mixed)")
    ap.add_argument("--batch-pass", type=int, default=0,
                    help="clouds merged into each launch of the batched form (0: all 64)")
    ap.add_argument("--weld-reps", type=int, default=20,
                    help="timed welds, mesh normals and point-normal estimations of the workload (0: skip)")
    ap.add_argument("--d9-steps", type=int, default=3,
                    help="timed reconstructions of the 2M-point depth-9 displaced sphere, each precision (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=80,
                    help="requests in the served Poisson-arrival stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=8.0, help="arrival rate of the served stream, per second")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain elementwise copy of MIB MiB three times (HBM counter calibration)")
    return ap.parse_args()


def copy_probe(mib):
    """A kernel of known traffic: y = x + 0 over mib MiB of float32 reads and
    writes, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.add(x, 0.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 3}))


def measured_traffic(workload, precision, kernel):
    """HBM bytes per solve of `kernel` ("a+b": the kernels of a solve that runs
    as two launches, summed) from the committed rocprofv3 counter summary of the
    same workload (FETCH_SIZE and WRITE_SIZE passes scaled by a copy kernel of
    known size), or None when there is none."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s_%s.md" % (PROFILE_TAG.get(workload, workload), precision))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith("TRAFFIC "):
            t = json.loads(line.strip()[len("TRAFFIC "):])
            parts = [t.get(k) for k in kernel.split("+")]
            return None if any(p is None for p in parts) else sum(parts)
    return None


def _cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                return ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "unknown"


def cpu_baseline(xyz, nrm, depth, reps, threads):
    """The FP32 port of the reference algorithm (same stages, quirks and CG
    rule: absolute 1e-5 on the recursive residual, at most 10000 iterations),
    compiled here with -march=native and OpenMP: one full-size warm-up, then
    the median of `reps` timed reconstructions on all usable cores and one on
    a single thread."""
    import statistics
    import tempfile
    import __graft_entry__ as ge
    orc = ge.load_oracle()
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    if threads <= 0:
        threads = min(usable, int(os.environ.get("OMP_NUM_THREADS", usable) or usable))
    with tempfile.TemporaryDirectory() as tmp:
        L = orc.native_port(tmp)
        orc.time_port(L, xyz, nrm, depth, threads)  # warm-up at full size
        ms = [orc.time_port(L, xyz, nrm, depth, threads) for _ in range(reps)]
        one = orc.time_port(L, xyz, nrm, depth, 1)
    med = statistics.median(ms)
    return {"value": 1000.0 / med, "unit": "reconstructions/sec", "cores": threads, "kind": "port",
            "one_thread_value": 1000.0 / one, "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "build": "g++ -O3 -march=native -fopenmp -ffp-contract=off",
            "sample": "median of %d whole FP32 reconstructions of the same input after a full-size warm-up "
                      "(%s ms), each stage timed from normalisation to the end of subdivision, basis tables "
                      "excluded; one_thread_value is one run on a single thread (%.0f ms)"
                      % (reps, ", ".join("%.0f" % m for m in ms), one)}





def cg_kernel_name(st, precision):
    """The CG kernel(s) of the solve: a mixed solve with several rows per
    thread and a depth 9 or deeper (the indefinite depth solved with MINRES,
    in a launch of its own) runs two kernels, named "a+b"."""
    if st["cg_row_form"]:
        return "cg_mixed_row" if precision == "mixed" else "cg_fp32_row"
    if precision == "mixed":
        return "cg_mixed_all_depths+cg_mixed_minres" if st["depth"] >= 9 else "cg_mixed_all_depths"
    return "cg_all_depths"


def cg_algorithmic_bytes(st, precision):
    """HBM bytes the CG launch must move, per depth with z stored operator
    entries (4 B value + 4 B column) and n rows, each array read once per pass.
    One row per thread (stencil held on chip, read once): FP32 8z + 16n (b in,
    records out, x out) + 16n per iteration (8-byte records {r, p} read and
    written); mixed 8z + 8n (b, x) + 28n per outer step (rhs, 12-byte records
    out; b, x, residual) + 28n per inner iteration (records {r, w, s} read and
    written, x out).  Several rows per thread (stencil read every sweep):
    FP32 (CGAlgorithm.cu:10-80) 8z + 48n per iteration; mixed 8n, then
    16z + 32n per outer step and 8z + 40n per inner iteration."""
    total = 0
    row = bool(st["cg_row_form"])
    for k, z, n, o in zip(st["cg_iterations"], st["cg_nnz"], st["nodes_per_depth"], st["cg_outer"]):
        if row and precision == "mixed":
            total += 8 * z + 8 * n + o * 28 * n + k * 28 * n
        elif row:
            total += 8 * z + 16 * n + k * 16 * n
        elif precision == "mixed":
            total += 8 * n + o * (16 * z + 32 * n) + k * (8 * z + 40 * n)
        else:
            total += k * (8 * z + 48 * n)
    return total


def measure(pkg, tx, tn, depth, precision, steps, warmup):
    """Times `steps` whole reconstructions after `warmup` untimed ones; the CG
    kernel's launch time comes from the library's hipEvents on its stream,
    summed by the library over the timed runs (read once afterwards, so no
    statistics call sits inside the timed region).  The iteration counts, and
    so the algorithmic bytes, are those of every run (the input is fixed)."""
    import torch
    prec = pkg.binding.PRECISION_MIXED if precision == "mixed" else pkg.binding.PRECISION_FP32
    rec = pkg.PoissonReconstruction(depth=depth, precision=prec)
    torch.cuda.synchronize()
    for _ in range(warmup):
        rec.SolvePoissionReconstructionMesh(tx, tn)
    torch.cuda.synchronize()
    before = rec.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        rec.SolvePoissionReconstructionMesh(tx, tn)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = rec.stats()
    rec.close()
    assert st["n_runs"] - before["n_runs"] == steps
    cg_ms = (st["ms_cg_kernel_sum"] - before["ms_cg_kernel_sum"]) / steps
    return {"value": steps / dt, "ms_per_step": dt * 1e3 / steps, "cg_ms": cg_ms,
            "cg_bytes": float(cg_algorithmic_bytes(st, precision)), "stats": st}


# Non-CG algorithmic bytes of one reconstruction of the 200k-point torus at
# depth 7 (octree build, splat, divergence, point values, marching cubes and
# subdivision, each array read or written once), and the per-iteration CG
# bytes 8 nnz + 48 n at fixed iteration counts: the whole-reconstruction
# roofline figure taken with the FP32 form's iteration counts.
NON_CG_BYTES_TORUS_D7 = 156e6


def fixed_k_roofline(st_fp32, ms_per_step):
    total = NON_CG_BYTES_TORUS_D7
    for k, z, n in zip(st_fp32["cg_iterations"], st_fp32["cg_nnz"], st_fp32["nodes_per_depth"]):
        total += k * (8 * z + 48 * n)
    achieved = total / (ms_per_step * 1e-3) / 1e9
    return {"bytes_per_reconstruction": total, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS}


def measure_depth9(pkg, steps):
    """The 2,000,000-point displaced sphere at depth 9, mixed precision (FP16
    vector field, MINRES at the indefinite depth 9, CG elsewhere) and FP32 (the
    reference's CG rule, which runs into its 10000-iteration cap at depth 9)."""
    import torch
    xyz, nrm, depth = pkg.synth.workload("displaced")
    tx, tn = torch.from_numpy(xyz).cuda(), torch.from_numpy(nrm).cuda()
    out = {"workload": WORKLOADS["displaced"]}
    for name in ("mixed", "fp32"):
        r = measure(pkg, tx, tn, depth, name, steps, 1)
        st = r["stats"]
        out[name] = {"recon_per_s": r["value"], "ms_per_step": r["ms_per_step"], "cg_ms": r["cg_ms"],
                     "cg_iterations": st["cg_iterations"], "ms_solve": st["ms_solve"],
                     "true_relres": st["cg_true_relres"] if name == "mixed" else None}
    return out


def measure_serve(pkg, count, rate):
    """A seeded Poisson-arrival stream of `count` requests (10k-1M points,
    depth 6-9 by size, ellipsoids and tori) served with HBM-budget admission
    into batched calls, mixed precision: once with every request waiting from
    the start (the saturated rate; also the warm-up), then on the wall clock
    at `rate` arrivals per second (latency = completion - arrival)."""
    import torch
    reqs = pkg.synth.request_stream(count, rate)
    dev = [(a, torch.from_numpy(x).cuda(), torch.from_numpy(n).cuda(), d) for a, x, n, d in reqs]
    srv = pkg.serve.Server(precision=pkg.binding.PRECISION_MIXED)
    drain = pkg.serve.summary(srv.serve(dev, realtime=False), rate)
    live = pkg.serve.summary(srv.serve(dev, realtime=True), rate)
    budget = srv.budget
    srv.close()
    return {"stream": "%d requests, log-uniform 10k-1M points, depth 6-9 by size" % count,
            "arrival_rate_hz": rate, "recon_per_s": live["recon_per_s"], "p50_ms": live["p50_ms"],
            "p99_ms": live["p99_ms"], "mean_batch": live["mean_batch"], "saturated_recon_per_s": drain["recon_per_s"],
            "per_depth": live["per_depth"], "points": live["points"], "hbm_budget_gb": budget / 1e9}


def measure_weld(pkg, tx, tn, depth, reps):
    """Times psr_weld_mesh on the workload's mixed-precision mesh and
    psr_mesh_normals on the welded mesh (both calls return with their results
    final, so each timing includes its wait)."""
    import torch
    rec = pkg.PoissonReconstruction(depth=depth, precision=pkg.binding.PRECISION_MIXED)
    mesh = rec.SolvePoissionReconstructionMesh(tx, tn)
    nv, nt = mesh.n_vertices, mesh.n_triangles
    rec.WeldMesh()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        w = rec.WeldMesh()
    dt = time.perf_counter() - t0
    # per-vertex normals of the welded mesh (psr_mesh_normals returns once written)
    import ctypes
    nbuf = torch.empty((max(w.n_vertices, 1), 3), dtype=torch.float32, device="cuda")
    L = pkg.binding.lib()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        assert L.psr_mesh_normals(rec._ctx, ctypes.byref(w), nbuf.data_ptr()) == 0
    dn = time.perf_counter() - t1
    # colours of the welded vertices (in the points' frame) from the points, with seeded colours
    norm = rec.export("NORM_PARAMS")
    wv = torch.from_numpy(rec.mesh_to_host(w)[0] * norm[3] + norm[:3]).float().cuda()
    cols = torch.rand((tx.shape[0], 3), generator=torch.Generator().manual_seed(11)).cuda()
    cbuf = torch.empty_like(wv)
    wm = pkg.binding.Mesh(wv.data_ptr(), None, wv.shape[0], 0)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(reps):
        assert L.psr_vertex_colors(rec._ctx, ctypes.byref(wm), tx.data_ptr(), cols.data_ptr(), tx.shape[0], 12,
                                   cbuf.data_ptr()) == 0
    dc = time.perf_counter() - t2
    out = {"ms": dt * 1e3 / reps, "vertices": nv, "welded_vertices": w.n_vertices, "triangles": nt,
           "normals_ms": dn * 1e3 / reps, "colors_ms": dc * 1e3 / reps, "reps": reps}
    rec.close()
    return out


def measure_point_normals(pkg, tx, reps):
    """Times psr_estimate_normals (k = 10, viewpoint at the origin) on the
    workload's points; the call returns with the normals written."""
    import ctypes
    import torch
    rec = pkg.PoissonReconstruction(depth=7)
    n = tx.shape[0]
    nrm = torch.empty((n, 3), dtype=torch.float32, device=tx.device)
    cur = torch.empty((n,), dtype=torch.float32, device=tx.device)
    vp = (ctypes.c_float * 3)(0.0, 0.0, 0.0)
    L = pkg.binding.lib()
    call = lambda: L.psr_estimate_normals(rec._ctx, tx.data_ptr(), n, 12, 10, vp, nrm.data_ptr(), cur.data_ptr(),
                                          None)
    assert call() == 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        assert call() == 0
    dt = time.perf_counter() - t0
    rec.close()
    return {"ms": dt * 1e3 / reps, "points": int(n), "k": 10, "reps": reps}


def measure_batch(pkg, reps, per_pass, precision="fp32"):
    """The 64-cloud batch (sizes log-uniform in [5k, 150k], ellipsoids and tori,
    depth 7): 64 sequential single-cloud calls against one batched call (one
    launch per stage for all clouds), both in the given precision form.  The
    batch's warm-up calls let its tuner try its CG plans (one per call after an
    untimed first) and settle on the fastest."""
    import torch
    prec = pkg.binding.PRECISION_MIXED if precision == "mixed" else pkg.binding.PRECISION_FP32
    clouds = pkg.synth.batch64()
    dev = [(torch.from_numpy(x).cuda(), torch.from_numpy(n).cuda()) for x, n in clouds]
    torch.cuda.synchronize()
    single = pkg.PoissonReconstruction(depth=7, precision=prec)
    for x, n in dev[:4]:
        single.SolvePoissionReconstructionMesh(x, n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for x, n in dev:
            single.SolvePoissionReconstructionMesh(x, n)
    torch.cuda.synchronize()
    seq = (time.perf_counter() - t0) / reps
    single.close()
    batch = pkg.PoissonReconstruction(depth=7, precision=prec, batch_workers=per_pass)
    for _ in range(10):
        batch.ReconstructBatch(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.ReconstructBatch(dev)
    torch.cuda.synchronize()
    bat = (time.perf_counter() - t0) / reps
    plan = batch.batch_cg_plan()
    batch.close()
    return {"workload": "64 clouds, 5k-150k points (%d total), ellipsoids and tori, depth 7, %s"
                        % (sum(int(x.shape[0]) for x, _ in dev), "mixed" if precision == "mixed" else "FP32"),
            "sequential_recon_per_s": 64 / seq, "batched_recon_per_s": 64 / bat, "speedup": seq / bat,
            "clouds_per_pass": per_pass or 64, "cg_plan": plan, "reps": reps}


def leg(fn, *args):
    """Runs one of the line's secondary measurements; a library error there
    is recorded in its place rather than losing the line (the headline
    measurement is not wrapped)."""
    import torch
    try:
        return fn(*args)
    except Exception as e:  # noqa: BLE001 - reported in the line
        torch.cuda.synchronize()
        return {"error": "%s: %s" % (type(e).__name__, e)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only; multi-GPU runs are not part of this build "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import __graft_entry__ as ge
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    pkg = ge.load_package()
    if not torch.cuda.is_available():
        raise RuntimeError("bench.py needs a GPU")
    xyz, nrm, depth = pkg.synth.workload(a.workload)
    dev = torch.device("cuda:0")
    tx = torch.from_numpy(xyz).to(dev)
    tn = torch.from_numpy(nrm).to(dev)
    forms = ["fp32", "mixed"] if a.precision == "both" else [a.precision]
    res = {f: measure(pkg, tx, tn, depth, f, a.steps, a.warmup) for f in forms}
    head = forms[-1]  # the headline form: mixed precision when both run
    h = res[head]
    st = h["stats"]
    kernel = cg_kernel_name(st, head)
    achieved = h["cg_bytes"] / (h["cg_ms"] * 1e-3) / 1e9
    metric = ("reconstructions/sec at depth 7, 200k pts; mixed-prec vs FP32 at <=1e-5 residual" if a.workload == "torus"
              else "reconstructions/sec, %s" % WORKLOADS[a.workload])
    line = {
        "metric": metric,
        "value": h["value"],
        "unit": "reconstructions/sec",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed" if head == "mixed" else "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[a.workload], "points": int(len(xyz)), "depth": depth,
                   "precision": head, "nodes": st["n_nodes"], "cg_iterations": st["cg_iterations"],
                   "mesh_triangles": st["mesh_triangles"]},
        "roofline": {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(a.workload, head, kernel),
                     "ms_per_launch": h["cg_ms"], "bytes_per_launch": h["cg_bytes"]},
        "stage_ms": {k: st[k] for k in ("ms_octree", "ms_field", "ms_solve", "ms_mesh", "ms_total")},
    }
    if head == "mixed":
        line["mixed"] = {"true_relres_max": max(st["cg_true_relres"]),
                         "storage": "FP16 vector field; FP32 CG sweeps with compensated dots; FP64 refinement"}
    if len(forms) == 2:
        f = res["fp32"]
        line["fp32"] = {"value": f["value"], "ms_per_step": f["ms_per_step"], "cg_ms": f["cg_ms"],
                        "cg_iterations": f["stats"]["cg_iterations"]}
        line["mixed_vs_fp32"] = h["value"] / f["value"]
        if a.workload == "torus":
            # the same whole-reconstruction figure with the FP32 form's iteration counts
            line["roofline_fixed_k"] = fixed_k_roofline(f["stats"], h["ms_per_step"])
            line["roofline_fixed_k"]["frac_fp32_form"] = fixed_k_roofline(f["stats"], f["ms_per_step"])["frac"]
    if a.weld_reps > 0:
        line["weld"] = leg(measure_weld, pkg, tx, tn, depth, a.weld_reps)
        line["point_normals"] = leg(measure_point_normals, pkg, tx, a.weld_reps)
    if a.batch_reps > 0:
        forms = [f.strip() for f in a.batch_forms.split(",") if f.strip()]
        if "fp32" in forms:
            line["batch64"] = leg(measure_batch, pkg, a.batch_reps, a.batch_pass)
        if "mixed" in forms:
            line["batch64_mixed"] = leg(measure_batch, pkg, a.batch_reps, a.batch_pass, "mixed")
    if a.d9_steps > 0 and a.workload != "displaced":
        line["depth9"] = leg(measure_depth9, pkg, a.d9_steps)
    if a.serve_requests > 0:
        line["serve"] = leg(measure_serve, pkg, a.serve_requests, a.serve_rate)
    if a.cpu_sample > 0:
        line["cpu_baseline"] = cpu_baseline(xyz, nrm, depth, a.cpu_sample, a.cpu_threads)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
