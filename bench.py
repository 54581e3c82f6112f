"""Benchmark: HTFETI load-case solves/s on one MI355X (plain FP64 method).

Workload (default): 2x2x2 clusters x 4x4x4 subdomains x 10^3 hex8 elements
(512 subdomains, 3,993 DOF each, ~1.6M global DOF, 588,846 dual rows), linear
elasticity E=2.1e5, nu=0.3, all DOFs clamped on x=0, body force -1 in z, HTFETI
with face-averaged rigid-body cluster gluing, Dirichlet preconditioner, PCPG to a
1e-8 relative dual residual.  A step is one load-case solve (dual rhs, PCPG loop,
displacement recovery) against the set-up mesh resident in HBM; the setup
(assembly, factorisations, explicit operators, coarse problems) is reported
separately.

Prints ONE JSON line (see the repository's benchmark contract).
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

CFG1 = dict(Cx=2, Cy=2, Cz=2, Nx=4, Ny=4, Nz=4, nx=10, ny=10, nz=10, bc="clamp_x0", material="uniform",
            young_modulus="2.1e5", poissons_ratio="0.3", load="body", body_force="-1.0", typeBc="ker",
            preconditioner="Dirichlet_explicit")
# CPU sample: the same 10^3-element subdomains, one cluster of 2x2x2
SAMPLE = dict(Nx=2, Ny=2, Nz=2, nx=10, ny=10, nz=10, bc="clamp_x0", material="uniform", young_modulus="2.1e5",
              poissons_ratio="0.3", load="body", body_force="-1.0", typeBc="ker",
              preconditioner="Dirichlet_explicit")
HBM_PEAK_GBS = 8000.0


def cpu_sample():
    """Time the oracle (numpy/scipy restatement of the reference method: sparse LU
    of every K_reg, implicit K^+ in every PCPG iteration as in Cluster.cpp:1126)
    solving one load case of SAMPLE on one core."""
    from threadpoolctl import threadpool_limits
    from oracle.mesh import opts
    from oracle.feti import HTFETI
    h = HTFETI(opts(eps_iter="1e-8", **SAMPLE), sparse=True)
    reps = 0
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        while True:  # about 10 s of single-core solve time
            r = h.pcpg(eps=1e-8)
            reps += 1
            if time.perf_counter() - t0 > 10.0:
                break
        t = (time.perf_counter() - t0) / reps
    print(json.dumps({"seconds": t, "iterations": r["iterations"], "nsub": h.nsub, "reps": reps}))


def run_cpu_baseline(gpu_iterations, nsub_target):
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    t0 = time.time()
    p = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-sample"], capture_output=True, text=True,
                       env=env, timeout=900)
    if p.returncode != 0:
        return {"value": None, "unit": "solves/s", "cores": 1, "kind": "port",
                "sample": "failed: " + p.stderr[-300:]}
    s = json.loads(p.stdout.strip().splitlines()[-1])
    # per subdomain-iteration cost of the sample, scaled to the workload
    per = s["seconds"] / (s["nsub"] * max(1, s["iterations"] + 3))
    t_full = per * nsub_target * (gpu_iterations + 3)
    return {"value": 1.0 / t_full, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": ("oracle (numpy/scipy, SuperLU K_reg, implicit K+ per iteration) on 2x2x2 subdomains of "
                       "10^3 elements: %.3f s per solve (mean of %d) for %d iterations; scaled per subdomain-iteration to %d subdomains "
                       "x %d iterations" % (s["seconds"], s["reps"], s["iterations"], nsub_target, gpu_iterations)),
            "wall_s": round(time.time() - t0, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", action="store_true")
    ap.add_argument("--small", action="store_true", help="2x2x2 x 5^3 smoke-sized workload")
    a = ap.parse_args()
    if a.cpu_sample:
        cpu_sample()
        return 0
    if a.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path (see DESIGN.md, sharding).",
              file=sys.stderr)
        return 2
    import hfls_amd
    cfg = dict(CFG1)
    workload = "htfeti_cube_8cl_512sub_10e3"
    if a.small:
        cfg.update(Cx=1, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, nx=5, ny=5, nz=5)
        workload = "htfeti_cube_1cl_8sub_5e3"
    t0 = time.perf_counter()
    mesh = hfls_amd.Mesh(device=0, **cfg)
    setup_s = time.perf_counter() - t0
    print("setup %.2f s (factor %.2f s), nlam %d, device bytes %.2f GB" %
          (setup_s, mesh.factor_ms / 1e3, mesh.nlam, mesh.device_bytes / 1e9), file=sys.stderr, flush=True)
    for i in range(a.warmup):
        r = mesh.pcpg(eps=a.eps, max_iter=500, want_lambda=False)
        print("warmup solve: %d iterations, residual %.2e, %.1f ms" % (r["iterations"], r["residual"], r["solve_ms"]),
              file=sys.stderr, flush=True)
    lat = []
    its = []
    t0 = time.perf_counter()
    for i in range(a.steps):
        s0 = time.perf_counter()
        r = mesh.pcpg(eps=a.eps, max_iter=500, want_lambda=False)
        lat.append(time.perf_counter() - s0)
        its.append(r["iterations"])
        assert r["converged"], "PCPG did not converge"
    total = time.perf_counter() - t0
    print("timed %d solves: %.3f s, iterations %s" % (a.steps, total, its), file=sys.stderr, flush=True)
    # roofline of the dominant kernel: interface gather + K+_bb stream (inside F)
    ms, nbytes = mesh.time_stream(0, reps=20)
    ms_s, nbytes_s = mesh.time_stream(1, reps=20)
    achieved = nbytes / (ms * 1e-3) / 1e9
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "load-case solves/s at 1e-8 dual residual (plain FP64)",
        "value": a.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": float(np.mean(lat_ms)),
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (structured cube mesh generated in-process)",
        "config": {"workload": workload, "subdomains": mesh.nsub, "dual_rows": mesh.nlam,
                   "global_dof": 3 * mesh.nPoints, "iterations": its[-1], "eps": a.eps, "precision": "fp64",
                   "parallelism": "single GPU", "setup_s": round(setup_s, 3)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_symv_tiles (K+_bb, packed lower tiles)", "kernel_ms": ms, "bytes_per_launch": nbytes,
                     "S_stream_GBs": nbytes_s / (ms_s * 1e-3) / 1e9},
    }
    mesh.close()
    if not a.no_cpu:
        out["cpu_baseline"] = run_cpu_baseline(its[-1], out["config"]["subdomains"])
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
