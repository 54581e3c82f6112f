"""Benchmark: HTFETI load-case solves/s on one MI355X, batched mixed precision vs
the plain FP64 form.

Workload (default): the 512-subdomain cube, 2x2x2 clusters x 4x4x4 subdomains x
10^3 hex8 elements (3,993 DOF per subdomain, ~1.6M global DOF, 588,846 dual
rows), linear elasticity E=2.1e5, nu=0.3, all DOFs clamped on x=0; HTFETI with
face-averaged rigid-body cluster gluing, Dirichlet preconditioner, PCPG to a 1e-8
relative dual residual.  Load cases: seeded random smoothed tractions on the x=1
face (one load case = one right-hand side).

  * batched mixed precision (the headline `value`): 64 load cases per block,
    K+_bb and S_i applied on the FP16 matrix cores in split FP16 (each value as
    hi + lo FP16 with power-of-two scales, FP32 accumulation), everything else
    FP64, every case certified on its FP64 true residual;
  * plain FP64 (`plain`): the same load cases solved one at a time.

A step is one solve (a block of 64 cases, or one case for the plain form) against
the set-up mesh resident in HBM, loads already on the device: `value` uses the
device time of the solves (HIP events on the library stream); the host-side
throughput including the PCIe transfer of loads and displacements is reported
separately.  The setup (assembly, factorisations, explicit operators, coarse
problems) is reported separately.

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


def cube(e):
    return dict(Cx=2, Cy=2, Cz=2, Nx=4, Ny=4, Nz=4, nx=e, ny=e, nz=e, bc="clamp_x0", material="uniform",
                young_modulus="2.1e5", poissons_ratio="0.3", load="body", body_force="-1.0", typeBc="ker",
                preconditioner="Dirichlet_explicit")


HBM_PEAK_GBS = 8000.0
FP16_MFMA_PEAK_TFS = 2500.0  # dense FP16 matrix cores, published
SEED = 20240611
# HBM bytes per launch of the split-FP16 GEMM on the default workload, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over `bench.py --kernel-only`,
# each scaled by the counter's reading for kb_copy of known size (FETCH x2.000,
# WRITE x1.000).  Row form (k_tgemm_h16, K chunk depth 64): 7.249e9 fetched +
# 2.222e8 written.  Band form: k_tgemm_h16b 3.676e9 + 0.865e9, k_tgemm_h16b_reduce
# 0.827e9 + 0.189e9.  Valid for the default workload only.
GEMM_TRAFFIC_BYTES = {"rows": 7.47e9, "bands": 5.56e9}
GEMM_KERNEL_NAME = {
    "rows": "k_tgemm_h16 (deflated K+_bb, split-FP16 packed symmetric 64x64 tiles, 64 columns, tiles read per row block)",
    "bands": "k_tgemm_h16b + k_tgemm_h16b_reduce (deflated K+_bb, split-FP16 packed symmetric 64x64 tiles, 64 "
             "columns, every tile read once; time of both launches)"}


def cpu_port_run(cfg):
    """CPU baseline process (one per bench run): the C++ FP64 restatement of the
    reference's PARDISO path (oracle/cpu/htfeti_cpu.cpp: analytic kernel, fixing
    DOFs, band Cholesky of K_reg and K_rr, Dirichlet_implicit, implicit K_c^+ in
    every PCPG iteration as in Cluster.cpp:1126, OpenMP over subdomains) on the
    bench's own mesh and first load case.
      * all cores: set-up and one solve of the whole mesh, both measured;
      * one core: the same solve on 1 thread stopped after 3 PCPG iterations
        (initialisation, 3 iterations, recovery measured; the iterations scaled to
        the all-core run's count), and the 1-thread set-up of one cluster of the
        mesh scaled by the number of clusters."""
    from oracle.mesh import opts
    from oracle.cpu_port import CpuHTFETI
    from oracle.loads import traction_loads
    ncores = min(16, len(os.sched_getaffinity(0)))
    o = opts(**dict(cfg, preconditioner="Dirichlet_implicit"))
    t = time.perf_counter()
    c = CpuHTFETI(o, threads=ncores)
    prep_setup = time.perf_counter() - t
    f = traction_loads(c.mesh, c.lite, SEED, [0])[0]
    ra = c.solve(f, eps=1e-8, max_iter=1000, threads=ncores)
    r1 = c.solve(f, eps=1e-8, max_iter=3, threads=1)
    c.close()
    one = r1["times"]
    one_s = one["init"] + one["recover"] + one["pcpg"] * ra["iterations"] / 3.0
    cl = dict(cfg, Cx=1, Cy=1, Cz=1)
    ncl = int(cfg.get("Cx", 1)) * int(cfg.get("Cy", 1)) * int(cfg.get("Cz", 1))
    s1 = CpuHTFETI(opts(**dict(cl, preconditioner="Dirichlet_implicit")), threads=1)
    set1 = s1.setup_times
    s1.close()
    print(json.dumps({"ncores": ncores, "iterations": ra["iterations"], "residual": ra["residual"],
                      "all_setup": c.setup_times, "all_solve": ra["times"], "one_solve": one,
                      "one_solve_s": one_s, "one_setup_cluster": set1, "clusters": ncl,
                      "index_structures_s": prep_setup - c.setup_times["total"], "nsub": c.nsub}))


def run_cpu_baseline(cfg):
    """Runs cpu_port_run in a child process and reports it as `cpu_baseline`."""
    env = dict(os.environ, OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    env.pop("OMP_NUM_THREADS", None)
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-sample", "--cpu-config", json.dumps(cfg)]
    t0 = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=1200)
    if p.returncode != 0:
        return {"value": None, "unit": "solves/s", "cores": 1, "kind": "port",
                "sample": "failed: " + p.stderr[-300:]}
    s = json.loads(p.stdout.strip().splitlines()[-1])
    nsub = s["nsub"]
    out = {"value": 1.0 / s["one_solve_s"], "unit": "solves/s", "cores": 1, "kind": "port",
           "sample": ("C++ FP64 restatement of the reference's PARDISO path (oracle/cpu/htfeti_cpu.cpp; band "
                      "Cholesky in l2g order, no MKL on the host) on the bench mesh (%d subdomains) and its first "
                      "load case: one thread for initialisation, 3 PCPG iterations and recovery, measured; the "
                      "iteration time scaled to the %d iterations of the all-core solve" % (nsub, s["iterations"])),
           "solve_s": s["one_solve_s"], "solve_phases_s": s["one_solve"],
           "setup_s": s["one_setup_cluster"]["total"] * s["clusters"],
           "setup_sample": "1-thread set-up of one cluster (%d subdomains) x %d clusters: %.2f s per cluster"
                           % (nsub // max(1, s["clusters"]), s["clusters"], s["one_setup_cluster"]["total"]),
           "all_cores": {"value": 1.0 / s["all_solve"]["total"], "cores": s["ncores"],
                         "sample": "the whole bench mesh on %d OpenMP threads, one solve of its first load case, "
                                   "measured" % s["ncores"],
                         "solve_s": s["all_solve"]["total"], "solve_phases_s": s["all_solve"],
                         "iterations": s["iterations"], "setup_s": s["all_setup"]["total"],
                         "setup_phases_s": s["all_setup"]},
           "index_structures_s": s["index_structures_s"],
           "preconditioner": "Dirichlet_implicit (the reference default; same operator as the GPU's explicit S)"}
    out["host"] = {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0)), "cpu_model": cpu_model()}
    out["wall_s"] = round(time.time() - t0, 1)
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_fixture(name):
    """Oracle displacements of a full-size workload (tests/golden/<name>_u.npz: the
    oracle's sparse-LU PCPG at eps 1e-10, u at a seeded sample of DOFs), or None."""
    path = os.path.join(GOLDEN, name + "_u.npz")
    if not os.path.exists(path):
        return None
    with np.load(path) as f:
        fx = {k: f[k].copy() for k in f.files}
    fx["cases"] = json.loads(str(fx["cases"]))
    fx["options"] = json.loads(str(fx["options"]))
    return fx


def fixture_error(fx, u, case):
    """|u - u_oracle| / |u_oracle| on the fixture's sampled DOFs for load case `case`."""
    j = fx["cases"].index(str(case))
    ref = fx["u_sample"][j]
    return float(np.linalg.norm(u[fx["sample"]] - ref) / np.linalg.norm(ref))


def gemm_time(mesh, gemm_kernel, k):
    """Device time of one K+_bb GEMM launch (HIP events on the library stream) and its
    algorithmic HBM bytes: the mean over the GEMM launches inside one eager block
    solve of k load cases (events around each launch), the kernel as it runs in the
    solve; beside it the mean of 10 back-to-back launches of the GEMM alone."""
    b2b_ms, g_bytes = mesh.time_stream(7, reps=10)
    g_ms, _ = mesh.time_stream(9, reps=k)
    how = ("HIP events around each K+_bb GEMM launch (%s) of one eager %d-case block solve, mean over the "
           "launches that ran" % ("band kernel + reduction" if gemm_kernel == "bands" else "row form", k))
    return g_ms, g_bytes, b2b_ms, how


def roofline_of(mesh, gemm_kernel, k=64):
    """The split-FP16 tile GEMM over 64 columns with K+_bb (the dominant kernel):
    device time per launch and its algorithmic bytes."""
    g_ms, g_bytes, b2b_ms, how = gemm_time(mesh, gemm_kernel, k)
    _, g_flops = mesh.time_stream(2, reps=1)
    achieved = g_bytes / (g_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "kernel": GEMM_KERNEL_NAME[gemm_kernel].split(" (")[0], "kernel_ms": g_ms, "timing": how,
            "kernel_ms_back_to_back": b2b_ms, "algorithmic_bytes": g_bytes,
            "fp16_mfma_frac": 3.0 * g_flops / (g_ms * 1e-3) / 1e12 / FP16_MFMA_PEAK_TFS}


def block_workload(hfls_amd, cfg, k, warmup, steps, plain_steps, eps, F=None, fixture=None, fixture_cases=(),
                   gemm_kernel="bands"):
    """Set the mesh up, solve `warmup + steps` blocks of k load cases in mixed precision
    and `warmup + plain_steps` single cases in FP64; per-block device times, rates,
    iterations and, with a fixture, the error of the fixture's cases (columns
    `fixture_cases` of the first block, and the plain solves of the same loads)
    against the oracle."""
    t0 = time.perf_counter()
    mesh = hfls_amd.Mesh(device=0, **cfg)
    setup_s = time.perf_counter() - t0
    mesh.configure(gemm_kernel=gemm_kernel)
    nblk = warmup + steps
    if F is None:
        F = mesh.traction(SEED, 0, k * nblk)
    plain_ms, plain_it, err_plain = [], [], []
    for i in range(warmup + plain_steps):
        r = mesh.pcpg(eps=eps, max_iter=2000, f_packed=F[i], want_lambda=False)
        assert r["converged"], "plain PCPG did not converge"
        if i >= warmup:
            plain_ms.append(r["solve_ms"])
            plain_it.append(r["iterations"])
        if fixture is not None and i < len(fixture_cases):
            err_plain.append(fixture_error(fixture, r["u"], fixture_cases[i]))
    blk_ms, its, res, err_mixed = [], [], [], []
    for b in range(nblk):
        r = mesh.pcpg_block(F[b * k:(b + 1) * k], eps=eps, max_iter=2000, precision=32)
        assert np.all(r["status"] == 2), "a load case did not converge"
        if b == 0 and fixture is not None:
            err_mixed = [fixture_error(fixture, r["u"][j], c) for j, c in enumerate(fixture_cases)]
        if b >= warmup:
            blk_ms.append(r["solve_ms"])
            its.append(r["iterations"])
            res.append(float(r["residual"].max()))
    its = np.concatenate(its)
    rate = k * len(blk_ms) / (float(np.sum(blk_ms)) / 1e3)
    plain_rate = 1e3 / float(np.mean(plain_ms))
    out = {"solves_per_s": rate, "ms_per_step": float(np.mean(blk_ms)), "load_cases_per_block": k,
           "subdomains": mesh.nsub, "dual_rows": mesh.nlam, "global_dof": 3 * mesh.nPoints,
           "iterations_min_max": [int(its.min()), int(its.max())], "max_true_residual": max(res),
           "plain": {"solves_per_s": plain_rate, "ms_per_solve": float(np.mean(plain_ms)), "iterations": plain_it[-1]},
           "speedup_vs_plain": rate / plain_rate, "setup_s": round(setup_s, 3), "device_bytes": mesh.device_bytes,
           "roofline": roofline_of(mesh, gemm_kernel, k)}
    if fixture is not None:
        out["error_vs_fp64_oracle"] = {"cases": [str(c) for c in fixture_cases], "mixed": err_mixed,
                                       "plain_fp64": err_plain}
    mesh.close()
    return out


FAMILIES = {64: dict(Cx=2, Cy=2, Cz=2, Nx=2, Ny=2, Nz=2), 216: dict(Cx=3, Cy=3, Cz=3, Nx=2, Ny=2, Nz=2),
            512: dict(Cx=2, Cy=2, Cz=2, Nx=4, Ny=4, Nz=4), 1000: dict(Cx=5, Cy=5, Cz=5, Nx=2, Ny=2, Nz=2)}


def serve_mesh(nsub, e, seed):
    """A cube family member with the uneven-batching materials: lognormal E per
    subdomain (sigma 1.5, clipped to a 1e3 contrast), x=0 clamped."""
    o = dict(FAMILIES[nsub], nx=e, ny=e, nz=e, bc="clamp_x0", material="lognormal", young_modulus="2.1e5",
             mat_sigma="1.5", mat_contrast="1e3", seed=str(seed), poissons_ratio="0.3", load="body", typeBc="ker",
             preconditioner="Dirichlet_explicit")
    return o


def run_serving(hfls_amd, families, edges, nreq, rate, seed, eps, budget_gb, warm, host_cache_gb=0.0):
    """Request stream with Poisson arrivals: each request draws its mesh uniformly
    from families x edges and 1..64 load cases log-uniformly (seeded).  Meshes are
    set up on first use and kept on the device while they fit the HBM budget
    (least recently used idle meshes are evicted); with warm, one single-case
    request per mesh runs before the stream starts.  With host_cache_gb, evicted
    meshes are kept in that much pinned host memory and restored from it.  Latency
    is submission to the last case recovered, set-up time of cold meshes included."""
    meshes = [serve_mesh(f, e, 100 + i) for i, (f, e) in enumerate((f, e) for f in families for e in edges)]
    out = run_stream(hfls_amd, meshes, nreq, rate, seed, eps, budget_gb, warm, host_cache_gb)
    out["meshes"] = ["%dx%d^3" % (f, e) for f in families for e in edges]
    return out


def run_stream(hfls_amd, meshes, nreq, rate, seed, eps, budget_gb, warm, host_cache_gb=0.0):
    rng = np.random.default_rng(seed)
    srv = hfls_amd.Server(slots=64, precision=32, eps=eps, max_iter=1000, chunk=8, hbm_budget_gb=budget_gb,
                          host_cache_gb=host_cache_gb)
    t0 = time.perf_counter()
    if warm:
        for m in meshes:
            rid = srv.submit(m, k=1, seed=1, case0=0)
            while srv.step():
                pass
            srv.result(rid)
    warm_s = time.perf_counter() - t0
    st0 = srv.stats()
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    picks = rng.integers(0, len(meshes), nreq)
    ks = np.minimum(64, np.floor(np.exp(rng.uniform(0.0, np.log(65.0), nreq)))).astype(int)
    ids = []
    start = time.perf_counter()
    nxt = 0
    busy = False
    while nxt < nreq or busy:
        now = time.perf_counter() - start
        while nxt < nreq and arrivals[nxt] <= now:
            ids.append(srv.submit(meshes[picks[nxt]], k=int(ks[nxt]), seed=1000 + nxt, case0=0))
            nxt += 1
        busy = srv.step()
        if not busy and nxt < nreq:
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - start)))
    wall = time.perf_counter() - start
    lat, cases, iters, worst = [], 0, [], 0.0
    for rid in ids:
        r = srv.result(rid)
        assert np.all(r["status"] == 2), "a served load case did not converge"
        lat.append(r["latency_ms"])
        cases += len(r["iterations"])
        iters.extend(r["iterations"].tolist())
        worst = max(worst, float(r["residual"].max()))
    st = srv.stats()
    srv.close()
    lat = np.array(lat)
    return {"requests": nreq, "load_cases": cases, "arrival_rate_per_s": rate, "wall_s": wall,
            "solves_per_s": cases / wall, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()), "hbm_budget_gb": budget_gb,
            "cache": "warmed before the stream" if warm else "cold start (set-ups inside the latencies)",
            "cache_warm_s": warm_s, "setups_in_stream": int(st["setups"] - st0["setups"]),
            "setup_s_in_stream": (st["setup_ms"] - st0["setup_ms"]) / 1e3,
            "evictions_in_stream": int(st["evictions"] - st0["evictions"]),
            "host_cache": {"gb": host_cache_gb, "restores_in_stream": int(st["restores"] - st0["restores"]),
                           "restore_s_in_stream": (st["restore_ms"] - st0["restore_ms"]) / 1e3,
                           "saves": int(st["saves"]), "save_s": st["save_ms"] / 1e3,
                           "meshes_held_at_end": int(st["host_meshes"]), "bytes_held_at_end": st["host_bytes"],
                           "dropped": int(st["host_drops"])},
            "iterations_min_max": [int(min(iters)), int(max(iters))], "max_true_residual": worst,
            "slot_joins": int(st["joins"]), "rounds": int(st["steps"] - st0["steps"]),
            "host_ms": {"join": st["join_ms"] - st0["join_ms"], "sweeps": st["sweep_ms"] - st0["sweep_ms"],
                        "finish": st["finish_ms"] - st0["finish_ms"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed block solves (64 load cases each)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--k", type=int, default=64, help="load cases per block")
    ap.add_argument("--plain-steps", type=int, default=3, help="timed plain FP64 single-case solves")
    ap.add_argument("--elements", type=int, default=10, help="elements per subdomain edge (10: 512 x 10^3)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", action="store_true")
    ap.add_argument("--cpu-config", default="{}", help="mesh options of the CPU baseline process (internal)")
    ap.add_argument("--small", action="store_true", help="2x2x2 x 5^3 smoke-sized workload")
    ap.add_argument("--serve-requests", type=int, default=32, help="requests in the serving run (0: none)")
    ap.add_argument("--serve-rate", type=float, default=0.5, help="Poisson arrival rate of requests per second")
    ap.add_argument("--serve-families", default="64,216,512,1000", help="subdomain counts of the served cube families")
    ap.add_argument("--serve-edges", default="8,9,10,11,12", help="elements per subdomain edge of the served meshes")
    ap.add_argument("--serve-budget-gb", type=float, default=250.0, help="HBM budget of the server")
    ap.add_argument("--serve-warm", action="store_true", help="set every mesh up before the stream starts")
    ap.add_argument("--serve-host-cache-gb", type=float, default=128.0,
                    help="pinned host memory for evicted meshes of the serving run (0: set evicted meshes up again)")
    ap.add_argument("--serve-compare-no-cache", action="store_true",
                    help="also run the serving stream without the host cache")
    ap.add_argument("--configs", default="3,2",
                    help="extra workloads: 3 = 512 x 12^3 cube, 2 = 512 uneven subdomains ('' for none)")
    ap.add_argument("--config-steps", type=int, default=2, help="timed blocks of each extra workload")
    ap.add_argument("--serve-saturated", type=int, default=48,
                    help="requests of the saturated serving run on the cached 512 x 10^3 cube (0: none)")
    ap.add_argument("--no-f32", action="store_true", help="skip the FP32 tile format (block_gemm=f32) run")
    ap.add_argument("--gemm-kernel", choices=("bands", "rows"), default="bands",
                    help="form of the split-FP16 GEMM: every tile read once (bands) or once per row block (rows)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the iteration kernels directly instead of replaying captured HIP graphs")
    ap.add_argument("--kernel-only", action="store_true",
                    help="setup, then only the dominant kernel (FP32 tile GEMM, 4 launches) and a device copy of "
                         "known size (4 launches; calibration of HBM traffic counters)")
    a = ap.parse_args()
    if a.cpu_sample:
        cpu_port_run(json.loads(a.cpu_config))
        return 0
    if a.gpus != 1:
        print("bench.py runs on exactly one GPU: the library has no multi-GPU path.",
              file=sys.stderr)
        return 2
    import hfls_amd
    cfg = cube(a.elements)
    workload = "htfeti_cube_8cl_512sub_%de3_%dcases" % (a.elements, a.k)
    if a.small:
        cfg.update(Cx=1, Cy=1, Cz=1, Nx=2, Ny=2, Nz=2, nx=5, ny=5, nz=5)
        workload = "htfeti_cube_1cl_8sub_5e3_%dcases" % a.k
    t0 = time.perf_counter()
    mesh = hfls_amd.Mesh(device=0, **cfg)
    setup_s = time.perf_counter() - t0
    if a.no_graph:
        mesh.configure(hip_graph=0)
    mesh.configure(gemm_kernel=a.gemm_kernel)
    log("setup %.2f s (factor %.2f s), nlam %d, device bytes %.2f GB" %
        (setup_s, mesh.factor_ms / 1e3, mesh.nlam, mesh.device_bytes / 1e9))
    if a.kernel_only:
        g_ms, g_flops = mesh.time_stream(2, reps=3)
        c_ms, c_bytes = mesh.time_stream(4, reps=3)
        print(json.dumps({"kernel_ms": g_ms, "flops": g_flops, "copy_ms": c_ms, "copy_bytes_each_way": c_bytes}))
        mesh.close()
        return 0
    nblk = a.warmup + a.steps
    F = mesh.traction(SEED, 0, a.k * max(1, nblk))
    log("loads: %d seeded face tractions" % F.shape[0])

    # oracle displacements of the 512 x 10^3 cube (cases 0 and 1 are F[0], F[1])
    fx = load_fixture("cube512_10e3") if (a.elements == 10 and not a.small and a.k >= 2) else None
    err_plain, err_mixed = [], []

    # ---- plain FP64: one load case at a time
    plain_ms, plain_it, plain_res, plain_pg = [], [], [], []
    for i in range(a.warmup + a.plain_steps):
        r = mesh.pcpg(eps=a.eps, max_iter=1000, f_packed=F[i], want_lambda=False)
        assert r["converged"], "plain PCPG did not converge"
        if fx is not None and i < 2:
            err_plain.append(fixture_error(fx, r["u"], i))
        if i >= a.warmup:
            plain_ms.append(r["solve_ms"])
            plain_it.append(r["iterations"])
            plain_res.append(r["residual"])
            plain_pg.append(r["pg_rel"])
    log("plain fp64: %s ms, iterations %s" % (["%.1f" % x for x in plain_ms], plain_it))
    plain_rate = 1e3 / float(np.mean(plain_ms)) if plain_ms else None  # --plain-steps 0: not timed

    # ---- batched mixed precision: blocks of k load cases.  The loads of a block are
    # placed in a reused page-locked buffer before its timer starts and the
    # displacements land in another (hfls_host_alloc), as a serving loop keeps them
    Fp = hfls_amd.pinned_empty((a.k, mesh.npacked))
    up = hfls_amd.pinned_empty((a.k, 3 * mesh.nPoints))
    blk_ms, blk_wall, its, res, certs, sweeps = [], [], [], [], [], []
    for s in range(nblk):
        Fp[:] = F[s * a.k:(s + 1) * a.k]
        w0 = time.perf_counter()
        r = mesh.pcpg_block(Fp, eps=a.eps, max_iter=1000, precision=32, out=up)
        w1 = time.perf_counter()
        assert np.all(r["status"] == 2), "a load case did not converge"
        if s == 0 and fx is not None:
            err_mixed = [fixture_error(fx, r["u"][j], j) for j in range(2)]
        log("block %d: %.1f ms device (init %.1f, pcpg %.1f, recover %.1f), wall %.1f ms, iterations %d..%d, "
            "max true residual %.2e, FP64 refinement points %d" %
            (s, r["solve_ms"], r["init_ms"], r["pcpg_ms"], r["recover_ms"], (w1 - w0) * 1e3, r["iterations"].min(),
             r["iterations"].max(), r["residual"].max(), r["cert_steps"]))
        if s >= a.warmup:
            blk_ms.append(r["solve_ms"])
            blk_wall.append(w1 - w0)
            its.append(r["iterations"])
            res.append(r["residual"].max())
            certs.append(r["cert_steps"])
            sweeps.append(r["sweeps"])
    total_ms = float(np.sum(blk_ms))
    value = a.k * len(blk_ms) / (total_ms / 1e3)
    its = np.concatenate(its)
    # one more block through pageable numpy arrays (a slice of the loads, a fresh
    # result array): the copies then go through the runtime's staging buffers
    w0 = time.perf_counter()
    mesh.pcpg_block(F[a.warmup * a.k:(a.warmup + 1) * a.k], eps=a.eps, max_iter=1000, precision=32)
    pageable_wall = time.perf_counter() - w0
    # two blocks at once in two block contexts (own streams, shared mesh operators)
    conc = None
    if nblk >= 2 and not a.small:
        F2 = F[:2 * a.k]
        mesh.pcpg_blocks(F2, k=a.k, eps=a.eps, max_iter=1000, precision=32)  # graphs of the second context
        rc = mesh.pcpg_blocks(F2, k=a.k, eps=a.eps, max_iter=1000, precision=32)
        assert np.all(rc["status"] == 2), "a load case did not converge"
        conc = {"blocks": 2, "span_ms": rc["solve_ms"], "solves_per_s": 2 * a.k / (rc["solve_ms"] / 1e3)}
        log("2 concurrent blocks: span %.1f ms -> %.1f solves/s" % (rc["solve_ms"], conc["solves_per_s"]))
    # plain vs block on the same case (first timed case): displacement agreement
    r1 = mesh.pcpg(eps=a.eps, max_iter=1000, f_packed=F[a.warmup * a.k], want_lambda=False)
    rb = mesh.pcpg_block(F[a.warmup * a.k:a.warmup * a.k + 1], eps=a.eps, max_iter=1000, precision=32)
    agree = float(np.linalg.norm(rb["u"][0] - r1["u"]) / np.linalg.norm(r1["u"]))

    # roofline of the dominant kernel: the split-FP16 tile GEMM over 64 columns,
    # bound by HBM (its matrix-core time at the published FP16 peak is a fraction
    # of the time to stream the packed tiles once)
    g_ms, g_bytes, b2b_ms, g_how = gemm_time(mesh, a.gemm_kernel, a.k)
    _, g_flops = mesh.time_stream(2, reps=1)
    s_ms, s_bytes = mesh.time_stream(8, reps=10)
    # sustained matrix-core rates on register data with random operands (no memory
    # traffic): the ceiling the FP32 tile GEMM runs against on real data
    ceil = {}
    for name, which in (("f64_16x16x4", 5), ("f32_32x32x2", 6)):
        c_ms, c_fl = mesh.time_stream(which, reps=3)
        ceil[name] = c_fl / (c_ms * 1e-3) / 1e12
    tune = mesh.tuning()
    achieved = g_bytes / (g_ms * 1e-3) / 1e9
    lat = np.array(blk_ms)
    out = {
        "metric": "load-case solves/s at 1e-8 dual residual, batched mixed precision (64 cases/block) vs plain FP64",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": float(np.mean(blk_ms)),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16x2+f64",
        "data": "synthetic (structured cube mesh and seeded face tractions generated in-process)",
        "config": {"workload": workload, "subdomains": mesh.nsub, "dual_rows": mesh.nlam,
                   "global_dof": 3 * mesh.nPoints, "load_cases_per_block": a.k, "eps": a.eps,
                   "precision": "K+_bb, S_i as split FP16 (hi + lo, FP32 accumulation) on MFMA; factorisation, "
                                "kernel, coarse, corrections, dots, projector, certification FP64",
                   "iterations_min_max": [int(its.min()), int(its.max())], "max_true_residual": float(max(res)),
                   "fp64_refinement_points_per_block": float(np.mean(certs)), "sweeps_per_block": float(np.mean(sweeps)),
                   "parallelism": "single GPU", "setup_s": round(setup_s, 3)},
        "plain": (None if plain_rate is None else
                  {"value": plain_rate, "unit": "solves/s", "ms_per_solve": float(np.mean(plain_ms)),
                   "iterations": plain_it[-1], "dtype": "f64", "residual_max": float(max(plain_res)),
                   "projected_gradient_rel_max": float(max(plain_pg))}),
        "speedup_vs_plain": value / plain_rate if plain_rate else None,
        "concurrent_blocks": conc,
        "host_inclusive_solves_per_s": a.k * len(blk_wall) / float(np.sum(blk_wall)),
        "host_inclusive_pageable_solves_per_s": a.k / pageable_wall,
        "block_vs_plain_displacement_rel_diff": agree,
        "error_vs_fp64_oracle": (None if fx is None else
                                 {"reference": "oracle (tests/golden/cube512_10e3_u.npz: sparse-LU PCPG at eps 1e-10, "
                                               "16,384 sampled DOFs)",
                                  "cases": ["traction 0", "traction 1"], "mixed_block": err_mixed,
                                  "plain_fp64": err_plain, "max": max(err_mixed + err_plain)}),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": GEMM_TRAFFIC_BYTES[a.gemm_kernel] if (a.elements == 10 and not a.small) else None,
                     "kernel": GEMM_KERNEL_NAME[a.gemm_kernel],
                     "kernel_ms": g_ms, "timing": g_how, "kernel_ms_back_to_back": b2b_ms,
                     "algorithmic_bytes": g_bytes, "flops_per_launch": g_flops,
                     "algorithmic_TFs": g_flops / (g_ms * 1e-3) / 1e12,
                     "fp16_mfma_frac": 3.0 * g_flops / (g_ms * 1e-3) / 1e12 / FP16_MFMA_PEAK_TFS,
                     "S_gemm_ms": s_ms, "S_gemm_GBs": s_bytes / (s_ms * 1e-3) / 1e9,
                     "measured_mfma_ceiling_TFs": ceil,
                     "autotune": {"k_chunk": 32 * tune["row_tiles"],
                                  "candidate_ms": [tune["ms"][1], tune["ms"][2]], "band_ms": tune["ms"][3]}},
    }
    mesh.close()
    if not a.small and not a.no_f32:
        # the FP32 tile format (block_gemm=f32) of the same workload
        cf = dict(cfg, block_gemm="f32")
        r32 = block_workload(hfls_amd, cf, a.k, 1, 2, 1, a.eps, gemm_kernel=a.gemm_kernel)
        out["block_gemm_f32"] = {"solves_per_s": r32["solves_per_s"], "ms_per_step": r32["ms_per_step"],
                                 "iterations_min_max": r32["iterations_min_max"]}
        log("block_gemm=f32: %.1f solves/s" % r32["solves_per_s"])
    extra = {}
    for c in [x for x in a.configs.split(",") if x] if not a.small else []:
        if c == "3":  # 512 x 12^3, 64 face tractions per block
            extra["cfg3_512sub_12e3_64cases"] = block_workload(hfls_amd, cube(12), a.k, 1, a.config_steps, 1, a.eps,
                                                               gemm_kernel=a.gemm_kernel)
        elif c == "2":  # 512 uneven subdomains, lognormal E; column 0 the body force (oracle fixture)
            fu = load_fixture("uneven512")
            if fu is not None:
                cu = dict(fu["options"])  # the fixture's own options
            else:
                cu = dict(cube(10), slab_choices="6,8,10,12,14", seed=str(SEED), material="lognormal",
                          mat_sigma="1.5", mat_contrast="1e3")
                for key in ("nx", "ny", "nz"):
                    cu.pop(key)
            m2 = hfls_amd.Mesh(device=0, **cu)
            Fu = m2.traction(SEED, 0, a.k * (1 + a.config_steps))
            Fu[0] = m2.load_vector()
            m2.close()
            extra["cfg2_512sub_uneven_64cases"] = block_workload(
                hfls_amd, cu, a.k, 1, a.config_steps, 1, a.eps, F=Fu, fixture=fu,
                fixture_cases=("body",) if fu is not None else (), gemm_kernel=a.gemm_kernel)
        else:
            raise SystemExit("unknown extra config " + c)
        log("config", c, json.dumps(extra[list(extra)[-1]]))
    if extra:
        out["configs"] = extra
    if a.serve_saturated > 0 and not a.small:
        # the served form at saturation: requests for the cached 512 x 10^3 cube arrive
        # faster than they are served (Poisson, 200 per second), 1..64 cases each
        sat = run_stream(hfls_amd, [cfg], a.serve_saturated, 200.0, SEED + 1, a.eps, 250.0, True)
        sat["meshes"] = [workload]
        sat["speedup_vs_plain"] = sat["solves_per_s"] / plain_rate if plain_rate else None
        out["serving_saturated"] = sat
        log("saturated serving:", json.dumps(sat))
    if a.serve_requests > 0:
        fam = [int(x) for x in a.serve_families.split(",")]
        edg = [int(x) for x in a.serve_edges.split(",")]
        out["serving"] = run_serving(hfls_amd, fam, edg, a.serve_requests, a.serve_rate, SEED, a.eps,
                                     a.serve_budget_gb, a.serve_warm, a.serve_host_cache_gb)
        log("serving:", json.dumps(out["serving"]))
        if a.serve_compare_no_cache and a.serve_host_cache_gb > 0:
            out["serving_no_host_cache"] = run_serving(hfls_amd, fam, edg, a.serve_requests, a.serve_rate, SEED,
                                                       a.eps, a.serve_budget_gb, a.serve_warm, 0.0)
            log("serving without host cache:", json.dumps(out["serving_no_host_cache"]))
    if not a.no_cpu:
        out["cpu_baseline"] = run_cpu_baseline(cfg)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
