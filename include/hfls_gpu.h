/* C ABI of the MI355X HTFETI hot path (libhfls.so).
 *
 * Conventions (mirroring the reference's Matrix/Cluster layout):
 *   - every function returns 0 on success and a negative code on failure; the
 *     message is available from hfls_last_error() (thread local).  Nothing calls
 *     exit(), unlike the reference's PARDISO wrapper (Matrix.cpp:1174-1178).
 *   - options are the reference's options2 map (htfeti.cpp:12-57) passed as
 *     parallel key/value string arrays; unknown keys are rejected.
 *   - dual vectors (lambda, g, w) have nlam entries in the reference's Bf row
 *     order (Cluster.cpp:846-924); coarse vectors have 6*n_clusters entries.
 *   - "packed local" vectors concatenate per-subdomain vectors in the l2g order of
 *     Data.cpp:111-144 (sorted global DOFs), subdomains ascending.
 *   - all arrays are caller owned host memory; FP64.
 *   - a handle is not thread safe; one handle per host thread.
 */
#ifndef HFLS_GPU_H
#define HFLS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfls_problem_s* hfls_problem;  /* host-side mesh + gluing (integer work) */
typedef struct hfls_mesh_s* hfls_mesh;        /* device-resident set-up mesh */

typedef struct {
  int iterations;    /* number of F applications in the PCPG loop (Cluster.cpp:1422-1450) */
  int converged;     /* 1 if sqrt(g'Pz) < eps*sqrt(g0'Pz0) (Cluster.cpp:1425) */
  double residual;   /* sqrt(g'Pz)/sqrt(g0'Pz0) at exit */
  double solve_ms;   /* device time of the whole load-case solve */
  double pcpg_ms;    /* device time of the iteration loop */
  double pg_rel;     /* |P g| / |P g0| at exit (projected gradient, reported beside residual) */
} hfls_stats;

const char* hfls_last_error(void);

/* ---- host problem: Mesh::createMesh (Mesh.cpp:69-289), Data::buildMappingStruct
 *      (Data.cpp:111-144), Cluster::create_cluster_constraints (Cluster.cpp:682-769) */
int hfls_problem_create(int nopt, const char* const* keys, const char* const* vals, hfls_problem* out);
int hfls_problem_destroy(hfls_problem p);
/* sizes[0..7] = nsub, nlam, n_glue, nPoints, n_clusters, nsub_per_cluster, n_dirichlet, n_corner_dofs */
int hfls_problem_sizes(hfls_problem p, int64_t* sizes);
/* info[0..7] = n, n_r, n_s, n_bs, n_ss, cluster, n_bf_entries, E bits unused (0) */
int hfls_problem_sub_info(hfls_problem p, int d, int* info);
/* compressed Bf_d (Cluster.cpp:931-959): global row, local DOF, sign (+1/-1) */
int hfls_problem_sub_bf(hfls_problem p, int d, int* rows, int* dofs, int* signs);
/* j_col_cmpr (Cluster.cpp:756-765), s-set used by the kernel detection, l2g */
int hfls_problem_sub_bdofs(hfls_problem p, int d, int* bdofs);
int hfls_problem_sub_sset(hfls_problem p, int d, int* sset);
int hfls_problem_sub_l2g(hfls_problem p, int d, int64_t* l2g);
int hfls_problem_sub_young(hfls_problem p, int d, double* E);
/* legacy ASCII VTK of a displacement field u (3 * nPoints, global DOF order):
 * points, hex8 cells, displacements, decomposition and material id per cell;
 * file <folder>/box.vtk (iter < 0) or <folder>/box<iter>.vtk.  Replaces
 * Mesh::SaveVTK (Mesh.cpp:292-351) as called from Cluster::printVTK
 * (Cluster.cpp:1648-1678). */
int hfls_problem_write_vtk(hfls_problem p, const double* u, const char* folder, int iter);
/* dual rows: subdomain of the +1 entry, of the -1 entry (-1 for Dirichlet rows),
 * global DOF, multiplicity weight (Cluster.cpp:893-895,918) */
int hfls_problem_rows(hfls_problem p, int* sub_a, int* sub_b, int64_t* dof, double* weight);
int hfls_problem_dirichlet(hfls_problem p, int64_t* dofs);
int hfls_problem_corners(hfls_problem p, int64_t* dofs);

/* ---- device mesh: replaces Cluster::initialization's factorisation and coarse
 *      setup (Cluster.cpp:68-352) and the PARDISO seam (Matrix.cpp:1029-1287) */
int hfls_mesh_create(int device, int nopt, const char* const* keys, const char* const* vals, hfls_mesh* out);
int hfls_mesh_destroy(hfls_mesh m);
/* sizes[0..5] = nsub, nlam, ncol_G, packed_local_length, nPoints, device_bytes */
int hfls_mesh_sizes(hfls_mesh m, int64_t* sizes);
/* setup timings (ms, t holds 5): [0] whole setup, [1] factorisation and kernel
 * detection, [2] index structures, [3] scratch allocation and element assembly,
 * [4] Bc, cluster and coarse problems */
int hfls_mesh_setup_times(hfls_mesh m, double* t);
/* ---- mesh cache (no counterpart in the reference, which assembles and factorises
 *      on every run, Driver.cpp:20-38; its MatrixMarket dumps, Matrix.cpp:556-632,
 *      are never read back).  hfls_mesh_save writes the set-up state of a mesh (its
 *      options, the factors W = L_rr^-1, the K+_bb and S_i tiles, kernels, gluing and
 *      coarse inverses) to `path` (through `path`.part, renamed when complete);
 *      hfls_mesh_restore builds a mesh from such a file on `device` without the
 *      factorisation: the index structures are rebuilt from the saved options and
 *      every saved size and index is checked against them.  A restored mesh gives
 *      bit-identical results to the saved one; hfls_mesh_setup_times reports the
 *      original set-up and *restore_ms the load. */
int hfls_mesh_save(hfls_mesh m, const char* path);
int hfls_mesh_restore(int device, const char* path, hfls_mesh* out, double* restore_ms);
/* kernel R (packed local, n x 6 column major per subdomain), fixing DOFs (6 local
 * DOFs per subdomain), detected defect per subdomain, body-force loads (packed local) */
int hfls_mesh_kernel(hfls_mesh m, double* R, int* fix, int* defect);
/* the same VTK output for a set-up mesh (e.g. the u of hfls_solve) */
int hfls_mesh_write_vtk(hfls_mesh m, const double* u, const char* folder, int iter);
int hfls_mesh_load(hfls_mesh m, double* f);

/* runtime settings of a set-up mesh (no counterpart in the reference):
 * "hip_graph" "1"/"0": replay the PCPG iteration chunks from captured HIP graphs
 * (default on); "refine_tol": mixed-precision refinement factor (default 1e-4);
 * "poll_chunk": iterations between host polls of the stopping flags (default 8);
 * "gemm_rows": "1", "2" or "auto" (default): tile shape of the mixed-precision block
 * GEMM, overriding the autotuner (FP32 format: 1 or 2 output row tiles per workgroup;
 * split-FP16 format, the default, option block_gemm: K chunks of 32 or 64);
 * "gemm_kernel": "bands" (default) or "rows": the split-FP16 GEMM reads every stored
 * tile once (workgroups own bands of 4 row blocks, partial sums added in a fixed
 * order), or once per row block that needs it; the two forms round differently,
 * each is bitwise reproducible */
int hfls_mesh_configure(hfls_mesh m, const char* key, const char* value);
/* autotuner result (after the first block call): [0] tile shape chosen for the
 * mixed-precision block GEMM (1 or 2, as for "gemm_rows"), [1], [2] measured ms per
 * launch of candidates 1 and 2, [3] ms of the band form of the split-FP16 GEMM
 * (0 in the FP32 format); out holds 4 */
int hfls_mesh_tuning(hfls_mesh m, double* out);

/* operators: Cluster::mult_Ff (Cluster.cpp:1114-1188), Projection (1314-1329,
 * alpha as returned there), Preconditioning (1043-1112), mult_Kplus_f (967-1040) */
int hfls_apply_F(hfls_mesh m, const double* w, double* Fw);
int hfls_project(hfls_mesh m, const double* x, double* Px, double* alpha);
int hfls_precondition(hfls_mesh m, const double* x, double* Mx);
int hfls_kplus(hfls_mesh m, const double* f_local, double* x_local);
/* coarse operators: lambda = Gf alpha (Cluster::mult_Gf, Cluster.cpp:1275-1283),
 * e = Gf^T lambda (mult_GfT, 1285-1293), e = Rf^T f (mult_RfT, 1295-1312); alpha and
 * e have 6 entries per cluster (Rf_d = R_d H_d, H the kernel of the cluster's
 * Gc^T Gc, Cluster.cpp:1332-1352) */
int hfls_mult_Gf(hfls_mesh m, const double* alpha, double* lambda);
int hfls_mult_GfT(hfls_mesh m, const double* lambda, double* e);
int hfls_mult_RfT(hfls_mesh m, const double* f_local, double* e);
/* recovery of Cluster::printVTK (Cluster.cpp:1648-1678): x_local = K_c^+ (f -
 * Bf^T lambda) - Rf alpha per subdomain (packed local) and u (3*nPoints) its scatter
 * by l2g where the last subdomain holding a DOF writes it; alpha as returned by
 * hfls_project.  x_local or u may be NULL. */
int hfls_recover(hfls_mesh m, const double* f_local, const double* lambda, const double* alpha, double* x_local,
                 double* u);
/* Projection on device memory with no allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * result as hfls_project, bit for bit: Px = x - Gf (Gf^T Gf)^-1 Gf^T x and alpha =
 * -(Gf^T Gf)^-1 Gf^T x, with x_dev, Px_dev (nlam doubles) and alpha_dev (6 per
 * cluster, may be NULL) the caller's device memory.  Its partial sums (12 doubles
 * per subdomain) live in one file-scope device array per GPU of at most 196,608
 * doubles (1.5 MiB, held until the process ends): meshes of more than 16,384
 * subdomains are rejected with an error.  stream NULL: the library's stream,
 * synchronised before return; otherwise (a hipStream_t) the work is only enqueued
 * on that stream.  Calls that overlap on one GPU race on the shared array. */
int hfls_project_noalloc(hfls_mesh m, const double* x_dev, double* Px_dev, double* alpha_dev, void* stream);
/* Rf (Cluster.cpp:1332-1352): packed local, n_d x 6 column major per subdomain
 * (the 6 columns of the subdomain's cluster) */
int hfls_mesh_rf(hfls_mesh m, double* Rf);
/* the getters of Cluster.hpp:26-31: [0] nSubClst, [1] neqClst (sum of n_d),
 * [2] nLam_c, [3] nLam_f, [4] nRBM_f (6 per cluster), [5] nRBM_c (6 per subdomain) */
int hfls_mesh_counts(hfls_mesh m, int64_t* counts);

/* one load case: Cluster::pcpg (Cluster.cpp:1363-1458) + the recovery of
 * printVTK (Cluster.cpp:1648-1678).  f_local may be NULL (body force of the
 * options); u has 3*nPoints entries, lambda nlam (either may be NULL).  With the
 * option path2data set, the final displacements are written to <path2data>/box.vtk
 * and, with vtkWithinIter=true, those of every iteration to box<it>.vtk (one
 * eager iteration per host poll, as Cluster.cpp:1448-1449). */
int hfls_solve(hfls_mesh m, double eps, int max_iter, const double* f_local, double* u, double* lambda,
               hfls_stats* stats);

/* ---- blocks of load cases (no single counterpart in the reference: its
 *      Cluster::pcpg (Cluster.cpp:1363-1458) solves one load case; a block runs
 *      that iteration per column, each with its own scalars and stopping test).
 * Column-major blocks: load case j of f_local at j*packed_local_length, of u at
 * j*3*nPoints, of lambda at j*nlam; 1 <= k <= 64.
 * precision 64: local operators applied in FP64 (same method as hfls_solve);
 * precision 32: F_i and S_i applied in FP32 on the matrix cores with FP32
 * accumulation, everything else in FP64 (factorisation, kernel, coarse problem,
 * dual vectors, dot products, projector, initialisation, recovery), and every
 * column's stopping test is certified on its FP64 true residual F lambda - d
 * (a column that fails continues from the true residual). */
typedef struct {
  int iterations;    /* F applications of this column's PCPG loop */
  int status;        /* 2 converged (certified), 3 iteration limit, 6 stopped on a
                        non-finite residual (NaN/Inf in the column's loads or
                        iterates; the other columns are unaffected), 7 (server
                        only) the mesh's set-up failed */
  double residual;   /* FP64 true residual sqrt(g'PMPg)/sqrt(g0'PMPg0) at exit */
} hfls_case_stats;
typedef struct {
  double solve_ms, init_ms, pcpg_ms, recover_ms; /* device times of the block */
  int sweeps;        /* block iterations executed */
  int cert_steps;    /* FP64 true-residual evaluations inside the loop */
} hfls_block_stats;
int hfls_solve_block(hfls_mesh m, double eps, int max_iter, int k, int precision, const double* f_local, double* u,
                     double* lambda, hfls_case_stats* cases, hfls_block_stats* stats);
/* nblocks (1..8) blocks of k load cases solved concurrently on the mesh, one block
 * context (own buffers, scalars and HIP stream, shared mesh operators) per block:
 * f_local holds nblocks*k load cases, u and cases nblocks*k results; the block
 * stats report the device time span of all blocks (no counterpart in the
 * reference: a throughput form of hfls_solve_block). */
int hfls_solve_blocks(hfls_mesh m, double eps, int max_iter, int nblocks, int k, int precision, const double* f_local,
                      double* u, hfls_case_stats* cases, hfls_block_stats* stats);
/* page-locked host memory (hipHostMalloc) for the f_local and u arrays of the solve
 * entry points: their copies to and from the device then run as DMA at the link
 * rate instead of through the runtime's staging of pageable memory.  No counterpart
 * in the reference (its arrays never leave the host); for callers that reuse their
 * load and result buffers across solves.  Release with hfls_host_free. */
int hfls_host_alloc(int64_t bytes, void** out);
int hfls_host_free(void* p);
/* block operators on k dual vectors (column major, ld nlam): op 0 mult_Ff
 * (Cluster.cpp:1114-1188), 1 Preconditioning (1043-1112), 2 Projection
 * (1314-1329); precision 32 or 64 selects the F_i / S_i arithmetic */
int hfls_apply_block(hfls_mesh m, int op, int k, int precision, const double* x, double* y);
/* k seeded random smoothed tractions on the x = 1 face (load cases case0..case0+k-1)
 * as packed local loads, column major; the definition is in problem.cpp */
int hfls_problem_traction(hfls_problem p, uint64_t seed, int case0, int k, double* f_local);
int hfls_mesh_traction(hfls_mesh m, uint64_t seed, int case0, int k, double* f_local);

/* ---- request loop with continuous batching (no counterpart in the reference,
 *      whose Driver solves one mesh and one load case per run, Driver.cpp:20-38).
 * A server caches set-up meshes on the device (admitted by HBM footprint, idle
 * ones evicted least recently used), gives each mesh `slots` (<= 64) block
 * columns, starts queued load cases of any request in free slots between chunks
 * of `chunk` PCPG iterations, runs different meshes on their own streams, and
 * recovers and releases a slot as soon as its case stops.  A case's result does
 * not depend on the other cases in flight.  Driven by the caller: step() runs
 * one scheduling round. */
typedef struct hfls_server_s* hfls_server;
int hfls_server_create(int device, double hbm_budget_gb, int slots, int precision, double eps, int max_iter,
                       int chunk, hfls_server* out);
int hfls_server_destroy(hfls_server s);
/* packed local length, 3*nPoints and the estimated device footprint (bytes, the
 * admission figure) of a mesh given by its options */
int hfls_server_sizes(hfls_server s, int nopt, const char* const* keys, const char* const* vals, int64_t* npacked,
                      int64_t* nglob, int64_t* bytes);
/* k load cases (column major packed local loads, or NULL for k seeded face
 * tractions seed, case0..case0+k-1) on the mesh given by the options */
int hfls_server_submit(hfls_server s, int nopt, const char* const* keys, const char* const* vals, int k,
                       const double* f_local, uint64_t seed, int case0, int64_t* id);
/* one scheduling round (admission, joins, one chunk per busy mesh, recoveries);
 * *busy = 0 when nothing is queued or running */
int hfls_server_step(hfls_server s, int* busy);
/* 0 queued, 1 admitted, 2 done */
int hfls_server_state(hfls_server s, int64_t id, int* state);
/* displacements (k x 3*nPoints), per-case stats, latency_ms[0] = submit -> done,
 * latency_ms[1] = submit -> admitted; releases the request */
int hfls_server_result(hfls_server s, int64_t id, double* u, hfls_case_stats* cases, double* latency_ms);
/* [0] cached meshes, [1] device bytes held by them, [2] mesh setups, [3] evictions,
 * [4] cases done, [5] requests done, [6] slot joins, [7] summed setup ms,
 * [8] failed set-ups, [9] cancelled requests, [10] scheduling rounds, [11] [12]
 * [13] host wall ms spent in joins, sweeps (to the stream sync) and finishes
 * (out holds 14) */
int hfls_server_stats(hfls_server s, double* out);
/* drop a request (no counterpart in the reference): queued cases are removed,
 * running cases leave their slots at once without recovery; the id is unknown
 * afterwards.  A request whose mesh set-up failed ends with status 7 per case and
 * hfls_server_result returns the set-up error (after writing the statuses). */
int hfls_server_cancel(hfls_server s, int64_t id);
/* pinned host memory (GB) for evicted meshes (no counterpart in the reference;
 * 0 = off, the default): an evicted mesh's set-up state (as hfls_mesh_save writes
 * it) is copied to host memory on the set-up worker thread, least recently used
 * saved meshes are dropped when it is full, and a later request for the mesh
 * restores the state over the host link instead of setting the mesh up again
 * (bit-identical results either way) */
int hfls_server_host_cache(hfls_server s, double host_budget_gb);
/* [0] meshes held in host memory, [1] their bytes, [2] saves, [3] summed save ms
 * (worker thread), [4] restores, [5] summed restore ms (request for the mesh to
 * mesh cached, as [7] of hfls_server_stats for set-ups), [6] saved meshes dropped
 * to make room (out holds 7) */
int hfls_server_cache_stats(hfls_server s, double* out);

/* Measurement hook (no counterpart in the reference): average device time of
 * the dominant kernel over reps back-to-back launches timed with HIP events on
 * the library's stream.  which = 0 / 1: the packed-tile GEMV streaming K+_bb
 * (inside mult_Ff) / S_i (inside Preconditioning), *bytes = algorithmic bytes of
 * one launch; which = 2 / 3: the tile GEMM of the mixed-precision block solve
 * (split-FP16 or FP32 format, option block_gemm) over 64 columns with K+_bb / S_i,
 * *bytes = algorithmic flops of one launch; which = 7 / 8: the same GEMMs, *bytes =
 * algorithmic HBM bytes of one launch (packed tiles, right-hand sides and results
 * each moved once); which = 4:
 * a plain device copy (reference for traffic counters), *bytes = bytes read;
 * which = 5 / 6: the sustained FP64 / FP32 matrix-core rate on register data with
 * random operands (no memory traffic), *bytes = flops of one launch; which = 9: the
 * mixed-precision K+_bb GEMM (as configured: band form with its reduction, row form,
 * or FP32 tiles) inside one mixed-precision block solve of reps seeded face-traction
 * load cases, launched eagerly with events around every GEMM, *ms = mean over the
 * launches that ran (not skipped through their device gate), *bytes = algorithmic
 * HBM bytes of one launch. */
int hfls_time_stream(hfls_mesh m, int which, int reps, double* ms, double* bytes);

#ifdef __cplusplus
}
#endif
#endif /* HFLS_GPU_H */
