"""Benchmark: QPs/sec of the batched mixed-precision device IPM on one MI355X.

Workload (one "step"): solve a batch of 4096 uneven condensed-MPC QPs
(T ~ U{8..64}, nx ~ U{4..16}, nu ~ U{1..6}, instance i from seed 1000+i) to
qtqp's default tolerances (atol 1e-7, rtol 1e-8), inputs resident in HBM,
device-side equilibration included.  Printed as one JSON line with the
roofline of the dominant kernel, the plain FP64 GPU path (one QP per call,
FP64 factor and solves, same kernels, no batching, no graph) on a sample,
and the CPU baseline (the FP64 oracle, a process pool on this host's cores)
on a bounded sample of the same workload.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--count 4096]
"""

from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "QPs/sec at 1e-8 rel KKT residual: mixed-prec batched vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(*a):
  print("[bench]", *a, file=sys.stderr, flush=True)


def _cpu_worker(args):
  seed_lo, seed_hi = args
  from qtqp_amd import problems
  from oracle import qtqp_oracle
  qps = [problems.uneven_mpc_batch(1, base_seed=s)[0] for s in range(seed_lo, seed_hi)]
  t0 = time.perf_counter()
  st = [qtqp_oracle.solve_qp(q.a, q.b, q.c, q.z, q.p)["status"] for q in qps]
  return time.perf_counter() - t0, len(qps), st


_BLAS_VARS = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")


def cpu_baseline(base_seed: int, budget_s: float = 20.0):
  """FP64 oracle (dense Gram/Cholesky IPM) over a pool of this host's cores.

  One process per core with BLAS threads = 1; each worker solves a disjoint
  slice of the same uneven-MPC family.  QPs/s = QPs / slowest worker time.
  """
  try:
    cores = len(os.sched_getaffinity(0))
  except AttributeError:
    cores = os.cpu_count() or 1
  cores = max(1, min(cores, 16))
  saved = {k: os.environ.get(k) for k in _BLAS_VARS}
  for k in _BLAS_VARS:
    os.environ[k] = "1"
  try:
    ctx = mp.get_context("spawn")
    with ctx.Pool(cores) as pool:
      # calibrate: one small slice per worker
      t0 = time.perf_counter()
      cal = pool.map(_cpu_worker, [(base_seed + 100 + 2 * w, base_seed + 102 + 2 * w)
                                   for w in range(cores)])
      per_qp = max(r[0] for r in cal) / 2
      log(f"cpu baseline: {cores} workers, ~{per_qp * 1e3:.0f} ms/QP, "
          f"calibrated in {time.perf_counter() - t0:.1f}s")
      per_worker = int(max(1, min(256, budget_s / max(per_qp, 1e-4))))
      jobs = [(base_seed + 1000 + w * per_worker, base_seed + 1000 + (w + 1) * per_worker)
              for w in range(cores)]
      res = []
      t0 = time.perf_counter()
      for r in pool.imap_unordered(_cpu_worker, jobs):
        res.append(r)
        log(f"cpu baseline: worker done ({len(res)}/{cores}) in {r[0]:.1f}s")
      wall = time.perf_counter() - t0
  finally:
    for k, v in saved.items():
      if v is None:
        os.environ.pop(k, None)
      else:
        os.environ[k] = v
  solved = sum(r[1] for r in res)
  solve_time = max(r[0] for r in res)
  return {"value": solved / solve_time, "unit": "QPs/s", "cores": cores, "kind": "port",
          "sample": f"{solved} QPs of the same uneven-MPC family (seeds {base_seed + 1000}..), "
                    f"FP64 oracle IPM + dense Gram/Cholesky (scipy LAPACK), one process per core, "
                    f"BLAS threads 1; pool wall {wall:.1f}s",
          "cpu_model": _cpu_model()}


def _cpu_model():
  try:
    for line in open("/proc/cpuinfo"):
      if line.startswith("model name"):
        return line.split(":", 1)[1].strip()
  except OSError:
    pass
  return "unknown"


VEC_PASSES_PER_IR = 4
VEC_PASSES_PER_ITER = 25


def algorithmic_work(n, m, counts, iters, a_bytes=None, m_full=None):
  """Algorithmic FLOPs and HBM bytes of the executed work of every QP.

  Per QP, from the engine's own counters (m is the rows of A the engine
  stores: a row that is the exact negation of another is stored once and
  serves both): each factorisation costs
  m n^2 + n^3/3 flops (Gram product + LDL') and reads A once (FP32 or
  FP64); each refinement pass (the rest of a correction, the FP64 residual
  and the next right-hand side, fused) reads A and P once in FP64; each
  factor solve reads the triangular factor twice (n^2/2 entries each way)
  in the factor's precision; each IPM iteration's termination test reads A
  and P once in FP64.  A read of A costs a_bytes (FP64; half that in FP32)
  when given -- the bytes inside the stored rows' nonzero column extents,
  which is what the engine's passes read -- else 8 m n.  Pointwise work is
  counted in vector passes of 8 (n + 2 m_full) bytes (m_full: all rows):
  four per refinement pass (its row inputs and outputs, P-pass vectors and
  the closing vector update) and VEC_PASSES_PER_ITER per IPM iteration (the
  iterate's normalisation, regularised diagonal, right-hand sides, step
  lengths, update and termination vectors).
  """
  f32, f64, kmv, fs32, fs64 = (counts[:, k].astype(np.float64) for k in range(5))
  it = iters.astype(np.float64)
  mn, nn = m.astype(np.float64) * n, n.astype(np.float64) ** 2
  vec = 8.0 * (n.astype(np.float64) + 2.0 * m_full) if m_full is not None else 0.0
  if a_bytes is not None:
    mn = np.asarray(a_bytes, dtype=np.float64) / 8.0
  fac_flops = m.astype(np.float64) * nn + nn * n / 3.0
  flops32 = float(np.sum(f32 * fac_flops))
  flops64 = float(np.sum(f64 * fac_flops))
  bytes_ = np.sum(f32 * 4 * mn + f64 * 8 * mn + (kmv + it) * (8 * mn + 8 * nn)
                  + fs32 * 4 * nn + fs64 * 8 * nn
                  + (kmv * VEC_PASSES_PER_IR + it * VEC_PASSES_PER_ITER) * vec)
  return flops32, flops64, float(bytes_)


def kkt_rel_residual(x, y, s, a, b, c, p, z):
  """max(pres/prel, dres/drel, gap/min|cost|) of a solution (tests/test_qtqp.py:283-309 form)."""
  pres = np.linalg.norm(a @ x + s - b, np.inf)
  dres = np.linalg.norm(p @ x + a.T @ y + c, np.inf)
  gap = abs(c @ x + b @ y + x @ p @ x)
  prel = max(np.linalg.norm(a @ x, np.inf), np.linalg.norm(s, np.inf), np.linalg.norm(b, np.inf))
  drel = max(np.linalg.norm(p @ x, np.inf), np.linalg.norm(a.T @ y, np.inf), np.linalg.norm(c, np.inf))
  cost = min(abs(c @ x + 0.5 * x @ p @ x), abs(b @ y + 0.5 * x @ p @ x))
  return max(pres / prel, dres / drel, gap / max(cost, 1e-300))


def warm_start_leg(n_train, n_eval, steps, k_star=7):
  """Learned warm start on the parametric MPC family (T=32, nx=8, nu=3).

  Cold and warm solves run on one resident family batch (A and P uploaded
  once; each solve uploads b and c): cold from the default start, warm from
  the network's proposal decoded on the device.  The warm wall time includes
  the inference.  Parity: a sample of the warm-started QPs is re-solved by
  the FP64 oracle from the same proposed start (same status, x/y/s error),
  and compared with the oracle's default-start solution.
  """
  from qtqp_amd import problems, warmstart
  from oracle import qtqp_oracle
  fam = warmstart.Family()
  x_tr = problems.parametric_mpc_x0(n_train, seed=11)
  x_ev = problems.parametric_mpc_x0(n_eval, seed=22)
  solver = warmstart.FamilySolver(fam, chunk=n_eval)
  t0 = time.perf_counter()
  solver.solve(x_ev, max_iter=1)  # creates the resident batch
  t_create = time.perf_counter() - t0
  t_cold = []
  for _ in range(3):
    t0 = time.perf_counter()
    cold = solver.solve(x_ev)
    t_cold.append(time.perf_counter() - t0)
  log(f"warm start: cold solve of {n_eval} QPs: {cold['iters'].mean():.2f} iterations mean, "
      f"{min(t_cold):.3f}s")
  ws = warmstart.WarmStart(fam, k_star=k_star)
  t0 = time.perf_counter()
  t, _ = ws.targets(x_tr)
  t_data = time.perf_counter() - t0
  log(f"warm start: {n_train} solver targets in {t_data:.1f}s; training")
  t0 = time.perf_counter()
  losses = ws.fit(x_tr, t, steps=steps)
  t_train = time.perf_counter() - t0
  cap = int(np.percentile(cold["iters"], 99))
  t_warm = []
  for _ in range(3):
    t0 = time.perf_counter()
    warm = warmstart.solve_with_warm_start(ws, x_ev, cap=cap, solver=solver)
    t_warm.append(time.perf_counter() - t0)
  log(f"warm start: warm solve {warm['iters'].mean():.2f} iterations mean, {min(t_warm):.3f}s")
  solver.free()
  idx = np.linspace(0, n_eval - 1, 32).astype(int)
  qps = fam.qps(x_ev[idx])
  px, py, ps, pt, _, valid = ws.propose(x_ev[idx])
  names = {0: "unfinished", 1: "solved", 2: "infeasible", 3: "unbounded"}
  res, d_same, d_default = [], {k: [] for k in "xys"}, {k: [] for k in "xys"}
  ok_same = ok_default = True
  rel = lambda u, v: float(np.max(np.abs(u - v)) / max(1.0, np.max(np.abs(v))))
  for j, (i, (a, b, c, z, p)) in enumerate(zip(idx, qps)):
    got = names.get(int(warm["status"][i]), "other")
    ref = qtqp_oracle.solve_qp(a, b, c, z, p)
    ok_default &= ref["status"] == got
    same = None
    if valid[j] and not warm["fallback"][i]:
      same = qtqp_oracle.solve_qp(a, b, c, z, p, max_iter=cap,
                                  start=(px[j], py[j], ps[j], pt[j]))
      ok_same &= same["status"] == got
    if ref["status"] != "solved":
      continue
    res.append(kkt_rel_residual(warm["x"][i], warm["y"][i], warm["s"][i], a, b, c, p, z))
    for k in "xys":
      d_default[k].append(rel(warm[k][i], ref[k]))
      if same is not None:
        d_same[k].append(rel(warm[k][i], same[k]))
  ws.mlp.free()
  return {"family": "condensed MPC T=32 nx=8 nu=3, x0 ~ N(0, 0.5^2 I)", "k_star": k_star,
          "train_qps": n_train, "eval_qps": n_eval, "mlp": "8-256-256-1506 GELU, BF16 MFMA, Adam",
          "train_steps": steps, "loss_first_last": [float(losses[:10].mean()), float(losses[-50:].mean())],
          "data_gen_s": t_data, "train_s": t_train,
          "cold_iters_mean": float(cold["iters"].mean()), "cold_iters_p99": float(np.percentile(cold["iters"], 99)),
          "warm_iters_mean": float(warm["iters"].mean()), "warm_iters_p99": float(np.percentile(warm["iters"], 99)),
          "iteration_cut": float(1.0 - warm["iters"].mean() / cold["iters"].mean()),
          "statuses_identical": bool(np.array_equal(warm["status"], cold["status"])),
          "fallback_rate": float(warm["fallback"].mean()),
          "resident_batch": f"one family batch of {n_eval} QPs, A and P uploaded once "
                            f"(created in {t_create:.3f}s, outside the timings); each solve "
                            "uploads b and c and downloads the solutions",
          "cold_qps_per_s": n_eval / min(t_cold), "warm_qps_per_s_incl_inference": n_eval / min(t_warm),
          "warm_over_cold_qps": min(t_cold) / min(t_warm),
          "cold_device_ms": float(cold["device_ms"]), "warm_device_ms": float(warm["device_ms"]),
          "device_time_cut": float(1.0 - warm["device_ms"] / cold["device_ms"]),
          "oracle_sample": len(qps),
          "oracle_same_start": {"status_match": bool(ok_same),
                                "max_rel_err": {k: float(max(v)) if v else None for k, v in d_same.items()}},
          "oracle_default_start": {"status_match": bool(ok_default),
                                   "max_rel_err": {k: float(max(v)) for k, v in d_default.items()}},
          "max_rel_kkt_residual": float(max(res))}


def single_instance_leg(name, qp, oracle_ref=None):
  """One QP through the drop-in QTQP.solve, mixed vs plain FP64 HIP backends (the
  device-resident iteration loop; the mixed solve is repeated with the host loop).

  Reports per-iteration factor time (HIP events around the factorisation
  kernels), its MFMA utilisation (executed Gram-route flops m n^2 + n^3/3
  per factorisation over the FP32 or FP64 dense MFMA peak), refinement step
  counts, and the x/y/s error against the planted optimum and/or the FP64
  oracle.
  """
  import types
  import scipy.sparse as sp
  import qtqp_amd
  from qtqp_amd import direct
  n, m = qp.n, qp.m
  flops = m * n * n + n ** 3 / 3
  out = {"n": n, "m": m, "z": qp.z, "N": n + m, "factor_flops": flops}
  a_csc, p_csc = sp.csc_matrix(qp.a), sp.csc_matrix(qp.p)
  for prec, cls, peak in (("mixed", direct.HipDenseSolver, FP32_PEAK_TFLOPS),
                          ("fp64", direct.HipDenseFp64Solver, FP64_PEAK_TFLOPS)):
    infos = []

    class Probe(cls):  # records the engine counters before a handle is released
      def device_ipm(self, *args, **kw):
        ipm = super().device_ipm(*args, **kw)
        release = ipm.free

        def free():
          if ipm._h is not None:
            infos.append(ipm.info())
          release()
        ipm.free = free
        return ipm

      def free(self):
        if getattr(self, "_h", None) is not None:
          infos.append(self.info())
        super().free()

    sel = types.SimpleNamespace(name=f"HIP_DENSE[{prec}]", value=Probe)
    log(f"{name}: {prec} solve (n={n}, m={m})")
    prob = qtqp_amd.QTQP(a=a_csc, b=qp.b, c=qp.c, z=qp.z, p=p_csc)
    t0 = time.perf_counter()
    sol = prob.solve(linear_solver=sel, verbose=False, collect_stats=True)
    wall = time.perf_counter() - t0
    host_wall = None
    if prec == "mixed":
      # The same solve with the iteration loop on the host (the reference's
      # call pattern over the LinearSolver protocol), for comparison.
      Probe.device_driver = False
      log(f"{name}: {prec} solve, host iteration loop")
      prob = qtqp_amd.QTQP(a=a_csc, b=qp.b, c=qp.c, z=qp.z, p=p_csc)
      t1 = time.perf_counter()
      hsol = prob.solve(linear_solver=sel, verbose=False)
      host_wall = {"wall_s": time.perf_counter() - t1, "status": hsol.status.value,
                   "max_rel_diff_xys_vs_device_loop": max(
                       float(np.max(np.abs(g - w)) / max(1.0, np.max(np.abs(w))))
                       for g, w in zip((hsol.x, hsol.y, hsol.s), (sol.x, sol.y, sol.s)))}
      infos.pop()
    inf = infos[0] if prec == "mixed" else infos[-1]
    nf = max(1, inf["factorizations"])
    f_ms = inf["factor_ms"] / nf
    ir = [st.get("solves", 0) for it in sol.stats if isinstance(it, dict)
          for k, st in it.items() if k.endswith("lin_sys_stats")]
    r = {"status": sol.status.value, "iterations": len(sol.stats), "wall_s": wall,
         "factor_ms_per_iter": f_ms, "factor_tflops": flops / (f_ms / 1e3) / 1e12,
         "factor_mfma_frac": flops / (f_ms / 1e3) / 1e12 / peak,
         "peak_tflops": peak, "solve_ms_total": inf["solve_ms"],
         "refinement_solves_per_system": {"mean": float(np.mean(ir)) if ir else None,
                                          "max": int(max(ir)) if ir else None},
         "backend_solves": inf["backend_solves"], "fp32_inner_steps": inf["inner_steps"],
         "fp64_fallbacks": inf["fp64_fallbacks"]}
    if host_wall is not None:
      r["host_iteration_loop"] = host_wall
    for tag, ref in (("planted", (qp.x_star, qp.y_star, qp.s_star)),
                     ("oracle", None if oracle_ref is None else
                      (oracle_ref["x"], oracle_ref["y"], oracle_ref["s"]))):
      if ref is None or ref[0] is None:
        continue
      r[f"max_rel_err_xys_vs_{tag}"] = max(
          float(np.max(np.abs(g - w)) / max(1.0, np.max(np.abs(w))))
          for g, w in zip((sol.x, sol.y, sol.s), ref))
    r["rel_kkt_residual"] = kkt_rel_residual(sol.x, sol.y, sol.s, qp.a, qp.b, qp.c, qp.p, qp.z)
    out[prec] = r
  out["factor_speedup_mixed_vs_fp64"] = (out["fp64"]["factor_ms_per_iter"] /
                                         out["mixed"]["factor_ms_per_iter"])
  return out


def timed_oracle(qtqp_oracle, qp, threads=None):
  """The FP64 oracle (the reference algorithm on scipy LAPACK) on one QP: (result, seconds).

  threads=None uses every BLAS thread the host offers; 1 pins BLAS to one thread.
  """
  from threadpoolctl import threadpool_limits
  with threadpool_limits(limits=threads):
    t0 = time.perf_counter()
    ref = qtqp_oracle.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p)
    return ref, time.perf_counter() - t0


def single_legs(problems, qtqp_oracle):
  """The single-instance configurations: the n=200 dense QP (CPU reference case),
  the n=4000 dense QP and the T=64 condensed MPC QP, each with the FP64 oracle's
  CPU time (all BLAS threads and one) and the GPU drop-in backends."""
  try:
    cores = len(os.sched_getaffinity(0))
  except AttributeError:
    cores = os.cpu_count()
  out = {}
  # first use of the GPU libraries (code-object loading) outside the timings
  import scipy.sparse as sp
  import qtqp_amd
  w = problems.random_dense_qp(40, 60, 10, seed=1)
  for ls in (qtqp_amd.LinearSolver.HIP_DENSE, qtqp_amd.LinearSolver.HIP_DENSE_FP64):
    qtqp_amd.QTQP(a=sp.csc_matrix(w.a), b=w.b, c=w.c, z=w.z, p=sp.csc_matrix(w.p)).solve(
        linear_solver=ls, verbose=False)
  for key, label, qp in (
      ("single_small_dense", "n=200 m=300 z=50 dense QP (seed 0)",
       problems.random_dense_qp(200, 300, 50, seed=0)),
      ("single_large_dense", "n=4000 m=6000 z=1000 dense QP (seed 0)",
       problems.random_dense_qp(4000, 6000, 1000, seed=0)),
      ("single_mpc", "condensed MPC T=64 nx=12 nu=4 (seed 0)", problems.mpc_qp(64, 12, 4, seed=0))):
    log(f"{label}: FP64 oracle on the CPU")
    ref, t_all = timed_oracle(qtqp_oracle, qp)
    cpu = {"status": ref["status"], "iterations": ref["iters"], "seconds_all_threads": t_all,
           "threads": cores}
    if qp.n <= 1000:
      _, t_one = timed_oracle(qtqp_oracle, qp, threads=1)
      cpu["seconds_one_thread"] = t_one
    leg = single_instance_leg(label, qp, ref)
    leg["cpu_oracle"] = cpu
    out[key] = leg
  return out


def sparse_leg(m=6000, n=4000, z=1000, density=0.002, seed=0, reps=10, cpu=True, qtqp=True):
  """The sparse KKT backend (HIP_SPARSE, supernodal LDL') on the reference's sparse
  QP family (tests/test_qtqp.py:211-235) at n=4000, m=6000, z=1000.

  Times, with HIP events on the backend's stream, the numeric factorisation and
  one solve of the KKT matrix K_reg of a mid-solve IPM iterate (mu = 0.3, s/y
  drawn in [0.005, 200]), then one whole QTQP.solve with HIP_SPARSE.  The CPU
  comparison is scipy.sparse.linalg.factorized on the full symmetric K_reg, which
  is what the reference's SCIPY backend calls (solvers_sparse.py:96-121)."""
  import scipy.sparse as sp
  import scipy.sparse.linalg as spla
  import qtqp_amd
  from qtqp_amd import direct, problems
  qp = problems.random_sparse_qp(m, n, z, seed=seed, density=density)
  kkt = sp.bmat([[sp.triu(qp.p, format="csc") + sp.eye(n, format="csc"), qp.a.T],
                 [None, sp.eye(m, format="csc")]], format="csc")
  rng = np.random.default_rng(seed + 1)
  mu = 0.3
  s_over_y = rng.uniform(0.1, 2.0, m) / rng.uniform(0.01, 2.0, m)
  diag = np.concatenate([np.asarray(qp.p.diagonal()) + mu, np.full(z, mu), s_over_y[z:] + mu])
  diag[n:] *= -1.0
  log(f"sparse KKT: analysis of N={n + m}")
  t0 = time.perf_counter()
  be = direct.HipSparseSolver()
  be.set_kkt(kkt)
  create_s = time.perf_counter() - t0
  be.update_diag(diag)
  be.factorize()  # first use (code objects) outside the timings
  rhs = rng.normal(size=n + m)
  sol = be.solve(rhs)
  info0 = be.info()
  for _ in range(reps):
    be.factorize()
  for _ in range(reps):
    sol = be.solve(rhs)
  info = be.info()
  fac_ms = (info["factor_ms"] - info0["factor_ms"]) / reps
  sol_ms = (info["solve_ms"] - info0["solve_ms"]) / reps
  u = sp.triu(kkt, k=1)
  full = sp.csc_matrix(u + u.T + sp.diags(diag))
  res = float(np.max(np.abs(full @ sol - rhs)) / np.max(np.abs(rhs)))
  be.free()
  flops = info["factor_flops"]
  # L's algorithmic bytes for one solve: the factor read twice (forward, backward)
  solve_bytes = 2.0 * 8.0 * info["nnz_l"]
  out = {"workload": f"reference sparse QP family, n={n} m={m} z={z}, A density {density}, "
                     f"seed {seed}",
         "N": n + m, "nnz_kkt_upper": info["nnz_kkt"], "nnz_l": info["nnz_l"],
         "supernodes": info["supernodes"], "levels": info["levels"],
         "update_tiles": info["update_tiles"], "analysis_ms": info["analyze_ms"],
         "create_s": create_s,
         "factor_ms": fac_ms, "factor_flops": flops,
         "factor_tflops": flops / (fac_ms / 1e3) / 1e12,
         "factor_frac_of_fp64_peak": flops / (fac_ms / 1e3) / 1e12 / FP64_PEAK_TFLOPS,
         "solve_ms": sol_ms, "solve_gbs": solve_bytes / (sol_ms / 1e3) / 1e9,
         "solve_frac_of_hbm_peak": solve_bytes / (sol_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
         "solve_rel_residual": res}
  if qtqp:
    log("sparse KKT: QTQP.solve with HIP_SPARSE")
    t0 = time.perf_counter()
    qs = qtqp_amd.QTQP(a=qp.a, b=qp.b, c=qp.c, z=z, p=qp.p).solve(
        linear_solver=qtqp_amd.LinearSolver.HIP_SPARSE, verbose=False, collect_stats=True)
    out["qtqp_solve"] = {"status": qs.status.value, "iterations": len(qs.stats),
                         "wall_s": time.perf_counter() - t0,
                         "rel_kkt_residual": kkt_rel_residual(qs.x, qs.y, qs.s, qp.a, qp.b, qp.c,
                                                              qp.p, z)}
  if cpu:
    log("sparse KKT: scipy.sparse.linalg.factorized on the CPU")
    t0 = time.perf_counter()
    f = spla.factorized(full)
    t_fac = time.perf_counter() - t0
    t0 = time.perf_counter()
    xc = f(rhs)
    t_sol = time.perf_counter() - t0
    out["cpu_scipy_factorized"] = {
        "factor_ms": t_fac * 1e3, "solve_ms": t_sol * 1e3, "cores": 1,
        "max_rel_diff_vs_gpu": float(np.max(np.abs(xc - sol)) / np.max(np.abs(xc))),
        "factor_speedup_gpu": t_fac * 1e3 / fac_ms}
    # Like for like: SuperLU in symmetric mode with a minimum-degree ordering of
    # A' + A (diagonal pivots), its ordering and numeric phases timed apart: the
    # numeric phase alone is a second factorisation of the matrix permuted by
    # the first one's column order with the natural order.
    log("sparse KKT: SuperLU, MMD_AT_PLUS_A, symmetric mode, on the CPU")
    opts = dict(SymmetricMode=True)
    t0 = time.perf_counter()
    lu = spla.splu(full, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0, options=opts)
    t_all = time.perf_counter() - t0
    ipc = np.argsort(lu.perm_c)  # perm_c[i]: the position of column i
    permuted = sp.csc_matrix(full[ipc][:, ipc])
    t0 = time.perf_counter()
    lu2 = spla.splu(permuted, permc_spec="NATURAL", diag_pivot_thresh=0.0, options=opts)
    t_num = time.perf_counter() - t0
    t0 = time.perf_counter()
    xm = lu.solve(rhs)
    t_sol2 = time.perf_counter() - t0
    out["cpu_superlu_mmd_symmetric"] = {
        "ordering_and_factor_ms": t_all * 1e3, "numeric_factor_ms": t_num * 1e3,
        "ordering_ms": max(0.0, (t_all - t_num) * 1e3), "solve_ms": t_sol2 * 1e3, "cores": 1,
        "nnz_l_plus_u": int(lu.L.nnz + lu.U.nnz), "nnz_l_plus_u_numeric_run": int(lu2.L.nnz + lu2.U.nnz),
        "max_rel_diff_vs_gpu": float(np.max(np.abs(xm - sol)) / np.max(np.abs(xm))),
        "numeric_factor_speedup_gpu": t_num * 1e3 / fac_ms,
        "gpu_analysis_ms": info["analyze_ms"], "gpu_factor_ms": fac_ms}
  return out


def dropin_fp64(qps, reps=1):
  """The drop-in path: qtqp_amd.QTQP.solve with the HIP_DENSE_FP64 backend, one QP per
  call (the device-resident iteration loop)."""
  import scipy.sparse as sp
  import qtqp_amd
  q0 = qps[0]
  qtqp_amd.QTQP(a=sp.csc_matrix(q0.a), b=q0.b, c=q0.c, z=q0.z, p=sp.csc_matrix(q0.p)).solve(
      linear_solver=qtqp_amd.LinearSolver.HIP_DENSE_FP64, verbose=False)
  t0 = time.perf_counter()
  st = []
  for q in qps:
    sol = qtqp_amd.QTQP(a=sp.csc_matrix(q.a), b=q.b, c=q.c, z=q.z, p=sp.csc_matrix(q.p)).solve(
        linear_solver=qtqp_amd.LinearSolver.HIP_DENSE_FP64, verbose=False)
    st.append(sol.status.value)
  return len(qps) / (time.perf_counter() - t0), st


def calibrate_copy(mib: int = 1024):
  """A plain streaming kernel of known traffic for scaling HBM counters.

  out = in * 1 over `mib` MiB of FP32: one elementwise kernel that reads and
  writes exactly mib MiB each.
  """
  import torch
  x = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
  y = torch.empty_like(x)
  for _ in range(3):
    torch.mul(x, 1.0, out=y)
  torch.cuda.synchronize()
  print(json.dumps({"copy_kernel": "torch.mul(x, 1.0, out=y), FP32", "bytes_read": mib << 20,
                    "bytes_written": mib << 20, "dispatches": 3}), flush=True)


TRAFFIC_FILE = ROOT / "profiles" / "k_ipm_solve_mpc4096_traffic.md"


def measured_traffic():
  """(HBM bytes per k_ipm_solve launch, the launch time in ms the counters
  were taken at) from the committed counter summary, or (None, None)."""
  import re
  try:
    txt = TRAFFIC_FILE.read_text()
  except OSError:
    return None, None
  mt = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
  ml = re.search(r"traffic_launch_ms:\s*([0-9.eE+]+)", txt)
  return (float(mt.group(1)) if mt else None), (float(ml.group(1)) if ml else None)


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=3)
  ap.add_argument("--warmup", type=int, default=1)
  ap.add_argument("--count", type=int, default=4096)
  ap.add_argument("--base-seed", type=int, default=1000)
  ap.add_argument("--plain-sample", type=int, default=32)
  ap.add_argument("--cpu-budget", type=float, default=20.0)
  ap.add_argument("--no-cpu", action="store_true")
  ap.add_argument("--ws-train", type=int, default=100000)
  ap.add_argument("--ws-eval", type=int, default=16384)
  ap.add_argument("--ws-steps", type=int, default=8000)
  ap.add_argument("--no-ws", action="store_true")
  ap.add_argument("--no-single", action="store_true")
  ap.add_argument("--calibrate-copy", action="store_true")
  ap.add_argument("--no-plain", action="store_true")
  ap.add_argument("--single-only", action="store_true")
  ap.add_argument("--no-sparse", action="store_true")
  ap.add_argument("--no-sparse-cpu", action="store_true")
  ap.add_argument("--sparse-only", action="store_true",
                  help="only the sparse KKT backend leg (for its kernel profiles)")
  ap.add_argument("--sparse-sweep", action="store_true",
                  help="the sparse leg at several sizes of the reference's sparse family, "
                       "each with the scipy.sparse.linalg.factorized comparison")
  ap.add_argument("--sparse-reps", type=int, default=10,
                  help="timed factorisations and solves of the sparse leg")
  ap.add_argument("--sparse-no-qtqp", action="store_true",
                  help="sparse leg without its whole QTQP.solve (for counter runs)")
  ap.add_argument("--oracle-sample", type=int, default=64)
  ap.add_argument("--profile-qp", choices=("dense4000", "mpc64"), default="dense4000",
                  help="the QP --profile-single solves")
  ap.add_argument("--profile-single", choices=("mixed", "fp64"),
                  help="only solve the n=4000 m=6000 z=1000 dense QP once through QTQP.solve "
                       "(for kernel profiles of the device-resident loop)")
  args = ap.parse_args()
  if args.gpus != 1:
    print(json.dumps({"error": "this build measures one GPU only; the batched engine shards "
                               "QPs across GPUs by design but no multi-GPU run is supported"}))
    sys.exit(2)
  if args.calibrate_copy:
    calibrate_copy()
    return
  if args.profile_single:
    import scipy.sparse as sp
    import qtqp_amd
    from qtqp_amd import problems
    if args.profile_qp == "mpc64":
      qp, label = problems.mpc_qp(64, 12, 4, seed=0), "condensed MPC T=64 nx=12 nu=4 (seed 0)"
    else:
      qp, label = problems.random_dense_qp(4000, 6000, 1000, seed=0), \
          "n=4000 m=6000 z=1000 dense QP (seed 0)"
    ls = {"mixed": qtqp_amd.LinearSolver.HIP_DENSE,
          "fp64": qtqp_amd.LinearSolver.HIP_DENSE_FP64}[args.profile_single]
    q = qtqp_amd.QTQP(a=sp.csc_matrix(qp.a), b=qp.b, c=qp.c, z=qp.z, p=sp.csc_matrix(qp.p))
    times = []
    for _ in range(2):  # the first call also loads code objects and captures graphs
      t0 = time.perf_counter()
      sol = q.solve(linear_solver=ls, verbose=False, collect_stats=True)
      times.append(time.perf_counter() - t0)
    print(json.dumps({"workload": label, "linear_solver": ls.name, "status": sol.status.value,
                      "iterations": len(sol.stats), "solve_s": times}), flush=True)
    return

  if args.sparse_sweep:
    sweep = []
    for m_, n_, z_, dens in ((150, 100, 10, 0.1), (1500, 1000, 100, 0.01),
                             (3000, 2000, 100, 0.01), (6000, 4000, 1000, 0.002)):
      leg = sparse_leg(m=m_, n=n_, z=z_, density=dens, reps=args.sparse_reps,
                       cpu=not args.no_sparse_cpu)
      log(f"sparse sweep N={m_ + n_}: factor {leg['factor_ms']:.2f} ms")
      sweep.append(leg)
    print(json.dumps({"sparse_sweep": sweep}), flush=True)
    return
  if args.sparse_only:
    print(json.dumps(sparse_leg(reps=args.sparse_reps, cpu=not args.no_sparse_cpu,
                                qtqp=not args.sparse_no_qtqp)), flush=True)
    return
  from qtqp_amd import batch, problems
  from oracle import qtqp_oracle
  if args.single_only:
    print(json.dumps(single_legs(problems, qtqp_oracle)), flush=True)
    return

  log(f"generating {args.count} uneven MPC QPs")
  t0 = time.perf_counter()
  qps = problems.uneven_mpc_batch(args.count, base_seed=args.base_seed)
  n = np.array([q.n for q in qps])
  m = np.array([q.m for q in qps])
  log(f"generated in {time.perf_counter() - t0:.1f}s; uploading")
  bq = batch.BatchQP(qps)
  log("autotuning Gram tile shapes per size bucket")
  t0 = time.perf_counter()
  tuned = bq.autotune(sample=32, iters=3, precision="mixed")
  tuned["wall_s"] = time.perf_counter() - t0
  log(f"autotune {tuned['wall_s']:.2f}s: FP32 {tuned['fp32']['table']}, "
      f"FP64 {tuned['fp64']['table']}")
  log("warmup")
  for _ in range(args.warmup):
    bq.solve_device(precision="mixed")
  log("timed steps")
  times, cnts = [], []
  for k in range(args.steps):
    t0 = time.perf_counter()
    cnt = bq.solve_device(precision="mixed")
    times.append(time.perf_counter() - t0)
    cnts.append(cnt)
    log(f"step {k}: {times[-1] * 1e3:.1f} ms wall, {cnt['device_ms']:.1f} ms device, "
        f"{cnt['iterations']} iterations")
  sols = bq.results()
  work = bq.work_counts()
  ms_per_step = 1e3 * float(np.mean(times))
  value = args.count / (ms_per_step / 1e3)
  statuses = [s.status.value for s in sols]
  iters = np.array([s.iters for s in sols])
  launch_ms = bq.launch_times_ms()
  inst_ms = bq.instance_ms()
  log("per-launch ms: " + " ".join(f"{t:.1f}" for t in launch_ms))
  # the same solve with one launch per IPM iteration (graph replayed per iteration)
  log("per-iteration schedule")
  t0 = time.perf_counter()
  cnt_pi = bq.solve_device(precision="mixed", per_iteration=True)
  per_iter = {"ms": 1e3 * (time.perf_counter() - t0), "device_ms": cnt_pi["device_ms"],
              "launches": len(bq.launch_times_ms()),
              "qps_per_s": args.count / (time.perf_counter() - t0)}
  per_iter["statuses_equal"] = [s_.status.value for s_ in bq.results()] == statuses

  def timed_solves(reps, **kw):
    ts, cn = [], None
    for _ in range(reps):
      t0 = time.perf_counter()
      cn = bq.solve_device(**kw)
      ts.append(time.perf_counter() - t0)
    return 1e3 * float(np.mean(ts)), cn

  # the same 4096 QPs with the plain FP64 factor (the precision pattern's effect)
  log("batched FP64 (same batch, FP64 factor)")
  bq.solve_device(precision="fp64")
  ms64, cnt64 = timed_solves(args.steps, precision="fp64")
  sols64 = bq.results()
  work64 = bq.work_counts()
  batched_fp64 = {
      "value": args.count / (ms64 / 1e3), "unit": "QPs/s", "ms_per_solve": ms64,
      "statuses_equal": [s_.status.value for s_ in sols64] == statuses,
      "iterations_mean": float(np.mean([s_.iters for s_ in sols64])),
      "factorizations": {"fp32": int(work64[:, 0].sum()), "fp64": int(work64[:, 1].sum())},
      "refinement_passes": int(work64[:, 2].sum()),
      "mixed_over_fp64": value / (args.count / (ms64 / 1e3))}
  # tighter stopping tolerances, where every solved QP is meant to reach a
  # relative KKT residual of 1e-8 (the metric's accuracy) on the unscaled data
  tight = {"atol": 1e-10, "rtol": 1e-9}
  log(f"tight tolerances {tight}")
  bq.solve_device(precision="mixed", **tight)
  ms_t, cnt_t = timed_solves(args.steps, precision="mixed", **tight)
  sols_t = bq.results()
  kkt_t = np.array([kkt_rel_residual(s_.x, s_.y, s_.s, q.a, q.b, q.c, q.p, q.z)
                    for s_, q in zip(sols_t, qps) if s_.status.value == "solved"])
  st_t = [s_.status.value for s_ in sols_t]
  tight_leg = {"value": args.count / (ms_t / 1e3), "unit": "QPs/s", "ms_per_solve": ms_t,
               "tolerances": tight, "statuses": {k: st_t.count(k) for k in set(st_t)},
               "iterations_mean": float(np.mean([s_.iters for s_ in sols_t])),
               "rel_kkt_residual_solved": {"max": float(kkt_t.max()),
                                           "median": float(np.median(kkt_t)),
                                           "frac_le_1e-8": float(np.mean(kkt_t <= 1e-8))}}

  # roofline of the dominant kernel (the per-iteration IPM kernel)
  mc = bq.stored_rows()
  a_bytes = bq.row_bytes()
  flops32, flops64, bytes_ = algorithmic_work(n, mc, work, iters, a_bytes, m)
  _, _, bytes_dense = algorithmic_work(n, mc, work, iters, None, m)
  iter_ms = float(np.sum(launch_ms))
  t_mem = bytes_ / (HBM_PEAK_GBS * 1e9) * 1e3
  t_mfma = flops32 / (FP32_PEAK_TFLOPS * 1e12) * 1e3 + flops64 / (FP64_PEAK_TFLOPS * 1e12) * 1e3
  nl = max(1, len(launch_ms))
  if t_mem >= t_mfma:
    achieved = bytes_ / nl / (iter_ms / nl / 1e3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None}
  else:
    achieved = (flops32 + flops64) / nl / (iter_ms / nl / 1e3) / 1e12
    roof = {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None}
  # the counters come from a separate rocprofv3 run of the same workload,
  # committed under profiles/; its launch time shows whether it is current
  roof["traffic"], prof_ms = measured_traffic()
  roof["traffic_profile_launch_ms"] = prof_ms
  roof["traffic_profile_current"] = (
      prof_ms is not None and abs(prof_ms - iter_ms / nl) <= 0.05 * prof_ms)
  roof.update({"kernel": "k_ipm_solve", "launches": nl, "avg_launch_ms": iter_ms / nl,
               "algorithmic_bytes_per_launch": bytes_ / nl,
               "algorithmic_flops_per_launch": (flops32 + flops64) / nl,
               "stored_rows_of_a": float(mc.sum() / m.sum()),
               "a_read_bytes_over_dense_stored_rows": float(a_bytes.sum() / (8.0 * (mc * n).sum())),
               "algorithmic_bytes_per_launch_dense_stored_rows": bytes_dense / nl,
               "combined_frac": (t_mem + t_mfma) / iter_ms})

  # plain FP64 GPU paths, one QP per call, FP64 factor/solves, no batching/graph:
  # (1) the batched engine's kernels on a 1-QP batch, (2) the drop-in driver.
  if args.no_plain:
    plain = None
  else:
    pidx = np.linspace(0, args.count - 1, args.plain_sample).astype(int)
    sample = [qps[i] for i in pidx]
    log(f"plain FP64: {len(sample)} QPs (evenly spaced over the batch) one per call")
    singles = [batch.BatchQP([q]) for q in sample]
    for s_ in singles[:2]:
      s_.solve_device(precision="fp64", use_graph=False)
    t0 = time.perf_counter()
    for s_ in singles:
      s_.solve_device(precision="fp64", use_graph=False)
    plain_qps = len(sample) / (time.perf_counter() - t0)
    plain_status_match = all(s_.results()[0].status.value == statuses[i]
                             for i, s_ in zip(pidx, singles))
    for s_ in singles:
      s_.free()
    didx = pidx[np.linspace(0, len(pidx) - 1, 8).astype(int)]
    drop_qps, drop_st = dropin_fp64([qps[i] for i in didx])
    drop_match = all(st == statuses[i] for i, st in zip(didx, drop_st))
    denom = max(plain_qps, drop_qps)
    plain = {"value": denom, "unit": "QPs/s",
             "device_1qp_per_call": {"value": plain_qps, "sample": len(sample),
                                     "status_match": plain_status_match},
             "dropin_driver_1qp_per_call": {"value": drop_qps, "sample": len(didx),
                                            "status_match": drop_match},
             "sample_indices": "evenly spaced over the batch",
             "speedup_of_batched_mixed": value / denom}

  # accuracy against the FP64 oracle on a sample
  log(f"oracle check on {args.oracle_sample} instances")
  errs = []
  status_ok = True
  for i in np.linspace(0, args.count - 1, args.oracle_sample).astype(int):
    q = qps[i]
    ref = qtqp_oracle.solve_qp(q.a, q.b, q.c, q.z, q.p)
    status_ok &= ref["status"] == statuses[i]
    for k in "xys":
      got, want = getattr(sols[i], k), ref[k]
      errs.append(float(np.max(np.abs(got - want)) / max(1.0, np.max(np.abs(want)))))
  # achieved relative KKT residual of every solved instance
  kkt = [kkt_rel_residual(s_.x, s_.y, s_.s, q.a, q.b, q.c, q.p, q.z)
         for s_, q in zip(sols, qps) if s_.status.value == "solved"]
  # per-instance completion time: equilibration/start, then the device wall clock
  # from the start of the IPM launch to the end of the instance's last iteration
  prep_ms = max(0.0, cnts[-1]["device_ms"] - float(np.sum(launch_ms)))
  done_ms = prep_ms + inst_ms

  line = {
      "metric": METRIC, "value": value, "unit": "QPs/s", "n_gpus": 1, "steps": args.steps,
      "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
      "scaling": "weak", "vs_baseline": None,
      "dtype": "f64 refinement; f32 factor while it needs <= 3 refinement steps per "
               "IPM iteration, f64 factor after",
      "data": "synthetic (seeded condensed MPC)",
      "config": {"workload": f"batched uneven condensed MPC, {args.count} QPs, T~U{{8..64}}, "
                             f"nx~U{{4..16}}, nu~U{{1..6}}, seeds {args.base_seed}..",
                 "count": args.count, "sum_mn": int(np.sum(m * n)),
                 "mean_N": float(np.mean(n + m)), "precision": "mixed",
                 "factorizations": {"fp32": int(work[:, 0].sum()), "fp64": int(work[:, 1].sum())},
                 "tolerances": "atol 1e-7, rtol 1e-8"},
      "roofline": roof,
      "schedule": {"used": "one persistent launch per solve (k_ipm_solve)",
                   "per_iteration_launches": per_iter},
      "plain_fp64_gpu": plain,
      "iterations": {"mean": float(iters.mean()), "p99": float(np.percentile(iters, 99)),
                     "max": int(iters.max())},
      "instance_completion_ms": {"p50": float(np.percentile(done_ms, 50)),
                                 "p99": float(np.percentile(done_ms, 99)),
                                 "prepare_ms": prep_ms},
      "rel_kkt_residual_solved": {"max": float(np.max(kkt)), "median": float(np.median(kkt)),
                                  "frac_le_1e-8": float(np.mean(np.array(kkt) <= 1e-8))},
      "batched_fp64": batched_fp64,
      "tight_tolerance": tight_leg,
      "statuses": {s: statuses.count(s) for s in set(statuses)},
      "oracle_check": {"sample": len(errs) // 3, "status_match": bool(status_ok),
                       "max_rel_err_xys": max(errs) if errs else None},
      "counters": cnts[-1],
      "autotune": tuned,
  }
  bq.free()
  if not args.no_single:
    line.update(single_legs(problems, qtqp_oracle))
  if not args.no_sparse:
    line["sparse_kkt"] = sparse_leg(cpu=not args.no_sparse_cpu)
  if not args.no_ws:
    log("warm start leg")
    line["warm_start"] = warm_start_leg(args.ws_train, args.ws_eval, args.ws_steps)
  if not args.no_cpu:
    log("cpu baseline")
    line["cpu_baseline"] = cpu_baseline(args.base_seed, args.cpu_budget)
  print(json.dumps(line), flush=True)


if __name__ == "__main__":
  main()
