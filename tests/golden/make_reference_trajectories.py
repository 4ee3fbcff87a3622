"""Writes reference_trajectories.json: exact IPM trajectories of qtqp v0.0.5.

The per-iteration trajectory (iteration count, mu, sigma, alpha, tau, tau
method, IR solve counts) and the final x/y/s of qtqp's IPM are not pinned by
any closed form, so this script imports the reference package from
/root/reference/src in the build container only, solves seeded instances with
its dense backend (LinearSolver.SCIPY_DENSE) and stores the outputs.  Inputs
are not stored: they are regenerated from the seeds by qtqp_amd.problems and
by the generator restatements in tests/qp_generators.py.

Usage (build container only): python tests/golden/make_reference_trajectories.py
"""
import json
import pathlib
import sys

import numpy as np
import scipy.sparse as sp

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, "/root/reference/src")

import qtqp  # noqa: E402  (reference, build container only)
from qtqp_amd import problems  # noqa: E402
import qp_generators  # noqa: E402

OUT = pathlib.Path(__file__).with_name("reference_trajectories.json")


def cases():
  readme = json.loads(pathlib.Path(__file__).with_name("readme_example.json").read_text())["problem"]
  yield "readme", dict(kind="readme"), (np.array(readme["a"]), np.array(readme["b"]),
                                          np.array(readme["c"]), readme["z"], np.array(readme["p"]))
  qp = problems.random_dense_qp(200, 300, 50, seed=0)
  yield "dense_n200_m300_z50_s0", dict(kind="dense", n=200, m=300, z=50, seed=0), (qp.a, qp.b, qp.c, qp.z, qp.p)
  for seed in (1, 2):
    qp = problems.random_dense_qp(40, 60, 10, seed=seed)
    yield f"dense_n40_m60_z10_s{seed}", dict(kind="dense", n=40, m=60, z=10, seed=seed), (qp.a, qp.b, qp.c, qp.z, qp.p)
  qp = problems.mpc_qp(8, 4, 2, seed=3)
  yield "mpc_T8_nx4_nu2_s3", dict(kind="mpc", T=8, nx=4, nu=2, seed=3), (qp.a, qp.b, qp.c, qp.z, qp.p)
  for kind, gen in (("feasible", qp_generators.gen_feasible),
                    ("infeasible", qp_generators.gen_infeasible),
                    ("unbounded", qp_generators.gen_unbounded)):
    for (m, n, z), seed in (((30, 20, 4), 11), ((60, 40, 6), 12)):
      a, b, c, p = gen(m, n, z, seed)
      yield f"{kind}_m{m}_n{n}_z{z}_s{seed}", dict(kind=kind, m=m, n=n, z=z, seed=seed), (a, b, c, z, p)


def run(a, b, c, z, p, equilibrate):
  sol = qtqp.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=z, p=sp.csc_matrix(p)).solve(
      linear_solver=qtqp.LinearSolver.SCIPY_DENSE, verbose=False,
      collect_stats=True, equilibrate=equilibrate)
  its = []
  for st in sol.stats:
    its.append({
        "mu": st["mu"], "sigma": float(st["sigma"]), "alpha": st["alpha"],
        "tau": st["tau"], "pres": st["pres"], "dres": st["dres"], "gap": st["gap"],
        "solves": [st[k]["solves"] for k in ("q_lin_sys_stats", "predictor_lin_sys_stats",
                                             "corrector_lin_sys_stats")],
        "tau_method": [st[k].get("tau_method") for k in ("predictor_lin_sys_stats",
                                                         "corrector_lin_sys_stats")],
    })
  tolist = lambda v: [None if not np.isfinite(t) else float(t) for t in v]
  return {"status": sol.status.value, "iters": len(sol.stats), "trajectory": its,
          "x": tolist(sol.x), "y": tolist(sol.y), "s": tolist(sol.s)}


def main():
  out = {"reference": f"qtqp {qtqp.__version__} (LinearSolver.SCIPY_DENSE)", "cases": {}}
  for name, spec, (a, b, c, z, p) in cases():
    for eq in (True, False):
      key = f"{name}_eq{int(eq)}"
      out["cases"][key] = {"spec": spec, "equilibrate": eq, **run(a, b, c, z, p, eq)}
      print(key, out["cases"][key]["status"], out["cases"][key]["iters"])
  OUT.write_text(json.dumps(out, indent=0) + "\n")
  print("wrote", OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
  main()
