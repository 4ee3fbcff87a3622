"""Regenerates the inputs of the golden trajectory fixtures from their seeds."""
import json
import pathlib

import numpy as np

from qtqp_amd import problems
import qp_generators

GOLDEN = pathlib.Path(__file__).with_name("golden")


def load_trajectories():
  return json.loads((GOLDEN / "reference_trajectories.json").read_text())["cases"]


def inputs(spec):
  """(a, b, c, z, p) dense arrays for a fixture spec."""
  kind = spec["kind"]
  if kind == "readme":
    pr = json.loads((GOLDEN / "readme_example.json").read_text())["problem"]
    return (np.array(pr["a"]), np.array(pr["b"]), np.array(pr["c"]), pr["z"],
            np.array(pr["p"]))
  if kind == "dense":
    qp = problems.random_dense_qp(spec["n"], spec["m"], spec["z"], seed=spec["seed"])
    return qp.a, qp.b, qp.c, qp.z, qp.p
  if kind == "mpc":
    qp = problems.mpc_qp(spec["T"], spec["nx"], spec["nu"], seed=spec["seed"])
    return qp.a, qp.b, qp.c, qp.z, qp.p
  gen = {"feasible": qp_generators.gen_feasible,
         "infeasible": qp_generators.gen_infeasible,
         "unbounded": qp_generators.gen_unbounded}[kind]
  a, b, c, p = gen(spec["m"], spec["n"], spec["z"], spec["seed"])
  return a, b, c, spec["z"], p
