"""Writes readme_example.json from the reference README's example run.

Reads README.md (lines 76-112 of qtqp v0.0.5) as text: the example problem
and the printed iteration table / final vectors.  Nothing is imported or run
from the reference.  Usage: python tests/golden/make_readme_fixture.py
"""
import json
import pathlib
import re

SRC = pathlib.Path("/root/reference/README.md")
OUT = pathlib.Path(__file__).with_name("readme_example.json")


def main():
  text = SRC.read_text()
  rows = []
  for line in text.splitlines():
    m = re.match(r"^\|\s+(\d+)\s+\|(.*)\|\s*$", line)
    if not m:
      continue
    cells = [c.strip() for c in m.group(2).split("|")]
    rows.append({
        "iter": int(m.group(1)),
        "pcost": float(cells[0]), "dcost": float(cells[1]),
        "pres": float(cells[2]), "dres": float(cells[3]),
        "gap": float(cells[4]), "infeas": float(cells[5]),
        "mu": float(cells[6]),
        "solves": [int(v) for v in cells[7].split(",")],
    })

  def vec(name):
    m = re.search(r"sol\." + name + r"=array\(\[([^\]]*)\]\)", text)
    return [float(v) for v in m.group(1).split(",")]

  fixture = {
      "source": "qtqp README.md:82-112 (example run, v0.0.3 banner)",
      "problem": {
          "p": [[3.0, -1.0], [-1.0, 2.0]],
          "a": [[-1.0, 1.0], [1.0, 0.0], [0.0, 1.0]],
          "b": [-1.0, 0.3, -0.5], "c": [-1.0, -1.0], "z": 1,
      },
      "iterations": rows,
      "x": vec("x"), "y": vec("y"), "s": vec("s"),
      "status": "solved",
  }
  OUT.write_text(json.dumps(fixture, indent=1) + "\n")
  print(f"wrote {OUT} with {len(rows)} iterations")


if __name__ == "__main__":
  main()
