"""Writes tests/golden/flame1d.json from the reference's own flame1d script.

Runs `node tests/flame1d.mjs B=2e10` and `node tests/flame1d.mjs calibrate=1`
from the reference checkout (read-only, this container only: the GPU box has
no reference) and records the printed S_L tables with the wall time of each
run.  The reference script prints S_L to 3 decimals; those are the vectors.

usage: python tests/golden/make_flame1d_golden.py [/path/to/reference]
"""
import json
import os
import re
import subprocess
import sys
import time

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "flame1d.json")


def run(args):
    t0 = time.perf_counter()
    out = subprocess.run(["node", "tests/flame1d.mjs"] + args, cwd=REF, check=True,
                         capture_output=True, text=True).stdout
    return out, time.perf_counter() - t0


def main():
    node = subprocess.run(["node", "--version"], capture_output=True, text=True).stdout.strip()
    res = {"source": "tests/flame1d.mjs (reference), run under node " + node, "runs": []}
    out, wall = run(["B=2e10"])
    rows = []
    for m in re.finditer(r"dx=([0-9.]+)mm F=([0-9.]+): S=([0-9.]+) m/s", out):
        rows.append({"dx_mm": float(m.group(1)), "F": float(m.group(2)), "SL": float(m.group(3))})
    res["runs"].append({"args": ["B=2e10"], "stdout": out, "wall_s": wall, "rows": rows})
    out, wall = run(["calibrate=1"])
    rows = []
    for m in re.finditer(r"B=([0-9.e+]+) Ta=(\d+) resolved: S_L=([0-9.]+) m/s Tb=(\d+)", out):
        rows.append({"B": float(m.group(1)), "Ta": float(m.group(2)), "SL": float(m.group(3)),
                     "Tb": float(m.group(4))})
    res["runs"].append({"args": ["calibrate=1"], "stdout": out, "wall_s": wall, "rows": rows})
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
