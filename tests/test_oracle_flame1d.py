"""The flame1d port (oracle/flame1d.c) against the reference script's own
printed table (tests/golden/flame1d.json, written by make_flame1d_golden.py
from tests/flame1d.mjs) and against the README's calibration claim."""
import json
import os

import pytest

from oracle.pyoracle import flame_speed

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "flame1d.json")))


def _cases():
    # tests/flame1d.mjs:85-88: dx list, F = 1 and max(1, 3 dx / 5e-4) for dx > 5e-5, L = max(0.03, 60 dx)
    out = []
    for dx in [5e-5, 2e-4, 1e-3, 3.125e-3]:
        for F in ([1, max(1, 3 * dx / 5e-4)] if dx > 5e-5 else [1]):
            out.append((dx, F))
    return out


def test_table_rows_match_reference_output():
    rows = GOLD["runs"][0]["rows"]
    cases = _cases()
    assert len(rows) == len(cases)
    for (dx, F), row in zip(cases, rows):
        assert abs(row["dx_mm"] - dx * 1000) < 1e-9
        r = flame_speed(2e10, 15000, dx, F, 1.0, max(0.03, 60 * dx), 3.0)
        # the reference prints 3 decimals (toFixed(3))
        assert abs(r["SL"] - row["SL"]) <= 0.0005 + 1e-12, (dx, F, r["SL"], row["SL"])


def test_calibration_rows_match_reference_output():
    rows = GOLD["runs"][1]["rows"]
    for row in rows:
        r = flame_speed(row["B"], 15000, 5e-5, 1, 1.0, 0.03, 1.0)
        assert abs(r["SL"] - row["SL"]) <= 0.0005 + 1e-12, (row, r)
        assert abs(r["Tb"] - row["Tb"]) <= 0.5 + 1e-9, (row, r)


def test_flame_speed_within_2pct_of_readme():
    # README.md:89: B = 2e10 m^3/(kg s) with Ta = 15000 K gives 0.38 m/s
    r = flame_speed(2e10, 15000, 5e-5, 1, 1.0, 0.03, 3.0)
    assert abs(r["SL"] - 0.38) / 0.38 < 0.02


@pytest.mark.parametrize("B", [3e9, 4e10])
def test_speed_grows_like_sqrt_B(B):
    # laminar flame theory: S_L ~ sqrt(B) at fixed Ta (closed-form scaling)
    r0 = flame_speed(1e10, 15000, 5e-5, 1, 1.0, 0.03, 1.0)["SL"]
    r1 = flame_speed(B, 15000, 5e-5, 1, 1.0, 0.03, 1.0)["SL"]
    ratio = r1 / r0
    expect = (B / 1e10) ** 0.5
    # within 15 %: a dropped or linear B dependence (ratio 1 or 4) fails
    assert abs(ratio / expect - 1) < 0.15
