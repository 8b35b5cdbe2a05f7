"""Precision study (1024^2 after 2000 warm-up steps): each schedule's one-step
error against the FP64 oracle step, per field."""
import json
import sys

import pytest

from precision_study import study

pytestmark = pytest.mark.gpu


def test_precision_study_1024():
    res = study(1024, 2000, 1, log=lambda *a: print(*a, flush=True))
    print(json.dumps(res), flush=True)
    sys.stdout.flush()
    e = res["errors"]
    # converged schedules: every field at FP32 round-off of the converged step
    for name in ("fp32_dc_1e-6", "fp32_dc_1e-6_direct_coarse"):
        assert res["resid"][name] < 1e-6, name
        for k, v in e[name].items():
            assert v < 1e-4, (name, k, v)
    # the reference schedule (one V-cycle) leaves the pressure unconverged,
    # which the converged schedules remove
    assert e["fp32_reference"]["p"] > 10 * e["fp32_dc_1e-6"]["p"]
    # the GPU's FP32 reference step is the FP32 oracle's to round-off
    for k in ("T", "Y_F", "Y_O"):
        assert abs(e["fp32_reference"][k] - e["oracle_fp32_reference"][k]) <= 0.1 * e["oracle_fp32_reference"][k] + 1e-7
