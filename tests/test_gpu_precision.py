"""Precision study at test size: deterministic FP32 Kahan sums stay within the
1e-5 parity bound of the FP64 oracle with no sign flips; plain FP32 atomics are
reported beside them (their error is larger and run-to-run nondeterministic)."""
import json

import pytest

from tests.precision_study import study

pytestmark = pytest.mark.gpu


def test_precision_study_thin_wedge(gpu_ctx):
    r = study(gpu_ctx, 0.0625)
    print("\n" + json.dumps(r))
    k, a = r["modes"]["fp32_kahan_sorted"], r["modes"]["fp32_atomic"]
    assert k["occupancy_rel_err_max"] <= 1e-5 and k["sign_flips"] == 0
    assert k["hausdorff_rel"] <= 1e-6
    assert a["occupancy_rel_err_max"] <= 1e-2
