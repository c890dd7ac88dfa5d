"""The J^T J probe's two forms (vector reduction and matrix cores) against
numpy sums of the same generated rows."""
import pytest

pytestmark = pytest.mark.gpu


def test_jtj_forms_agree_with_numpy():
    from pmilos import jtj_probe
    r = jtj_probe.run(blocks=64, reps=3)
    # FP32 rows: products of 4 terms in FP32 summed in FP64 (vector), FP32
    # accumulation over 120 samples (MFMA); FP64 rows: rounding only
    assert r["fp32"]["max_rel_err_vector"] < 1e-6
    assert r["fp32"]["max_rel_err_mfma"] < 1e-5
    assert r["fp64"]["max_rel_err_vector"] < 1e-13
    assert r["fp64"]["max_rel_err_mfma"] < 1e-13
