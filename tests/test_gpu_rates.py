"""sigpde_issue_rates on the GPU: the probe's figures are the right order of magnitude for a
gfx950 SIMD (one wave64 FMA takes 1..16 ticks, one 16x16x32 f16 MFMA 8..64) and the implied
peaks stay within the device's published rates (MI355X: FP32 vector 157.3, FP64 vector 78.6,
FP16 MFMA 2500 TFLOP/s) plus 10%."""
import pytest

from sigpde_amd import _native

pytestmark = pytest.mark.gpu


def test_issue_rates_are_plausible():
    r = _native.issue_rates()
    assert 1.0 <= r["ticks_per_fp32_fma"] <= 16.0, r
    assert 1.0 <= r["ticks_per_fp64_fma"] <= 16.0, r
    assert 8.0 <= r["ticks_per_mfma_f32_16x16x32_f16"] <= 64.0, r
    assert 0.3 <= r["mfma_plus_valu_over_sum"] <= 2.0, r
    # waves of one role each: the slower role takes at least its time alone (the full-overlap
    # ratio) and, give or take timing noise, at most the sum of the two
    assert 0.3 <= r["split_full_overlap_over_sum"] <= 1.0, r
    assert r["split_full_overlap_over_sum"] * 0.9 <= r["split_1x1_over_sum"] <= 1.3, r
    assert 0.3 <= r["split_2x2_over_sum"] <= 1.3, r
    assert 1.0 <= r["ticks_per_fp32_add_dpp"] <= 16.0, r
    assert 2.0 <= r["latency_fp32_fma_ticks"] <= 64.0, r
    assert 2.0 <= r["latency_fp32_add_dpp_ticks"] <= 64.0, r
    assert r["fp32_valu_tflops"] <= 157.3 * 1.1, r
    assert r["fp64_valu_tflops"] <= 78.6 * 1.1, r
    assert r["fp16_mfma_tflops"] <= 2500 * 1.1, r
