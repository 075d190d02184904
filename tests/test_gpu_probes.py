"""The measurement entry points behind the roofline figures: gh_measure_bandwidth
(streaming read and copy), gh_measure_issue (int8 MFMA and VALU issue
rates) and gh_measure_valu (single VALU instruction kinds). Their numbers vary with the device, so the checks are sanity bounds:
positive, below the published peaks, and the argument checks."""
import pytest

pytestmark = pytest.mark.gpu


def test_bandwidth_probe_below_peak(ctx):
    read, copy = ctx.measure_bandwidth(64 << 20, 2)
    assert 0 < read < 8000 * 1.05 and 0 < copy < 8000 * 1.05


def test_issue_probe_rates(ctx):
    import lattica_amd as la
    r = ctx.measure_issue(2)
    assert all(v > 0 for v in r["us"].values())
    # dense int8 peak 5.0 POP/s; allow clock variation
    assert 0 < r["mfma_32x32x32_tops"] < 5000 * 1.1
    assert 0 < r["mfma_16x16x64_tops"] < 5000 * 1.1
    # one wave instruction per 4 cycles per SIMD at most
    assert 0 < r["valu_ginst_per_s"] < ctx.num_cus * 4 * 3.0 / 4
    # the interleaved loop takes at least as long as its longer part
    assert r["mixed_over_max"] > 0.95
    with pytest.raises(la.GridHipError):
        ctx.measure_issue(0)
    with pytest.raises(la.GridHipError):
        ctx.measure_issue(9)


def test_valu_probe_rates(ctx):
    import lattica_amd as la
    r = ctx.measure_valu(2)
    rates = r["G_wave_inst_per_s"]
    assert set(rates) == set(ctx.VALU_KINDS)
    # at most one wave instruction per cycle per SIMD (clock below 3 GHz)
    for k, v in rates.items():
        assert 0 < v < ctx.num_cus * 4 * 3.0, k
    with pytest.raises(la.GridHipError):
        ctx.measure_valu(0)
    with pytest.raises(la.GridHipError):
        ctx.measure_valu(9)
