"""The VALU throughput probe (gt_hip_debug_valu_rate) that bench.py prices the
evaluation against: a plausible rate for 256 CUs (between a tenth and twice
the 4-cycles-per-wave64 rate at 2.4 GHz, 39.3e12 lane operations/s), and
slower with a v_rsq_f32 in every four operations than with FMAs alone."""
import pytest

pytestmark = pytest.mark.gpu


def test_valu_rate_is_plausible(hip):
    fma = hip.valu_rate()
    rsq = hip.valu_rate(rsq=True)
    print(f"FMA {fma / 1e12:.1f}e12, FMA + rsq {rsq / 1e12:.1f}e12 lane operations/s", flush=True)
    assert 3.9e12 < fma < 7.9e13
    assert rsq < fma
