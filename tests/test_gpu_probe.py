"""The measurement helper behind the deposition's roofline
(sfh_probe_valu_issue, no reference counterpart): it returns a rate between
1 T lane-instructions/s and the dual-issue peak at the 2.4 GHz peak clock
(256 CUs x 4 SIMDs x 64 lanes x 2 instructions per 4 cycles), and refuses
iteration counts outside 1 .. 2^24."""
import pytest

pytestmark = pytest.mark.gpu


def test_valu_issue_probe(sfh, gpu_ctx):
    r = gpu_ctx.valu_issue_rate(2000)
    print(f"VALU issue rate of the strip's mix: {r / 1e12:.1f} T lane-instructions/s")
    assert 1e12 < r <= 256 * 4 * 32 * 2.4e9 * 1.01
    for bad in (0, (1 << 24) + 1):
        with pytest.raises(sfh.ArgumentError):
            gpu_ctx.valu_issue_rate(bad)
