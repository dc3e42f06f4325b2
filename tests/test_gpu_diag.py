"""Device ceilings measured by the library's probes."""
import pytest

import xray_projection_render_amd as X

pytestmark = pytest.mark.gpu


def test_l2_read_bandwidth_is_plausible(ctx):
    """The L2 probe streams 2 MiB per XCD with 16-byte loads.  Sanity range:
    above half the HBM peak (L2 hits should beat HBM), below 128 B/clk per CU
    over 256 CUs at 2.4 GHz (79 TB/s)."""
    gbps = X.diag_l2_bandwidth(ctx)
    print(f"L2 read bandwidth {gbps:.0f} GB/s", flush=True)
    assert 4000.0 < gbps < 79000.0
