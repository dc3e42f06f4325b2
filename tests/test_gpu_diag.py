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


def test_hardware_sincos_error_within_guard_assumption(ctx):
    """The FP32 gyroid evaluators use the hardware sine/cosine on arguments
    reduced to [-0.5, 0.5] revolutions and size their guard bands for an
    absolute error of at most 1.3e-7 there; the library re-measures it over
    every FP32 argument in that range against FP64."""
    es, ec = X.diag_sincos_error(ctx)
    print(f"hardware sin/cos max abs error {es:.4e} / {ec:.4e}", flush=True)
    assert 0.0 < es <= 1.3e-7
    assert 0.0 < ec <= 1.3e-7


def test_hardware_deformation_functions_within_guard_assumption(ctx):
    """The FP32 deformations evaluate tanh, the logistic function and e^-x
    through the hardware base-2 exponential and reciprocal, and their
    position error bounds assume absolute errors of at most 3e-7, 2e-7 and
    2e-7 (kHwTanhErr, kHwSigmoidErr, kHwExpNegErr); the library re-measures
    them over every FP32 argument of their ranges against FP64."""
    et, es, ee = X.diag_deform_error(ctx)
    print(f"hardware tanh / logistic / exp(-x) max abs error {et:.4e} / {es:.4e} / {ee:.4e}",
          flush=True)
    assert 0.0 < et <= 3.0e-7
    assert 0.0 < es <= 2.0e-7
    assert 0.0 < ee <= 2.0e-7
