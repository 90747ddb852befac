"""Exhaustive GPU checks of the short exact sequences in pd-flow_amd/csrc/pdflow_pixel.hpp (the disc
projection's square root and reciprocal, and the step sizes' divisions of 0.5 and 1).

The iteration kernels' unit-disc projection (pdflow_cudalib.cu:849-859) takes 1/sqrt(x) on
x >= 1. It uses a square root without the range handling that the compiler wraps around
v_sqrt_f32, and a reciprocal that is the v_rcp_f32 estimate after one Newton step, without the
compiler's residual corrections and range handling. Every result must equal sqrtf and "/" bit for
bit. These tests evaluate every float of each range (about 10^9 inputs apiece) on the GPU and
count differing bits."""

import pytest

pytestmark = pytest.mark.gpu

ONE = 0x3F800000            # 1.0f
INF = 0x7F800000
TWO_TO_64 = 0x5F800000      # 2^64
TWO_TO_M94 = 0x10800000     # 2^-94
TWO_TO_124 = 0x7D800000     # 2^124


@pytest.fixture(scope="module")
def lib(gpu_available):
    from pdflow import lib
    return lib


@pytest.mark.parametrize("which,lo,hi,what", [
    (0, ONE, INF, "sqrt on [1, inf]"),
    (1, ONE, TWO_TO_64 - 1, "reciprocal on [1, 2^64)"),
    (2, ONE, INF, "1/sqrt on [1, inf]"),
    (3, TWO_TO_M94, TWO_TO_124, "1/d on [2^-94, 2^124]"),
    (4, TWO_TO_M94, TWO_TO_124, "0.5/d on [2^-94, 2^124]"),
])
def test_sequence_matches_ieee_on_its_whole_range(lib, which, lo, hi, what):
    n, first = lib.check_exact_sequence(which, lo, hi)
    assert n == 0, f"{what}: {n} inputs differ, first at bit pattern {first:#010x}"


def test_bad_arguments_are_refused(lib):
    with pytest.raises(lib.PDFlowError):
        lib.check_exact_sequence(5, ONE, ONE)
    with pytest.raises(lib.PDFlowError):
        lib.check_exact_sequence(0, INF, ONE)
