"""dugks_get_statistics_noalloc (library-held scratch, no allocation)
against dugks_get_statistics on the same states, through the C ABI: bit for
bit, on two calls in a row whose second has fewer wall-normal rows, and the
row limit rejected with an error."""
import ctypes

import numpy as np
import pytest

import dugks_amd
from dugks_amd import PREC_FP32, PREC_FP64, configs

pytestmark = pytest.mark.gpu


def _both(eng):
    a = eng.statistics()
    b = eng.statistics(noalloc=True)
    return a, b


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP32], ids=["fp64", "fp32"])
def test_noalloc_statistics_match_the_allocating_entry_point(precision):
    big = dugks_amd.Engine(configs.turbulent_channel(nx=32, ny=48, nz=8, flow_rate=0.0), precision=precision,
                           device=0)
    small = dugks_amd.Engine(configs.turbulent_channel(nx=16, ny=24, nz=4, flow_rate=0.0), precision=precision,
                             device=0)
    big.step(3)
    small.step(2)
    for eng, rows in ((big, 48), (small, 24)):  # the second call has fewer rows
        (pa, da), (pb, db) = _both(eng)
        assert pa.shape == (rows, 9) and pb.shape == (rows, 9)
        assert np.array_equal(pa, pb) and np.array_equal(da, db)
        assert np.abs(pa).max() > 0
    # the first state again after the smaller call: the scratch holds the
    # last call's contents, which must not leak into the next
    (pa, da), (pb, db) = _both(big)
    assert np.array_equal(pa, pb) and np.array_equal(da, db)


def test_noalloc_statistics_reject_more_rows_than_the_limit():
    eng = dugks_amd.Engine(configs.base(nx=4, ny=2056, nz=1, periods_y=0, ywidth=2.0), precision=PREC_FP32,
                           device=0)
    with pytest.raises(dugks_amd.DugksError):
        eng.statistics(noalloc=True)
    prof, _ = eng.statistics()  # the allocating entry point has no limit
    assert prof.shape == (2056, 9)
