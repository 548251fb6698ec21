"""dugks_advance_noalloc (no device allocation in the call; its time-step history lives in a
fixed per-GPU array) against dugks_advance on the same case: bitwise equal time steps and
fields over two calls in a row, the second shorter than the first, on the staged path and
both fused schedules; a history request beyond the array's 65536 steps is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fused,precision", [(0, 0), (2, 0), (3, 1)])
def test_advance_noalloc_matches_advance(fused, precision):
    from dugks_amd import cases, hip
    c = cases.cavity_case(24, "NC38", 10, 4.0, 0.1)
    a = hip.FvDVM(c, precision=precision, fused=fused)
    b = hip.FvDVM(c, precision=precision, fused=fused)
    for n in (40, 15):
        ha = a.run(n, record_dt=True)
        hb = b.run_noalloc(n, record_dt=True)
        assert len(hb) == n and np.array_equal(ha, hb), n
        assert a.dt == b.dt
        ga, gh = a.df()
        gb, hh = b.df()
        assert np.array_equal(ga, gb) and np.array_equal(gh, hh)
        ma, mb = a.cell_macros(), b.cell_macros()
        for k in ma:
            assert np.array_equal(ma[k], mb[k]), (n, k)
    b.run_noalloc(5)  # without a history: no limit applies, results still advance
    a.run(5)
    assert np.array_equal(a.df()[0], b.df()[0])
    with pytest.raises(hip.DugksError, match="65536"):
        b.run_noalloc(65537, record_dt=True)
    a.close()
    b.close()
