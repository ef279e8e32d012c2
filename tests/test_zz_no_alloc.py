"""vmx_stats_update_noalloc, the statistics update that allocates nothing, called through the C ABI:
the same kernels as vmx_stats_update, so its statistics equal vmx_stats_update's (bit for bit where the
reductions run in a fixed order).
Checked on two calls in a row, the second on a smaller system (the fixed scratch then still holds
the first call's partials), for one system and for a batch of replicas, FP64 and FP32, and its
refusal of statistics whose partials do not fit the scratch."""
import numpy as np
import pytest

from vampire_amd import _lib as L
from vampire_amd import system as S
from vampire_amd.context import FP32, FP64, Context

pytestmark = pytest.mark.gpu


def _update(ctx, noalloc):
    L.check(ctx.lib.vmx_stats_update_noalloc(ctx.h) if noalloc else ctx.lib.vmx_stats_update(ctx.h))


def _context(system, precision, spins):
    c = Context(system, precision=precision, seed=5)
    c.set_conditions(300.0, (0.0, 0.1, 0.4))
    c.set_spins(spins)
    n = system.n
    handles = (c.add_magnetization(np.zeros(n, np.int32), 2),                       # register histogram
               c.add_magnetization((np.arange(n) % 10).astype(np.int32), 10),       # LDS histogram
               c.add_energy((np.arange(n) % 3).astype(np.int32), 3))
    return c, handles


def _results(c, handles):
    hs, hm, he = handles
    return [c.magnetization(hs), c.magnetization(hm), c.energy(he)]


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_noalloc_update_equals_update(gpu, precision):
    batch = S.concat_systems([S.ferromagnet("sc", 5), S.ferromagnet("sc", 4)])
    for s in (S.ferromagnet("fcc", 12), S.ferromagnet("sc", 7), batch):   # 6912, then 343 atoms; replicas
        spins = S.random_spins(s.n, 7)
        ref, hr = _context(s, precision, spins)
        got, hg = _context(s, precision, spins)
        with ref, got:
            for call in range(2):
                _update(ref, False)
                _update(got, True)
                for k, (a, b) in enumerate(zip(_results(ref, hr), _results(got, hg))):
                    cur_a, mean_a, cnt_a = a
                    cur_b, mean_b, cnt_b = b
                    if k == 1:
                        # masks of more than 8 bins sum each workgroup's histogram in LDS with
                        # atomic adds, whose order varies from run to run: 1e-13 relative
                        np.testing.assert_allclose(cur_b, cur_a, rtol=1e-13, atol=1e-13 * np.abs(cur_a).max())
                        np.testing.assert_allclose(mean_b, mean_a, rtol=1e-13, atol=1e-13 * np.abs(mean_a).max())
                    else:             # register histograms, fixed-order reductions: bit for bit
                        assert np.array_equal(cur_a, cur_b) and np.array_equal(mean_a, mean_b)
                    assert cnt_a == cnt_b == call + 1
                ref.llg_heun(3)
                got.llg_heun(3)


def test_noalloc_update_refuses_partials_beyond_scratch(gpu):
    """SC 51^3 (132,651 atoms, 512 workgroups) with a 300-bin mask needs 512 x 300 x 32 bytes =
    4.9 MB of partials: refused before any work, while vmx_stats_update serves it."""
    s = S.ferromagnet("sc", 51)
    with Context(s, precision=FP32, seed=1) as c:
        c.set_spins(S.random_spins(s.n, 2))
        h = c.add_magnetization((np.arange(s.n) % 300).astype(np.int32), 300)
        rc = c.lib.vmx_stats_update_noalloc(c.h)
        assert rc == L.VMX_ERR_ARG
        assert b"no-allocation update holds at most" in c.lib.vmx_last_error()
        _, _, cnt = c.magnetization(h)
        assert cnt == 0
        _update(c, False)
        _, _, cnt = c.magnetization(h)
        assert cnt == 1
