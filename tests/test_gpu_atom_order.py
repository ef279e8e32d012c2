"""The device's basis-major atom order is invisible: for lattices with a basis (FCC q = 4, BCC
q = 2) every API result in FP64 equals the caller-order run bit for bit, and the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from vampire_amd import system as S
from vampire_amd.context import FP64, PLAIN, FUSED, Context
from vampire_amd.fit import Trajectory

pytestmark = pytest.mark.gpu

CASES = [("fcc", 8, 4), ("bcc", 10, 2), ("sc", 12, 1)]


@pytest.mark.parametrize("crystal,L,q", CASES)
def test_period_detected(gpu, crystal, L, q):
    s = S.ferromagnet(crystal, L)
    with Context(s) as c:
        assert c.atom_period == q
    with Context(s, atom_order="caller") as c:
        assert c.atom_period == 1


@pytest.mark.parametrize("kernel", [PLAIN, FUSED])
@pytest.mark.parametrize("crystal,L,q", CASES[:2])
def test_llg_identical_in_both_orders(gpu, crystal, L, q, kernel):
    s = S.ferromagnet(crystal, L)
    sp = S.random_spins(s.n, 2)
    out = {}
    for order in ("auto", "caller"):
        with Context(s, precision=FP64, kernel=kernel, seed=5, atom_order=order) as c:
            c.set_conditions(500.0, (0.0, 0.3, 0.4))
            c.set_spins(sp)
            c.llg_heun(40)
            out[order] = c.get_spins()
            hs, he = c.fields()
            out[order + "_f"] = np.concatenate([hs, he])
            out[order + "_n"] = c.thermal_noise(3)
    assert np.array_equal(out["auto"], out["caller"])
    assert np.array_equal(out["auto_f"], out["caller_f"])
    assert np.array_equal(out["auto_n"], out["caller_n"])
    o = O.Oracle(s)
    o.set_conditions(500.0, (0.0, 0.3, 0.4))
    assert np.abs(out["auto"] - o.llg_heun(sp, 40, seed=5)).max() <= 1e-10


def test_noise_buffer_and_stats_in_both_orders(gpu):
    s = S.ferromagnet("fcc", 8)
    sp = S.random_spins(s.n, 3)
    noise = np.stack([O.thermal_noise(9, s.n, k) for k in range(10)])
    mask = (np.arange(s.n) % 5).astype(np.int32)
    res = {}
    for order in ("auto", "caller"):
        with Context(s, precision=FP64, seed=1, atom_order=order) as c:
            c.set_conditions(300.0)
            c.set_spins(sp)
            c.llg_heun(10, noise=noise)
            hm = c.add_magnetization(mask, 5)
            he = c.add_energy(mask, 5)
            c.stats_update()
            res[order] = (c.get_spins(), c.magnetization(hm)[0], c.energy(he)[0])
    assert np.array_equal(res["auto"][0], res["caller"][0])
    np.testing.assert_allclose(res["auto"][1], res["caller"][1], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(res["auto"][2], res["caller"][2], rtol=1e-12, atol=1e-9)
    o = O.Oracle(s)
    o.set_conditions(300.0)
    ref = o.llg_heun(sp, 10, noise=noise)
    assert np.abs(res["auto"][0] - ref).max() <= 1e-10


def test_metropolis_in_both_orders(gpu):
    s = S.ferromagnet("bcc", 6)
    sp = S.random_spins(s.n, 4)
    out = {}
    for order in ("auto", "caller"):
        with Context(s, precision=FP64, seed=3, atom_order=order) as c:
            assert c.atom_period == (2 if order == "auto" else 1)
            c.set_conditions(400.0)
            c.set_spins(sp)
            c.mc_sweep(5)
            out[order] = (c.get_spins(), c.mc_state())
    assert np.array_equal(out["auto"][0], out["caller"][0])
    o = O.Oracle(s)
    o.set_conditions(400.0)
    ref, sigma, moves, rej = o.mc_checkerboard(sp, 5, seed=3)
    assert np.abs(out["auto"][0] - ref).max() <= 1e-12
    assert out["auto"][1][2][0] == rej


def test_adjoint_in_both_orders(gpu):
    s = S.ferromagnet("bcc", 6)
    rng = np.random.default_rng(1)
    sp = np.array([[0.5], [0.0], [0.866]]) + 0.1 * rng.standard_normal((3, s.n))
    sp /= np.linalg.norm(sp, axis=0)
    h = 0.02 * rng.standard_normal((3, s.n))
    res = {}
    for order in ("auto", "caller"):
        t = Trajectory(s, sp, nsteps=60, record_every=5, checkpoint_every=20, H=(0, 0, 1.0))
        t.ctx.close()
        t.ctx = Context(s, precision=FP64, atom_order=order)
        mz = t.forward(S.J_CO, S.KU_CO, 0.1, h)
        g = t.backward(np.linspace(-1, 1, mz.size))
        res[order] = (mz, g)
        t.close()
    np.testing.assert_allclose(res["auto"][0], res["caller"][0], rtol=1e-13, atol=1e-15)
    for a, b in zip(res["auto"][1][:3], res["caller"][1][:3]):
        assert abs(a - b) <= 1e-10 * abs(b)
    np.testing.assert_allclose(res["auto"][1][3], res["caller"][1][3], rtol=1e-9, atol=1e-14)
