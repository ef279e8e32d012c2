"""Checkpoint / resume of the device state: a context restored from a checkpoint continues with
the same spins, Philox counters, conditions and Metropolis state as the one that wrote it, so the
continued run is bit-identical (both precisions, LLG and Metropolis, replicas)."""
import numpy as np
import pytest

from vampire_amd import system as S
from vampire_amd._lib import VmxError
from vampire_amd.context import FP32, FP64, FUSED, PLAIN, Context

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [FP64, FP32])
@pytest.mark.parametrize("kernel", [FUSED, PLAIN])
def test_llg_resume_is_bit_identical(gpu, precision, kernel, tmp_path):
    s = S.ferromagnet("fcc", 6)
    sp = S.random_spins(s.n, 4)
    with Context(s, precision=precision, kernel=kernel, seed=9) as a:
        a.set_conditions(450.0, (0.1, 0.0, 0.3))
        a.set_spins(sp)
        a.llg_heun(30)
        a.save_checkpoint(str(tmp_path / "a.chk"))
        a.llg_heun(30)
        ref, ref_step = a.get_spins(), a.step
    with Context(s, precision=precision, kernel=kernel, seed=9) as b:
        b.load_checkpoint(str(tmp_path / "a.chk"))
        assert b.step == 30
        b.llg_heun(30)
        assert b.step == ref_step
        assert np.array_equal(b.get_spins(), ref)


def test_metropolis_and_replicas_resume(gpu):
    reps = [S.ferromagnet("sc", L) for L in (5, 7, 6)]
    s = S.concat_systems(reps)
    sp = S.random_spins(s.n, 5)
    temps = (200.0, 500.0, 800.0)
    with Context(s, precision=FP64, seed=3) as a:
        for r, T in enumerate(temps):
            a.set_conditions(T, replica=r)
        a.set_spins(sp)
        a.mc_sweep(6)
        a.llg_heun(10)
        chk = a.checkpoint()
        a.mc_sweep(6)
        a.llg_heun(10)
        ref, ref_mc = a.get_spins(), a.mc_state()
    with Context(s, precision=FP64, seed=3) as b:
        b.restore(chk)                       # conditions come from the checkpoint
        b.mc_sweep(6)
        b.llg_heun(10)
        assert np.array_equal(b.get_spins(), ref)
        for x, y in zip(b.mc_state(), ref_mc):
            assert np.array_equal(x, y)


def test_checkpoint_rejects_other_contexts(gpu):
    s = S.ferromagnet("sc", 6)
    with Context(s, precision=FP64, seed=1) as a:
        a.set_conditions(300.0)
        a.set_spins(S.random_spins(s.n, 1))
        chk = a.checkpoint()
    with Context(s, precision=FP64, seed=2) as b:
        with pytest.raises(VmxError, match="seed"):
            b.restore(chk)
    with Context(s, precision=FP32, seed=1) as c:
        with pytest.raises(VmxError, match="precision"):
            c.restore(chk)
    with Context(s, precision=FP64, seed=1) as d:
        with pytest.raises(VmxError):
            d.restore(chk[: len(chk) // 2])
        with pytest.raises(VmxError, match="not a libvmx checkpoint"):
            d.restore(b"X" * len(chk))


def test_step_and_sweep_counters_drive_the_noise(gpu):
    """The Heun step and Metropolis sweep counters are the Philox counters: they count the calls,
    can be set (the drop-in starts both at the restored sim::time), and two contexts with equal
    spins and counters draw equal noise, while a different counter draws different noise."""
    s = S.ferromagnet("sc", 6)
    sp = S.random_spins(s.n, 8)

    def run(step, sweep):
        with Context(s, precision=FP64, seed=5) as c:
            c.set_conditions(400.0, (0.0, 0.0, 0.2))
            c.set_spins(sp)
            c.step, c.sweep = step, sweep
            assert (c.step, c.sweep) == (step, sweep)
            c.mc_sweep(2)
            c.llg_heun(3)
            assert (c.step, c.sweep) == (step + 3, sweep + 2)
            return c.get_spins()

    base = run(0, 0)
    assert np.array_equal(run(0, 0), base)
    assert np.array_equal(run(40, 17), run(40, 17))
    assert np.abs(run(40, 0) - base).max() > 1e-3       # Heun noise from another counter
    assert np.abs(run(0, 17) - base).max() > 1e-3       # Metropolis moves from another counter
