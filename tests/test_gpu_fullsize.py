"""Parity at the benchmark workloads' full sizes.  The oracle finishes a few steps of the 4M-atom
FCC system in about a second on the host's threads, so the device is compared with it directly;
replica batches and the adjoint are checked through properties that do not depend on size
(a replica run in a batch equals the same replica run alone; the adjoint equals central finite
differences)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from vampire_amd import system as S
from vampire_amd.context import FP32, FP64, Context
from vampire_amd.fit import Trajectory

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def c1():
    return S.config_c1(100)


@pytest.mark.timeout(300)
def test_c1_full_size_equals_oracle(gpu, c1):
    """FCC 100^3 (4,000,000 atoms, z = 12), 600 K, spins +z, 3 Heun steps.  On the Philox stream
    (seed 7) FP64 equals the oracle to 1e-10 per spin component and the FP32 magnetisation to
    1e-5.  Fed the oracle's normals, FP32 equals the oracle to 2e-6 per spin component (its
    arithmetic alone: on its own Philox stream the single-precision uniforms of the rare draws
    near zero move a few spins by ~1e-4 at this temperature) with |S| = 1 to 1e-6."""
    s = c1
    assert s.n == 4_000_000
    spins = np.zeros((3, s.n))
    spins[2] = 1.0
    o = O.Oracle(s)
    o.set_conditions(600.0)
    ref = o.llg_heun(spins, 3, seed=7, step0=0, threads=THREADS)
    out = {}
    for precision in (FP64, FP32):
        with Context(s, precision=precision, seed=7) as c:
            c.set_conditions(600.0)
            c.set_spins(spins)
            c.llg_heun(3)
            out[precision] = c.get_spins()
    assert np.abs(out[FP64] - ref).max() <= 1e-10
    assert np.abs(out[FP32].mean(axis=1) - ref.mean(axis=1)).max() <= 1e-5
    noise = np.stack([O.thermal_noise(7, s.n, q) for q in range(3)])
    with Context(s, precision=FP32, seed=7) as c:
        c.set_conditions(600.0)
        c.set_spins(spins)
        c.llg_heun(3, noise=noise)
        dev = c.get_spins()
    assert np.abs(dev - ref).max() <= 2e-6
    assert np.abs(np.linalg.norm(dev, axis=0) - 1.0).max() <= 1e-6


@pytest.mark.timeout(300)
def test_c4_full_size_equals_oracle(gpu):
    """SC film 256 x 256 x 8 with 10 % dilution (~472k atoms), two materials, DMI tensor exchange,
    300 K, random spins, 5 FP64 Heun steps: device = oracle to 1e-10."""
    s = S.config_c4()
    spins = S.random_spins(s.n, 4)
    o = O.Oracle(s)
    o.set_conditions(300.0)
    ref = o.llg_heun(spins, 5, seed=3, step0=0, threads=THREADS)
    with Context(s, precision=FP64, seed=3) as c:
        c.set_conditions(300.0)
        c.set_spins(spins)
        c.llg_heun(5)
        dev = c.get_spins()
    assert np.abs(dev - ref).max() <= 1e-10


@pytest.mark.timeout(300)
def test_c4_full_size_fp32_pulse_equals_oracle(gpu):
    """The film workload at its own precision: FP32 spins and fields, DMI tensor exchange, 10 %
    dilution (~472k atoms), random spins, with the temperature set at every step along the
    rising flank of a Gaussian 100 fs pulse from 300 K toward 1200 K (the per-step temperature of
    temperature_pulse.cpp:288-298).  Fed the oracle's normals, 20 steps agree with the FP64 oracle
    to 1e-5 per spin component (FP32 rounding, ~2e-7 per step), |S| = 1 to 1e-6."""
    s = S.config_c4()
    spins = S.random_spins(s.n, 4)
    temps = [300.0 + 900.0 * np.exp(-((k * 1e-16 - 70e-15) / 100e-15) ** 2) for k in range(20)]
    assert len(set(temps)) == len(temps)
    o = O.Oracle(s)
    ref = spins
    with Context(s, precision=FP32, seed=3) as c:
        c.set_spins(spins)
        for k, T in enumerate(temps):
            noise = O.thermal_noise(3, s.n, k)
            o.set_conditions(T)
            ref = o.llg_heun(ref, 1, noise=noise[None], threads=THREADS)
            c.set_conditions(T)
            c.llg_heun(1, noise=noise[None])
        dev = c.get_spins()
        assert c.topology[0] == "ell"
    err = np.abs(dev - ref).max()
    assert err <= 1e-5, err
    assert np.abs(np.linalg.norm(dev, axis=0) - 1.0).max() <= 1e-6


@pytest.mark.timeout(300)
def test_c1_full_size_fp32_same_normals_equals_fp64(gpu, c1):
    """FCC 100^3 at 600 K from +z: FP32 with the FP64 generator's normals (fp64_normals) against
    the FP64 step over 20 steps: within 2e-5 per spin component, M/Ms within 1e-6."""
    spins = np.zeros((3, c1.n))
    spins[2] = 1.0
    out = {}
    for key, prec in (("fp64", FP64), ("fp32", FP32)):
        with Context(c1, precision=prec, seed=7, fp64_normals=True) as c:
            c.set_conditions(600.0)
            c.set_spins(spins)
            c.llg_heun(20)
            out[key] = c.get_spins()
    assert np.abs(out["fp32"] - out["fp64"]).max() <= 2e-5
    assert np.abs(out["fp32"].mean(axis=1) - out["fp64"].mean(axis=1)).max() <= 1e-6


@pytest.mark.timeout(300)
def test_c2_full_batch_replicas_equal_standalone(gpu):
    """The 256-replica batch (SC, L in [8, 48], 8.67M atoms, T in [0, 1.5 Tc], alpha = 1, FP32):
    after 20 batched steps, the smallest, the largest and a middle replica equal the same replica
    run alone with key seed + r, bit for bit."""
    systems, temps = S.config_c2()
    batch = S.concat_systems(systems)
    spins = S.random_spins(batch.n, 12)
    with Context(batch, precision=FP32, seed=100) as c:
        for r, T in enumerate(temps):
            c.set_conditions(T, replica=r)
        c.set_spins(spins)
        c.llg_heun(20)
        got = c.get_spins()
    off = batch.replica_offset
    sizes = np.diff(off)
    for r in (int(np.argmin(sizes)), int(np.argmax(sizes)), len(systems) // 2):
        with Context(systems[r], precision=FP32, seed=100 + r) as c:
            c.set_conditions(temps[r])
            c.set_spins(spins[:, off[r]:off[r + 1]])
            c.llg_heun(20)
            assert np.array_equal(c.get_spins(), got[:, off[r]:off[r + 1]]), r


@pytest.mark.timeout(300)
def test_c2_full_batch_metropolis_replicas_equal_standalone(gpu):
    """The 256-replica batch under FP32 checkerboard Metropolis (the launches of a sweep mix the
    replicas' colours to even out their sizes): after 5 batched sweeps, the smallest, the largest,
    a middle replica and the first odd and even lattices equal the same replica swept alone with
    key seed + r, bit for bit."""
    systems, temps = S.config_c2()
    batch = S.concat_systems(systems)
    spins = S.random_spins(batch.n, 13)
    with Context(batch, precision=FP32, seed=200) as c:
        for r, T in enumerate(temps):
            c.set_conditions(T, replica=r)
        c.set_spins(spins)
        c.mc_sweep(5)
        got = c.get_spins()
    off = batch.replica_offset
    sizes = np.diff(off)
    side = np.rint(np.cbrt(sizes)).astype(int)
    picks = {int(np.argmin(sizes)), int(np.argmax(sizes)), len(systems) // 2,
             int(np.argmax(side % 2 == 1)), int(np.argmax(side % 2 == 0))}
    for r in sorted(picks):
        with Context(systems[r], precision=FP32, seed=200 + r) as c:
            c.set_conditions(temps[r])
            c.set_spins(spins[:, off[r]:off[r + 1]])
            c.mc_sweep(5)
            assert np.array_equal(c.get_spins(), got[:, off[r]:off[r + 1]]), r


@pytest.mark.timeout(300)
def test_c3_full_size_adjoint_matches_finite_differences(gpu):
    """BCC 32^3 (65,536 atoms), T = 0, Hz = 1 T, 30 degree tilt, 2000 steps recorded every 100 with
    checkpoints every 100: the adjoint gradient of a trajectory loss in J, Ku and alpha equals
    central finite differences of the FP64 forward to 1e-4 relative."""
    s = S.config_c3()
    rng = np.random.default_rng(3)
    th = np.radians(30.0 + 8.0 * rng.standard_normal(s.n))
    ph = 2 * np.pi * rng.random(s.n) * 0.05
    sp = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    J0, KU0, A0 = S.J_CO, S.KU_CO, 0.1
    traj = Trajectory(s, sp, nsteps=2000, record_every=100, checkpoint_every=100, H=(0.0, 0.0, 1.0))
    target = traj.forward(J0 * 1.1, KU0 * 0.85, A0 * 1.15)

    def loss(J, Ku, a):
        r = traj.forward(J, Ku, a) - target
        return float((r * r).mean()), 2.0 * r / r.size

    _, dmz = loss(J0, KU0, A0)
    dJ, dKu, da, _ = traj.backward(dmz)
    for name, g, x, f in (("J", dJ, J0, lambda v: loss(v, KU0, A0)[0]),
                          ("Ku", dKu, KU0, lambda v: loss(J0, v, A0)[0]),
                          ("alpha", da, A0, lambda v: loss(J0, KU0, v)[0])):
        eps = 1e-4 * abs(x)
        fd = (f(x + eps) - f(x - eps)) / (2 * eps)
        assert abs(g - fd) <= 1e-4 * abs(fd), (name, g, fd)
    traj.close()
