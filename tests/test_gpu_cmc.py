"""Constrained Monte Carlo on the device against the oracle's colour-paired sweep (FP64, same Philox
streams and permutations) and against the reference's own cmc_step run (statistically)."""
import numpy as np
import pytest

import refcases as RC
from oracle import oracle as O
from vampire_amd import system as S
from vampire_amd.context import FP32, FP64, Context

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("topology", ["csr", "ell", "stencil"])
@pytest.mark.parametrize("angle", [False, True])
def test_cmc_matches_oracle(gpu, topology, angle):
    """FP64 constrained sweeps from random spins equal the oracle's colour-paired sweeps to 1e-12,
    with the same accepted / energy-rejected / sphere-rejected pair counts, for the cone move and
    the temperature-scaled angle move, on every neighbour layout."""
    s = S.ferromagnet("sc", 6)
    o = O.Oracle(s)
    o.set_conditions(400.0, (0.0, 0.0, 0.2))
    k = o.cmc(70.0, 120.0, move=1 if angle else 0, sigma=45.0)
    spins = S.random_spins(s.n, 31)
    ref, counts = o.cmc_checkerboard(spins, 6, k, seed=91, sweep0=0)
    with Context(s, precision=FP64, seed=91, topology=topology) as c:
        c.set_conditions(400.0, (0.0, 0.0, 0.2))
        c.set_spins(spins)
        c.mc_set_sigma(45.0)
        c.cmc_set_constraint(70.0, 120.0)
        c.cmc_set_move(angle)
        c.cmc_sweep(6)
        dev = c.get_spins()
        dcounts = c.cmc_counts()
    print("counts", counts, dcounts[0])
    assert np.abs(dev - ref).max() <= 1e-12
    assert np.array_equal(dcounts[0], counts)
    assert counts[0] > 0 and counts[1] > 0 and counts[2] > 0


@pytest.mark.parametrize("topology", ["csr", "ell", "stencil"])
def test_cmc_odd_colour_runs_match_oracle(gpu, topology):
    """SC 7^3 needs four colours of 117, 117, 55 and 54 atoms: runs of odd length, whose pairs
    the device finds through the inverse of the permutation's stride (one slot per entry, the
    first entries of pairs active).  FP64 sweeps equal the oracle's to 1e-12 with equal counts."""
    s = S.ferromagnet("sc", 7)
    o = O.Oracle(s)
    _, _, start = o.colouring()
    assert sorted(np.diff(start) % 2) == [0, 1, 1, 1]
    o.set_conditions(500.0, (0.1, 0.0, 0.3))
    k = o.cmc(40.0, 30.0, sigma=50.0)
    spins = S.random_spins(s.n, 12)
    ref, counts = o.cmc_checkerboard(spins, 5, k, seed=17, sweep0=3)
    with Context(s, precision=FP64, seed=17, topology=topology) as c:
        c.set_conditions(500.0, (0.1, 0.0, 0.3))
        c.set_spins(spins)
        c.sweep = 3
        c.mc_set_sigma(50.0)
        c.cmc_set_constraint(40.0, 30.0)
        c.cmc_sweep(5)
        assert np.abs(c.get_spins() - ref).max() <= 1e-12
        assert np.array_equal(c.cmc_counts()[0], counts)
    assert counts[0] > 0 and counts[1] > 0 and counts[2] > 0


def test_cmc_smallest_runs_match_oracle(gpu):
    """Colour runs of 1, 2, 3, 5 and 7 atoms (isolated spins, one colour each): a run of one atom
    has no pair and keeps its spin; the others equal the oracle's sweeps to 1e-12 with equal counts,
    alone and as replicas of one batch (which equals the runs alone bit for bit)."""
    sizes = [1, 2, 3, 5, 7]
    systems = [RC.isolated_spins(n) for n in sizes]
    spins = []
    for r, n in enumerate(sizes):
        sp = S.random_spins(n, 3 + r)
        sp[2] = np.abs(sp[2]) + 0.5
        spins.append(sp / np.linalg.norm(sp, axis=0))
    alone = []
    for r, x in enumerate(systems):
        o = O.Oracle(x)
        o.set_conditions(300.0, (0.0, 0.0, 0.5))
        ref, counts = o.cmc_checkerboard(spins[r], 20, o.cmc(40.0, 30.0, sigma=40.0), seed=5 + r)
        with Context(x, precision=FP64, seed=5 + r) as c:
            c.set_conditions(300.0, (0.0, 0.0, 0.5))
            c.set_spins(spins[r])
            c.mc_set_sigma(40.0)
            c.cmc_set_constraint(40.0, 30.0)
            c.cmc_sweep(20)
            dev, dcounts = c.get_spins(), c.cmc_counts()[0]
        assert np.abs(dev - ref).max() <= 1e-12
        assert np.array_equal(dcounts, counts)
        if sizes[r] == 1:
            assert np.array_equal(dev, spins[r]) and not counts.any()
        else:
            assert counts[0] > 0
        alone.append((dev, dcounts))
    batch = S.concat_systems(systems)
    with Context(batch, precision=FP64, seed=5) as c:
        c.set_conditions(300.0, (0.0, 0.0, 0.5))
        c.set_spins(np.concatenate(spins, axis=1))
        c.mc_set_sigma(40.0)
        c.cmc_set_constraint(40.0, 30.0)
        c.cmc_sweep(20)
        got, gcounts = c.get_spins(), c.cmc_counts()
    off = batch.replica_offset
    for r in range(len(sizes)):
        assert np.array_equal(got[:, off[r]:off[r + 1]], alone[r][0])
        assert np.array_equal(gcounts[r], alone[r][1])


def test_cmc_two_materials_tensor_exchange_matches_oracle(gpu):
    """The thin film with two materials (moments 1.6 and 7.6 mu_B: the compensating spin scales by
    mu1/mu2) and tensor exchange with DMI: device equals the oracle to 1e-12 with equal counts."""
    s = RC.system_of("film_dmi")
    o = O.Oracle(s)
    o.set_conditions(300.0, (0.0, 0.0, 0.5))
    k = o.cmc(30.0, 10.0, sigma=30.0)
    spins = S.random_spins(s.n, 4)
    ref, counts = o.cmc_checkerboard(spins, 5, k, seed=3)
    with Context(s, precision=FP64, seed=3) as c:
        c.set_conditions(300.0, (0.0, 0.0, 0.5))
        c.set_spins(spins)
        c.mc_set_sigma(30.0)
        c.cmc_set_constraint(30.0, 10.0)
        c.cmc_sweep(5)
        assert np.abs(c.get_spins() - ref).max() <= 1e-12
        assert np.array_equal(c.cmc_counts()[0], counts)


def test_cmc_replicas_equal_standalone(gpu):
    """A batch of replicas with their own constraints and temperatures equals each run alone with
    seed + r, bit for bit, counts included."""
    sizes = [4, 6, 5]
    cons = [(20.0, 0.0), (60.0, 45.0), (90.0, 270.0)]
    temps = [200.0, 400.0, 600.0]
    systems = [S.ferromagnet("sc", L) for L in sizes]
    batch = S.concat_systems(systems)
    spins = [S.random_spins(x.n, 40 + r) for r, x in enumerate(systems)]
    with Context(batch, precision=FP64, seed=12) as c:
        for r in range(3):
            c.set_conditions(temps[r], replica=r)
            c.cmc_set_constraint(*cons[r], replica=r)
        c.set_spins(np.concatenate(spins, axis=1))
        c.cmc_sweep(7)
        got, gcounts = c.get_spins(), c.cmc_counts()
    off = batch.replica_offset
    for r, x in enumerate(systems):
        with Context(x, precision=FP64, seed=12 + r) as c:
            c.set_conditions(temps[r])
            c.cmc_set_constraint(*cons[r])
            c.set_spins(spins[r])
            c.cmc_sweep(7)
            assert np.array_equal(c.get_spins(), got[:, off[r]:off[r + 1]])
            assert np.array_equal(c.cmc_counts()[0], gcounts[r])


def test_cmc_host_and_device_permutations_agree(gpu):
    """A context of few replicas forms each sweep's permutation table on the host (up to 1 KiB);
    a larger batch forms it on the device (k_cmc_perm).  A batch of 24 SC 4^3 replicas (two
    colours: 96 entries, 1.5 KiB) equals each replica run alone (4 entries), bit for bit."""
    systems = [S.ferromagnet("sc", 4) for _ in range(24)]
    batch = S.concat_systems(systems)
    spins = [S.random_spins(x.n, 70 + r) for r, x in enumerate(systems)]
    with Context(batch, precision=FP64, seed=40) as c:
        for r in range(len(systems)):
            c.set_conditions(150.0 + 20.0 * r, replica=r)
        c.cmc_set_constraint(35.0, 80.0)
        c.set_spins(np.concatenate(spins, axis=1))
        c.cmc_sweep(6)
        got, gcounts = c.get_spins(), c.cmc_counts()
    off = batch.replica_offset
    for r in (0, 7, 23):
        with Context(systems[r], precision=FP64, seed=40 + r) as c:
            c.set_conditions(150.0 + 20.0 * r)
            c.cmc_set_constraint(35.0, 80.0)
            c.set_spins(spins[r])
            c.cmc_sweep(6)
            assert np.array_equal(c.get_spins(), got[:, off[r]:off[r + 1]])
            assert np.array_equal(c.cmc_counts()[0], gcounts[r])
    assert gcounts[:, 0].sum() > 0


@pytest.mark.parametrize("precision", [FP64, FP32])
@pytest.mark.parametrize("case", ["sc_batch", "fcc", "film_dmi", "bcc_rotated"])
def test_cmc_colour_order_equals_storage_order(gpu, case, precision):
    """Calls of several constrained sweeps run on the colour-ordered copy of the spins (pair list,
    moments and tables in colour order, Philox counters from the caller's indices, the
    magnetisation summed in storage order); single sweeps run in storage order.  Eight single
    sweeps and one call of eight give the same spins and outcome counts bit for bit: replicas of
    odd and even size, FCC and BCC stored basis-major, two materials with DMI tensors."""
    if case == "sc_batch":
        s = S.concat_systems([S.ferromagnet("sc", L) for L in (5, 4, 7, 6)])
    elif case == "fcc":
        s = S.ferromagnet("fcc", 4)
    elif case == "film_dmi":
        s = RC.system_of("film_dmi")
    else:
        s = RC.system_of("bcc4_rotated")
    spins = S.random_spins(s.n, 18)
    spins[2] = np.abs(spins[2]) + 0.3
    spins /= np.linalg.norm(spins, axis=0)
    out = []
    for calls in ([1] * 8, [8]):
        with Context(s, precision=precision, seed=6, topology="stencil") as c:
            for r in range(s.num_replicas):
                c.set_conditions(200.0 + 100.0 * r, (0.0, 0.0, 0.3), replica=r)
            c.mc_set_sigma(40.0)
            c.cmc_set_constraint(30.0, 50.0)
            c.set_spins(spins)
            for k in calls:
                c.cmc_sweep(k)
            out.append((c.get_spins(), c.cmc_counts()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][1][:, 0].sum() > 0


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_cmc_curve_matches_reference(gpu, precision):
    """Statistical parity with the reference's montecarlo::cmc_step on SC 8^3 with uniaxial
    anisotropy, magnetisation held at phi = 40, theta = 30 degrees: <m.v>(T) and <E>(T) at
    100-650 K within combined error bars, the moment on the constraint axis throughout, and the
    accepted / energy-rejected / sphere-rejected fractions within 0.01; the four temperatures run
    as replicas of one launch, in FP64 and in FP32."""
    f = RC.load("cmc_sc8")
    rows = f["rows"]
    temps = rows[:, 0]
    phi, theta = float(f["phi"]), float(f["theta"])
    base = S.ferromagnet("sc", 8)
    n = base.n
    batch = S.concat_systems([S.ferromagnet("sc", 8) for _ in temps])
    o = O.Oracle(base)
    k = o.cmc(phi, theta)
    v, P = np.array(k.v), np.array(k.P).reshape(3, 3)
    mask = np.zeros(n, np.int32)
    sweeps, equil, every = 6000, 1000, 5
    with Context(batch, precision=precision, seed=2025) as c:
        for r, T in enumerate(temps):
            c.set_conditions(float(T), replica=r)
        c.cmc_set_constraint(phi, theta)
        c.set_spins(np.tile(v[:, None], (1, batch.n)))
        c.cmc_sweep(equil)
        ms, es = [[] for _ in temps], [[] for _ in temps]
        trans = np.zeros(len(temps))
        for _ in range((sweeps - equil) // every):
            c.cmc_sweep(every)
            sp = c.get_spins().astype(np.float64)
            for r, T in enumerate(temps):
                x = sp[:, r * n:(r + 1) * n]
                m = x.sum(axis=1) / n
                ms[r].append(m @ v)
                trans[r] = max(trans[r], float(np.abs((P @ m)[:2]).max()))
                o.set_conditions(float(T))
                es[r].append(o.energy(x, mask, 2)[0][0])
        counts = c.cmc_counts()
    # N pair trials per sweep; outcome fractions as in the reference run (random pairs there)
    assert np.array_equal(counts.sum(axis=1), np.full(len(temps), float(sweeps * n)))
    np.testing.assert_allclose(counts / (sweeps * n), f["counts"][:, :3] / f["counts"][:, 3:], atol=0.01)
    got = []
    for r, T in enumerate(temps):
        mm, dm = RC.block_mean(ms[r])
        me, de = RC.block_mean(es[r])
        # FP32 spins keep the transverse moment to single-precision rounding only
        got.append([T, mm, dm, trans[r] if precision == FP64 else 0.0, me, de])
        print(T, mm, dm, trans[r], me, de)
    RC.assert_cmc_curve(np.array(got))
    if precision == FP32:
        assert trans.max() < 1e-4
