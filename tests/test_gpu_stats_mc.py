"""Device statistics (FP64 reductions) and checkerboard Metropolis against the oracle and the
reference's golden vectors."""
import numpy as np
import pytest

import refcases as RC
from oracle import oracle as O
from vampire_amd import system as S
from vampire_amd.context import FP32, FP64, FUSED, Context

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("topology", ["csr", "ell", "stencil"])
@pytest.mark.parametrize("name", ["sc6_uniaxial", "fcc3_cubic", "bcc4_rotated", "film_dmi", "sc6_vectorial",
                                  "film_vectorial", "film_local"])
def test_stats_match_reference(gpu, name, topology):
    """System and material magnetisation and system energies after the reference's steps
    (stats_after = system M [8], material M [4(nmat+1)], system E [5x2])."""
    f = RC.load(name)
    s = RC.system_of(name)
    T, H = RC.conditions(name)
    nmat = len(s.materials)
    with Context(s, precision=FP64, topology=topology) as c:
        c.set_conditions(T, H, **RC.flags(name))
        c.set_spins(f["snaps"][-1])
        hs = c.add_magnetization(np.zeros(s.n, np.int32), 2)
        hm = c.add_magnetization(s.type, nmat + 1)
        he = c.add_energy(np.zeros(s.n, np.int32), 2)
        c.stats_update()
        m_sys, _, cnt = c.magnetization(hs)
        m_mat, _, _ = c.magnetization(hm)
        e_sys, _, _ = c.energy(he)
    ref = f["stats_after"]
    assert cnt == 1
    np.testing.assert_allclose(m_sys, ref[:8], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(m_mat, ref[8:8 + 4 * (nmat + 1)], rtol=1e-12, atol=1e-14)
    e_ref = ref[8 + 4 * (nmat + 1):].reshape(5, 2)
    np.testing.assert_allclose(e_sys, e_ref, rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_stats_running_means(gpu, precision):
    s = S.ferromagnet("bcc", 5)
    o = O.Oracle(s)
    o.set_conditions(200.0, (0.1, 0.0, 0.3))
    mask = (np.arange(s.n) % 3).astype(np.int32)   # ids 0,1 magnetic, 2 collects the rest
    with Context(s, precision=precision, seed=2) as c:
        c.set_conditions(200.0, (0.1, 0.0, 0.3))
        c.set_spins(S.random_spins(s.n, 8))
        hm = c.add_magnetization(mask, 3)
        he = c.add_energy(mask, 3)
        acc_m, acc_e = np.zeros(12), np.zeros((5, 3))
        for _ in range(4):
            c.llg_heun(5)
            c.stats_update()
            sp = c.get_spins()
            rm = o.magnetization(sp, mask, 3)
            re = o.energy(sp, mask, 3)
            cur, _, _ = c.magnetization(hm)
            ecur, _, _ = c.energy(he)
            tol = 1e-12 if precision == FP64 else 1e-6
            np.testing.assert_allclose(cur, rm, rtol=tol, atol=tol)
            np.testing.assert_allclose(ecur, re, rtol=tol, atol=tol * np.abs(re).max())
            acc_m += rm
            acc_e[:, :2] += re[:, :2]
        _, mean, cnt = c.magnetization(hm)
        _, emean, ecnt = c.energy(he)
        assert cnt == 4 and ecnt == 4
        np.testing.assert_allclose(mean, acc_m, rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(emean[:, :2], acc_e[:, :2], rtol=1e-6, atol=1e-6)
        c.stats_reset()
        _, mean, cnt = c.magnetization(hm)
        assert cnt == 0 and not mean.any()


@pytest.mark.parametrize("topology", ["csr", "ell", "stencil"])
def test_mc_checkerboard_matches_oracle(gpu, topology):
    """FP64 checkerboard sweeps on the device reproduce the oracle's colour-ordered Metropolis
    (same colouring, Philox streams and acceptance), including the adaptive cone width."""
    s = S.ferromagnet("sc", 6)
    o = O.Oracle(s)
    o.set_conditions(400.0, (0.0, 0.0, 0.2))
    spins = S.random_spins(s.n, 21)
    ref, sigma, moves, rej = o.mc_checkerboard(spins, 8, seed=77, sweep0=0, sigma=60.0)
    with Context(s, precision=FP64, seed=77, topology=topology) as c:
        c.set_conditions(400.0, (0.0, 0.0, 0.2))
        c.set_spins(spins)
        assert c.mc_num_colours() == 2
        c.mc_sweep(8)
        dev = c.get_spins()
        dsig, dmov, drej = c.mc_state()
    assert np.abs(dev - ref).max() <= 1e-12
    assert dmov[0] == moves and drej[0] == rej
    assert abs(dsig[0] - sigma) <= 1e-12 * sigma


def test_mc_smallest_runs_match_oracle(gpu):
    """Isolated spins (5 K, 5 T) in colour runs of 1, 2, 3, 5 and 7 atoms: FP64 sweeps equal the
    oracle's to 1e-12 with equal move and rejection counts and cone width, in calls of one sweep
    (storage order) and of eight (colour order); the batch of all five equals the runs alone bit for
    bit."""
    sizes = [1, 2, 3, 5, 7]
    systems = [RC.isolated_spins(n) for n in sizes]
    spins = [S.random_spins(n, 60 + r) for r, n in enumerate(sizes)]
    alone = []
    for r, x in enumerate(systems):
        o = O.Oracle(x)
        o.set_conditions(5.0, (0.0, 0.0, 5.0))
        ref, sigma, moves, rej = o.mc_checkerboard(spins[r], 8, seed=9 + r, sweep0=0, sigma=60.0)
        assert rej > 0
        for per_call in (1, 8):
            with Context(x, precision=FP64, seed=9 + r) as c:
                c.set_conditions(5.0, (0.0, 0.0, 5.0))
                c.set_spins(spins[r])
                for _ in range(8 // per_call):
                    c.mc_sweep(per_call)
                dev = c.get_spins()
                dsig, dmov, drej = c.mc_state()
            assert np.abs(dev - ref).max() <= 1e-12
            assert dmov[0] == moves and drej[0] == rej
            assert abs(dsig[0] - sigma) <= 1e-12 * sigma
        alone.append(dev)
    batch = S.concat_systems(systems)
    with Context(batch, precision=FP64, seed=9) as c:
        c.set_conditions(5.0, (0.0, 0.0, 5.0))
        c.set_spins(np.concatenate(spins, axis=1))
        c.mc_sweep(8)
        got = c.get_spins()
    off = batch.replica_offset
    for r in range(len(sizes)):
        assert np.array_equal(got[:, off[r]:off[r + 1]], alone[r])


@pytest.mark.parametrize("precision", [FP64, FP32])
@pytest.mark.parametrize("case", ["sc_batch", "fcc", "film_dmi", "bcc_rotated"])
def test_mc_colour_order_equals_storage_order(gpu, case, precision):
    """Calls of several sweeps run with the spins in colour order (each replica's atoms colour by
    colour); single sweeps run in storage order.  Eight single sweeps and one call of eight give
    the same spins, cone widths and rejection totals bit for bit."""
    if case == "sc_batch":
        s = S.concat_systems([S.ferromagnet("sc", L) for L in (5, 4, 7, 6)])
    elif case == "fcc":
        s = S.ferromagnet("fcc", 4)
    elif case == "film_dmi":
        s = RC.system_of("film_dmi")
    else:
        s = RC.system_of("bcc4_rotated")
    spins = S.random_spins(s.n, 8)
    out = []
    for calls in ([1] * 8, [8]):
        with Context(s, precision=precision, seed=5, topology="stencil") as c:
            for r in range(s.num_replicas):
                c.set_conditions(250.0 + 100.0 * r, (0.0, 0.0, 0.3), replica=r)
            c.set_spins(spins)
            for k in calls:
                c.mc_sweep(k)
            out.append((c.get_spins(), c.mc_state()))
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)


def test_mc_replicas_equal_standalone(gpu):
    """A batch of even and odd periodic lattices (two colours and four): the sweep's launches carry
    different replicas' colours side by side, each replica's colours in order, so every replica
    equals the same lattice swept alone with seed + r, bit for bit, cone widths and rejection
    totals included."""
    sizes = [5, 4, 6, 3, 7, 8]
    temps = [300.0, 500.0, 200.0, 700.0, 400.0, 600.0]
    systems = [S.ferromagnet("sc", L) for L in sizes]
    batch = S.concat_systems(systems)
    spins = [S.random_spins(x.n, 60 + r) for r, x in enumerate(systems)]
    with Context(batch, precision=FP64, seed=33) as c:
        for r, T in enumerate(temps):
            c.set_conditions(T, replica=r)
        c.set_spins(np.concatenate(spins, axis=1))
        c.mc_sweep(6)
        got = c.get_spins()
        gsig, gmov, grej = c.mc_state()
    off = batch.replica_offset
    ncol = []
    for r, x in enumerate(systems):
        with Context(x, precision=FP64, seed=33 + r) as c:
            c.set_conditions(temps[r])
            c.set_spins(spins[r])
            c.mc_sweep(6)
            ncol.append(c.mc_num_colours())
            assert np.array_equal(c.get_spins(), got[:, off[r]:off[r + 1]]), r
            sig, mov, rej = c.mc_state()
            assert sig[0] == gsig[r] and mov[0] == gmov[r] and rej[0] == grej[r], r
    assert min(ncol) == 2 and max(ncol) > 2


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_mc_curve_matches_reference(gpu, precision):
    """Statistical parity: <|m|>(T) of checkerboard Metropolis on SC 8^3 against the reference
    random-site montecarlo::mc_step (its own MT19937 stream) within combined error bars
    (4 sigma + 0.01 for finite-run bias), temperatures 100-900 K, and the pseudo-critical
    temperature (the <|m|> = 1/2 crossing) within 4 combined sigma + 5 K; in FP64 and in FP32
    (the precision of the replica-batching workload)."""
    f = RC.load("mc_sc8")
    rows = f["M_T"]
    temps = rows[:, 0]
    systems = [S.ferromagnet("sc", 8, Ku=0.0) for _ in temps]
    batch = S.concat_systems(systems)
    n = systems[0].n
    sweeps, equil, every = 6000, 1000, 5
    with Context(batch, precision=precision, seed=2024) as c:
        for r, T in enumerate(temps):
            c.set_conditions(float(T), replica=r)
        sp = np.zeros((3, batch.n))
        sp[2] = 1.0
        c.set_spins(sp)
        c.mc_sweep(equil)
        samples = []
        for _ in range((sweeps - equil) // every):
            c.mc_sweep(every)
            x = c.get_spins().reshape(3, len(temps), n).sum(axis=2)
            samples.append(np.sqrt((x * x).sum(axis=0)) / n)
    mean, err = RC.block_mean(samples)
    RC.assert_mc_curve(temps, mean, err)


@pytest.mark.parametrize("nbins", [2, 4, 6, 10])
def test_stats_mask_sizes_match_oracle(gpu, nbins):
    """Masks whose id changes at every atom: per-thread register histograms (magnetisation up to
    8 bins, energy up to 4) and workgroup LDS histograms above; magnetisation and energy
    statistics equal the oracle's."""
    s = S.ferromagnet("fcc", 4)
    o = O.Oracle(s)
    o.set_conditions(250.0, (0.0, 0.2, 0.1))
    mask = (np.arange(s.n) % nbins).astype(np.int32)
    sp = S.random_spins(s.n, 12)
    with Context(s, precision=FP64, seed=4) as c:
        c.set_conditions(250.0, (0.0, 0.2, 0.1))
        c.set_spins(sp)
        hm = c.add_magnetization(mask, nbins)
        he = c.add_energy(mask, nbins)
        c.stats_update()
        cur, _, _ = c.magnetization(hm)
        ecur, _, _ = c.energy(he)
    np.testing.assert_allclose(cur, o.magnetization(sp, mask, nbins), rtol=1e-12, atol=1e-14)
    re = o.energy(sp, mask, nbins)
    np.testing.assert_allclose(ecur, re, rtol=1e-12, atol=1e-12 * np.abs(re).max())
