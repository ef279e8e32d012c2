"""Device macrocell dipole field (tensor solver) against the oracle, itself equal bit for bit to the
reference's dipole::calculate_field on the same cells and tensor (tests/golden/ref_dipole_sc10.npz:
SC 10x10x6, open boundaries, two materials, 75 occupied cells of 100, refresh every 10 steps)."""
import numpy as np
import pytest

import refcases as RC
from oracle import oracle as O
from vampire_amd._lib import VmxError
from vampire_amd.context import FP32, FP64, FUSED, PLAIN, Context

pytestmark = pytest.mark.gpu

NAME = "dipole_sc10"


def _dipole_kwargs(f):
    return dict(num_cells=int(f["num_cells"]), local_cell=f["local_cell"], atom_cell=f["atom_cell"],
                volume=f["volume"], tensor=f["tensor"], update_rate=int(f["update_rate"]))


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_dipole_field_equals_reference(gpu, precision):
    """Cell moments, cell fields and atom fields of the device for the reference's two spin
    sets: FP64 bit for bit; FP32 (single-precision spins summed in double) within 1e-6 of the
    largest field."""
    f = RC.load(NAME)
    s = RC.system_of(NAME)
    with Context(s, precision=precision, seed=1) as c:
        c.set_dipole(**_dipole_kwargs(f), spins=f["spins_in"])
        for k, sp in enumerate(("spins_in", "spins2")):
            if k:
                c.set_spins(f[sp])
                c.update_dipole()
            B, Hd = c.dipole_field()
            if precision == FP64:
                assert np.array_equal(B, f[f"atom_B{k}"]) and np.array_equal(Hd, f[f"atom_Hd{k}"])
            else:
                scale = np.abs(f[f"atom_Hd{k}"]).max()
                assert np.abs(B - f[f"atom_B{k}"]).max() <= 1e-6 * scale
                assert np.abs(Hd - f[f"atom_Hd{k}"]).max() <= 1e-6 * scale


@pytest.mark.parametrize("kernel", [FUSED, PLAIN])
def test_dipole_trajectory_equals_oracle(gpu, kernel):
    """30 FP64 Heun steps at 300 K with three refreshes (steps 10, 20, 30): the device equals the
    oracle to 1e-10 per spin component, while dropping the dipole field moves the spins by far
    more than that."""
    f = RC.load(NAME)
    s = RC.system_of(NAME)
    T, H = RC.conditions(NAME)
    seed = 21
    o = O.Oracle(s)
    o.set_conditions(T, H)
    o.set_dipole(**RC.dipole_args(NAME))
    ref = o.llg_heun(f["spins_in"], 30, seed=seed, step0=0)
    o_plain = O.Oracle(s)
    o_plain.set_conditions(T, H)
    no_dip = o_plain.llg_heun(f["spins_in"], 30, seed=seed, step0=0)
    with Context(s, precision=FP64, kernel=kernel, seed=seed) as c:
        c.set_conditions(T, H)
        c.set_dipole(**_dipole_kwargs(f), spins=f["spins_in"])
        c.llg_heun(30)
        dev = c.get_spins()
        B, _ = c.dipole_field()
    assert np.abs(dev - ref).max() <= 1e-10
    assert np.abs(no_dip - ref).max() > 1e-6
    # the field held after the run is the one of the spins at step 30 (refreshed there)
    o.dipole_update(dev)
    assert np.abs(B - o.dipole_fields()[0]).max() <= 1e-12 * np.abs(B).max()


def test_dipole_refresh_cadence(gpu):
    """The field changes only at multiples of the update rate: after 9 steps it is still the
    time-0 field; after the 10th it is the field of the step-10 spins."""
    f = RC.load(NAME)
    s = RC.system_of(NAME)
    T, H = RC.conditions(NAME)
    o = O.Oracle(s)
    with Context(s, precision=FP64, seed=5) as c:
        c.set_conditions(T, H)
        c.set_dipole(**_dipole_kwargs(f), spins=f["spins_in"])
        B0, _ = c.dipole_field()
        c.llg_heun(9)
        assert np.array_equal(c.dipole_field()[0], B0)
        c.llg_heun(1)
        B10, _ = c.dipole_field()
        o.set_dipole(**dict(RC.dipole_args(NAME), spins=c.get_spins()))
        assert not np.array_equal(B10, B0)
        assert np.array_equal(B10, o.dipole_fields()[0])


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_uploaded_atom_field_equals_device_solver(gpu, precision):
    """The per-atom upload used by the gpu:: transfer hook gives the same first 9 steps as the
    device solver holding the same time-0 field (no refresh before step 10)."""
    f = RC.load(NAME)
    s = RC.system_of(NAME)
    T, H = RC.conditions(NAME)
    out, B0 = [], None
    for native in (True, False):
        with Context(s, precision=precision, seed=8) as c:
            c.set_conditions(T, H)
            c.set_spins(f["spins_in"])
            if native:
                c.set_dipole(**_dipole_kwargs(f))
                B0 = c.dipole_field()[0]
            else:
                c.set_atom_field(B0)
            c.llg_heun(9)
            out.append(c.get_spins())
    assert np.array_equal(out[0], out[1])


def test_dipole_metropolis_matches_oracle(gpu):
    """FP64 checkerboard sweeps with the device dipole field: the single-spin energy takes
    -mu0Hd.S last (sim::calculate_spin_energy) and the field is refreshed every 10 sweeps, as
    sim::increment_time does after a Monte Carlo step; 25 sweeps equal the oracle's colour-ordered
    Metropolis to 1e-12, with the same rejection count and cone width."""
    f = RC.load(NAME)
    s = RC.system_of(NAME)
    o = O.Oracle(s)
    o.set_conditions(300.0, (0.0, 0.0, 0.2))
    o.set_dipole(**RC.dipole_args(NAME))
    ref, sigma, moves, rej = o.mc_checkerboard(f["spins_in"], 25, seed=77, sweep0=0, sigma=60.0)
    with Context(s, precision=FP64, seed=77) as c:
        c.set_conditions(300.0, (0.0, 0.0, 0.2))
        c.set_dipole(**_dipole_kwargs(f), spins=f["spins_in"])
        c.mc_sweep(25)
        dev = c.get_spins()
        dsig, dmov, drej = c.mc_state()
        B, _ = c.dipole_field()
    assert np.abs(dev - ref).max() <= 1e-12
    assert dmov[0] == moves and drej[0] == rej
    assert abs(dsig[0] - sigma) <= 1e-12 * sigma
    # both refreshed from their sweep-20 spins
    assert np.abs(B - o.dipole_fields()[0]).max() <= 1e-12 * np.abs(B).max()


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_dipole_energy_statistics(gpu, precision):
    """Energy statistics with the device dipole field: the magnetostatic row,
    1/2 sum -mu0Hd.S mu per mask (statistics/energy.cpp:198-208), is non-zero and equals the
    oracle's, as do the total and the other terms (1e-12 FP64, 1e-6 FP32, relative)."""
    f = RC.load(NAME)
    s = RC.system_of(NAME)
    T, H = RC.conditions(NAME)
    mask = np.ascontiguousarray(s.type, dtype=np.int32)     # one mask per material, id 2 empty
    o = O.Oracle(s)
    o.set_conditions(T, H)
    o.set_dipole(**RC.dipole_args(NAME))
    re = o.energy(f["spins_in"], mask, 3)
    with Context(s, precision=precision, seed=3) as c:
        c.set_conditions(T, H)
        c.set_dipole(**_dipole_kwargs(f), spins=f["spins_in"])
        he = c.add_energy(mask, 3)
        c.stats_update()
        ecur, _, _ = c.energy(he)
    assert np.abs(re[4, :2]).min() > 0
    tol = 1e-12 if precision == FP64 else 1e-6
    np.testing.assert_allclose(ecur, re, rtol=tol, atol=tol * np.abs(re).max())


def test_dipole_refusals(gpu):
    """Paths that would not refresh or account for the field refuse with an error; clearing the
    field restores them."""
    f = RC.load(NAME)
    s = RC.system_of(NAME)
    with Context(s, precision=FP64, seed=2) as c:
        c.set_conditions(300.0)
        c.set_dipole(**_dipole_kwargs(f), spins=f["spins_in"])
        with pytest.raises(VmxError, match="dipole"):
            c.llg_heun_graph(10, 1)
        with pytest.raises(VmxError, match="dipole"):
            c.checkpoint()
        with Context(s, precision=FP64, seed=2) as plain:
            chk = plain.checkpoint()
        with pytest.raises(VmxError, match="dipole"):
            c.restore(chk)
        with pytest.raises(VmxError, match="dipole solver"):
            c.set_atom_field(f["atom_B0"])
        bad = dict(_dipole_kwargs(f))
        bad["atom_cell"] = np.full(s.n, int(f["num_cells"]), dtype=np.int32)
        with pytest.raises(VmxError, match="cell"):
            c.set_dipole(**bad)
        c.clear_dipole()
        c.llg_heun_graph(10, 1)
        assert len(c.checkpoint()) > 0
        # an uploaded field has no demag field: Metropolis and device energies refuse
        c.set_atom_field(f["atom_B0"])
        with pytest.raises(VmxError, match="demag"):
            c.mc_sweep(1)
        c.add_energy(np.zeros(s.n, np.int32), 2)
        with pytest.raises(VmxError, match="demag"):
            c.stats_update()
        c.set_atom_field(None)
        c.mc_sweep(1)
        c.stats_update()
