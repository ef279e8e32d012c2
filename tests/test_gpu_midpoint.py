"""Device LLG midpoint integrator (sim::LLG_Midpoint, src/simulate/LLGMidpoint.cpp) against the
reference's own trajectories (tests/golden/ref_*_mid.npz) and the oracle."""
import numpy as np
import pytest

import refcases as RC
from oracle import oracle as O
from vampire_amd import system as S
from vampire_amd.context import FP32, FP64, FUSED, PLAIN, Context

pytestmark = pytest.mark.gpu

MID_CASES = ["sc6_uniaxial_mid", "film_dmi_mid"]


@pytest.mark.parametrize("keep", [False, True])
@pytest.mark.parametrize("topology", ["csr", "ell", "stencil"])
@pytest.mark.parametrize("name", MID_CASES)
def test_midpoint_matches_reference(gpu, name, topology, keep):
    """Device FP64 fed the reference's noise stream reproduces sim::LLG_Midpoint: <= 1e-10 per
    spin, for every neighbour-list layout, with the external field regenerated or kept."""
    f = RC.load(name)
    s = RC.system_of(name)
    T, H = RC.conditions(name)
    seed = int(f["noise_seed"])
    with Context(s, precision=FP64, kernel=FUSED, topology=topology, keep_external_field=keep) as c:
        c.set_conditions(T, H)
        c.set_spins(f["spins_in"])
        last = 0
        for k, st in enumerate(int(v) for v in f["snap_steps"]):
            noise = np.stack([O.thermal_noise(seed, s.n, q) for q in range(last + 1, st + 1)])
            c.llg_midpoint(st - last, noise=noise)
            last = st
            assert np.abs(c.get_spins() - f["snaps"][k]).max() <= 1e-10


def test_midpoint_philox_matches_oracle(gpu):
    """C0 (SC 20^3, 300 K), 500 midpoint steps with the device's Philox noise against the
    oracle's stream: <= 1e-10 per spin (FP64); the run shares the step counter with the Heun
    integrator, so Heun steps before and after use the next Philox steps."""
    s = S.config_c0()
    spins = S.random_spins(s.n, 1)
    o = O.Oracle(s)
    o.set_conditions(300.0)
    ref = o.llg_heun(spins, 20, seed=7, step0=0)
    ref = o.llg_midpoint(ref, 500, seed=7, step0=20)
    ref = o.llg_heun(ref, 20, seed=7, step0=520)
    with Context(s, precision=FP64, seed=7) as c:
        c.set_conditions(300.0)
        c.set_spins(spins)
        c.llg_heun(20)
        c.llg_midpoint(500)
        c.llg_heun(20)
        dev = c.get_spins()
        assert c.step == 540
    assert np.abs(dev - ref).max() <= 1e-10


def test_midpoint_fp32_one_step(gpu):
    """FP32 midpoint fed the same injected normals as the FP64 oracle: one step and ten steps
    agree to single-precision rounding (a wrong term would show at dt |H| ~ 1e-2)."""
    s = S.config_c0()
    spins = S.random_spins(s.n, 3)
    o = O.Oracle(s)
    o.set_conditions(300.0)
    noise = np.stack([O.thermal_noise(5, s.n, q) for q in range(10)])
    with Context(s, precision=FP32, seed=5) as c:
        c.set_conditions(300.0)
        c.set_spins(spins)
        c.llg_midpoint(1, noise=noise[:1])
        one = c.get_spins()
        c.llg_midpoint(9, noise=noise[1:])
        ten = c.get_spins()
    assert np.abs(one - o.llg_midpoint(spins, 1, noise=noise[:1])).max() <= 2e-6
    assert np.abs(ten - o.llg_midpoint(spins, 10, noise=noise)).max() <= 2e-5


def test_midpoint_fp32_tracks_fp64_like_heun(gpu):
    """200 steps at 300 K on the device's Philox stream: the FP32 midpoint run departs from the
    FP64 one by no more than the FP32 Heun run departs from FP64 Heun under the same conditions
    (times 4; both grow chaotically from rounding and single-precision normals), and |S| stays 1
    to 1e-5 without renormalisation."""
    s = S.config_c0()
    spins = S.random_spins(s.n, 3)
    out = {}
    for integ in ("heun", "mid"):
        for prec in (FP64, FP32):
            with Context(s, precision=prec, seed=5) as c:
                c.set_conditions(300.0)
                c.set_spins(spins)
                (c.llg_heun if integ == "heun" else c.llg_midpoint)(200)
                out[integ, prec] = c.get_spins()
    d_heun = np.abs(out["heun", FP32] - out["heun", FP64]).max()
    d_mid = np.abs(out["mid", FP32] - out["mid", FP64]).max()
    assert d_mid <= 4 * d_heun + 1e-6, (d_mid, d_heun)
    assert np.abs(np.linalg.norm(out["mid", FP32], axis=0) - 1.0).max() <= 1e-5


def test_midpoint_with_dipole_field(gpu):
    """The dipole field refreshes after midpoint steps too (sim::increment_time follows every
    integrator): 30 FP64 steps equal the oracle's midpoint run with the same field."""
    f = RC.load("dipole_sc10")
    s = RC.system_of("dipole_sc10")
    T, H = RC.conditions("dipole_sc10")
    o = O.Oracle(s)
    o.set_conditions(T, H)
    o.set_dipole(**RC.dipole_args("dipole_sc10"))
    ref = o.llg_midpoint(f["spins_in"], 30, seed=4, step0=0)
    with Context(s, precision=FP64, kernel=PLAIN, seed=4) as c:
        c.set_conditions(T, H)
        c.set_dipole(num_cells=int(f["num_cells"]), local_cell=f["local_cell"], atom_cell=f["atom_cell"],
                     volume=f["volume"], tensor=f["tensor"], update_rate=int(f["update_rate"]), spins=f["spins_in"])
        c.llg_midpoint(30)
        dev = c.get_spins()
    assert np.abs(dev - ref).max() <= 1e-10


def test_midpoint_replicas_equal_standalone(gpu):
    """A batch of uneven replicas at different temperatures and fields, integrated with the
    midpoint scheme in one launch, equals each replica run alone with key seed + r, bit for bit."""
    systems = [S.ferromagnet("sc", L) for L in (3, 5, 4)]
    batch = S.concat_systems(systems)
    spins = [S.random_spins(x.n, 20 + r) for r, x in enumerate(systems)]
    temps = [100.0, 300.0, 600.0]
    with Context(batch, precision=FP64, seed=60) as c:
        for r, T in enumerate(temps):
            c.set_conditions(T, (0.0, 0.0, 0.5 * r), replica=r)
        c.set_spins(np.concatenate(spins, axis=1))
        c.llg_midpoint(40)
        got = c.get_spins()
    off = batch.replica_offset
    for r, x in enumerate(systems):
        with Context(x, precision=FP64, seed=60 + r) as c:
            c.set_conditions(temps[r], (0.0, 0.0, 0.5 * r))
            c.set_spins(spins[r])
            c.llg_midpoint(40)
            assert np.array_equal(c.get_spins(), got[:, off[r]:off[r + 1]])


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_midpoint_resume_is_bit_identical(gpu, precision, tmp_path):
    """A checkpoint taken between midpoint steps resumes to the same spins and step counter."""
    s = S.ferromagnet("fcc", 5)
    sp = S.random_spins(s.n, 6)
    with Context(s, precision=precision, seed=13) as a:
        a.set_conditions(500.0, (0.0, 0.2, 0.0))
        a.set_spins(sp)
        a.llg_midpoint(15)
        a.save_checkpoint(str(tmp_path / "m.chk"))
        a.llg_midpoint(15)
        ref = a.get_spins()
    with Context(s, precision=precision, seed=13) as b:
        b.load_checkpoint(str(tmp_path / "m.chk"))
        assert b.step == 15
        b.llg_midpoint(15)
        assert np.array_equal(b.get_spins(), ref)
