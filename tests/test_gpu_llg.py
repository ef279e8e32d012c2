"""Device parity of the LLG-Heun path (libvmx.so through the C ABI) against the oracle and the
reference golden vectors.  Tolerances are stated per test; FP64 kernels compute in the reference's
operation order without FMA contraction, so most comparisons are exact or at the ulp level."""
import numpy as np
import pytest

import refcases as RC
from oracle import oracle as O
from vampire_amd import system as S
from vampire_amd.context import FP32, FP64, FUSED, PLAIN, Context

pytestmark = pytest.mark.gpu

GOLDEN_CASES = ["sc6_uniaxial", "fcc3_cubic", "bcc4_rotated", "film_dmi", "sc6_vectorial", "film_vectorial",
                "film_local"]


TOPOLOGIES = ["csr", "ell", "stencil"]


# Fused-step variants: "fused" keeps dS for the corrector and regenerates H_ext there (the
# default); "fused_keep" also keeps H_ext; "fused_recompute" keeps nothing (dS recomputed too).
KERNELS = [PLAIN, FUSED, "fused_keep", "fused_recompute"]


def _ctx(system, precision=FP64, kernel=FUSED, seed=0, block=0, topology="auto"):
    return Context(system, precision=precision, kernel=PLAIN if kernel == PLAIN else FUSED, seed=seed,
                   block_size=block, topology=topology, recompute_euler=kernel == "fused_recompute",
                   keep_external_field=kernel == "fused_keep")


@pytest.mark.parametrize("topology", TOPOLOGIES)
def test_topology_is_honoured(gpu, topology):
    s = S.ferromagnet("fcc", 5)
    with _ctx(s, topology=topology) as c:
        assert c.topology[0] == topology
    with _ctx(s) as c:   # a periodic lattice: automatic choice is the stencil layout
        name, npat = c.topology
        assert name == "stencil" and 4 <= npat <= 4 * 27


def test_device_and_library(gpu):
    from vampire_amd import context
    assert context.device_count() >= 1


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_philox_noise_matches_oracle(gpu, precision):
    s = S.ferromagnet("sc", 6)
    with _ctx(s, precision=precision, seed=12345) as c:
        for step in (0, 1, 7, 2**33 + 5):
            dev = c.thermal_noise(step)
            ref = O.thermal_noise(12345, s.n, step)
            # FP64: same uniforms, device log/sincos within a few ulp of libm;
            # FP32: the same uniforms rounded to float, hardware log2/sin/cos: below 2e-5
            tol = 1e-13 if precision == FP64 else 2e-5
            assert np.abs(dev - ref).max() <= tol * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_fields_match_reference(gpu, name):
    f = RC.load(name)
    s = RC.system_of(name)
    T, H = RC.conditions(name)
    with _ctx(s) as c:
        c.set_conditions(T, H, **RC.flags(name))
        c.set_spins(f["spins_in"])
        hs, he = c.fields(noise=O.thermal_noise(int(f["noise_seed"]), s.n, 0))
    ref = f["fields"]
    # bit-exact expected; allow 1 ulp of the field scale
    scale = np.abs(ref).max()
    assert np.abs(hs - ref[:3]).max() <= 2.3e-16 * scale
    assert np.abs(he - ref[3:]).max() <= 2.3e-16 * scale


@pytest.mark.parametrize("topology", TOPOLOGIES)
@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_trajectory_matches_reference(gpu, name, kernel, topology):
    """Device FP64 fed the reference's noise stream reproduces sim::LLG_Heun: <= 1e-10 per spin,
    for every neighbour-list layout."""
    f = RC.load(name)
    s = RC.system_of(name)
    T, H = RC.conditions(name)
    seed = int(f["noise_seed"])
    steps = [int(v) for v in f["snap_steps"]]
    with _ctx(s, kernel=kernel, topology=topology) as c:
        c.set_conditions(T, H, **RC.flags(name))
        c.set_spins(f["spins_in"])
        last = 0
        for k, st in enumerate(steps):
            noise = np.stack([O.thermal_noise(seed, s.n, q) for q in range(last + 1, st + 1)])
            c.llg_heun(st - last, noise=noise)
            last = st
            assert np.abs(c.get_spins() - f["snaps"][k]).max() <= 1e-10


@pytest.mark.parametrize("kernel", KERNELS)
def test_c0_1000_steps_philox(gpu, kernel):
    """C0 (SC 20^3, 300 K): device Philox noise vs the oracle's Philox stream, 1000 steps,
    per-spin agreement <= 1e-10 (FP64)."""
    s = S.config_c0()
    spins = S.random_spins(s.n, 1)
    o = O.Oracle(s)
    o.set_conditions(300.0)
    ref = o.llg_heun(spins, 1000, seed=7, step0=0)
    with _ctx(s, kernel=kernel, seed=7) as c:
        c.set_conditions(300.0)
        c.set_spins(spins)
        c.llg_heun(1000)
        dev = c.get_spins()
    assert np.abs(dev - ref).max() <= 1e-10


def test_c0_golden_1000_steps(gpu):
    """Against the reference run itself: C0 after 1000 steps with the injected stream."""
    f = RC.load("c0_sc20")
    s = S.config_c0()
    with _ctx(s) as c:
        c.set_conditions(300.0)
        c.set_spins(f["spins_in"])
        for blk in range(10):
            noise = np.stack([O.thermal_noise(1, s.n, q) for q in range(blk * 100 + 1, blk * 100 + 101)])
            c.llg_heun(100, noise=noise)
        assert np.abs(c.get_spins() - f["snaps"][0]).max() <= 1e-10


@pytest.mark.parametrize("kernel", [FUSED, PLAIN])
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_fp32_matches_reference_same_normals(gpu, name, kernel):
    """FP32 spins and fields fed the normals the reference consumed: the first snapshot (one
    step, or the first ten) within 2e-6 per spin component of the reference's FP64 spins, every
    later snapshot within 5e-5 (FP32 rounding, 6e-8 per operation, grows with the steps), and
    |S| = 1 to 1e-6."""
    f = RC.load(name)
    s = RC.system_of(name)
    T, H = RC.conditions(name)
    seed = int(f["noise_seed"])
    with _ctx(s, precision=FP32, kernel=kernel) as c:
        c.set_conditions(T, H, **RC.flags(name))
        c.set_spins(f["spins_in"])
        last = 0
        for k, st in enumerate(int(v) for v in f["snap_steps"]):
            noise = np.stack([O.thermal_noise(seed, s.n, q) for q in range(last + 1, st + 1)])
            c.llg_heun(st - last, noise=noise)
            last = st
            got = c.get_spins()
            err = np.abs(got - f["snaps"][k]).max()
            assert err <= (2e-6 if st <= 10 else 5e-5), (st, err)
            assert np.abs(np.linalg.norm(got, axis=0) - 1.0).max() <= 1e-6


def test_fp64_normals_mode_feeds_fp32_the_fp64_stream(gpu):
    """fp64_normals: an FP32 context draws the normals of an FP64 context with the same seed.
    One step from the same state then differs by FP32 arithmetic alone (<= 2e-6), and the FP32
    generator's own normals stay within 2e-6 of them in that step (full-width uniforms)."""
    s = S.config_c0()
    spins = S.random_spins(s.n, 3)
    out = {}
    for key, prec, f64n in (("fp64", FP64, False), ("fp32_same", FP32, True), ("fp32_own", FP32, False)):
        c = Context(s, precision=prec, seed=9, fp64_normals=f64n)
        with c:
            c.set_conditions(300.0)
            c.set_spins(spins)
            c.llg_heun(1)
            out[key] = c.get_spins()
            if key == "fp32_same":
                with pytest.raises(RuntimeError):
                    c.llg_heun_graph(2, 1)
    assert np.abs(out["fp32_same"] - out["fp64"]).max() <= 2e-6
    assert np.abs(out["fp32_own"] - out["fp64"]).max() <= 2e-6


def test_fp32_drift_same_normals_1000_steps(gpu):
    """FP32 against FP64 over 1000 steps of C0 (SC 20^3, 300 K) with the same normals.  The
    trajectories are chaotic: one 2^-24 perturbation of the FP64 start grows ~700-fold in 1000
    steps.  So the per-spin floor of any FP32 spin storage is what FP64 arithmetic with the spins
    rounded to float after every step reaches, and the FP32 step stays within 2x of it; the
    magnetisation agrees to 1e-5 of Ms (measured ~3e-7)."""
    s = S.config_c0()
    spins = S.random_spins(s.n, 3)
    rng = np.random.default_rng(1)
    pert = spins * (1.0 + rng.choice([-1.0, 1.0], spins.shape) * 2.0 ** -24)
    pert /= np.linalg.norm(pert, axis=0)
    out = {}
    for key, prec, start in (("fp64", FP64, spins), ("fp64_pert", FP64, pert), ("fp32", FP32, spins)):
        with Context(s, precision=prec, seed=9, fp64_normals=prec == FP32) as c:
            c.set_conditions(300.0)
            c.set_spins(start)
            c.llg_heun(1000)
            out[key] = c.get_spins()
    with Context(s, precision=FP64, seed=9) as c:      # FP64 arithmetic, spins stored as float
        c.set_conditions(300.0)
        c.set_spins(spins.astype(np.float32).astype(np.float64))
        for _ in range(1000):
            c.llg_heun(1)
            c.set_spins(c.get_spins().astype(np.float32).astype(np.float64))
        out["fp64_store32"] = c.get_spins()
    d32 = np.abs(out["fp32"] - out["fp64"]).max()
    dpert = np.abs(out["fp64_pert"] - out["fp64"]).max()
    dstore = np.abs(out["fp64_store32"] - out["fp64"]).max()
    dM = np.abs(out["fp32"].mean(axis=1) - out["fp64"].mean(axis=1)).max()
    print(f"fp32 same-normals max|dm| {d32:.3e}; fp64 one perturbation {dpert:.3e}; "
          f"fp64 stored as float {dstore:.3e}; dM/Ms {dM:.3e}")
    assert dpert >= 100.0 * 2.0 ** -24          # the growth that makes the floor
    assert dM <= 1e-5
    assert d32 <= 2.0 * dstore, (d32, dstore)
    assert np.abs(np.linalg.norm(out["fp32"], axis=0) - 1.0).max() < 1e-6


def test_graph_equals_eager(gpu):
    """Graph replay equals eager steps; a recorded graph is dropped when the time step or the
    field flags (kernel arguments) change, so later replays integrate with the new ones."""
    s = S.ferromagnet("fcc", 6)
    spins = S.random_spins(s.n, 4)
    res = []
    for use_graph in (False, True):
        with _ctx(s, seed=3) as c:
            c.set_conditions(400.0, (0.0, 0.0, 1.0))
            c.set_spins(spins)
            if use_graph:
                c.llg_heun_graph(10, 3)
            else:
                c.llg_heun(30)
            c.set_time_step(0.5 * s.dt)
            c.set_conditions(400.0, (0.0, 0.0, 1.0), thermal=False)
            if use_graph:
                c.llg_heun_graph(10, 2)
            else:
                c.llg_heun(20)
            c.set_conditions(200.0, (0.0, 0.5, 1.0))
            if use_graph:
                c.llg_heun_graph(10, 1)
            else:
                c.llg_heun(10)
            res.append(c.get_spins())
            assert c.step == 60
    assert np.array_equal(res[0], res[1])


def test_eager_and_graph_interleaved_equal_eager(gpu):
    """Eager calls carry their step as an offset from the device step counter and leave it where
    it is; a graph replay reads and advances the counter itself.  Eager calls, replays and a jump
    of the step counter interleaved equal the same steps taken eagerly, bit for bit, with the
    same thermal noise (the step is the Philox counter), in both precisions."""
    s = S.ferromagnet("fcc", 6)
    spins = S.random_spins(s.n, 9)
    for precision in (FP64, FP32):
        res = []
        for use_graph in (False, True):
            with _ctx(s, seed=5, precision=precision) as c:
                def tens(n):     # n steps, as graph replays of 10 or eagerly
                    if use_graph:
                        c.llg_heun_graph(10, n // 10)
                    else:
                        c.llg_heun(n)
                c.set_conditions(500.0, (0.0, 0.0, 0.5))
                c.set_spins(spins)
                c.llg_heun(7)
                tens(20)
                c.llg_heun(5)
                tens(10)
                assert c.step == 42
                c.step = 1000
                c.llg_heun(3)
                tens(10)
                c.llg_heun(1)
                assert c.step == 1014
                res.append(c.get_spins())
        assert np.array_equal(res[0], res[1])


def test_repeated_and_changed_conditions_match_oracle(gpu):
    """gpu::llg_heun sets the conditions before every one-step call.  Repeated values write
    nothing (the device table already holds them), changed ones are written: a per-step sequence of
    temperatures and fields with repeats, on/off thermal flags included, equals the oracle stepped
    with the same sequence (FP64, Philox noise at the same step counters, <= 1e-12)."""
    s = S.ferromagnet("fcc", 5)
    spins = S.random_spins(s.n, 12)
    seq = [(300.0, (0.0, 0.0, 1.0), True), (300.0, (0.0, 0.0, 1.0), True), (700.0, (0.0, 0.0, 1.0), True),
           (700.0, (0.0, 0.0, 1.0), True), (700.0, (0.0, 0.5, 1.0), True), (700.0, (0.0, 0.5, 1.0), False),
           (700.0, (0.0, 0.5, 1.0), False), (200.0, (0.0, 0.5, 1.0), True), (200.0, (0.0, 0.5, 1.0), True)]
    o = O.Oracle(s)
    ref = spins
    for k, (T, H, th) in enumerate(seq):
        o.set_conditions(T, H, thermal=th)
        ref = o.llg_heun(ref, 1, seed=21, step0=k)
    with _ctx(s, seed=21) as c:
        c.set_spins(spins)
        for T, H, th in seq:
            c.set_conditions(T, H, thermal=th)
            c.llg_heun(1)
        assert np.abs(c.get_spins() - ref).max() <= 1e-12


def test_step_timing_of_calls(gpu):
    """Calls of two or more steps and graph replays are timed with events on the stream; a call of
    one step is not (its events would cost more than its launches) and reports 0, not the timing
    of an earlier call."""
    s = S.ferromagnet("fcc", 6)
    with _ctx(s, seed=2, precision=FP32) as c:
        c.set_conditions(300.0)
        c.set_spins(S.random_spins(s.n, 2))
        c.llg_heun(20)
        assert c.last_step_time_ms() > 0.0
        c.llg_heun(1)
        assert c.last_step_time_ms() == 0.0
        c.llg_heun_graph(4, 2)
        assert c.last_step_time_ms() > 0.0
        c.llg_midpoint(1)
        assert c.last_step_time_ms() == 0.0
        c.llg_midpoint(3)
        assert c.last_step_time_ms() > 0.0


def test_replicas_equal_standalone(gpu):
    """A batch of uneven replicas equals each replica run alone with seed + r."""
    sizes = [3, 5, 4]
    systems = [S.ferromagnet("sc", L) for L in sizes]
    batch = S.concat_systems(systems)
    spins = [S.random_spins(x.n, 10 + r) for r, x in enumerate(systems)]
    temps = [100.0, 300.0, 600.0]
    with _ctx(batch, seed=50) as c:
        for r, T in enumerate(temps):
            c.set_conditions(T, (0.0, 0.0, 0.5 * r), replica=r)
        c.set_spins(np.concatenate(spins, axis=1))
        c.llg_heun(50)
        got = c.get_spins()
    off = batch.replica_offset
    for r, x in enumerate(systems):
        with _ctx(x, seed=50 + r) as c:
            c.set_conditions(temps[r], (0.0, 0.0, 0.5 * r))
            c.set_spins(spins[r])
            c.llg_heun(50)
            assert np.array_equal(c.get_spins(), got[:, off[r]:off[r + 1]])


def test_applied_field_precession_analytic(gpu):
    """Reference test tests/physical/AppliedField: single spin, T=0, H=10 T along z, alpha=1,
    S(0)=x; analytic Sx=sech(la)cos a, Sy=sech(la)sin a, Sz=tanh(la), a=t gamma H/(1+l^2)
    (applied_field_errors.py:12-25); tolerance 1e-6 (run_tests.sh:53).  256 copies of the spin
    with no neighbours run together."""
    n = 256
    mat = S.Material(mu_s_muB=1.72, alpha=1.0)
    dt = 1e-16 * S.GAMMA_SI
    sysm = S.System(n=n, type=np.zeros(n, np.int32), moment=np.full(n, 1.72), nbr_start=np.zeros(n, np.int32),
                    nbr_end=np.full(n, -1, np.int32), nbr=np.zeros(0, np.int32), nbr_iid=np.zeros(0, np.int32),
                    exchange_type=0, exchange=np.zeros((1, 1)), materials=[S.derived_material(mat, dt)],
                    anisotropy_flags=0, dt=dt, half_dt=0.5 * dt)
    spins = np.zeros((3, n))
    spins[0] = 1.0
    with _ctx(sysm) as c:
        c.set_conditions(0.0, (0.0, 0.0, 10.0))
        c.set_spins(spins)
        worst = 0.0
        for k in range(1, 101):
            c.llg_heun(100)
            t = k * 100 * 1e-16
            lam = 1.0
            a = t * 1.76e11 * 10.0 / (1 + lam * lam)
            exp = np.array([np.cos(a) / np.cosh(lam * a), np.sin(a) / np.cosh(lam * a), np.tanh(lam * a)])
            worst = max(worst, np.abs(c.get_spins()[:, 0] - exp).max())
    assert worst < 1e-6


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("dims,lds", [((16, 16, 4), True), ((24, 24, 6), False)])
def test_ragged_stencils_small_and_large_tables(gpu, dims, lds, kernel):
    """Diluted two-material films have ragged rows and hundreds to thousands of stencils: the
    fused kernels read the offset table from LDS (<= 4096 entries) or from global memory.  Both
    reproduce the oracle fed the same noise, FP64, <= 1e-10 per spin over 20 steps."""
    nx, ny, nz = dims
    s = S.config_c4(nx, ny, nz, dilution=0.1, seed=3, dmi=0.0)
    spins = S.random_spins(s.n, 5)
    noise = np.stack([O.thermal_noise(11, s.n, q) for q in range(20)])
    o = O.Oracle(s)
    o.set_conditions(400.0, (0.0, 0.0, 0.5))
    ref = o.llg_heun(spins, 20, noise=noise)
    with _ctx(s, kernel=kernel, topology="stencil") as c:
        name, npat = c.topology
        assert name == "stencil"
        assert (npat * 6 <= 4096) == lds
        c.set_conditions(400.0, (0.0, 0.0, 0.5))
        c.set_spins(spins)
        c.llg_heun(20, noise=noise)
        assert np.abs(c.get_spins() - ref).max() <= 1e-10


@pytest.mark.parametrize("kernel", KERNELS)
def test_thin_film_demag_matches_reference(gpu, kernel):
    """The thin-film demag field (vmx_set_demag_field; calculate_applied_fields,
    src/simulate/fields.cpp:219-239) set at every step to the field the reference formed from its
    system magnetisation statistic: the device step reproduces the reference's trajectory
    (<= 1e-10, FP64), fields included, for every kernel variant."""
    f = RC.load("film_demag")
    s = RC.system_of("film_demag")
    T, H = RC.conditions("film_demag")
    hd = f["demag_field"]
    seed = int(f["noise_seed"])
    snaps = [int(v) for v in f["snap_steps"]]
    with _ctx(s, kernel=kernel) as c:
        c.set_conditions(T, H)
        c.set_spins(f["spins_in"])
        c.set_demag_field(hd[0])
        hs, he = c.fields(noise=O.thermal_noise(seed, s.n, 0))
        scale = np.abs(f["fields"]).max()
        assert np.abs(he - f["fields"][3:]).max() <= 2.3e-16 * scale
        k = 0
        for st in range(1, snaps[-1] + 1):
            c.set_demag_field(hd[st - 1])
            c.llg_heun(1, noise=O.thermal_noise(seed, s.n, st)[None])
            if st == snaps[k]:
                assert np.abs(c.get_spins() - f["snaps"][k]).max() <= 1e-10, st
                k += 1
