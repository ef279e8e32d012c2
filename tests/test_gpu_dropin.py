"""The drop-in: a VAMPIRE executable built from the reference sources with its src/gpu dispatchers
replaced by vampire_amd/shim/gpu_shim.cpp over libvmx.so.  The same executable runs an input deck
with gpu:acceleration=off (the reference CPU integrator) and =on (the device path through the
gpu:: interface); the binary spin snapshots must agree.  The executable is linked in the
development container (make -C vampire_amd/shim); without it these tests are skipped."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "vampire_amd", "shim", "_build", "vampire-vmx")

CO = """material:num-materials=1
material[1]:material-name=Co
material[1]:damping-constant={alpha}
material[1]:exchange-matrix[1]=6.064e-21
material[1]:atomic-spin-moment=1.72 !muB
material[1]:second-order-uniaxial-anisotropy-constant={ku}
material[1]:initial-spin-direction={init}
"""

LATTICE = """create:crystal-structure={crystal}
create:periodic-boundaries-x
create:periodic-boundaries-y
create:periodic-boundaries-z
dimensions:unit-cell-size = 3.54 !A
dimensions:system-size-x = {size} !A
dimensions:system-size-y = {size} !A
dimensions:system-size-z = {size} !A
"""

RUN = """material:file=Co.mat
sim:temperature={T}
sim:equilibration-time-steps=0
sim:time-steps-increment={inc}
sim:total-time-steps={steps}
sim:time-step=1.0E-16
sim:applied-field-strength={H} !T
sim:applied-field-unit-vector={Hdir}
sim:program=time-series
sim:integrator=llg-heun
gpu:acceleration={acc}
config:atoms=end
config:output-format=binary
output:time-steps
output:magnetisation
"""


def run_deck(tmp, name, head, mat, **kw):
    d = tmp / name
    d.mkdir()
    (d / "input").write_text(head + RUN.format(**kw))
    (d / "Co.mat").write_text(mat)
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    raw = (d / "spins-00000000.data").read_bytes()
    n = np.frombuffer(raw[:8], dtype=np.uint64)[0]
    return np.frombuffer(raw[8:], dtype=np.float64).copy(), int(n)


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        pytest.skip("drop-in VAMPIRE executable not built (make -C vampire_amd/shim)")
    return EXE


@pytest.mark.parametrize("crystal,size", [("sc", 21.23), ("fcc", 10.61)])
def test_dropin_gpu_equals_reference_cpu(gpu, exe, tmp_path, crystal, size):
    """T = 0, exchange + uniaxial anisotropy + 0.5 T field from the reference's own random initial
    spins, 1000 steps: device path through gpu:: == reference sim::LLG_Heun (<= 1e-10)."""
    head = LATTICE.format(crystal=crystal, size=size)
    mat = CO.format(alpha=0.1, ku=4.644e-24, init="random")
    kw = dict(T=0.0, inc=100, steps=1000, H=0.5, Hdir="0,0,1")
    cpu, n1 = run_deck(tmp_path, "cpu", head, mat, acc="off", **kw)
    dev, n2 = run_deck(tmp_path, "gpu", head, mat, acc="on", **kw)
    assert n1 == n2 and n1 > 100
    assert np.abs(dev - cpu).max() <= 1e-10


def test_dropin_applied_field_physics(gpu, exe, tmp_path):
    """The reference's tests/physical/AppliedField deck (single spin, alpha = 1, 10 T along z,
    S0 = x) run through the device path: final spin against the analytic damped precession
    (tolerance 1e-6, run_tests.sh:53)."""
    head = "create:single-spin\n"
    mat = CO.format(alpha=1.0, ku=0.0, init="1,0,0").replace("exchange-matrix[1]=6.064e-21",
                                                             "exchange-matrix[1]=11.2e-21")
    spins, n = run_deck(tmp_path, "gpu", head, mat, acc="on", T=0.0, inc=100, steps=10000, H=10.0,
                        Hdir="0,0,1")
    a = 10000 * 1e-16 * 1.76e11 * 10.0 / 2.0
    exp = np.array([np.cos(a) / np.cosh(a), np.sin(a) / np.cosh(a), np.tanh(a)])
    assert n == 1
    assert np.abs(spins[:3] - exp).max() < 1e-6


def test_dropin_dipole_equals_reference_cpu(gpu, exe, tmp_path):
    """Macrocell dipole field (tensor solver, 2x2x2-cell macrocells, refreshed every 10 steps by
    dipole::calculate_field on the host and handed to the device through
    gpu::transfer_dipole_fields_from_cpu_to_gpu): SC 10x10x6 block with open boundaries, T = 0,
    0.5 T, 200 steps.  Device path == reference CPU (<= 1e-10); the same deck without the dipole
    field ends more than 1e-6 away, so the field is in the comparison."""
    head = ("create:crystal-structure=sc\ndimensions:unit-cell-size = 3.54 !A\n"
            "dimensions:system-size-x = 35.39 !A\ndimensions:system-size-y = 35.39 !A\n"
            "dimensions:system-size-z = 21.23 !A\n")
    dip = "dipole:solver=tensor\ndipole:field-update-rate=10\ncells:macro-cell-size = 7.08 !A\n"
    mat = CO.format(alpha=0.1, ku=4.644e-24, init="random")
    kw = dict(T=0.0, inc=100, steps=200, H=0.5, Hdir="0,0,1")
    cpu, n1 = run_deck(tmp_path, "cpu", head + dip, mat, acc="off", **kw)
    dev, n2 = run_deck(tmp_path, "gpu", head + dip, mat, acc="on", **kw)
    nodip, _ = run_deck(tmp_path, "nodip", head, mat, acc="off", **kw)
    assert n1 == n2 == 600
    assert np.abs(dev - cpu).max() <= 1e-10
    assert np.abs(nodip - cpu).max() > 1e-6


DECKS = os.path.join(ROOT, "tests", "golden", "decks")


def run_reference_deck(tmp, name, deck, acc, edits=()):
    """Runs one of the reference's own test decks (tests/golden/decks/<deck>: input and material
    file as the reference holds them) with gpu:acceleration appended and the given line edits;
    returns the rows of its `output` file."""
    d = tmp / name
    d.mkdir()
    src = os.path.join(DECKS, deck)
    text = open(os.path.join(src, "input")).read()
    for old, new in edits:
        assert old in text, old
        text = text.replace(old, new)
    (d / "input").write_text(text + f"\ngpu:acceleration={acc}\n")
    for f in os.listdir(src):
        if f != "input":
            (d / f).write_text(open(os.path.join(src, f)).read())
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=280)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = [[float(v) for v in line.split()] for line in (d / "output").read_text().splitlines()
            if line.strip() and not line.startswith("#")]
    return np.array(rows)


@pytest.mark.timeout(300)
def test_reference_curie_deck_on_device(gpu, exe, tmp_path):
    """The reference's tests/curie-temperature deck, unmodified but for gpu:acceleration=on (SC
    22^3 cells, J = 11.2e-21 J, 0-1800 K in 100 K steps, 10k + 30k Heun steps each), run by
    sim::run through gpu:: on the device.  Its output (columns: time, T, mean |m| over the
    temperature's loop, m direction and |m|, |m|, field alignment) must show the ferromagnet's
    M(T): |m| = 1 at 0 K, falling with T, crossing 1/2 below the infinite-lattice
    Tc = 1.443 J/kB = 1171 K and near zero above it."""
    out = run_reference_deck(tmp_path, "gpu", "curie-temperature", "on")
    T, m = out[:, 1], out[:, 6]
    assert np.array_equal(T, np.arange(0.0, 1801.0, 100.0))
    assert m[0] == pytest.approx(1.0, abs=1e-12)
    below = T <= 1200.0
    assert np.all(np.diff(m[below]) < 0.02)
    assert np.all(m[T >= 1400.0] < 0.05)
    k = int(np.nonzero((m[:-1] >= 0.5) & (m[1:] < 0.5))[0][0])
    t_half = T[k] + (m[k] - 0.5) * 100.0 / (m[k] - m[k + 1])
    assert 800.0 < t_half < 1171.0


@pytest.mark.timeout(300)
def test_reference_curie_deck_device_vs_cpu(gpu, exe, tmp_path):
    """A shortened copy of the same deck (5 nm cube, 0-1800 K in 300 K steps, 2000 + 4000 steps)
    run by the reference CPU integrator and by the device: their noise streams differ (MT19937 vs
    Philox), so the mean |m| over each temperature's loop (reset after equilibration by
    stats::mag_m_reset, src/program/curie_temperature.cpp:111) must agree statistically (0.03)."""
    edits = [("system-size-x = 7.7", "system-size-x = 5.0"), ("system-size-y = 7.7", "system-size-y = 5.0"),
             ("system-size-z = 7.7", "system-size-z = 5.0"), ("temperature-increment=100", "temperature-increment=300"),
             ("equilibration-time-steps=10000", "equilibration-time-steps=2000"),
             ("loop-time-steps=30000", "loop-time-steps=4000")]
    cpu = run_reference_deck(tmp_path, "cpu", "curie-temperature", "off", edits)
    dev = run_reference_deck(tmp_path, "gpu", "curie-temperature", "on", edits)
    assert np.array_equal(cpu[:, 1], dev[:, 1]) and len(cpu) == 7
    assert np.abs(cpu[:, 2] - dev[:, 2]).max() <= 0.03, (cpu[:, 2], dev[:, 2])


def test_reference_time_series_deck_brownian_relaxation(gpu, exe, tmp_path):
    """The reference's tests/time-series deck (SC 22^3, J = 0, alpha = 1, 1000 K, spins +z: free
    spins in a heat bath) with dt = 0.1 fs over 200 steps, run by the device and by the reference
    CPU integrator.  Each spin diffuses freely on the sphere, so |m|(t) = exp(-t/tau) with
    tau = (1 + alpha^2) mu_s / (2 alpha gamma k_B T) = 6.6 fs (Brown's rotational diffusion; the
    thermal-field amplitude and the damping normalisation fix it).  Both runs follow it within 0.03
    (about three times the 1/sqrt(N) noise) and agree with each other within 0.03."""
    edits = [("sim:time-step=1.0E-15", "sim:time-step=1.0E-16"),
             ("sim:total-time-steps=10000", "sim:total-time-steps=200"),
             ("sim:time-steps-increment=1000", "sim:time-steps-increment=10")]
    cpu = run_reference_deck(tmp_path, "cpu", "time-series", "off", edits)
    dev = run_reference_deck(tmp_path, "gpu", "time-series", "on", edits)
    t = cpu[:, 0]
    assert np.array_equal(t, dev[:, 0]) and len(t) >= 4
    alpha, gamma, kB, T, mu_s = 1.0, 1.76e11, 1.3806503e-23, 1000.0, 1.72 * 9.27400915e-24
    tau = (1.0 + alpha ** 2) * mu_s / (2.0 * alpha * gamma * kB * T)
    theory = np.exp(-t / tau)
    for run in (cpu, dev):
        assert np.abs(run[:, 5] - theory).max() <= 0.03, (run[:, 5], theory)
    assert np.abs(cpu[:, 5] - dev[:, 5]).max() <= 0.03


def run_checkpoint_series(tmp, acc):
    """The same 300 K deck (FCC 3x3x3 cells, random start) run 200 steps without a break (A), and
    run 100 steps with sim:save-checkpoint=end (B1) and then from its vampire0.chk to step 200 with
    sim:load-checkpoint=continue (B2) or =restart (R); returns the final spins of A, B2 and R."""
    head = LATTICE.format(crystal="fcc", size=10.61)
    mat = CO.format(alpha=0.1, ku=4.644e-24, init="random")
    runs = {"A": (200, ""), "B1": (100, "sim:save-checkpoint=end\n"),
            "B2": (200, "sim:load-checkpoint=continue\n"), "R": (200, "sim:load-checkpoint=restart\n")}
    out = {}
    for name, (steps, extra) in runs.items():
        d = tmp / f"{acc}_{name}"
        d.mkdir()
        kw = dict(T=300.0, inc=50, steps=steps, H=0.5, Hdir="0,0,1", acc=acc)
        (d / "input").write_text(head + RUN.format(**kw) + extra)
        (d / "Co.mat").write_text(mat)
        if name in ("B2", "R"):
            (d / "vampire0.chk").write_bytes((tmp / f"{acc}_B1" / "vampire0.chk").read_bytes())
        r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        snap = sorted(d.glob("spins-*.data"))[-1]      # the file counter continues from the checkpoint
        out[name] = np.frombuffer(snap.read_bytes()[8:], dtype=np.float64).copy()
    return out["A"], out["B2"], out["R"]


def test_dropin_checkpoint_continue_resumes_noise_stream(gpu, exe, tmp_path):
    """The reference's own checkpoint (src/utility/checkpoint.cpp, vampire<rank>.chk) through the
    drop-in at 300 K.  The reference CPU path resumes bit for bit: its MT19937 state is in the file.
    The device path resumes bit for bit too: the spins come back through gpu::finalize before
    save_checkpoint (sim.cpp:521-524), and gpu::initialize starts the Philox counter at the restored
    sim::time.  A restart from the same file (sim::time = 0) ends elsewhere, so the noise stream is
    what the comparison sees."""
    for acc in ("off", "on"):
        a, b, r = run_checkpoint_series(tmp_path, acc)
        assert a.shape == (324,)
        assert np.array_equal(a, b), (acc, np.abs(a - b).max())
        assert np.abs(a - r).max() > 1e-3, acc


FILM = """create:crystal-structure=sc
create:periodic-boundaries-x
create:periodic-boundaries-y
dimensions:unit-cell-size = 3.54 !A
dimensions:system-size-x = 21.23 !A
dimensions:system-size-y = 21.23 !A
dimensions:system-size-z = 14.15 !A
"""

TWO = """material:num-materials=2
material[1]:material-name=A
material[1]:damping-constant=0.1
material[1]:exchange-matrix[1]=6.064e-21
material[1]:exchange-matrix[2]=-2.0e-21
material[1]:atomic-spin-moment=1.72 !muB
material[1]:second-order-uniaxial-anisotropy-constant=4.644e-24
material[1]:initial-spin-direction=random
material[1]:maximum-height=0.5
material[2]:material-name=B
material[2]:damping-constant=0.05
material[2]:exchange-matrix[1]=-2.0e-21
material[2]:exchange-matrix[2]=5.0e-21
material[2]:atomic-spin-moment=2.5 !muB
material[2]:initial-spin-direction=random
material[2]:minimum-height=0.5
"""

STATS_OUT = """output:height-magnetisation
output:material-height-magnetisation
output:material-magnetisation
output:total-energy
output:exchange-energy
output:anisotropy-energy
output:applied-field-energy
output:mean-total-energy
output:material-total-energy
output:material-mean-total-energy
output:mean-magnetisation-length
"""


def assert_columns_close(dev, cpu, rtol_scale, rtol_print=1e-5):
    """The output file prints six significant digits, so two runs that agree to 1e-10 can still
    print one unit apart in the last digit: each value must match within rtol_print of itself plus
    rtol_scale of its column's largest value; failures name the columns."""
    scale = np.abs(cpu).max(axis=0)
    tol = rtol_print * np.abs(cpu) + rtol_scale * scale + 1e-300
    diff = np.abs(dev - cpu)
    bad = [(k, float(diff[:, k].max()), float(scale[k])) for k in range(cpu.shape[1]) if (diff[:, k] > tol[:, k]).any()]
    assert not bad, f"columns (index, max difference, column scale) beyond tolerance: {bad}"


def run_film(tmp, name, acc, extra="", env=None, T=0.0, steps=500, inc=50, H=0.5, Hdir="0.3,0,1", check=True):
    d = tmp / name
    d.mkdir()
    kw = dict(T=T, inc=inc, steps=steps, H=H, Hdir=Hdir, acc=acc)
    (d / "input").write_text(FILM + RUN.format(**kw) + extra)
    (d / "Co.mat").write_text(TWO)
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=200,
                       env=None if env is None else {**os.environ, **env})
    if not check:
        return r, d
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = np.array([[float(v) for v in line.split()] for line in (d / "output").read_text().splitlines()
                     if line.strip() and not line.startswith("#")])
    raw = (d / "spins-00000000.data").read_bytes()
    return rows, np.frombuffer(raw[8:], dtype=np.float64).copy()


@pytest.mark.parametrize("stats", ["host", "device"])
def test_dropin_statistics_equal_reference_cpu(gpu, exe, tmp_path, stats):
    """Two-material layered film (open in z, antiparallel inter-layer coupling), T = 0: every
    statistic in the output file (system, material, height and material-height magnetisation;
    total, exchange, anisotropy, applied-field and material energies and their means) equals the
    reference CPU run's, with statistics on the host (the backend's default, gpu:cpu-statistics)
    and on the device (gpu:calculate-statistics-on-cpu=false: magnetisation and energy sums
    reduced on the device, running means handed back through set_magnetization and
    set_*_energy / update_mean_counter)."""
    extra = STATS_OUT + ("gpu:calculate-statistics-on-cpu=false\n" if stats == "device" else "")
    cpu, s_cpu = run_film(tmp_path, "cpu", "off", STATS_OUT)
    dev, s_dev = run_film(tmp_path, "gpu", "on", extra)
    assert cpu.shape == dev.shape and cpu.shape[0] == 10 and cpu.shape[1] > 30
    assert np.abs(s_dev - s_cpu).max() <= 1e-10
    assert_columns_close(dev, cpu, 1e-9)


@pytest.mark.parametrize("stats", ["host", "device"])
def test_dropin_thin_film_demag_equals_reference_cpu(gpu, exe, tmp_path, stats):
    """sim:demagnetisation-factor (calculate_applied_fields, src/simulate/fields.cpp:219-239): the
    field follows the system magnetisation statistic, updated every 10 steps.  Device path ==
    reference CPU (<= 1e-10) with host and with device statistics; without the term the run ends
    elsewhere, so the term is in the comparison."""
    demag = "sim:demagnetisation-factor = 0.1,0.1,0.8\n"
    dev_stats = "gpu:calculate-statistics-on-cpu=false\n" if stats == "device" else ""
    kw = dict(steps=200, inc=10)
    cpu, s_cpu = run_film(tmp_path, "cpu", "off", demag, **kw)
    dev, s_dev = run_film(tmp_path, "gpu", "on", demag + dev_stats, **kw)
    plain, s_plain = run_film(tmp_path, "plain", "off", "", **kw)
    assert np.abs(s_dev - s_cpu).max() <= 1e-10
    assert_columns_close(dev, cpu, 1e-9)
    assert np.abs(s_plain - s_cpu).max() > 1e-3


def test_dropin_single_precision_against_reference_cpu(gpu, exe, tmp_path):
    """VMX_PRECISION=single selects the FP32 step through gpu:: (the reference backend's CUDA_SP
    build choice, src/cuda/makefile:36-38, made at run time), with VMX_AUTOTUNE=1: 500 steps of the
    film at T = 0 stay within 5e-5 per spin component of the reference CPU run (about a hundred
    FP32 roundings, 6e-8 each; measured 1.2e-5), and the output's statistics within 5e-5
    relative."""
    cpu, s_cpu = run_film(tmp_path, "cpu", "off", STATS_OUT)
    dev, s_dev = run_film(tmp_path, "gpu", "on", STATS_OUT, env={"VMX_PRECISION": "single", "VMX_AUTOTUNE": "1"})
    err = np.abs(s_dev - s_cpu).max()
    assert 1e-10 < err <= 5e-5, err           # FP32 (not the FP64 step) and close
    assert_columns_close(dev, cpu, 5e-5)
    r, _ = run_film(tmp_path, "bad", "on", env={"VMX_PRECISION": "half"}, check=False)
    assert r.returncode != 0 and "VMX_PRECISION" in r.stderr


def test_dropin_refuses_terms_the_device_does_not_compute(gpu, exe, tmp_path):
    """A deck with material-specific temperatures (sim::local_temperature, fields.cpp:264) is
    refused by gpu::initialize instead of running without the term."""
    r, _ = run_film(tmp_path, "local", "on", "sim:temperature=300\n", check=False)
    assert r.returncode == 0
    mat = TWO.replace("material[2]:minimum-height=0.5", "material[2]:minimum-height=0.5\nmaterial[2]:temperature=600")
    d = tmp_path / "refused"
    d.mkdir()
    (d / "input").write_text(FILM + RUN.format(T=300.0, inc=50, steps=100, H=0.5, Hdir="0,0,1", acc="on"))
    (d / "Co.mat").write_text(mat)
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "does not compute material-specific temperatures" in r.stderr


def test_dropin_continuous_checkpoints_hold_device_spins(gpu, exe, tmp_path):
    """sim:save-checkpoint=continuous writes vampire0.chk from the host spin arrays at output steps
    (src/vio/datalog.cpp:487, after gpu::stats::get); with device statistics nothing else copies
    the spins back, so the drop-in synchronises them first.  A 250-step run at 300 K with output
    every 50 steps and checkpoints every 100 leaves the step-200 checkpoint; resuming it with
    sim:load-checkpoint=continue ends bit for bit where the uninterrupted run ends."""
    extra = ("gpu:calculate-statistics-on-cpu=false\nsim:save-checkpoint=continuous\n"
             "sim:save-checkpoint-rate=100\n")
    _, a = run_film(tmp_path, "A", "on", extra, T=300.0, steps=250, inc=50)
    d = tmp_path / "B"
    d.mkdir()
    (d / "vampire0.chk").write_bytes((tmp_path / "A" / "vampire0.chk").read_bytes())
    kw = dict(T=300.0, inc=50, steps=250, H=0.5, Hdir="0.3,0,1", acc="on")
    (d / "input").write_text(FILM + RUN.format(**kw) + "gpu:calculate-statistics-on-cpu=false\n"
                             "sim:load-checkpoint=continue\n")
    (d / "Co.mat").write_text(TWO)
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    b = np.frombuffer(sorted(d.glob("spins-*.data"))[-1].read_bytes()[8:], dtype=np.float64)
    assert np.array_equal(a, b), np.abs(a - b).max()


HARNESS_VMX = os.path.join(ROOT, "vampire_amd", "shim", "_build", "ref_harness_vmx")


def test_dropin_constrained_monte_carlo_against_reference(gpu, tmp_path):
    """Constrained Monte Carlo through the gpu:: boundary: the oracle harness linked to the shim
    runs gpu::initialize, then gpu::cmc_step per sweep (its first call runs the reference's CMCinit
    on the host spins), samples the host arrays after gpu::transfer_spin_positions_from_gpu_to_cpu,
    and reads the montecarlo::cmc counters after gpu::finalize.  Against the reference's own
    cmc_step run of the same deck (SC 8^3, Co, phi = 40, theta = 30 degrees, 100-650 K):
    <m.v> and <E> within combined error bars, the moment on the constraint axis, N pair trials
    per sweep and the accepted / energy-rejected / sphere-rejected fractions within 0.01."""
    if not os.path.exists(HARNESS_VMX):
        pytest.skip("harness over the shim not built (make -C vampire_amd/shim)")
    import refcases as RC
    f = RC.load("cmc_sc8")
    phi, theta = float(f["phi"]), float(f["theta"])
    sweeps, every, equil = int(f["sweeps"]), int(f["every"]), int(f["equil"])
    L, a = int(f["L"]), float(f["a"])
    size = L * a - 0.01
    p, t = np.radians(phi), np.radians(theta)
    v = np.array([np.sin(p) * np.cos(t), np.sin(p) * np.sin(t), np.cos(p)])
    got = []
    for T, ref_counts in zip(f["rows"][:, 0], f["counts"]):
        d = tmp_path / f"T{int(T)}"
        d.mkdir()
        (d / "input").write_text(
            "create:crystal-structure=sc\n"
            + "".join(f"create:periodic-boundaries-{x}\n" for x in "xyz")
            + f"dimensions:unit-cell-size = {a} !A\n"
            + "".join(f"dimensions:system-size-{x} = {size} !A\n" for x in "xyz")
            + f"material:file=test.mat\nsim:temperature={T}\nsim:time-step=1.0E-16\n"
            + "sim:program=time-series\nsim:integrator=llg-heun\nsim:total-time-steps=0\n"
            + "output:total-energy\noutput:material-magnetisation\n")
        (d / "test.mat").write_text(CO.format(alpha=0.1, ku=4.644e-24, init="random"))
        print("gpucmc", T, flush=True)
        r = subprocess.run([HARNESS_VMX, str(d), "gpucmc", str(sweeps), str(T), str(every), str(phi), str(theta)],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        raw = np.fromfile(d / "cmc.bin", dtype=np.float64)
        data, cnt = raw[:-4].reshape(-1, 5), raw[-4:]
        keep = data[equil // every:]
        m = keep[:, :3] * keep[:, 3:4]
        mv = m @ v
        trans = np.abs(m - np.outer(mv, v)).max()
        mm, dm = RC.block_mean(mv)
        me, de = RC.block_mean(keep[:, 4])
        got.append([T, mm, dm, trans, me, de])
        assert cnt[3] == sweeps * L ** 3 and cnt[:3].sum() == cnt[3]
        np.testing.assert_allclose(cnt[:3] / cnt[3], ref_counts[:3] / ref_counts[3], atol=0.01)
    RC.assert_cmc_curve(np.array(got))


def _fixture_script():
    """tests/golden/make_ref_vectors.py as a module: the decks the golden vectors were made from."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_ref_vectors",
                                                  os.path.join(ROOT, "tests", "golden", "make_ref_vectors.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("name", ["sc6_uniaxial", "film_local", "sc6_uniaxial_mid", "film_dmi_mid"])
def test_dropin_heun_and_midpoint_through_shim_equal_reference(gpu, tmp_path, name):
    """gpu::initialize / gpu::llg_heun / gpu::llg_midpoint / gpu::transfer_spin_positions_from_gpu_to_cpu
    on the oracle harness linked to the shim, on the deck a golden trajectory was made from: with
    mtrandom::integration_seed = the noise seed and sim::time = 1 the device draws the Philox stream
    the reference was fed, so its snapshots equal the reference's sim::LLG_Heun / sim::LLG_Midpoint
    trajectory to 1e-10 (FP64)."""
    if not os.path.exists(HARNESS_VMX):
        pytest.skip("harness over the shim not built (make -C vampire_amd/shim)")
    import refcases as RC
    mrv = _fixture_script()
    kw = mrv.llg_cases()[name]
    f = RC.load(name)
    inp, mat = mrv.deck(kw["crystal"], kw["L"], kw["a"], kw["materials"], kw.get("pbc", (True, True, True)),
                        kw["T"], kw.get("H", 0.0), kw.get("Hdir", (0, 0, 1)), kw.get("dmi_cutoff"), kw.get("Lz"),
                        kw.get("extra", ()))
    d = tmp_path / name
    d.mkdir()
    (d / "input").write_text(inp)
    (d / "test.mat").write_text(mat)
    np.ascontiguousarray(f["spins_in"], dtype=np.float64).tofile(d / "spins_in.bin")
    steps = [int(v) for v in f["snap_steps"]]
    integ = "mid" if kw.get("integrator") == "midpoint" else "heun"
    print("gpullg", name, integ, flush=True)
    r = subprocess.run([HARNESS_VMX, str(d), "gpullg", str(steps[-1]), "1", integ, str(kw.get("seed", 11))],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    n = f["spins_in"].shape[1]
    traj = np.fromfile(d / "traj.bin", dtype=np.float64).reshape(steps[-1], 3, n)
    for k, st in enumerate(steps):
        assert np.abs(traj[st - 1] - f["snaps"][k]).max() <= 1e-10, (name, st)


def test_dropin_metropolis_through_shim_matches_reference(gpu, tmp_path):
    """gpu::mc_step (one storage-order checkerboard sweep per call) on the harness linked to the
    shim, on the deck of the reference's mc_step curve (SC 8^3, no anisotropy, from +z, 6000
    sweeps, samples every 5 after 1000): <|m|>(T) at 100-900 K and the 1/2 crossing within the
    combined error bars of the reference run."""
    if not os.path.exists(HARNESS_VMX):
        pytest.skip("harness over the shim not built (make -C vampire_amd/shim)")
    import refcases as RC
    mrv = _fixture_script()
    f = RC.load("mc_sc8")
    L, a = int(f["L"]), float(f["a"])
    sweeps, equil, every = int(f["sweeps"]), int(f["equil"]), 5
    mat = dict(mrv.CO, **{"second-order-uniaxial-anisotropy-constant": 0.0, "initial-spin-direction": "0,0,1"})
    temps = f["M_T"][:, 0]
    means, errs = [], []
    for T in temps:
        d = tmp_path / f"T{int(T)}"
        d.mkdir()
        inp, mtxt = mrv.deck("sc", L, a, [mat], (True, True, True), float(T))
        (d / "input").write_text(inp)
        (d / "test.mat").write_text(mtxt)
        print("gpumc", T, flush=True)
        r = subprocess.run([HARNESS_VMX, str(d), "gpumc", str(sweeps), str(T), str(every)],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        m = np.fromfile(d / "mc.bin", dtype=np.float64).reshape(-1, 8)[:, 3]
        mean, err = RC.block_mean(m[equil // every:])
        means.append(mean)
        errs.append(err)
    RC.assert_mc_curve(temps, np.array(means), np.array(errs))


def test_dropin_constrained_mc_follows_a_moved_constraint(gpu, tmp_path):
    """cmc_anisotropy moves the constraint between its loops and re-runs CMCinit on the host
    (src/program/cmc_anisotropy.cpp:114-118, 163-165).  Through gpu::cmc_step the device must then
    start from CMCinit's spins and hold the new direction: SC 8^3 at 300 K, phi/theta 40/30 for
    3000 sweeps, then 70/120.  Pair moves conserve the transverse moment exactly, so a device that
    kept its old spins would carry the old moment's large transverse part for ever; here it is
    below 1e-12 about each constraint in turn, and <m.v> under the new constraint equals the
    reference's under the old one within 0.02 (the weak anisotropy barely changes the magnitude)."""
    if not os.path.exists(HARNESS_VMX):
        pytest.skip("harness over the shim not built (make -C vampire_amd/shim)")
    import refcases as RC
    mrv = _fixture_script()
    f = RC.load("cmc_sc8")
    L, a = int(f["L"]), float(f["a"])
    ref300 = f["rows"][np.where(f["rows"][:, 0] == 300.0)[0][0]]
    d = tmp_path / "move"
    d.mkdir()
    inp, mtxt = mrv.deck("sc", L, a, [mrv.CO], (True, True, True), 300.0)
    (d / "input").write_text(inp)
    (d / "test.mat").write_text(mtxt)
    print("gpucmc move", flush=True)
    r = subprocess.run([HARNESS_VMX, str(d), "gpucmc", "6000", "300.0", "5", "40", "30", "70", "120", "3001"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    raw = np.fromfile(d / "cmc.bin", dtype=np.float64)
    data = raw[:-4].reshape(-1, 5)
    m = data[:, :3] * data[:, 3:4]

    def axis(phi, theta):
        p, t = np.radians(phi), np.radians(theta)
        return np.array([np.sin(p) * np.cos(t), np.sin(p) * np.sin(t), np.cos(p)])
    v1, v2 = axis(40, 30), axis(70, 120)
    before, after = m[:600], m[600:]                  # samples at sweeps 5..3000 and 3005..6000
    assert np.abs(before - np.outer(before @ v1, v1)).max() < 1e-12
    assert np.abs(after - np.outer(after @ v2, v2)).max() < 1e-12
    mv, dm = RC.block_mean(after[200:] @ v2)
    assert abs(mv - ref300[1]) <= 0.02 + 4 * np.hypot(dm, ref300[2]), (mv, ref300[1])
