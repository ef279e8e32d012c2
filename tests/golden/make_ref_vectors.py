"""Writes the golden vectors under tests/golden/ from the reference VAMPIRE CPU code.

Run in the development container (needs /root/reference and oracle/_ref/ref_harness, built by
`make -C oracle ref`):

    python tests/golden/make_ref_vectors.py

For each case an input deck and material file are written to a scratch directory, the reference
builds the system itself (cs::create, exchange unrolling, anisotropy set-up) and the harness
dumps what it built, the spin and external fields, the spins after some sim::LLG_Heun steps fed
an injected noise stream, and the masked magnetisation / energy statistics.  The noise stream
is Philox-4x32-10 normals (oracle.thermal_noise) so it can be regenerated from (seed, step) and
is not stored.  A Monte Carlo case records |m| per sweep of the reference montecarlo::mc_step
with its own MT19937 generator at several temperatures.

Fixtures are .npz files of inputs and expected outputs only.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT = os.path.dirname(os.path.abspath(__file__))


def deck(crystal, L, a, materials, pbc=(True, True, True), T=0.0, H=0.0, Hdir=(0, 0, 1), dmi_cutoff=None,
         Lz=None, extra=()):
    lines = [
        f"create:crystal-structure={crystal}",
        f"dimensions:unit-cell-size = {a} !A",
        f"dimensions:system-size-x = {L * a - 0.01} !A",
        f"dimensions:system-size-y = {L * a - 0.01} !A",
        f"dimensions:system-size-z = {(Lz or L) * a - 0.01} !A",
        "material:file=test.mat",
        f"sim:temperature={T}",
        "sim:time-step=1.0E-16",
        f"sim:applied-field-strength={H} !T",
        "sim:applied-field-unit-vector=" + ",".join(str(v) for v in Hdir),
        "sim:program=time-series",
        "sim:integrator=llg-heun",
        "sim:total-time-steps=0",
        "output:total-energy",
        "output:material-magnetisation",
    ]
    for axis, p in zip("xyz", pbc):
        if p:
            lines.append(f"create:periodic-boundaries-{axis}")
    if dmi_cutoff is not None:
        lines.append(f"exchange:dmi-cutoff-range = {dmi_cutoff} !A")
    lines.extend(extra)
    mat = [f"material:num-materials={len(materials)}"]
    for k, m in enumerate(materials, start=1):
        for key, val in m.items():
            mat.append(f"material[{k}]:{key}={val}")
    return "\n".join(lines) + "\n", "\n".join(mat) + "\n"


def run(rundir, *args):
    r = subprocess.run([HARNESS, rundir, *map(str, args)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"harness failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")


def read_meta(rundir):
    meta = {"materials": []}
    for ln in open(os.path.join(rundir, "meta.txt")):
        p = ln.split()
        if p[0] == "material":
            d = {p[k]: float(p[k + 1]) for k in range(2, len(p), 2)}
            meta["materials"].append(d)
        else:
            meta[p[0]] = float(p[1]) if "." in p[1] or "e" in p[1] else int(p[1])
    return meta


def read_system(rundir, meta):
    n, nnb, nint = meta["num_atoms"], meta["num_nbr"], meta["num_interactions"]
    K = {0: 1, 1: 3, 2: 9}[meta["exchange_type"]]
    raw = open(os.path.join(rundir, "system.bin"), "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off).copy()
        off += a.nbytes
        return a

    sysd = dict(
        type=take(np.int32, n), nbr_start=take(np.int32, n), nbr_end=take(np.int32, n),
        nbr=take(np.int32, nnb), nbr_iid=take(np.int32, nnb), moment=take(np.float64, n),
        coords=np.stack([take(np.float64, n), take(np.float64, n), take(np.float64, n)], axis=1),
        spins0=np.stack([take(np.float64, n), take(np.float64, n), take(np.float64, n)]),
        exchange=take(np.float64, nint * K).reshape(nint, K),
    )
    assert off == len(raw), (off, len(raw))
    return sysd


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def case(name, crystal, L, a, materials, steps, snaps, T, H=0.0, Hdir=(0, 0, 1), pbc=(True, True, True),
         seed=11, spin_seed=5, dmi_cutoff=None, Lz=None, slim=False, integrator="heun", extra=()):
    """slim: store the neighbour list and exchange table as a SHA-256 digest (plus the distinct J
    values) instead of the arrays; the builder must reproduce them.  integrator: "heun"
    (sim::LLG_Heun), "midpoint" (sim::LLG_Midpoint) or "demag" (Heun with the statistics updated
    before every step, so the thin-film demag field of a sim:demagnetisation-factor deck follows the
    spins; that field is stored per step as demag_field)."""
    with tempfile.TemporaryDirectory(prefix="vref_") as d:
        inp, mat = deck(crystal, L, a, materials, pbc, T, H, Hdir, dmi_cutoff, Lz, extra)
        open(os.path.join(d, "input"), "w").write(inp)
        open(os.path.join(d, "test.mat"), "w").write(mat)
        run(d, "dump")
        meta = read_meta(d)
        sysd = read_system(d, meta)
        n = meta["num_atoms"]
        rng = np.random.default_rng(spin_seed)
        g = rng.standard_normal((3, n))
        spins = g / np.sqrt((g * g).sum(axis=0))
        spins.tofile(os.path.join(d, "spins_in.bin"))
        # noise: block for the field dump at step "-1" is Philox step 0, then steps 1..nsteps
        blocks = [O.thermal_noise(seed, n, s) for s in range(steps + 1)]
        np.concatenate([b.reshape(-1) for b in blocks]).tofile(os.path.join(d, "noise.bin"))
        run(d, "llg", steps, 1, *({"midpoint": ["mid"], "demag": ["demag"]}.get(integrator, [])))
        fields = np.fromfile(os.path.join(d, "fields.bin"), dtype=np.float64).reshape(6, n)
        extra_arrays = {}
        if integrator == "demag":
            extra_arrays["demag_field"] = np.fromfile(os.path.join(d, "demag.bin"), dtype=np.float64).reshape(steps, 3)
        traj = np.fromfile(os.path.join(d, "traj.bin"), dtype=np.float64).reshape(steps, 3, n)
        stats_after = np.fromfile(os.path.join(d, "stats.bin"), dtype=np.float64)
    big = {k: v for k, v in sysd.items() if k != "spins0"}
    if slim:
        big = dict(type=big["type"], moment=big["moment"],
                   csr_digest=digest(big["nbr_start"], big["nbr_end"], big["nbr"]),
                   exchange_values=np.unique(big["exchange"]))
        fields = fields[:, :0]
    np.savez_compressed(
        os.path.join(OUT, f"ref_{name}.npz"),
        crystal=crystal, L=L, a=a, T=T, H=H, Hdir=np.array(Hdir, dtype=float), pbc=np.array(pbc), integrator=integrator,
        noise_seed=seed, spins_in=spins, snap_steps=np.array(snaps),
        snaps=traj[np.array(snaps) - 1], fields=fields, stats_after=stats_after, **extra_arrays,
        dt=meta["dt"], half_dt=meta["half_dt"], exchange_type=meta["exchange_type"],
        materials=np.array([[m[k] for k in ("alpha", "gamma_rel", "mu_s_SI", "one_oneplusalpha_sq",
                                            "alpha_oneplusalpha_sq", "H_th_sigma", "ku2", "mu_s_muB")]
                            for m in meta["materials"]]),
        **big,
    )
    print(f"{name}: n={n} nnz={meta['num_nbr']} exchange_type={meta['exchange_type']} "
          f"interactions={meta['num_interactions']}")


def read_dipole(path, n, nc, nl):
    raw = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off).copy()
        off += a.nbytes
        return a

    d = dict(local_cell=take(np.int32, nl), atom_cell=take(np.int32, n), cell_count=take(np.int32, nc),
             volume=take(np.float64, nc), tensor=take(np.float64, 6 * nl * nc).reshape(6, nl, nc),
             cell_mag=take(np.float64, 3 * nc).reshape(3, nc), cell_B=take(np.float64, 3 * nc).reshape(3, nc),
             cell_Hd=take(np.float64, 3 * nc).reshape(3, nc), atom_B=take(np.float64, 3 * n).reshape(3, n),
             atom_Hd=take(np.float64, 3 * n).reshape(3, n))
    assert off == len(raw), (off, len(raw))
    return d


def dipole_case(name, L, Lz, a, materials, cell_size, rate, steps, snaps, T, seed=13):
    """Macrocell dipole field (tensor solver, open boundaries): the reference's cells, tensor and
    fields for two random spin sets, and an LLG trajectory with the field refreshed every `rate`
    steps as sim::increment_time does."""
    extra = ["dipole:solver=tensor", f"dipole:field-update-rate={rate}", f"cells:macro-cell-size = {cell_size} !A"]
    with tempfile.TemporaryDirectory(prefix="vref_dip_") as d:
        inp, mat = deck("sc", L, a, materials, (False, False, False), T, 0.0, (0, 0, 1), None, Lz, extra)
        open(os.path.join(d, "input"), "w").write(inp)
        open(os.path.join(d, "test.mat"), "w").write(mat)
        run(d, "dump")
        meta = read_meta(d)
        sysd = read_system(d, meta)
        n = meta["num_atoms"]
        rng = np.random.default_rng(seed)
        sp = []
        for _ in range(2):
            g = rng.standard_normal((3, n))
            sp.append(g / np.sqrt((g * g).sum(axis=0)))
        sp[0].tofile(os.path.join(d, "spins_in.bin"))
        sp[1].tofile(os.path.join(d, "spins2.bin"))
        run(d, "dipole")
        dm = {k.split()[0]: k.split()[1] for k in open(os.path.join(d, "meta_dipole.txt"))}
        nc, nl = int(dm["num_cells"]), int(dm["num_local_cells"])
        d0 = read_dipole(os.path.join(d, "dipole.bin"), n, nc, nl)
        d1 = read_dipole(os.path.join(d, "dipole2.bin"), n, nc, nl)
        blocks = [O.thermal_noise(seed, n, s) for s in range(steps + 1)]
        np.concatenate([b.reshape(-1) for b in blocks]).tofile(os.path.join(d, "noise.bin"))
        run(d, "llg", steps, 1)
        traj = np.fromfile(os.path.join(d, "traj.bin"), dtype=np.float64).reshape(steps, 3, n)
    for k in ("local_cell", "atom_cell", "cell_count", "volume", "tensor"):
        assert np.array_equal(d0[k], d1[k]), k
    np.savez_compressed(
        os.path.join(OUT, f"ref_{name}.npz"),
        L=L, Lz=Lz, a=a, T=T, noise_seed=seed, update_rate=rate, macro_cell_size=float(dm["macro_cell_size"]),
        spins_in=sp[0], spins2=sp[1], snap_steps=np.array(snaps), snaps=traj[np.array(snaps) - 1],
        dt=meta["dt"], half_dt=meta["half_dt"], exchange_type=meta["exchange_type"],
        materials=np.array([[m[k] for k in ("alpha", "gamma_rel", "mu_s_SI", "one_oneplusalpha_sq",
                                            "alpha_oneplusalpha_sq", "H_th_sigma", "ku2", "mu_s_muB")]
                            for m in meta["materials"]]),
        type=sysd["type"], moment=sysd["moment"], nbr_start=sysd["nbr_start"], nbr_end=sysd["nbr_end"],
        nbr=sysd["nbr"], nbr_iid=sysd["nbr_iid"], exchange=sysd["exchange"], coords=sysd["coords"],
        num_cells=nc, local_cell=d0["local_cell"], atom_cell=d0["atom_cell"], cell_count=d0["cell_count"],
        volume=d0["volume"], tensor=d0["tensor"],
        **{f"{k}0": d0[k] for k in ("cell_mag", "cell_B", "cell_Hd", "atom_B", "atom_Hd")},
        **{f"{k}1": d1[k] for k in ("cell_mag", "cell_B", "cell_Hd", "atom_B", "atom_Hd")},
    )
    print(f"{name}: n={n} cells={nc} local={nl} rate={rate}")


def mc_case(name, L, a, materials, temps, sweeps, every, equil):
    rows = []
    for T in temps:
        with tempfile.TemporaryDirectory(prefix="vref_mc_") as d:
            inp, mat = deck("sc", L, a, materials, (True, True, True), T)
            open(os.path.join(d, "input"), "w").write(inp)
            open(os.path.join(d, "test.mat"), "w").write(mat)
            run(d, "mc", sweeps, T, every)
            m = np.fromfile(os.path.join(d, "mc.bin"), dtype=np.float64).reshape(-1, 8)
        mlen = m[:, 3]
        keep = mlen[equil // every:]
        # error bar from 10 blocks (autocorrelation-aware)
        blocks = np.array_split(keep, 10)
        bm = np.array([b.mean() for b in blocks])
        rows.append((T, keep.mean(), bm.std(ddof=1) / np.sqrt(len(bm))))
        print(f"  T={T:7.1f}  <|m|>={keep.mean():.4f} +- {rows[-1][2]:.4f}")
    np.savez_compressed(os.path.join(OUT, f"ref_{name}.npz"), L=L, a=a, sweeps=sweeps, equil=equil,
                        M_T=np.array(rows))


def cmc_case(name, L, a, materials, temps, phi, theta, sweeps, every, equil):
    """Constrained Monte Carlo (montecarlo::cmc_step) at constraint (phi, theta): per temperature
    the mean magnetisation along the constraint and the mean total energy over the samples after
    `equil` sweeps, with block errors, the mean transverse moment in the constraint frame, and the
    outcome counts (accepted, energy and sphere rejections, trials)."""
    rows, counts = [], []
    P = np.empty(9)
    Pt = np.empty(9)
    v = np.empty(3)
    O.load().or_cmc_frame(float(phi), float(theta), O._d(P), O._d(Pt), O._d(v))
    for T in temps:
        with tempfile.TemporaryDirectory(prefix="vref_cmc_") as d:
            inp, mat = deck("sc", L, a, materials, (True, True, True), T)
            open(os.path.join(d, "input"), "w").write(inp)
            open(os.path.join(d, "test.mat"), "w").write(mat)
            run(d, "cmc", sweeps, T, every, phi, theta)
            raw = np.fromfile(os.path.join(d, "cmc.bin"), dtype=np.float64)
        data, cnt = raw[:-4].reshape(-1, 5), raw[-4:]
        keep = data[equil // every:]
        m = keep[:, :3] * keep[:, 3:4]                 # magnetisation vector / Ms
        mv = m @ v
        trans = np.abs(m @ P.reshape(3, 3).T[:, :2]).max()
        bm = np.array([b.mean() for b in np.array_split(mv, 10)])
        be = np.array([b.mean() for b in np.array_split(keep[:, 4], 10)])
        rows.append((T, mv.mean(), bm.std(ddof=1) / np.sqrt(10), keep[:, 4].mean(), be.std(ddof=1) / np.sqrt(10), trans))
        counts.append(cnt)
        print(f"  T={T:7.1f}  <m.v>={mv.mean():.4f} +- {rows[-1][2]:.4f}  <E>={rows[-1][3]:.5g}  "
              f"max transverse {trans:.1e}  counts {cnt}")
    np.savez_compressed(os.path.join(OUT, f"ref_{name}.npz"), L=L, a=a, sweeps=sweeps, equil=equil, every=every,
                        phi=phi, theta=theta, rows=np.array(rows), counts=np.array(counts))


CO = {"material-name": "Co", "damping-constant": 0.1, "exchange-matrix[1]": 6.064e-21,
      "atomic-spin-moment": "1.72 !muB", "second-order-uniaxial-anisotropy-constant": 4.644e-24,
      "uniaxial-anisotropy-direction": "0,0,1", "initial-spin-direction": "random"}


def llg_cases():
    """The trajectory decks: name -> the keywords of case() (deck, steps, snapshots, conditions)."""
    c = {}
    # SC, scalar exchange, uniaxial anisotropy, applied field, thermal noise
    c["sc6_uniaxial"] = dict(crystal="sc", L=6, a=3.54, materials=[CO], steps=200, snaps=[1, 10, 200], T=300.0,
                             H=0.5)
    # FCC, cubic anisotropy, alpha 0.05, T = 600 K
    fcc = dict(CO, **{"damping-constant": 0.05, "second-order-uniaxial-anisotropy-constant": 0.0,
                      "fourth-order-cubic-anisotropy-constant": 3.0e-24})
    c["fcc3_cubic"] = dict(crystal="fcc", L=3, a=3.54, materials=[fcc], steps=100, snaps=[1, 100], T=600.0, H=0.0)
    # BCC, rotated cubic basis, field along (1,0,1)/sqrt2
    bcc = dict(CO, **{"second-order-uniaxial-anisotropy-constant": 0.0,
                      "fourth-order-cubic-anisotropy-constant": -2.0e-24,
                      "cubic-anisotropy-direction-1": "1,1,0", "cubic-anisotropy-direction-2": "-1,1,0"})
    c["bcc4_rotated"] = dict(crystal="bcc", L=4, a=2.87, materials=[bcc], steps=100, snaps=[1, 100], T=100.0,
                             H=2.0, Hdir=(1, 0, 1))
    # two-material SC film, open in z, ferrimagnetic inter-material exchange and interfacial DMI:
    # tensorial exchange with the reference's DMI tensors
    m1 = {"material-name": "A", "damping-constant": 0.1, "exchange-matrix[1]": 5e-21,
          "exchange-matrix[2]": -1.2e-21, "atomic-spin-moment": "1.6 !muB", "dmi-constant[2]": 0.3e-21,
          "minimum-height": 0.0, "maximum-height": 0.5, "initial-spin-direction": "random"}
    m2 = {"material-name": "B", "damping-constant": 0.1, "exchange-matrix[1]": -1.2e-21,
          "exchange-matrix[2]": 5e-21, "atomic-spin-moment": "7.6 !muB", "dmi-constant[1]": 0.3e-21,
          "minimum-height": 0.5, "maximum-height": 1.0, "initial-spin-direction": "random"}
    c["film_dmi"] = dict(crystal="sc", L=5, a=2.5, materials=[m1, m2], steps=100, snaps=[1, 100], T=300.0, H=0.0,
                         pbc=(True, True, False), dmi_cutoff=3.6, Lz=4)
    # vectorial exchange: a three-value exchange-matrix selects exchange::vectorial
    # (src/exchange/interface.cpp:86-95, initialize.cpp:63-64), diagonal Jxx, Jyy, Jzz per edge
    # (exchange_fields.cpp:85-116).  One material (every edge the same J) ...
    cov = dict(CO, **{"exchange-matrix[1]": "6.064e-21,5.2e-21,7.3e-21"})
    c["sc6_vectorial"] = dict(crystal="sc", L=6, a=3.54, materials=[cov], steps=200, snaps=[1, 10, 200], T=300.0,
                              H=0.5, Hdir=(1, 1, 1))
    # ... and two materials in layers with antiparallel inter-layer coupling (per-edge J ids)
    va = {"material-name": "A", "damping-constant": 0.1, "exchange-matrix[1]": "5e-21,4.1e-21,6.2e-21",
          "exchange-matrix[2]": "-1.2e-21,-0.9e-21,-1.6e-21", "atomic-spin-moment": "1.6 !muB",
          "second-order-uniaxial-anisotropy-constant": 2.0e-24, "uniaxial-anisotropy-direction": "0,0,1",
          "minimum-height": 0.0, "maximum-height": 0.5, "initial-spin-direction": "random"}
    vb = {"material-name": "B", "damping-constant": 0.05, "exchange-matrix[1]": "-1.2e-21,-0.9e-21,-1.6e-21",
          "exchange-matrix[2]": "5.5e-21,6.1e-21,4.4e-21", "atomic-spin-moment": "2.4 !muB",
          "second-order-uniaxial-anisotropy-constant": 0.0,
          "minimum-height": 0.5, "maximum-height": 1.0, "initial-spin-direction": "random"}
    c["film_vectorial"] = dict(crystal="sc", L=6, a=2.5, materials=[va, vb], steps=100, snaps=[1, 100], T=200.0,
                               H=1.0, Hdir=(0, 1, 1), pbc=(True, True, False), Lz=4)
    # material-specific applied fields (material[n]:applied-field-strength / -unit-vector set
    # sim::local_applied_field, src/vio/match.cpp:2160-2230): each atom takes H + its material's
    # field (fields.cpp:193-205); two layered materials, global field along z
    la = {"material-name": "A", "damping-constant": 0.1, "exchange-matrix[1]": 5e-21, "exchange-matrix[2]": 2e-21,
          "atomic-spin-moment": "1.6 !muB", "second-order-uniaxial-anisotropy-constant": 1.0e-24,
          "uniaxial-anisotropy-direction": "0,0,1", "applied-field-strength": "0.8 !T",
          "applied-field-unit-vector": "1,0,0", "minimum-height": 0.0, "maximum-height": 0.5,
          "initial-spin-direction": "random"}
    lb = {"material-name": "B", "damping-constant": 0.05, "exchange-matrix[1]": 2e-21, "exchange-matrix[2]": 4e-21,
          "atomic-spin-moment": "2.0 !muB", "second-order-uniaxial-anisotropy-constant": 0.0,
          "applied-field-strength": "1.5 !T", "applied-field-unit-vector": "0,1,1", "minimum-height": 0.5,
          "maximum-height": 1.0, "initial-spin-direction": "random"}
    c["film_local"] = dict(crystal="sc", L=6, a=2.5, materials=[la, lb], steps=100, snaps=[1, 100], T=300.0, H=0.5,
                           Hdir=(0, 0, 1), pbc=(True, True, False), Lz=4)
    # thin-film demagnetising field (sim:demagnetisation-factor, fields.cpp:219-239): an SC film
    # open in z, statistics before every step so the field follows the system magnetisation
    c["film_demag"] = dict(crystal="sc", L=6, a=3.54, materials=[CO], steps=50, snaps=[1, 10, 50], T=300.0, H=0.5,
                           Hdir=(1, 0, 0), pbc=(True, True, False), Lz=3, integrator="demag",
                           extra=["sim:demagnetisation-factor = 0,0.2,0.8"])
    # the LLG midpoint integrator (sim::LLG_Midpoint) on the SC uniaxial and two-material DMI decks
    c["sc6_uniaxial_mid"] = dict(crystal="sc", L=6, a=3.54, materials=[CO], steps=100, snaps=[1, 10, 100],
                                 T=300.0, H=0.5, integrator="midpoint")
    c["film_dmi_mid"] = dict(crystal="sc", L=5, a=2.5, materials=[m1, m2], steps=100, snaps=[1, 100], T=300.0,
                             H=0.0, pbc=(True, True, False), dmi_cutoff=3.6, Lz=4, integrator="midpoint")
    # the CPU-runnable workload: SC 20^3, 1000 steps at 300 K
    c["c0_sc20"] = dict(crystal="sc", L=20, a=3.54, materials=[CO], steps=1000, snaps=[1000], T=300.0, H=0.0,
                        seed=1, spin_seed=1, slim=True)
    return c


def main(only=()):
    """Writes every fixture, or those named in `only` (e.g. `dipole_sc10`)."""
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")

    def want(name):
        return not only or name in only

    for name, kw in llg_cases().items():
        if want(name):
            case(name, **kw)
    # macrocell dipole field, tensor solver: SC 10x10x6 block, open boundaries, two materials in
    # layers (1.72 and 2.5 mu_B), cells of 2x2x2 unit cells, field refreshed every 10 steps
    da = dict(CO, **{"maximum-height": 0.5, "exchange-matrix[2]": 6.064e-21})
    db = dict(CO, **{"material-name": "Co2", "atomic-spin-moment": "2.5 !muB", "minimum-height": 0.5,
                     "exchange-matrix[2]": 6.064e-21})
    if want("dipole_sc10"):
        dipole_case("dipole_sc10", 10, 6, 3.54, [da, db], cell_size=7.08, rate=10, steps=30,
                    snaps=[1, 9, 10, 11, 30], T=300.0)
    # Metropolis M(T) of the reference random-site algorithm, SC 8^3, no anisotropy
    mcm = dict(CO, **{"second-order-uniaxial-anisotropy-constant": 0.0, "initial-spin-direction": "0,0,1"})
    if want("mc_sc8"):
        mc_case("mc_sc8", 8, 3.54, [mcm], temps=[100.0, 300.0, 500.0, 600.0, 700.0, 900.0],
                sweeps=6000, every=5, equil=1000)
    # constrained Monte Carlo (montecarlo::cmc_step), SC 8^3 with uniaxial anisotropy along z,
    # magnetisation held at 40 degrees from z (azimuth 30 degrees), reference MT19937 stream
    if want("cmc_sc8"):
        cmc_case("cmc_sc8", 8, 3.54, [CO], temps=[100.0, 300.0, 500.0, 650.0], phi=40.0, theta=30.0,
                 sweeps=6000, every=5, equil=1000)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
