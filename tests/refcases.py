"""The golden-vector cases of tests/golden/ref_*.npz as Systems.

Each case pairs the arrays the reference built (neighbour list, interaction ids, exchange table,
material ids, moments) with the material parameters of its input deck (make_ref_vectors.py), run
through this package's derived-parameter code.  Loading only reads the .npz files.
"""
from __future__ import annotations

import os

import numpy as np

from vampire_amd import system as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CO = dict(mu_s_muB=1.72, alpha=0.1, Ku=4.644e-24, ku_vector=(0.0, 0.0, 1.0))

CASES = {
    "sc6_uniaxial": dict(mats=[S.Material(**CO)], ani=dict(uniaxial=True)),
    "fcc3_cubic": dict(mats=[S.Material(mu_s_muB=1.72, alpha=0.05, Kc=3.0e-24)], ani=dict(cubic=True)),
    "bcc4_rotated": dict(mats=[S.Material(mu_s_muB=1.72, alpha=0.1, Kc=-2.0e-24, kc_vector1=(1.0, 1.0, 0.0),
                                          kc_vector2=(-1.0, 1.0, 0.0))], ani=dict(cubic_rotated=True)),
    "film_dmi": dict(mats=[S.Material(mu_s_muB=1.6, alpha=0.1), S.Material(mu_s_muB=7.6, alpha=0.1)], ani={}),
    "c0_sc20": dict(mats=[S.Material(**CO)], ani=dict(uniaxial=True)),
    # vectorial exchange (three-value exchange-matrix): one material, then two layered materials
    # thin-film demagnetising field (sim:demagnetisation-factor = 0,0.2,0.8), film open in z
    "film_demag": dict(mats=[S.Material(**CO)], ani=dict(uniaxial=True)),
    "sc6_vectorial": dict(mats=[S.Material(**CO)], ani=dict(uniaxial=True),
                          J=[[[6.064e-21, 5.2e-21, 7.3e-21]]]),
    "film_vectorial": dict(mats=[S.Material(mu_s_muB=1.6, alpha=0.1, Ku=2.0e-24, ku_vector=(0.0, 0.0, 1.0)),
                                 S.Material(mu_s_muB=2.4, alpha=0.05)], ani=dict(uniaxial=True),
                           J=[[[5e-21, 4.1e-21, 6.2e-21], [-1.2e-21, -0.9e-21, -1.6e-21]],
                              [[-1.2e-21, -0.9e-21, -1.6e-21], [5.5e-21, 6.1e-21, 4.4e-21]]]),
    # material-specific applied fields (sim::local_applied_field): 0.8 T along x, 1.5 T along (0,1,1)
    "film_local": dict(mats=[S.Material(mu_s_muB=1.6, alpha=0.1, Ku=1.0e-24, applied_field_strength=0.8,
                                        applied_field_unit_vector=(1.0, 0.0, 0.0)),
                             S.Material(mu_s_muB=2.0, alpha=0.05, applied_field_strength=1.5,
                                        applied_field_unit_vector=(0.0, 1.0, 1.0))],
                       ani=dict(uniaxial=True), local_applied=True),
    "sc6_uniaxial_mid": dict(mats=[S.Material(**CO)], ani=dict(uniaxial=True)),
    "film_dmi_mid": dict(mats=[S.Material(mu_s_muB=1.6, alpha=0.1), S.Material(mu_s_muB=7.6, alpha=0.1)], ani={}),
    "dipole_sc10": dict(mats=[S.Material(**CO), S.Material(**dict(CO, mu_s_muB=2.5))], ani=dict(uniaxial=True)),
}


def load(name: str):
    return np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))


def system_of(name: str) -> S.System:
    """System with the reference's own arrays and this package's derived material constants."""
    f = load(name)
    spec = CASES[name]
    dt = float(f["dt"])
    mats = [S.derived_material(m, dt) for m in spec["mats"]]
    a = spec["ani"]
    flags = (1 if a.get("uniaxial") else 0) | (2 if a.get("cubic") else 0) | (4 if a.get("cubic_rotated") else 0)
    return S.System(
        n=int(f["type"].shape[0]), type=f["type"].astype(np.int32), moment=f["moment"].astype(np.float64),
        nbr_start=f["nbr_start"].astype(np.int32), nbr_end=f["nbr_end"].astype(np.int32),
        nbr=f["nbr"].astype(np.int32), nbr_iid=f["nbr_iid"].astype(np.int32),
        exchange_type=int(f["exchange_type"]), exchange=np.ascontiguousarray(f["exchange"], dtype=np.float64),
        materials=mats, anisotropy_flags=flags, dt=dt, half_dt=float(f["half_dt"]), coords=f["coords"])


def conditions(name: str):
    f = load(name)
    if "H" not in f.files:
        return float(f["T"]), np.zeros(3)
    return float(f["T"]), np.array(S.applied_field(float(f["H"]), f["Hdir"]))


def flags(name: str) -> dict:
    """Condition keywords a deck needs beyond (T, H): local_applied for material-specific fields."""
    return {"local_applied": True} if CASES[name].get("local_applied") else {}


def dipole_args(name: str, spins_key: str = "spins_in"):
    """Arguments of Oracle.set_dipole / Context.set_dipole for a dipole fixture."""
    f = load(name)
    return dict(spins=f[spins_key], num_cells=int(f["num_cells"]), local_cell=f["local_cell"],
                atom_cell=f["atom_cell"], volume=f["volume"], tensor=f["tensor"], update_rate=int(f["update_rate"]))


def isolated_spins(n: int) -> S.System:
    """n spins with no neighbours and uniaxial anisotropy (one colour of n atoms)."""
    mat = S.Material(mu_s_muB=1.72, alpha=1.0, Ku=5e-24)
    dt = 1e-16 * S.GAMMA_SI
    return S.System(n=n, type=np.zeros(n, np.int32), moment=np.full(n, 1.72), nbr_start=np.zeros(n, np.int32),
                    nbr_end=np.full(n, -1, np.int32), nbr=np.zeros(0, np.int32), nbr_iid=np.zeros(0, np.int32),
                    exchange_type=0, exchange=np.zeros((1, 1)), materials=[S.derived_material(mat, dt)],
                    anisotropy_flags=1, dt=dt, half_dt=0.5 * dt)


def block_mean(samples, nblocks: int = 10):
    """Mean over axis 0 and its error from `nblocks` block means (autocorrelation-aware)."""
    bm = np.array([b.mean(axis=0) for b in np.array_split(np.asarray(samples), nblocks)])
    return bm.mean(axis=0), bm.std(axis=0, ddof=1) / np.sqrt(len(bm))


def crossing(temps, m, dm, level=0.5):
    """Temperature where the linear interpolation of m(T) crosses `level`, and its error from the
    errors of the two bracketing points."""
    k = int(np.nonzero((m[:-1] >= level) & (m[1:] < level))[0][0])
    t1, t2, m1, m2 = temps[k], temps[k + 1], m[k], m[k + 1]
    tc = t1 + (m1 - level) * (t2 - t1) / (m1 - m2)
    d1 = (t2 - t1) * (level - m2) / (m1 - m2) ** 2
    d2 = (t2 - t1) * (m1 - level) / (m1 - m2) ** 2
    return tc, float(np.hypot(d1 * dm[k], d2 * dm[k + 1]))


def assert_mc_curve(temps, mean, err, name: str = "mc_sc8"):
    """<|m|>(T) against the reference's montecarlo::mc_step curve within 4 combined sigma + 0.01
    (finite-run bias), and the <|m|> = 1/2 crossing within 4 combined sigma + 5 K."""
    rows = load(name)["M_T"]
    assert np.array_equal(np.asarray(temps, dtype=float), rows[:, 0])
    for r in range(len(temps)):
        tol = 4.0 * np.hypot(err[r], rows[r, 2]) + 0.01
        assert abs(mean[r] - rows[r, 1]) <= tol, (temps[r], mean[r], rows[r, 1], tol)
    tc_ref, dtc_ref = crossing(rows[:, 0], rows[:, 1], rows[:, 2])
    tc, dtc = crossing(rows[:, 0], np.asarray(mean), np.asarray(err))
    assert 500.0 < tc_ref < 700.0
    assert abs(tc - tc_ref) <= 4.0 * np.hypot(dtc_ref, dtc) + 5.0, (tc, tc_ref, dtc, dtc_ref)


def cmc_curve(run_sweeps, n, v, P, temps_rows, equil=1000, sweeps=6000, every=5, energy=None):
    """<m.v> and <E> per temperature of a constrained Monte Carlo driver, sampled as the reference
    fixture was: run_sweeps(T, spins, nsweeps, sweep0) -> spins; energy(spins) -> total energy."""
    out = []
    for T in temps_rows[:, 0]:
        sp = np.tile(np.asarray(v, dtype=float)[:, None], (1, n))
        sp = run_sweeps(float(T), sp, equil, 0)
        ms, es, trans = [], [], 0.0
        for w in range((sweeps - equil) // every):
            sp = run_sweeps(float(T), sp, every, equil + every * w)
            m = sp.sum(axis=1) / n
            ms.append(m @ v)
            trans = max(trans, float(np.abs((np.asarray(P).reshape(3, 3) @ m)[:2]).max()))
            if energy is not None:
                es.append(energy(float(T), sp))
        mm, dm = block_mean(ms)
        row = [T, mm, dm, trans]
        if energy is not None:
            me, de = block_mean(es)
            row += [me, de]
        out.append(row)
    return np.array(out)


def assert_cmc_curve(got, name="cmc_sc8"):
    """<m.v>(T) within 4 combined sigma + 0.01 of the reference's cmc_step, <E>(T) within 4
    combined sigma + 0.5 %, transverse moment in the constraint frame below 1e-12."""
    rows = load(name)["rows"]
    for r, g in zip(rows, got):
        assert g[0] == r[0]
        assert abs(g[1] - r[1]) <= 4.0 * np.hypot(g[2], r[2]) + 0.01, (r[0], g[1], r[1])
        assert g[3] <= 1e-12, (r[0], g[3])
        if len(g) > 4:
            assert abs(g[4] - r[3]) <= 4.0 * np.hypot(g[5], r[4]) + 0.005 * abs(r[3]), (r[0], g[4], r[3])
