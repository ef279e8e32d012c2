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


def dipole_args(name: str, spins_key: str = "spins_in"):
    """Arguments of Oracle.set_dipole / Context.set_dipole for a dipole fixture."""
    f = load(name)
    return dict(spins=f[spins_key], num_cells=int(f["num_cells"]), local_cell=f["local_cell"],
                atom_cell=f["atom_cell"], volume=f["volume"], tensor=f["tensor"], update_rate=int(f["update_rate"]))
