"""ctypes binding of oracle/liboracle.so — TEST INFRASTRUCTURE.

The oracle is what the device path is checked against: only tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg use it.  See vmx_oracle.c for the reference functions it restates.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class or_material(C.Structure):
    _fields_ = [
        ("alpha", C.c_double), ("gamma_rel", C.c_double), ("mu_s_SI", C.c_double),
        ("one_oneplusalpha_sq", C.c_double), ("alpha_oneplusalpha_sq", C.c_double), ("H_th_sigma", C.c_double),
        ("temperature_rescaling_alpha", C.c_double), ("temperature_rescaling_Tc", C.c_double),
        ("ku2", C.c_double), ("ku_vector", C.c_double * 3), ("kc4", C.c_double),
        ("kc_vector1", C.c_double * 3), ("kc_vector2", C.c_double * 3), ("kc_vector3", C.c_double * 3),
        ("applied_field_strength", C.c_double), ("applied_field_unit_vector", C.c_double * 3),
    ]


_P = C.POINTER
_IP = _P(C.c_int)
_DP = _P(C.c_double)


class or_dipole(C.Structure):
    _fields_ = [
        ("ncells", C.c_int), ("nlocal", C.c_int), ("local_cell", _IP), ("atom_cell", _IP), ("magnetic", _IP),
        ("volume", _DP), ("tensor", _DP), ("update_rate", C.c_int), ("time", C.c_longlong), ("atom_B", _DP),
        ("atom_Hd", _DP), ("cell_mag", _DP), ("cell_B", _DP), ("cell_Hd", _DP),
    ]


class or_sys(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("nmat", C.c_int), ("type", _IP), ("moment", _DP), ("nbr_start", _IP),
        ("nbr_end", _IP), ("nbr", _IP), ("iid", _IP), ("exchange_type", C.c_int), ("exchange", _DP),
        ("mat", _P(or_material)), ("ani", C.c_uint), ("dip", _P(or_dipole)),
    ]


class or_cond(C.Structure):
    _fields_ = [("temperature", C.c_double), ("H", C.c_double * 3), ("fld", C.c_uint), ("dt", C.c_double),
                ("half_dt", C.c_double), ("Hd", C.c_double * 3)]


class or_cmc(C.Structure):
    _fields_ = [("P", C.c_double * 9), ("Pt", C.c_double * 9), ("v", C.c_double * 3), ("move", C.c_int),
                ("sigma", C.c_double)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        lib = C.CDLL(LIB_PATH)
        lib.or_philox4x32_10.argtypes = [_P(C.c_uint32), _P(C.c_uint32), _P(C.c_uint32)]
        lib.or_normals3.argtypes = [C.c_uint64, C.c_int, C.c_uint64, C.c_uint32, _DP]
        lib.or_thermal_noise.argtypes = [C.c_uint64, C.c_int, C.c_uint64, _DP]
        lib.or_spin_fields.argtypes = [_P(or_sys), _DP, _DP, _DP, _DP, _DP, _DP]
        lib.or_external_fields.argtypes = [_P(or_sys), _P(or_cond), _DP, C.c_uint64, C.c_uint64, _DP, _DP, _DP]
        lib.or_llg_heun.argtypes = [_P(or_sys), _P(or_cond), _DP, _DP, _DP, C.c_int, _DP, C.c_uint64,
                                    C.c_uint64, C.c_int]
        lib.or_llg_heun.restype = C.c_int
        lib.or_llg_midpoint.argtypes = lib.or_llg_heun.argtypes
        lib.or_llg_midpoint.restype = C.c_int
        lib.or_magnetization.argtypes = [C.c_int, _DP, _DP, _DP, _DP, _IP, C.c_int, _DP]
        lib.or_energy.argtypes = [_P(or_sys), _P(or_cond), _DP, _DP, _DP, _IP, C.c_int, _DP]
        lib.or_spin_energy.argtypes = [_P(or_sys), _P(or_cond), _DP, _DP, _DP, C.c_int, C.c_double, C.c_double,
                                       C.c_double]
        lib.or_spin_energy.restype = C.c_double
        lib.or_mc_checkerboard.argtypes = [_P(or_sys), _P(or_cond), _DP, _DP, _DP, _IP, _IP, C.c_int, C.c_int,
                                           C.c_uint64, C.c_uint64, _DP, _DP, _DP]
        lib.or_mc_random_site.argtypes = [_P(or_sys), _P(or_cond), _DP, _DP, _DP, C.c_int, C.c_uint64,
                                          C.c_uint64, _DP, _DP, _DP]
        lib.or_dipole_update.argtypes = [_P(or_sys), _DP, _DP, _DP]
        lib.or_cmc_frame.argtypes = [C.c_double, C.c_double, _DP, _DP, _DP]
        lib.or_cmc_random_pairs.argtypes = [_P(or_sys), _P(or_cond), _P(or_cmc), _DP, _DP, _DP, C.c_int,
                                            C.c_uint64, C.c_uint64, _DP]
        lib.or_cmc_checkerboard.argtypes = [_P(or_sys), _P(or_cond), _P(or_cmc), _DP, _DP, _DP, _IP, _IP, C.c_int,
                                            C.c_int, C.c_uint64, C.c_uint64, _DP]
        lib.or_greedy_colour.argtypes = [_P(or_sys), _IP]
        lib.or_greedy_colour.restype = C.c_int
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(_DP)


def _i(a):
    return a.ctypes.data_as(_IP)


def philox4x32_10(ctr: Sequence[int], key: Sequence[int]):
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.or_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def thermal_noise(seed: int, n: int, step: int) -> np.ndarray:
    out = np.empty(3 * n)
    load().or_thermal_noise(seed, n, step, _d(out))
    return out.reshape(3, n)


class Oracle:
    """The reference CPU path on one system (one replica)."""

    def __init__(self, system):
        self.lib = load()
        s = system
        self.sys_ = s
        self.n = s.n
        self._keep = dict(
            type=np.ascontiguousarray(s.type, dtype=np.int32),
            moment=np.ascontiguousarray(s.moment, dtype=np.float64),
            start=np.ascontiguousarray(s.nbr_start, dtype=np.int32),
            end=np.ascontiguousarray(s.nbr_end, dtype=np.int32),
            nbr=np.ascontiguousarray(s.nbr, dtype=np.int32),
            iid=np.ascontiguousarray(s.nbr_iid, dtype=np.int32),
            ex=np.ascontiguousarray(s.exchange, dtype=np.float64).reshape(-1),
        )
        mats = (or_material * len(s.materials))()
        for m, d in enumerate(s.materials):
            for key in ("alpha", "gamma_rel", "mu_s_SI", "one_oneplusalpha_sq", "alpha_oneplusalpha_sq",
                        "H_th_sigma", "temperature_rescaling_alpha", "temperature_rescaling_Tc", "ku2", "kc4",
                        "applied_field_strength"):
                setattr(mats[m], key, float(d[key]))
            for key in ("ku_vector", "kc_vector1", "kc_vector2", "kc_vector3", "applied_field_unit_vector"):
                for k in range(3):
                    getattr(mats[m], key)[k] = float(d[key][k])
        self._mats = mats
        k = self._keep
        self.sys = or_sys(n=s.n, nmat=len(s.materials), type=_i(k["type"]), moment=_d(k["moment"]),
                          nbr_start=_i(k["start"]), nbr_end=_i(k["end"]), nbr=_i(k["nbr"]), iid=_i(k["iid"]),
                          exchange_type=s.exchange_type, exchange=_d(k["ex"]), mat=mats,
                          ani=s.anisotropy_flags)
        self.cond = or_cond(temperature=0.0, fld=0, dt=s.dt, half_dt=s.half_dt)

    def set_dipole(self, spins: np.ndarray, num_cells: int, local_cell, atom_cell, volume, tensor,
                   update_rate: int, magnetic=None):
        """Macrocell dipole field, tensor solver: cells and tensor as the reference builds them
        (tensor [6][len(local_cell)][num_cells]); the field of `spins` is computed now (time 0) and
        refreshed every `update_rate` Heun steps or sweeps."""
        nl = len(local_cell)
        k = dict(local_cell=np.ascontiguousarray(local_cell, dtype=np.int32),
                 atom_cell=np.ascontiguousarray(atom_cell, dtype=np.int32),
                 magnetic=np.ascontiguousarray(np.ones(len(self.sys_.materials)) if magnetic is None else magnetic,
                                               dtype=np.int32),
                 volume=np.ascontiguousarray(volume, dtype=np.float64),
                 tensor=np.ascontiguousarray(tensor, dtype=np.float64).reshape(-1),
                 atom_B=np.zeros(3 * self.n), atom_Hd=np.zeros(3 * self.n), cell_mag=np.zeros(3 * num_cells),
                 cell_B=np.zeros(3 * num_cells), cell_Hd=np.zeros(3 * num_cells))
        assert k["tensor"].size == 6 * nl * num_cells and k["atom_cell"].size == self.n
        self._dip_keep = k
        self.dip = or_dipole(ncells=int(num_cells), nlocal=nl, local_cell=_i(k["local_cell"]),
                             atom_cell=_i(k["atom_cell"]), magnetic=_i(k["magnetic"]), volume=_d(k["volume"]),
                             tensor=_d(k["tensor"]), update_rate=int(update_rate), time=0, atom_B=_d(k["atom_B"]),
                             atom_Hd=_d(k["atom_Hd"]), cell_mag=_d(k["cell_mag"]), cell_B=_d(k["cell_B"]),
                             cell_Hd=_d(k["cell_Hd"]))
        self.sys.dip = C.pointer(self.dip)
        self.dipole_update(spins)

    def dipole_update(self, spins: np.ndarray):
        sp = [np.ascontiguousarray(spins[k], dtype=np.float64) for k in range(3)]
        self.lib.or_dipole_update(C.byref(self.sys), *[_d(a) for a in sp])

    def dipole_fields(self):
        """(atom B [3][n], atom mu0 Hd [3][n], cell moments, cell B, cell mu0 Hd [3][num_cells])."""
        k = self._dip_keep
        return tuple(k[x].reshape(3, -1).copy() for x in ("atom_B", "atom_Hd", "cell_mag", "cell_B", "cell_Hd"))

    def set_conditions(self, temperature: float, H=(0.0, 0.0, 0.0), thermal=True, applied=True,
                       local_applied=False, demag=(0.0, 0.0, 0.0)):
        """demag: the thin-film demagnetising field added after the applied field (sim::ext_demag)."""
        self.cond.temperature = float(temperature)
        for k in range(3):
            self.cond.H[k] = float(H[k])
            self.cond.Hd[k] = float(demag[k])
        self.cond.fld = (1 if thermal else 0) | (2 if applied else 0) | (4 if local_applied else 0)

    def spin_fields(self, spins: np.ndarray) -> np.ndarray:
        sp = [np.ascontiguousarray(spins[k], dtype=np.float64) for k in range(3)]
        h = [np.empty(self.n) for _ in range(3)]
        self.lib.or_spin_fields(C.byref(self.sys), *[_d(a) for a in sp], *[_d(a) for a in h])
        return np.stack(h)

    def external_fields(self, noise: Optional[np.ndarray] = None, seed: int = 0, step: int = 0) -> np.ndarray:
        h = [np.empty(self.n) for _ in range(3)]
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64).reshape(-1)
        self.lib.or_external_fields(C.byref(self.sys), C.byref(self.cond), None if nz is None else _d(nz), seed,
                                    step, *[_d(a) for a in h])
        return np.stack(h)

    def llg_heun(self, spins: np.ndarray, nsteps: int, noise: Optional[np.ndarray] = None, seed: int = 0,
                 step0: int = 0, threads: int = 1) -> np.ndarray:
        sp = [np.ascontiguousarray(spins[k], dtype=np.float64).copy() for k in range(3)]
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64).reshape(-1)
        rc = self.lib.or_llg_heun(C.byref(self.sys), C.byref(self.cond), *[_d(a) for a in sp], int(nsteps),
                                  None if nz is None else _d(nz), seed, step0, threads)
        if rc != 0:
            raise MemoryError("oracle allocation failed")
        return np.stack(sp)

    def llg_midpoint(self, spins: np.ndarray, nsteps: int, noise: Optional[np.ndarray] = None, seed: int = 0,
                     step0: int = 0, threads: int = 1) -> np.ndarray:
        sp = [np.ascontiguousarray(spins[k], dtype=np.float64).copy() for k in range(3)]
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64).reshape(-1)
        rc = self.lib.or_llg_midpoint(C.byref(self.sys), C.byref(self.cond), *[_d(a) for a in sp], int(nsteps),
                                      None if nz is None else _d(nz), seed, step0, threads)
        if rc != 0:
            raise MemoryError("oracle allocation failed")
        return np.stack(sp)

    def magnetization(self, spins: np.ndarray, mask: np.ndarray, mask_size: int) -> np.ndarray:
        sp = [np.ascontiguousarray(spins[k], dtype=np.float64) for k in range(3)]
        out = np.empty(4 * mask_size)
        m = np.ascontiguousarray(mask, dtype=np.int32)
        self.lib.or_magnetization(self.n, *[_d(a) for a in sp], _d(self._keep["moment"]), _i(m), mask_size,
                                  _d(out))
        return out

    def energy(self, spins: np.ndarray, mask: np.ndarray, mask_size: int) -> np.ndarray:
        sp = [np.ascontiguousarray(spins[k], dtype=np.float64) for k in range(3)]
        out = np.empty(5 * mask_size)
        m = np.ascontiguousarray(mask, dtype=np.int32)
        self.lib.or_energy(C.byref(self.sys), C.byref(self.cond), *[_d(a) for a in sp], _i(m), mask_size, _d(out))
        return out.reshape(5, mask_size)

    def spin_energy(self, spins: np.ndarray, atom: int, s=None) -> float:
        sp = [np.ascontiguousarray(spins[k], dtype=np.float64) for k in range(3)]
        if s is None:
            s = (sp[0][atom], sp[1][atom], sp[2][atom])
        return self.lib.or_spin_energy(C.byref(self.sys), C.byref(self.cond), *[_d(a) for a in sp], atom,
                                       float(s[0]), float(s[1]), float(s[2]))

    def colouring(self):
        col = np.empty(self.n, dtype=np.int32)
        nc = self.lib.or_greedy_colour(C.byref(self.sys), _i(col))
        lists = [np.nonzero(col == k)[0].astype(np.int32) for k in range(nc)]
        start = np.zeros(nc + 1, dtype=np.int32)
        for k in range(nc):
            start[k + 1] = start[k] + len(lists[k])
        return col, np.ascontiguousarray(np.concatenate(lists)), start

    def mc_checkerboard(self, spins, nsweeps, seed=0, sweep0=0, sigma=60.0):
        sp = [np.ascontiguousarray(spins[k], dtype=np.float64).copy() for k in range(3)]
        _, lst, start = self.colouring()
        sg, mv, rj = C.c_double(sigma), C.c_double(0.0), C.c_double(0.0)
        self.lib.or_mc_checkerboard(C.byref(self.sys), C.byref(self.cond), *[_d(a) for a in sp], _i(lst),
                                    _i(start), len(start) - 1, int(nsweeps), seed, sweep0, C.byref(sg),
                                    C.byref(mv), C.byref(rj))
        return np.stack(sp), sg.value, mv.value, rj.value

    def cmc(self, phi: float, theta: float, move: int = 0, sigma: float = 60.0):
        """Constraint of constrained Monte Carlo: polar angle phi and azimuth theta (degrees) of the
        magnetisation (sim::constraint_phi, constraint_theta); move 0 = cone of width sigma, 1 = the
        temperature-scaled angle move."""
        k = or_cmc()
        self.lib.or_cmc_frame(float(phi), float(theta), k.P, k.Pt, k.v)
        k.move = int(move)
        k.sigma = float(sigma)
        return k

    def cmc_random_pairs(self, spins, nsweeps, cmc, seed=0, sweep0=0):
        """montecarlo::cmc_step sweeps; returns (spins, [accepted, energy rejects, sphere rejects])."""
        sp = [np.ascontiguousarray(spins[k], dtype=np.float64).copy() for k in range(3)]
        counts = np.zeros(3)
        self.lib.or_cmc_random_pairs(C.byref(self.sys), C.byref(self.cond), C.byref(cmc), *[_d(a) for a in sp],
                                     int(nsweeps), seed, sweep0, _d(counts))
        return np.stack(sp), counts

    def cmc_checkerboard(self, spins, nsweeps, cmc, seed=0, sweep0=0):
        """The colour-paired constrained sweeps of the device; returns (spins, counts)."""
        sp = [np.ascontiguousarray(spins[k], dtype=np.float64).copy() for k in range(3)]
        _, lst, start = self.colouring()
        counts = np.zeros(3)
        self.lib.or_cmc_checkerboard(C.byref(self.sys), C.byref(self.cond), C.byref(cmc), *[_d(a) for a in sp],
                                     _i(lst), _i(start), len(start) - 1, int(nsweeps), seed, sweep0, _d(counts))
        return np.stack(sp), counts

    def mc_random_site(self, spins, nsweeps, seed=0, sweep0=0, sigma=60.0):
        sp = [np.ascontiguousarray(spins[k], dtype=np.float64).copy() for k in range(3)]
        sg, mv, rj = C.c_double(sigma), C.c_double(0.0), C.c_double(0.0)
        self.lib.or_mc_random_site(C.byref(self.sys), C.byref(self.cond), *[_d(a) for a in sp], int(nsweeps), seed,
                                   sweep0, C.byref(sg), C.byref(mv), C.byref(rj))
        return np.stack(sp), sg.value, mv.value, rj.value
