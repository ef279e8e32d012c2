"""Host-side construction of spin systems in the reference's data model.

This is the set-up half of the path: what VAMPIRE's ``cs::create`` + ``exchange::initialize`` +
``mp::set_derived_parameters`` + ``anisotropy::initialize`` leave in the ``atoms::`` and ``mp::``
globals before ``gpu::initialize`` hands them to the device.  Arrays follow the reference layout:
inclusive neighbour start/end indices, a neighbour array, a per-neighbour interaction id into an
exchange table, a material id per atom and a moment per atom.

Semantics followed (reference files):
  * unit cells: src/unitcell/sc.cpp:24-67, fcc.cpp:25-89, bcc.cpp:25-76
  * interaction template and its order: src/unitcell/interactions.cpp:48-187
  * atom order (z, y, x, unit-cell atom): src/create/cs_create_crystal_structure2.cpp:100-135
  * per-atom neighbour order = template order, periodic wrap: src/neighbours/generate.cpp:266-378
  * exchange unrolling J_ij = template_J * Jij_matrix[mi][mj]: src/exchange/unroll_normalised.cpp
  * derived constants: src/main/initialise_variables.cpp:352-451,
    src/anisotropy/initialize.cpp:29-143
  * DMI folded into the tensor: src/exchange/dmi.cpp:46-221
  * random initial spins (normalised Gaussian triple): src/create/cs_set_atom_vars2.cpp:84-119
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

MU_B = 9.27400915e-24        # src/main/initialise_variables.cpp:448, anisotropy/initialize.cpp:51
K_B = 1.3806503e-23          # initialise_variables.cpp:431
GAMMA_SI = 1.76e11           # mp::gamma_SI

ISOTROPIC, VECTORIAL, TENSORIAL = 0, 1, 2

UNIT_CELLS = {
    # fractional positions in reference order and nearest-neighbour cutoff (unit-cell units)
    "sc": ([(0.0, 0.0, 0.0)], 1.0),
    "fcc": ([(0.0, 0.0, 0.0), (0.5, 0.5, 0.0), (0.5, 0.0, 0.5), (0.0, 0.5, 0.5)], math.sqrt(2.0) / 2.0),
    "bcc": ([(0.0, 0.0, 0.0), (0.5, 0.5, 0.5)], math.sqrt(3.0) / 2.0),
}


def interaction_template(crystal: str):
    """Nearest-neighbour interaction list (i, j, dx, dy, dz) of a unit cell, in the order
    uc::internal::calculate_interactions generates it (interactions.cpp:48-130): the replicated
    atoms are enumerated x, y, z, atom; i runs over the middle cell, j over all atoms."""
    atoms, cutoff = UNIT_CELLS[crystal]
    rcut = cutoff * 1.0 * 1.001
    rcutsq = rcut * rcut
    nrep = 1 + 2 * int(math.ceil(rcut))
    ratoms = []
    for x in range(nrep):
        for y in range(nrep):
            for z in range(nrep):
                for a, (ax, ay, az) in enumerate(atoms):
                    ratoms.append(((ax + x) * 1.0, (ay + y) * 1.0, (az + z) * 1.0, a, x, y, z))
    mid = (nrep - 1) // 2
    out = []
    for i, ri in enumerate(ratoms):
        if not (ri[4] == mid and ri[5] == mid and ri[6] == mid):
            continue
        for j, rj in enumerate(ratoms):
            rx, ry, rz = rj[0] - ri[0], rj[1] - ri[1], rj[2] - ri[2]
            if rx * rx + ry * ry + rz * rz < rcutsq and i != j:
                out.append((ri[3], rj[3], rj[4] - ri[4], rj[5] - ri[5], rj[6] - ri[6]))
    return out


@dataclasses.dataclass
class Material:
    """Input-file parameters of one material (material[n]: keys) plus derived constants."""
    mu_s_muB: float = 1.72                    # atomic-spin-moment (mu_B)
    alpha: float = 0.1                        # damping-constant
    gamma_rel: float = 1.0                    # gyromagnetic-ratio
    Ku: float = 0.0                           # second-order-uniaxial-anisotropy-constant (J)
    ku_vector: Sequence[float] = (0.0, 0.0, 1.0)
    Kc: float = 0.0                           # fourth-order-cubic-anisotropy-constant (J)
    kc_vector1: Sequence[float] = (1.0, 0.0, 0.0)
    kc_vector2: Sequence[float] = (0.0, 1.0, 0.0)
    temperature_rescaling_alpha: float = 1.0
    temperature_rescaling_Tc: float = 0.0
    applied_field_strength: float = 0.0
    applied_field_unit_vector: Sequence[float] = (0.0, 0.0, 1.0)

    @property
    def mu_s_SI(self) -> float:
        return self.mu_s_muB * MU_B


def input_unit_vector(u: Sequence[float]) -> tuple:
    """Direction as the input parser stores it: divided by its length
    (vin::check_for_valid_unit_vector, src/vio/check.cpp:253-266)."""
    u = [float(v) for v in u]
    length = math.sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2])
    if length < 1.0e-9:
        raise ValueError("direction vector has zero length")
    return (u[0] / length, u[1] / length, u[2] / length)


def applied_field(strength: float, direction: Sequence[float]) -> tuple:
    """H = sim::H_vec * sim::H_applied with H_vec from the input parser."""
    u = input_unit_vector(direction)
    return (u[0] * strength, u[1] * strength, u[2] * strength)


def derived_material(m: Material, dt: float) -> dict:
    """mp::set_derived_parameters (initialise_variables.cpp:419-432) and the anisotropy unrolling
    (anisotropy/initialize.cpp:50-143) for one material; dt is the reduced step dt_SI*gamma."""
    mu_s_SI = m.mu_s_SI
    one = -m.gamma_rel / (1.0 + m.alpha * m.alpha)
    mu_s_array = mu_s_SI / 9.27400915e-24
    inverse_mu_s = 1.0 / (mu_s_array * MU_B)
    e1 = list(input_unit_vector(m.kc_vector1))
    e2 = list(input_unit_vector(m.kc_vector2))
    e3 = [e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]]
    n1 = math.sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2])
    n2 = math.sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2])
    n3 = math.sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2])
    return dict(
        alpha=m.alpha,
        gamma_rel=m.gamma_rel,
        mu_s_SI=mu_s_SI,
        one_oneplusalpha_sq=one,
        alpha_oneplusalpha_sq=m.alpha * one,
        H_th_sigma=math.sqrt(2.0 * m.alpha * 1.3806503e-23 / (mu_s_SI * m.gamma_rel * dt)),
        temperature_rescaling_alpha=m.temperature_rescaling_alpha,
        temperature_rescaling_Tc=m.temperature_rescaling_Tc,
        ku2=m.Ku * inverse_mu_s,
        ku_vector=input_unit_vector(m.ku_vector),
        kc4=m.Kc * inverse_mu_s,
        kc_vector1=tuple(v / n1 for v in e1),
        kc_vector2=tuple(v / n2 for v in e2),
        kc_vector3=tuple(v / n3 for v in e3),
        applied_field_strength=m.applied_field_strength,
        applied_field_unit_vector=input_unit_vector(m.applied_field_unit_vector),
        mu_s_muB=mu_s_array,
    )


@dataclasses.dataclass
class Lattice:
    n: int
    coords: np.ndarray          # [n,3] Angstrom
    uc_id: np.ndarray           # [n] unit-cell atom index
    cell: np.ndarray            # [n,3] unit-cell coordinates
    nbr_start: np.ndarray       # [n] inclusive
    nbr_end: np.ndarray         # [n] inclusive
    nbr: np.ndarray             # [nnz]
    tmpl: np.ndarray            # [nnz] template interaction index
    vec: np.ndarray             # [nnz,3] i->j vector (Angstrom), periodic images included
    template: list


def build_lattice(crystal: str, nx: int, ny: int, nz: int, a: float = 3.54,
                  pbc=(True, True, True), keep: Optional[np.ndarray] = None,
                  vectors: bool = False) -> Lattice:
    """Atoms of an nx*ny*nz block of unit cells (order z, y, x, atom) and their neighbour list
    in template order with periodic wrap (generate.cpp:266-378).  `keep` (bool per site) removes
    sites, compacting atom ids in order, as create_system_type does for dilution.  With `vectors`
    the i->j bond vectors (needed for DMI) are kept as well."""
    atoms, _ = UNIT_CELLS[crystal]
    na = len(atoms)
    tmpl = interaction_template(crystal)
    nsite = nx * ny * nz * na
    s = np.arange(nsite, dtype=np.int64)
    ua = (s % na).astype(np.int32)
    c = s // na
    cx = (c % nx).astype(np.int64)
    cy = ((c // nx) % ny).astype(np.int64)
    cz = (c // (nx * ny)).astype(np.int64)
    frac = np.asarray(atoms, dtype=np.float64)
    coords = np.stack([(cx + frac[ua, 0]) * a, (cy + frac[ua, 1]) * a, (cz + frac[ua, 2]) * a], axis=1)
    if keep is None:
        keep = np.ones(nsite, dtype=bool)
    keep = np.asarray(keep, dtype=bool)
    atom_id = np.cumsum(keep) - 1
    dims = (nx, ny, nz)
    # per unit-cell atom: the template rows that start there, in template order
    per_uc = [[t for t, it in enumerate(tmpl) if it[0] == u] for u in range(na)]
    zmax = max(len(p) for p in per_uc)
    nb = np.full((nsite, zmax), -1, dtype=np.int64)
    tid = np.full((nsite, zmax), -1, dtype=np.int32)
    vv = np.zeros((nsite, zmax, 3) if vectors else (1, 1, 3), dtype=np.float64)
    for u in range(na):
        su = s[ua == u]
        ccx, ccy, ccz = cx[ua == u], cy[ua == u], cz[ua == u]
        for k, t in enumerate(per_uc[u]):
            _, jb, dx, dy, dz = tmpl[t]
            ok = np.ones(su.shape, dtype=bool)
            shift = np.zeros((su.shape[0], 3))
            ncoord = []
            for axis, (cc, d) in enumerate(((ccx, dx), (ccy, dy), (ccz, dz))):
                q = cc + d
                if pbc[axis]:
                    hi = q >= dims[axis]
                    lo = q < 0
                    shift[hi, axis] += dims[axis] * a
                    shift[lo, axis] -= dims[axis] * a
                    q = np.where(hi, q - dims[axis], np.where(lo, q + dims[axis], q))
                else:
                    ok &= (q >= 0) & (q < dims[axis])
                    q = np.clip(q, 0, dims[axis] - 1)
                ncoord.append(q)
            js = ((ncoord[2] * ny + ncoord[1]) * nx + ncoord[0]) * na + jb
            ok &= keep[su] & keep[js]
            nb[su[ok], k] = js[ok]
            tid[su[ok], k] = t
            if vectors:
                vv[su[ok], k, :] = coords[js[ok]] - coords[su[ok]] + shift[ok]
    nb = nb[keep]
    tid = tid[keep]
    if vectors:
        vv = vv[keep]
    valid = nb >= 0
    counts = valid.sum(axis=1)
    n = int(keep.sum())
    row = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=row[1:])
    nbr = atom_id[nb[valid]].astype(np.int32)
    return Lattice(
        n=n,
        coords=coords[keep],
        uc_id=ua[keep],
        cell=np.stack([cx, cy, cz], axis=1)[keep].astype(np.int32),
        nbr_start=row[:-1].astype(np.int32),
        nbr_end=(row[1:] - 1).astype(np.int32),
        nbr=nbr,
        tmpl=tid[valid].astype(np.int32),
        vec=vv[valid] if vectors else np.zeros((0, 3)),
        template=tmpl,
    )


@dataclasses.dataclass
class System:
    """The atoms::/mp:: state handed to the device boundary."""
    n: int
    type: np.ndarray                 # int32 [n]
    moment: np.ndarray               # float64 [n] (mu_B)
    nbr_start: np.ndarray            # int32 [n]
    nbr_end: np.ndarray              # int32 [n]
    nbr: np.ndarray                  # int32 [nnz]
    nbr_iid: np.ndarray              # int32 [nnz]
    exchange_type: int
    exchange: np.ndarray             # float64 [num_interactions, 1|3|9]
    materials: List[dict]            # derived_material() dicts
    anisotropy_flags: int
    dt: float                        # reduced step
    half_dt: float
    replica_offset: Optional[np.ndarray] = None   # int32 [R+1]
    coords: Optional[np.ndarray] = None

    @property
    def nnz(self) -> int:
        return int(self.nbr.shape[0])

    @property
    def num_replicas(self) -> int:
        return 1 if self.replica_offset is None else int(self.replica_offset.shape[0] - 1)


def exchange_tables(lat: Lattice, types: np.ndarray, mats: List[Material], J_SI: np.ndarray,
                    exchange_type: int, dmi_SI: Optional[np.ndarray] = None,
                    dmi_cutoff: float = 2.6):
    """Exchange in the reference's normalised form (unroll_normalised.cpp:37-140).

    J_SI[mi][mj] (J) -> Jij_matrix = J_SI/mu_s_i (initialise_variables.cpp:427-429); the template
    J of a nearest-neighbour interaction is 1.  Isotropic/vectorial: one table entry per material
    pair (edges refer to it by id).  Tensorial (or with DMI): one 3x3 tensor per edge, with the DMI
    of dmi.cpp:46-221 added when dmi_SI is given."""
    nm = len(mats)
    mu_s = np.array([m.mu_s_SI for m in mats])
    J_SI = np.asarray(J_SI, dtype=np.float64)
    # J_SI[mi][mj] or, for vectorial exchange, J_SI[mi][mj][k] (Jxx, Jyy, Jzz: the three values of
    # material[i]:exchange-matrix[j], src/exchange/interface.cpp:86-95)
    Jm = J_SI / mu_s.reshape((nm,) + (1,) * (J_SI.ndim - 1))
    row_atom = np.repeat(np.arange(lat.n), lat.nbr_end - lat.nbr_start + 1)
    mi = types[row_atom]
    mj = types[lat.nbr]
    if dmi_SI is not None:
        exchange_type = TENSORIAL
    if exchange_type == ISOTROPIC:
        table = Jm.reshape(nm * nm, 1).copy()
        iid = (mi * nm + mj).astype(np.int32)
        return exchange_type, table, iid
    if exchange_type == VECTORIAL:
        # unroll_normalised.cpp:66-90: template J (1 for a nearest neighbour) times Jij_matrix[k]
        table = Jm.reshape(nm * nm, 3).copy() if Jm.ndim == 3 else np.repeat(Jm.reshape(nm * nm, 1), 3, axis=1)
        iid = (mi * nm + mj).astype(np.int32)
        return exchange_type, table, iid
    nnz = lat.nbr.shape[0]
    ten = np.zeros((nnz, 9))
    if Jm.ndim == 3:
        raise ValueError("tensorial exchange takes one J per material pair")
    jv = Jm[mi, mj]
    ten[:, 0] = 1.0 * jv
    ten[:, 4] = 1.0 * jv
    ten[:, 8] = 1.0 * jv
    if dmi_SI is not None:
        ten += dmi_tensors(lat, types, mats, np.asarray(dmi_SI, dtype=np.float64), dmi_cutoff)
    return TENSORIAL, ten, np.arange(nnz, dtype=np.int32)


def dmi_tensors(lat: Lattice, types: np.ndarray, mats: List[Material], dmi_SI: np.ndarray,
                cutoff: float) -> np.ndarray:
    """Per-edge DMI contribution of exchange::internal::calculate_dmi (dmi.cpp:46-221): for the
    pair i-j, every other neighbour k of i with |r_ik|, |r_jk| <= cutoff adds
    D_ij = 0.5 (D[mi][mk] + D[mj][mk]) / mu_s_i along e_ik x e_jk, antisymmetrically.
    Vectorised over (i, j); the mediating neighbours k are visited in row order, as in the
    reference loop, so each tensor entry is accumulated in the same order."""
    cutoff_sq = cutoff * cutoff
    n = lat.n
    cnt = (lat.nbr_end - lat.nbr_start + 1).astype(np.int64)
    zmax = int(cnt.max()) if n else 0
    out = np.zeros((lat.nbr.shape[0], 9))
    if zmax == 0:
        return out
    mu_s = np.array([m.mu_s_SI for m in mats])
    slot = np.arange(zmax)[None, :]
    P = lat.nbr_start[:, None].astype(np.int64) + slot
    valid = slot < cnt[:, None]
    P = np.where(valid, P, 0)
    V = lat.vec[P]                       # [n, z, 3]
    NB = np.where(valid, lat.nbr[P], -1)  # [n, z]
    imat = types[np.arange(n)]
    i_mu_s = 1.0 / mu_s[imat]            # [n]
    jmat = np.where(valid, types[np.maximum(NB, 0)], 0)
    T = np.zeros((n, zmax, 9))
    for kk in range(zmax):
        nk = NB[:, kk]
        vk = valid[:, kk]
        kmat = jmat[:, kk]
        eik = V[:, kk, :]                                     # [n, 3]
        mod_eik_sq = eik[:, 0] * eik[:, 0] + eik[:, 1] * eik[:, 1] + eik[:, 2] * eik[:, 2]
        ejk = eik[:, None, :] - V                             # [n, z, 3]
        mod_ejk_sq = ejk[..., 0] * ejk[..., 0] + ejk[..., 1] * ejk[..., 1] + ejk[..., 2] * ejk[..., 2]
        ok = (valid & vk[:, None] & (NB != nk[:, None]) & (NB != np.arange(n)[:, None])
              & (mod_eik_sq <= cutoff_sq)[:, None] & (mod_ejk_sq <= cutoff_sq))
        if not ok.any():
            continue
        with np.errstate(divide="ignore", invalid="ignore"):
            inv_rik = 1.0 / np.sqrt(mod_eik_sq)
            inv_rjk = 1.0 / np.sqrt(mod_ejk_sq)
            e_ik = eik * inv_rik[:, None]                     # [n, 3]
            e_jk = ejk * inv_rjk[..., None]                   # [n, z, 3]
        Dij = 0.5 * (dmi_SI[imat, kmat][:, None] + dmi_SI[jmat, kmat[:, None]]) * i_mu_s[:, None]
        a0, a1, a2 = e_ik[:, 0:1], e_ik[:, 1:2], e_ik[:, 2:3]
        Dx = a1 * e_jk[..., 2] - a2 * e_jk[..., 1]
        Dy = a2 * e_jk[..., 0] - a0 * e_jk[..., 2]
        Dz = a0 * e_jk[..., 1] - a1 * e_jk[..., 0]
        for idx, val in ((1, Dij * +Dz), (2, Dij * -Dy), (3, Dij * -Dz), (5, Dij * +Dx),
                         (6, Dij * +Dy), (7, Dij * -Dx)):
            T[..., idx] = np.where(ok, T[..., idx] - val, T[..., idx])
    out[P[valid]] = T[valid]
    return out


def make_system(lat: Lattice, types: np.ndarray, mats: List[Material], J_SI, dt_SI: float = 1e-16,
                exchange_type: int = ISOTROPIC, dmi_SI=None, dmi_cutoff: float = 2.6,
                uniaxial: bool = False, cubic: bool = False, cubic_rotated: bool = False) -> System:
    dt = dt_SI * GAMMA_SI
    dm = [derived_material(m, dt) for m in mats]
    et, table, iid = exchange_tables(lat, types, mats, np.asarray(J_SI, dtype=np.float64), exchange_type,
                                     dmi_SI, dmi_cutoff)
    flags = (1 if uniaxial else 0) | (2 if cubic else 0) | (4 if cubic_rotated else 0)
    moment = np.array([dm[t]["mu_s_muB"] for t in range(len(mats))])[types]
    return System(n=lat.n, type=np.ascontiguousarray(types, dtype=np.int32),
                  moment=np.ascontiguousarray(moment, dtype=np.float64),
                  nbr_start=lat.nbr_start, nbr_end=lat.nbr_end, nbr=lat.nbr, nbr_iid=iid,
                  exchange_type=et, exchange=np.ascontiguousarray(table, dtype=np.float64), materials=dm,
                  anisotropy_flags=flags, dt=dt, half_dt=0.5 * dt, coords=lat.coords)


def concat_systems(systems: List[System]) -> System:
    """Block-diagonal union of independent replicas (atom ids and neighbour ids offset)."""
    base = systems[0]
    for s in systems[1:]:
        if s.exchange_type != base.exchange_type or s.anisotropy_flags != base.anisotropy_flags:
            raise ValueError("replicas must share exchange type and anisotropy terms")
        if len(s.materials) != len(base.materials) or s.dt != base.dt:
            raise ValueError("replicas must share the material table and time step")
    off = np.zeros(len(systems) + 1, dtype=np.int64)
    for r, s in enumerate(systems):
        off[r + 1] = off[r] + s.n
    eoff = np.zeros(len(systems) + 1, dtype=np.int64)
    for r, s in enumerate(systems):
        eoff[r + 1] = eoff[r] + s.nnz
    toff = np.zeros(len(systems) + 1, dtype=np.int64)
    for r, s in enumerate(systems):
        toff[r + 1] = toff[r] + s.exchange.shape[0]
    cat = lambda xs: np.concatenate(xs)
    return System(
        n=int(off[-1]),
        type=cat([s.type for s in systems]).astype(np.int32),
        moment=cat([s.moment for s in systems]),
        nbr_start=cat([s.nbr_start + eoff[r] for r, s in enumerate(systems)]).astype(np.int32),
        nbr_end=cat([s.nbr_end + eoff[r] for r, s in enumerate(systems)]).astype(np.int32),
        nbr=cat([s.nbr + off[r] for r, s in enumerate(systems)]).astype(np.int32),
        nbr_iid=cat([s.nbr_iid + toff[r] for r, s in enumerate(systems)]).astype(np.int32),
        exchange_type=base.exchange_type,
        exchange=np.ascontiguousarray(cat([s.exchange for s in systems])),
        materials=base.materials,
        anisotropy_flags=base.anisotropy_flags,
        dt=base.dt, half_dt=base.half_dt,
        replica_offset=off.astype(np.int32),
    )


def random_spins(n: int, seed: int) -> np.ndarray:
    """Uniform directions on the sphere from normalised Gaussian triples
    (cs_set_atom_vars2.cpp:103-117), drawn from numpy's generator with `seed`."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((3, n))
    g /= np.sqrt((g * g).sum(axis=0))
    return np.ascontiguousarray(g)


# ---------------------------------------------------------------------------------------------
# Workloads of the benchmark and parity suite
J_CO = 6.064e-21
KU_CO = 4.644e-24


def co_material(alpha: float = 0.1, Ku: float = KU_CO) -> Material:
    return Material(mu_s_muB=1.72, alpha=alpha, gamma_rel=1.0, Ku=Ku, ku_vector=(0.0, 0.0, 1.0))


def ferromagnet(crystal: str, L: int, alpha: float = 0.1, J: float = J_CO, Ku: float = KU_CO,
                dt_SI: float = 1e-16, Ly: Optional[int] = None, Lz: Optional[int] = None,
                pbc=(True, True, True)) -> System:
    """Single-material ferromagnet on an L^3 (or L x Ly x Lz) block, scalar nearest-neighbour J,
    uniaxial anisotropy along z."""
    lat = build_lattice(crystal, L, Ly or L, Lz or L, pbc=pbc)
    types = np.zeros(lat.n, dtype=np.int32)
    return make_system(lat, types, [co_material(alpha, Ku)], [[J]], dt_SI=dt_SI, uniaxial=Ku != 0.0)


def config_c0() -> System:
    """SC 20^3 periodic ferromagnet, J=6.064e-21 J, Ku=4.644e-24 J along z, mu_s=1.72 mu_B,
    alpha=0.1, dt=1e-16 s (8000 atoms)."""
    return ferromagnet("sc", 20, alpha=0.1)


def config_c1(L: int = 100) -> System:
    """FCC 100^3 unit cells (4,000,000 atoms), 12 nearest neighbours, alpha=0.05."""
    return ferromagnet("fcc", L, alpha=0.05)


# Curie temperature of the classical Heisenberg simple-cubic ferromagnet with the Co-like C0
# parameters: k_B Tc = 1.443 J (high-temperature series), J = 6.064e-21 J -> Tc = 634 K.
TC_SC_CO = 1.443 * J_CO / K_B


def config_c2(n_replicas: int = 256, seed: int = 11, L_min: int = 8, L_max: int = 48):
    """Batch of periodic SC ferromagnets with the C0 parameters and alpha = 1: edge lengths
    uniform in [L_min, L_max] cells, temperatures uniform over [0, 1.5 Tc].  Returns the list of
    replica systems and their temperatures (K); concat_systems() gives the batched system."""
    rng = np.random.default_rng(seed)
    Ls = rng.integers(L_min, L_max + 1, size=n_replicas)
    temps = rng.uniform(0.0, 1.5 * TC_SC_CO, size=n_replicas)
    systems = [ferromagnet("sc", int(L), alpha=1.0) for L in Ls]
    return systems, [float(t) for t in temps]


def config_c3(L: int = 32) -> System:
    """BCC 32^3 unit cells (65,536 atoms), deterministic (T=0) fitting workload."""
    return ferromagnet("bcc", L, alpha=0.1)


def config_c4(nx: int = 256, ny: int = 256, nz: int = 8, dilution: float = 0.1, seed: int = 3,
              dmi: float = 0.3e-21, a: float = 2.5, dmi_cutoff: float = 3.6) -> System:
    """Two-sublattice ferrimagnetic SC film (periodic in-plane, open in z), layers alternate
    mu_s = 1.6 / 7.6 mu_B, intra-layer-material J=+5e-21 J, inter-material J=-1.2e-21 J,
    interfacial DMI between the two materials, 10% random site dilution."""
    rng = np.random.default_rng(seed)
    keep = rng.random(nx * ny * nz) >= dilution
    lat = build_lattice("sc", nx, ny, nz, a=a, pbc=(True, True, False), keep=keep, vectors=bool(dmi))
    types = (lat.cell[:, 2] % 2).astype(np.int32)
    mats = [Material(mu_s_muB=1.6, alpha=0.1), Material(mu_s_muB=7.6, alpha=0.1)]
    J = [[5e-21, -1.2e-21], [-1.2e-21, 5e-21]]
    D = None
    if dmi:
        D = [[0.0, dmi], [dmi, 0.0]]
    return make_system(lat, types, mats, J, exchange_type=TENSORIAL, dmi_SI=D, dmi_cutoff=dmi_cutoff)
