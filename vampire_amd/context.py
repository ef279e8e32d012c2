"""Python handle on a libvmx context (include/vmx.h).

``Context`` owns one device-side copy of a ``System`` and exposes the calls that VAMPIRE's
``gpu::`` boundary makes (hdr/gpu.hpp:33-56), plus the diagnostics the parity tests use.
Everything runs in libvmx.so; this module only marshals numpy arrays.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .system import System

FP64 = "fp64"
FP32 = "fp32"
PLAIN = "plain"
FUSED = "fused"
TOPOLOGIES = {"auto": 0, "csr": 1, "ell": 2, "stencil": 3}


def device_count() -> int:
    lib = L.load()
    c = C.c_int32(0)
    L.check(lib.vmx_device_count(C.byref(c)))
    return int(c.value)


def material_struct(d: dict) -> L.vmx_material:
    m = L.vmx_material()
    for key in ("alpha", "gamma_rel", "mu_s_SI", "one_oneplusalpha_sq", "alpha_oneplusalpha_sq", "H_th_sigma",
                "temperature_rescaling_alpha", "temperature_rescaling_Tc", "ku2", "kc4", "applied_field_strength"):
        setattr(m, key, float(d[key]))
    for key in ("ku_vector", "kc_vector1", "kc_vector2", "kc_vector3", "applied_field_unit_vector"):
        arr = getattr(m, key)
        for k in range(3):
            arr[k] = float(d[key][k])
    return m


class Context:
    """A device context for one system (or a batch of replicas).

    fp64_normals (FP32 only) draws each step's thermal normals with the FP64 generator, so an
    FP32 run and an FP64 run of the same seed consume the same normals; a measurement mode, one
    more kernel per step."""

    def __init__(self, system: System, precision: str = FP64, kernel: str = FUSED, block_size: int = 0,
                 seed: int = 0, device: int = -1, topology: str = "auto", recompute_euler: bool = False,
                 atom_order: str = "auto", keep_external_field: bool = False, fp64_normals: bool = False):
        self.lib = L.load()
        self.system = system
        self.n = system.n
        self.precision = precision
        self.kernel = kernel
        # keep the arrays alive for the duration of vmx_create
        s = system
        self._arrs = [np.ascontiguousarray(x) for x in (s.type, s.nbr_start, s.nbr_end, s.nbr, s.nbr_iid)]
        t, st, en, nb, iid = self._arrs
        mom = np.ascontiguousarray(s.moment, dtype=np.float64)
        ex = np.ascontiguousarray(s.exchange, dtype=np.float64)
        mats = (L.vmx_material * len(s.materials))(*[material_struct(d) for d in s.materials])
        roff = None if s.replica_offset is None else np.ascontiguousarray(s.replica_offset, dtype=np.int32)
        sysc = L.vmx_system(
            num_atoms=s.n, num_materials=len(s.materials), type=L.iptr(t), moment=L.dptr(mom),
            nbr_start=L.iptr(st), nbr_end=L.iptr(en), num_nbr=nb.shape[0], nbr=L.iptr(nb),
            nbr_iid=L.iptr(iid), exchange_type=s.exchange_type, num_interactions=ex.shape[0],
            exchange=L.dptr(ex), materials=mats, anisotropy_flags=s.anisotropy_flags,
            num_replicas=s.num_replicas, replica_offset=None if roff is None else L.iptr(roff))
        opt = L.vmx_options(precision=L.VMX_FP64 if precision == FP64 else L.VMX_FP32,
                            kernel=L.VMX_KERNEL_FUSED if kernel == FUSED else L.VMX_KERNEL_PLAIN,
                            block_size=block_size, device=device, seed=seed,
                            topology=TOPOLOGIES[topology],
                            recompute_euler=int(bool(recompute_euler)),
                            atom_order={"auto": 0, "caller": 1}[atom_order],
                            keep_external_field=int(bool(keep_external_field)),
                            fp64_normals=int(bool(fp64_normals)))
        h = C.c_void_p()
        L.check(self.lib.vmx_create(C.byref(sysc), C.byref(opt), C.byref(h)))
        self.h = h
        self._arrs = None
        L.check(self.lib.vmx_set_time_step(self.h, s.dt, s.half_dt))

    # --- lifetime -------------------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            L.check(self.lib.vmx_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- state ------------------------------------------------------------------------------
    def set_spins(self, spins: np.ndarray) -> None:
        sp = np.ascontiguousarray(spins, dtype=np.float64).reshape(3, self.n)
        x, y, z = (np.ascontiguousarray(sp[k]) for k in range(3))
        L.check(self.lib.vmx_set_spins(self.h, L.dptr(x), L.dptr(y), L.dptr(z)))

    def get_spins(self) -> np.ndarray:
        out = np.empty((3, self.n), dtype=np.float64)
        x, y, z = np.empty(self.n), np.empty(self.n), np.empty(self.n)
        L.check(self.lib.vmx_get_spins(self.h, L.dptr(x), L.dptr(y), L.dptr(z)))
        out[0], out[1], out[2] = x, y, z
        return out

    def set_conditions(self, temperature: float, H: Sequence[float] = (0.0, 0.0, 0.0), thermal: bool = True,
                       applied: bool = True, local_applied: bool = False, replica: int = -1) -> None:
        flags = ((L.VMX_FIELD_THERMAL if thermal else 0) | (L.VMX_FIELD_APPLIED if applied else 0)
                 | (L.VMX_FIELD_LOCAL_APPLIED if local_applied else 0))
        h = (C.c_double * 3)(*[float(v) for v in H])
        L.check(self.lib.vmx_set_conditions(self.h, replica, float(temperature), h, flags))

    def set_demag_field(self, H: Sequence[float], replica: int = -1) -> None:
        """Thin-film demagnetising field (T): added to the step's applied field, not to the
        applied-field energy (calculate_applied_fields, src/simulate/fields.cpp:219-239)."""
        h = (C.c_double * 3)(*[float(v) for v in H])
        L.check(self.lib.vmx_set_demag_field(self.h, replica, h))

    def set_time_step(self, dt: float, half_dt: Optional[float] = None) -> None:
        L.check(self.lib.vmx_set_time_step(self.h, dt, 0.5 * dt if half_dt is None else half_dt))

    @property
    def topology(self):
        """(layout name, number of stencils) chosen for the neighbour list."""
        t, p = C.c_int32(0), C.c_int32(0)
        L.check(self.lib.vmx_get_topology(self.h, C.byref(t), C.byref(p)))
        return {1: "csr", 2: "ell", 3: "stencil"}[t.value], int(p.value)

    @property
    def atom_period(self) -> int:
        """Lattice period q of the device's basis-major atom order (1: the caller's order)."""
        v = C.c_int32(0)
        L.check(self.lib.vmx_get_atom_period(self.h, C.byref(v)))
        return int(v.value)

    @property
    def step(self) -> int:
        v = C.c_int64(0)
        L.check(self.lib.vmx_get_step(self.h, C.byref(v)))
        return int(v.value)

    @step.setter
    def step(self, value: int) -> None:
        L.check(self.lib.vmx_set_step(self.h, int(value)))

    @property
    def sweep(self) -> int:
        """Metropolis sweeps taken so far (the Philox counter of mc_sweep)."""
        v = C.c_int64(0)
        L.check(self.lib.vmx_get_sweep(self.h, C.byref(v)))
        return int(v.value)

    @sweep.setter
    def sweep(self, value: int) -> None:
        L.check(self.lib.vmx_set_sweep(self.h, int(value)))

    # --- integration ------------------------------------------------------------------------
    def llg_heun(self, nsteps: int = 1, noise: Optional[np.ndarray] = None) -> None:
        if noise is None:
            L.check(self.lib.vmx_llg_heun(self.h, int(nsteps), None))
        else:
            nz = np.ascontiguousarray(noise, dtype=np.float64)
            if nz.size != nsteps * 3 * self.n:
                raise ValueError(f"noise must hold {nsteps}x3x{self.n} values")
            L.check(self.lib.vmx_llg_heun(self.h, int(nsteps), L.dptr(nz.reshape(-1))))

    def llg_midpoint(self, nsteps: int = 1, noise: Optional[np.ndarray] = None) -> None:
        """nsteps steps of the LLG midpoint integrator (sim::LLG_Midpoint); noise as llg_heun."""
        if noise is None:
            L.check(self.lib.vmx_llg_midpoint(self.h, int(nsteps), None))
        else:
            nz = np.ascontiguousarray(noise, dtype=np.float64)
            if nz.size != nsteps * 3 * self.n:
                raise ValueError(f"noise must hold {nsteps}x3x{self.n} values")
            L.check(self.lib.vmx_llg_midpoint(self.h, int(nsteps), L.dptr(nz.reshape(-1))))

    def llg_heun_graph(self, nsteps: int, nrepeat: int) -> None:
        L.check(self.lib.vmx_llg_heun_graph(self.h, int(nsteps), int(nrepeat)))

    def synchronize(self) -> None:
        L.check(self.lib.vmx_synchronize(self.h))

    def last_step_time_ms(self) -> float:
        v = C.c_double(0.0)
        L.check(self.lib.vmx_last_step_time_ms(self.h, C.byref(v)))
        return float(v.value)

    def time_kernels(self, nsteps: int):
        """Average device ms of each kernel of a Heun step (HIP events around every launch)."""
        ms = np.zeros(5)
        nk = C.c_int32(0)
        L.check(self.lib.vmx_time_kernels(self.h, int(nsteps), L.dptr(ms), C.byref(nk)))
        return ms[:nk.value].copy()

    # --- checkpoints --------------------------------------------------------------------------
    def checkpoint(self) -> bytes:
        """The device state (spins, Heun step and Metropolis sweep counters, conditions,
        Metropolis cone widths) as a self-describing byte string."""
        nb = C.c_uint64(0)
        L.check(self.lib.vmx_checkpoint_bytes(self.h, C.byref(nb)))
        buf = C.create_string_buffer(nb.value)
        L.check(self.lib.vmx_checkpoint_save(self.h, buf, nb.value))
        return buf.raw

    def restore(self, data: bytes) -> None:
        """Loads a checkpoint written by a context of the same system, precision and seed."""
        buf = C.create_string_buffer(data, len(data))
        L.check(self.lib.vmx_checkpoint_load(self.h, buf, len(data)))

    def save_checkpoint(self, path: str) -> None:
        with open(path, "wb") as f:
            f.write(self.checkpoint())

    def load_checkpoint(self, path: str) -> None:
        with open(path, "rb") as f:
            self.restore(f.read())

    @property
    def launch_config(self):
        """(threads per workgroup, XCD remapping on) in use."""
        b, x = C.c_int32(0), C.c_int32(0)
        L.check(self.lib.vmx_get_launch_config(self.h, C.byref(b), C.byref(x)))
        return int(b.value), bool(x.value)

    def autotune(self, steps_per_trial: int = 20) -> int:
        v = C.c_int32(0)
        L.check(self.lib.vmx_autotune(self.h, int(steps_per_trial), C.byref(v)))
        return int(v.value)

    # --- macrocell dipole field ---------------------------------------------------------------
    def set_dipole(self, num_cells: int, local_cell, atom_cell, volume, tensor, update_rate: int,
                   magnetic=None, spins: Optional[np.ndarray] = None) -> None:
        """Device dipole solver (tensor method) on the reference's cells and cell-cell tensor
        ([6][len(local_cell)][num_cells]); the field of the current spins (or of `spins`, set
        first) is formed now and refreshed every `update_rate` Heun steps."""
        if spins is not None:
            self.set_spins(spins)
        lc = np.ascontiguousarray(local_cell, dtype=np.int32)
        ac = np.ascontiguousarray(atom_cell, dtype=np.int32)
        vol = np.ascontiguousarray(volume, dtype=np.float64)
        ten = np.ascontiguousarray(tensor, dtype=np.float64).reshape(-1)
        if ac.size != self.n or vol.size != num_cells or ten.size != 6 * lc.size * num_cells:
            raise ValueError("dipole arrays do not match the system and cell counts")
        mag = None if magnetic is None else np.ascontiguousarray(magnetic, dtype=np.int32)
        d = L.vmx_dipole(num_cells=int(num_cells), num_local_cells=int(lc.size),
                         local_cell=lc.ctypes.data_as(C.POINTER(C.c_int32)),
                         atom_cell=ac.ctypes.data_as(C.POINTER(C.c_int32)),
                         magnetic=None if mag is None else mag.ctypes.data_as(C.POINTER(C.c_int32)),
                         volume=L.dptr(vol), tensor=L.dptr(ten), update_rate=int(update_rate))
        L.check(self.lib.vmx_set_dipole(self.h, C.byref(d)))

    def clear_dipole(self) -> None:
        L.check(self.lib.vmx_set_dipole(self.h, None))

    def update_dipole(self) -> None:
        L.check(self.lib.vmx_update_dipole(self.h))

    def dipole_field(self):
        """(B, mu0 Hd): per-atom [3][n] fields of the last refresh, caller's order."""
        out = [np.empty(self.n) for _ in range(6)]
        L.check(self.lib.vmx_get_dipole_field(self.h, *[L.dptr(a) for a in out]))
        return np.stack(out[:3]), np.stack(out[3:])

    def set_atom_field(self, h: Optional[np.ndarray]) -> None:
        """A per-atom field [3][n] (tesla) added to H_ext in the dipole field's place; None removes it."""
        if h is None:
            L.check(self.lib.vmx_set_atom_field(self.h, None, None, None))
            return
        hh = np.ascontiguousarray(h, dtype=np.float64).reshape(3, self.n)
        x, y, z = (np.ascontiguousarray(hh[k]) for k in range(3))
        L.check(self.lib.vmx_set_atom_field(self.h, L.dptr(x), L.dptr(y), L.dptr(z)))

    # --- Monte Carlo -------------------------------------------------------------------------
    def mc_sweep(self, nsweeps: int = 1) -> None:
        L.check(self.lib.vmx_mc_sweep(self.h, int(nsweeps)))

    def mc_state(self):
        R = self.system.num_replicas
        s, m, r = np.empty(R), np.empty(R), np.empty(R)
        L.check(self.lib.vmx_mc_get_state(self.h, L.dptr(s), L.dptr(m), L.dptr(r)))
        return s, m, r

    def mc_set_sigma(self, sigma: float, replica: int = -1) -> None:
        L.check(self.lib.vmx_mc_set_sigma(self.h, replica, float(sigma)))

    def mc_num_colours(self) -> int:
        v = C.c_int32(0)
        L.check(self.lib.vmx_mc_num_colours(self.h, C.byref(v)))
        return int(v.value)

    # --- constrained Monte Carlo ----------------------------------------------------------------
    def cmc_set_constraint(self, phi: float, theta: float, replica: int = -1) -> None:
        """Constraint direction: polar angle phi and azimuth theta (degrees) of the magnetisation."""
        L.check(self.lib.vmx_cmc_set_constraint(self.h, replica, float(phi), float(theta)))

    def cmc_set_move(self, angle: bool) -> None:
        """False: cone of the replica's Metropolis width (default); True: the angle move."""
        L.check(self.lib.vmx_cmc_set_move(self.h, L.VMX_CMC_MOVE_ANGLE if angle else L.VMX_CMC_MOVE_CONE))

    def cmc_sweep(self, nsweeps: int = 1) -> None:
        L.check(self.lib.vmx_cmc_sweep(self.h, int(nsweeps)))

    def cmc_counts(self) -> np.ndarray:
        """[replicas][3]: accepted pairs, energy rejections, sphere rejections."""
        out = np.empty(3 * self.system.num_replicas)
        L.check(self.lib.vmx_cmc_get_counts(self.h, L.dptr(out)))
        return out.reshape(-1, 3)

    # --- statistics -----------------------------------------------------------------------------
    def add_magnetization(self, mask: np.ndarray, mask_size: int) -> int:
        m = np.ascontiguousarray(mask, dtype=np.int32)
        h = C.c_int32(0)
        L.check(self.lib.vmx_stats_add_magnetization(self.h, int(mask_size), L.iptr(m), C.byref(h)))
        self._msize = getattr(self, "_msize", {})
        self._msize[("m", h.value)] = mask_size
        return int(h.value)

    def add_energy(self, mask: np.ndarray, mask_size: int) -> int:
        m = np.ascontiguousarray(mask, dtype=np.int32)
        h = C.c_int32(0)
        L.check(self.lib.vmx_stats_add_energy(self.h, int(mask_size), L.iptr(m), C.byref(h)))
        self._msize = getattr(self, "_msize", {})
        self._msize[("e", h.value)] = mask_size
        return int(h.value)

    def stats_update(self) -> None:
        L.check(self.lib.vmx_stats_update(self.h))

    def stats_reset(self) -> None:
        L.check(self.lib.vmx_stats_reset(self.h))

    def magnetization(self, handle: int = 0):
        ms = self._msize[("m", handle)]
        cur, mean = np.empty(4 * ms), np.empty(4 * ms)
        cnt = C.c_int64(0)
        L.check(self.lib.vmx_stats_get_magnetization(self.h, handle, L.dptr(cur), L.dptr(mean), C.byref(cnt)))
        return cur, mean, int(cnt.value)

    def energy(self, handle: int = 0):
        ms = self._msize[("e", handle)]
        cur, mean = np.empty(5 * ms), np.empty(5 * ms)
        cnt = C.c_int64(0)
        L.check(self.lib.vmx_stats_get_energy(self.h, handle, L.dptr(cur), L.dptr(mean), C.byref(cnt)))
        return cur.reshape(5, ms), mean.reshape(5, ms), int(cnt.value)

    # --- diagnostics ------------------------------------------------------------------------------
    def fields(self, noise: Optional[np.ndarray] = None):
        hs, he = np.empty(3 * self.n), np.empty(3 * self.n)
        nz = None
        if noise is not None:
            nz = np.ascontiguousarray(noise, dtype=np.float64).reshape(-1)
        L.check(self.lib.vmx_compute_fields(self.h, None if nz is None else L.dptr(nz), L.dptr(hs), L.dptr(he)))
        return hs.reshape(3, self.n), he.reshape(3, self.n)

    def thermal_noise(self, step: int) -> np.ndarray:
        out = np.empty(3 * self.n)
        L.check(self.lib.vmx_thermal_noise(self.h, int(step), L.dptr(out)))
        return out.reshape(3, self.n)
