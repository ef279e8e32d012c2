"""ctypes binding of libvmx.so (include/vmx.h).

The library is loaded from the package directory only; there is no fallback implementation.
A missing library raises ``VmxUnavailable`` at first use, so nothing silently runs elsewhere.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvmx.so")

VMX_OK, VMX_ERR_ARG, VMX_ERR_HIP, VMX_ERR_STATE, VMX_ERR_NOMEM = 0, 1, 2, 3, 4
VMX_EXCHANGE_ISOTROPIC, VMX_EXCHANGE_VECTORIAL, VMX_EXCHANGE_TENSORIAL = 0, 1, 2
VMX_FP64, VMX_FP32 = 0, 1
VMX_KERNEL_PLAIN, VMX_KERNEL_FUSED = 0, 1
VMX_TOPOLOGY_AUTO, VMX_TOPOLOGY_CSR, VMX_TOPOLOGY_ELL, VMX_TOPOLOGY_STENCIL = 0, 1, 2, 3
VMX_ANISOTROPY_UNIAXIAL_SECOND_ORDER = 0x1
VMX_ANISOTROPY_CUBIC_FOURTH_ORDER = 0x2
VMX_ANISOTROPY_CUBIC_FOURTH_ROTATION = 0x4
VMX_FIELD_THERMAL = 0x1
VMX_FIELD_APPLIED = 0x2
VMX_FIELD_LOCAL_APPLIED = 0x4
VMX_CMC_MOVE_CONE, VMX_CMC_MOVE_ANGLE = 0, 1


class VmxUnavailable(RuntimeError):
    """libvmx.so is not built or cannot be loaded."""


class VmxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vmx error {code}: {msg}")
        self.code = code


class vmx_material(C.Structure):
    _fields_ = [
        ("alpha", C.c_double),
        ("gamma_rel", C.c_double),
        ("mu_s_SI", C.c_double),
        ("one_oneplusalpha_sq", C.c_double),
        ("alpha_oneplusalpha_sq", C.c_double),
        ("H_th_sigma", C.c_double),
        ("temperature_rescaling_alpha", C.c_double),
        ("temperature_rescaling_Tc", C.c_double),
        ("ku2", C.c_double),
        ("ku_vector", C.c_double * 3),
        ("kc4", C.c_double),
        ("kc_vector1", C.c_double * 3),
        ("kc_vector2", C.c_double * 3),
        ("kc_vector3", C.c_double * 3),
        ("applied_field_strength", C.c_double),
        ("applied_field_unit_vector", C.c_double * 3),
    ]


_P = C.POINTER


class vmx_system(C.Structure):
    _fields_ = [
        ("num_atoms", C.c_int32),
        ("num_materials", C.c_int32),
        ("type", _P(C.c_int32)),
        ("moment", _P(C.c_double)),
        ("nbr_start", _P(C.c_int32)),
        ("nbr_end", _P(C.c_int32)),
        ("num_nbr", C.c_int64),
        ("nbr", _P(C.c_int32)),
        ("nbr_iid", _P(C.c_int32)),
        ("exchange_type", C.c_int32),
        ("num_interactions", C.c_int64),
        ("exchange", _P(C.c_double)),
        ("materials", _P(vmx_material)),
        ("anisotropy_flags", C.c_uint32),
        ("num_replicas", C.c_int32),
        ("replica_offset", _P(C.c_int32)),
    ]


class vmx_options(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("kernel", C.c_int32),
        ("block_size", C.c_int32),
        ("device", C.c_int32),
        ("seed", C.c_uint64),
        ("topology", C.c_int32),
        ("recompute_euler", C.c_int32),
        ("atom_order", C.c_int32),
        ("keep_external_field", C.c_int32),
        ("fp64_normals", C.c_int32),
    ]


class vmx_fit_params(C.Structure):
    _fields_ = [
        ("J", C.c_double),
        ("ku2", C.c_double),
        ("alpha", C.c_double),
        ("H", C.c_double * 3),
    ]


class vmx_dipole(C.Structure):
    _fields_ = [
        ("num_cells", C.c_int32),
        ("num_local_cells", C.c_int32),
        ("local_cell", _P(C.c_int32)),
        ("atom_cell", _P(C.c_int32)),
        ("magnetic", _P(C.c_int32)),
        ("volume", _P(C.c_double)),
        ("tensor", _P(C.c_double)),
        ("update_rate", C.c_int32),
    ]


_VP = C.c_void_p
_DP = _P(C.c_double)

# (name, restype, argtypes) for every symbol declared in include/vmx.h
SIGNATURES = [
    ("vmx_last_error", C.c_char_p, []),
    ("vmx_version", C.c_char_p, []),
    ("vmx_device_count", C.c_int, [_P(C.c_int32)]),
    ("vmx_create", C.c_int, [_P(vmx_system), _P(vmx_options), _P(_VP)]),
    ("vmx_destroy", C.c_int, [_VP]),
    ("vmx_get_topology", C.c_int, [_VP, _P(C.c_int32), _P(C.c_int32)]),
    ("vmx_get_atom_period", C.c_int, [_VP, _P(C.c_int32)]),
    ("vmx_set_spins", C.c_int, [_VP, _DP, _DP, _DP]),
    ("vmx_get_spins", C.c_int, [_VP, _DP, _DP, _DP]),
    ("vmx_set_time_step", C.c_int, [_VP, C.c_double, C.c_double]),
    ("vmx_set_conditions", C.c_int, [_VP, C.c_int32, C.c_double, _DP, C.c_uint32]),
    ("vmx_set_demag_field", C.c_int, [_VP, C.c_int32, _DP]),
    ("vmx_llg_heun", C.c_int, [_VP, C.c_int32, _DP]),
    ("vmx_llg_midpoint", C.c_int, [_VP, C.c_int32, _DP]),
    ("vmx_llg_heun_graph", C.c_int, [_VP, C.c_int32, C.c_int32]),
    ("vmx_get_step", C.c_int, [_VP, _P(C.c_int64)]),
    ("vmx_set_step", C.c_int, [_VP, C.c_int64]),
    ("vmx_get_sweep", C.c_int, [_VP, _P(C.c_int64)]),
    ("vmx_set_sweep", C.c_int, [_VP, C.c_int64]),
    ("vmx_mc_sweep", C.c_int, [_VP, C.c_int32]),
    ("vmx_mc_get_state", C.c_int, [_VP, _DP, _DP, _DP]),
    ("vmx_mc_set_sigma", C.c_int, [_VP, C.c_int32, C.c_double]),
    ("vmx_mc_num_colours", C.c_int, [_VP, _P(C.c_int32)]),
    ("vmx_cmc_set_constraint", C.c_int, [_VP, C.c_int32, C.c_double, C.c_double]),
    ("vmx_cmc_set_move", C.c_int, [_VP, C.c_int32]),
    ("vmx_cmc_sweep", C.c_int, [_VP, C.c_int32]),
    ("vmx_cmc_get_counts", C.c_int, [_VP, _DP]),
    ("vmx_stats_add_magnetization", C.c_int, [_VP, C.c_int32, _P(C.c_int32), _P(C.c_int32)]),
    ("vmx_stats_add_energy", C.c_int, [_VP, C.c_int32, _P(C.c_int32), _P(C.c_int32)]),
    ("vmx_stats_update", C.c_int, [_VP]),
    ("vmx_stats_update_noalloc", C.c_int, [_VP]),
    ("vmx_stats_get_magnetization", C.c_int, [_VP, C.c_int32, _DP, _DP, _P(C.c_int64)]),
    ("vmx_stats_get_energy", C.c_int, [_VP, C.c_int32, _DP, _DP, _P(C.c_int64)]),
    ("vmx_stats_reset", C.c_int, [_VP]),
    ("vmx_compute_fields", C.c_int, [_VP, _DP, _DP, _DP]),
    ("vmx_thermal_noise", C.c_int, [_VP, C.c_int64, _DP]),
    ("vmx_synchronize", C.c_int, [_VP]),
    ("vmx_last_step_time_ms", C.c_int, [_VP, _DP]),
    ("vmx_time_kernels", C.c_int, [_VP, C.c_int32, _DP, _P(C.c_int32)]),
    ("vmx_fit_forward", C.c_int, [_VP, _P(vmx_fit_params), _DP, _DP, C.c_int32, C.c_int32, C.c_int32, _DP]),
    ("vmx_fit_backward", C.c_int, [_VP, _DP, _DP, _DP]),
    ("vmx_autotune", C.c_int, [_VP, C.c_int32, _P(C.c_int32)]),
    ("vmx_checkpoint_bytes", C.c_int, [_VP, _P(C.c_uint64)]),
    ("vmx_checkpoint_save", C.c_int, [_VP, C.c_void_p, C.c_uint64]),
    ("vmx_checkpoint_load", C.c_int, [_VP, C.c_void_p, C.c_uint64]),
    ("vmx_get_launch_config", C.c_int, [_VP, _P(C.c_int32), _P(C.c_int32)]),
    ("vmx_set_dipole", C.c_int, [_VP, _P(vmx_dipole)]),
    ("vmx_update_dipole", C.c_int, [_VP]),
    ("vmx_get_dipole_field", C.c_int, [_VP, _DP, _DP, _DP, _DP, _DP, _DP]),
    ("vmx_set_atom_field", C.c_int, [_VP, _DP, _DP, _DP]),
]

_lib = None


def load() -> C.CDLL:
    """Loads libvmx.so from the package directory (raises VmxUnavailable if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VmxUnavailable(
            f"{LIB_PATH} is not built; run __graft_entry__.build() or python -m vampire_amd.build")
    try:
        lib = C.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - depends on the ROCm runtime
        raise VmxUnavailable(f"cannot load {LIB_PATH}: {e}") from e
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != VMX_OK:
        msg = load().vmx_last_error().decode(errors="replace")
        raise VmxError(rc, msg)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_P(C.c_int32))
