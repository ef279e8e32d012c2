"""The C ABI library: loads, exports every function include/vmx.h declares, and rejects bad
arguments on the host before any device work (no compute calls; runs without a GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from vampire_amd import _lib as L
from vampire_amd import system as S
from vampire_amd.context import material_struct

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "vmx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vmx_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = L.load()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(n for n, _, _ in L.SIGNATURES) == names


def test_exported_symbols_are_c_linkage():
    out = os.popen(f"nm -D --defined-only {L.LIB_PATH}").read()
    exported = set(re.findall(r" T (vmx_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported


def test_version_and_error_text():
    lib = L.load()
    assert b"gfx950" in lib.vmx_version()
    assert isinstance(lib.vmx_last_error(), bytes)


def _sys_struct(s, **over):
    keep = dict(t=np.ascontiguousarray(s.type, np.int32), st=np.ascontiguousarray(s.nbr_start, np.int32),
                en=np.ascontiguousarray(s.nbr_end, np.int32), nb=np.ascontiguousarray(s.nbr, np.int32),
                iid=np.ascontiguousarray(s.nbr_iid, np.int32), mom=np.ascontiguousarray(s.moment),
                ex=np.ascontiguousarray(s.exchange))
    keep.update(over)
    mats = (L.vmx_material * len(s.materials))(*[material_struct(d) for d in s.materials])
    st = L.vmx_system(num_atoms=s.n, num_materials=len(s.materials), type=L.iptr(keep["t"]),
                      moment=L.dptr(keep["mom"]), nbr_start=L.iptr(keep["st"]), nbr_end=L.iptr(keep["en"]),
                      num_nbr=keep["nb"].shape[0], nbr=L.iptr(keep["nb"]), nbr_iid=L.iptr(keep["iid"]),
                      exchange_type=s.exchange_type, num_interactions=keep["ex"].shape[0],
                      exchange=L.dptr(keep["ex"]), materials=mats, anisotropy_flags=s.anisotropy_flags,
                      num_replicas=1, replica_offset=None)
    return st, keep, mats


@pytest.mark.parametrize("bad", ["neighbour", "iid", "type", "range", "precision", "block", "empty"])
def test_create_rejects_bad_input_on_host(bad):
    lib = L.load()
    s = S.ferromagnet("sc", 3)
    over = {}
    opt = L.vmx_options(precision=L.VMX_FP64, kernel=L.VMX_KERNEL_FUSED, block_size=0, device=-1, seed=0)
    if bad == "neighbour":
        nb = s.nbr.copy(); nb[5] = s.n + 3; over["nb"] = nb
    elif bad == "iid":
        iid = s.nbr_iid.copy(); iid[0] = 99; over["iid"] = iid
    elif bad == "type":
        t = s.type.copy(); t[2] = 4; over["t"] = t
    elif bad == "range":
        en = s.nbr_end.copy(); en[-1] = s.nbr.shape[0] + 10; over["en"] = en
    elif bad == "precision":
        opt.precision = 7
    elif bad == "block":
        opt.block_size = 100
    st, keep, mats = _sys_struct(s, **over)
    if bad == "empty":
        st.num_atoms = 0
    h = C.c_void_p()
    rc = lib.vmx_create(C.byref(st), C.byref(opt), C.byref(h))
    assert rc == L.VMX_ERR_ARG
    assert lib.vmx_last_error()
    assert not h.value


def test_null_context_is_an_argument_error():
    lib = L.load()
    assert lib.vmx_llg_heun(None, 1, None) == L.VMX_ERR_ARG
    assert lib.vmx_stats_update(None) == L.VMX_ERR_ARG
    assert lib.vmx_destroy(None) == L.VMX_OK


def test_library_refuses_to_load_when_missing(tmp_path, monkeypatch):
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "libvmx.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(L.VmxUnavailable):
        L.load()
