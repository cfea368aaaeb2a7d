"""CPU checks of the C-ABI boundary: libmhx.so loads, exports every symbol
include/mhx.h declares, and the ctypes signatures match the header."""
import os
import re

import pytest

from magic_amd import sht

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    h = open(os.path.join(ROOT, "include", "mhx.h")).read()
    h = re.sub(r"/\*.*?\*/", "", h, flags=re.S)
    out = {}
    for m in re.finditer(r"(?:int|const char \*)\s*(mhx_\w+)\s*\(([^)]*)\)\s*;", h):
        args = [a.strip() for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        kinds = ""
        for a in args:
            if "mhx_plan" in a and "**" not in a:
                kinds += "P"
            elif "*" in a:
                kinds += "p"
            elif a.startswith("double"):
                kinds += "d"
            elif a.startswith("size_t"):
                kinds += "z"
            elif a.startswith("int"):
                kinds += "i"
            else:
                kinds += "?"
        out[m.group(1)] = kinds
    return out


def test_library_exports_every_header_symbol():
    if not os.path.exists(sht.LIB_PATH):
        pytest.skip("libmhx.so not built")
    L = sht.lib()
    funcs = header_functions()
    assert len(funcs) >= 25
    for name in funcs:
        assert hasattr(L, name), name


def test_ctypes_signatures_match_header():
    funcs = header_functions()
    for name, kinds in funcs.items():
        if name == "mhx_last_error":
            continue
        assert sht._SIGS.get(name) == kinds, (name, kinds, sht._SIGS.get(name))
    assert set(sht.exported_symbols()) == set(funcs)


def test_module_sht_names_are_mirrored():
    """Every public routine of the reference's module sht
    (src/sht_native.f90:16-20) has an entry point and a Plan method."""
    names = ["scal_to_spat", "scal_to_grad_spat", "pol_to_grad_spat", "torpol_to_spat",
             "pol_to_curlr_spat", "torpol_to_curl_spat", "torpol_to_dphspat", "scal_to_SH",
             "spat_to_sphertor", "torpol_to_spat_IC", "torpol_to_curl_spat_IC", "axi_to_spat",
             "spat_to_qst", "sphtor_to_spat", "toraxi_to_spat"]
    funcs = header_functions()
    for n in names:
        assert "mhx_" + n in funcs, n
        assert hasattr(sht.Plan, n), n
    assert hasattr(sht.Plan, "finalize_sht")


def test_plan_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    if not os.path.exists(sht.LIB_PATH):
        pytest.skip("libmhx.so not built")
    with pytest.raises(sht.MhxError):
        sht.Plan(16)


def test_level_entries_reject_null_plan():
    """The lm-space and boundary entries return an error (no GPU touched)
    for a null plan, with a message."""
    import ctypes
    L = sht.lib()
    vp = ctypes.c_void_p
    z = vp(0)
    calls = {
        "mhx_get_br_v_bcs": lambda: L.mhx_get_br_v_bcs(None, z, z, z, 0.0, 1.0, 1.0, z, z),
        "mhx_get_e_kin": lambda: L.mhx_get_e_kin(None, 1, z, z, z, 1, z, z, z),
        "mhx_get_td": lambda: L.mhx_get_td(None, 1, z, 1, z, z, z, z, 1.0, z, None),
        "mhx_nl_step_courant": lambda: L.mhx_nl_step_courant(None, 1, z, 1, z, z, z, z, 1.0, z,
                                                             1.0, 1.0, z, z, z, z, None),
    }
    for name, call in calls.items():
        rc = call()
        assert rc < 0, name
        msg = L.mhx_last_error()
        assert msg and b"null" in (msg if isinstance(msg, bytes) else msg.encode()), name
