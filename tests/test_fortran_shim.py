"""The Fortran drop-in module sht (magic_amd/fortran/sht_mhx.f90) through the
reference's own calling code.

oracle/_ref/shim_driver is the 'api' replay of oracle/ref_harness.f90
compiled against the drop-in instead of the reference's src/sht_native.f90,
with the reference's truncation, blocking and useful modules (built by
oracle/Makefile).  The GPU test feeds it the inputs of every golden api file
and compares with the outputs the reference's module sht wrote for them.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import load_golden, rel_l2
from api_cases import GOLDEN_API_FILES, API_OUTPUTS, api_levels, write_api_input, \
    read_api_output

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "oracle", "_ref", "shim_driver")
TOL = 1e-12

need_shim = pytest.mark.skipif(not os.path.exists(SHIM),
                               reason="oracle/_ref/shim_driver not built (make -C oracle)")


def run_shim(g, tmp_path):
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    lev = api_levels(g)
    write_api_input(fin, lev)
    r = subprocess.run([SHIM, str(int(g["l_max"])), str(int(g["minc"])), "20", fin, fout,
                        str(len(lev))], capture_output=True, text=True, timeout=120)
    return r, fout, len(lev)


@need_shim
def test_shim_fails_loudly_without_gpu(tmp_path):
    """With no GPU the drop-in stops through the reference's abortRun
    (src/useful.f90:271-302) at initialize_sht, before any transform."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    g = load_golden("api_L16_m1.npz")
    r, fout, _ = run_shim(g, tmp_path)
    text = r.stdout + r.stderr
    assert "Something went wrong, MagIC will stop now" in text
    assert "initialize_sht" in text
    assert not os.path.exists(fout)


@pytest.mark.gpu
@need_shim
@pytest.mark.parametrize("fname", GOLDEN_API_FILES)
def test_shim_matches_reference_module_sht(fname, tmp_path):
    g = load_golden(fname)
    print("shim_driver on", fname, flush=True)
    r, fout, nlev = run_shim(g, tmp_path)
    assert r.returncode == 0 and os.path.exists(fout), r.stdout + r.stderr
    hd, res = read_api_output(fout, nlev)
    assert hd["n_theta"] == int(g["n_theta"]) and hd["n_phi"] == int(g["n_phi"])
    for lev in range(nlev):
        for name, _kind, _cnt in API_OUTPUTS:
            want = g["out_%s_%d" % (name, lev)]
            got = res[lev][name]
            assert got.shape == want.shape, name
            err = rel_l2(got, want)
            assert err <= TOL, (fname, lev, name, err)
