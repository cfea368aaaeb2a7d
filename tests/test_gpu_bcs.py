"""Boundary products of a conducting mantle or inner core (mhx_get_br_v_bcs,
src/nonlinear_bcs.f90:24-68): against the reference's own get_br_v_bcs
(golden, l_max=16 and 64 with minc=4) and the FP64 oracle at l_max=255 with
host and device grids.  FP64: <= 1e-12 relative L2 per output.  The
nonlinear boundary terms built from them (mhx_get_b_nl_bcs,
src/nonlinear_bcs.f90:71-118): bit for bit against the reference's
get_b_nl_bcs('CMB') and ('ICB') on the same goldens."""
import ctypes

import numpy as np
import pytest

from conftest import load_golden, rel_l2
from magic_amd import sht, workload

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("fname", ["brv_L16_m1.npz", "brv_L64_m4.npz"])
def test_br_v_bcs_matches_reference_golden(fname):
    g = load_golden(fname)
    print("brv golden", fname, flush=True)
    p = sht.Plan(int(g["l_max"]), int(g["n_theta"]), int(g["n_phi"]), minc=int(g["minc"]))
    for lev in range(g["grids"].shape[0]):
        a, b = p.get_br_v_bcs(*g["grids"][lev], g["omega"][lev], g["or2"][lev], g["orho"][lev])
        assert rel_l2(a, g["br_vt_lm"][lev]) < TOL, (fname, lev)
        assert rel_l2(b, g["br_vp_lm"][lev]) < TOL, (fname, lev)
    p.close()


def test_br_v_bcs_matches_oracle_device_grids():
    """l_max=255 grids synthesised from (l+1)^-2 fields (the boundary
    level's br, vt, vp), passed as host arrays and as device tensors."""
    import torch
    from oracle import oracle
    L = 255
    nt, nph = workload.grid_size(L)
    print("brv oracle", L, flush=True)
    o = oracle.Plan(L, nt, nph)
    W, dW, Z = (workload.coefficients(L, 31, f, 0) for f in range(3))
    br = o.scal_to_spat(W * 40.0, L)
    vr, vt, vp = o.torpol_to_spat(W, dW, Z, L)
    omega, or2, orho = 0.37, 1.0 / 1.538 ** 2, 0.9
    ra, rb = o.get_br_v_bcs(br, vt, vp, omega, or2, orho)
    p = sht.Plan(L, nt, nph)
    a, b = p.get_br_v_bcs(br, vt, vp, omega, or2, orho)
    assert rel_l2(a, ra) < TOL and rel_l2(b, rb) < TOL
    dev = torch.device("cuda")
    tg = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (br, vt, vp)]
    ta = torch.zeros(p.lm_max, dtype=torch.complex128, device=dev)
    tb = torch.zeros_like(ta)
    vp_ = ctypes.c_void_p
    sht._check(p._L.mhx_get_br_v_bcs(p.h, vp_(tg[0].data_ptr()), vp_(tg[1].data_ptr()),
                                     vp_(tg[2].data_ptr()), ctypes.c_double(omega),
                                     ctypes.c_double(or2), ctypes.c_double(orho),
                                     vp_(ta.data_ptr()), vp_(tb.data_ptr())))
    assert rel_l2(ta.cpu().numpy(), ra) < TOL and rel_l2(tb.cpu().numpy(), rb) < TOL
    # the inputs are left untouched (the products go to the plan's workspace)
    assert np.array_equal(tg[0].cpu().numpy(), br) and np.array_equal(tg[2].cpu().numpy(), vp)
    p.close()


@pytest.mark.parametrize("fname", ["brv_L16_m1.npz", "brv_L64_m4.npz"])
def test_b_nl_bcs_matches_reference_golden(fname):
    import torch
    g = load_golden(fname)
    print("b_nl golden", fname, flush=True)
    p = sht.Plan(int(g["l_max"]), int(g["n_theta"]), int(g["n_phi"]), minc=int(g["minc"]))
    dev = torch.device("cuda")
    for lev in range(g["br_vt_lm"].shape[0]):
        a, b = g["br_vt_lm"][lev], g["br_vp_lm"][lev]
        bc, jc = p.get_b_nl_bcs("CMB", a, b, g["conductance_ma"][lev], g["prmag"][lev])
        nb, ji = p.get_b_nl_bcs("ICB", a, b, g["sigma_ratio"][lev], g["prmag"][lev])
        assert nb is None
        assert np.array_equal(bc, g["b_nl_cmb"][lev]) and np.array_equal(jc, g["aj_nl_cmb"][lev])
        assert np.array_equal(ji, g["aj_nl_icb"][lev])
        # device arrays in and out
        ta, tb = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (a, b))
        to, tj = torch.zeros_like(ta), torch.zeros_like(ta)
        vp_ = ctypes.c_void_p
        sht._check(p._L.mhx_get_b_nl_bcs(p.h, sht.BC_CMB, vp_(ta.data_ptr()), vp_(tb.data_ptr()),
                                         ctypes.c_double(g["conductance_ma"][lev]),
                                         ctypes.c_double(g["prmag"][lev]), vp_(to.data_ptr()),
                                         vp_(tj.data_ptr())))
        assert np.array_equal(to.cpu().numpy(), g["b_nl_cmb"][lev])
        assert np.array_equal(tj.cpu().numpy(), g["aj_nl_cmb"][lev])
    z = np.zeros(p.lm_max, dtype=np.complex128)
    with pytest.raises(sht.MhxError):
        sht._check(p._L.mhx_get_b_nl_bcs(p.h, 2, sht._ptr(z), sht._ptr(z), 1.0, 1.0, sht._ptr(z),
                                         sht._ptr(z)))
    p.close()
