"""The radial loop for any physics switches (mhx_rloop) against the
reference's own radial loop: for each fixture (one radialLoopG call of the
reference program on a restart state of its samples or on its own initial
fields after a few steps, tests/golden/make_rloop_golden.py), the interior
levels' explicit terms and Courant steps from the same inputs, within
1e-12 relative L2 per output (FP64 on the GPU).  Fixtures cover Boussinesq
MHD with curl advection (boussBenchSat, conducting rotating inner core),
u.grad(u) advection, non-magnetic convection with composition
(doubleDiffusion), an anelastic dynamo with viscous and Ohmic heating and
the centrifugal term, precession, and the phase field."""
import numpy as np
import pytest

from conftest import rel_l2
from magic_amd import sht
import rloop_cases as rc

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _run(d, device, precision=sht.PREC_FP64):
    import torch
    sel = rc.interior(d)
    nlev = sel.size
    p = sht.Plan(**rc.plan_args(d), precision=precision, max_levels=nlev)
    fields = rc.fields_of(d, sel)
    names = rc.outputs_of(d)
    if device:
        dev = torch.device("cuda")
        fields = {k: (torch.from_numpy(v).to(dev) if v is not None else None) for k, v in fields.items()}
        outs = {n: torch.zeros((nlev, p.lm_max), dtype=torch.complex128, device=dev) for n in names}
    else:
        outs = {n: np.zeros((nlev, p.lm_max), dtype=np.complex128) for n in names}
    dtr, dth = np.zeros(nlev), np.zeros(nlev)
    sht.rloop(p, rc.phys_of(d), nlev, rc.radial_of(d, sel), fields, outs, dtr, dth)
    if device:
        torch.cuda.synchronize()
        outs = {k: v.cpu().numpy() for k, v in outs.items()}
    p.close()
    return sel, outs, dtr, dth


@pytest.mark.parametrize("case", rc.CASES)
def test_rloop_matches_reference_radial_loop(case):
    d = rc.load(case)
    print("rloop", case, "l_max", int(d["l_max"]), "minc", int(d["minc"]), flush=True)
    sel, outs, dtr, dth = _run(d, device=False)
    for n, got in outs.items():
        want = d[n][sel]
        e = rel_l2(got, want)
        print("  %-8s %.2e" % (n, e), flush=True)
        assert e < TOL, (case, n, e)
    assert rel_l2(dtr, d["dtrkc"][sel]) < TOL, (case, dtr, d["dtrkc"][sel])
    assert rel_l2(dth, d["dthkc"][sel]) < TOL, (case, dth, d["dthkc"][sel])


def test_rloop_device_arrays_equal_host_arrays():
    """device tensors in and out give the same numbers as host arrays"""
    d = rc.load("anel_mhd")
    _, host, h1, h2 = _run(d, device=False)
    _, dev, d1, d2 = _run(d, device=True)
    for n in host:
        assert np.array_equal(host[n], dev[n]), n
    assert np.array_equal(h1, d1) and np.array_equal(h2, d2)


def test_rloop_refuses_missing_inputs():
    d = rc.load("doublediff")
    sel = rc.interior(d)
    p = sht.Plan(**rc.plan_args(d), max_levels=sel.size)
    fields = rc.fields_of(d, sel)
    fields["xi"] = None
    outs = {n: np.zeros((sel.size, p.lm_max), dtype=np.complex128) for n in rc.outputs_of(d)}
    with pytest.raises(sht.MhxError, match="xi"):
        sht.rloop(p, rc.phys_of(d), sel.size, rc.radial_of(d, sel), fields, outs, np.zeros(sel.size),
                  np.zeros(sel.size))
    p.close()


def test_rloop_split_plan_runs_the_fused_split_step():
    """A split-precision plan with the Boussinesq MHD switches and curl
    advection (the switch sets the fused step covers) runs the fused split
    step inside mhx_rloop: its explicit terms within 2e-5 and its Courant
    steps within 1e-5 of the reference's loop (the split form's bounds on
    this state), and device arrays give the host arrays' numbers."""
    d = rc.load("bouss_mhd")
    print("rloop bouss_mhd, split plan", flush=True)
    sel, outs, dtr, dth = _run(d, device=False, precision=sht.PREC_SPLIT16)
    for n, got in outs.items():
        e = rel_l2(got, d[n][sel])
        print("  %-8s %.2e" % (n, e), flush=True)
        assert e < 2e-5, (n, e)
    np.testing.assert_allclose(dtr, d["dtrkc"][sel], rtol=1e-5, atol=0)
    np.testing.assert_allclose(dth, d["dthkc"][sel], rtol=1e-5, atol=0)
    _, dev, d1, d2 = _run(d, device=True, precision=sht.PREC_SPLIT16)
    for n in outs:
        assert np.array_equal(outs[n], dev[n]), n
    assert np.array_equal(dtr, d1) and np.array_equal(dth, d2)
