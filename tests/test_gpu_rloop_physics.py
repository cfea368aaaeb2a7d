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


@pytest.mark.parametrize("precision", [sht.PREC_FP64, sht.PREC_SPLIT16])
def test_rloop_fused_route_skips_dudr_at_l255(precision):
    """mhx_rloop's fused route runs the step without the du/dr grids (they
    are not among the loop's outputs).  At l_max=255 (the theta-row tile grid
    stage in the split form, the slot stage in FP64), against
    magic_amd.rloop.RadialLoop, which runs the full step with them: in FP64
    the explicit terms, dVSrLM and Courant steps are identical (the skipped
    work feeds no output); in the split form each is as close to the FP64
    full step as the split full step is, within a factor 2 (the tile stage
    pairs components in one complex FFT, so the empty du/dr halves change
    FP32 rounding, and a field's error follows its partner's size: VxBp
    (dbdt) shares its forward FFT with Advr)."""
    import torch
    from magic_amd import rloop, workload
    L, nlev = 255, 5
    nt, nph = workload.grid_size(L)
    print("rloop fused route l_max=255, precision", precision, flush=True)
    ins = workload.nl_inputs(L, nlev, stream=7)
    r = np.linspace(0.6, 1.4, nlev)
    one, zero = np.ones(nlev), np.zeros(nlev)
    dxr, dxh = np.full(nlev, 1e-3), np.full(nlev, 2e-2)
    ph = {"l_conv": 1, "l_conv_nl": 1, "l_heat": 1, "l_heat_nl": 1, "l_mag": 1, "l_mag_LF": 1,
          "l_mag_nl": 1, "l_adv_curl": 1, "l_corr": 1, "l_cour_alf_damp": 1, "LFfac": 200.0,
          "CorFac": 1000.0, "opm": 0.2, "courfac": 0.8, "alffac": 0.35, "oek": 1000.0,
          "epsc": 0.5}
    rad = {"r": r, "or1": 1 / r, "or2": 1 / r ** 2, "or4": 1 / r ** 4, "orho1": one, "beta": zero,
           "otemp1": one, "visc": one, "lambda": one, "temp0": one, "epscProf": one,
           "delxr2": dxr, "delxh2": dxh, "nR": np.arange(2, nlev + 2, dtype=np.int32),
           "l_R": np.full(nlev, L, dtype=np.int32)}
    names = ("dwdt", "dzdt", "dpdt", "dsdt", "dbdt", "djdt", "dVxBhLM", "dVSrLM")

    def full_step(prec):
        p = sht.Plan(L, nt, nph, precision=prec, max_levels=nlev)
        rl = rloop.RadialLoop(p, nlev)
        td, nl, d1, d2 = rl.bulk(ins, rad["or2"], rad["or4"], rad["orho1"], ph["LFfac"], ph["CorFac"],
                                 ph["epsc"] * rad["epscProf"], rad["l_R"], ph["courfac"], ph["alffac"],
                                 ph["opm"], dxr, dxh)
        torch.cuda.synchronize()
        o = {n: td[k].cpu().numpy() for k, n in enumerate(sht.TD_OUTPUTS)}
        o["dVSrLM"] = nl[3].cpu().numpy()
        p.close()
        return o, d1, d2

    p = sht.Plan(L, nt, nph, precision=precision, max_levels=nlev)
    fields = dict(zip(sht.NL_INPUTS, ins))
    outs = {n: np.zeros((nlev, p.lm_max), dtype=np.complex128) for n in names}
    dtr, dth = np.zeros(nlev), np.zeros(nlev)
    sht.rloop(p, ph, nlev, rad, fields, outs, dtr, dth)
    p.close()
    want, dtr2, dth2 = full_step(precision)
    if precision == sht.PREC_FP64:
        for n in names:
            assert np.array_equal(outs[n], want[n]), n
        assert np.array_equal(dtr, dtr2) and np.array_equal(dth, dth2)
        return
    ref, dtr64, dth64 = full_step(sht.PREC_FP64)
    for n in names:
        e_route, e_full = rel_l2(outs[n], ref[n]), rel_l2(want[n], ref[n])
        print("  %-8s vs FP64: route %.2e, full split step %.2e" % (n, e_route, e_full), flush=True)
        assert e_route <= 2 * e_full + 1e-7, (n, e_route, e_full)
    for a, b, c in ((dtr, dtr2, dtr64), (dth, dth2, dth64)):
        assert np.max(np.abs(a / c - 1)) <= 2 * np.max(np.abs(b / c - 1)) + 1e-7


def test_rloop_own_kernels_at_l255_match_oracle():
    """The loop's own kernels at l_max=255 (n_phi = 768: the grid stage
    k_rl_grid_s<3, 1>, rows of 384 points by wave slots), with the u.grad u
    dump's switches (outside the fused step's sets), on two levels of
    synthetic fields against the oracle's radial loop (FP64, 1e-11)."""
    from oracle import oracle
    from rloop_reference import oracle_rloop
    from magic_amd import workload
    L, nlev = 255, 2
    nt, nph = workload.grid_size(L)
    d = rc.load("bouss_nocurl")
    ph = rc.phys_of(d)
    assert not ph["l_adv_curl"]
    print("rloop own kernels l_max=255 x", nlev, flush=True)
    ins = workload.nl_inputs(L, nlev, stream=11)
    fields = {n: None for n in sht.RLOOP_IN}
    fields.update(dict(zip(sht.NL_INPUTS, ins)))
    r = np.array([0.8, 1.2])
    one, zero = np.ones(nlev), np.zeros(nlev)
    rad = {"r": r, "or1": 1 / r, "or2": 1 / r ** 2, "or4": 1 / r ** 4, "orho1": one, "beta": zero,
           "otemp1": one, "visc": one, "lambda": one, "temp0": one, "epscProf": one,
           "delxr2": np.full(nlev, 1e-3), "delxh2": np.full(nlev, 2e-2),
           "nR": np.array([5, 9], dtype=np.int32), "l_R": np.full(nlev, L, dtype=np.int32)}
    names = rc.outputs_of(d)
    p = sht.Plan(L, nt, nph, max_levels=nlev)
    outs = {n: np.zeros((nlev, p.lm_max), dtype=np.complex128) for n in names}
    dtr, dth = np.zeros(nlev), np.zeros(nlev)
    sht.rloop(p, ph, nlev, rad, fields, outs, dtr, dth)
    p.close()
    o = oracle.Plan(L, nt, nph)
    want, wtr, wth = oracle_rloop(o, ph, nlev, rad, fields, 1)
    for n in names:
        e = rel_l2(outs[n], want[n])
        print("  %-8s %.2e" % (n, e), flush=True)
        assert e < 1e-11, (n, e)
    assert rel_l2(dtr, wtr) < 1e-11 and rel_l2(dth, wth) < 1e-11
