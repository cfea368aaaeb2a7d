"""Split-precision form (scaled FP16 x3 Legendre MFMA, FP32 accumulation,
FP32 Fourier data and grids): every output field within 1e-6 relative L2 of
the FP64 result, on the oracle at l_max=64 and against the FP64 GPU form
(itself held to 1e-12 of the oracle) at the full configuration
l_max=255, 384 x 768, 97 levels."""
import numpy as np
import pytest

from conftest import rel_l2
from magic_amd import sht, workload
from nl_reference import make_inputs, oracle_step, radial

pytestmark = pytest.mark.gpu

SPLIT_TOL = 1e-6


def _run(p, inputs, rad, lcut):
    import torch
    dev = torch.device("cuda")
    nlev = inputs[0].shape[0]
    tin = [torch.from_numpy(x).to(dev) for x in inputs]
    tout = [torch.zeros((nlev, p.lm_max), dtype=torch.complex128, device=dev) for _ in range(10)]
    sht.nl_step(p, tin, tout, rad[0], rad[1], rad[2], rad[3], lcut)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in tout]


def test_split_matches_oracle_l64():
    from oracle import oracle
    L, nlev = 64, 10
    nt, nph = workload.grid_size(L)
    p = sht.Plan(L, nt, nph, precision=sht.PREC_SPLIT16, max_levels=nlev)
    o = oracle.Plan(L, nt, nph)
    inputs = make_inputs(L, nlev)
    rad = radial(nlev)
    lcut = [L] * nlev
    lcut[3] = L - 7
    outs = _run(p, inputs, rad, lcut)
    _, ref = oracle_step(o, inputs, *rad, lcut)
    errs = [rel_l2(outs[k], ref[k]) for k in range(10)]
    print("split vs oracle l_max=64:", " ".join("%.2e" % e for e in errs), flush=True)
    for k, e in enumerate(errs):
        assert e < SPLIT_TOL, (sht.NL_OUTPUTS[k], e)


def test_split_full_config_vs_fp64():
    L, nlev = 255, 97
    nt, nph = workload.grid_size(L)
    inputs = make_inputs(L, nlev)
    rad = radial(nlev)
    lcut = [L] * nlev
    p64 = sht.Plan(L, nt, nph, max_levels=nlev)
    ref = _run(p64, inputs, rad, lcut)
    p64.close()
    p16 = sht.Plan(L, nt, nph, precision=sht.PREC_SPLIT16, max_levels=nlev)
    outs = _run(p16, inputs, rad, lcut)
    errs = [rel_l2(outs[k], ref[k]) for k in range(10)]
    print("split vs FP64 l_max=255 x 97:", " ".join("%.2e" % e for e in errs), flush=True)
    for k, e in enumerate(errs):
        assert e < SPLIT_TOL, (sht.NL_OUTPUTS[k], e)
    # level independence holds in the split form too
    sub = [x[5:6] for x in inputs]
    subrad = tuple(np.asarray(r)[5:6] if np.ndim(r) else r for r in rad)
    o1 = _run(p16, sub, subrad, [L])
    for k in range(10):
        assert rel_l2(o1[k][0], outs[k][5]) < 1e-12


def test_split_l511_dipole_vs_fp64():
    """The large-shell case: l_max=511, 768 x 1536, l^-5/3 spectra with the
    l = 1 poloidal magnetic field 100x larger, on 8 levels (one with a lower
    lcut); every output within 1e-6 of the FP64 form."""
    L, nlev = 511, 8
    nt, nph = workload.grid_size(L)
    inputs = workload.nl_inputs(L, nlev, stream=4, dipole=100.0)
    rad = radial(nlev)
    lcut = [L] * nlev
    lcut[2] = L - 40
    p64 = sht.Plan(L, nt, nph, max_levels=nlev)
    ref = _run(p64, inputs, rad, lcut)
    p64.close()
    p16 = sht.Plan(L, nt, nph, precision=sht.PREC_SPLIT16, max_levels=nlev)
    outs = _run(p16, inputs, rad, lcut)
    p16.close()
    errs = [rel_l2(outs[k], ref[k]) for k in range(10)]
    print("split vs FP64 l_max=511 x 8, dipole x100:", " ".join("%.2e" % e for e in errs), flush=True)
    for k, e in enumerate(errs):
        assert e < SPLIT_TOL, (sht.NL_OUTPUTS[k], e)


@pytest.mark.parametrize("L,nlev", [(511, 6), (1023, 2)])
def test_split_recurrence_tables(L, nlev):
    """PREC_SPLIT16_REC: the synthesis and analysis tables generated in-kernel
    by the FP64 recurrence from FP64 seeds (no stored FP16 tables) give every
    output within 1e-6 of the FP64 form, one level with a lower lcut; at
    l_max=511 they also agree with the stored-table split form to 5e-7 (the
    two differ only by the rounding of the table entries: each is ~3e-7 from
    FP64)."""
    print("split recurrence tables", L, nlev, flush=True)
    nt, nph = workload.grid_size(L)
    inputs = make_inputs(L, nlev)
    rad = radial(nlev)
    lcut = [L] * nlev
    lcut[-1] = L - 33
    p64 = sht.Plan(L, nt, nph, max_levels=nlev)
    ref = _run(p64, inputs, rad, lcut)
    p64.close()
    pr = sht.Plan(L, nt, nph, precision=sht.PREC_SPLIT16_REC, max_levels=nlev)
    outs = _run(pr, inputs, rad, lcut)
    pr.close()
    errs = [rel_l2(outs[k], ref[k]) for k in range(10)]
    print("recurrence split vs FP64:", " ".join("%.2e" % e for e in errs), flush=True)
    for k, e in enumerate(errs):
        assert e < SPLIT_TOL, (sht.NL_OUTPUTS[k], e)
    if L <= 511:
        ps = sht.Plan(L, nt, nph, precision=sht.PREC_SPLIT16, max_levels=nlev)
        st = _run(ps, inputs, rad, lcut)
        ps.close()
        d = [rel_l2(outs[k], st[k]) for k in range(10)]
        print("recurrence vs stored tables:", " ".join("%.2e" % e for e in d), flush=True)
        assert max(d) < 5e-7


def test_split_recurrence_plan_limits():
    """the recurrence form is refused outside 255 < l_max <= 1023"""
    with pytest.raises(sht.MhxError):
        sht.Plan(127, 192, 384, precision=sht.PREC_SPLIT16_REC)


def test_split_graph_replay_matches_eager():
    """The split step is captured into a graph on its second call with the same
    arrays and replayed on later calls.  Every replay, with the outputs zeroed
    and other device work and synchronisations in between, equals the eager
    first call bit for bit; after the inputs change in place, a replay equals
    the eager call of a fresh plan on the new inputs."""
    import torch
    L, nlev = 127, 24
    nt, nph = workload.grid_size(L)
    inputs = make_inputs(L, nlev)
    rad = radial(nlev)
    lcut = [L] * nlev
    lcut[5] = L - 11
    dev = torch.device("cuda")
    tin = [torch.from_numpy(x).to(dev) for x in inputs]
    p16 = sht.Plan(L, nt, nph, precision=sht.PREC_SPLIT16, max_levels=nlev)
    tout = [torch.zeros((nlev, p16.lm_max), dtype=torch.complex128, device=dev) for _ in range(10)]

    def step(plan, outs):
        sht.nl_step(plan, tin, outs, rad[0], rad[1], rad[2], rad[3], lcut)
        torch.cuda.synchronize()
        return [t.cpu().numpy() for t in outs]

    ref = step(p16, tout)
    for rep in range(4):
        a = torch.rand(512, 512, device=dev)
        (a @ a).sum().item()
        for t in tout:
            t.zero_()
        torch.cuda.synchronize()
        got = step(p16, tout)
        for k in range(10):
            assert np.array_equal(got[k], ref[k]), (rep, sht.NL_OUTPUTS[k])
    for t in tin:
        t.mul_(0.5).add_(t.roll(1, dims=1))
    got = step(p16, tout)
    fresh = sht.Plan(L, nt, nph, precision=sht.PREC_SPLIT16, max_levels=nlev)
    tout2 = [torch.zeros_like(t) for t in tout]
    want = step(fresh, tout2)
    fresh.close()
    p16.close()
    for k in range(10):
        assert np.array_equal(got[k], want[k]), sht.NL_OUTPUTS[k]
        assert not np.array_equal(got[k], ref[k]), sht.NL_OUTPUTS[k]


def test_grid_tile_form_matches_slot_form():
    """The split step's two grid-stage kernels on the l_max=255 grid: the
    theta-row tile kernel (the default for n_phi = 768) against the one-row
    slot kernel (MHX_GRID_FORM=slots at plan creation), same inputs: the ten
    outputs, the 18 synthesised grids and the Courant maxima agree to FP32
    rounding (both forms are held to 1e-6 of FP64 at the full size by
    test_split_full_config_vs_fp64, which runs the tile form)."""
    import os
    import torch
    L, nlev = 255, 6
    nt, nph = workload.grid_size(L)
    inputs = make_inputs(L, nlev)
    rad = radial(nlev)
    lcut = [L] * nlev
    lcut[2] = L - 5
    dev = torch.device("cuda")
    res = {}
    for form in ("slots", "tile"):
        if form == "slots":
            os.environ["MHX_GRID_FORM"] = "slots"
        try:
            p = sht.Plan(L, nt, nph, precision=sht.PREC_SPLIT16, max_levels=nlev)
        finally:
            os.environ.pop("MHX_GRID_FORM", None)
        print("grid form %s: l_max=%d x %d levels" % (form, L, nlev), flush=True)
        tin = [torch.from_numpy(x).to(dev) for x in inputs]
        tout = [torch.zeros((nlev, p.lm_max), dtype=torch.complex128, device=dev) for _ in range(10)]
        grids = torch.zeros((nlev, 18, nph, nt), dtype=torch.float32, device=dev)
        sht.nl_step(p, tin, tout, rad[0], rad[1], rad[2], rad[3], lcut, grids)
        torch.cuda.synchronize()
        outs = [t.cpu().numpy() for t in tout]
        g = grids.cpu().numpy()
        vr2, vh2 = sht.nl_step_courant(p, tin, tout, rad[0], rad[1], rad[2], rad[3], lcut, 2.5, 1.0,
                                       [0.01] * nlev, [0.02] * nlev)
        p.close()
        res[form] = (outs, g, vr2, vh2)
    (o1, g1, r1, h1), (o2, g2, r2, h2) = res["slots"], res["tile"]
    errs = [rel_l2(o2[k], o1[k]) for k in range(10)]
    gerr = [rel_l2(g2[:, c], g1[:, c]) for c in range(18)]
    print("tile vs slots outputs:", " ".join("%.1e" % e for e in errs), flush=True)
    print("tile vs slots grids:", " ".join("%.1e" % e for e in gerr), flush=True)
    # two FP32 FFT orderings (radices 12.4.4.2 against 16.4.12, two real
    # rows per complex FFT) each about 1.5e-7 from the FP64 result
    for k, e in enumerate(errs):
        assert e < 5e-7, (sht.NL_OUTPUTS[k], e)
    for c, e in enumerate(gerr):
        assert e < 5e-7, (c, e)
    np.testing.assert_allclose(r2, r1, rtol=1e-5)
    np.testing.assert_allclose(h2, h1, rtol=1e-5)


def test_recurrence_plan_builds_fp64_tables_on_first_use():
    """A recurrence-table plan (l_max = 1023) keeps no FP64 tables at
    creation (they would take 12.9 GB at this size): its split kernels
    generate theirs in-kernel, and its row factors come from the recurrence.
    The first per-call FP64 entry builds the tables; its result equals a
    plain FP64 plan's bit for bit."""
    import torch
    L = 1023
    nt, nph = workload.grid_size(L)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    p = sht.Plan(L, nt, nph, precision=sht.PREC_SPLIT16_REC, max_levels=2)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    print("recurrence plan l_max=1023: %.2f GB at creation" % ((free0 - free1) / 1e9), flush=True)
    assert free0 - free1 < 3e9, free0 - free1
    ls, ms = workload.lm_index(L)
    rng = np.random.default_rng(21)
    s = (rng.standard_normal(ls.size) + 1j * rng.standard_normal(ls.size)) / (ls + 1.0) ** 2
    s[ms == 0] = s[ms == 0].real
    got = p.scal_to_spat(s, L)
    torch.cuda.synchronize()
    free2 = torch.cuda.mem_get_info()[0]
    print("after the first FP64 entry: %.2f GB more" % ((free1 - free2) / 1e9), flush=True)
    assert free1 - free2 > 1e10, free1 - free2
    p.close()
    p64 = sht.Plan(L, nt, nph, max_levels=1)
    want = p64.scal_to_spat(s, L)
    p64.close()
    assert np.array_equal(got, want)
