"""Pins the FP64 oracle (oracle/mhx_oracle.c).

* against golden vectors written by the reference's own Fortran modules
  (tests/golden/make_golden.py through oracle/_ref/ref_harness);
* against numpy (Gauss-Legendre nodes, real FFT) and closed forms
  (low-degree P_lm, orthonormality under Gauss quadrature, round trips,
  the sin(theta) dP/dtheta identity) so that a dropped term, a wrong sign or
  a transposed index fails somewhere independent of the golden files.
"""
import numpy as np
import pytest

from conftest import load_golden, rel_l2
from api_cases import API_CALLS, GOLDEN_API_FILES, call, levels
from oracle import oracle
from magic_amd import workload


# ---------------------------------------------------------------- golden
@pytest.mark.parametrize("fname", GOLDEN_API_FILES)
def test_oracle_matches_reference_api(fname):
    g = load_golden(fname)
    p = oracle.Plan(int(g["l_max"]), int(g["n_theta"]), int(g["n_phi"]), minc=int(g["minc"]))
    for lev in range(levels(g)):
        for name in API_CALLS:
            got = call(p, name, g, lev)
            ref = g["out_%s_%d" % (name, lev)]
            assert len(got) == ref.shape[0]
            for k in range(ref.shape[0]):
                err = rel_l2(got[k], ref[k])
                assert err < 1e-12, (fname, lev, name, k, err)


@pytest.mark.parametrize("fname", ["tables_L16_m1.npz", "tables_L64_m4.npz"])
def test_oracle_tables_match_reference(fname):
    g = load_golden(fname)
    p = oracle.Plan(int(g["l_max"]), int(g["n_theta"]), int(g["n_phi"]), minc=int(g["minc"]))
    t = p.tables()
    nh = int(g["n_theta"]) // 2
    np.testing.assert_allclose(t["theta"], g["theta_ord"][:nh], rtol=0, atol=1e-14)
    np.testing.assert_allclose(t["gauss"], g["gauss"][0::2][:nh], rtol=1e-13)
    np.testing.assert_allclose(t["osin2"], g["osin2"], rtol=1e-13)
    for k in ("plm", "dplm", "wplm", "wdplm"):
        scale = np.abs(g[k]).max()
        assert np.abs(t[k] - g[k]).max() <= 1e-13 * scale, k
    assert np.array_equal(p.lm2l, g["lm2l"]) and np.array_equal(p.lm2m, g["lm2m"])


def test_oracle_get_nl_matches_reference():
    g = load_golden("nl_L16.npz")
    p = oracle.Plan(int(g["l_max"]), int(g["n_theta"]), int(g["n_phi"]))
    or2, or4, orho1, lff = g["radial"]
    out = p.get_nl(or2, or4, orho1, lff, list(g["in_grids"]))
    for k in range(12):
        assert rel_l2(out[k], g["out"][k]) < 1e-14, k


def test_oracle_nl_step_matches_reference():
    """The nonlinear-term step composed from the oracle in the reference's
    call order against the reference's own composition (ref_harness nlstep)."""
    from nl_reference import make_inputs, oracle_step, radial
    g = load_golden("nlstep_L16.npz")
    L, nlev = int(g["l_max"]), int(g["nlev"])
    o = oracle.Plan(L, int(g["n_theta"]), int(g["n_phi"]))
    G, outs = oracle_step(o, make_inputs(L, nlev), *radial(nlev), list(g["lcut"]))
    for c in range(18):
        assert rel_l2(G[:, c], g["grids"][:, c]) < 1e-14, c
    for k in range(10):
        assert rel_l2(outs[k], g["outs"][k]) < 1e-14, k


# ---------------------------------------------------------- independent
@pytest.mark.parametrize("fname", ["td_L16.npz", "td_L64.npz"])
def test_oracle_get_td_matches_reference(fname):
    """lm-space assembly (src/get_td.f90) against the reference's own
    nonlinear_lm_t run by ref_harness td, one level with l_R < l_max"""
    g = load_golden(fname)
    L = int(g["l_max"])
    nt, nph = workload.grid_size(L)
    o = oracle.Plan(L, nt, nph)
    for lev in range(int(g["nlev"])):
        outs = o.get_td(int(g["lcut"][lev]), g["or2"][lev], g["or4"][lev], float(g["corfac"]),
                        float(g["epsc"]), [g["inputs"][k, lev] for k in range(10)])
        for k in range(7):
            assert rel_l2(outs[k], g["outs"][k, lev]) < 1e-14, (fname, lev, k)


def test_get_td_coriolis_identities():
    """closed forms of the Coriolis terms: a solid-body rotation z_10 gives
    no Coriolis force on itself except through dTheta2A of l = 0; with all
    nonlinear terms zero and w = dw = 0, dpdt and dwdt at (l, m) come from
    z at l -/+ 1 only, with dTheta factors (l-1) c_lm and (l+2) c_{l+1,m}"""
    L = 16
    nt, nph = workload.grid_size(L)
    o = oracle.Plan(L, nt, nph)
    ls, ms = workload.lm_index(L)
    zero = np.zeros(ls.size, dtype=np.complex128)
    z = zero.copy()
    i10 = np.flatnonzero((ls == 1) & (ms == 0))[0]
    z[i10] = 1.0
    or2 = 0.81
    outs = o.get_td(L, or2, 1.0, 0.5, 0.0, [zero] * 7 + [zero, zero, z])
    dwdt, dzdt, dpdt = outs[0], outs[1], outs[2]
    c = lambda l, m: np.sqrt((l + m) * (l - m) / ((2 * l - 1.0) * (2 * l + 1.0)))  # noqa: E731
    # dwdt(l=0) = 2 CorFac or1 dTheta2A(0) z_10, dTheta2A(0) = 2 c_10
    assert abs(dwdt[0] - 2 * 0.5 * np.sqrt(or2) * 2 * c(1, 0)) < 1e-15
    # dwdt(l=2, m=0) = -2 CorFac or1 dTheta2S z_10 with dTheta2S = (2-1) c_20
    i20 = np.flatnonzero((ls == 2) & (ms == 0))[0]
    assert abs(dwdt[i20] + 2 * 0.5 * np.sqrt(or2) * 1 * c(2, 0)) < 1e-15
    # dpdt(l=2, m=0) = 2 CorFac or2 dTheta3S z_10, dTheta3S = (2-1)(2+1) c_20
    assert abs(dpdt[i20] - 2 * 0.5 * or2 * 3 * c(2, 0)) < 1e-15
    # z_10 itself: the toroidal Coriolis term is i m z = 0 at m = 0
    assert abs(dzdt[i10]) < 1e-15
    # every (l, m) other than l = 0, 2 at m = 0 is untouched
    others = np.ones(ls.size, bool)
    others[[0, i20]] = False
    assert np.abs(dwdt[others]).max() == 0.0


def test_oracle_courant_matches_reference():
    """Courant maxima (src/courant.f90, magnetic branch, Alfven damping) on
    the oracle's step grids against the reference's courant run by
    ref_harness courant on the same inputs: dtrkc, dthkc to 1e-14"""
    from nl_reference import make_inputs, oracle_step, radial
    g = load_golden("courant_L16.npz")
    L, nlev = int(g["l_max"]), int(g["nlev"])
    nt, nph = workload.grid_size(L)
    o = oracle.Plan(L, nt, nph)
    rad = radial(nlev)
    G, _ = oracle_step(o, make_inputs(L, nlev), *rad, list(g["lcut"]))
    opm = float(g["opm"])
    for lev in range(nlev):
        dr2, dh2 = g["delxr2"][lev], g["delxh2"][lev]
        vri2, vhi2 = (0.5 * (1 + opm)) ** 2 / dr2, (0.5 * (1 + opm)) ** 2 / dh2
        a, b = o.courant(rad[0][lev], rad[1][lev], rad[2][lev], rad[3], float(g["courfac"]),
                         float(g["alffac"]), vri2, vhi2, [G[lev, c] for c in (0, 1, 2, 9, 10, 11)])
        assert abs(np.sqrt(dr2 / a) / g["dtrkc"][lev] - 1) < 1e-14
        assert abs(np.sqrt(dh2 / b) / g["dthkc"][lev] - 1) < 1e-14


@pytest.mark.parametrize("fname", ["brv_L16_m1.npz", "brv_L64_m4.npz"])
def test_oracle_get_br_v_bcs_matches_reference(fname):
    """Boundary products br*vt, br*(vp - omega sin^2) analysed to
    (0, br_vt_lm, br_vp_lm) (src/nonlinear_bcs.f90 get_br_v_bcs) against the
    reference's own routine run by ref_harness brv: 1e-12 relative"""
    g = load_golden(fname)
    p = oracle.Plan(int(g["l_max"]), int(g["n_theta"]), int(g["n_phi"]), minc=int(g["minc"]))
    for lev in range(g["grids"].shape[0]):
        a, b = p.get_br_v_bcs(*g["grids"][lev], g["omega"][lev], g["or2"][lev], g["orho"][lev])
        assert rel_l2(a, g["br_vt_lm"][lev]) < 1e-12, (fname, lev)
        assert rel_l2(b, g["br_vp_lm"][lev]) < 1e-12, (fname, lev)


@pytest.mark.parametrize("fname", ["brv_L16_m1.npz", "brv_L64_m4.npz"])
def test_oracle_get_b_nl_bcs_matches_reference(fname):
    """orc_get_b_nl_bcs against the reference's get_b_nl_bcs('CMB') and
    ('ICB') on its own get_br_v_bcs products (ref_harness brv): bit for bit
    (one product per entry, fac = conductance_ma prmag or sigma_ratio prmag);
    another bc is refused like the reference's abortRun."""
    g = load_golden(fname)
    p = oracle.Plan(int(g["l_max"]), int(g["n_theta"]), int(g["n_phi"]), minc=int(g["minc"]))
    for lev in range(g["br_vt_lm"].shape[0]):
        a, b = g["br_vt_lm"][lev], g["br_vp_lm"][lev]
        bc, jc = p.get_b_nl_bcs("CMB", a, b, g["conductance_ma"][lev], g["prmag"][lev])
        _, ji = p.get_b_nl_bcs("ICB", a, b, g["sigma_ratio"][lev], g["prmag"][lev])
        assert np.array_equal(bc, g["b_nl_cmb"][lev]) and np.array_equal(jc, g["aj_nl_cmb"][lev])
        assert np.array_equal(ji, g["aj_nl_icb"][lev])
        assert bc[0] == 0 and jc[0] == 0 and ji[0] == 0
    z = np.zeros(p.lm_max, dtype=np.complex128)
    assert p._L.orc_get_b_nl_bcs(int(p.lm_max), 2, oracle._ptr(z), oracle._ptr(z), 1.0, 1.0,
                                 oracle._ptr(z), oracle._ptr(z)) == -1


def test_br_v_bcs_solid_body_rotation():
    """A uniform br = 1 over a fluid at rest on a boundary rotating at
    omega: br_vp = -omega sin^2(theta) is solid-body rotation, so the
    analysis holds only the toroidal l=1, m=0 coefficient and resynthesises
    the grid; br_vt_lm vanishes."""
    L = 16
    nt, nph = workload.grid_size(L)
    p = oracle.Plan(L, nt, nph)
    one, zero = np.ones((nph, nt)), np.zeros((nph, nt))
    a, b = p.get_br_v_bcs(one, zero, zero, 0.9, 2.0, 1.0)
    assert np.abs(a).max() < 1e-15
    assert abs(b[1]) > 0.1 and np.abs(np.delete(b, 1)).max() < 1e-14
    t = p.tables()
    s2 = np.empty(nt)
    s2[0::2] = s2[1::2] = np.sin(t["theta"]) ** 2
    vt, vp = p.sphtor_to_spat(a, b, L)
    np.testing.assert_allclose(vt, 0.0, atol=1e-14)
    np.testing.assert_allclose(vp, np.broadcast_to(-0.9 * s2, vp.shape), atol=1e-14)


def test_oracle_radial_loop_matches_reference():
    """The oracle's nonlinear step, Courant maxima and get_td chained over
    the interior levels of the boussBenchSat restart state against the
    reference's own chain (ref_harness rloop): 1e-12"""
    import restart_state
    from nl_reference import oracle_step
    g = load_golden("rloop_boussBenchSat.npz")
    h, ins, P = restart_state.load(g["levels"])
    L = h["l_max"]
    nt, nph = workload.grid_size(L, h["nalias"], h["minc"])
    o = oracle.Plan(L, nt, nph, minc=h["minc"])
    G, outs = oracle_step(o, ins, P["or2"], P["or4"], P["orho1"], P["LFfac"], P["lcut"])
    for i in range(ins[0].shape[0]):
        arr = [outs[k][i] for k in (0, 1, 2, 4, 6, 7, 8)] + [ins[0][i], ins[1][i], ins[3][i]]
        td = o.get_td(L, P["or2"][i], P["or4"][i], P["CorFac"], P["epsc"][i], arr)
        for k in range(7):
            assert rel_l2(td[k], g["td"][k][i]) < 1e-12, (i, k)
        vri2 = (0.5 * (1 + P["opm"])) ** 2 / P["delxr2"][i]
        vhi2 = (0.5 * (1 + P["opm"])) ** 2 / P["delxh2"][i]
        a, b = o.courant(P["or2"][i], P["or4"][i], P["orho1"][i], P["LFfac"], P["courfac"],
                         P["alffac"], vri2, vhi2, [G[i, c] for c in (0, 1, 2, 9, 10, 11)])
        assert abs(np.sqrt(P["delxr2"][i] / a) / g["dtrkc"][i] - 1) < 1e-13
        assert abs(np.sqrt(P["delxh2"][i] / b) / g["dthkc"][i] - 1) < 1e-13


def test_gauleg_vs_numpy():
    for n in (24, 96, 384):
        th, w = oracle.gauleg(n)
        x, wx = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(np.cos(th), x[::-1], atol=1e-14)
        np.testing.assert_allclose(w, wx[::-1], rtol=0, atol=5e-14)
        assert abs(w.sum() - 2.0) < 1e-13


def test_plm_closed_forms():
    th = 0.7
    p, d = oracle.plm_theta(th, 3)
    c, s = np.cos(th), np.sin(th)
    # m-major: (0,0) (1,0) (2,0) (3,0) (1,1) (2,1) (3,1) (2,2) (3,2) (3,3)
    y00 = 1 / np.sqrt(4 * np.pi)
    y10 = np.sqrt(3 / (4 * np.pi)) * c
    y20 = np.sqrt(5 / (16 * np.pi)) * (3 * c * c - 1)
    y11 = np.sqrt(3 / (8 * np.pi)) * s          # no Condon-Shortley phase
    y22 = np.sqrt(15 / (32 * np.pi)) * s * s
    np.testing.assert_allclose(p[[0, 1, 2, 4, 7]], [y00, y10, y20, y11, y22], rtol=1e-14)
    # sin(theta) d/dtheta of the same
    np.testing.assert_allclose(d[[0, 1, 2, 4, 7]],
                               [0.0, -np.sqrt(3 / (4 * np.pi)) * s * s,
                                np.sqrt(5 / (16 * np.pi)) * (-6 * c * s * s),
                                np.sqrt(3 / (8 * np.pi)) * c * s,
                                np.sqrt(15 / (32 * np.pi)) * 2 * s * s * c], atol=1e-14)


def test_dplm_is_sin_theta_derivative():
    l_max, h = 40, 1e-5
    th = 1.1
    p_plus, _ = oracle.plm_theta(th + h, l_max)
    p_minus, _ = oracle.plm_theta(th - h, l_max)
    _, d = oracle.plm_theta(th, l_max)
    fd = np.sin(th) * (p_plus - p_minus) / (2 * h)
    assert np.abs(fd - d).max() < 1e-6 * np.abs(d).max()


def test_plm_orthonormal_under_gauss():
    l_max, nt = 32, 48
    p = oracle.Plan(l_max, nt, 96)
    t = p.tables()
    P = np.concatenate([t["plm"], t["plm"] * ((-1.0) ** (p.lm2l - p.lm2m))[None, :]])
    w = np.concatenate([t["gauss"], t["gauss"]])
    for m in (0, 5, 17, 32):
        sel = p.lm2m == m
        G = 2 * np.pi * (P[:, sel] * w[:, None]).T @ P[:, sel]
        np.testing.assert_allclose(G, np.eye(sel.sum()), atol=1e-13)


def test_fft_vs_numpy():
    p = oracle.Plan(16, 24, 48)
    rng = np.random.default_rng(3)
    g = rng.standard_normal((48, 24))
    F = p.fft_forward(g)
    ref = np.fft.rfft(g, axis=0) / 48
    ref[0] = ref[0].real
    ref[-1] = ref[-1].real
    np.testing.assert_allclose(F, ref, atol=1e-15)
    back = p.fft_inverse(F)
    np.testing.assert_allclose(back, g, atol=1e-13)
    # inverse is the unnormalised Hermitian synthesis; imaginary parts of
    # modes 0 and n/2 are ignored
    F2 = F.copy()
    F2[0] += 0.3j
    F2[-1] -= 0.1j
    np.testing.assert_allclose(p.fft_inverse(F2), np.fft.irfft(F, n=48, axis=0) * 48, atol=1e-13)


def _direct_synthesis(p, S):
    """f(theta_j, phi_k) = sum_lm S_lm P_lm(theta_j) e^{i m phi_k} + c.c., brute force."""
    t = p.tables()
    th = np.empty(p.n_theta)
    th[0::2] = t["theta"]
    th[1::2] = np.pi - t["theta"]
    phi = 2 * np.pi * np.arange(p.n_phi) / (p.n_phi * p.minc)
    sign = (-1.0) ** (p.lm2l - p.lm2m)
    Pfull = np.empty((p.n_theta, p.lm_max))
    Pfull[0::2] = t["plm"]
    Pfull[1::2] = t["plm"] * sign
    f = np.zeros((p.n_phi, p.n_theta))
    for lm in range(p.lm_max):
        m = p.lm2m[lm]
        e = np.exp(1j * m * p.minc * 0 + 1j * m * phi)
        term = np.outer(e, Pfull[:, lm]) * S[lm]
        f += term.real if m == 0 else 2 * term.real
    return f


def test_scalar_synthesis_brute_force():
    p = oracle.Plan(12, 24, 48)
    S = workload.coefficients(12, 3, 0, 0, potential=False)
    got = p.scal_to_spat(S, 12)
    ref = _direct_synthesis(p, S)
    assert rel_l2(got, ref) < 1e-13


def test_round_trips():
    l_max = 32
    nt, nphi = workload.grid_size(l_max)
    p = oracle.Plan(l_max, nt, nphi)
    S = workload.coefficients(l_max, 4, 0, 0, potential=False)
    assert rel_l2(p.scal_to_SH(p.scal_to_spat(S, l_max), l_max), S) < 1e-13
    Sp = workload.coefficients(l_max, 4, 1, 0)
    T = workload.coefficients(l_max, 4, 2, 0)
    vt, vp = p.sphtor_to_spat(Sp, T, l_max)
    s2, t2 = p.spat_to_sphertor(vt, vp, l_max)
    assert rel_l2(s2, Sp) < 1e-13 and rel_l2(t2, T) < 1e-13


def test_qst_components_are_robert_form():
    """v_theta*sin(theta) of a pure l=1,m=0 spheroidal field: S_10 dP_10."""
    p = oracle.Plan(8, 12, 24)
    S = p.coef()
    S[1] = 1.0
    Z = p.coef()
    vt, vp = p.sphtor_to_spat(S, Z, 8)
    t = p.tables()
    th = np.empty(p.n_theta)
    th[0::2] = t["theta"]
    th[1::2] = np.pi - t["theta"]
    expect = -np.sqrt(3 / (4 * np.pi)) * np.sin(th) ** 2
    np.testing.assert_allclose(vt, np.broadcast_to(expect, vt.shape), atol=1e-14)
    np.testing.assert_allclose(vp, 0.0, atol=1e-14)
