"""The radial loop of any physics switches composed from the FP64 oracle, per
interior level, in the reference's call order (src/rIter.f90:188-444 with
nBc = 0: transform_to_grid_space :466-622, get_nl, transform_to_lm_space
:624-712, courant, nonlinear_lm_t): the check for mhx_rloop.  Test
infrastructure."""
import numpy as np

from oracle import oracle as O


def _flags(ph, lcr):
    f = [0] * 16
    f[O.F_CONV_NL] = ph["l_conv_nl"]
    f[O.F_ADV_CURL] = ph["l_adv_curl"]
    f[O.F_HEAT_NL] = ph["l_heat"] and ph["l_heat_nl"]
    f[O.F_CHEM] = ph["l_chemical_conv"]
    f[O.F_PHASE] = ph["l_phase_field"]
    f[O.F_PREC] = ph["l_precession"]
    f[O.F_CENTR] = ph["l_centrifuge"]
    f[O.F_MAG_LF] = ph["l_mag_LF"]
    f[O.F_MAG_NL] = ph["l_mag_nl"]
    f[O.F_ANEL] = ph["l_anel"] and ph["l_heat"]
    f[O.F_LCR] = lcr
    f[O.F_CORR] = ph["l_corr"]
    f[O.F_SINGLE] = ph["l_single_matrix"]
    f[O.F_DCURL] = ph["l_double_curl"]
    f[O.F_ANEL_LIQ] = ph["l_anelastic_liquid"]
    f[O.F_MAG_DBDT] = ph["l_mag_nl"] or ph["l_mag_kin"]
    return f


def _par(ph, minc):
    p = [0.0] * 18
    p[O.P_LFFAC] = ph["LFfac"]
    p[O.P_VISCHEAT] = ph["ViscHeatFac"]
    p[O.P_OHMLOSS] = ph["OhmLossFac"]
    p[O.P_EPSPHASE] = ph["epsPhase"]
    p[O.P_PENALTY] = ph["penaltyFac"]
    p[O.P_PHASEDIFF] = ph["phaseDiffFac"]
    p[O.P_TMELT] = ph["tmelt"]
    p[O.P_PO] = ph["po"]
    p[O.P_PREC_ANGLE] = ph["prec_angle"]
    p[O.P_OEK] = ph["oek"]
    p[O.P_TIME] = ph["timeStage"]
    p[O.P_DILUTION] = ph["dilution_fac"]
    p[O.P_RA] = ph["ra"]
    p[O.P_OPR] = ph["opr"]
    p[O.P_MINC] = float(minc)
    p[O.P_CORFAC] = ph["CorFac"]
    p[O.P_EPSC] = ph["epsc"]
    p[O.P_EPSCXI] = ph["epscXi"]
    return p


def oracle_rloop(o, ph, nlev, rad, fld, minc):
    """ph: mhx_rloop_phys values (dict); rad: per-level arrays (RLOOP_RAD +
    nR, l_R); fld: (nlev, lm_max) inputs by name.  Returns (outputs by
    name, dtrkc, dthkc)."""
    conv = ph["l_conv"] or ph["l_mag_kin"]
    heat, chem, phase = ph["l_heat"], ph["l_chemical_conv"], ph["l_phase_field"]
    mag = ph["l_mag"] or ph["l_mag_LF"]
    curl = ph["l_adv_curl"]
    anel = ph["l_anel"] and heat
    need_grad = (not curl) or anel
    dcurl = ph["l_double_curl"]
    names = ["dwdt", "dzdt"]
    if not dcurl or ph["lPressNext"]:
        names.append("dpdt")
    if heat:
        names += ["dsdt", "dVSrLM"]
    if chem:
        names += ["dxidt", "dVXirLM"]
    if phase:
        names.append("dphidt")
    if ph["l_mag"]:
        names += ["dbdt", "djdt", "dVxBhLM"]
    if dcurl:
        names.append("dVxVhLM")
    outs = {n: np.zeros((nlev, o.lm_max), dtype=np.complex128) for n in names}
    dtr, dth = np.zeros(nlev), np.zeros(nlev)
    for i in range(nlev):
        lc = int(rad["l_R"][i])
        or2 = float(rad["or2"][i])
        f = {k: (v[i] if v is not None else None) for k, v in fld.items()}
        g = {}
        if conv:
            g[O.RI_VR], g[O.RI_VT], g[O.RI_VP] = o.torpol_to_spat(f["w"], f["dw"], f["z"], lc)
            if curl:
                g[O.RI_CVR], g[O.RI_CVT], g[O.RI_CVP] = o.torpol_to_curl_spat(
                    or2, f["w"], f["ddw"], f["z"], f["dz"], lc)
            if need_grad:
                if not curl:
                    g[O.RI_CVR] = o.pol_to_curlr_spat(f["z"], lc)
                g[O.RI_DVRDR], g[O.RI_DVTDR], g[O.RI_DVPDR] = o.torpol_to_spat(
                    f["dw"], f["ddw"], f["dz"], lc)
                g[O.RI_DVRDT], g[O.RI_DVRDP] = o.pol_to_grad_spat(f["w"], lc)
                g[O.RI_DVTDP], g[O.RI_DVPDP] = o.torpol_to_dphspat(f["dw"], f["z"], lc)
        if heat:
            g[O.RI_S] = o.scal_to_spat(f["s"], lc)
        if chem:
            g[O.RI_XI] = o.scal_to_spat(f["xi"], lc)
        if phase:
            g[O.RI_PHI] = o.scal_to_spat(f["phi"], lc)
        if mag:
            g[O.RI_BR], g[O.RI_BT], g[O.RI_BP] = o.torpol_to_spat(f["b"], f["db"], f["aj"], lc)
            g[O.RI_CBR], g[O.RI_CBT], g[O.RI_CBP] = o.torpol_to_curl_spat(
                or2, f["b"], f["ddb"], f["aj"], f["dj"], lc)
        lcr = int(rad["nR"][i]) <= int(ph["n_r_LCR"])
        radv = [0.0] * 11
        for k, n in ((O.R_R, "r"), (O.R_OR1, "or1"), (O.R_OR2, "or2"), (O.R_OR4, "or4"),
                     (O.R_ORHO1, "orho1"), (O.R_BETA, "beta"), (O.R_OTEMP1, "otemp1"),
                     (O.R_VISC, "visc"), (O.R_LAMBDA, "lambda"), (O.R_TEMP0, "temp0"),
                     (O.R_EPSCPROF, "epscProf")):
            radv[k] = float(rad[n][i])
        fl = _flags(ph, lcr)
        pa = _par(ph, minc)
        want = []
        adv = ph["l_conv_nl"] or ph["l_mag_LF"]
        if adv:
            want += [O.RO_ADVR, O.RO_ADVT, O.RO_ADVP]
        if heat:
            want += [O.RO_VSR, O.RO_VST, O.RO_VSP]
        if chem:
            want += [O.RO_VXIR, O.RO_VXIT, O.RO_VXIP]
        if ph["l_mag_nl"]:
            want += [O.RO_VXBR, O.RO_VXBT, O.RO_VXBP]
        if anel:
            want.append(O.RO_HEAT)
        if phase:
            want.append(O.RO_PHI)
        pr = o.rl_products(fl, pa, radv, g, want)
        cour_mag = ph["l_mag"] and ph["l_mag_LF"] and not ph["l_mag_kin"]
        cg = [g.get(O.RI_VR), g.get(O.RI_VT), g.get(O.RI_VP), g.get(O.RI_BR), g.get(O.RI_BT),
              g.get(O.RI_BP)]
        dtr[i], dth[i] = o.rl_courant(cour_mag, [ph["courfac"], ph["alffac"], ph["LFfac"], ph["opm"],
                                                 ph["l_cour_alf_damp"]],
                                      [or2, rad["or4"][i], rad["orho1"][i], rad["delxr2"][i],
                                       rad["delxh2"][i]], cg)
        nl = {}
        if adv:
            nl[O.NL_ADVR], nl[O.NL_ADVT], nl[O.NL_ADVP] = o.spat_to_qst(
                pr[O.RO_ADVR], pr[O.RO_ADVT], pr[O.RO_ADVP], lc)
        if heat:
            q, st, _ = o.spat_to_qst(pr[O.RO_VSR], pr[O.RO_VST], pr[O.RO_VSP], lc)
            outs["dVSrLM"][i] = q
            nl[O.NL_VST] = st
        if chem:
            q, st, _ = o.spat_to_qst(pr[O.RO_VXIR], pr[O.RO_VXIT], pr[O.RO_VXIP], lc)
            outs["dVXirLM"][i] = q
            nl[O.NL_VXIT] = st
        if ph["l_mag_nl"]:
            if lcr:
                nl[O.NL_VXBR] = np.zeros(o.lm_max, dtype=np.complex128)
                nl[O.NL_VXBT], nl[O.NL_VXBP] = o.spat_to_sphertor(pr[O.RO_VXBT], pr[O.RO_VXBP], lc)
            else:
                nl[O.NL_VXBR], nl[O.NL_VXBT], nl[O.NL_VXBP] = o.spat_to_qst(
                    pr[O.RO_VXBR], pr[O.RO_VXBT], pr[O.RO_VXBP], lc)
        if anel:
            nl[O.NL_HEAT] = o.scal_to_SH(pr[O.RO_HEAT], lc)
        if phase:
            outs["dphidt"][i] = o.scal_to_SH(pr[O.RO_PHI], lc)
        tw = {O.TO_DWDT: "dwdt", O.TO_DZDT: "dzdt", O.TO_DPDT: "dpdt", O.TO_DSDT: "dsdt",
              O.TO_DXIDT: "dxidt", O.TO_DBDT: "dbdt", O.TO_DJDT: "djdt", O.TO_DVXBH: "dVxBhLM",
              O.TO_DVXVH: "dVxVhLM"}
        want_td = [k for k, n in tw.items() if n in outs]
        if O.TO_DBDT in want_td:
            want_td += [O.TO_DJDT, O.TO_DVXBH]
        td = o.rl_td(fl, pa, radv, lc, nl, [f["w"], f["dw"], f["ddw"], f["z"], f["dz"]],
                     sorted(set(want_td)))
        for k, arr in td.items():
            outs[tw[k]][i] = arr
    return outs, dtr, dth
