"""ORACLE - test infrastructure only.

ctypes front end of ``liboracle.so`` (the FP64 C restatement of MagIC's
native transform path, ``mhx_oracle.c``) and of the reference harness
``_ref/ref_harness`` (the reference's own Fortran modules).  Only the tests,
``__graft_entry__.smoke`` and the CPU-baseline leg of ``bench.py`` may import
this module, and only to check or time; the product never does.

Array conventions follow the reference: coefficient vectors are complex128 of
length lm_max in m-major order; grids are float64 arrays of shape
(n_phi, n_theta) in C order, i.e. theta fastest like Fortran ``f(nt, np)``.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        D = ctypes.c_double
        I = ctypes.c_int
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_plan_create.restype = P
        L.orc_plan_create.argtypes = [I, I, I, I, I]
        L.orc_plan_destroy.argtypes = [P]
        L.orc_plan_lm_max.argtypes = [P]
        L.orc_plan_n_m_max.argtypes = [P]
        L.orc_gauleg.argtypes = [I, dp, dp]
        L.orc_plm_theta.argtypes = [D, I, I, I, I, dp, dp]
        L.orc_plan_tables.argtypes = [P] + [dp] * 7
        for name in ("orc_fft_forward", "orc_fft_inverse"):
            getattr(L, name).argtypes = [P, dp, dp]
        L.orc_get_nl.argtypes = [P, D, D, D, D, ctypes.POINTER(dp), ctypes.POINTER(dp)]
        L.orc_get_td.argtypes = [P, I, D, D, D, D, ctypes.POINTER(dp), ctypes.POINTER(dp)]
        L.orc_courant.argtypes = [P, D, D, D, D, D, D, D, D, ctypes.POINTER(dp), dp, dp]
        L.orc_get_br_v_bcs.argtypes = [P, dp, dp, dp, D, D, D, dp, dp]
        L.orc_get_b_nl_bcs.argtypes = [I, I, dp, dp, D, D, dp, dp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_rl_products.argtypes = [P, ip, dp, dp, ctypes.POINTER(dp), ctypes.POINTER(dp)]
        L.orc_rl_courant.argtypes = [P, I, dp, dp, ctypes.POINTER(dp), dp, dp]
        L.orc_rl_td.argtypes = [P, ip, dp, dp, I, ctypes.POINTER(dp), ctypes.POINTER(dp),
                                ctypes.POINTER(dp)]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def gauleg(n):
    th = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauleg(n, _ptr(th), _ptr(w))
    return th, w


def plm_theta(theta, l_max, m_max=None, minc=1):
    m_max = l_max if m_max is None else m_max
    lm_max = sum(l_max - m + 1 for m in range(0, m_max + 1, minc))
    p = np.zeros(lm_max)
    d = np.zeros(lm_max)
    lib().orc_plm_theta(float(theta), l_max, 0, m_max, minc, _ptr(p), _ptr(d))
    return p, d


def lm_index(l_max, m_max=None, minc=1):
    """(lm2l, lm2m) of the standard m-major order (src/blocking.f90:309-317)."""
    m_max = l_max if m_max is None else min(m_max, l_max)
    ls, ms = [], []
    for m in range(0, m_max + 1, minc):
        for l in range(m, l_max + 1):
            ls.append(l)
            ms.append(m)
    return np.array(ls), np.array(ms)


class Plan:
    """FP64 oracle plan for one truncation and grid."""

    def __init__(self, l_max, n_theta, n_phi, m_max=None, minc=1):
        m_max = l_max if m_max is None else min(m_max, l_max)
        m_max = (m_max // minc) * minc
        self.l_max, self.m_max, self.minc = l_max, m_max, minc
        self.n_theta, self.n_phi = n_theta, n_phi
        self._L = lib()
        self.h = self._L.orc_plan_create(l_max, m_max, minc, n_theta, n_phi)
        self.lm_max = self._L.orc_plan_lm_max(self.h)
        self.n_m_max = self._L.orc_plan_n_m_max(self.h)
        self.lm2l, self.lm2m = lm_index(l_max, m_max, minc)

    def __del__(self):
        try:
            self._L.orc_plan_destroy(self.h)
        except Exception:
            pass

    # -- helpers
    def grid(self):
        return np.zeros((self.n_phi, self.n_theta))

    def coef(self):
        return np.zeros(self.lm_max, dtype=np.complex128)

    def tables(self):
        nh = self.n_theta // 2
        th = np.zeros(nh)
        gw = np.zeros(nh)
        t = [np.zeros((nh, self.lm_max)) for _ in range(4)]
        o2 = np.zeros(self.n_theta)
        self._L.orc_plan_tables(self.h, _ptr(th), _ptr(gw), *[_ptr(x) for x in t], _ptr(o2))
        return dict(theta=th, gauss=gw, plm=t[0], dplm=t[1], wplm=t[2], wdplm=t[3], osin2=o2)

    def _c(self, a, n=None):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.shape != ((self.lm_max,) if n is None else (n,)):
            raise ValueError("coefficient array has wrong shape %r" % (a.shape,))
        return a

    def _g(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.n_phi, self.n_theta):
            raise ValueError("grid has wrong shape %r" % (a.shape,))
        return a

    def _call(self, name, *args):
        conv = []
        for a in args:
            if isinstance(a, np.ndarray):
                conv.append(_ptr(a))
            elif isinstance(a, float):
                conv.append(ctypes.c_double(a))
            else:
                conv.append(ctypes.c_int(int(a)))
        getattr(self._L, name)(ctypes.c_void_p(self.h), *conv)

    # -- FFT
    def fft_forward(self, g):
        F = np.zeros((self.n_phi // 2 + 1, self.n_theta), dtype=np.complex128)
        self._call("orc_fft_forward", self._g(g), F)
        return F

    def fft_inverse(self, F):
        F = np.ascontiguousarray(F, dtype=np.complex128)
        g = self.grid()
        self._call("orc_fft_inverse", F, g)
        return g

    # -- module sht API (src/sht_native.f90)
    def scal_to_spat(self, S, lcut):
        f = self.grid()
        self._call("orc_scal_to_spat", self._c(S), f, lcut)
        return f

    def scal_to_grad_spat(self, S, lcut):
        a, b = self.grid(), self.grid()
        self._call("orc_scal_to_grad_spat", self._c(S), a, b, lcut)
        return a, b

    def pol_to_grad_spat(self, S, lcut):
        a, b = self.grid(), self.grid()
        self._call("orc_pol_to_grad_spat", self._c(S), a, b, lcut)
        return a, b

    def torpol_to_spat(self, W, dW, Z, lcut):
        a, b, c = self.grid(), self.grid(), self.grid()
        self._call("orc_torpol_to_spat", self._c(W), self._c(dW), self._c(Z), a, b, c, lcut)
        return a, b, c

    def pol_to_curlr_spat(self, Q, lcut):
        a = self.grid()
        self._call("orc_pol_to_curlr_spat", self._c(Q), a, lcut)
        return a

    def torpol_to_curl_spat(self, or2, B, ddB, J, dJ, lcut):
        a, b, c = self.grid(), self.grid(), self.grid()
        self._call("orc_torpol_to_curl_spat", float(or2), self._c(B), self._c(ddB),
                   self._c(J), self._c(dJ), a, b, c, lcut)
        return a, b, c

    def torpol_to_dphspat(self, dW, Z, lcut):
        a, b = self.grid(), self.grid()
        self._call("orc_torpol_to_dphspat", self._c(dW), self._c(Z), a, b, lcut)
        return a, b

    def sphtor_to_spat(self, S, T, lcut):
        a, b = self.grid(), self.grid()
        self._call("orc_sphtor_to_spat", self._c(S), self._c(T), a, b, lcut)
        return a, b

    def qst_to_spat(self, Q, S, T, lcut):
        a, b, c = self.grid(), self.grid(), self.grid()
        self._call("orc_qst_to_spat", self._c(Q), self._c(S), self._c(T), a, b, c, lcut)
        return a, b, c

    def torpol_to_spat_IC(self, r, r_icb, W, dW, Z):
        a, b, c = self.grid(), self.grid(), self.grid()
        self._call("orc_torpol_to_spat_IC", float(r), float(r_icb), self._c(W), self._c(dW),
                   self._c(Z), a, b, c)
        return a, b, c

    def torpol_to_curl_spat_IC(self, r, r_icb, dB, ddB, J, dJ):
        a, b, c = self.grid(), self.grid(), self.grid()
        self._call("orc_torpol_to_curl_spat_IC", float(r), float(r_icb), self._c(dB),
                   self._c(ddB), self._c(J), self._c(dJ), a, b, c)
        return a, b, c

    def axi_to_spat(self, Sl):
        f = np.zeros(self.n_theta)
        self._call("orc_axi_to_spat", self._c(Sl, self.l_max + 1), f)
        return f

    def toraxi_to_spat(self, Tl, lcut):
        a, b = np.zeros(self.n_theta), np.zeros(self.n_theta)
        self._call("orc_toraxi_to_spat", self._c(Tl, self.l_max + 1), a, b, lcut)
        return a, b

    def scal_to_SH(self, f, lcut):
        out = self.coef()
        self._call("orc_scal_to_SH", self._g(f), out, lcut)
        return out

    def spat_to_sphertor(self, f, g, lcut):
        a, b = self.coef(), self.coef()
        self._call("orc_spat_to_sphertor", self._g(f), self._g(g), a, b, lcut)
        return a, b

    def spat_to_qst(self, f, g, h, lcut):
        a, b, c = self.coef(), self.coef(), self.coef()
        self._call("orc_spat_to_qst", self._g(f), self._g(g), self._g(h), a, b, c, lcut)
        return a, b, c

    def get_nl(self, or2, or4, orho1, LFfac, grids):
        """grids: 13 arrays (vr vt vp cvr cvt cvp br bt bp cbr cbt cbp s).
        Returns 12 arrays Adv(3) LF(3) VxB(3) VS(3)."""
        ins = [self._g(g) for g in grids]
        outs = [self.grid() for _ in range(12)]
        arr_in = (ctypes.POINTER(ctypes.c_double) * 13)(*[_ptr(a) for a in ins])
        arr_out = (ctypes.POINTER(ctypes.c_double) * 12)(*[_ptr(a) for a in outs])
        self._L.orc_get_nl(ctypes.c_void_p(self.h), float(or2), float(or4), float(orho1),
                           float(LFfac), arr_in, arr_out)
        return outs


    def get_td(self, lR, or2, or4, CorFac, epsc, arrays):
        """arrays: 10 complex lm arrays (AdvrLM AdvtLM AdvpLM VStLM VxBrLM
        VxBtLM VxBpLM w dw z).  Returns dwdt dzdt dpdt dsdt dbdt djdt
        dVxBhLM (src/get_td.f90, interior level)."""
        ins = [self._c(a) for a in arrays]
        outs = [self.coef() for _ in range(7)]
        arr_in = (ctypes.POINTER(ctypes.c_double) * 10)(*[_ptr(a) for a in ins])
        arr_out = (ctypes.POINTER(ctypes.c_double) * 7)(*[_ptr(a) for a in outs])
        self._L.orc_get_td(ctypes.c_void_p(self.h), int(lR), float(or2), float(or4),
                           float(CorFac), float(epsc), arr_in, arr_out)
        return outs


    def courant(self, or2, or4, orho1, LFfac, courfac, alffac, valri2, valhi2, grids):
        """grids: vr vt vp br bt bp (n_phi, n_theta).  Returns (vr2max, vh2max)
        of src/courant.f90's magnetic branch."""
        ins = [self._g(g) for g in grids]
        arr = (ctypes.POINTER(ctypes.c_double) * 6)(*[_ptr(a) for a in ins])
        a, b = np.zeros(1), np.zeros(1)
        self._L.orc_courant(ctypes.c_void_p(self.h), float(or2), float(or4), float(orho1),
                            float(LFfac), float(courfac), float(alffac), float(valri2),
                            float(valhi2), arr, _ptr(a), _ptr(b))
        return float(a[0]), float(b[0])

    def get_br_v_bcs(self, br, vt, vp, omega, O_r_E_2, O_rho):
        """br_vt_lm, br_vp_lm of src/nonlinear_bcs.f90 get_br_v_bcs."""
        a, b = self.coef(), self.coef()
        self._L.orc_get_br_v_bcs(ctypes.c_void_p(self.h), _ptr(self._g(br)), _ptr(self._g(vt)),
                                 _ptr(self._g(vp)), float(omega), float(O_r_E_2), float(O_rho),
                                 _ptr(a), _ptr(b))
        return a, b

    def get_b_nl_bcs(self, bc, br_vt_lm, br_vp_lm, cond_or_sigma, prmag):
        """(b_nl_bc, aj_nl_bc) of src/nonlinear_bcs.f90 get_b_nl_bcs, bc
        'CMB' or 'ICB' (b_nl_bc None at the ICB)."""
        code = {"CMB": 0, "ICB": 1}[bc]
        a = np.ascontiguousarray(br_vt_lm, dtype=np.complex128)
        b = np.ascontiguousarray(br_vp_lm, dtype=np.complex128)
        bo, jo = self.coef(), self.coef()
        rc = self._L.orc_get_b_nl_bcs(int(self.lm_max), code, _ptr(a), _ptr(b), float(cond_or_sigma),
                                      float(prmag), _ptr(bo), _ptr(jo))
        assert rc == 0
        return (bo if code == 0 else None), jo


    # ---- the radial loop of any physics switches, one interior level
    def rl_products(self, flags, par, rad, grids, want):
        """src/get_nl.f90:240-436 merged as src/rIter.f90:644-680.  ``grids``:
        dict ORC_RI index -> grid; ``want``: ORC_RO indices to compute.
        Returns dict ORC_RO index -> grid."""
        P = ctypes.POINTER(ctypes.c_double)
        ins = (P * RL_NI)(*[(_ptr(self._g(grids[k])) if k in grids else None) for k in range(RL_NI)])
        outs = {k: self.grid() for k in want}
        arr_out = (P * RL_NO)(*[(_ptr(outs[k]) if k in outs else None) for k in range(RL_NO)])
        fl = (ctypes.c_int * len(flags))(*[int(x) for x in flags])
        pa = (ctypes.c_double * len(par))(*[float(x) for x in par])
        ra = (ctypes.c_double * len(rad))(*[float(x) for x in rad])
        self._L.orc_rl_products(ctypes.c_void_p(self.h), fl, pa, ra, ins, arr_out)
        return outs

    def rl_courant(self, cour_mag, par, rad, grids):
        """(dtrkc, dthkc) of src/courant.f90 on vr vt vp br bt bp"""
        P = ctypes.POINTER(ctypes.c_double)
        z = self.grid()
        arr = (P * 6)(*[_ptr(self._g(g)) if g is not None else _ptr(z) for g in grids])
        pa = (ctypes.c_double * 5)(*[float(x) for x in par])
        ra = (ctypes.c_double * 5)(*[float(x) for x in rad])
        a, b = np.zeros(1), np.zeros(1)
        self._L.orc_rl_courant(ctypes.c_void_p(self.h), int(cour_mag), pa, ra, arr, _ptr(a), _ptr(b))
        return float(a[0]), float(b[0])

    def rl_td(self, flags, par, rad, lR, nl, fld, want):
        """src/get_td.f90 nonlinear_lm_t at one interior level: ``nl`` dict
        ORC_NL index -> coefficients, ``fld`` = w dw ddw z dz, ``want``
        ORC_TO indices.  Returns dict ORC_TO index -> coefficients."""
        P = ctypes.POINTER(ctypes.c_double)
        nla = (P * RL_NNL)(*[(_ptr(self._c(nl[k])) if k in nl else None) for k in range(RL_NNL)])
        fa = (P * 5)(*[_ptr(self._c(f)) for f in fld])
        outs = {k: self.coef() for k in want}
        oa = (P * RL_NTO)(*[(_ptr(outs[k]) if k in outs else None) for k in range(RL_NTO)])
        fl = (ctypes.c_int * len(flags))(*[int(x) for x in flags])
        pa = (ctypes.c_double * len(par))(*[float(x) for x in par])
        ra = (ctypes.c_double * len(rad))(*[float(x) for x in rad])
        self._L.orc_rl_td(ctypes.c_void_p(self.h), fl, pa, ra, int(lR), nla, fa, oa)
        return outs


# index sets of the radial-loop oracle (mhx_oracle.h)
RL_NI, RL_NO, RL_NNL, RL_NTO = 22, 14, 9, 9
(RI_VR, RI_VT, RI_VP, RI_CVR, RI_CVT, RI_CVP, RI_DVRDR, RI_DVTDR, RI_DVPDR, RI_DVRDT, RI_DVRDP,
 RI_DVTDP, RI_DVPDP, RI_S, RI_XI, RI_PHI, RI_BR, RI_BT, RI_BP, RI_CBR, RI_CBT, RI_CBP) = range(22)
(RO_ADVR, RO_ADVT, RO_ADVP, RO_VSR, RO_VST, RO_VSP, RO_VXIR, RO_VXIT, RO_VXIP, RO_VXBR, RO_VXBT,
 RO_VXBP, RO_HEAT, RO_PHI) = range(14)
(F_CONV_NL, F_ADV_CURL, F_HEAT_NL, F_CHEM, F_PHASE, F_PREC, F_CENTR, F_MAG_LF, F_MAG_NL, F_ANEL,
 F_LCR, F_CORR, F_SINGLE, F_DCURL, F_ANEL_LIQ, F_MAG_DBDT) = range(16)
(P_LFFAC, P_VISCHEAT, P_OHMLOSS, P_EPSPHASE, P_PENALTY, P_PHASEDIFF, P_TMELT, P_PO, P_PREC_ANGLE,
 P_OEK, P_TIME, P_DILUTION, P_RA, P_OPR, P_MINC, P_CORFAC, P_EPSC, P_EPSCXI) = range(18)
(R_R, R_OR1, R_OR2, R_OR4, R_ORHO1, R_BETA, R_OTEMP1, R_VISC, R_LAMBDA, R_TEMP0,
 R_EPSCPROF) = range(11)
(NL_ADVR, NL_ADVT, NL_ADVP, NL_VST, NL_VXIT, NL_VXBR, NL_VXBT, NL_VXBP, NL_HEAT) = range(9)
(TO_DWDT, TO_DZDT, TO_DPDT, TO_DSDT, TO_DXIDT, TO_DBDT, TO_DJDT, TO_DVXBH, TO_DVXVH) = range(9)


def have_ref():
    return os.path.exists(REF_HARNESS)


def run_ref(mode, l_max, minc, nalias, infile, outfile, nlev=1, nrep=1, threads=None,
            cwd=None):
    """Runs the reference harness (large automatic arrays: unlimited stack)."""
    env = dict(os.environ)
    env.setdefault("OMP_STACKSIZE", "1G")
    if threads is not None:
        env["OMP_NUM_THREADS"] = str(threads)
    cmd = "ulimit -s unlimited; %s %s %d %d %d %s %s %d %d" % (
        REF_HARNESS, mode, l_max, minc, nalias, infile, outfile, nlev, nrep)
    return subprocess.run(["bash", "-c", cmd], env=env, cwd=cwd, check=True,
                          capture_output=True, text=True)


# ---------------------------------------------------------------- radial / energies
def cheb_nodes(N):
    return np.cos(np.pi * np.arange(N) / (N - 1))


def _cheb_interp(N):
    """T[k, n] = T_n(x_k) at the Gauss-Lobatto nodes and its inverse (the
    interpolant's coefficients of nodal values)."""
    th = np.pi * np.arange(N) / (N - 1)
    T = np.cos(np.outer(th, np.arange(N)))
    return T, np.linalg.inv(T)


def cheb_dr(f, r):
    """d f / d r of nodal values f[..., k] on a linearly mapped Gauss-Lobatto
    grid r (outer boundary first), by differentiating the interpolating
    Chebyshev series: the reference's get_dr for the Chebyshev scheme
    (src/radial_derivatives.f90, via costf1 and get_dcheb) in exact
    arithmetic.  T_n'(x) = n sin(n th)/sin(th), n^2 at x = 1, (-1)^(n+1) n^2
    at x = -1."""
    N = r.size
    T, Ti = _cheb_interp(N)
    n = np.arange(N, dtype=np.float64)
    th = np.pi * np.arange(N) / (N - 1)
    Tp = np.empty((N, N))
    Tp[0] = n ** 2
    Tp[-1] = (-1.0) ** (n + 1) * n ** 2
    Tp[1:-1] = n[None, :] * np.sin(np.outer(th[1:-1], n)) / np.sin(th[1:-1])[:, None]
    drx = 2.0 / (r[0] - r[-1])
    D = drx * (Tp @ Ti)
    return np.tensordot(D, f, axes=([1], 0))


def rint_interp(f, r):
    """Integral over [r[-1], r[0]] of the polynomial interpolating f at the
    radii (src/integration.f90:56-96, rInt_R, Chebyshev scheme), from the
    interpolant's coefficients and int_{-1}^{1} T_n = 2/(1-n^2) (n even)."""
    N = r.size
    _, Ti = _cheb_interp(N)
    n = np.arange(N)
    wt = np.zeros(N)
    wt[n % 2 == 0] = 2.0 / (1.0 - n[n % 2 == 0].astype(np.float64) ** 2)
    return (wt @ Ti) @ np.asarray(f).T * (0.5 * (r[0] - r[-1]))


def e_kin_levels(w, dw, z, or2, orho1, l_max, minc=1):
    """Per-level sums of src/kinetic_energy.f90:119-160 (get_e_kin):
    rows e_p, e_t, e_p_as, e_t_as, e_p_es, e_t_es, e_p_eas, e_t_eas."""
    out = np.zeros((8, w.shape[0]))
    lm = 0
    for m in range(0, l_max + 1, minc):
        for l in range(m, l_max + 1):
            dLh = float(l * (l + 1))
            cw, cd, cz = w[:, lm], dw[:, lm], z[:, lm]
            if m == 0:
                nw, nd, nz = cw.real ** 2, cd.real ** 2, cz.real ** 2
            else:
                nw = 2.0 * (cw.real ** 2 + cw.imag ** 2)
                nd = 2.0 * (cd.real ** 2 + cd.imag ** 2)
                nz = 2.0 * (cz.real ** 2 + cz.imag ** 2)
            ep = orho1 * dLh * (dLh * or2 * nw + nd)
            et = orho1 * dLh * nz
            out[0] += ep
            out[1] += et
            if m == 0:
                out[2] += ep
                out[3] += et
                if l % 2 == 0:
                    out[6] += ep
                else:
                    out[7] += et
            if (l + m) % 2 == 0:
                out[4] += ep
            else:
                out[5] += et
            lm += 1
    return out


def e_mag_levels(b, db, aj, or2, l_max, minc=1):
    """Per-level sums of src/magnetic_energy.f90:264-340 (get_e_mag, outer
    core): rows e_p, e_t, e_p_as, e_t_as, e_p_es, e_t_es, e_p_eas, e_t_eas,
    e_dipole, e_dipole_ax (the equatorial parities opposite to the kinetic
    ones: the poloidal potential of an equatorially symmetric field has odd
    l + m)."""
    out = np.zeros((10, b.shape[0]))
    lm = 0
    for m in range(0, l_max + 1, minc):
        for l in range(m, l_max + 1):
            if l == 0:
                lm += 1
                continue
            dLh = float(l * (l + 1))
            cb, cd, cj = b[:, lm], db[:, lm], aj[:, lm]
            if m == 0:
                nb, nd, nj = cb.real ** 2, cd.real ** 2, cj.real ** 2
            else:
                nb = 2.0 * (cb.real ** 2 + cb.imag ** 2)
                nd = 2.0 * (cd.real ** 2 + cd.imag ** 2)
                nj = 2.0 * (cj.real ** 2 + cj.imag ** 2)
            ep = dLh * (dLh * or2 * nb + nd)
            et = dLh * nj
            out[0] += ep
            out[1] += et
            if m == 0:
                out[2] += ep
                out[3] += et
                if l % 2 == 1:
                    out[6] += ep
                else:
                    out[7] += et
            if (l + m) % 2 == 1:
                out[4] += ep
            else:
                out[5] += et
            if l == 1:
                out[8] += ep
                if m == 0:
                    out[9] += ep
            lm += 1
    return out
