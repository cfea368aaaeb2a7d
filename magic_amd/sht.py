"""Host-side mirror of the reference's transform API (module sht,
src/sht_native.f90:16-20) over the HIP library libmhx.so.

``Plan`` exposes one method per public routine of module sht with the same
name, argument meaning and order; arrays may be numpy arrays (host) or torch
tensors on the GPU (device).  Coefficients are complex128 vectors of length
lm_max in m-major order, grids float64 arrays of shape (n_phi, n_theta) in C
order (theta fastest, north/south interleaved), exactly as the reference.

There is no CPU fallback: if libmhx.so cannot be loaded, or no GPU is
present, every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmhx.so")

PREC_FP64 = 0
PREC_SPLIT16 = 1
# the split form with the nonlinear step's tables for l_max > 255 recomputed
# in-kernel (FP64 recurrence from FP64 seeds); 255 < l_max <= 1023
PREC_SPLIT16_REC = 2

_lib = None

# (name, argument kinds) for every symbol of include/mhx.h; kinds:
#   P plan, p pointer, i int, d double, z size_t, pp pointer-to-pointer
_SIGS = {
    "mhx_plan_create": "iiiiiiip",
    "mhx_plan_destroy": "P",
    "mhx_plan_info": "Ppppp",
    "mhx_plan_tables": "Ppppp",
    "mhx_scal_to_spat": "Pppi",
    "mhx_scal_to_grad_spat": "Ppppi",
    "mhx_pol_to_grad_spat": "Ppppi",
    "mhx_torpol_to_spat": "Pppppppi",
    "mhx_sphtor_to_spat": "Pppppi",
    "mhx_torpol_to_curl_spat_IC": "Pddppppppp",
    "mhx_torpol_to_spat_IC": "Pddpppppp",
    "mhx_torpol_to_dphspat": "Pppppi",
    "mhx_pol_to_curlr_spat": "Pppi",
    "mhx_torpol_to_curl_spat": "Pdpppppppi",
    "mhx_scal_to_SH": "Pppi",
    "mhx_spat_to_qst": "Pppppppi",
    "mhx_spat_to_sphertor": "Pppppi",
    "mhx_get_br_v_bcs": "Ppppdddpp",
    "mhx_get_b_nl_bcs": "Pippddpp",
    "mhx_axi_to_spat": "Ppp",
    "mhx_toraxi_to_spat": "Ppppi",
    "mhx_torpol_to_spat_batch": "Pipppzpppzpp",
    "mhx_spat_to_qst_batch": "Pipppzpppzpp",
    "mhx_nl_step": "Pipzppppdpp",
    "mhx_nl_step_grids": "Pipzppppdppp",
    "mhx_get_td": "Pipzppppdpp",
    "mhx_get_e_kin": "Pipppzppp",
    "mhx_get_e_mag": "Pipppzpp",
    "mhx_get_e_kin_noalloc": "Pipppzpppp",
    "mhx_nl_step_td": "Pipzpppppdpdpp",
    "mhx_rloop": "Ppippzpp",
    "mhx_nl_step_courant": "Pipzppppdpddppppp",
    "mhx_set_timing": "Pi",
    "mhx_set_dudr": "Pi",
    "mhx_stage_times": "Pp",
    "mhx_launch_times": "Ppi",
    "mhx_batch_create": "piip",
    "mhx_batch_execute": "pp",
    "mhx_batch_tuning": "pppppi",
    "mhx_batch_shapes": "ppppppppi",
    "mhx_batch_status": "p",
    "mhx_batch_destroy": "p",
}
_CT = {"P": ctypes.c_void_p, "p": ctypes.c_void_p, "i": ctypes.c_int, "d": ctypes.c_double,
       "z": ctypes.c_size_t}


class MhxError(RuntimeError):
    pass


def lib():
    """Loads libmhx.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MhxError("libmhx.so is not built (python -m magic_amd.build)")
        # torch ships its own HIP runtime (SONAME libamdhip64.so.7); loading
        # torch first makes libmhx bind to that same runtime instead of a
        # second copy, so tensors and kernels share one device context
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, kinds in _SIGS.items():
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [_CT.get(k, ctypes.c_void_p) for k in kinds]
        L.mhx_last_error.restype = ctypes.c_char_p
        L.mhx_last_error.argtypes = []
        _lib = L
    return _lib


def exported_symbols():
    return ["mhx_last_error"] + list(_SIGS)


def _check(rc):
    if rc != 0:
        raise MhxError("mhx error %d: %s" % (rc, lib().mhx_last_error().decode()))


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _ptr(a):
    if a is None:
        return None
    if _is_torch(a):
        if not a.is_contiguous():
            raise ValueError("torch tensors must be contiguous")
        return ctypes.c_void_p(a.data_ptr())
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays must be C-contiguous")
    return ctypes.c_void_p(a.ctypes.data)


class Plan:
    """A transform plan: initialize_sht for one truncation and grid
    (src/sht_native.f90:24-32).  n_theta / n_phi default to the reference's
    sizing for nalias = 20 (src/truncation.f90:66-71)."""

    def __init__(self, l_max, n_theta=None, n_phi=None, m_max=None, minc=1,
                 precision=PREC_FP64, max_levels=1):
        if n_theta is None or n_phi is None:
            from .workload import grid_size
            n_theta, n_phi = grid_size(l_max, 20, minc)
        self._L = lib()
        h = ctypes.c_void_p()
        _check(self._L.mhx_plan_create(l_max, 0 if m_max is None else m_max, minc, n_theta,
                                       n_phi, precision, max_levels, ctypes.byref(h)))
        self.h = h
        lm, nm = ctypes.c_int(), ctypes.c_int()
        _check(self._L.mhx_plan_info(h, ctypes.byref(lm), ctypes.byref(nm), None, None))
        self.l_max, self.minc = l_max, minc
        self.n_theta, self.n_phi = n_theta, n_phi
        self.lm_max, self.n_m_max = lm.value, nm.value
        self.m_max = (self.n_m_max - 1) * minc

    def close(self):
        if getattr(self, "h", None):
            self._L.mhx_plan_destroy(self.h)
            self.h = None

    finalize_sht = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- array helpers (host numpy)
    def grid(self):
        return np.zeros((self.n_phi, self.n_theta))

    def coef(self):
        return np.zeros(self.lm_max, dtype=np.complex128)

    def _c(self, a, n=None):
        if _is_torch(a):
            return a
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.shape != ((self.lm_max,) if n is None else (n,)):
            raise ValueError("coefficient array has shape %r" % (a.shape,))
        return a

    def _g(self, a):
        if _is_torch(a):
            return a
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.n_phi, self.n_theta):
            raise ValueError("grid has shape %r" % (a.shape,))
        return a

    def tables(self):
        nh = self.n_theta // 2
        th, gw = np.zeros(nh), np.zeros(nh)
        P = np.zeros((nh, self.lm_max))
        D = np.zeros((nh, self.lm_max))
        _check(self._L.mhx_plan_tables(self.h, _ptr(th), _ptr(gw), _ptr(P), _ptr(D)))
        return dict(theta=th, gauss=gw, plm=P, dplm=D)

    # -- module sht (src/sht_native.f90)
    def scal_to_spat(self, Slm, lcut, fieldc=None):
        f = self.grid() if fieldc is None else fieldc
        _check(self._L.mhx_scal_to_spat(self.h, _ptr(self._c(Slm)), _ptr(f), lcut))
        return f

    def scal_to_grad_spat(self, Slm, lcut):
        a, b = self.grid(), self.grid()
        _check(self._L.mhx_scal_to_grad_spat(self.h, _ptr(self._c(Slm)), _ptr(a), _ptr(b), lcut))
        return a, b

    def pol_to_grad_spat(self, Slm, lcut):
        a, b = self.grid(), self.grid()
        _check(self._L.mhx_pol_to_grad_spat(self.h, _ptr(self._c(Slm)), _ptr(a), _ptr(b), lcut))
        return a, b

    def torpol_to_spat(self, Wlm, dWlm, Zlm, lcut):
        a, b, c = self.grid(), self.grid(), self.grid()
        _check(self._L.mhx_torpol_to_spat(self.h, _ptr(self._c(Wlm)), _ptr(self._c(dWlm)),
                                          _ptr(self._c(Zlm)), _ptr(a), _ptr(b), _ptr(c), lcut))
        return a, b, c

    def sphtor_to_spat(self, dWlm, Zlm, lcut):
        a, b = self.grid(), self.grid()
        _check(self._L.mhx_sphtor_to_spat(self.h, _ptr(self._c(dWlm)), _ptr(self._c(Zlm)),
                                          _ptr(a), _ptr(b), lcut))
        return a, b

    def torpol_to_curl_spat_IC(self, r, r_ICB, dBlm, ddBlm, Jlm, dJlm):
        a, b, c = self.grid(), self.grid(), self.grid()
        _check(self._L.mhx_torpol_to_curl_spat_IC(
            self.h, float(r), float(r_ICB), _ptr(self._c(dBlm)), _ptr(self._c(ddBlm)),
            _ptr(self._c(Jlm)), _ptr(self._c(dJlm)), _ptr(a), _ptr(b), _ptr(c)))
        return a, b, c

    def torpol_to_spat_IC(self, r, r_ICB, Wlm, dWlm, Zlm):
        a, b, c = self.grid(), self.grid(), self.grid()
        _check(self._L.mhx_torpol_to_spat_IC(
            self.h, float(r), float(r_ICB), _ptr(self._c(Wlm)), _ptr(self._c(dWlm)),
            _ptr(self._c(Zlm)), _ptr(a), _ptr(b), _ptr(c)))
        return a, b, c

    def torpol_to_dphspat(self, dWlm, Zlm, lcut):
        a, b = self.grid(), self.grid()
        _check(self._L.mhx_torpol_to_dphspat(self.h, _ptr(self._c(dWlm)), _ptr(self._c(Zlm)),
                                             _ptr(a), _ptr(b), lcut))
        return a, b

    def pol_to_curlr_spat(self, Qlm, lcut):
        a = self.grid()
        _check(self._L.mhx_pol_to_curlr_spat(self.h, _ptr(self._c(Qlm)), _ptr(a), lcut))
        return a

    def torpol_to_curl_spat(self, or2, Blm, ddBlm, Jlm, dJlm, lcut):
        a, b, c = self.grid(), self.grid(), self.grid()
        _check(self._L.mhx_torpol_to_curl_spat(
            self.h, float(or2), _ptr(self._c(Blm)), _ptr(self._c(ddBlm)), _ptr(self._c(Jlm)),
            _ptr(self._c(dJlm)), _ptr(a), _ptr(b), _ptr(c), lcut))
        return a, b, c

    def scal_to_SH(self, f, lcut):
        out = self.coef()
        _check(self._L.mhx_scal_to_SH(self.h, _ptr(self._g(f)), _ptr(out), lcut))
        return out

    def spat_to_qst(self, f, g, h, lcut):
        a, b, c = self.coef(), self.coef(), self.coef()
        _check(self._L.mhx_spat_to_qst(self.h, _ptr(self._g(f)), _ptr(self._g(g)),
                                       _ptr(self._g(h)), _ptr(a), _ptr(b), _ptr(c), lcut))
        return a, b, c

    def spat_to_sphertor(self, f, g, lcut):
        a, b = self.coef(), self.coef()
        _check(self._L.mhx_spat_to_sphertor(self.h, _ptr(self._g(f)), _ptr(self._g(g)),
                                            _ptr(a), _ptr(b), lcut))
        return a, b

    def get_br_v_bcs(self, br, vt, vp, omega, O_r_E_2, O_rho):
        """br_vt_lm, br_vp_lm: the induction products at a conducting
        boundary analysed at l_max (src/nonlinear_bcs.f90:24-68)."""
        a, b = self.coef(), self.coef()
        _check(self._L.mhx_get_br_v_bcs(self.h, _ptr(self._g(br)), _ptr(self._g(vt)),
                                        _ptr(self._g(vp)), float(omega), float(O_r_E_2),
                                        float(O_rho), _ptr(a), _ptr(b)))
        return a, b

    def get_b_nl_bcs(self, bc, br_vt_lm, br_vp_lm, cond_or_sigma, prmag):
        """(b_nl_bc, aj_nl_bc): the nonlinear term of the magnetic boundary
        condition at bc 'CMB' (cond_or_sigma = conductance_ma) or 'ICB'
        (sigma_ratio; b_nl_bc None) (src/nonlinear_bcs.f90:71-118)."""
        code = {"CMB": BC_CMB, "ICB": BC_ICB}[bc]
        bo, jo = self.coef(), self.coef()
        a = self._c(br_vt_lm, self.lm_max) if code == BC_CMB else None
        _check(self._L.mhx_get_b_nl_bcs(self.h, code, _ptr(a) if a is not None else None,
                                        _ptr(self._c(br_vp_lm, self.lm_max)), float(cond_or_sigma),
                                        float(prmag), _ptr(bo) if code == BC_CMB else None, _ptr(jo)))
        return (bo if code == BC_CMB else None), jo

    def axi_to_spat(self, fl_ax):
        f = np.zeros(self.n_theta)
        _check(self._L.mhx_axi_to_spat(self.h, _ptr(self._c(fl_ax, self.l_max + 1)), _ptr(f)))
        return f

    def toraxi_to_spat(self, fl_ax, lcut):
        a, b = np.zeros(self.n_theta), np.zeros(self.n_theta)
        _check(self._L.mhx_toraxi_to_spat(self.h, _ptr(self._c(fl_ax, self.l_max + 1)),
                                          _ptr(a), _ptr(b), lcut))
        return a, b

    # -- batched over radial levels (torch tensors on the GPU)
    def torpol_to_spat_batch(self, W, dW, Z, lcut, vr, vt, vp):
        nlev = W.shape[0]
        lc = (ctypes.c_int * nlev)(*[int(x) for x in lcut])
        _check(self._L.mhx_torpol_to_spat_batch(
            self.h, nlev, _ptr(W), _ptr(dW), _ptr(Z), self.lm_max, _ptr(vr), _ptr(vt), _ptr(vp),
            self.n_phi * self.n_theta, lc, None))

    def spat_to_qst_batch(self, f, g, h, lcut, q, s, t):
        nlev = f.shape[0]
        lc = (ctypes.c_int * nlev)(*[int(x) for x in lcut])
        _check(self._L.mhx_spat_to_qst_batch(
            self.h, nlev, _ptr(f), _ptr(g), _ptr(h), self.n_phi * self.n_theta, _ptr(q), _ptr(s),
            _ptr(t), self.lm_max, lc, None))

    def set_timing(self, on=True):
        _check(self._L.mhx_set_timing(self.h, 1 if on else 0))

    def set_dudr(self, on=True):
        """whether the nonlinear step computes the du/dr grids (mhx_set_dudr)"""
        _check(self._L.mhx_set_dudr(self.h, 1 if on else 0))

    def stage_times(self):
        ms = (ctypes.c_double * 5)()
        _check(self._L.mhx_stage_times(self.h, ms))
        return list(ms)

    def launch_times(self):
        ms = (ctypes.c_double * 16)()
        n = self._L.mhx_launch_times(self.h, ms, 16)
        if n < 0:
            _check(n)
        return list(ms)[:n]


NL_INPUTS = ("w", "dw", "ddw", "z", "dz", "s", "b", "db", "ddb", "aj", "dj")
NL_OUTPUTS = ("AdvrLM", "AdvtLM", "AdvpLM", "VSrLM", "VStLM", "VSpLM", "VxBrLM", "VxBtLM",
              "VxBpLM", "uGsLM")
NL_GRIDS = ("vr", "vt", "vp", "cvr", "cvt", "cvp", "dvr", "dvt", "dvp", "br", "bt", "bp",
            "cbr", "cbt", "cbp", "s", "dsdt", "dsdp")


def nl_step(plan, inputs, outputs, or2, or4, orho1, LFfac, lcut, grids=None, stream=None):
    """The nonlinear-term path of one radial loop for all levels at once
    (src/rIter.f90:466-712, bulk levels).  ``inputs``: 11 device tensors
    (nlev, lm_max) complex128 in NL_INPUTS order; ``outputs``: 10 such tensors
    (NL_OUTPUTS order), written in place.  ``grids`` (optional): a float64
    device tensor (nlev, 18, n_phi, n_theta) that receives the inverse grids.
    ``stream`` (optional, a torch.cuda.Stream): the step is ordered on it and
    the call returns without waiting; without it the call returns when the
    outputs are written."""
    L = plan._L
    nlev = inputs[0].shape[0]
    ins = (ctypes.c_void_p * 11)(*[t.data_ptr() for t in inputs])
    outs = (ctypes.c_void_p * 10)(*[t.data_ptr() for t in outputs])
    d = lambda v: (ctypes.c_double * nlev)(*[float(x) for x in v])  # noqa: E731
    lc = (ctypes.c_int * nlev)(*[int(x) for x in lcut])
    st = ctypes.c_void_p(stream.cuda_stream) if stream is not None else None
    if grids is None:
        _check(L.mhx_nl_step(plan.h, nlev, ins, plan.lm_max, outs, d(or2), d(or4), d(orho1),
                             float(LFfac), lc, st))
    else:
        _check(L.mhx_nl_step_grids(plan.h, nlev, ins, plan.lm_max, outs, d(or2), d(or4),
                                   d(orho1), float(LFfac), lc, _ptr(grids), st))


def nl_step_td(plan, inputs, outputs, td_outputs, or2, or4, orho1, LFfac, epsc, CorFac, lcut,
               stream=None):
    """nl_step followed by the lm-space assembly of get_td on its outputs in
    one call (mhx_nl_step_td): ``td_outputs`` are 7 device tensors (nlev,
    lm_max) complex128 in TD_OUTPUTS order, written besides ``outputs``;
    ``epsc`` per level as get_td's (epsc * epscProf), ``CorFac`` as the
    reference's."""
    L = plan._L
    nlev = inputs[0].shape[0]
    ins = (ctypes.c_void_p * 11)(*[t.data_ptr() for t in inputs])
    outs = (ctypes.c_void_p * 10)(*[t.data_ptr() for t in outputs])
    tds = (ctypes.c_void_p * 7)(*[t.data_ptr() for t in td_outputs])
    d = lambda v: (ctypes.c_double * nlev)(*[float(x) for x in v])  # noqa: E731
    lc = (ctypes.c_int * nlev)(*[int(x) for x in lcut])
    st = ctypes.c_void_p(stream.cuda_stream) if stream is not None else None
    _check(L.mhx_nl_step_td(plan.h, nlev, ins, plan.lm_max, outs, tds, d(or2), d(or4), d(orho1),
                            float(LFfac), d(epsc), float(CorFac), lc, st))


def nl_step_courant(plan, inputs, outputs, or2, or4, orho1, LFfac, lcut, courfac, alffac,
                    valri2, valhi2):
    """nl_step that also returns the per-level Courant maxima (vr2max,
    vh2max) of src/courant.f90's magnetic branch, reduced inside the grid
    stage; dtrkc = sqrt(delxr2 / vr2max), dthkc = sqrt(delxh2 / vh2max)."""
    import numpy as np
    L = plan._L
    nlev = inputs[0].shape[0]
    ins = (ctypes.c_void_p * 11)(*[t.data_ptr() for t in inputs])
    outs = (ctypes.c_void_p * 10)(*[t.data_ptr() for t in outputs])
    d = lambda v: (ctypes.c_double * nlev)(*[float(x) for x in v])  # noqa: E731
    lc = (ctypes.c_int * nlev)(*[int(x) for x in lcut])
    vr2 = (ctypes.c_double * nlev)()
    vh2 = (ctypes.c_double * nlev)()
    _check(L.mhx_nl_step_courant(plan.h, nlev, ins, plan.lm_max, outs, d(or2), d(or4), d(orho1),
                                 float(LFfac), lc, float(courfac), float(alffac), d(valri2),
                                 d(valhi2), vr2, vh2, None))
    return np.array(vr2[:]), np.array(vh2[:])


TD_INPUTS = ("AdvrLM", "AdvtLM", "AdvpLM", "VStLM", "VxBrLM", "VxBtLM", "VxBpLM", "w", "dw", "z")
TD_OUTPUTS = ("dwdt", "dzdt", "dpdt", "dsdt", "dbdt", "djdt", "dVxBhLM")


def get_e_kin_levels(plan, w, dw, z, or2, orho1):
    """Per-level kinetic-energy sums (src/kinetic_energy.f90:119-160) of the
    (nlev, lm_max) complex arrays w, dw, z (numpy or device tensors): an
    (8, nlev) array e_p, e_t, e_p_as, e_t_as, e_p_es, e_t_es, e_p_eas,
    e_t_eas per level."""
    nlev = int(w.shape[0])
    arrs = []
    for a in (w, dw, z):
        if hasattr(a, "data_ptr"):
            arrs.append(a)
        else:
            arrs.append(np.ascontiguousarray(a, dtype=np.complex128).reshape(nlev, plan.lm_max))
    ptr = lambda a: ctypes.c_void_p(a.data_ptr() if hasattr(a, "data_ptr")  # noqa: E731
                                    else a.ctypes.data)
    d = lambda v: (ctypes.c_double * nlev)(*[float(x) for x in v])  # noqa: E731
    out = np.zeros((8, nlev))
    _check(plan._L.mhx_get_e_kin(plan.h, nlev, ptr(arrs[0]), ptr(arrs[1]), ptr(arrs[2]),
                                 plan.lm_max, d(or2), d(orho1), ptr(out)))
    return out


def get_e_kin_noalloc(plan, w, dw, z, or2, orho1, sums, stream=None):
    """The kinetic-energy sums of get_e_kin_levels through the entry that
    allocates nothing (mhx_get_e_kin_noalloc): w, dw, z (nlev, lm_max)
    complex, or2 and orho1 (nlev) float64 and sums (8, nlev) float64, all
    contiguous device tensors.  Queued on `stream` (a torch stream, or the
    null stream); sums are valid once that stream has run."""
    nlev = int(w.shape[0])
    for a in (w, dw, z, or2, orho1, sums):
        if not _is_torch(a) or not a.is_cuda:
            raise ValueError("get_e_kin_noalloc takes device tensors only")
    if tuple(sums.shape) != (8, nlev) or or2.numel() != nlev or orho1.numel() != nlev:
        raise ValueError("sums must be (8, nlev) and or2, orho1 nlev long")
    st = ctypes.c_void_p(stream.cuda_stream) if stream is not None else None
    _check(plan._L.mhx_get_e_kin_noalloc(plan.h, nlev, _ptr(w), _ptr(dw), _ptr(z), w.shape[1],
                                         _ptr(or2), _ptr(orho1), _ptr(sums), st))


def get_e_mag_levels(plan, b, db, aj, or2):
    """Per-level magnetic-energy sums of the outer core
    (src/magnetic_energy.f90:264-340) of the (nlev, lm_max) complex arrays b,
    db, aj (numpy or device tensors): a (10, nlev) array e_p, e_t, e_p_as,
    e_t_as, e_p_es, e_t_es, e_p_eas, e_t_eas, e_dipole, e_dipole_ax."""
    nlev = int(b.shape[0])
    arrs = []
    for a in (b, db, aj):
        if hasattr(a, "data_ptr"):
            arrs.append(a)
        else:
            arrs.append(np.ascontiguousarray(a, dtype=np.complex128).reshape(nlev, plan.lm_max))
    ptr = lambda a: ctypes.c_void_p(a.data_ptr() if hasattr(a, "data_ptr")  # noqa: E731
                                    else a.ctypes.data)
    out = np.zeros((10, nlev))
    _check(plan._L.mhx_get_e_mag(plan.h, nlev, ptr(arrs[0]), ptr(arrs[1]), ptr(arrs[2]),
                                 plan.lm_max, (ctypes.c_double * nlev)(*[float(x) for x in or2]),
                                 ptr(out)))
    return out


def get_td(plan, inputs, outputs, or2, or4, epsc, CorFac, lcut, stream=None):
    """lm-space assembly of the explicit terms at interior levels
    (src/get_td.f90 get_dwdt, get_dzdt, get_dpdt, get_dsdt, get_dbdt; rotating
    Boussinesq MHD, no double curl) for all levels at once.  ``inputs``: 10
    complex128 tensors (nlev, lm_max) in TD_INPUTS order (host or device);
    ``outputs``: 7 such tensors (TD_OUTPUTS order), written in place; per level
    ``or2``, ``or4``, ``epsc`` (epsc times epscProf) and ``lcut`` (l_R)."""
    L = plan._L
    nlev = inputs[0].shape[0]
    ins = (ctypes.c_void_p * 10)(*[t.data_ptr() for t in inputs])
    outs = (ctypes.c_void_p * 7)(*[t.data_ptr() for t in outputs])
    d = lambda v: (ctypes.c_double * nlev)(*[float(x) for x in v])  # noqa: E731
    lc = (ctypes.c_int * nlev)(*[int(x) for x in lcut])
    st = ctypes.c_void_p(stream.cuda_stream) if stream is not None else None
    _check(L.mhx_get_td(plan.h, nlev, ins, plan.lm_max, outs, d(or2), d(or4), d(epsc),
                        float(CorFac), lc, st))


# ---------------------------------------------------------------- radial loop
RLOOP_INT = ("l_conv", "l_conv_nl", "l_heat", "l_heat_nl", "l_chemical_conv", "l_phase_field",
             "l_mag", "l_mag_LF", "l_mag_nl", "l_mag_kin", "l_anel", "l_anelastic_liquid",
             "l_adv_curl", "l_double_curl", "l_corr", "l_single_matrix", "l_precession",
             "l_centrifuge", "l_cour_alf_damp", "lPressNext", "n_r_LCR")
RLOOP_DBL = ("LFfac", "CorFac", "epsc", "epscXi", "ViscHeatFac", "OhmLossFac", "opm", "courfac",
             "alffac", "epsPhase", "penaltyFac", "phaseDiffFac", "tmelt", "po", "prec_angle", "oek",
             "timeStage", "dilution_fac", "ra", "opr")
RLOOP_RAD = ("r", "or1", "or2", "or4", "orho1", "beta", "otemp1", "visc", "lambda", "temp0",
             "epscProf", "delxr2", "delxh2")
RLOOP_RAD_INT = ("nR", "l_R")
RLOOP_IN = ("w", "dw", "ddw", "z", "dz", "s", "xi", "phi", "b", "db", "ddb", "aj", "dj")
RLOOP_OUT = ("dwdt", "dzdt", "dpdt", "dsdt", "dVSrLM", "dxidt", "dVXirLM", "dphidt", "dbdt", "djdt",
             "dVxBhLM", "dVxVhLM")


class _RloopPhys(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in RLOOP_INT] + [(n, ctypes.c_double) for n in RLOOP_DBL]


class _RloopRadial(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in RLOOP_RAD + RLOOP_RAD_INT]


class _RloopFields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in RLOOP_IN]


class _RloopOut(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in RLOOP_OUT + ("dtrkc", "dthkc")]


def rloop(plan, phys, nlev, radial, fields, outputs, dtrkc, dthkc, ld_lm=None, stream=None):
    """The radial loop of interior levels (mhx_rloop) for any physics
    switches.  ``phys``: dict of the switches (RLOOP_INT, 0/1) and constants
    (RLOOP_DBL), missing entries 0; ``radial``: dict of per-level host arrays
    (RLOOP_RAD float64, RLOOP_RAD_INT int32) of nlev entries; ``fields`` and
    ``outputs``: dicts of (nlev, ld_lm) complex128 arrays (numpy or device
    tensors) by the names of RLOOP_IN / RLOOP_OUT (absent ones None);
    ``dtrkc``, ``dthkc``: float64 arrays of nlev written with the Courant
    steps."""
    p = _RloopPhys()
    for n in RLOOP_INT:
        setattr(p, n, int(phys.get(n, 0)))
    for n in RLOOP_DBL:
        setattr(p, n, float(phys.get(n, 0.0)))
    keep = []
    r = _RloopRadial()
    for n in RLOOP_RAD:
        a = np.ascontiguousarray(radial[n], dtype=np.float64)
        assert a.shape == (nlev,), (n, a.shape)
        keep.append(a)
        setattr(r, n, a.ctypes.data)
    for n in RLOOP_RAD_INT:
        a = np.ascontiguousarray(radial[n], dtype=np.int32)
        assert a.shape == (nlev,), (n, a.shape)
        keep.append(a)
        setattr(r, n, a.ctypes.data)
    ld = plan.lm_max if ld_lm is None else int(ld_lm)

    def ptr(a):
        if a is None:
            return None
        if _is_torch(a):
            assert a.is_contiguous() and a.shape[0] == nlev and a.shape[-1] >= ld
            return a.data_ptr()
        assert a.flags["C_CONTIGUOUS"] and a.dtype == np.complex128 and a.shape[0] == nlev
        return a.ctypes.data

    f = _RloopFields()
    for n in RLOOP_IN:
        setattr(f, n, ptr(fields.get(n)))
    o = _RloopOut()
    for n in RLOOP_OUT:
        setattr(o, n, ptr(outputs.get(n)))
    for n, a in (("dtrkc", dtrkc), ("dthkc", dthkc)):
        if _is_torch(a):
            setattr(o, n, a.data_ptr())
        else:
            assert a.flags["C_CONTIGUOUS"] and a.dtype == np.float64 and a.shape == (nlev,)
            setattr(o, n, a.ctypes.data)
    st = ctypes.c_void_p(stream.cuda_stream) if stream is not None else None
    _check(plan._L.mhx_rloop(plan.h, ctypes.byref(p), nlev, ctypes.byref(r), ctypes.byref(f), ld,
                             ctypes.byref(o), st))


JOB_SCALAR, JOB_VECTOR = 0, 1
BC_CMB, BC_ICB = 0, 1


class _Job(ctypes.Structure):
    _fields_ = [("l_max", ctypes.c_int), ("nlev", ctypes.c_int), ("kind", ctypes.c_int),
                ("nfields", ctypes.c_int), ("inp", ctypes.c_void_p), ("out", ctypes.c_void_p)]


class Batch:
    """An uneven batch of independent transform jobs run from one captured
    graph (mhx_batch_*): one persistent launch over a work list of all the
    jobs, or (MHX_BATCH_FORM=graph) the jobs' launches on graph branches.  ``jobs``: (l_max, kind, inp, out) with ``inp`` and
    ``out`` complex128 device tensors of shape (ncomp, nfields, nlev, lm_max),
    ncomp 1 (JOB_SCALAR: s) or 3 (JOB_VECTOR: W, dW, Z in; Q, S, T out).
    Each job is the round trip synthesis -> grids -> analysis of its fields
    (scal_to_spat + scal_to_SH, or torpol_to_spat + spat_to_qst)."""

    def __init__(self, jobs):
        L = lib()
        self._keep = []
        arr = (_Job * len(jobs))()
        for i, (l_max, kind, inp, out) in enumerate(jobs):
            nc, nf, nlev, lm = inp.shape
            if out.shape != inp.shape or nc != (3 if kind == JOB_VECTOR else 1):
                raise MhxError("job %d: bad array shapes" % i)
            if lm != (l_max + 1) * (l_max + 2) // 2:
                raise MhxError("job %d: lm_max does not match l_max" % i)
            arr[i] = _Job(l_max, nlev, kind, nf, inp.data_ptr(), out.data_ptr())
            self._keep += [inp, out]
        h = ctypes.c_void_p()
        _check(L.mhx_batch_create(arr, len(jobs), PREC_SPLIT16, ctypes.byref(h)))
        self.h = h

    def execute(self, stream=None):
        """Replays the batch's graph; with ``stream`` (a torch.cuda.Stream) it
        returns at once, ordered on the stream."""
        st = ctypes.c_void_p(stream.cuda_stream) if stream is not None else None
        _check(lib().mhx_batch_execute(self.h, st))

    def tuning(self):
        """[(l_max, synthesis form, analysis form, row-FFT rows per block)]
        as chosen when the batch was made (forms: 1 whole, 2 chunked)."""
        n = 16
        a = [(ctypes.c_int * n)() for _ in range(4)]
        k = lib().mhx_batch_tuning(self.h, *a, n)
        if k < 0:
            _check(k)
        return [tuple(x[i] for x in a) for i in range(min(k, n))]

    def shapes(self):
        """(form, [(l_max, orders per item, level groups per item below the
        split order, level groups per item from it on, split order index,
        theta rows per row item)]): form 1 is the one persistent launch (the
        default), 0 the per-job launches (MHX_BATCH_FORM=graph)."""
        n = 16
        form = ctypes.c_int()
        a = [(ctypes.c_int * n)() for _ in range(6)]
        k = lib().mhx_batch_shapes(self.h, ctypes.byref(form), *a, n)
        if k < 0:
            _check(k)
        return form.value, [tuple(x[i] for x in a) for i in range(min(k, n))]

    def status(self):
        """Waits for the batch's executions so far; raises MhxError if a work
        item of the persistent launch gave up waiting for its inputs."""
        _check(lib().mhx_batch_status(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().mhx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
