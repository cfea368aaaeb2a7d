"""Kinetic-energy diagnostic of the reference (src/kinetic_energy.f90
get_e_kin, :69-276) on the GPU path.

The lm sums per radial level run on the device (``mhx_get_e_kin``); the
radial integral of each profile is the reference's ``rInt_R`` for the
Chebyshev scheme (src/integration.f90:56-96): the integral over [r_icb,
r_cmb] of the polynomial interpolating f at the Gauss-Lobatto radii, i.e.
Clenshaw-Curtis quadrature in x = the mapped Chebyshev coordinate.
"""
import numpy as np

from . import sht

E_KIN_NAMES = ("e_p", "e_t", "e_p_as", "e_t_as", "e_p_es", "e_t_es", "e_p_eas", "e_t_eas")


def rint_cheb(f, r):
    """Integral over r of the profile f given at the Gauss-Lobatto radii r
    (outer boundary first) of a linearly mapped Chebyshev grid.

    With x_k = cos(pi k / (N-1)) and g = f dr/dx, the interpolant's
    coefficients are a_n = 2/(N-1) sum''_k g_k cos(pi n k / (N-1)) (first and
    last terms halved) and its integral over [-1, 1] is
    sum''_n even a_n 2 / (1 - n^2)."""
    f = np.asarray(f, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    N = r.size
    if f.shape[-1] != N or N < 3:
        raise ValueError("profile and radius sizes differ (or fewer than 3 radii)")
    dr_dx = 0.5 * (r[0] - r[-1])
    x = np.cos(np.pi * np.arange(N) / (N - 1))
    if not np.allclose(r, r[-1] + dr_dx * (1.0 + x), rtol=0, atol=1e-12 * abs(r[0])):
        raise ValueError("radii are not a linearly mapped Gauss-Lobatto grid")
    g = f * dr_dx
    k = np.arange(N)
    C = np.cos(np.pi * np.outer(k, k) / (N - 1))
    hw = np.ones(N)
    hw[0] = hw[-1] = 0.5
    a = (2.0 / (N - 1)) * ((g * hw) @ C)
    wn = np.zeros(N)
    ev = k[k % 2 == 0].astype(np.float64)
    wn[k % 2 == 0] = 2.0 / (1.0 - ev ** 2)
    wn *= hw
    return a @ wn


def get_e_kin(plan, w, dw, z, r, orho1=None, e_scale=1.0):
    """Kinetic energies of the outer core as get_e_kin writes them
    (e_kin.TAG columns 2-9): poloidal, toroidal, their axisymmetric,
    equatorially symmetric and equatorially + axially symmetric parts.
    ``w``, ``dw``, ``z``: (n_r, lm_max) complex at radii ``r`` (outer first,
    numpy or device tensors); ``orho1`` defaults to 1 (Boussinesq)."""
    r = np.asarray(r, dtype=np.float64)
    or2 = 1.0 / r ** 2
    if orho1 is None:
        orho1 = np.ones_like(r)
    sums = sht.get_e_kin_levels(plan, w, dw, z, or2, orho1)
    e = 0.5 * e_scale * rint_cheb(sums, r)
    return dict(zip(E_KIN_NAMES, (float(v) for v in e)))
