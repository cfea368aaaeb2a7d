"""The calls the golden files record, per module sht routine.

Each entry maps a routine name to a function (impl, A, G, lcut, r, r_icb, or2)
-> tuple of outputs, written against the names and argument order of the
reference's module sht (src/sht_native.f90:16-20).  ``impl`` is anything
exposing that API: the oracle Plan or the HIP-backed magic_amd.sht.Plan.
"""

API_CALLS = {
    "scal_to_spat": lambda p, A, G, lc, r, ri, o2: (p.scal_to_spat(A[0], lc),),
    "scal_to_grad_spat": lambda p, A, G, lc, r, ri, o2: p.scal_to_grad_spat(A[0], lc),
    "pol_to_grad_spat": lambda p, A, G, lc, r, ri, o2: p.pol_to_grad_spat(A[0], lc),
    "torpol_to_spat": lambda p, A, G, lc, r, ri, o2: p.torpol_to_spat(A[0], A[1], A[2], lc),
    "pol_to_curlr_spat": lambda p, A, G, lc, r, ri, o2: (p.pol_to_curlr_spat(A[0], lc),),
    "torpol_to_curl_spat": lambda p, A, G, lc, r, ri, o2:
        p.torpol_to_curl_spat(o2, A[0], A[1], A[2], A[3], lc),
    "torpol_to_dphspat": lambda p, A, G, lc, r, ri, o2: p.torpol_to_dphspat(A[0], A[1], lc),
    "sphtor_to_spat": lambda p, A, G, lc, r, ri, o2: p.sphtor_to_spat(A[0], A[1], lc),
    "torpol_to_spat_IC": lambda p, A, G, lc, r, ri, o2:
        p.torpol_to_spat_IC(r, ri, A[0], A[1], A[2]),
    "torpol_to_curl_spat_IC": lambda p, A, G, lc, r, ri, o2:
        p.torpol_to_curl_spat_IC(r, ri, A[0], A[1], A[2], A[3]),
    "axi_to_spat": lambda p, A, G, lc, r, ri, o2: (p.axi_to_spat(A[0][:p.l_max + 1]),),
    "toraxi_to_spat": lambda p, A, G, lc, r, ri, o2: p.toraxi_to_spat(A[0][:p.l_max + 1], lc),
    "scal_to_SH": lambda p, A, G, lc, r, ri, o2: (p.scal_to_SH(G[0], lc),),
    "spat_to_sphertor": lambda p, A, G, lc, r, ri, o2: p.spat_to_sphertor(G[0], G[1], lc),
    "spat_to_qst": lambda p, A, G, lc, r, ri, o2: p.spat_to_qst(G[0], G[1], G[2], lc),
}

GOLDEN_API_FILES = ["api_L16_m1.npz", "api_L16_m1_cplxm0.npz", "api_L16_m4.npz",
                    "api_L64_m1.npz", "api_L64_m4.npz"]


def levels(g):
    return len(g["lcut"])


def call(impl, name, g, lev):
    return API_CALLS[name](impl, g["in_coefs_%d" % lev], g["in_grids_%d" % lev],
                           int(g["lcut"][lev]), float(g["r"]), float(g["r_icb"]),
                           float(g["or2"]))
