"""The calls the golden files record, per module sht routine.

Each entry maps a routine name to a function (impl, A, G, lcut, r, r_icb, or2)
-> tuple of outputs, written against the names and argument order of the
reference's module sht (src/sht_native.f90:16-20).  ``impl`` is anything
exposing that API: the oracle Plan or the HIP-backed magic_amd.sht.Plan.
"""
import numpy as np

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


# order of results written by the 'api' mode of oracle/_ref/ref_harness (and
# oracle/_ref/shim_driver), per level:
# (name, kind, count)  kind: g = grid, a = n_theta vector, c = coefficient vector
API_OUTPUTS = [
    ("scal_to_spat", "g", 1), ("scal_to_grad_spat", "g", 2), ("pol_to_grad_spat", "g", 2),
    ("torpol_to_spat", "g", 3), ("pol_to_curlr_spat", "g", 1),
    ("torpol_to_curl_spat", "g", 3), ("torpol_to_dphspat", "g", 2),
    ("sphtor_to_spat", "g", 2), ("torpol_to_spat_IC", "g", 3),
    ("torpol_to_curl_spat_IC", "g", 3), ("axi_to_spat", "a", 1),
    ("toraxi_to_spat", "a", 2), ("scal_to_SH", "c", 1), ("spat_to_sphertor", "c", 2),
    ("spat_to_qst", "c", 3),
]


def read_header(f):
    h = np.fromfile(f, dtype="<i4", count=6)
    return dict(zip(("l_max", "m_max", "minc", "n_theta", "n_phi", "lm_max"), h.tolist()))


def write_api_input(path, levels):
    with open(path, "wb") as f:
        for L in levels:
            np.array([L["lcut"]], dtype="<i4").tofile(f)
            np.array([L["r"], L["r_icb"], L["or2"]], dtype="<f8").tofile(f)
            for c in L["coefs"]:
                np.ascontiguousarray(c, dtype="<c16").tofile(f)
            np.ascontiguousarray(L["grids"], dtype="<f8").tofile(f)


def read_api_output(path, nlev):
    res = []
    with open(path, "rb") as f:
        hd = read_header(f)
        nt, npz, lm = hd["n_theta"], hd["n_phi"], hd["lm_max"]
        for _ in range(nlev):
            out = {}
            for name, kind, cnt in API_OUTPUTS:
                arrs = []
                for _k in range(cnt):
                    if kind == "g":
                        arrs.append(np.fromfile(f, "<f8", nt * npz).reshape(npz, nt))
                    elif kind == "a":
                        arrs.append(np.fromfile(f, "<f8", nt))
                    else:
                        arrs.append(np.fromfile(f, "<c16", lm))
                out[name] = np.stack(arrs)
            res.append(out)
        assert f.read() == b"", "trailing data in harness output"
    return hd, res


def api_levels(g):
    """The per-level inputs of a golden api file, as write_api_input takes them."""
    return [dict(lcut=int(g["lcut"][i]), r=float(g["r"]), r_icb=float(g["r_icb"]),
                 or2=float(g["or2"]), coefs=list(g["in_coefs_%d" % i]),
                 grids=g["in_grids_%d" % i]) for i in range(levels(g))]
