"""GPU parity of the HIP path (libmhx.so through the C ABI) against the
reference's own outputs (golden files) and the FP64 oracle.

The bar is the FP64 plain form: <= 1e-12 relative L2 per output field.
"""
import numpy as np
import pytest

from conftest import load_golden, rel_l2
from api_cases import API_CALLS, GOLDEN_API_FILES, call, levels
from magic_amd import sht, workload

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _plan_for(g):
    return sht.Plan(int(g["l_max"]), int(g["n_theta"]), int(g["n_phi"]), minc=int(g["minc"]))


@pytest.mark.parametrize("fname", GOLDEN_API_FILES)
def test_api_matches_reference_golden(fname):
    g = load_golden(fname)
    print("golden", fname, flush=True)
    p = _plan_for(g)
    for lev in range(levels(g)):
        for name in API_CALLS:
            got = call(p, name, g, lev)
            ref = g["out_%s_%d" % (name, lev)]
            assert len(got) == ref.shape[0], name
            for k in range(ref.shape[0]):
                err = rel_l2(got[k], ref[k])
                assert err < TOL, (fname, lev, name, k, err)


def test_tables_match_oracle():
    from oracle import oracle
    for L, minc in ((16, 1), (64, 4), (255, 1)):
        nt, nph = workload.grid_size(L, 20, minc)
        print("tables", L, minc, flush=True)
        p = sht.Plan(L, nt, nph, minc=minc)
        o = oracle.Plan(L, nt, nph, minc=minc)
        a, b = p.tables(), o.tables()
        np.testing.assert_allclose(a["theta"], b["theta"], atol=1e-15)
        np.testing.assert_allclose(a["gauss"], b["gauss"], rtol=1e-14)
        for k in ("plm", "dplm"):
            scale = np.abs(b[k]).max()
            assert np.abs(a[k] - b[k]).max() <= 1e-13 * scale, (L, k)


@pytest.mark.parametrize("L", [64, 255])
def test_api_matches_oracle_large(L):
    """l_max=64 and l_max=255 shapes, every routine, one level."""
    from oracle import oracle
    nt, nph = workload.grid_size(L)
    print("oracle compare", L, nt, nph, flush=True)
    p = sht.Plan(L, nt, nph)
    o = oracle.Plan(L, nt, nph)
    A = [workload.coefficients(L, 2, f, 0, "l^-5/3") for f in range(4)]
    u = workload.uniform01(workload.stream_seed(2, 77, 0), 3 * nph * nt)
    G = (2 * u - 1).reshape(3, nph, nt)
    g = {"in_coefs_0": np.stack(A), "in_grids_0": G, "lcut": np.array([L]), "r": 0.3,
         "r_icb": 0.35, "or2": 1.0 / 0.81}
    names = list(API_CALLS) if L <= 64 else ["scal_to_spat", "torpol_to_spat",
                                             "torpol_to_curl_spat", "spat_to_qst",
                                             "scal_to_grad_spat", "spat_to_sphertor"]
    for name in names:
        got = call(p, name, g, 0)
        ref = call(o, name, g, 0)
        for k in range(len(ref)):
            err = rel_l2(got[k], ref[k])
            assert err < TOL, (L, name, k, err)


def test_lcut_and_errors():
    p = sht.Plan(16)
    S = workload.coefficients(16, 5, 0, 0, potential=False)
    with pytest.raises(sht.MhxError):
        p.scal_to_spat(S, 17)
    with pytest.raises(sht.MhxError):
        p.scal_to_spat(S, -1)
    f0 = p.scal_to_spat(S, 0)   # only l = 0, m = 0 survives
    assert np.allclose(f0, S[0].real / np.sqrt(4 * np.pi), atol=1e-15)


def test_batched_equals_per_level():
    import torch
    L = 32
    nt, nph = workload.grid_size(L)
    nlev = 11  # not a multiple of the 8-level column tile
    p = sht.Plan(L, nt, nph, max_levels=nlev)
    lc = [L if i % 3 else 20 for i in range(nlev)]
    W = np.stack([workload.coefficients(L, 3, 0, i) for i in range(nlev)])
    dW = np.stack([workload.coefficients(L, 3, 1, i) for i in range(nlev)])
    Z = np.stack([workload.coefficients(L, 3, 2, i) for i in range(nlev)])
    dev = torch.device("cuda")
    tW, tdW, tZ = (torch.from_numpy(x).to(dev) for x in (W, dW, Z))
    vr, vt, vp = (torch.zeros((nlev, nph, nt), dtype=torch.float64, device=dev) for _ in range(3))
    p.torpol_to_spat_batch(tW, tdW, tZ, lc, vr, vt, vp)
    torch.cuda.synchronize()
    q, s, t = (torch.zeros((nlev, p.lm_max), dtype=torch.complex128, device=dev) for _ in range(3))
    p.spat_to_qst_batch(vr, vt, vp, lc, q, s, t)
    torch.cuda.synchronize()
    for i in range(nlev):
        ref = p.torpol_to_spat(W[i], dW[i], Z[i], lc[i])
        for a, b in zip((vr, vt, vp), ref):
            assert rel_l2(a[i].cpu().numpy(), b) < 1e-14, i
        refq = p.spat_to_qst(vr[i].cpu().numpy(), vt[i].cpu().numpy(), vp[i].cpu().numpy(), lc[i])
        for a, b in zip((q, s, t), refq):
            assert rel_l2(a[i].cpu().numpy(), b) < 1e-14, i
