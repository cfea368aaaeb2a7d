"""Courant maxima fused into the grid stage (mhx_nl_step_courant,
src/courant.f90 magnetic branch): FP64 against the reference's courant
(golden, l_max=16) and the oracle (l_max=64, slot and generic grid stages);
the step's outputs are those of mhx_nl_step; the split form's maxima (FP32
grids) within 1e-5; repeated calls (graph replay) give the same maxima."""
import numpy as np
import pytest

from conftest import load_golden, rel_l2
from magic_amd import sht, workload
from nl_reference import make_inputs, oracle_step, radial

pytestmark = pytest.mark.gpu


def _vals(opm, dr2, dh2):
    return (0.5 * (1 + opm)) ** 2 / np.asarray(dr2), (0.5 * (1 + opm)) ** 2 / np.asarray(dh2)


def _run(p, inputs, rad, lcut, cf, af, vri2, vhi2):
    import torch
    dev = torch.device("cuda")
    nlev = inputs[0].shape[0]
    tin = [torch.from_numpy(x).to(dev) for x in inputs]
    tout = [torch.zeros((nlev, p.lm_max), dtype=torch.complex128, device=dev) for _ in range(10)]
    vr2, vh2 = sht.nl_step_courant(p, tin, tout, rad[0], rad[1], rad[2], rad[3], lcut, cf, af,
                                   vri2, vhi2)
    return vr2, vh2, [t.cpu().numpy() for t in tout], tin, tout


def test_courant_matches_reference_golden():
    g = load_golden("courant_L16.npz")
    L, nlev = int(g["l_max"]), int(g["nlev"])
    nt, nph = workload.grid_size(L)
    print("courant golden", flush=True)
    p = sht.Plan(L, nt, nph, max_levels=nlev)
    vri2, vhi2 = _vals(float(g["opm"]), g["delxr2"], g["delxh2"])
    vr2, vh2, _, _, _ = _run(p, make_inputs(L, nlev), radial(nlev), list(g["lcut"]),
                             float(g["courfac"]), float(g["alffac"]), vri2, vhi2)
    p.close()
    assert np.allclose(np.sqrt(g["delxr2"] / vr2), g["dtrkc"], rtol=1e-12, atol=0)
    assert np.allclose(np.sqrt(g["delxh2"] / vh2), g["dthkc"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("n_phi", [192, 200])
def test_courant_matches_oracle(n_phi):
    """n_phi = 192: slot grid stage; 200: generic grid stage"""
    import torch
    from oracle import oracle
    L, nt, nlev = 64, 96, 5
    print("courant oracle", n_phi, flush=True)
    inputs = make_inputs(L, nlev)
    rad = radial(nlev)
    lcut = [L] * nlev
    lcut[2] = L - 6
    cf, af, opm = 2.5, 1.0, 0.2
    dr2 = 1e-3 * (1.0 + np.arange(nlev))
    dh2 = 2e-2 / (1.0 + np.arange(nlev))
    vri2, vhi2 = _vals(opm, dr2, dh2)
    o = oracle.Plan(L, nt, n_phi)
    G, _ = oracle_step(o, inputs, *rad, lcut)
    ref = [o.courant(rad[0][i], rad[1][i], rad[2][i], rad[3], cf, af, vri2[i], vhi2[i],
                     [G[i, c] for c in (0, 1, 2, 9, 10, 11)]) for i in range(nlev)]
    ref = np.array(ref)
    p = sht.Plan(L, nt, n_phi, max_levels=nlev)
    vr2, vh2, outs, tin, tout = _run(p, inputs, rad, lcut, cf, af, vri2, vhi2)
    assert np.allclose(vr2, ref[:, 0], rtol=1e-12, atol=0)
    assert np.allclose(vh2, ref[:, 1], rtol=1e-12, atol=0)
    # the outputs are those of the plain step
    t2 = [torch.zeros_like(t) for t in tout]
    sht.nl_step(p, tin, t2, rad[0], rad[1], rad[2], rad[3], lcut)
    torch.cuda.synchronize()
    for k in range(10):
        assert np.array_equal(t2[k].cpu().numpy(), outs[k])
    # repeated calls on the same arrays (the second captures the step's
    # graph, the third replays it)
    for _ in range(2):
        a, b = sht.nl_step_courant(p, tin, tout, rad[0], rad[1], rad[2], rad[3], lcut, cf, af,
                                   vri2, vhi2)
        assert np.array_equal(a, vr2) and np.array_equal(b, vh2)
    p.close()
    p16 = sht.Plan(L, nt, n_phi, precision=sht.PREC_SPLIT16, max_levels=nlev)
    a, b, _, _, _ = _run(p16, inputs, rad, lcut, cf, af, vri2, vhi2)
    p16.close()
    assert np.allclose(a, ref[:, 0], rtol=1e-5, atol=0)
    assert np.allclose(b, ref[:, 1], rtol=1e-5, atol=0)
