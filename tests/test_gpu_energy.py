"""Kinetic-energy diagnostic on the GPU (mhx_get_e_kin + magic_amd.energy,
src/kinetic_energy.f90 get_e_kin): the energies of the reference's
boussBenchSat and doubleDiffusion restart states against the first row of its e_kin reference
output (2e-8 relative, the printed precision), the per-level sums against
the oracle (1e-13) with host and device arrays, and at l_max=255 x 97."""
import os

import numpy as np
import pytest

from magic_amd import checkpoint, energy, sht, workload
from oracle import oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("sample", ["boussBenchSat", "doubleDiffusion"])
def test_e_kin_matches_reference_output(sample):
    import torch
    print("e_kin", sample, flush=True)
    h = checkpoint.read_checkpoint(os.path.join(GOLDEN, sample + "_checkpoint_end.start"))
    ref = np.loadtxt(os.path.join(GOLDEN, sample + "_e_kin_reference.out"))
    r, w, z = h["r"], h["w"], h["z"]
    dw = oracle.cheb_dr(w, r)
    nt, nph = workload.grid_size(h["l_max"], h["nalias"], h["minc"])
    p = sht.Plan(h["l_max"], nt, nph, minc=h["minc"])
    e = energy.get_e_kin(p, w, dw, z, r)
    np.testing.assert_allclose([e[k] for k in energy.E_KIN_NAMES], ref[0, 1:9], rtol=2e-8, atol=0)
    want = oracle.e_kin_levels(w, dw, z, 1.0 / r ** 2, np.ones_like(r), h["l_max"], h["minc"])
    got = sht.get_e_kin_levels(p, w, dw, z, 1.0 / r ** 2, np.ones_like(r))
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    dev = torch.device("cuda")
    tw, tdw, tz = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (w, dw, z))
    got_d = sht.get_e_kin_levels(p, tw, tdw, tz, 1.0 / r ** 2, np.ones_like(r))
    np.testing.assert_allclose(got_d, got, rtol=1e-15, atol=0)
    p.close()


def test_e_kin_levels_large():
    L, nlev = 255, 97
    print("e_kin levels", L, nlev, flush=True)
    w, dw, _, z = workload.nl_inputs(L, nlev, stream=2, kind="l^-5/3")[:4]
    r = np.linspace(0.54, 1.54, nlev)
    orho = 1.0 / (1.0 + 0.1 * r)
    p = sht.Plan(L)
    got = sht.get_e_kin_levels(p, w, dw, z, 1.0 / r ** 2, orho)
    want = oracle.e_kin_levels(w, dw, z, 1.0 / r ** 2, orho, L, 1)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())
    p.close()
