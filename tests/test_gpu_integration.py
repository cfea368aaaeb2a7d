"""The 200-step integration check: the reference program with its radial
loop replaced by the GPU one (oracle/_ref/magic_mhx: the reference's own
time stepper, implicit solves and boundary levels on the host, every
interior level of every stage through mhx_rloop) against the reference
program itself (golden series written by tests/golden/make_integration_golden.py).

Per case (tests/integration_cases.py: the boussBenchSat dynamo benchmark
state, and an anelastic dynamo growing from the reference's initial fields):
* the logged kinetic and magnetic energies of all 21 log steps equal the
  reference's to their printed precision (9 digits, 2e-8 relative);
* the kinetic and outer-core magnetic energies of the final state, from the
  full-precision restart files of both runs (oracle.e_kin_levels,
  e_mag_levels, pinned by test_energy.py), agree to 1e-10 relative;
* the final fields agree to 1e-9 relative L2."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import rel_l2
import integration_cases as ic
from magic_amd import checkpoint
from oracle import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAGIC_MHX = os.path.join(ROOT, "oracle", "_ref", "magic_mhx")


def _energies(h, fields):
    r = h["r"]
    L, minc = h["l_max"], h["minc"]
    w, z = fields["w"], fields["z"]
    ek = oracle.e_kin_levels(w, oracle.cheb_dr(w, r), z, 1.0 / r ** 2, np.ones_like(r), L, minc)
    b, aj = fields["b"], fields["aj"]
    em = oracle.e_mag_levels(b, oracle.cheb_dr(b, r), aj, 1.0 / r ** 2, L, minc)
    return 0.5 * oracle.rint_interp(ek, r), 0.5 * oracle.rint_interp(em[:8], r)


@pytest.mark.parametrize("case", list(ic.CASES))
def test_200_steps_match_reference(case):
    if not os.path.exists(MAGIC_MHX):
        pytest.skip("oracle/_ref/magic_mhx is built only where the reference sources are")
    with np.load(os.path.join(ic.GOLDEN, "integ_%s.npz" % case), allow_pickle=False) as z:
        gold = {k: z[k] for k in z.files}
    with tempfile.TemporaryDirectory() as td:
        tag = ic.prepare(case, td)
        print("integration", case, "in", td, flush=True)
        env = dict(os.environ, OMP_NUM_THREADS="8", OMP_STACKSIZE="1G")
        r = subprocess.run(["bash", "-c", "ulimit -s unlimited; %s input.nml" % MAGIC_MHX],
                           cwd=td, env=env, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
        ek = np.loadtxt(os.path.join(td, "e_kin." + tag))
        em = np.loadtxt(os.path.join(td, "e_mag_oc." + tag))
        h = checkpoint.read_checkpoint(os.path.join(td, "checkpoint_end." + tag))
    # logged series, to print precision
    assert ek.shape == gold["e_kin"].shape
    np.testing.assert_allclose(ek[:, 0], gold["e_kin"][:, 0], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(ek[:, 1:], gold["e_kin"][:, 1:], rtol=2e-8, atol=1e-300)
    np.testing.assert_allclose(em[:, 1:], gold["e_mag_oc"][:, 1:], rtol=2e-8, atol=1e-300)
    # final state in full precision
    assert abs(h["time"] - float(gold["time"])) < 1e-12 * abs(float(gold["time"])) + 1e-14
    for k in ic.FIELDS:
        e = rel_l2(h[k], gold[k])
        print("  final %-3s rel L2 %.2e" % (k, e), flush=True)
        assert e < 1e-9, (case, k, e)
    ekg, emg = _energies(h, gold)
    ekr, emr = _energies(h, h)
    dk = np.max(np.abs(ekr - ekg) / np.abs(ekg).clip(1e-300))
    dm = np.max(np.abs(emr - emg) / np.abs(emg).clip(1e-300))
    print("  final energies: kinetic %.2e, magnetic %.2e (max relative difference)" % (dk, dm),
          flush=True)
    assert dk < 1e-10 and dm < 1e-10, (dk, dm)


def test_200_steps_split_plan_stay_near_reference():
    """The boussBenchSat case with MHX_PRECISION=split: the plugin's plan is
    split, so its interior levels run the fused split step (the switch set
    the fused step covers).  After 200 steps the logged kinetic and magnetic
    energies stay within 2e-6 of the reference program's and the final
    fields within 2e-6 relative L2 (measured: 5.6e-7 and 1.7e-7 for the
    energies, 4e-7 at most for the fields)."""
    if not os.path.exists(MAGIC_MHX):
        pytest.skip("oracle/_ref/magic_mhx is built only where the reference sources are")
    case = "bouss200"
    with np.load(os.path.join(ic.GOLDEN, "integ_%s.npz" % case), allow_pickle=False) as z:
        gold = {k: z[k] for k in z.files}
    with tempfile.TemporaryDirectory() as td:
        tag = ic.prepare(case, td)
        print("integration", case, "split plan, in", td, flush=True)
        env = dict(os.environ, OMP_NUM_THREADS="8", OMP_STACKSIZE="1G", MHX_PRECISION="split")
        r = subprocess.run(["bash", "-c", "ulimit -s unlimited; %s input.nml" % MAGIC_MHX],
                           cwd=td, env=env, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
        ek = np.loadtxt(os.path.join(td, "e_kin." + tag))
        em = np.loadtxt(os.path.join(td, "e_mag_oc." + tag))
        h = checkpoint.read_checkpoint(os.path.join(td, "checkpoint_end." + tag))
    assert ek.shape == gold["e_kin"].shape
    np.testing.assert_allclose(ek[:, 1:], gold["e_kin"][:, 1:], rtol=2e-6, atol=1e-300)
    np.testing.assert_allclose(em[:, 1:], gold["e_mag_oc"][:, 1:], rtol=2e-6, atol=1e-300)
    for k in ic.FIELDS:
        e = rel_l2(h[k], gold[k])
        print("  final %-3s rel L2 %.2e" % (k, e), flush=True)
        assert e < 2e-6, (k, e)
