"""Pins the oracle's radial loop of any physics switches (orc_rl_products,
orc_rl_courant, orc_rl_td composed in the reference's call order,
tests/rloop_reference.py) against the reference's own radial loop: the
fixtures are dumps of one radialLoopG call of the reference program under
six physics configurations (tests/golden/make_rloop_golden.py).  Interior
levels, every output and the Courant steps within 1e-12 relative L2.
A planted mistake in a product formula is caught."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle
import rloop_cases as rc
from rloop_reference import oracle_rloop

TOL = 1e-12


def _oracle_case(d):
    sel = rc.interior(d)
    pa = rc.plan_args(d)
    o = oracle.Plan(pa["l_max"], pa["n_theta"], pa["n_phi"], m_max=pa["m_max"], minc=pa["minc"])
    return sel, oracle_rloop(o, rc.phys_of(d), sel.size, rc.radial_of(d, sel), rc.fields_of(d, sel),
                             pa["minc"])


@pytest.mark.parametrize("case", rc.CASES)
def test_oracle_rloop_matches_reference(case):
    d = rc.load(case)
    sel, (outs, dtr, dth) = _oracle_case(d)
    assert set(outs) == set(rc.outputs_of(d))
    for n, got in outs.items():
        e = rel_l2(got, d[n][sel])
        assert e < TOL, (case, n, e)
        assert np.linalg.norm(d[n][sel]) > 0, (case, n)
    assert rel_l2(dtr, d["dtrkc"][sel]) < TOL
    assert rel_l2(dth, d["dthkc"][sel]) < TOL


def test_fixtures_cover_the_physics_switches():
    seen = {}
    for case in rc.CASES:
        d = rc.load(case)
        for k in ("l_mag", "l_adv_curl", "l_chemical_conv", "l_anel", "l_precession",
                  "l_phase_field", "l_centrifuge", "l_cond_ic"):
            seen.setdefault(k, set()).add(int(d[k]))
    for k, v in seen.items():
        assert v == {0, 1}, (k, v)


def test_planted_product_mistake_is_caught():
    """a sign flip in one advection term of the oracle's inputs shows up"""
    d = rc.load("anel_mhd")
    sel = rc.interior(d)
    pa = rc.plan_args(d)
    o = oracle.Plan(pa["l_max"], pa["n_theta"], pa["n_phi"], m_max=pa["m_max"], minc=pa["minc"])
    f = rc.fields_of(d, sel)
    f["dz"] = -f["dz"]  # dv_p/dr enters u.grad(u) and the viscous heating
    outs, _, _ = oracle_rloop(o, rc.phys_of(d), sel.size, rc.radial_of(d, sel), f, pa["minc"])
    assert rel_l2(outs["dzdt"], d["dzdt"][sel]) > 1e-3
