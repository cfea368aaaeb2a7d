"""Test helpers for the radial loop of any physics switches (mhx_rloop):
reads a reference dump fixture (tests/golden/rloop_<case>.npz, written by
tests/golden/make_rloop_golden.py from the reference's own radial loop) into
the switch, radial, field and output dictionaries of magic_amd.sht.rloop."""
import os

import numpy as np

from magic_amd import sht

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("bouss_mhd", "bouss_nocurl", "doublediff", "anel_mhd", "precession", "phase")


def load(case):
    with np.load(os.path.join(GOLDEN, "rloop_%s.npz" % case), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def phys_of(d):
    """the mhx_rloop_phys values of a dump"""
    ph = {}
    for n in sht.RLOOP_INT:
        if n == "n_r_LCR":
            ph[n] = int(d["n_r_LCR"])
        elif n == "lPressNext":
            ph[n] = int(d["lPressNext"])
        else:
            ph[n] = int(d[n]) if n in d else 0
    for n in sht.RLOOP_DBL:
        if n == "timeStage":
            ph[n] = float(d["timeStage"])
        else:
            ph[n] = float(d[n]) if n in d else 0.0
    return ph


def interior(d):
    """indices (into the dump's kept levels) of the interior levels"""
    lev = d["levels"]
    n_r = int(d["n_r_max"])
    return np.nonzero((lev > 1) & (lev < n_r))[0]


def radial_of(d, sel):
    rad = {n: np.ascontiguousarray(d[n][sel], dtype=np.float64) for n in sht.RLOOP_RAD}
    rad["nR"] = np.ascontiguousarray(d["levels"][sel], dtype=np.int32)
    rad["l_R"] = np.ascontiguousarray(d["l_R"][sel], dtype=np.int32)
    return rad


def fields_of(d, sel):
    return {n: np.ascontiguousarray(d[n][sel]) if n in d else None for n in sht.RLOOP_IN}


# outputs the reference writes at interior levels under the dump's switches
def outputs_of(d):
    names = ["dwdt", "dzdt"]
    if not d["l_double_curl"] or d["lPressNext"]:
        names.append("dpdt")
    if d["l_heat"]:
        names += ["dsdt", "dVSrLM"]
    if d["l_chemical_conv"]:
        names += ["dxidt", "dVXirLM"]
    if d["l_phase_field"]:
        names.append("dphidt")
    if d["l_mag"]:
        names += ["dbdt", "djdt", "dVxBhLM"]
    if d["l_double_curl"]:
        names.append("dVxVhLM")
    return names


def plan_args(d):
    return dict(l_max=int(d["l_max"]), n_theta=int(d["n_theta_max"]), n_phi=int(d["n_phi_max"]),
                m_max=int(d["m_max"]), minc=int(d["minc"]))
