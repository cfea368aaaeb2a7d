"""Inputs of the nonlinear step and the radial loop built from the
reference's boussBenchSat restart file (a saturated Boussinesq dynamo,
l_max=64, minc=4, 33 Chebyshev levels): the 11 coefficient fields with their
radial derivatives from the Chebyshev interpolant (oracle.cheb_dr) and the
run's physical parameters (samples/boussBenchSat/input.nml: E = 1e-3,
Pm = 5, Boussinesq)."""
import os

import numpy as np

from magic_amd import checkpoint
from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CKPT = os.path.join(GOLDEN, "boussBenchSat_checkpoint_end.start")


def load(levels=None):
    """(header, 11 fields (nlev, lm_max), params) at the given level indices
    (default: all)."""
    h = checkpoint.read_checkpoint(CKPT)
    r = h["r"]
    d = lambda f: oracle.cheb_dr(f, r)  # noqa: E731
    w, z, s, b, aj = h["w"], h["z"], h["s"], h["b"], h["aj"]
    dw, dz, db, dj = d(w), d(z), d(b), d(aj)
    ins = [w, dw, d(dw), z, dz, s, b, db, d(db), aj, dj]
    sel = np.arange(r.size) if levels is None else np.asarray(levels)
    ins = [np.ascontiguousarray(a[sel]) for a in ins]
    rs = r[sel]
    L = h["l_max"]
    # radial spacing as the smaller neighbour gap, horizontal r^2/(l(l+1))
    gap = np.abs(np.diff(r))
    gmin = np.minimum(np.concatenate([[gap[0]], gap]), np.concatenate([gap, [gap[-1]]]))
    params = dict(r=rs, or2=1.0 / rs ** 2, or4=1.0 / rs ** 4, orho1=np.ones_like(rs),
                  LFfac=1.0 / (h["ek"] * h["prmag"]), CorFac=2.0 / h["ek"],
                  epsc=np.zeros_like(rs), lcut=[L] * rs.size, courfac=2.5, alffac=1.0,
                  opm=1.0 / h["prmag"], delxr2=gmin[sel] ** 2,
                  delxh2=rs ** 2 / (L * (L + 1.0)))
    return h, ins, params
