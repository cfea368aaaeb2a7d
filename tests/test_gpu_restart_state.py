"""The nonlinear step on a real dynamo state: the reference's boussBenchSat
restart file (l_max=64, minc=4, 33 radial levels; read by
magic_amd.checkpoint) with radial derivatives from the Chebyshev
interpolant.  FP64 against the oracle (1e-12 relative L2 per output) and the
split form against FP64 (1e-6) on every level."""
import os

import numpy as np
import pytest

from conftest import rel_l2
from magic_amd import checkpoint, sht, workload
from nl_reference import oracle_step
from oracle import oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _inputs():
    h = checkpoint.read_checkpoint(os.path.join(GOLDEN, "boussBenchSat_checkpoint_end.start"))
    r = h["r"]
    d = lambda f: oracle.cheb_dr(f, r)  # noqa: E731
    w, z, s, b, aj = h["w"], h["z"], h["s"], h["b"], h["aj"]
    dw, dz, db, dj = d(w), d(z), d(b), d(aj)
    ins = [w, dw, d(dw), z, dz, s, b, db, d(db), aj, dj]
    rad = (1.0 / r ** 2, 1.0 / r ** 4, np.ones_like(r), 1.0 / (h["ek"] * h["prmag"]))
    return h, [np.ascontiguousarray(a) for a in ins], rad


def _run(p, inputs, rad, lcut):
    import torch
    dev = torch.device("cuda")
    nlev = inputs[0].shape[0]
    tin = [torch.from_numpy(x).to(dev) for x in inputs]
    tout = [torch.zeros((nlev, p.lm_max), dtype=torch.complex128, device=dev) for _ in range(10)]
    sht.nl_step(p, tin, tout, rad[0], rad[1], rad[2], rad[3], lcut)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in tout]


def test_nl_step_on_restart_state():
    h, ins, rad = _inputs()
    L, minc, nlev = h["l_max"], h["minc"], ins[0].shape[0]
    nt, nph = workload.grid_size(L, h["nalias"], minc)
    print("restart state nl_step", L, minc, nlev, nt, nph, flush=True)
    lcut = [L] * nlev
    p64 = sht.Plan(L, nt, nph, minc=minc, max_levels=nlev)
    out64 = _run(p64, ins, rad, lcut)
    p64.close()
    sel = [0, 1, 8, 16, 24, 31, 32]
    o = oracle.Plan(L, nt, nph, minc=minc)
    sub = [a[sel] for a in ins]
    _, ref = oracle_step(o, sub, *(x[sel] for x in rad[:3]), rad[3], [L] * len(sel))
    for k in range(10):
        e = rel_l2(out64[k][sel], ref[k])
        assert e < 1e-12, (sht.NL_OUTPUTS[k], e)
    p16 = sht.Plan(L, nt, nph, minc=minc, precision=sht.PREC_SPLIT16, max_levels=nlev)
    out16 = _run(p16, ins, rad, lcut)
    p16.close()
    errs = [rel_l2(out16[k], out64[k]) for k in range(10)]
    print("split vs fp64 on the restart state:", " ".join("%.2e" % e for e in errs), flush=True)
    for k, e in enumerate(errs):
        assert e < 1e-6, (sht.NL_OUTPUTS[k], e)
