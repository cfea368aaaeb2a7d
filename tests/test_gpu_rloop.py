"""The bulk levels of the radial loop (magic_amd.rloop.RadialLoop:
mhx_nl_step_courant then mhx_get_td, src/rIter.f90:188-470 with nBc = 0) on
the interior levels of the reference's boussBenchSat restart state against
the reference's own chain of nl_level, courant and get_td (ref_harness
rloop golden): FP64 to 1e-12 (explicit terms, relative L2 per output;
Courant steps, relative).  The split form: explicit terms to 2e-5, Courant
steps to 1e-5.  Its forward outputs meet 1e-6 on this state
(test_gpu_restart_state.py: 1.5-3.4e-7); the explicit terms multiply them by
l(l+1) (dzdt, dpdt, dsdt, dbdt) or l(l+1)/r^4 (djdt), which weights the
high degrees, where the split form's relative error is largest (measured:
5e-8 to 7.6e-6)."""
import numpy as np
import pytest

from conftest import load_golden, rel_l2
from magic_amd import rloop, sht, workload
import restart_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol,tol_dt", [(sht.PREC_FP64, 1e-12, 1e-12),
                                                  (sht.PREC_SPLIT16, 2e-5, 1e-5)])
def test_radial_loop_bulk_matches_reference(precision, tol, tol_dt):
    g = load_golden("rloop_boussBenchSat.npz")
    h, ins, P = restart_state.load(g["levels"])
    nlev = ins[0].shape[0]
    nt, nph = workload.grid_size(h["l_max"], h["nalias"], h["minc"])
    print("radial loop bulk, precision", precision, nlev, flush=True)
    p = sht.Plan(h["l_max"], nt, nph, minc=h["minc"], precision=precision, max_levels=nlev)
    rl = rloop.RadialLoop(p, nlev)
    for rep in range(2):  # the second pass replays the step's graph
        td, _, dtr, dth = rl.bulk(ins, P["or2"], P["or4"], P["orho1"], P["LFfac"], P["CorFac"],
                                  P["epsc"], P["lcut"], P["courfac"], P["alffac"], P["opm"],
                                  P["delxr2"], P["delxh2"])
        errs = [rel_l2(td[k].cpu().numpy(), g["td"][k]) for k in range(7)]
        print("  td errors:", " ".join("%.2e" % e for e in errs), flush=True)
        for k, e in enumerate(errs):
            assert e < tol, (sht.TD_OUTPUTS[k], rep, e)
        np.testing.assert_allclose(dtr, g["dtrkc"], rtol=tol_dt, atol=0)
        np.testing.assert_allclose(dth, g["dthkc"], rtol=tol_dt, atol=0)
    p.close()


def test_radial_loop_device_fields_and_shape_check():
    """Device-tensor fields give the same result as host arrays; a field of
    the wrong shape is refused before any launch."""
    import torch
    g = load_golden("rloop_boussBenchSat.npz")
    h, ins, P = restart_state.load(g["levels"][:6])
    nlev = ins[0].shape[0]
    nt, nph = workload.grid_size(h["l_max"], h["nalias"], h["minc"])
    print("radial loop device fields", nlev, flush=True)
    p = sht.Plan(h["l_max"], nt, nph, minc=h["minc"], max_levels=nlev)
    rl = rloop.RadialLoop(p, nlev)
    args = (P["or2"], P["or4"], P["orho1"], P["LFfac"], P["CorFac"], P["epsc"], P["lcut"],
            P["courfac"], P["alffac"], P["opm"], P["delxr2"], P["delxh2"])
    td_h, _, dtr_h, dth_h = rl.bulk(ins, *args)
    td_h = [t.cpu().numpy() for t in td_h]
    dev = [torch.from_numpy(a).cuda() for a in ins]
    td_d, _, dtr_d, dth_d = rl.bulk(dev, *args)
    for k in range(7):
        np.testing.assert_array_equal(td_d[k].cpu().numpy(), td_h[k])
    np.testing.assert_array_equal(dtr_d, dtr_h)
    np.testing.assert_array_equal(dth_d, dth_h)
    for k in range(7):
        assert rel_l2(td_h[k], g["td"][k][:6]) < 1e-12
    with pytest.raises(ValueError):
        rl.bulk([a[:, :-1] for a in ins], *args)
    p.close()
