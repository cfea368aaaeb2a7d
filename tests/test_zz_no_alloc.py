"""The allocation-free kinetic-energy entry (mhx_get_e_kin_noalloc) against
the allocating one (mhx_get_e_kin): the same kernels over the same device
arrays, so the sums must agree bit for bit.  Two calls in a row, the second
with fewer levels and other data (the fixed scratch array keeps the first
call's partial sums, none of which the second may read), then the refusals:
more levels than the fixed array holds, and host arrays."""
import ctypes

import numpy as np
import pytest

from magic_amd import sht, workload

pytestmark = pytest.mark.gpu


def _fields(rng, L, nlev, dev):
    import torch
    ls, ms = workload.lm_index(L)
    out = []
    for _ in range(3):
        x = (rng.standard_normal((nlev, ls.size)) + 1j * rng.standard_normal((nlev, ls.size)))
        x /= (ls + 1.0) ** 1.5
        x[:, ms == 0] = x[:, ms == 0].real
        out.append(torch.from_numpy(np.ascontiguousarray(x)).to(dev))
    return out


def test_noalloc_e_kin_equals_allocating_entry():
    import torch
    dev = torch.device("cuda:0")
    L = 255
    nt, nph = workload.grid_size(L)
    p = sht.Plan(L, nt, nph, max_levels=1)
    rng = np.random.default_rng(11)
    st = torch.cuda.Stream()
    for nlev in (97, 33):
        print("noalloc e_kin: l_max=%d, %d levels" % (L, nlev), flush=True)
        w, dw, z = _fields(rng, L, nlev, dev)
        or2 = rng.uniform(0.5, 3.0, nlev)
        orho1 = rng.uniform(0.8, 1.2, nlev)
        want = sht.get_e_kin_levels(p, w, dw, z, or2, orho1)
        sums = torch.full((8, nlev), float("nan"), dtype=torch.float64, device=dev)
        st.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(st):
            sht.get_e_kin_noalloc(p, w, dw, z, torch.from_numpy(or2).to(dev),
                                  torch.from_numpy(orho1).to(dev), sums, stream=st)
        st.synchronize()
        got = sums.cpu().numpy()
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, want), np.max(np.abs(got - want) / np.abs(want))
    p.close()


def test_noalloc_e_kin_refusals():
    import torch
    dev = torch.device("cuda:0")
    L = 16
    nt, nph = workload.grid_size(L)
    p = sht.Plan(L, nt, nph, max_levels=1)
    rng = np.random.default_rng(12)
    # more levels than the fixed array holds
    nlev = 2049
    w, dw, z = _fields(rng, L, nlev, dev)
    r = torch.ones(nlev, dtype=torch.float64, device=dev)
    sums = torch.zeros((8, nlev), dtype=torch.float64, device=dev)
    with pytest.raises(sht.MhxError, match="2048 levels"):
        sht.get_e_kin_noalloc(p, w, dw, z, r, r, sums)
    # a host array
    nlev = 3
    w, dw, z = _fields(rng, L, nlev, dev)
    r = torch.ones(nlev, dtype=torch.float64, device=dev)
    sums = torch.zeros((8, nlev), dtype=torch.float64, device=dev)
    host = np.zeros((nlev, p.lm_max), dtype=np.complex128)
    rc = p._L.mhx_get_e_kin_noalloc(p.h, nlev, ctypes.c_void_p(host.ctypes.data),
                                    ctypes.c_void_p(dw.data_ptr()), ctypes.c_void_p(z.data_ptr()),
                                    p.lm_max, ctypes.c_void_p(r.data_ptr()),
                                    ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(sums.data_ptr()),
                                    None)
    assert rc != 0
    assert "device arrays only" in sht.lib().mhx_last_error().decode()
    torch.cuda.synchronize()
    assert float(sums.abs().sum()) == 0.0
    p.close()
