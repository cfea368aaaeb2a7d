"""Uneven batch of transform jobs (mhx_batch_*, split form): every job is the
round trip synthesis -> grids -> analysis of its fields, which for
band-limited inputs on the Gauss grid is the identity (scalar: s -> s;
vector: W, dW, Z -> Q = l(l+1) W, S = dW, T = Z for l >= 1,
src/sht_native.f90:114-123, :358-364).  Held to 1e-6 relative L2 per job
and component; a replay of the batch's graph reproduces it bit for bit."""
import numpy as np
import pytest

from conftest import rel_l2
from magic_amd import sht

pytestmark = pytest.mark.gpu

SPLIT_TOL = 1e-6


def lm_index(L):
    ls, ms = [], []
    for m in range(L + 1):
        for l in range(m, L + 1):
            ls.append(l)
            ms.append(m)
    return np.array(ls), np.array(ms)


def job_input(rng, L, ncomp, nfields, nlev):
    """Coefficients with amplitude (l+1)^-2, random phases, real at m = 0."""
    ls, ms = lm_index(L)
    shape = (ncomp, nfields, nlev, ls.size)
    amp = (ls + 1.0) ** -2
    x = amp * np.exp(2j * np.pi * rng.random(shape))
    x = np.where(ms == 0, x.real + 0j, x)
    return x.astype(np.complex128)


def expected(kind, L, x):
    ls, _ = lm_index(L)
    if kind == sht.JOB_SCALAR:
        return x, np.ones(ls.size, bool)
    y = x.copy()
    y[0] = ls * (ls + 1.0) * x[0]
    return y, ls >= 1


def test_batch_round_trips_and_replays():
    import torch
    rng = np.random.default_rng(3)
    dev = torch.device("cuda")
    specs = [(63, sht.JOB_SCALAR, 2, 5), (63, sht.JOB_VECTOR, 1, 7), (127, sht.JOB_VECTOR, 3, 3),
             (127, sht.JOB_SCALAR, 1, 4), (191, sht.JOB_SCALAR, 2, 3), (191, sht.JOB_VECTOR, 1, 5),
             (255, sht.JOB_VECTOR, 2, 3), (255, sht.JOB_SCALAR, 1, 6)]
    host, jobs = [], []
    for L, kind, nf, nlev in specs:
        nc = 3 if kind == sht.JOB_VECTOR else 1
        x = job_input(rng, L, nc, nf, nlev)
        host.append(x)
        jobs.append((L, kind, torch.from_numpy(x).to(dev), torch.zeros(x.shape, dtype=torch.complex128,
                                                                        device=dev)))
    b = sht.Batch(jobs)
    tuning = b.tuning()
    assert sorted(t[0] for t in tuning) == [63, 127, 191, 255]
    for _, syn, ana, rows in tuning:
        assert syn in (1, 2) and ana in (1, 2) and rows in (1, 2, 4, 8)
    b.execute()
    first = [j[3].cpu().numpy() for j in jobs]
    for (L, kind, nf, nlev), x, got in zip(specs, host, first):
        want, mask = expected(kind, L, x)
        for c in range(x.shape[0]):
            sel = mask if c > 0 else (mask if kind == sht.JOB_VECTOR else np.ones_like(mask))
            e = rel_l2(got[c][..., sel], want[c][..., sel])
            print("l_max=%d kind=%d comp=%d: %.2e" % (L, kind, c, e), flush=True)
            assert e < SPLIT_TOL, (L, kind, c, e)
    for j in jobs:
        j[3].zero_()
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    b.execute(st)
    torch.cuda.synchronize()
    for j, f in zip(jobs, first):
        assert np.array_equal(j[3].cpu().numpy(), f)
    b.close()
