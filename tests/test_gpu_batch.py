"""Uneven batch of transform jobs (mhx_batch_*, split form): every job is the
round trip synthesis -> grids -> analysis of its fields, which for
band-limited inputs on the Gauss grid is the identity (scalar: s -> s;
vector: W, dW, Z -> Q = l(l+1) W, S = dW, T = Z for l >= 1,
src/sht_native.f90:114-123, :358-364).  Held to 1e-6 relative L2 per job
and component, in the persistent launch (default) and the per-job graph
form; a replay of the batch's graph reproduces it bit for bit, and the two
forms agree to the rounding of their kernel choices."""
import os

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
        assert syn in (1, 2) and ana in (1, 2) and rows in (1, 2, 4, 8, 16, 32)
    form, shapes = b.shapes()
    print("form %d, shapes %s" % (form, shapes), flush=True)
    assert form == 1
    for L, orders, groups, groups_hi, m_split, rows in shapes:
        assert orders in (1, 2) and groups in (1, 2, 4) and groups_hi in (1, 2, 4) and rows in (4, 8, 16, 32)
        assert groups_hi % groups == 0 and 0 < m_split <= L + 1
    b.execute()
    b.status()
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
    b.status()
    for j, f in zip(jobs, first):
        assert np.array_equal(j[3].cpu().numpy(), f)
    b.close()


def test_batch_persistent_launch_matches_graph_form():
    """The same jobs through the one persistent launch and through the
    per-job launches (MHX_BATCH_FORM=graph): within 4e-7 of each other per
    job and component, and both within 1e-6 of the identity.  The forms
    round differently: wave-slot row FFTs (radix 12 first) against the
    generic LDS passes, and the graph form may pick the chunked Legendre
    kernels for a group; each form lies within about 2e-7 of the identity."""
    import torch
    rng = np.random.default_rng(11)
    dev = torch.device("cuda")
    specs = [(63, sht.JOB_VECTOR, 4, 17), (127, sht.JOB_SCALAR, 3, 9), (191, sht.JOB_VECTOR, 2, 5),
             (255, sht.JOB_VECTOR, 1, 11), (255, sht.JOB_SCALAR, 2, 3)]
    host = []
    for L, kind, nf, nlev in specs:
        nc = 3 if kind == sht.JOB_VECTOR else 1
        host.append(job_input(rng, L, nc, nf, nlev))
    outs = {}
    for name in ("graph", "persistent"):
        jobs = [(L, kind, torch.from_numpy(x).to(dev), torch.zeros(x.shape, dtype=torch.complex128, device=dev))
                for (L, kind, nf, nlev), x in zip(specs, host)]
        old = os.environ.pop("MHX_BATCH_FORM", None)
        if name == "graph":
            os.environ["MHX_BATCH_FORM"] = "graph"
        try:
            b = sht.Batch(jobs)
        finally:
            os.environ.pop("MHX_BATCH_FORM", None)
            if old is not None:
                os.environ["MHX_BATCH_FORM"] = old
        assert b.shapes()[0] == (0 if name == "graph" else 1)
        b.execute()
        b.status()
        outs[name] = [j[3].cpu().numpy() for j in jobs]
        b.close()
    for (L, kind, nf, nlev), x, g, p in zip(specs, host, outs["graph"], outs["persistent"]):
        want, mask = expected(kind, L, x)
        for c in range(x.shape[0]):
            sel = mask if c > 0 else (mask if kind == sht.JOB_VECTOR else np.ones_like(mask))
            d = rel_l2(p[c][..., sel], g[c][..., sel])
            e = rel_l2(p[c][..., sel], want[c][..., sel])
            print("l_max=%d kind=%d comp=%d: persistent vs graph %.2e, vs identity %.2e" % (L, kind, c, d, e),
                  flush=True)
            assert d < 4e-7, (L, kind, c, d)
            assert e < SPLIT_TOL, (L, kind, c, e)


def test_batch_small_and_ragged_jobs():
    """Small truncations (l_max 7, 15, 23, 31: n_phi 24 .. 96, row FFTs on
    the generic LDS passes rather than wave slots; n_theta = 3 (l_max + 1) / 2
    must be a multiple of 4), single levels and fields, a level group cut
    short (rows not a multiple of 8), a job of two whole level groups:
    persistent launch against the identity and against the per-job graph
    form."""
    import torch
    rng = np.random.default_rng(17)
    dev = torch.device("cuda")
    specs = [(7, sht.JOB_SCALAR, 1, 1), (15, sht.JOB_VECTOR, 1, 2), (23, sht.JOB_VECTOR, 2, 1),
             (31, sht.JOB_SCALAR, 3, 3), (31, sht.JOB_VECTOR, 1, 16), (63, sht.JOB_SCALAR, 1, 1)]
    host = []
    for L, kind, nf, nlev in specs:
        nc = 3 if kind == sht.JOB_VECTOR else 1
        host.append(job_input(rng, L, nc, nf, nlev))
    outs = {}
    for name in ("graph", "persistent"):
        jobs = [(L, kind, torch.from_numpy(x).to(dev), torch.zeros(x.shape, dtype=torch.complex128, device=dev))
                for (L, kind, nf, nlev), x in zip(specs, host)]
        old = os.environ.pop("MHX_BATCH_FORM", None)
        if name == "graph":
            os.environ["MHX_BATCH_FORM"] = "graph"
        try:
            b = sht.Batch(jobs)
        finally:
            os.environ.pop("MHX_BATCH_FORM", None)
            if old is not None:
                os.environ["MHX_BATCH_FORM"] = old
        form, shapes = b.shapes()
        print(name, "form", form, "shapes", shapes, flush=True)
        assert form == (0 if name == "graph" else 1)
        b.execute()
        b.execute()
        b.status()
        outs[name] = [j[3].cpu().numpy() for j in jobs]
        b.close()
    for (L, kind, nf, nlev), x, g, p in zip(specs, host, outs["graph"], outs["persistent"]):
        want, mask = expected(kind, L, x)
        for c in range(x.shape[0]):
            sel = mask if c > 0 else (mask if kind == sht.JOB_VECTOR else np.ones_like(mask))
            d = rel_l2(p[c][..., sel], g[c][..., sel])
            e = rel_l2(p[c][..., sel], want[c][..., sel])
            print("l_max=%d kind=%d nf=%d nlev=%d comp=%d: persistent vs graph %.2e, vs identity %.2e"
                  % (L, kind, nf, nlev, c, d, e), flush=True)
            assert d < 4e-7, (L, kind, c, d)
            assert e < SPLIT_TOL, (L, kind, c, e)
