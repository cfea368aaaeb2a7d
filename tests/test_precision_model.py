"""Where the split form's per-l-band error comes from (CPU model).

A scalar stand-in for u x B: coefficients a and b with l^-5/3 spectra, b's
l = 1 entries x100 (a dipole-dominated field), the product f = a b formed on
the Gauss grid and analysed back (scal_to_spat, FFTs, scal_to_SH of
src/sht_native.f90).  Each stage can be switched from FP64 to the form the
split step computes it in:

* synthesis / analysis: scaled FP16 x3 Legendre products -- one power-of-two
  scale per column (coefficient vector, or theta column of Sym/Anti Fourier
  values) and per table row, hi/lo FP16 halves, the products hi*hi, hi*lo,
  lo*hi accumulated in FP32 per chunk of 32 (one MFMA's K);
* the inverse and forward FFTs in FP32, the grid product in FP32.

The FP64 path is the oracle's to rounding.  Every reduced stage adds noise of
roughly its precision relative to the field, so a band's relative error grows
as the band's share of the field falls; the stages contribute comparably, and
the split stages' error is the same with FP16's exponent range lifted (the
significand, not the range, sets it)."""
import numpy as np
import pytest

from oracle import oracle
from magic_amd import workload

L = 127


def _f16(x):
    return x.astype(np.float16).astype(np.float64)


def _f11(x):  # 11 significant bits, no exponent limit
    m, e = np.frexp(x)
    return np.ldexp(np.round(m * 2048.0) / 2048.0, e)


def _p2(mx):  # power of two s with mx s in [2^14, 2^15)
    mx = np.where(mx > 0, mx, 1.0)
    e = np.clip(np.floor(np.log2(mx)) + 1, -100, 125)
    return 2.0 ** (15 - e)


class Model:
    def __init__(self):
        self.nt, self.N = workload.grid_size(L)
        self.o = oracle.Plan(L, self.nt, self.N)
        t = self.o.tables()
        self.plm, self.wplm = t["plm"], t["wplm"]
        self.nh = self.nt // 2
        self.ls, self.ms = workload.lm_index(L)
        self.par = (self.ls - self.ms) % 2
        self.half = _f16

    def split_dot(self, A, v):
        """A @ v in the split form: A (rows x K) row-scaled, v one column."""
        half = self.half
        if A.shape[1] == 0:
            return np.zeros(A.shape[0])
        rs = _p2(np.abs(A).max(axis=1))[:, None]
        As = A * rs
        Ah, Al = half(As), half(As - half(As))
        s = _p2(np.abs(v).max()) if np.abs(v).max() > 0 else 1.0
        vs = v * s
        vh, vl = half(vs), half(vs - half(vs))
        acc = np.zeros(A.shape[0], dtype=np.float32)
        for k0 in range(0, A.shape[1], 32):
            k = slice(k0, k0 + 32)
            acc = (acc + (Ah[:, k] @ vl[k] + Al[:, k] @ vh[k] + Ah[:, k] @ vh[k]).astype(np.float32)).astype(np.float32)
        return acc.astype(np.float64) / (rs[:, 0] * s)

    def syn(self, c, split):
        F = np.zeros((L + 1, self.nt), dtype=np.complex128)
        for m in range(L + 1):
            sel = self.ms == m
            P, x, pr = self.plm[:, sel], c[sel], self.par[sel]
            sa = []
            for pp in (0, 1):
                Pp, xp = P[:, pr == pp], x[pr == pp]
                sa.append(self.split_dot(Pp, xp.real) + 1j * self.split_dot(Pp, xp.imag) if split else Pp @ xp)
            F[m, 0::2], F[m, 1::2] = sa[0] + sa[1], sa[0] - sa[1]
        return F

    def grid(self, F, f32):
        X = np.zeros((self.N // 2 + 1, self.nt), dtype=np.complex128)
        X[:L + 1] = F
        X[0] = X[0].real
        if f32:
            return (np.fft.irfft(X.astype(np.complex64), n=self.N, axis=0) * self.N).astype(np.float32).astype(np.float64)
        return np.fft.irfft(X, n=self.N, axis=0) * self.N

    def fourier(self, g, f32):
        if f32:
            X = np.fft.rfft(g.astype(np.float32), axis=0) / np.float32(self.N)
            return X.astype(np.complex64).astype(np.complex128)[:L + 1]
        return (np.fft.rfft(g, axis=0) / self.N)[:L + 1]

    def ana(self, F, split):
        out = np.zeros(self.ls.size, dtype=np.complex128)
        for m in range(L + 1):
            sel = self.ms == m
            W, pr = self.wplm[:, sel], self.par[sel]
            Nn, Ss = F[m, 0::2], F[m, 1::2]
            res = np.zeros(sel.sum(), dtype=np.complex128)
            for pp in (0, 1):
                col = Nn + Ss if pp == 0 else Nn - Ss
                idx = np.where(pr == pp)[0]
                Wt = W[:, idx].T
                res[idx] = (self.split_dot(Wt, col.real) + 1j * self.split_dot(Wt, col.imag)) if split else Wt @ col
            out[sel] = res
        return out

    def product(self, a, b, s_syn=False, f_inv=False, f_prod=False, f_fwd=False, s_ana=False):
        ga, gb = self.grid(self.syn(a, s_syn), f_inv), self.grid(self.syn(b, s_syn), f_inv)
        g = (ga.astype(np.float32) * gb.astype(np.float32)).astype(np.float64) if f_prod else ga * gb
        return self.ana(self.fourier(g, f_fwd), s_ana)


@pytest.fixture(scope="module")
def model():
    M = Model()
    a = workload.coefficients(L, 9, 0, 0, kind="l^-5/3", potential=False)
    b = workload.coefficients(L, 9, 1, 0, kind="l^-5/3", potential=False)
    b[M.ls == 1] *= 100.0
    exact = M.product(a, b)

    def bands(x):
        r = []
        for b0 in range(0, L + 1, 32):
            s = (M.ls >= b0) & (M.ls < b0 + 32)
            r.append(np.linalg.norm(x[s] - exact[s]) / np.linalg.norm(exact[s]))
        return np.array(r)

    return M, a, b, exact, bands


def test_fp64_path_is_the_oracle(model):
    M, a, _, _, _ = model
    g = M.o.scal_to_spat(a, L)
    assert np.linalg.norm(M.grid(M.syn(a, False), False) - g) < 1e-13 * np.linalg.norm(g)
    c = M.o.scal_to_SH(g, L)
    assert np.linalg.norm(M.ana(M.fourier(g, False), False) - c) < 1e-13 * np.linalg.norm(c)


def test_every_reduced_stage_adds_a_band_error_of_its_own(model):
    M, a, b, exact, bands = model
    stages = {"split synthesis": dict(s_syn=True), "FP32 inverse FFT": dict(f_inv=True),
              "FP32 product": dict(f_prod=True), "FP32 forward FFT": dict(f_fwd=True),
              "split analysis": dict(s_ana=True)}
    top = {}
    for name, kw in stages.items():
        e = bands(M.product(a, b, **kw))
        print("%-18s bands %s" % (name, " ".join("%.1e" % x for x in e)))
        # noise of FP32 class: the field's low band near 1e-7, the bands'
        # errors rising as their share of the field falls
        assert e[0] < 3e-7 and e[-1] > 3 * e[0], (name, e)
        top[name] = e[-1]
    e_all = bands(M.product(a, b, s_syn=True, f_inv=True, f_prod=True, f_fwd=True, s_ana=True))
    print("all stages         bands %s" % " ".join("%.1e" % x for x in e_all))
    # comparable shares: none of the five is negligible or alone
    big = max(top.values())
    assert all(v > big / 10 for v in top.values()), top
    assert e_all[-1] > 1.2 * big, (e_all[-1], big)


def test_split_error_is_the_significand_not_the_fp16_range(model):
    M, a, b, exact, bands = model
    e16 = bands(M.product(a, b, s_syn=True, s_ana=True))
    M.half = _f11
    try:
        e11 = bands(M.product(a, b, s_syn=True, s_ana=True))
    finally:
        M.half = _f16
    print("FP16 halves", e16, "11-bit halves without range limit", e11)
    assert np.allclose(e16, e11, rtol=0.05)
