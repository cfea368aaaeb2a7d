"""Synthetic spectral inputs and grid sizing.

Coefficients are c_lm = A_l exp(i phi) with phi uniform in [0, 2 pi) drawn
from a SplitMix64 stream keyed by (seed, field, level); m = 0 keeps the real
part only; l = 0 is zero for poloidal/toroidal potentials.  The generator is
portable integer arithmetic so that C/C++/Python writers produce identical
inputs.

Grid sizing follows the reference's truncation rules
(src/truncation.f90:66-71, prime_decomposition :162-192): n_theta =
30 l_max / nalias, n_phi = 2 n_theta rounded up to a 2^i 3^j 5^k multiple of
minc, then n_theta = n_phi_tot / 2.
"""
import concurrent.futures
import functools
import os

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
SEED_BASE = 0x4D41474943000000


def splitmix64(seed, n):
    """n outputs of SplitMix64 started at ``seed`` (uint64 array)."""
    with np.errstate(over="ignore"):
        s = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = s + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n):
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def stream_seed(stream, field, level):
    return (SEED_BASE + stream) ^ ((field + 1) * 0x100000001B3) ^ ((level + 1) * 0xC2B2AE3D27D4EB4F)


def lm_index(l_max, m_max=None, minc=1):
    """(l, m) of every lm in the standard m-major order (src/blocking.f90:309-317)."""
    m_max = l_max if m_max is None else min(m_max, l_max)
    ls, ms = _lm_index(l_max, m_max, minc)
    return ls.copy(), ms.copy()


@functools.lru_cache(maxsize=16)
def _lm_index(l_max, m_max, minc):
    m = np.arange(0, m_max + 1, minc, dtype=np.int64)
    cnt = l_max - m + 1
    ms = np.repeat(m, cnt)
    start = np.repeat(np.cumsum(cnt) - cnt, cnt)
    ls = ms + np.arange(ms.size, dtype=np.int64) - start
    return ls, ms


def spectrum(l, kind):
    l = np.asarray(l, dtype=np.float64)
    if kind == "l+1^-2":
        return (l + 1.0) ** -2
    if kind == "l^-5/3":
        return np.maximum(l, 1.0) ** (-5.0 / 3.0)
    if kind == "white-to-l^-3":
        # flat for l << 64, l^-3 for l >> 64
        return (1.0 + (l / 64.0) ** 2) ** -1.5
    if kind.startswith("l+1^-"):
        return (l + 1.0) ** (-float(kind[5:]))
    raise ValueError(kind)


def coefficients(l_max, stream, field, level, kind="l+1^-2", potential=True, m_max=None,
                 minc=1, amp=1.0):
    ls, ms = lm_index(l_max, m_max, minc)
    ph = 2.0 * np.pi * uniform01(stream_seed(stream, field, level), ls.size)
    c = amp * spectrum(ls, kind) * np.exp(1j * ph)
    c[ms == 0] = c[ms == 0].real
    if potential:
        c[ls == 0] = 0.0
    return c.astype(np.complex128)


# potential -> field l-factor of each input of the nonlinear step (w dw ddw z
# dz s b db ddb aj dj): v_r ~ l(l+1) W, the horizontal components ~ l dW,
# l Z, ..., the temperature itself
NL_FIELD_ORDER = ("w", "dw", "ddw", "z", "dz", "s", "b", "db", "ddb", "aj", "dj")


def nl_inputs(l_max, nlev, stream=2, kind="l^-5/3", minc=1, dipole=1.0):
    """Coefficient inputs of the nonlinear step whose synthesised fields
    (velocity, vorticity, magnetic field, current density, temperature) have
    amplitude spectrum `kind` with random phases: each potential is drawn with
    that spectrum and divided by the l-factor its transform applies (l(l+1)
    for the poloidal potentials w and b, l for the others, 1 for s).
    `dipole` multiplies l = 1 of the poloidal magnetic potential b and its
    radial derivatives db, ddb (a dipole-dominated field)."""
    ls, _ = lm_index(l_max, None, minc)
    lsafe = np.maximum(ls, 1).astype(np.float64)
    out = []
    for f, name in enumerate(NL_FIELD_ORDER):
        # "slopes-0-3": field f has amplitude (l+1)^-s with s = f mod 4
        kf = "l+1^-%d" % (f % 4) if kind == "slopes-0-3" else kind
        if name in ("w", "b"):
            fac = 1.0 / (lsafe * (lsafe + 1.0))
        elif name == "s":
            fac = np.ones_like(lsafe)
        else:
            fac = 1.0 / lsafe
        if dipole != 1.0 and name in ("b", "db", "ddb"):
            fac = fac * np.where(ls == 1, dipole, 1.0)
        arr = np.empty((nlev, ls.size), dtype=np.complex128)

        def fill(lev, f=f, kf=kf, fac=fac, name=name, arr=arr):
            arr[lev] = coefficients(l_max, stream, f, lev, kf, potential=(name != "s"),
                                    minc=minc) * fac
        # levels are independent streams: filled by a thread pool (numpy's
        # array operations release the interpreter lock)
        if nlev * ls.size >= 1 << 20:
            with concurrent.futures.ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
                list(ex.map(fill, range(nlev)))
        else:
            for lev in range(nlev):
                fill(lev)
        out.append(arr)
    return out


def prime_decomposition(n):
    """Smallest m >= n that factors as 2^i 3^j 5^k (src/truncation.f90:162-192)."""
    m = n
    while True:
        k = m
        for f in (2, 3, 5):
            while k % f == 0:
                k //= f
        if k == 1:
            return m
        m += 1


def grid_size(l_max, nalias=20, minc=1):
    """(n_theta, n_phi) the reference derives from l_max (src/truncation.f90:66-71),
    with the consistency checks of checkTruncation (:126-160)."""
    n_theta = (30 * l_max) // nalias
    n_phi_tot = prime_decomposition(2 * n_theta)
    n_theta, n_phi = n_phi_tot // 2, n_phi_tot // minc
    if n_phi_tot % minc or n_phi % 4 or n_phi_tot % 16 or n_theta % 4:
        raise ValueError("l_max=%d, minc=%d gives an invalid grid (%d x %d)"
                         % (l_max, minc, n_theta, n_phi))
    return n_theta, n_phi


def batch_specs(njobs=64, seed=3):
    """The uneven batch of the batched-transform workload: per job (l_max,
    kind, nfields, nlev) with l_max from {63, 127, 191, 255}, n_r uniform in
    17..65, scalar or vector with equal odds, 1..4 fields."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(njobs):
        L = int(rng.choice([63, 127, 191, 255]))
        kind = int(rng.integers(0, 2))
        out.append((L, kind, int(rng.integers(1, 5)), int(rng.integers(17, 66))))
    return out


def batch_coeffs(rng, L, ncomp, nfields, nlev):
    """(ncomp, nfields, nlev, lm_max) complex coefficients, amplitude
    (l+1)^-2, random phases, real at m = 0 (m-major, l fastest)."""
    ls, ms = [], []
    for m in range(L + 1):
        for l in range(m, L + 1):
            ls.append(l)
            ms.append(m)
    ls, ms = np.array(ls), np.array(ms)
    x = (ls + 1.0) ** -2 * np.exp(2j * np.pi * rng.random((ncomp, nfields, nlev, ls.size)))
    return np.where(ms == 0, x.real + 0j, x).astype(np.complex128)
