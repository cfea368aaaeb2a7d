"""Reader of the reference's restart files (checkpoint_*.start).

The file is a little-endian Fortran stream written by ``store``
(src/storeCheckPoints.f90:45-277) and read back by ``readStartFields``
(src/readCheckPoints.f90:767-1468); the record order below follows those two
routines for the stream versions 2 to 5.  Coefficient arrays are complex
``(lm_max, n_r_max)`` in Fortran order with the m-major standard lm order
(src/readCheckPoints.f90:2522-2527), returned here as ``(n_r_max, lm_max)``
C-order arrays: one row per radial level, the layout of the nonlinear step's
inputs.  The radial grid ``r`` runs from the outer boundary (index 0) to the
inner one.  Only reading is provided; mapping onto another truncation is the
caller's business.
"""
import numpy as np

FIELDS_OC = ("w", "z", "p", "s", "xi", "phi", "b", "aj")


class _Stream:
    def __init__(self, data):
        self.b = data
        self.o = 0

    def take(self, dtype, n=1):
        dt = np.dtype(dtype)
        nbytes = dt.itemsize * n
        if self.o + nbytes > len(self.b):
            raise ValueError("checkpoint truncated at byte %d (need %d more)" % (self.o, nbytes))
        a = np.frombuffer(self.b, dtype=dt, count=n, offset=self.o)
        self.o += nbytes
        return a

    def i4(self, n=1):
        a = self.take("<i4", n)
        return int(a[0]) if n == 1 else a.astype(np.int64)

    def f8(self, n=1):
        a = self.take("<f8", n)
        return float(a[0]) if n == 1 else a.copy()

    def chars(self, n):
        s = bytes(self.take("u1", n)).decode("ascii", "replace")
        return s.strip()

    def logical(self, n):
        return [bool(x) for x in self.take("<i4", n)]

    def field(self, lm_max, n_r):
        a = self.take("<c16", lm_max * n_r)
        return a.reshape(n_r, lm_max).copy()


def lm_count(l_max, m_min, m_max, minc):
    return sum(l_max - m + 1 for m in range(m_min, m_max + 1, minc))


def read_checkpoint(path, with_history=False):
    """Parses a restart file.  Returns a dict with the header (version, time,
    family, nexp/nimp/nold, dt, n_time_step, the physical parameters, the
    grid sizes, l_max/m_min/m_max, the radial scheme, r, the rotation
    rates, the switches) and one ``(n_r_max, lm_max)`` complex array per
    stored field (w z p s xi phi b aj; b_ic aj_ic with ``n_r_ic_max`` rows).
    ``with_history`` keeps the multistep scheme's stored explicit /
    implicit / old terms under ``"<field>_hist"`` (list of arrays)."""
    with open(path, "rb") as f:
        S = _Stream(f.read())
    h = {}
    version = S.i4()
    if version < 2 or version > 5:
        raise ValueError("unsupported checkpoint version %d" % version)
    h["version"] = version
    h["time"] = S.f8()
    fam = S.chars(10)
    nexp, nimp, nold = S.i4(), S.i4(), S.i4()
    h.update(family=fam, nexp=nexp, nimp=nimp, nold=nold)
    if fam == "MULTISTEP":
        h["dt"] = np.atleast_1d(S.f8(nexp))
    elif fam == "DIRK":
        h["dt"] = np.atleast_1d(S.f8(1))
    else:
        raise ValueError("unknown time-scheme family %r" % fam)
    h["n_time_step"] = S.i4()
    names = ["ra", "pr", "raxi", "sc", "prmag", "ek"] + (["stef"] if version > 2 else []) + \
            ["radratio", "sigma_ratio"]
    h.update(zip(names, S.f8(len(names))))
    h.setdefault("stef", 0.0)
    n_r_max, n_theta_max, n_phi_tot, minc, nalias, n_r_ic_max = (int(x) for x in S.i4(6))
    h.update(n_r_max=n_r_max, n_theta_max=n_theta_max, n_phi_tot=n_phi_tot, minc=minc,
             nalias=nalias, n_r_ic_max=n_r_ic_max)
    if version > 3:
        l_max, m_min, m_max = (int(x) for x in S.i4(3))
    else:
        l_max = nalias * n_phi_tot // 60 if n_phi_tot > 1 else nalias * n_theta_max // 30
        m_min, m_max = 0, (l_max // minc) * minc if n_phi_tot > 1 else 0
    h.update(l_max=l_max, m_min=m_min, m_max=m_max)
    h["rscheme"] = S.chars(72)
    h["n_in"], h["n_in_2"] = S.i4(), S.i4()
    h["ratio1"], h["ratio2"] = S.f8(), S.f8()
    h["r"] = S.f8(n_r_max)
    # rotation-rate histories (multistep only), then the rotation parameters
    nscal = 2 if version >= 5 else 4
    if fam == "MULTISTEP":
        S.f8(nscal * ((nexp - 1) + (nimp - 1) + (nold - 1)))
    h["omega"] = S.f8(12)
    if version == 2:
        sw = S.logical(5)
        l_heat, l_chem, l_mag, l_press, l_cond_ic = sw
        l_phase = False
    else:
        l_heat, l_chem, l_phase, l_mag, l_press, l_cond_ic = S.logical(6)
    h.update(l_heat=l_heat, l_chemical_conv=l_chem, l_phase_field=l_phase, l_mag=l_mag,
             l_press_store=l_press, l_cond_ic=l_cond_ic)
    lm_max = lm_count(l_max, m_min, m_max, minc)
    h["lm_max"] = lm_max
    nhist = (nexp - 1) + (nimp - 1) + (nold - 1) if fam == "MULTISTEP" else 0

    def one(name, n_r):
        h[name] = S.field(lm_max, n_r)
        hist = [S.field(lm_max, n_r) for _ in range(nhist)]
        if with_history:
            h[name + "_hist"] = hist

    present = dict(w=True, z=True, p=l_press, s=l_heat, xi=l_chem, phi=l_phase, b=l_mag,
                   aj=l_mag)
    for name in FIELDS_OC:
        if present[name]:
            one(name, n_r_max)
    if l_mag and l_cond_ic:
        one("b_ic", n_r_ic_max)
        one("aj_ic", n_r_ic_max)
    if S.o != len(S.b):
        raise ValueError("checkpoint has %d trailing bytes" % (len(S.b) - S.o))
    return h
