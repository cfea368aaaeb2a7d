"""The bulk levels of the reference's radial loop on the GPU.

``rIter_t%radialLoop`` (contract: src/rIteration.f90:34-79; body of the
single-level iterator: src/rIter.f90:188-470) visits every radial level and,
for a level away from the boundaries (nBc = 0), transforms the fields to the
grid, forms the nonlinear products (get_nl), takes the Courant limits on the
same grids (src/courant.f90), transforms the products back and assembles the
explicit terms (src/get_td.f90).  ``RadialLoop.bulk`` does that for all
levels at once with two calls of the C ABI (``mhx_nl_step_courant``, then
``mhx_get_td``), keeping every intermediate on the device.  The boundary
levels, the time scheme and the implicit solves stay with the caller.
"""
import numpy as np

from . import sht

_TD_FROM_NL = (0, 1, 2, 4, 6, 7, 8)  # AdvrLM AdvtLM AdvpLM VStLM VxBrLM VxBtLM VxBpLM


class RadialLoop:
    """Device buffers for ``nlev`` levels of one plan (its precision picks
    the FP64 or split nonlinear step; the lm-space assembly is FP64)."""

    def __init__(self, plan, nlev):
        import torch
        self.plan, self.nlev = plan, nlev
        dev = torch.device("cuda")
        z = lambda: torch.zeros((nlev, plan.lm_max), dtype=torch.complex128, device=dev)  # noqa: E731
        self.fields = [z() for _ in range(11)]
        self.nl = [z() for _ in range(10)]
        self.td = [z() for _ in range(7)]

    def bulk(self, fields, or2, or4, orho1, LFfac, CorFac, epsc, lcut, courfac, alffac,
             opm, delxr2, delxh2):
        """One pass over the levels.  ``fields``: 11 arrays (nlev, lm_max) in
        ``sht.NL_INPUTS`` order (numpy or device tensors); per level ``or2``,
        ``or4``, ``orho1``, ``epsc`` (times epscProf), ``lcut`` (l_R),
        ``delxr2``, ``delxh2``; scalars as the reference's namelists give them.
        Returns (td, nl, dtrkc, dthkc): the 7 explicit terms in
        ``sht.TD_OUTPUTS`` order and the 10 forward outputs (device tensors,
        valid until the next call) and the per-level Courant steps, 1e10
        where the level has no flow (src/rIter.f90:214-215, src/courant.f90:
        271-272)."""
        import torch
        n = self.nlev
        for dst, src in zip(self.fields, fields):
            if tuple(src.shape) != (n, self.plan.lm_max):
                raise ValueError("field shape %r, expected %r" % (tuple(src.shape),
                                                                   (n, self.plan.lm_max)))
            dst.copy_(src if isinstance(src, torch.Tensor) else torch.from_numpy(np.asarray(src)))
        delxr2 = np.asarray(delxr2, dtype=np.float64)
        delxh2 = np.asarray(delxh2, dtype=np.float64)
        # src/courant.f90:211-212 (magnetic run)
        valri2 = (0.5 * (1.0 + opm)) ** 2 / delxr2
        valhi2 = (0.5 * (1.0 + opm)) ** 2 / delxh2
        vr2, vh2 = sht.nl_step_courant(self.plan, self.fields, self.nl, or2, or4, orho1, LFfac,
                                       lcut, courfac, alffac, valri2, valhi2)
        td_in = [self.nl[k] for k in _TD_FROM_NL] + [self.fields[0], self.fields[1],
                                                     self.fields[3]]
        sht.get_td(self.plan, td_in, self.td, or2, or4, epsc, CorFac, lcut)
        with np.errstate(divide="ignore"):
            dtrkc = np.where(vr2 != 0.0, np.minimum(1e10, np.sqrt(delxr2 / vr2)), 1e10)
            dthkc = np.where(vh2 != 0.0, np.minimum(1e10, np.sqrt(delxh2 / vh2)), 1e10)
        return self.td, self.nl, dtrkc, dthkc
