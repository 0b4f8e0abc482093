"""The reference's calling surface over the HIP kernels.

``getSinograms`` / ``getCircusFunctions`` keep the argument meaning of
c++/src/sinogram.hpp:57-59 and c++/src/circus.hpp:69-71: matrices are
[rows, cols] tensors with the reference's (row, col) indexing, a sinogram is
[n, A] (row = column of the rotated image, column = angle) and a circus
function is a length-A vector.  ``Transformer`` mirrors
c++/src/transform.hpp:24-38 without file output.
"""
from __future__ import annotations

import math

import torch

from . import core
from .functionals import PFunctional, pfuncs_of, tcodes_of


def _as_device_matrix(m, device=None) -> torch.Tensor:
    if not isinstance(m, torch.Tensor):
        m = torch.as_tensor(m)
    if device is None:
        device = m.device if m.is_cuda else torch.device("cuda", torch.cuda.current_device())
    return m.to(device=device, dtype=torch.float32)


def getSinograms(input, angle_stepsize, tfunctionals):
    """c++/src/sinogram.cpp:67-178: one [n, A] sinogram per T-functional of
    the padded square ``input``, A = floor(360 / angle_stepsize)."""
    img = _as_device_matrix(input)
    if img.dim() != 2 or img.shape[0] != img.shape[1]:
        raise ValueError("getSinograms expects a square (padded) image")
    n_angles = core.angle_count(angle_stepsize)
    if n_angles < 1:
        raise ValueError(f"invalid angle step {angle_stepsize}")
    sino = core.sinograms(img.t().contiguous(), n_angles, float(angle_stepsize),
                          tcodes_of(tfunctionals))
    return [sino[t].t() for t in range(sino.shape[0])]


def getCircusFunctions(input, pfunctionals):
    """c++/src/circus.cpp:104-177: one length-A circus function per
    P-functional of the [n, A] sinogram ``input``."""
    sino = _as_device_matrix(input)
    if sino.dim() != 2:
        raise ValueError("getCircusFunctions expects an [n, A] sinogram")
    out = core.circus(sino.t().contiguous(), pfunctionals)
    return [out[p] for p in range(out.shape[0])]


class Transformer:
    """c++/src/transform.cpp:28-112 without the CSV/PGM output.

    ``orthonormal`` (all P-functionals Hermite, c++/src/demo.cpp:153-169)
    stretches the image to ceil(ceil(360/step)/sqrt(2)) squared before
    padding, replaces each sinogram by its nearest orthonormal sinogram and
    sets every Hermite P-functional's centre from it."""

    def __init__(self, image, basename: str = "", angle_stepsize: float = 1,
                 orthonormal: bool = False):
        img = _as_device_matrix(image)
        if img.dim() != 2:
            raise ValueError("image must be a 2-D matrix")
        rows, cols = img.shape
        img_cm = img.t().contiguous()
        if orthonormal:
            ndiag = int(math.ceil(360.0 / angle_stepsize))
            nsize = int(math.ceil(ndiag / math.sqrt(2)))
            img_cm = core.resize(img_cm, rows, cols, nsize, nsize)
            rows = cols = nsize
        self._image = core.pad(img_cm, rows, cols)  # [n, n] column-major
        self._basename = basename
        self._orthonormal = orthonormal
        self._angle_stepsize = angle_stepsize

    @property
    def image(self) -> torch.Tensor:
        """Padded image as an [n, n] matrix."""
        return self._image.t()

    def getTransform(self, tfunctionals, pfunctionals, write_data: bool = False):
        """Signatures [A, nt*np] (column t*np + p = zscore of circus p of T t);
        Hermite centres are written back into ``pfunctionals`` wrappers in
        orthonormal mode, as c++/src/transform.cpp:80-84 does."""
        n_angles = core.angle_count(self._angle_stepsize)
        codes = tcodes_of(tfunctionals)
        pf = pfuncs_of(pfunctionals)
        if any(p[0] == int(PFunctional.Hermite) for p in pf) and not self._orthonormal:
            raise ValueError("Hermite P-functionals need the orthonormal mode")
        sino = core.sinograms(self._image, n_angles, float(self._angle_stepsize), codes)
        if not pf:
            return None
        if not self._orthonormal:
            circ = core.circus(sino, pf)  # [nt, np, A]
            return core.zscore(circ.reshape(-1, n_angles)).t()
        sig = torch.empty((len(codes), len(pf), n_angles), dtype=torch.float32,
                          device=sino.device)
        nos, centers = core.nearest_orthonormal_sinograms(sino)
        for t in range(len(codes)):
            center = centers[t]
            pf_t = [(c, o, center) if c == int(PFunctional.Hermite) else (c, o, ce)
                    for (c, o, ce) in pf]
            for w in pfunctionals:
                if hasattr(w, "arguments") and w.functional == PFunctional.Hermite:
                    w.arguments.center = center
            sig[t] = core.circus(nos[t], pf_t)  # [np, A]
        return core.zscore(sig.reshape(-1, n_angles)).t()
