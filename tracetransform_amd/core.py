"""Device-level entry points over torch tensors (HBM-resident buffers).

Every call runs the HIP kernels of ``libtt_amd.so`` on the current torch
stream of the tensor's device.  Memory layouts are the reference's
column-major ones, spelled as C-ordered torch shapes:

  padded image  [n, n]   memory (col, row): tensor[col, row] = image(row, col)
  sinograms     [nt, A, n]  (one angle's n values contiguous)
  circus        [n_sino, np, A]
  index outputs [A, n]
"""
from __future__ import annotations

import ctypes as C
import threading

import torch

from . import _lib
from .functionals import pfuncs_of, tcodes_of

_tls = threading.local()


class Context:
    """Owns one tt_context (stream binding, device tables, workspaces)."""

    def __init__(self, device: int = 0):
        L = _lib.load()
        h = C.c_void_p()
        _lib.check("tt_context_create", L.tt_context_create(int(device), C.byref(h)))
        self.handle = h
        self.device = int(device)

    def close(self):
        if self.handle:
            _lib.load().tt_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bind_current_stream(self):
        s = torch.cuda.current_stream(self.device)
        _lib.check("tt_context_set_stream",
                   _lib.load().tt_context_set_stream(self.handle, C.c_void_p(s.cuda_stream)))

    def profile(self, enable: bool):
        _lib.check("tt_profile_enable",
                   _lib.load().tt_profile_enable(self.handle, int(bool(enable))))

    def profile_read(self, kind: int):
        """(summed kernel ms, launch count) of one kernel kind since the last
        read (see TT_KERNEL_* in include/tt.h)."""
        ms, cnt = C.c_double(), C.c_int64()
        _lib.check("tt_profile_read",
                   _lib.load().tt_profile_read(self.handle, int(kind), C.byref(ms),
                                               C.byref(cnt)))
        return ms.value, cnt.value

    def synchronize(self):
        _lib.check("tt_context_synchronize", _lib.load().tt_context_synchronize(self.handle))


def get_context(device: int | None = None) -> Context:
    if device is None:
        device = torch.cuda.current_device()
    cache = getattr(_tls, "ctx", None)
    if cache is None:
        cache = _tls.ctx = {}
    if device not in cache:
        cache[device] = Context(device)
    return cache[device]


def _dptr(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _require_cuda_f32(name: str, t: torch.Tensor):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != torch.float32:
        raise ValueError(f"{name} must be float32 (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def angle_count(step_deg: float) -> int:
    return _lib.load().tt_angle_count(float(step_deg))


def pad_geometry(rows: int, cols: int):
    n, dx, dy = C.c_int(), C.c_int(), C.c_int()
    _lib.check("tt_pad_geometry",
               _lib.load().tt_pad_geometry(int(rows), int(cols), C.byref(n),
                                           C.byref(dx), C.byref(dy)))
    return n.value, dx.value, dy.value


def pad(image_cm: torch.Tensor, rows: int, cols: int) -> torch.Tensor:
    """Zero-pad a column-major rows x cols image ([cols, rows] tensor) to the
    reference's n x n square (c++/src/auxiliary.cpp:257-275)."""
    _require_cuda_f32("image", image_cm)
    if image_cm.numel() != rows * cols:
        raise ValueError("image size does not match rows x cols")
    n, _, _ = pad_geometry(rows, cols)
    out = torch.empty((n, n), dtype=torch.float32, device=image_cm.device)
    ctx = get_context(image_cm.device.index)
    ctx.bind_current_stream()
    _lib.check("tt_pad", _lib.load().tt_pad(ctx.handle, _dptr(image_cm), rows, cols,
                                            _dptr(out)))
    return out


def sinograms(image_cm: torch.Tensor, n_angles: int, step_deg: float, tfunctionals,
              out: torch.Tensor | None = None, with_aux: bool = False):
    """Sinograms of a padded n x n image ([n, n] column-major tensor).

    Returns ``sino`` [nt, A, n], or ``(sino, aux)`` with int32 [A, n] index
    tensors ``median_x`` and ``median_sqrt`` (plus ``t6_rank``, ``t6_src``
    and ``t7_rank``, ``t7_src`` when T6/T7 are requested) when ``with_aux``."""
    _require_cuda_f32("image", image_cm)
    if image_cm.dim() != 2 or image_cm.shape[0] != image_cm.shape[1]:
        raise ValueError("image must be a square padded [n, n] tensor")
    n = image_cm.shape[0]
    codes = tcodes_of(tfunctionals)
    nt = len(codes)
    dev = image_cm.device
    if out is None:
        out = torch.empty((nt, n_angles, n), dtype=torch.float32, device=dev)
    else:
        _require_cuda_f32("out", out)
        if tuple(out.shape) != (nt, n_angles, n):
            raise ValueError(f"out must have shape {(nt, n_angles, n)}")
    aux = None
    aux_s = None
    if with_aux:
        names = ["median_x", "median_sqrt"]
        if 6 in codes:
            names += ["t6_rank", "t6_src"]
        if 7 in codes:
            names += ["t7_rank", "t7_src"]
        aux = {k: torch.full((n_angles, n), -1, dtype=torch.int32, device=dev)
               for k in names}
        ptr = {k: (aux[k].data_ptr() if k in aux else None)
               for k in ("median_x", "median_sqrt", "t6_rank", "t6_src", "t7_rank",
                         "t7_src")}
        aux_s = _lib.tt_sino_aux(ptr["median_x"], ptr["median_sqrt"], ptr["t6_rank"],
                                 ptr["t6_src"], ptr["t7_rank"], ptr["t7_src"])
    ctx = get_context(dev.index)
    ctx.bind_current_stream()
    carr = (C.c_int32 * max(nt, 1))(*codes)
    _lib.check("tt_sinograms",
               _lib.load().tt_sinograms(ctx.handle, _dptr(image_cm), n, int(n_angles),
                                        float(step_deg), carr, nt,
                                        _dptr(out) if nt else None,
                                        C.byref(aux_s) if aux_s is not None else None))
    return (out, aux) if with_aux else out


def circus(sino: torch.Tensor, pfunctionals, out: torch.Tensor | None = None,
           with_p2_rank: bool = False):
    """P-functionals over sinogram columns.

    ``sino`` is [n_sino, n_cols, n] (or [n_cols, n] for one sinogram).
    Returns ``circus`` [n_sino, np, n_cols] (and the P2 rank tensor)."""
    _require_cuda_f32("sino", sino)
    single = sino.dim() == 2
    s3 = sino.unsqueeze(0) if single else sino
    if s3.dim() != 3:
        raise ValueError("sino must be [n_sino, n_cols, n] or [n_cols, n]")
    n_sino, n_cols, n = s3.shape
    pf = pfuncs_of(pfunctionals)
    npf = len(pf)
    dev = sino.device
    if out is None:
        out = torch.empty((n_sino, npf, n_cols), dtype=torch.float32, device=dev)
    else:
        _require_cuda_f32("out", out)
        if tuple(out.shape) != (n_sino, npf, n_cols):
            raise ValueError(f"out must have shape {(n_sino, npf, n_cols)}")
    rank = None
    if with_p2_rank:
        rank = torch.full((n_sino, npf, n_cols), -1, dtype=torch.int32, device=dev)
    arr = (_lib.tt_pfunc * max(npf, 1))(*[_lib.tt_pfunc(*t) for t in pf])
    ctx = get_context(dev.index)
    ctx.bind_current_stream()
    _lib.check("tt_circus",
               _lib.load().tt_circus(ctx.handle, _dptr(s3), n, n_cols, n_sino, arr, npf,
                                     _dptr(out), _dptr(rank) if rank is not None else None))
    if single:
        out = out[0]
        rank = rank[0] if rank is not None else None
    return (out, rank) if with_p2_rank else out


class Batch:
    """A batch of padded images with preallocated outputs for
    ``tt_forward_batched``.  Reusing the same Batch (same buffers) replays the
    captured HIP graph; only the image contents may change between calls."""

    def __init__(self, images, n_angles: int, step_deg: float, tfunctionals,
                 pfunctionals=(), save_for_backward: bool = False):
        self.images = [im for im in images]
        for k, im in enumerate(self.images):
            _require_cuda_f32(f"images[{k}]", im)
            if im.dim() != 2 or im.shape[0] != im.shape[1]:
                raise ValueError(f"images[{k}] must be a square padded [n, n] tensor")
        self.n_angles = int(n_angles)
        self.step = float(step_deg)
        self.codes = tcodes_of(tfunctionals)
        self.pf = pfuncs_of(pfunctionals)
        dev = self.images[0].device
        nt, npf, A = len(self.codes), len(self.pf), self.n_angles
        self.sino = [torch.empty((nt, A, im.shape[0]), dtype=torch.float32, device=dev)
                     for im in self.images]
        # one [nimg, nt, np, A] tensor so callers can form losses in one op
        self.circus_all = torch.empty((len(self.images), nt, npf, A), dtype=torch.float32,
                                      device=dev) if npf else None
        self.circus = [self.circus_all[k] if npf else None
                       for k in range(len(self.images))]
        # with save_for_backward, per image the line data a BatchBackward built
        # from this batch reads instead of recomputing the forward
        self.saved = None
        if save_for_backward:
            L = _lib.load()
            self.saved = [torch.empty(int(L.tt_saved_bytes(im.shape[0], A, nt)),
                                      dtype=torch.uint8, device=dev) for im in self.images]
        arr = (_lib.tt_image * len(self.images))()
        for k, im in enumerate(self.images):
            arr[k].d_image = im.data_ptr()
            arr[k].n = im.shape[0]
            arr[k].d_sino = self.sino[k].data_ptr()
            arr[k].d_circus = self.circus[k].data_ptr() if npf else None
            arr[k].d_saved = self.saved[k].data_ptr() if self.saved else None
        self._arr = arr
        self._codes = (C.c_int32 * nt)(*self.codes)
        self._pf = (_lib.tt_pfunc * max(npf, 1))(*[_lib.tt_pfunc(*t) for t in self.pf])
        self.device = dev.index

    def run(self, autotune: bool | None = None):
        ctx = get_context(self.device)
        L = _lib.load()
        if autotune is not None:
            _lib.check("tt_context_set_autotune",
                       L.tt_context_set_autotune(ctx.handle, int(bool(autotune))))
        ctx.bind_current_stream()
        _lib.check("tt_forward_batched",
                   L.tt_forward_batched(ctx.handle, self._arr, len(self.images),
                                        self.n_angles, self.step, self._codes,
                                        len(self.codes), self._pf, len(self.pf)))
        return self.sino, self.circus

    def saved_medians(self, k: int):
        """Image k's weighted medians (findWeightedMedian of x and of
        sqrt(x), [A, n] int32 tensors, None where the functionals did not
        need them) from its saved line data, as of the last run."""
        if self.saved is None:
            raise ValueError("the Batch was not built with save_for_backward")
        n, A, nt = self.images[k].shape[0], self.n_angles, len(self.codes)
        offs = [C.c_size_t() for _ in range(5)]
        _lib.check("tt_saved_offsets",
                   _lib.load().tt_saved_offsets(n, A, nt, *[C.byref(o) for o in offs]))
        raw = self.saved[k]

        def part(off):
            return raw[off:off + 4 * A * n].view(torch.int32).view(A, n)

        need_x = any(c in (1, 2) for c in self.codes)
        need_s = any(c in (3, 4, 5) for c in self.codes)
        return (part(offs[0].value) if need_x else None,
                part(offs[1].value) if need_s else None)

    def launch_shape(self):
        """(waves per workgroup, workgroups, tile angles, tile columns) of
        the line kernel of the last planned batch."""
        ctx = get_context(self.device)
        w, b = C.c_int(), C.c_int()
        _lib.check("tt_batched_launch_shape",
                   _lib.load().tt_batched_launch_shape(ctx.handle, C.byref(w), C.byref(b)))
        t = batched_tiles(self.device)
        return w.value, b.value, t["line_angles"], t["line_columns"]


def set_tiles(device: int = 0, line=None, gather=None):
    """Forces the batched paths' tiles on ``device``'s context:
    ``line`` = (angles, columns) of a line-kernel wave item, ``gather`` =
    (tile side, waves per workgroup) of the backward's transposed rotation;
    None leaves the default (or the tuned choice)."""
    t = _lib.tt_tiles()
    if line is not None:
        t.line_angles, t.line_columns = int(line[0]), int(line[1])
    if gather is not None:
        t.gather_tile, t.gather_waves = int(gather[0]), int(gather[1])
    ctx = get_context(device)
    _lib.check("tt_context_set_tiles", _lib.load().tt_context_set_tiles(ctx.handle, C.byref(t)))


def batched_tiles(device: int = 0) -> dict:
    """Tiles of the last planned batched forward (line_*) and backward
    (gather_*, zero before any)."""
    t = _lib.tt_tiles()
    ctx = get_context(device)
    _lib.check("tt_batched_tiles", _lib.load().tt_batched_tiles(ctx.handle, C.byref(t)))
    return {f: getattr(t, f) for f, _ in t._fields_}


# ---------------------------------------------------------------- adjoint

def circus_backward(sino: torch.Tensor, pfunctionals, dcircus: torch.Tensor):
    """J_circus^T dcircus: ``sino`` [n_sino, n_cols, n], ``dcircus``
    [n_sino, np, n_cols] -> dsino [n_sino, n_cols, n]."""
    _require_cuda_f32("sino", sino)
    _require_cuda_f32("dcircus", dcircus)
    n_sino, n_cols, n = sino.shape
    pf = pfuncs_of(pfunctionals)
    dsino = torch.empty_like(sino)
    arr = (_lib.tt_pfunc * max(len(pf), 1))(*[_lib.tt_pfunc(*t) for t in pf])
    ctx = get_context(sino.device.index)
    ctx.bind_current_stream()
    _lib.check("tt_circus_backward",
               _lib.load().tt_circus_backward(ctx.handle, _dptr(sino), n, n_cols, n_sino,
                                              arr, len(pf), _dptr(dcircus), _dptr(dsino)))
    return dsino


def sinograms_backward(image_cm: torch.Tensor, n_angles: int, step_deg: float,
                       tfunctionals, dsino: torch.Tensor):
    """J_sinograms^T dsino ([nt, A, n]) -> dimage [n, n] (column-major)."""
    _require_cuda_f32("image", image_cm)
    _require_cuda_f32("dsino", dsino)
    n = image_cm.shape[0]
    codes = tcodes_of(tfunctionals)
    dimg = torch.empty_like(image_cm)
    ctx = get_context(image_cm.device.index)
    ctx.bind_current_stream()
    carr = (C.c_int32 * len(codes))(*codes)
    _lib.check("tt_sinograms_backward",
               _lib.load().tt_sinograms_backward(ctx.handle, _dptr(image_cm), n,
                                                 int(n_angles), float(step_deg), carr,
                                                 len(codes), _dptr(dsino), _dptr(dimg)))
    return dimg


def backward(image_cm: torch.Tensor, n_angles: int, step_deg: float, tfunctionals,
             pfunctionals, dcircus: torch.Tensor):
    """J^T dcircus ([nt, np, A]) of image -> circus -> dimage [n, n]."""
    _require_cuda_f32("image", image_cm)
    _require_cuda_f32("dcircus", dcircus)
    n = image_cm.shape[0]
    codes = tcodes_of(tfunctionals)
    pf = pfuncs_of(pfunctionals)
    dimg = torch.empty_like(image_cm)
    ctx = get_context(image_cm.device.index)
    ctx.bind_current_stream()
    carr = (C.c_int32 * len(codes))(*codes)
    arr = (_lib.tt_pfunc * max(len(pf), 1))(*[_lib.tt_pfunc(*t) for t in pf])
    _lib.check("tt_backward",
               _lib.load().tt_backward(ctx.handle, _dptr(image_cm), n, int(n_angles),
                                       float(step_deg), carr, len(codes), arr, len(pf),
                                       _dptr(dcircus), _dptr(dimg)))
    return dimg


class TraceTransformFunction(torch.autograd.Function):
    """Thin autograd wrapper (used for gradient checking): padded image
    [n, n] (column-major) -> circus functions [nt, np, A]."""

    @staticmethod
    def forward(ctx, image_cm, n_angles, step_deg, tcodes, pfuncs):
        sino = sinograms(image_cm, n_angles, step_deg, tcodes)
        circ = circus(sino, pfuncs)
        ctx.save_for_backward(image_cm)
        ctx.args = (n_angles, step_deg, tcodes, pfuncs)
        return circ

    @staticmethod
    def backward(ctx, dcirc):
        (image_cm,) = ctx.saved_tensors
        n_angles, step_deg, tcodes, pfuncs = ctx.args
        dimg = backward(image_cm, n_angles, step_deg, tcodes, pfuncs,
                        dcirc.contiguous().float())
        return dimg, None, None, None, None


def trace_transform(image_cm, n_angles, step_deg, tfunctionals, pfunctionals):
    return TraceTransformFunction.apply(image_cm, n_angles, step_deg,
                                        tcodes_of(tfunctionals), pfuncs_of(pfunctionals))


# ------------------------------------------------ steps either side of the path

def resize(image_cm: torch.Tensor, in_rows: int, in_cols: int, rows: int, cols: int):
    """c++/src/auxiliary.cpp:198-221 on a column-major image ([cols, rows]
    tensor in, [cols', rows'] out)."""
    _require_cuda_f32("image", image_cm)
    out = torch.empty((cols, rows), dtype=torch.float32, device=image_cm.device)
    ctx = get_context(image_cm.device.index)
    ctx.bind_current_stream()
    _lib.check("tt_resize", _lib.load().tt_resize(ctx.handle, _dptr(image_cm), in_rows,
                                                  in_cols, _dptr(out), rows, cols))
    return out


def zscore(x: torch.Tensor):
    """c++/src/auxiliary.cpp:305-318 over the last dimension."""
    _require_cuda_f32("x", x)
    n = x.shape[-1]
    out = torch.empty_like(x)
    ctx = get_context(x.device.index)
    ctx.bind_current_stream()
    _lib.check("tt_zscore", _lib.load().tt_zscore(ctx.handle, _dptr(x), n,
                                                  x.numel() // n, _dptr(out)))
    return out


def nearest_orthonormal_sinogram(sino: torch.Tensor):
    """c++/src/circus.cpp:68-102 for one sinogram ([A, n] tensor, i.e. the
    n x A column-major matrix).  Returns (nos [A, n'], center)."""
    _require_cuda_f32("sino", sino)
    A, n = sino.shape
    cap = 3 * n
    out = torch.empty((A, cap), dtype=torch.float32, device=sino.device)
    rows, center = C.c_int(), C.c_int()
    ctx = get_context(sino.device.index)
    ctx.bind_current_stream()
    _lib.check("tt_nearest_orthonormal_sinogram",
               _lib.load().tt_nearest_orthonormal_sinogram(
                   ctx.handle, _dptr(sino), n, A, _dptr(out), cap, C.byref(rows),
                   C.byref(center)))
    r = rows.value
    # the kernel wrote an r x A column-major matrix into the first r*A floats
    nos = out.view(-1)[: r * A].view(A, r).clone()
    return nos, center.value


def nearest_orthonormal_sinograms(sino: torch.Tensor, out: torch.Tensor | None = None):
    """nearest_orthonormal_sinogram of every sinogram of ``sino``
    ([n_sino, A, n], as ``sinograms`` returns them) in one batched solve.
    Returns (list of [A, n'_t] views, list of centres); ``out`` is an
    optional [n_sino, A * 3n] workspace the views point into."""
    _require_cuda_f32("sino", sino)
    if sino.dim() != 3:
        raise ValueError("sino must be [n_sino, A, n]")
    n_sino, A, n = sino.shape
    cap = 3 * n
    if out is None:
        out = torch.empty((n_sino, A * cap), dtype=torch.float32, device=sino.device)
    else:
        _require_cuda_f32("out", out)
        if tuple(out.shape) != (n_sino, A * cap):
            raise ValueError(f"out must have shape {(n_sino, A * cap)}")
    rows = (C.c_int * n_sino)()
    centers = (C.c_int * n_sino)()
    ctx = get_context(sino.device.index)
    ctx.bind_current_stream()
    _lib.check("tt_nearest_orthonormal_sinograms",
               _lib.load().tt_nearest_orthonormal_sinograms(
                   ctx.handle, _dptr(sino), n, A, n_sino, _dptr(out), cap, rows, centers))
    # sinogram t's factor is a rows[t] x A column-major matrix at the start
    # of out[t]
    views = [out[t, : rows[t] * A].view(A, rows[t]) for t in range(n_sino)]
    return views, [centers[t] for t in range(n_sino)]


class BatchBackward:
    """A batch for ``tt_backward_batched``: per image an upstream gradient
    [nt, np, A] in, the image gradient [n, n] out; reusing the object replays
    the captured graph."""

    def __init__(self, images, n_angles: int, step_deg: float, tfunctionals,
                 pfunctionals, forward: Batch | None = None):
        """``forward``: a Batch of the same images, angles and functionals
        built with ``save_for_backward``; its line data and sinograms from its
        last run replace the forward recompute (the images must not change
        between that run and this one)."""
        self.images = list(images)
        self.n_angles = int(n_angles)
        self.step = float(step_deg)
        self.codes = tcodes_of(tfunctionals)
        self.pf = pfuncs_of(pfunctionals)
        dev = self.images[0].device
        nt, npf = len(self.codes), len(self.pf)
        # one [nimg, nt, np, A] tensor; dcircus[k] is image k's view
        self.dcircus_all = torch.zeros((len(self.images), nt, npf, self.n_angles),
                                       dtype=torch.float32, device=dev)
        self.dcircus = [self.dcircus_all[k] for k in range(len(self.images))]
        self.dimage = [torch.empty_like(im) for im in self.images]
        if forward is not None and forward.saved is None:
            raise ValueError("the forward Batch was not built with save_for_backward")
        self.forward = forward
        arr = (_lib.tt_image_grad * len(self.images))()
        for k, im in enumerate(self.images):
            _require_cuda_f32(f"images[{k}]", im)
            arr[k].d_image = im.data_ptr()
            arr[k].n = im.shape[0]
            arr[k].d_dcircus = self.dcircus[k].data_ptr()
            arr[k].d_dimage = self.dimage[k].data_ptr()
            if forward is not None:
                arr[k].d_sino = forward.sino[k].data_ptr()
                arr[k].d_saved = forward.saved[k].data_ptr()
        self._arr = arr
        self._codes = (C.c_int32 * nt)(*self.codes)
        self._pf = (_lib.tt_pfunc * npf)(*[_lib.tt_pfunc(*t) for t in self.pf])
        self.device = dev.index

    def run(self):
        ctx = get_context(self.device)
        ctx.bind_current_stream()
        _lib.check("tt_backward_batched",
                   _lib.load().tt_backward_batched(ctx.handle, self._arr, len(self.images),
                                                   self.n_angles, self.step, self._codes,
                                                   len(self.codes), self._pf, len(self.pf)))
        return self.dimage
