"""Differentiable separable-Gaussian image formation on the GPU.

Model (pixel units): emitter (x, y, z, I) adds I * exp(-|z| * inv_depth) *
G(i; x) G(j; y) to pixel (i, j) of its frame over a (2R+1)^2 window, G being
the unit Gaussian of width sigma integrated over each pixel -- the Gaussian
radial profile of src/scopyon/_epifm.py:128-134 with the TIRF amplitude of
_epifm.py:1289.  Forward, Poisson negative log-likelihood and the
hand-written backward (x, y, z, I and sigma) run in libscopyon_hip.so;
``render`` and ``poisson_nll`` wrap them in torch.autograd.Function so the
usual optimisers drive them.

    model = GaussianModel([(256, 256)] * 64, sigma=1.2)
    mu = render(model, x, y, z, I, sigma, frame_id)      # (sum of pixels,) float32
    loss = poisson_nll(model, mu, k).sum()                # float64
    loss.backward()
"""
import ctypes

import numpy
import torch

from . import _device, _lib
from ._lib import FovT, GaussParamsT

__all__ = ['GaussianModel', 'render', 'poisson_nll']

TS = 32   # tile side of the forward pass (smx_splat_tile_size())


class GaussianModel:
    """Frame geometry, model constants and device workspaces for a batch of
    frames of possibly different sizes."""

    def __init__(self, frame_shapes, *, sigma=1.2, inv_depth=1.0 / 0.377, background=10.0,
                 radius=6, method='mfma-f32', device=None):
        self.device = device or _device.require_gpu()
        if method not in self.METHODS:
            raise ValueError('method must be one of {}'.format(sorted(self.METHODS)))
        self.method = method
        if not 1 <= int(radius) <= 7:
            raise ValueError('radius must be in 1..7 (window of at most 15 pixels)')
        if not sigma > 0:
            raise ValueError('sigma must be positive')
        self.shapes = [(int(a), int(b)) for a, b in frame_shapes]
        fovs, pix, tiles = [], 0, 0
        for nw, nh in self.shapes:
            f = FovT()
            f.nw, f.nh = nw, nh
            f.tiles_r, f.tiles_c = (nw + TS - 1) // TS, (nh + TS - 1) // TS
            f.pix_off, f.tile_off = pix, tiles
            f.pixel_length = 1.0
            fovs.append(f)
            pix += nw * nh
            tiles += f.tiles_r * f.tiles_c
        self.nframes, self.n_pix, self.n_tiles = len(fovs), pix, tiles
        self.pix_offsets = numpy.array([f.pix_off for f in fovs] + [pix])
        raw = b''.join(bytes(f) for f in fovs)
        self.frames_dev = torch.from_numpy(numpy.frombuffer(raw, dtype=numpy.uint8).copy()).to(
            self.device)
        self.sigma = float(sigma)
        self.inv_depth = float(inv_depth)
        self.background = float(background)
        self.radius = int(radius)
        self.max_frame_pixels = max(a * b for a, b in self.shapes)
        self.status = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.loss = torch.zeros(self.nframes, dtype=torch.float64, device=self.device)
        self.gsigma = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._ws = None
        self._cap = 0
        self._bin_key = None   # what the workspace's tile binning was built for
        lib = _lib.lib()
        self._nll_ws = torch.empty(
            max(1, lib.smx_gauss_nll_workspace_bytes(self.nframes, self.max_frame_pixels)),
            dtype=torch.uint8, device=self.device)

    def params(self, sigma=None):
        p = GaussParamsT()
        p.sigma = float(self.sigma if sigma is None else sigma)
        p.inv_depth = self.inv_depth
        p.background = self.background
        p.radius = self.radius
        return p

    def frame(self, buf, f):
        a, b = self.pix_offsets[f], self.pix_offsets[f + 1]
        return buf[a:b].view(*self.shapes[f])

    def _workspace(self, n):
        cap = max(1, 4 * n)
        lib = _lib.lib()
        need = max(lib.smx_gauss_workspace_bytes(n, self.n_tiles, cap),
                   lib.smx_gauss_backward_workspace_bytes(n))
        if self._ws is None or self._ws.numel() < need or self._cap < cap:
            self._ws = torch.empty(int(need), dtype=torch.uint8, device=self.device)
            self._cap = cap
        return self._ws, cap

    @staticmethod
    def _check(t, name):
        if t is None or t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise TypeError('{} must be a contiguous float32 CUDA tensor'.format(name))

    def _frame_ids(self, frame_id, n):
        if frame_id is None:
            if self.nframes != 1:
                raise ValueError('frame_id is required with more than one frame')
            return None
        if frame_id.dtype != torch.int32 or frame_id.numel() != n or not frame_id.is_cuda:
            raise TypeError('frame_id must be an int32 CUDA tensor with one entry per emitter')
        return frame_id

    # ---- raw kernels (no autograd) ----
    METHODS = {'lds': None, 'mfma-f32': 0, 'mfma-f16': 1, 'mfma-bf16': 2}

    def forward(self, x, y, z, I, frame_id=None, sigma=None, out=None, method=None):
        """Expected image.  method: 'mfma-f32' (the default) / 'mfma-f16' /
        'mfma-bf16' (the separable contraction on the matrix cores with
        float32, float16 or bfloat16 operands and float32 accumulation; z may
        be None) or 'lds' (float32 products added in LDS; same accuracy as
        'mfma-f32', slower at these densities)."""
        method = method or self.method
        if method not in self.METHODS:
            raise ValueError('method must be one of {}'.format(sorted(self.METHODS)))
        for t, nm in ((x, 'x'), (y, 'y'), (I, 'I')):
            self._check(t, nm)
        if z is not None or method == 'lds':
            self._check(z, 'z')
        n = x.numel()
        fid = self._frame_ids(frame_id, n)
        ws, cap = self._workspace(n)
        mu = out if out is not None else torch.empty(self.n_pix, dtype=torch.float32,
                                                     device=self.device)
        p = self.params(sigma)
        self._bin_key = None
        if method == 'lds':
            _lib.call('smx_gauss_forward', n, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                      I.data_ptr(), _device.ptr(fid), self.frames_dev.data_ptr(), self.nframes,
                      self.n_tiles, ctypes.byref(p), cap, mu.data_ptr(), ws.data_ptr(),
                      ws.numel(), self.status.data_ptr(), _device.stream())
        else:
            _lib.call('smx_sep_gemm', n, x.data_ptr(), y.data_ptr(), _device.ptr(z), I.data_ptr(),
                      _device.ptr(fid), self.frames_dev.data_ptr(), self.nframes, self.n_tiles,
                      ctypes.byref(p), self.METHODS[method], cap, mu.data_ptr(), ws.data_ptr(),
                      ws.numel(), self.status.data_ptr(), _device.stream())
        self._bin_key = self._binning_key(x, y, fid, p, n, cap)
        return mu

    def forward_nll(self, x, y, z, I, k, frame_id=None, grad_mu=None, sigma=None, mu=None,
                    method=None):
        """Forward image and Poisson NLL in one pass on the matrix cores
        (method 'mfma-f32' / 'mfma-f16' / 'mfma-bf16'): per-frame float64 loss
        sum(mu - k log mu) (the same per-pixel terms as nll(), summed per tile
        in a fixed order) and, into grad_mu, dL/dmu = 1 - k / mu, computed from
        the tile's accumulators; the image itself is written only into `mu`
        when given."""
        method = method or self.method
        if method not in self.METHODS or method == 'lds':
            raise ValueError('forward_nll runs the matrix-core forms (mfma-*)')
        for t, nm in ((x, 'x'), (y, 'y'), (I, 'I'), (k, 'k')):
            self._check(t, nm)
        if z is not None:
            self._check(z, 'z')
        n = x.numel()
        fid = self._frame_ids(frame_id, n)
        ws, cap = self._workspace(n)
        p = self.params(sigma)
        self._bin_key = None
        _lib.call('smx_sep_gemm_nll', n, x.data_ptr(), y.data_ptr(), _device.ptr(z), I.data_ptr(),
                  _device.ptr(fid), self.frames_dev.data_ptr(), self.nframes, self.n_tiles,
                  ctypes.byref(p), self.METHODS[method], cap, _device.ptr(mu), k.data_ptr(),
                  _device.ptr(grad_mu), self.loss.data_ptr(), ws.data_ptr(), ws.numel(),
                  self.status.data_ptr(), _device.stream())
        self._bin_key = self._binning_key(x, y, fid, p, n, cap)
        return self.loss

    def nll(self, mu, k, grad_mu=None):
        """Per-frame float64 loss (device tensor) and, into grad_mu, 1 - k/mu."""
        self._check(mu, 'mu')
        self._check(k, 'k')
        _lib.call('smx_gauss_poisson_nll', mu.data_ptr(), k.data_ptr(), self.frames_dev.data_ptr(),
                  self.nframes, self.max_frame_pixels, self.loss.data_ptr(),
                  _device.ptr(grad_mu), self._nll_ws.data_ptr(), self._nll_ws.numel(),
                  _device.stream())
        return self.loss

    @staticmethod
    def _binning_key(x, y, fid, p, n, cap):
        return (x.data_ptr(), x._version, y.data_ptr(), y._version,
                None if fid is None else (fid.data_ptr(), fid._version), p.sigma, p.radius, n, cap)

    def backward(self, x, y, z, I, grad_mu, frame_id=None, sigma=None, outputs=None,
                 binned=None):
        """dL/dx, dL/dy, dL/dz, dL/dI (per emitter) and dL/dsigma (float64
        device scalar) for grad_mu = dL/dmu.  binned=None reuses the tile
        binning of this model's last forward when it was built for these
        very x, y, frame ids and sigma (smx_gauss_backward_binned), and runs
        the standalone backward otherwise; True requires the binned form,
        False forces the standalone one.  Both give identical per-emitter
        gradients."""
        n = x.numel()
        fid = self._frame_ids(frame_id, n)
        self._check(grad_mu, 'grad_mu')
        for t, nm in ((x, 'x'), (y, 'y'), (z, 'z'), (I, 'I')):
            self._check(t, nm)
        if outputs is None:
            outputs = [torch.empty(n, dtype=torch.float32, device=self.device) for _ in range(4)]
        gx, gy, gz, gI = outputs
        ws, cap = self._workspace(n)
        p = self.params(sigma)
        reuse = self._bin_key is not None and self._bin_key == self._binning_key(x, y, fid, p, n,
                                                                                 cap)
        if binned and not reuse:
            raise RuntimeError('no forward binning of these emitters and sigma to reuse')
        if reuse and binned is not False:
            _lib.call('smx_gauss_backward_binned', n, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                      I.data_ptr(), _device.ptr(fid), self.frames_dev.data_ptr(), self.nframes,
                      self.n_tiles, ctypes.byref(p), cap, grad_mu.data_ptr(), gx.data_ptr(),
                      gy.data_ptr(), gz.data_ptr(), gI.data_ptr(), self.gsigma.data_ptr(),
                      ws.data_ptr(), ws.numel(), self.status.data_ptr(), _device.stream())
        else:
            # the standalone backward uses the workspace as scratch: the
            # forward's binning is gone afterwards
            self._bin_key = None
            _lib.call('smx_gauss_backward', n, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                      I.data_ptr(), _device.ptr(fid), self.frames_dev.data_ptr(), self.nframes,
                      ctypes.byref(p), grad_mu.data_ptr(), gx.data_ptr(), gy.data_ptr(),
                      gz.data_ptr(), gI.data_ptr(), self.gsigma.data_ptr(), ws.data_ptr(),
                      ws.numel(), _device.stream())
        return gx, gy, gz, gI, self.gsigma

    def check_status(self):
        s = int(self.status.item())
        if s & 2:
            raise RuntimeError('tile list capacity exceeded')
        if s & 16:
            raise RuntimeError('binned backward ran on the binning of another forward')
        return s


class _Render(torch.autograd.Function):

    @staticmethod
    def forward(ctx, model, x, y, z, I, sigma, frame_id):
        s = float(sigma.detach().item()) if torch.is_tensor(sigma) else float(sigma)
        ctx.model, ctx.s, ctx.frame_id = model, s, frame_id
        ctx.sigma_is_tensor = torch.is_tensor(sigma)
        ctx.save_for_backward(x, y, z, I)
        return model.forward(x.detach().contiguous(), y.detach().contiguous(),
                             z.detach().contiguous(), I.detach().contiguous(), frame_id, s)

    @staticmethod
    def backward(ctx, grad_mu):
        x, y, z, I = ctx.saved_tensors
        gx, gy, gz, gI, gs = ctx.model.backward(x.contiguous(), y.contiguous(), z.contiguous(),
                                                I.contiguous(), grad_mu.contiguous().float(),
                                                ctx.frame_id, ctx.s)
        gsig = gs.to(torch.float32).reshape(()) if ctx.sigma_is_tensor else None
        return None, gx, gy, gz, gI, gsig, None


class _NLL(torch.autograd.Function):

    @staticmethod
    def forward(ctx, model, mu, k):
        grad = torch.empty_like(mu)
        loss = model.nll(mu.detach().contiguous(), k.contiguous(), grad).clone()
        ctx.save_for_backward(grad)
        ctx.model = model
        return loss

    @staticmethod
    def backward(ctx, g_loss):
        (grad,) = ctx.saved_tensors
        model = ctx.model
        # per-frame upstream weights (float64) onto that frame's pixels
        scale = torch.repeat_interleave(
            g_loss.to(torch.float32),
            torch.as_tensor(numpy.diff(model.pix_offsets), device=grad.device))
        return None, grad * scale, None


def render(model, x, y, z, I, sigma=None, frame_id=None):
    """Expected image mu (float32, frames concatenated) with autograd support
    for x, y, z, I and sigma (a 0-d tensor; a float is held constant)."""
    return _Render.apply(model, x, y, z, I, model.sigma if sigma is None else sigma, frame_id)


def poisson_nll(model, mu, k):
    """Per-frame float64 Poisson negative log-likelihood sum(mu - k log mu)."""
    return _NLL.apply(model, mu, k)
