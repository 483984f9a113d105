"""Device plumbing on top of PyTorch-ROCm: current device/stream, uploads,
and a per-device scratch arena (so kernels never allocate on the hot path)."""
import numpy
import torch

from . import _lib


def require_gpu():
    """Load the HIP library and check that a GPU is visible (no CPU fallback)."""
    _lib.lib()
    if not torch.cuda.is_available():
        raise RuntimeError('scopyon_amd: no HIP device is visible; the GPU path has no CPU fallback')
    return torch.device('cuda', torch.cuda.current_device())


def stream():
    """Raw hipStream_t of torch's current stream (for the C-ABI)."""
    return torch.cuda.current_stream().cuda_stream


def ptr(t):
    return None if t is None else t.data_ptr()


def upload(a, dtype=torch.float64, device=None):
    device = device or require_gpu()
    a = numpy.ascontiguousarray(a)
    return torch.from_numpy(a).to(device=device, dtype=dtype, non_blocking=False)


class Arena:
    """Grow-only named scratch buffers, one set per device."""

    def __init__(self):
        self._bufs = {}

    def get(self, name, nbytes, device):
        key = (name, device.index if hasattr(device, 'index') else device)
        buf = self._bufs.get(key)
        nbytes = max(int(nbytes), 1)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(int(nbytes * 1.25) + 256, dtype=torch.uint8, device=device)
            self._bufs[key] = buf
        return buf

    def typed(self, name, n, dtype, device):
        itemsize = torch.empty((), dtype=dtype).element_size()
        raw = self.get(name, max(int(n), 1) * itemsize, device)
        return raw[: max(int(n), 1) * itemsize].view(dtype)[: int(n)]


ARENA = Arena()
