"""HBM buffers for the C ABI's device entry points, held by PyTorch-ROCm (plumbing only:
every computation runs in the HIP library)."""
import numpy as np


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device visible to PyTorch")
    return torch


class Buffer:
    """A device copy of a numpy record array (or raw bytes) with its element dtype."""

    def __init__(self, tensor, dtype, count):
        self.t = tensor
        self.dtype = np.dtype(dtype)
        self.count = int(count)

    @property
    def ptr(self):
        return self.t.data_ptr()

    def numpy(self):
        torch = _torch()
        torch.cuda.synchronize()
        return self.t.cpu().numpy().view(self.dtype)[: self.count]

    def rows(self, index):
        """A new buffer holding the records at `index` (a device or host index array)."""
        torch = _torch()
        idx = torch.as_tensor(np.asarray(index, np.int64) if not hasattr(index, "device") else index,
                              device=self.t.device)
        rec = self.t.view(-1, self.dtype.itemsize).index_select(0, idx).reshape(-1)
        return Buffer(rec, self.dtype, idx.numel())


def upload(array):
    torch = _torch()
    a = np.ascontiguousarray(array)
    t = torch.from_numpy(a.reshape(-1).view(np.uint8).copy()).to("cuda")
    return Buffer(t, a.dtype, a.size)


def empty(dtype, count):
    torch = _torch()
    dtype = np.dtype(dtype)
    t = torch.zeros(max(1, count) * dtype.itemsize, dtype=torch.uint8, device="cuda")
    return Buffer(t, dtype, count)


def stream_handle():
    return _torch().cuda.current_stream().cuda_stream


def synchronize():
    _torch().cuda.synchronize()
