"""Device-memory plumbing: numpy limb arrays <-> torch device buffers.

Field elements are (n, 4) uint64 numpy arrays (halo2's Fr([u64; 4]) bytes);
on the device they live in int64 torch tensors of the same shape, which the
kernels receive as raw pointers.
"""
import numpy as np
import torch


def device():
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def to_dev(arr):
    arr = np.ascontiguousarray(arr, dtype=np.uint64)
    return torch.from_numpy(arr.view(np.int64)).to(device())


def from_dev(t):
    return t.detach().cpu().contiguous().numpy().view(np.uint64)


def empty(n, words=4):
    return torch.empty((n, words), dtype=torch.int64, device=device())


def zeros(n, words=4):
    return torch.zeros((n, words), dtype=torch.int64, device=device())


def ptr(t):
    return None if t is None else t.data_ptr()


def ptr_array(tensors):
    """device array of device pointers (u64), as the reference passes lists"""
    arr = np.array([t.data_ptr() for t in tensors], dtype=np.uint64)
    return torch.from_numpy(arr.view(np.int64)).to(device())


class Arena:
    """Keeps device tensors alive while asynchronous kernels may still use
    them (a pointer taken from a temporary tensor would otherwise be handed
    back to the caching allocator and could alias the next allocation)."""

    def __init__(self):
        self.items = []

    def hold(self, t):
        self.items.append(t)
        return t.data_ptr()

    def put(self, arr):
        return self.hold(to_dev(arr))

    def ptrs(self, tensors):
        for t in tensors:
            self.items.append(t)
        return self.hold(ptr_array(tensors))

    def put_all(self, arrs):
        return self.ptrs([to_dev(a) for a in arrs])
