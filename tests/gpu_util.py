"""Host <-> device helpers for the GPU tests: raw bytes move through torch int tensors unchanged."""
import numpy as np


def to_dev(a):
    import torch
    a = np.ascontiguousarray(a)
    raw = a.view(np.uint8).reshape(-1)
    t = torch.from_numpy(raw.copy()).to("cuda")
    return t


def empty_dev(nbytes, fill=None):
    import torch
    t = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")
    if fill is not None:
        t.fill_(fill)
    return t


def from_dev(t, dtype, shape=None):
    import torch
    torch.cuda.synchronize()
    a = t.cpu().numpy().view(dtype)
    if shape is not None:
        a = a[:int(np.prod(shape))].reshape(shape)
    return a
