"""Helpers shared by the GPU tests (device upload, names printed before each case)."""
import sys

import numpy as np


def cuda():
    import torch
    return torch.device("cuda")


def up(a, dtype=None):
    import torch
    t = torch.as_tensor(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


def down(t):
    return t.detach().cpu().numpy()


def announce(*what):
    """Print and flush what names the case before the GPU runs it."""
    print("[gpu-case]", *what, flush=True)
    sys.stdout.flush()
