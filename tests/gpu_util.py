"""Helpers shared by the GPU tests."""
import sys

import numpy as np


def note(msg):
    """Name the case on stdout before a GPU step (survives a crashed run)."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def supported_by_planner(n):
    """Lengths the device planner schedules (every divisor of p - 1)."""
    if n <= 0:
        return False
    while n % 2 == 0:
        n //= 2
    for q in (3, 5, 17, 257, 65537):
        if n % q == 0:
            n //= q
    return n == 1


def to_torch(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x).view(np.int64)).cuda()


def from_torch(t):
    return t.cpu().numpy().view(np.uint64)
