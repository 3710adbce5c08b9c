"""Shared test helpers: numpy <-> torch block conversion and seeded inputs."""
import numpy as np


def rng(seed):
    return np.random.default_rng(seed)


def rand_blocks(r, n):
    return r.integers(0, 2**63, size=(n, 2), dtype=np.int64).view(np.uint64) ^ (
        r.integers(0, 2, size=(n, 2), dtype=np.uint64) << np.uint64(63))


def to_dev(a):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a.copy()).cuda()


def dev_blocks(n):
    import torch

    return torch.zeros((max(n, 1), 2), dtype=torch.int64, device="cuda")


def to_np(t, dtype=np.uint64):
    a = t.detach().cpu().numpy()
    if dtype == np.uint64 and a.dtype == np.int64:
        return a.view(np.uint64)
    return a
