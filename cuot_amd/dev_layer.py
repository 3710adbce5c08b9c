"""Python mirror of emp-ot Ferret's GPU layer (ferret/emp-ot/emp-ot/ferret/dev_layer.h:9-18).

Same names, argument meaning and blocking host-in/host-out behaviour as the
reference's ``cuda_*`` calls; ``Mat`` arguments become lists of torch device
tensors (one per GPU; this build drives one GPU), ``blk*`` host arrays become
numpy uint64 arrays of shape (N, 2).  All work runs in libcuot_amd.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

P = C.c_void_p


def _dev_ptrs(tensors):
    arr = (P * len(tensors))(*[t.data_ptr() for t in tensors])
    return arr


def _host(a, dtype=np.uint64):
    assert a.flags["C_CONTIGUOUS"] and a.dtype == dtype
    return a.ctypes.data_as(P)


def cuda_setdev(gpu: int):
    check(_lib.load().cuot_setdev(gpu), "cuot_setdev")


def cuda_mpcot_sender(ctx, expanded, buffer, sep, lSum_h, rSum_h, secret_sum, t, depth, delta, ngpu=1):
    """Ferret MPCOT sender (dev_layer.h:11-12).  expanded[0]: leaves out; its
    first t[0] blocks hold the roots on entry."""
    tt = (C.c_int * len(t))(*t)
    d = np.ascontiguousarray(delta, dtype=np.uint64).reshape(2)
    check(_lib.load().cuot_mpcot_sender(ctx.handle, _dev_ptrs(expanded), _dev_ptrs(buffer), _dev_ptrs(sep),
                                        _host(lSum_h), _host(rSum_h), _host(secret_sum), tt, depth, _host(d), ngpu),
          "cuot_mpcot_sender")


def cuda_mpcot_recver(ctx, expanded, buffer, sep, cSum_h, secret_sum, t, depth, choices, ngpu=1):
    """Ferret MPCOT receiver (dev_layer.h:14-15); choices: bool per tree*depth+level."""
    tt = (C.c_int * len(t))(*t)
    ch = np.ascontiguousarray(choices, dtype=np.uint8)
    check(_lib.load().cuot_mpcot_recver(ctx.handle, _dev_ptrs(expanded), _dev_ptrs(buffer), _dev_ptrs(sep),
                                        _host(cSum_h), _host(secret_sum), tt, depth, _host(ch, np.uint8), ngpu),
          "cuot_mpcot_recver")


def cuda_primal_lpn(ctx, role, pubMats, d, n, k, key, expanded, nn, kk_d, kk, ngpu=1):
    """10-local primal LPN (dev_layer.h:17-18); key: PRP round keys (first 16 B = seed)."""
    nn_arr = (C.c_int64 * len(n))(*n)
    key = np.ascontiguousarray(key, dtype=np.uint32)
    check(_lib.load().cuot_primal_lpn(ctx.handle, role, _dev_ptrs(pubMats) if pubMats else None, d, nn_arr, k,
                                      _host(key, np.uint32), _dev_ptrs(expanded), _host(nn), _dev_ptrs(kk_d),
                                      _host(kk), ngpu), "cuot_primal_lpn")
