"""Per-role event log in the reference's result-file format (gpu/logger.cu:20-95,
gpu/logger.h:7-27), used by the SOT roles (silent.py) and the Ferret extend
loop (ferret.py) the way the reference's drivers use it
(ferret/emp-ot/test/ferret.cpp:61-66, ferret_cot.hpp:84-92).

File layout, one file per role (Sender = 0, Recver = 1):

    0 CudaInit                 one line per event: index and name
    1 BaseOT
    2 SeedExp
    3 LPN
    4 Neural
    --------------------
    s <event> <ms>             start, milliseconds since the log's time origin
    e <event> <ms>             end
    ...
    t <event> <ms>             at close: summed duration / sample size
    m <event> <bytes>          at close: largest device-memory growth inside the event

Times follow the reference's arithmetic: whole seconds times 1000 plus
nanoseconds / 1e6, held in 32-bit floats, printed with 6 significant digits
(the C++ stream default).  With ignore_init the time origin is the first start
of an event other than CudaInit, and CudaInit is not timed.  The memory figure
is device memory in use (total - free) at start, raised by `mem` calls.
The reference's `mem` reads an uninitialised free/total pair (its
cudaMemGetInfo call is commented out, logger.cu:91-92); here it queries the
device.
"""
from __future__ import annotations

import time

import numpy as np

CudaInit, BaseOT, SeedExp, LPN, Neural = range(5)
NUM_EVENTS = 5
EVENT_STRING = ("CudaInit", "BaseOT", "SeedExp", "LPN", "Neural")
Sender, Recver = 0, 1


def _device_used() -> int:
    try:
        import torch

        if torch.cuda.is_available():
            free, total = torch.cuda.mem_get_info()
            return int(total - free)
    except Exception:
        pass
    return 0


def _fmt(x) -> str:
    """A 32-bit float as std::ostream prints it by default (%g, precision 6)."""
    return "%g" % float(np.float32(x))


class Log:
    """Static per-role state, as the reference's class of static members."""

    clock_ns = staticmethod(time.monotonic_ns)  # CLOCK_MONOTONIC
    mem_used = staticmethod(_device_used)

    _file = [None, None]
    _init_ns = [0, 0]
    _event_start = np.zeros((2, NUM_EVENTS), dtype=np.float32)
    _event_dur = np.zeros((2, NUM_EVENTS), dtype=np.float32)
    _mem_start = np.zeros((2, NUM_EVENTS), dtype=np.uint64)
    _mem_curr = np.zeros((2, NUM_EVENTS), dtype=np.uint64)
    _mem_max = np.zeros((2, NUM_EVENTS), dtype=np.uint64)
    _opened = [False, False]
    _ignore_init = [False, False]
    _init_set = [False, False]
    _sample = 1

    @classmethod
    def _since_init(cls, role) -> np.float32:
        # timespec fields of both instants, then the reference's float sum
        now_s, now_ns = divmod(cls.clock_ns(), 1_000_000_000)
        ini_s, ini_ns = divmod(cls._init_ns[role], 1_000_000_000)
        t = np.float32((now_s - ini_s) * 1000)
        return np.float32(float(t) + (now_ns - ini_ns) / 1000000.0)

    @classmethod
    def open(cls, role: int, filename: str, ignore_init: bool, sample_size: int = 1):
        cls._sample = sample_size
        cls._opened[role] = True
        cls._ignore_init[role] = ignore_init
        f = open(filename, "w")
        cls._file[role] = f
        for i in range(NUM_EVENTS):
            f.write(f"{i} {EVENT_STRING[i]}\n")
        f.write("--------------------\n")
        if not ignore_init:
            cls._init_ns[role] = cls.clock_ns()
            cls._init_set[role] = True
        # the reference clears both roles' memory tables (logger.cu:36-37)
        cls._mem_curr[:] = 0
        cls._mem_max[:] = 0

    @classmethod
    def close(cls, role: int):
        if cls._file[role] is None:
            return
        cls._opened[role] = False
        f = cls._file[role]
        for ev in range(NUM_EVENTS):
            f.write(f"t {ev} {_fmt(cls._event_dur[role, ev] / np.float32(cls._sample))}\n")
        for ev in range(NUM_EVENTS):
            f.write(f"m {ev} {int(cls._mem_max[role, ev])}\n")
        f.close()
        cls._file[role] = None

    @classmethod
    def start(cls, role: int, event: int):
        if not cls._opened[role]:
            return
        if not cls._init_set[role] and cls._ignore_init[role] and event > CudaInit:
            cls._init_ns[role] = cls.clock_ns()
            cls._init_set[role] = True
        used = cls.mem_used()
        cls._mem_start[role, event] = used
        cls._mem_curr[role, event] = used
        if not cls._ignore_init[role] or event != CudaInit:
            t = cls._since_init(role)
            cls._file[role].write(f"s {event} {_fmt(t)}\n")
            cls._event_start[role, event] = t

    @classmethod
    def end(cls, role: int, event: int):
        if not cls._opened[role]:
            return
        if not cls._ignore_init[role] or event != CudaInit:
            t = cls._since_init(role)
            cls._file[role].write(f"e {event} {_fmt(t)}\n")
            cls._event_dur[role, event] = np.float32(cls._event_dur[role, event] + (t - cls._event_start[role, event]))
        grown = int(cls._mem_curr[role, event]) - int(cls._mem_start[role, event])
        cls._mem_max[role, event] = max(grown, int(cls._mem_max[role, event]))

    @classmethod
    def mem(cls, role: int, event: int):
        if not cls._opened[role]:
            return
        cls._mem_curr[role, event] = max(cls.mem_used(), int(cls._mem_curr[role, event]))

    @classmethod
    def is_open(cls, role: int) -> bool:
        return cls._opened[role]

    @classmethod
    def reset(cls):
        """Clears every table (the reference's statics live for the process)."""
        for r in (0, 1):
            if cls._file[r] is not None:
                cls._file[r].close()
            cls._file[r] = None
            cls._opened[r] = cls._ignore_init[r] = cls._init_set[r] = False
            cls._init_ns[r] = 0
        for a in (cls._event_start, cls._event_dur, cls._mem_start, cls._mem_curr, cls._mem_max):
            a[:] = 0
        cls._sample = 1
