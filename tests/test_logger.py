"""The event log (cuot_amd.logger) against the reference's result-file format
(gpu/logger.cu:20-95) on a scripted clock and memory counter: header, the
ignore_init rule (CudaInit untimed, time origin at the first other start),
float32 millisecond arithmetic printed with 6 significant digits, per-sample
durations and memory growth at close."""
import numpy as np

from cuot_amd import logger as L


class Script:
    def __init__(self, times_ns, mems):
        self.t = list(times_ns)
        self.m = list(mems)

    def clock(self):
        return self.t.pop(0)

    def mem(self):
        return self.m.pop(0)


def run(tmp_path, ignore_init, times, mems, calls, sample=1):
    sc = Script(times, mems)
    L.Log.reset()
    L.Log.clock_ns = staticmethod(sc.clock)
    L.Log.mem_used = staticmethod(sc.mem)
    try:
        f = tmp_path / "log.txt"
        L.Log.open(L.Sender, str(f), ignore_init, sample)
        for what, ev in calls:
            getattr(L.Log, what)(L.Sender, ev)
        L.Log.close(L.Sender)
        return f.read_text().splitlines()
    finally:
        L.Log.clock_ns = staticmethod(__import__("time").monotonic_ns)
        L.Log.mem_used = staticmethod(L._device_used)
        L.Log.reset()


HEADER = ["0 CudaInit", "1 BaseOT", "2 SeedExp", "3 LPN", "4 Neural", "--------------------"]


def test_ignore_init(tmp_path):
    s = 1_000_000_000
    # CudaInit start/end: memory read at start only, no clock read;
    # SeedExp start sets the origin (one clock read) then stamps 0
    times = [6 * s + 100_000_000, 6 * s + 100_000_000, 6 * s + 334_567_800, 7 * s, 7 * s + 500_000_000]
    mems = [10, 100, 180, 300]
    calls = [("start", L.CudaInit), ("end", L.CudaInit), ("start", L.SeedExp), ("mem", L.SeedExp),
             ("end", L.SeedExp), ("start", L.LPN), ("end", L.LPN)]
    lines = run(tmp_path, True, times, mems, calls, sample=2)
    assert lines == HEADER + ["s 2 0", "e 2 234.568", "s 3 900", "e 3 1400",
                              "t 0 0", "t 1 0", "t 2 117.284", "t 3 250", "t 4 0",
                              "m 0 0", "m 1 0", "m 2 80", "m 3 0", "m 4 0"]


def test_timed_init_and_large_values(tmp_path):
    s = 1_000_000_000
    # origin at open; the second instant is 1234.5678 s later (float32 ms)
    times = [5 * s + 900_000_000, 5 * s + 900_000_000 + 250_000_000, 1240 * s + 467_800_000]
    calls = [("start", L.CudaInit), ("end", L.CudaInit)]
    lines = run(tmp_path, False, times, [0], calls)
    assert lines[6:8] == ["s 0 250", "e 0 1.23457e+06"]
    assert lines[8] == "t 0 1.23432e+06"


def test_float32_ms():
    # whole seconds * 1000 as float, plus the nanosecond difference / 1e6
    L.Log.reset()
    L.Log._init_ns[0] = 5 * 1_000_000_000 + 900_000_000
    L.Log.clock_ns = staticmethod(lambda: 6 * 1_000_000_000 + 100_000_000)
    try:
        v = L.Log._since_init(0)
        assert isinstance(v, np.float32) and v == np.float32(200.0)
    finally:
        L.Log.clock_ns = staticmethod(__import__("time").monotonic_ns)
        L.Log.reset()
