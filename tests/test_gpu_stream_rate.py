"""kh_stream_rate: the plain one-pass stream the update kernels' measured
traffic is set against runs and reports a rate below the 8 TB/s peak for
every read/write mix it takes (256 MiB arrays: even one read and one write
exceed the 256 MB last-level cache, so the rate is HBM's)."""
import ctypes as C
import sys

import pytest

from khronos_hip import _lib

pytestmark = pytest.mark.gpu


def test_stream_rate_every_mix():
    lib = _lib.load()
    out = C.c_double()
    for nr in range(1, 7):
        for nw in range(1, 4):
            _lib.check(lib.kh_stream_rate(nr, nw, 64 * 1024 * 1024, 3, C.byref(out)))
            print(f"[stream rate] {nr} reads, {nw} writes, 256 MiB arrays: {out.value:.1f} GB/s", flush=True)
            sys.stdout.flush()
            assert 10.0 < out.value < 8000.0, (nr, nw, out.value)
