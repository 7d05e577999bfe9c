"""The library's serving scheduler, host side: admission by slot count and
footprint budget, cancellation, and polling with nothing submitted (no
device work).  Stepping through it is covered by tests/test_gpu_serving.py."""
import ctypes as C

from khronos_hip import _lib


def _server(budget, slots, chunk=16):
    lib = _lib.load()
    s = C.c_void_p()
    _lib.check(lib.kh_server_create(float(budget), slots, chunk, C.byref(s)))
    return lib, s


def _admit(lib, s, fp):
    a = C.c_int32()
    _lib.check(lib.kh_server_admit(s, float(fp), C.byref(a)))
    return a.value


def test_admission_budget_and_slots():
    lib, s = _server(100.0, 2)
    try:
        assert _admit(lib, s, 60) == 1
        assert _admit(lib, s, 50) == 0          # 110 > budget
        assert _admit(lib, s, 30) == 1
        assert _admit(lib, s, 5) == 0           # both slots taken
        _lib.check(lib.kh_server_cancel(s, 60.0))
        assert _admit(lib, s, 10) == 1
        done = (C.c_int64 * 4)()
        n, infl = C.c_int32(), C.c_int32()
        _lib.check(lib.kh_server_poll(s, done, 4, C.byref(n), C.byref(infl)))
        assert n.value == 0 and infl.value == 2
    finally:
        _lib.check(lib.kh_server_destroy(s))


def test_oversized_request_runs_alone_and_errors():
    lib, s = _server(100.0, 4)
    try:
        assert _admit(lib, s, 500) == 1         # nothing else in flight
        assert _admit(lib, s, 1) == 0           # the budget is exceeded
        assert lib.kh_server_submit(s, None, 10, 500.0, 1) != 0   # null context
        _lib.check(lib.kh_server_cancel(s, 500.0))
        assert lib.kh_server_cancel(s, 1.0) != 0                  # nothing reserved
        assert lib.kh_server_admit(s, -1.0, C.byref(C.c_int32())) != 0
    finally:
        _lib.check(lib.kh_server_destroy(s))
    bad = C.c_void_p()
    assert lib.kh_server_create(0.0, 1, 1, C.byref(bad)) != 0
    assert lib.kh_server_create(1.0, 0, 1, C.byref(bad)) != 0
