"""The server's completer cannot wait forever on a launch: its wait
(csrc/fa_serve_wait.hpp, built for the host with a stub event and a simulated
clock) ends at the limit when the launch never completes, on the poll that
finds it complete, and on a runtime error.  On a time-out the completer fails
the launch's and the class's queued requests with FA_EHIP (csrc/fa_serve.hip,
complete_loop), so fa_server_wait returns instead of blocking."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpu_build", "serve_wait_host.cpp")


@pytest.fixture(scope="module")
def sw(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("serve_wait") / "serve_wait.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC", "-o", out, SRC], check=True)
    L = C.CDLL(out)
    L.sw_run.restype = C.c_int
    L.sw_run.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_double, C.POINTER(C.c_uint64),
                         C.POINTER(C.c_double)]
    return L


def _run(L, ready_after, error_at, step, limit):
    polls, elapsed = C.c_uint64(), C.c_double()
    r = L.sw_run(ready_after, error_at, step, limit, C.byref(polls), C.byref(elapsed))
    return r, polls.value, elapsed.value


def test_a_launch_that_never_completes_times_out(sw):
    r, polls, elapsed = _run(sw, -1, -1, 0.001, 60.0)
    assert r == 2
    assert 60.0 < elapsed <= 60.0 + 0.001 + 1e-9
    assert polls == 60002            # one poll per 1 ms step, the last after the limit


def test_completion_ends_the_wait_at_once(sw):
    for ready in (0, 1, 7, 5000):
        r, polls, elapsed = _run(sw, ready, -1, 0.001, 60.0)
        assert r == 0 and polls == ready + 1


def test_a_runtime_error_ends_the_wait(sw):
    r, polls, _ = _run(sw, -1, 3, 0.001, 60.0)
    assert r == 1 and polls == 4


def test_the_completer_uses_the_bounded_wait():
    src = open(os.path.join(ROOT, "fast-atari-rs_amd", "csrc", "fa_serve.hip")).read()
    body = src[src.index("void complete_loop("):src.index("void finish_launch(LaunchCtx *ctx) {")]
    assert "fa::wait_launch(" in body and "hipEventSynchronize" not in body
    assert "WaitEnd::TimedOut" in body and "s->cv_done.notify_all()" in body
