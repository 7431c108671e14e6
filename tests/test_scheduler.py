"""CPU tests of the serving scheduler with a mock device (tests/sched_mock.cpp).

The scheduler (q-pso_amd/csrc/qpso_sched.h) is the policy part of the
library's serving loop: admission in arrival order under an HBM budget and a
particle-slot limit; two lanes of resident solves that requests join by size
(like sizes balanced over both, far lighter ones apart), each stepped one
iteration at a time as soon as it can take a step; results posted when a
solve's own last step completes; memory released in the order of the lane
that ran it; cancellation, time-outs and ticket bookkeeping.  The same
template runs here against a mock device with a virtual clock, a fake
allocator, injected launch failures and seeded Poisson arrivals, built with
AddressSanitizer and UndefinedBehaviorSanitizer (every scenario) and
ThreadSanitizer (the scenarios with a server thread and blocked waiters).
"""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "tests", "sched_mock.cpp")
INC = ["-I", os.path.join(REPO, "include"), "-I", os.path.join(REPO, "q-pso_amd", "csrc")]

SCENARIOS = ["fifo_and_budget", "budget_refusal", "table_charged_once", "slots_and_oversize", "iterations_exact",
             "cancel_queued_and_resident", "timeouts", "create_and_init_failures", "step_failure_ends_one_lane",
             "poisson_stream", "light_lane_latency", "like_sizes_balanced", "interleaved_mix", "freed_bytes_admit",
             "threaded_shutdown_releases_waiters"]


def _build(tmp, name, flags):
    exe = os.path.join(tmp, name)
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", *flags, *INC, "-o", exe, SRC,
                           "-lpthread"])
    return exe


@pytest.fixture(scope="module")
def asan_exe(tmp_path_factory):
    return _build(str(tmp_path_factory.mktemp("sched")), "sched_asan",
                  ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"])


@pytest.mark.parametrize("scenario", SCENARIOS)
def test_scheduler_scenario(asan_exe, scenario):
    p = subprocess.run([asan_exe, scenario], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0 and f"PASS {scenario}" in p.stdout, p.stdout + p.stderr


def test_scheduler_threads_under_tsan(tmp_path):
    exe = _build(str(tmp_path), "sched_tsan", ["-fsanitize=thread"])
    for scenario in ("threaded_shutdown_releases_waiters", "cancel_queued_and_resident"):
        p = subprocess.run([exe, scenario], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0 and f"PASS {scenario}" in p.stdout, p.stdout + p.stderr
        assert "ThreadSanitizer" not in p.stderr, p.stderr
