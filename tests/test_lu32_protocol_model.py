"""The FP32 LU-SGS sweep's synchronisation, checked on a host model
(tests/lu32_protocol_model.cpp; CPU only).

The model runs the blocked, skewed sweep of k_lu_sweep32 with real threads
(one consumer and producers per workgroup, three streams sharing a limited
number of CU slots over several XCDs) and compares every result with a
sequential evaluation, bit for bit.  It contrasts the mailbox protocol the
kernel used before (a sentinel reset by a separate prep kernel, workgroup x
= block x) with the one it uses now (mail words tagged with the sweep number,
each 8-byte word written whole, blocks assigned by an atomic ticket):

* with the prep's reset written back late across the kernel boundary, the
  sentinel protocol takes a previous sweep's value as mail and the late
  reset overwrites fresh results with the NaN pattern -- a non-finite case,
  the symptom recorded for overlapped FP32 batches -- while the tagged
  protocol stays exact;
* with CU slots scarce, workgroups that wait on a block not yet dispatched
  hold the slots it needs (timeouts under the fixed block order); the
  ticket order never waits on a workgroup that has not started;
* under ThreadSanitizer the tagged protocol (LDS ring hand-off and mail) has
  no data race, and a ring refilled one step early is reported.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "lu32_protocol_model.cpp")


def _build(tmp_path_factory, tsan):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("lu32model") / ("model_tsan" if tsan else "model"))
    flags = ["-O1", "-g", "-fsanitize=thread"] if tsan else ["-O2"]
    subprocess.run(["g++", *flags, "-std=c++17", "-pthread", "-o", out, SRC], check=True)
    return out


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    return _build(tmp_path_factory, False)


@pytest.fixture(scope="module")
def model_tsan(tmp_path_factory):
    return _build(tmp_path_factory, True)


def _run(exe, *args, env=None):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=600, env=env)
    line = [x for x in r.stdout.splitlines() if x.startswith("protocol=")]
    counts = {}
    if line:
        for kv in line[-1].split():
            k, v = kv.split("=")
            counts[k] = v
    return r.returncode, counts, r.stdout + r.stderr


def test_tagged_protocol_exact_under_late_writeback(model):
    rc, c, out = _run(model, "--tagged", "--late-writeback", "--sweeps=40")
    assert rc == 0, out
    assert c["wrong"] == c["nan"] == c["stale_accepts"] == c["timeouts"] == "0"


def test_sentinel_protocol_fails_under_late_writeback(model):
    rc, c, out = _run(model, "--sentinel", "--late-writeback", "--sweeps=40")
    assert rc == 1, out
    assert int(c["stale_accepts"]) > 0 and int(c["nan"]) > 0, out


def test_sentinel_protocol_exact_when_reset_is_visible(model):
    """The control: with the reset written back before the sweep starts, the
    sentinel protocol is exact too (its failure needs the late reset)."""
    rc, c, out = _run(model, "--sentinel", "--sweeps=40")
    assert rc == 0, out


def test_fixed_block_order_stalls_when_slots_are_scarce(model):
    args = ["--xcds=3", "--slots=1", "--nk=15", "--sweeps=30", "--wait-ms=200"]
    rc, c, out = _run(model, "--sentinel", *args)
    assert rc == 1 and int(c["timeouts"]) > 0, out
    rc, c, out = _run(model, "--tagged", *args)
    assert rc == 0 and c["timeouts"] == "0", out


@pytest.mark.parametrize("args", [["--late-writeback"], ["--xcds=3", "--slots=1", "--nk=15"]])
def test_tagged_protocol_race_free_under_tsan(model_tsan, args):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    rc, c, out = _run(model_tsan, "--tagged", "--sweeps=10", *args, env=env)
    assert rc == 0 and "ThreadSanitizer" not in out, out


def test_tsan_reports_a_ring_slot_refilled_early(model_tsan):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    rc, c, out = _run(model_tsan, "--tagged", "--plant-ring-bug", "--sweeps=5", env=env)
    assert rc == 66 and "data race" in out, out
