"""A C++ program compiled against include/*.h only (tests/abi_consumer/consumer.cpp,
built by __graft_entry__.build()) drives the simulator and the drop-in launchers
through the C ABI, as the reference's host code would after the switch."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi_consumer", "consumer")


def test_cpp_consumer_runs(dev):
    assert os.path.exists(BIN), f"{BIN} is missing: run __graft_entry__.build()"
    print("running", BIN, flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.strip().splitlines()[-1].startswith("OK"), r.stdout + r.stderr
