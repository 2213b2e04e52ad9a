"""The GRAPE-6 ABI from a compiled C caller (starlab_amd/abi_bench, built
from starlab_amd/csrc/abi_bench.c against include/grape6.h): kira's per-block
sequence -- set_j for the previous block, set_ti, firsthalf, lasthalf2 --
runs for several block sizes; the program checks pot < 0 and 0 <= nn < nj
for every i and exits non-zero otherwise."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_c_caller_block_sequence():
    exe = os.path.join(ROOT, "starlab_amd", "abi_bench")
    r = subprocess.run([exe, "8192", "40", "1", "17", "256", "2048"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [json.loads(line) for line in r.stdout.splitlines()]
    assert [row["ni"] for row in rows] == [1, 17, 256, 2048]
    for row in rows:
        assert row["nj"] == 8192 and row["blocks"] == 40
        assert 0 < row["us_per_block"] < 1e5 and row["pairs_per_s"] > 0
