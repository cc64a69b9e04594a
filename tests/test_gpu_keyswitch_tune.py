"""The fused key-switch's tuned configuration: plan creation times the window
split, the K2 limb group and the K1/K3 tile (32 or 16 columns per workgroup)
once per (GPU architecture, N, L, dnum), appends them to the cache file
(HEAL_AMD_TUNE_CACHE) and the next process reads them back without timing.
Plans whose cached tile is 16 compute what the six-call chain computes, alone,
with the serving accumulate fused into their last kernel, and as members of a
group (whose kernels use 32-column tiles), at first/last-kernel windows of 6, 7
and 8 bits.  Each case runs in a child process
(keyswitch_tile_child.py): the cache file is read once per process."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import pytest
import torch

pytestmark = pytest.mark.gpu

CHILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "keyswitch_tile_child.py")


def _child(cache, mode):
    env = dict(os.environ, HEAL_AMD_TUNE_CACHE=str(cache))
    print(f"[tune] child {mode} with cache {cache}", flush=True)
    r = subprocess.run([sys.executable, "-u", CHILD, mode], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    return [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]


def _arch():
    return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]


def _lines(cache):
    return [line.split() for line in cache.read_text().splitlines() if line.strip()]


def test_tuned_choice_is_persisted_and_read_back(gpu, tmp_path):
    cache = tmp_path / "keyswitch_tune.txt"
    first = _child(cache, "tune")
    assert len(first) == 1
    cfg = first[0]["config"]
    assert cfg["tile"] in (16, 32) and cfg["low_bits"] == 6
    lines = _lines(cache)
    assert lines == [[_arch(), "4096", "3", "3", str(cfg["low_bits"]), str(cfg["limb_group"]), str(cfg["tile"])]]
    again = _child(cache, "tune")
    assert again[0]["config"] == cfg
    assert _lines(cache) == lines  # read back, not timed and appended again


def test_plans_with_sixteen_column_tiles(gpu, tmp_path):
    cache = tmp_path / "keyswitch_tune.txt"
    arch = _arch()
    # first a line of the earlier six-field form (tile 32), then seven-field lines
    rows = [f"{arch} 8192 4 3 6 1"]
    rows += [f"{arch} 4096 {k} 3 6 1 16" for k in range(1, 9)]
    rows.append(f"{arch} 65536 24 3 8 4 16")
    rows.append(f"{arch} 32768 16 3 8 4 16")  # 8 low bits: K1/K3 windows of 7 bits
    cache.write_text("\n".join(rows) + "\n")
    results = _child(cache, "check")
    assert [r["check"] for r in results] == ["fused", "fused", "fused", "rmac", "rmac", "rmac", "group"]
    for r in results:
        assert r["equal"], r
        if r["check"] == "group":
            assert r["tiles"] == [16], r
        else:
            assert r["config"]["tile"] == 16, r
            if r["n"] == 32768:
                assert r["config"]["low_bits"] == 8, r
    assert len(_lines(cache)) == len(rows)  # every plan found its shape in the file
