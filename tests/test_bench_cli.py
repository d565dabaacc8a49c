"""bench.py's contract pieces that need no GPU: it refuses more than one GPU
with a JSON message, and the committed counter summaries it reads parse into
the fields it uses."""

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_counter_summaries_parse():
    import bench
    tr = bench.traffic_from_summary()
    assert tr and tr["bytes_per_launch"] > 0
    assert abs(tr["read_bytes"] + tr["write_bytes"] - tr["bytes_per_launch"]) <= 2  # rounded separately
    vs = bench.traffic_from_summary(bench.VALU_SUMMARY)
    assert vs and vs["valu_insts_per_launch"] > 0 and vs["entries_per_launch"] > 0
    # 1024 SIMDs at 2.4 GHz, one wave64 VALU instruction per 4 cycles
    assert vs["peak_valu_insts_per_s"] == 1024 * 2.4e9 / 4
    for path in (bench.TRAFFIC_SUMMARY, bench.VALU_SUMMARY):
        text = open(path).read()
        assert "rocprofv3" in text and "bench.py --profile-run" in text
