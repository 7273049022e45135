"""The benchmark's contract on the CPU: more than one GPU is refused with a
message, and the HBM traffic it reports comes from the committed rocprofv3
summaries (FETCH_SIZE + WRITE_SIZE per call)."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_traffic_from_committed_profiles():
    bt = bench.measured_traffic("btree_b16_n2p30_u64_permute_traffic.txt")
    bp = bench.measured_traffic("batch_permutes_traffic.txt")
    # B-tree b=16 at 2^30 moves about 2.2 times its 17.18 GB of algorithmic
    # bytes, the batch's permutations about 2.2 times their 21.3 GB, the BST
    # about 2.05 passes and the mvEB about one
    assert bt is not None and 2.0 * 17.18e9 < bt < 2.5 * 17.18e9
    assert bp is not None and 1.8 * 21.3e9 < bp < 3.0 * 21.3e9
    bst = bench.layout_traffic("bst", 1 << 30)
    mveb = bench.layout_traffic("mveb", 1 << 30)
    assert bst is not None and 1.8 * 17.18e9 < bst < 2.4 * 17.18e9
    assert mveb is not None and 0.95 * 17.18e9 < mveb < 1.2 * 17.18e9
    assert bench.layout_traffic("mveb", 1 << 20) is None
    assert bench.measured_traffic("no_such_profile.txt") is None
