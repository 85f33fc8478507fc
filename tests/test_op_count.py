"""The operation count bench.py prices the VALU rate with is the count that
oracle/count_ops.cpp measures through the oracle restatement."""
import os
import re
import subprocess

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_op_count_matches_counted_oracle():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "count"])
    out = subprocess.check_output([os.path.join(ROOT, "oracle", "count_ops")], text=True)
    per_face = int(re.search(r"per interface \(one WENO\) (\d+)", out).group(1))
    per_cell = int(re.search(r"algorithmic per cell-update (\d+)", out).group(1))
    as_written = float(re.search(r"as written per cell-update (\d+)", out).group(1))
    assert bench.ALG_OPS_PER_FACE == per_face
    assert bench.ALG_OPS_PER_CELL_UPDATE == per_cell
    # the reference evaluates five WENO reconstructions per interface
    assert 1.4 * per_cell < as_written < 1.8 * per_cell
