"""bench.py's command line on the CPU: the multi-GPU request is refused with
a message, and tree splits parse as F forests on S streams."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr and r.stdout == ""


def test_tree_splits():
    import bench
    assert bench.parse_tree_splits("4,8") == ((4, 4), (8, 8))
    assert bench.parse_tree_splits("4x2, 1, 2") == ((4, 2), (2, 2))
    assert bench.parse_tree_splits("") == ()
    with pytest.raises(ValueError):
        bench.parse_tree_splits("4x0")
