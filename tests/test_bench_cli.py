"""bench.py's command line without a GPU: more than one GPU is refused with
a message (this build runs the whole mesh on one device); the algorithmic
byte model matches its written form."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_stage_bytes_model():
    import bench
    # 2D mixed with SA: 3*4*4 + 4*8 + 2*2*2 + 1 + (3*4 + 8 + 4) = 113 B per cell, 17 B per face
    assert bench.stage_bytes("mixed", 2, True, 1, 0) == 113
    assert bench.stage_bytes("mixed", 2, True, 0, 1) == 17
    # the last stage adds the FP64 masters: 4*8 + 8
    assert bench.stage_bytes("mixed", 2, True, 1, 0, last=True) == 113 + 40
    # 2D FP64 with SA: 96 + 32 + 8 + 1 + 24 + 8 + 8 = 177; fused first stage reads 16 B more per face
    assert bench.stage_bytes("fp64", 2, True, 1, 0) == 177
    assert bench.stage_bytes("fp64", 2, True, 0, 1, first=True) == 33
