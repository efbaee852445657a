"""bench.py's contract, checked on the CPU: it measures one GPU (a request
for more exits with status 2 and says why), and the algorithmic bytes per
lattice update behind its roofline."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("argv,env", [(["--gpus", "2"], {}), ([], {"WORLD_SIZE": "4"})],
                         ids=["gpus2", "world4"])
def test_more_than_one_gpu_is_refused(argv, env):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *argv],
                       capture_output=True, text=True, timeout=300,
                       env={**os.environ, **env})
    assert r.returncode == 2, r.stderr
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_bytes_per_lattice_update():
    import bench
    # τ = 1: read h, ux, uy; write 9 f, h, ux, uy, p (16 words)
    assert bench.bytes_per_lup(4, True) == 64
    assert bench.bytes_per_lup(8, True) == 128
    # τ ≠ 1 also reads the 9 populations f_t (25 words)
    assert bench.bytes_per_lup(4, False) == 100
    assert bench.bytes_per_lup(8, False) == 200


def test_workloads_are_named_with_their_shapes():
    import bench
    assert bench.WORKLOADS["rupture4096_f32"] == (4096, 4096, "f32", True)
    assert set(bench.WORKLOADS).isdisjoint(bench.BATCH_WORKLOADS)
