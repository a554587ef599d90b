"""bench.py's accounting, without a GPU: the per-kernel canonical bytes add
up to the plain substep's B(N, C, R) = N (508 + 136 R) + 16 C (integrate takes
the grid count), and to N (524 + 136 R) + 16 C with grid_count and integrate's
own 120 bytes (a batch); the committed counter summary names every kernel
class; more than one GPU is refused."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

SUBSTEP_ONCE = ("grid_scatter", "density_div", "div_apply", "forces", "integrate")
PER_PASS = ("den_kappa", "den_apply")


@pytest.mark.parametrize("refine", [1, 2, 5, 8])
def test_kernel_bytes_add_up_to_the_substep(refine):
    n, cells = 1_048_576, 357_911
    total = sum(bench.kernel_bytes(k, n, cells) for k in SUBSTEP_ONCE)
    total += refine * sum(bench.kernel_bytes(k, n, cells) for k in PER_PASS)
    total += bench.kernel_bytes("grid_scan", n, cells)
    assert total == bench.canonical_bytes(n, cells, refine)
    unfused = total - bench.kernel_bytes("integrate", n, cells) + bench.INTEGRATE_UNFUSED * n
    unfused += bench.kernel_bytes("grid_count", n, cells)
    assert unfused == bench.canonical_bytes(n, cells, refine, count_fused=False)


def test_the_counter_summary_covers_every_kernel_class():
    traffic = bench.load_traffic()
    for k in SUBSTEP_ONCE + PER_PASS + ("grid_scan",):
        assert traffic.get(k, 0) > 0, k
    # the list sweeps move more than their canonical bytes (the list itself)
    for k in PER_PASS:
        assert traffic[k] > bench.kernel_bytes(k, 1_048_576, 0)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
