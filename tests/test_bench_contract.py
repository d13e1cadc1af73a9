"""CPU checks of bench.py's reporting: the roofline's kernel is the one the
committed rocprofv3 summaries name, and more than one GPU is refused."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_roofline_kernel_is_in_the_committed_profiles():
    # the HBM-traffic summary holds calibrated read + write bytes for it
    traffic = bench.profiled_traffic(bench.EVAL_KERNEL)
    assert traffic is not None and traffic > 0
    # the kernel trace summary times the same kernel
    trace = open(os.path.join(ROOT, "profiles", "bench_3d256_k3_kernel_trace.md")).read()
    assert "`%s`" % bench.EVAL_KERNEL in trace


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in r.stderr + r.stdout
