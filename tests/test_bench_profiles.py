"""bench.py's roofline line reads each dominant kernel's HBM bytes per launch from the committed
rocprofv3 summary: every kernel it can name must be found there (template kernels are listed
with their return type), with a traffic at least the kernel's algorithmic bytes."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402

CELLS = 512 * 256 * 256  # the profiled workload, taylor_green_512x256x256


@pytest.mark.parametrize("kernel", sorted(bench.PROFILE_NAMES))
def test_profile_traffic_found_for_every_named_kernel(kernel):
    t = bench.measured_traffic(kernel)
    assert t is not None, kernel
    algorithmic = bench.BYTES_PER_CELL[kernel] * CELLS
    # the counters see at least the bytes the kernel must move, and not several times more
    assert 0.99 * algorithmic <= t <= 2.0 * algorithmic, (kernel, t, algorithmic)
