"""bench.py's host-side pieces (no GPU): the --gpus guard and the traffic
figures read from the committed rocprofv3 summary."""
import subprocess
import sys

import bench


def test_multi_gpu_request_exits_with_a_message():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], capture_output=True, text=True,
                       cwd=bench.ROOT)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_traffic_from_committed_profile_covers_every_storage():
    """The summary covers the kernels each storage's level starts on (256x1
    blocks for FP64, 64x2 for shifted FP32, the FP32-arithmetic collide
    for scaled FP16)."""
    cells = 256 ** 3
    t64 = bench.measured_traffic("f64", ["k_qpass_x256", "k_collide_x256"])
    t32 = bench.measured_traffic("f32s", ["k_qpass_y2", "k_collide_y2"])
    t16 = bench.measured_traffic("f16s", ["k_qpass", "k_collide_h32"])
    # at least the algorithmic bytes (read + write of every population once)
    assert t64 >= 864 * cells and t32 >= 432 * cells and t16 >= 216 * cells
    assert bench.measured_traffic("f64", ["k_not_profiled"]) is None
