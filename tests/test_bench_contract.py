"""CPU checks of bench.py's contract: a multi-GPU request is refused with a
message before anything touches a GPU, the roofline uses the fixed
152-byte-per-cell-step accounting against the 8 TB/s HBM peak, and the
committed counter summary (rocprofv3 FETCH_SIZE / WRITE_SIZE /
SQ_INSTS_VALU) is consistent with its own calibration factors."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_roofline_accounting():
    sys.path.insert(0, ROOT)
    import bench

    # state: read Q^n + write Q1; read Q1, Q^n + write Q2; read Q2, Q^n + write
    # Q^n+1 -- eight 16-byte streams; plus one 8-byte vertex per stage
    assert bench.ALG_BYTES_PER_CELL_STEP == 8 * 16 + 3 * 8 == 152
    assert bench.HBM_PEAK_GBPS == 8000.0


def test_counter_summary_is_consistent():
    with open(os.path.join(ROOT, "profiles", "stage_kernel_counters.json")) as f:
        c = json.load(f)
    cal = c["calibration"]
    for math_mode in ("exact", "fast"):
        t = c["per_math"][math_mode]
        fetch = t["fetch_kib_raw"] * 1024 * cal["fetch_scale"]
        write = t["write_kib_raw"] * 1024 * cal["write_scale"]
        assert abs(t["traffic_bytes_per_launch"] - (fetch + write)) <= 2
        # the per-stage medians average to the raw figures
        assert abs(sum(t["per_stage_fetch_kib"]) / 3 - t["fetch_kib_raw"]) < 0.01
        assert abs(sum(t["per_stage_write_kib"]) / 3 - t["write_kib_raw"]) < 0.01
        assert abs(sum(t["per_stage_valu_insts"]) / 3 - t["valu_insts_per_launch"]) <= 1
        # HBM-side bytes of a 512 x 384 stage launch against its algorithmic bytes
        alg = 152.0 / 3 * 512 * 384
        assert 0.8 < t["traffic_bytes_per_launch"] / alg < 1.5
        # every wavefront of the launch issues the stage's instructions
        assert t["waves_per_launch"] == 1728
    # exact math issues more instructions than fast math for the same work
    assert c["per_math"]["exact"]["valu_insts_per_launch"] > c["per_math"]["fast"]["valu_insts_per_launch"]
