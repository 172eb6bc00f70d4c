"""bench.py's contract pieces that need no GPU, and the consistency of the
measurement files it and the profiles hold:
- more than one GPU is refused with a message, before anything touches a device;
- the committed HBM traffic figure parses and lies near the fused kernel's
  algorithmic 52 bytes per cell of the SEAM-shaped grid;
- the committed batch and edges lines agree with themselves."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (module import only: no device work at import time)

SEAM_CELLS = 4552 * 3927


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120,
                       env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_committed_traffic_is_near_the_algorithmic_bytes():
    lines = [l for l in open(bench.TRAFFIC_FILE) if l.startswith("traffic_bytes_per_launch:")]
    assert len(lines) == 1
    traffic = int(lines[0].split(":")[1])
    per_cell = traffic / SEAM_CELLS
    # the counters see what the caches let through: a little under the 52
    # algorithmic bytes (halo rows hit in L2), never a re-read of whole fields
    assert 0.85 * bench.BYTES_FUSED_DERIVED < per_cell < 1.1 * bench.BYTES_FUSED_DERIVED


def test_committed_double_step_traffic_is_near_its_algorithmic_bytes():
    # the two-step interior launch reads 5 fields and 5 stored coefficient
    # arrays and writes 5 fields once per interior cell (60 B); its tiles
    # re-read halo columns and rows, but not a whole field
    lines = [l for l in open(bench.TRAFFIC_FILE_DOUBLE) if l.startswith("traffic_bytes_per_launch:")]
    assert len(lines) == 1
    per_cell = int(lines[0].split(":")[1]) / bench.split_interior_cells(4552, 3927, 50, 4)
    assert 60.0 < per_cell < 1.25 * 60.0


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_batch_line_is_self_consistent(storage):
    d = json.load(open(os.path.join(ROOT, "profiles", "batch_marmousi64_bench.json")))[storage]
    shots = int(d["config"]["workload"].split()[0])
    seconds = d["ms_per_step"] * d["steps"] / 1e3
    assert d["value"] == pytest.approx(shots / seconds * 3600.0, rel=2e-3)
    assert d["mcell_updates_per_s"] == pytest.approx(
        d["config"]["padded_cells_per_shot"] * shots / (d["ms_per_step"] / 1e3) / 1e6, rel=2e-3)
    assert d["storage"] == storage and d["gathers_finite"]


def test_storage_byte_counts_follow_the_array_accesses():
    # 15 array accesses per cell (read 5 fields + 5 coefficients, write 5 fields)
    width = {"f32": (4, 4), "bf16": (2, 2), "f16": (2, 2), "f16c": (4, 2)}  # (field, coefficient) bytes
    for st, (fb, cb) in width.items():
        assert bench.BYTES_FUSED_STORAGE[st] == 10 * fb + 5 * cb
    assert bench.BYTES_FUSED_DERIVED == 10 * 4 + 3 * 4


def test_edges_line_is_self_consistent():
    d = json.load(open(os.path.join(ROOT, "profiles", "edges_seam_bench.json")))
    s = d["setup_seconds"]
    assert d["setup_speedup"] == pytest.approx(s["host"] / s["device"], rel=2e-2)
    for key in ("frames", "frames_double"):
        f = d[key]
        assert f["overhead"] == pytest.approx(f["ms_per_step_with_frames"] / f["ms_per_step_plain"] - 1, abs=2e-3)
        assert f["frame_shape"] == [-(-3927 // f["downsample"]), -(-4552 // f["downsample"])]
        assert f["last_frame_finite_nonzero"]
    # the double step takes its frames inside double steps only for an even skip
    assert d["frames_double"]["skip"] % 2 == 0


def test_timed_kernels_run_at_a_step_inside_their_shot():
    """Every per-kernel timing launch runs at a step k with 1 <= k <= nt of
    the shot it runs on (a kernel launched past its shot's last step returns
    at once), and the autotuner's warm-up and timed steps fit the timing shot."""
    from fomo_julia_amd import tune as T
    for nt in (1, 2, 5, 20, 31, 32, 33, 64, 11520):
        k = bench.stage_step(nt)
        assert 1 <= k <= nt
    assert bench.stage_step(bench.TIMING_NT) == bench.STAGE_STEP
    with pytest.raises(ValueError):
        bench.stage_step(0)
    # the autotuner's defaults (4 warm-up steps, 32 timed) inside the timing shot
    steps = T.trial_steps(bench.TIMING_NT, 32, 4)
    assert steps == 32 and 4 + steps <= bench.TIMING_NT
    # short members: the trial shrinks to what the shots hold, or is refused
    assert T.trial_steps(20, 32, 4) == 16
    assert T.trial_steps(9, 32, 4) == 4
    for nt in (1, 4, 5):
        with pytest.raises(ValueError):
            T.trial_steps(nt, 32, 4)
    for nt in range(6, 200):
        s = T.trial_steps(nt, 32, 4)
        assert s >= 2 and s % 2 == 0 and 4 + s <= nt


def test_roofline_check_refuses_idle_kernels():
    bench.check_roofline(0.5, 0.233, 0.2955)
    bench.check_roofline(1.0, 0.2, 0.2)
    for frac, kms, step in ((47.5, 0.00245, 0.3018), (0.0, 0.2, 0.3), (-0.1, 0.2, 0.3),
                            (0.5, 0.31, 0.30), (0.5, 0.0, 0.3)):
        with pytest.raises(SystemExit):
            bench.check_roofline(frac, kms, step)
    # the double step's interior launch advances two steps: up to two step times
    bench.check_roofline(0.5, 0.37, 0.193, 2)
    with pytest.raises(SystemExit):
        bench.check_roofline(0.5, 0.39, 0.193, 2)
    with pytest.raises(SystemExit):
        bench.check_roofline(0.5, 0.37, 0.193)


def test_cpu_share_is_the_usable_cpus():
    s = bench.cpu_share()
    assert 1 <= s["usable"] <= s["affinity"] <= (s["host_cpus"] or s["affinity"])
    if s["cgroup_quota"] is not None:
        assert s["usable"] <= max(1, int(s["cgroup_quota"] + 0.999))
