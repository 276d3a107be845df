"""bench.py on the CPU: the sampling of HIP-event timing over the timed window, and the refusal
of more than one GPU (the engine runs one FID per GPU)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("steps,call_steps,every", [(40, 5, 4), (1000, 25, 4), (400, 25, 4),
                                                    (80, 5, 4), (40, 5, 3), (48, 5, 4)])
def test_sampled_calls_mirrored(steps, call_steps, every):
    n = -(-steps // call_steps)
    calls = bench.sampled_calls(steps, call_steps, every)
    assert calls == sorted(set(calls)) and all(0 <= c < n for c in calls)
    # one call in `every` (within one), and the sample's mean position is the window's centre
    assert abs(len(calls) - n / every) <= 1
    assert sum(calls) / len(calls) == pytest.approx((n - 1) / 2)
    # mirrored: c is sampled exactly when n - 1 - c is
    assert set(calls) == {n - 1 - c for c in calls}


def test_sampled_calls_headline_window():
    assert bench.sampled_calls(40, 5, 4) == [1, 6]


@pytest.mark.parametrize("steps,call_steps", [(5, 5), (10, 5), (20, 5), (35, 5), (3, 5)])
def test_short_windows_time_every_call(steps, call_steps):
    n = -(-steps // call_steps)
    assert bench.sampled_calls(steps, call_steps, 4) == list(range(n))


def test_every_call_when_asked():
    assert bench.sampled_calls(40, 5, 1) == list(range(8))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one FID per GPU" in msg["error"]


def _trace_csv(path, launches):
    with open(path, "w") as f:
        f.write('"Kind","Kernel_Name","Start_Timestamp","End_Timestamp"\n')
        for name, s, e in launches:
            f.write(f'"KERNEL_DISPATCH","void (anonymous namespace)::{name}((anonymous '
                    f'namespace)::BatchDev, (anonymous namespace)::PassMap const*)",{s},{e}\n')


def test_trace_summary_keeps_only_the_window(tmp_path):
    # two warmup rotations, then a window of three; each rotation: count then write
    launches, t = [], 0
    for r in range(5):
        dur = 1000 if r < 2 else 5000  # window launches are longer
        for name in ("k_rot_count<1, false>", "k_rot_write<1>"):
            launches.append((name, t, t + dur))
            t += dur + 10
    _trace_csv(tmp_path / "run_kernel_trace.csv", launches)
    md = bench.summarize_trace(str(tmp_path), rotations=3, steps=1)
    rows = {line.split("|")[1].strip(): line.split("|") for line in md.splitlines()
            if line.startswith("| k_")}
    assert set(rows) == {"k_rot_count<1, false>", "k_rot_write<1>"}
    for r in rows.values():
        assert int(r[2]) == 3 and float(r[4]) == pytest.approx(5.0)


def test_pmc_summary_averages_the_last_launches(tmp_path):
    with open(tmp_path / "run_counter_collection.csv", "w") as f:
        f.write('"Dispatch_Id","Kernel_Name","Counter_Name","Counter_Value"\n')
        for d in range(6):
            f.write(f'{d},"void k_rot_write<1>(BatchDev)","FETCH_SIZE",{100 if d < 2 else 400}\n')
            f.write(f'{d + 100},"void k_part_scan(BatchDev)","FETCH_SIZE",7\n')
    md = bench.summarize_pmc([str(tmp_path)], rotations=4)
    assert "| k_rot_write<1> | FETCH_SIZE | 4 | 400.0 |" in md
    assert "k_part_scan" not in md
