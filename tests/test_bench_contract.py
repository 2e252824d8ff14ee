"""bench.py's command-line contract and the profile figures its JSON line
carries (CPU only: no compute)."""

import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_profiled_traffic_reads_the_committed_counters():
    b = _bench()
    rep = b.profiled_traffic("k_report<1, false>")
    hist, scat = b.profiled_traffic("rs_hist<Seed>"), b.profiled_traffic("rs_scatter<Seed>")
    # calibrated bytes per launch: 2 x FETCH_SIZE KiB + WRITE_SIZE KiB
    assert rep is not None and 5e9 < rep < 5e10
    # a seed-sort pass reads and writes at least the 71 M x 16-byte seeds once
    assert hist is not None and scat is not None and scat > 2 * 71e6 * 16 and hist > 71e6 * 16
    assert b.profiled_traffic("no_such_kernel") is None


def test_profiled_valu_per_wave_step_reads_the_instruction_counters():
    b = _bench()
    v = b.profiled_valu_per_wave_step()
    # SQ_INSTS_VALU per launch over the launch's 2.7e8 aligner wave steps: the committed summary's column
    assert v is not None and 50 < v < 1000
    # the issue ceiling the bench line compares against: one wave64 instruction per CU per cycle
    assert abs(b.VALU_ISSUE_PEAK_G - 256 * 2.4) < 1e-9
