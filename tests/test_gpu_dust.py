"""DBdust on the GPU (marvel_amd/bin/DBdust over mi_dust) writes the same
.dust.anno / .dust.data as the compiled reference DBdust for every golden
case: default and non-default window, threshold and minimum length, the
biased scores, a DB block, a DB that grew since its last scan, and the
5 Mbp / 30x read set.  The unbiased scan runs in segments with a warm-up
each; short segments without warm-up force most segment boundaries through
the re-run of the true state, and the tracks stay the same."""

import json
import os

import pytest

import dust_cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "marvel_amd", "bin", "DBdust")
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "dust", "goldens.json")))


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(dust_cases.CASES))
def test_gpu_dbdust_matches_reference(case, tmp_path):
    assert os.path.exists(EXE), "DBdust not built (python -m marvel_amd.build)"
    print(f"[gpu DBdust] case {case}", flush=True)
    assert dust_cases.run(EXE, case, str(tmp_path)) == GOLD[case]


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["rep_default", "rep_w100_t3", "rep_w256_t12_m4", "rep_grow", "edges_default",
                                  "c2_default"])
def test_gpu_dbdust_segment_rerun(case, tmp_path):
    assert os.path.exists(EXE), "DBdust not built (python -m marvel_amd.build)"
    print(f"[gpu DBdust] case {case}, 64-step segments, no warm-up", flush=True)
    log = []
    env = {"MI_DUST_SEGMENT": "64", "MI_DUST_WARMUP": "0", "MI_DUST_STATS": "1"}
    assert dust_cases.run(EXE, case, str(tmp_path), env=env, log=log) == GOLD[case]
    stats = [ln for ln in "".join(log).splitlines() if ln.startswith("mi_dust:")]
    assert stats, log
    rerun = int(stats[-1].split(",")[-1].split()[0])
    assert rerun > 0, stats


def test_dbdust_options_are_checked(tmp_path):
    import subprocess
    for args, msg in ((["-w0", "X"], b"Window must be positive"), (["-t0", "X"], b"Threshold must be positive"),
                      (["-m1", "X"], b"Minimum hit size must be positive"), ([], b"DB is required"),
                      (["-x", "X"], b"Unsupported option")):
        r = subprocess.run([EXE] + args, cwd=tmp_path, capture_output=True)
        assert r.returncode == 1 and msg in r.stderr, (args, r.stderr)
