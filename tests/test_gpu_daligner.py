"""GPU daligner (marvel_amd/bin/daligner over libmarvl_align) writes .las
files byte-identical to the compiled reference daligner's for every golden
case; on a mismatch the oracle restatement's output names the first
differing record."""

import hashlib
import json
import os
import subprocess
import tempfile

import pytest

import las_cases
import oracle_lib
from marvel_amd import dbio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "marvel_amd", "bin", "daligner")
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "las", "goldens.json")))


def _shas(d):
    out = {}
    for dp, _, fs in os.walk(d):
        for f in fs:
            if f.endswith(".las"):
                p = os.path.join(dp, f)
                out[os.path.relpath(p, d)] = hashlib.sha256(open(p, "rb").read()).hexdigest()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(las_cases.CASES))
def test_gpu_daligner_matches_reference(case):
    assert os.path.exists(EXE), "GPU daligner not built (python -m marvel_amd.build)"
    print(f"[gpu daligner] case {case}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        las_cases.make_db(case, d)
        _, _, flags, blocks = las_cases.CASES[case]
        r = subprocess.run([EXE] + flags + blocks, cwd=d, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stderr[-2000:]
        got = _shas(d)
        if got != GOLD[case]["las"]:
            oracle_lib.ensure_built()
            with tempfile.TemporaryDirectory() as d2:
                for f in os.listdir(d):
                    if f.endswith((".db", ".idx", ".bps")):
                        os.symlink(os.path.join(d, f), os.path.join(d2, f))
                subprocess.run([oracle_lib.ORACLE_CLI] + flags + blocks, cwd=d2, capture_output=True, timeout=600)
                diffs = {rel: dbio.first_difference(os.path.join(d, rel), os.path.join(d2, rel))
                         for rel in GOLD[case]["las"] if os.path.exists(os.path.join(d, rel))}
            pytest.fail(f"{case}: {got} != {GOLD[case]['las']}; first differences vs oracle: {diffs}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["c1_j8", "rep_j8", "rep_mdust_mrnd", "at_b", "c2_j8"])
def test_gpu_daligner_in_process(case, tmp_path):
    """Same comparison through the library loaded in this process (Python
    mirror of Set_Filter_Params / Sort_Kmers / Match_Filter /
    Write_Overlap_Buffer), with the block indexed from HBM-resident reads."""
    from marvel_amd.daligner import Daligner

    _, _, flags, blocks = las_cases.CASES[case]
    assert blocks[0] == blocks[1], "self comparisons only"
    opts = {f[1]: f[2:] for f in flags}
    masks = [f[2:] for f in flags if f.startswith("-m")]
    las_cases.make_db(case, str(tmp_path))
    dal = Daligner()
    nthreads = int(opts.get("j", "4"))
    dal.configure(nthreads=nthreads, mem_limit_gb=int(opts["M"]), minover=2 * int(opts.get("l", "1000")),
                  biased=int("b" in opts))
    dal.set_filter_params(int(opts.get("k", "14")), int(opts.get("w", "6")), int(opts.get("t", "0")),
                          int(opts.get("h", "35")))
    A = dal.open_block(str(tmp_path / blocks[0]))
    if masks:
        dal.load_mask_tracks(A, masks)   # the device block keeps the mask
    spec = dal.new_align_spec(float(opts.get("e", ".70")), int(opts.get("s", "100")), list(A.freq), nthreads=nthreads)
    dev = dal.block_device(A)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        dal.compare_self(A, spec, 1, blocks[0], dev=dev)
    finally:
        os.chdir(cwd)
        dal.free_index(dev)
    assert _shas(str(tmp_path)) == GOLD[case]["las"]


INTEG = os.path.join(ROOT, "oracle", "_ref", "daligner_on_libmarvl")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(INTEG), reason="integration binary not built (make -C oracle integ)")
@pytest.mark.parametrize("case", ["c1_j8", "c1_blocks_12", "rep_mdust_mrnd", "rep_mask_blocks", "at_b_mrnd", "srep_O",
                                  "c2_j8"])
def test_reference_driver_on_library(case):
    """The reference's own daligner.c and DB code, linked against
    libmarvl_align instead of its filter.c / align.c, writes the golden .las.
    With -m the reference's read_DB / complement_DB build the mask tracks and
    the library's Sort_Kmers reads them from HITS_DB.tracks."""
    print(f"[reference driver on libmarvl_align] case {case}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        las_cases.make_db(case, d)
        _, _, flags, blocks = las_cases.CASES[case]
        r = subprocess.run([INTEG] + flags + blocks, cwd=d, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stderr[-2000:]
        assert _shas(d) == GOLD[case]["las"]


@pytest.mark.gpu
def test_band_edge_cell_rerun(tmp_path):
    """Pairs in the outermost diagonal bands: an alignment in the lowest band
    raises the band-array cell that is also the highest band's lastp, and the
    pairs of the report chunk after it are re-run with the carried value
    (edge events and a fix-up must occur), with the reference's .las."""
    from marvel_amd.daligner import Daligner

    case = "edge_j1"
    _, _, flags, blocks = las_cases.CASES[case]
    print(f"[gpu daligner] {case} in process", flush=True)
    las_cases.make_db(case, str(tmp_path))
    dal = Daligner()
    dal.configure(nthreads=1, mem_limit_gb=16, minover=2000)
    dal.set_filter_params(14, 6, 0, 35)
    A = dal.open_block(str(tmp_path / blocks[0]))
    spec = dal.new_align_spec(0.70, 100, list(A.freq), nthreads=1)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        s1, s2 = dal.compare_self(A, spec, 1, blocks[0])
    finally:
        os.chdir(cwd)
    assert _shas(str(tmp_path)) == GOLD[case]["las"]
    assert s1["edge_events"] + s2["edge_events"] > 0
    assert s1["edge_fixups"] + s2["edge_fixups"] >= 1


@pytest.mark.gpu
def test_output_arena_overflow_reruns(tmp_path, monkeypatch):
    """A tiny initial output arena: pairs whose records do not fit are re-run
    after the arena grows, and only each pair's last complete run counts."""
    from marvel_amd.daligner import Daligner

    case = "c1_j8"
    _, _, flags, blocks = las_cases.CASES[case]
    print(f"[gpu daligner] {case} with a 64-record initial arena", flush=True)
    monkeypatch.setenv("MI_REPORT_ARENA", "64")
    las_cases.make_db(case, str(tmp_path))
    dal = Daligner()
    dal.configure(nthreads=8, mem_limit_gb=16, minover=2000)
    dal.set_filter_params(14, 6, 0, 35)
    A = dal.open_block(str(tmp_path / blocks[0]))
    spec = dal.new_align_spec(0.70, 100, list(A.freq), nthreads=8)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        s1, s2 = dal.compare_self(A, spec, 1, blocks[0])
    finally:
        os.chdir(cwd)
    assert s1["retries"] + s2["retries"] > 0
    assert s1["report_launches"] + s2["report_launches"] > 2
    assert _shas(str(tmp_path)) == GOLD[case]["las"]


@pytest.mark.gpu
@pytest.mark.parametrize("level", [1, 2, 3, 4])
def test_every_report_window_matches_reference(tmp_path, monkeypatch, level):
    """Every pair started at aligner window level 1-4 (128 / 256 / 1,024
    diagonals, then the large pools and arenas) writes the reference's .las:
    the wide report kernels, otherwise run only for the few pairs that
    outgrow the 64-lane window, on a whole workload."""
    from marvel_amd.daligner import Daligner

    case = "rep_j8"
    _, _, flags, blocks = las_cases.CASES[case]
    print(f"[gpu daligner] {case} from window level {level}", flush=True)
    monkeypatch.setenv("MI_REPORT_FIRST_LEVEL", str(level))
    las_cases.make_db(case, str(tmp_path))
    dal = Daligner()
    dal.configure(nthreads=8, mem_limit_gb=16, minover=2000)
    dal.set_filter_params(14, 6, 0, 35)
    A = dal.open_block(str(tmp_path / blocks[0]))
    spec = dal.new_align_spec(0.70, 100, list(A.freq), nthreads=8)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        dal.compare_self(A, spec, 1, blocks[0])
    finally:
        os.chdir(cwd)
    assert _shas(str(tmp_path)) == GOLD[case]["las"]


@pytest.mark.gpu
def test_kmer_keys_that_do_not_fit_are_an_error(tmp_path):
    """k = 20 on the small read set needs 2k + 25 = 65 key bits: the GPU
    daligner stops with a message instead of writing wrong output."""
    las_cases.make_db("c1_k19", str(tmp_path))
    r = subprocess.run([EXE, "-k20", "-j4", "-M16", "C1.1", "C1.1"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1 and "64-bit k-mer keys" in r.stderr, r.stderr[-1000:]


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["rep_j8"])
def test_redundancy_branches_fire(case, tmp_path, monkeypatch):
    """Handle_Redundancies (dalign/filter.c:1403-1566) takes its fusion and
    replacement branches on these goldens: the counters of the report
    kernel's lane-0 restatement are non-zero for fusions of an earlier-
    starting kept alignment with a later one reaching further (with and
    without COMP, i.e. in both passes) and for containment / equal starts,
    and the .las is still the reference's.  (The other order, a new
    alignment starting before a kept one that reaches further, needs a
    reverse wave running back past an earlier alignment's start; it did not
    occur on these inputs.  On the iid 5 Mbp reads the function fires once
    in a pass: repeats are what make alignments of one pair overlap.)"""
    from marvel_amd.daligner import Daligner

    monkeypatch.setenv("MI_REDUND_STATS", "1")
    print(f"[redundancy branches] case {case}", flush=True)
    _, _, flags, blocks = las_cases.CASES[case]
    opts = {f[1]: f[2:] for f in flags}
    las_cases.make_db(case, str(tmp_path))
    dal = Daligner()
    dal.configure(nthreads=int(opts["j"]), mem_limit_gb=int(opts["M"]), minover=2000)
    dal.set_filter_params(14, 6, 0, 35)
    A = dal.open_block(str(tmp_path / blocks[0]))
    spec = dal.new_align_spec(0.70, 100, list(A.freq), nthreads=int(opts["j"]))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        s1, s2 = dal.compare_self(A, spec, 1, blocks[0])
    finally:
        os.chdir(cwd)
    assert _shas(str(tmp_path)) == GOLD[case]["las"]
    br = [a + b for a, b in zip(s1["redund"], s2["redund"])]
    print(f"[redundancy branches] {case}: {br}", flush=True)
    assert br[6] + br[7] > 0, br   # k first, j reaches further: fused, A and B views
    assert br[4] > 0 and br[5] > 0, br   # j contains k; same start, j longer
    assert s1["redund"][7] > 0 and s2["redund"][6] > 0, (s1["redund"], s2["redund"])   # not COMP / COMP


def _in_process(dal, case, d, tuning=None):
    """One in-process self comparison of a golden case in directory d; returns (shas, stats)."""
    _, _, flags, blocks = las_cases.CASES[case]
    opts = {f[1]: f[2:] for f in flags}
    nthreads = int(opts.get("j", "4"))
    dal.configure(nthreads=nthreads, mem_limit_gb=int(opts["M"]), minover=2000)
    dal.set_filter_params(14, 6, 0, 35)
    A = dal.open_block(os.path.join(d, blocks[0]))
    spec = dal.new_align_spec(0.70, 100, list(A.freq), nthreads=nthreads)
    cwd = os.getcwd()
    os.chdir(d)
    try:
        for f in os.listdir(d):
            if f.startswith("001_"):
                for g in os.listdir(os.path.join(d, f)):
                    os.remove(os.path.join(d, f, g))
        stats = dal.compare_self(A, spec, 1, blocks[0])
    finally:
        os.chdir(cwd)
    return _shas(d), stats


@pytest.mark.gpu
def test_graph_replays_match_reference(tmp_path):
    """The same block pair three times in one process, another block pair in
    between: the first run records its radix sorts and scans as HIP graphs,
    the later ones replay them between the program's own eager work and
    synchronisations; every .las is the reference's."""
    from marvel_amd.daligner import Daligner

    print("[graphs] c1_j8 x3 with rep_j8 between", flush=True)
    d1, d2 = str(tmp_path / "c1"), str(tmp_path / "rep")
    las_cases.make_db("c1_j8", d1)
    las_cases.make_db("rep_j8", d2)
    dal = Daligner()
    replayed = []
    for i, case in enumerate(["c1_j8", "rep_j8", "c1_j8", "c1_j8"]):
        shas, (s1, s2) = _in_process(dal, case, d1 if case == "c1_j8" else d2)
        assert shas == GOLD[case]["las"], (i, case)
        replayed.append(s1["graphs_replayed"] + s2["graphs_replayed"])
        print(f"[graphs] run {i} {case}: captured {s1['graphs_captured'] + s2['graphs_captured']}, replayed "
              f"{replayed[-1]} ({s1['launches_replayed'] + s2['launches_replayed']} launches)", flush=True)
    assert replayed[2] > 0 and replayed[3] > 0


@pytest.mark.gpu
def test_tuning_keeps_output(tmp_path):
    """Launch shapes other than the defaults (fewer waves, smaller pebble
    pools: more re-runs at wider windows) leave the .las and the batched
    aligner's results unchanged."""
    import ctypes as C

    import numpy as np

    from marvel_amd.align import gen_tasks_device, la_batch_device
    from marvel_amd.daligner import Daligner
    from marvel_amd.lib import MiTuning, lib

    print("[tuning] c1_j8 and rep_j8 with small grids and pools", flush=True)
    dt = gen_tasks_device(3000, 7)
    ref = la_batch_device(dt)
    t = MiTuning(12, 13, 4, 8, 3)
    assert lib.mi_set_tuning(C.byref(t)) == 0
    try:
        dal = Daligner()
        for case in ("c1_j8", "rep_j8"):
            d = str(tmp_path / case)
            las_cases.make_db(case, d)
            shas, (s1, s2) = _in_process(dal, case, d)
            assert shas == GOLD[case]["las"], case
            print(f"[tuning] {case}: retries {s1['retries'] + s2['retries']}", flush=True)
        got = la_batch_device(dt)
    finally:
        lib.mi_set_tuning(None)
    a, b = ref.fields(), got.fields()
    for k in ("status", "abpos", "aepos", "bbpos", "bepos", "diffs", "atlen", "btlen"):
        assert np.array_equal(a[k], b[k]), k
    ta, tb = ref.traces(), got.traces()
    for i in range(dt.n):
        assert np.array_equal(ta[a["atrace"][i]:a["atrace"][i] + a["atlen"][i]],
                              tb[b["atrace"][i]:b["atrace"][i] + b["atlen"][i]]), i
