#!/usr/bin/env python3
"""Benchmark of the GPU daligner overlap path (one MI355X).

Headline workload: the 5 Mbp / 30x configuration -- a 5 Mbp iid genome,
reads at 30x with lognormal lengths (mean 10 kb, sd 6 kb, min 1 kb) and
9/4/2 % insertion/deletion/substitution errors (seed 2), stored as one
~158 Mbp DB block, compared against itself with daligner -k14 -w6 -h35 -e.70
-l1000 -s100 -j8 -M32 (both strands).  Synthetic reads, regenerated from the
seed.

A step is one block-vs-itself comparison with the reads already resident in
HBM: k-mer index build and its reverse-complement index, both Match_Filter
passes (merge-join, seed sort, report + alignment), the records back on the
host, the stable sort and the .las write.  The produced .las is checked
byte for byte against the SHA-256 of the reference daligner's output; on a
mismatch the line carries value null and the script exits with status 3.

Further legs on the same line:
  * cpu_baseline: the reference daligner (compiled from source) on the same
    block with -j the largest power of two <= the host's CPUs;
  * pcie_inclusive: one step whose index build uploads the block from host
    memory (the headline keeps the reads in HBM);
  * batched_aligner: 2,000,000 Local_Alignment tasks generated in HBM
    (lengths log-uniform 1-60 kb, pair error 5-20 %, seeds up to 200 bp off
    the true diagonal) aligned from one device queue;
  * scorer_gated: the repeat-rich 50 Mbp genome's first 200 Mbp block
    against itself, plain GPU path vs the learned scorer gating it (shipped
    weights, trained on other data): overlaps/s of both and record recall.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-gated] [--no-batch]
"""

from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import os
import shutil
import struct
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "overlaps/s & aligned Gbp/s, bit-exact .las; scorer-gated at >=99.5% recall"
FLAGS = ["-k14", "-w6", "-h35", "-e.70", "-l1000", "-s100", "-j8", "-M32"]
GOLD_SHA = "2b379780f4257dd29f63fb4318baaa11b9d53ea3bf98de09e580c0401e0e363f"   # reference daligner, C2.1 x C2.1
HBM_PEAK_GBS = 8000.0
# int32 VALU lane-operations: 256 CUs x 64 lanes x 2.4 GHz
VALU_PEAK_GOPS = 256 * 64 * 2.4
# wave64 VALU instruction issue: one per CU per cycle (each SIMD one every 4 cycles), G instructions/s
VALU_ISSUE_PEAK_G = 256 * 2.4
GATED_FLAGS = ["-k14", "-w6", "-h35", "-e.70", "-l1000", "-s100", "-j8", "-M32"]
GATED_SHA = "8fb6b42443d1ed9a54b11316ad6e0d328cfce06719f0b89aad1a114ad94a9bd1"   # reference daligner, R.1 x R.1


def las_summary(path):
    """(records, Σ(aepos-abpos) over A-perspective records) of a self-comparison .las."""
    with open(path, "rb") as f:
        raw = f.read()
    n, tspace = struct.unpack_from("<qi", raw, 0)
    tb = 1 if tspace <= 125 else 2
    o, bases = 12, 0
    rec = struct.Struct("<iiiiiiIiii")
    for _ in range(n):
        tlen, _, ab, _, ae, _, _, ar, br, _ = rec.unpack_from(raw, o)
        if ar > br:   # the pair seeded with A = the higher read (ovla of report_thread)
            bases += ae - ab
        o += 40 + tlen * tb
    return n, bases, hashlib.sha256(raw).hexdigest()


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary (FETCH_SIZE / WRITE_SIZE in
    KiB, scaled by the factors a plain 1 GiB copy gives).  None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_report_pmc_5mbp.md")
    if not os.path.exists(path):
        return None
    fetch_scale = write_scale = None
    row = None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 3 and cells[0] == "FETCH_SIZE (KiB)":
            fetch_scale = float(cells[2].split()[0])
        elif len(cells) >= 3 and cells[0] == "WRITE_SIZE (KiB)":
            write_scale = float(cells[2].split()[0])
        elif len(cells) >= 3 and cells[0].startswith(f"`{kernel}`") and row is None:
            row = cells
    if row is None or fetch_scale is None or write_scale is None:
        return None
    kib = lambda c: float(c.replace(",", ""))
    return int((kib(row[1]) * fetch_scale + kib(row[2]) * write_scale) * 1024)


def profiled_valu_per_wave_step():
    """VALU instructions per aligner wave step of k_report<1, false> from the committed counter summary
    (SQ_INSTS_VALU per launch over the launch's wave steps).  None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_report_pmc_5mbp.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 3 and cells[0] == "SQ_INSTS_VALU":
            try:
                return float(cells[2].split()[0])
            except ValueError:
                return None
    return None


def run_steps(dal, A, spec, dev, cA, d, name, steps, warmup):
    """`warmup` untimed then `steps` timed block-vs-itself comparisons (reads
    resident in HBM); returns (seconds per step, stats of the last step)."""
    cwd = os.getcwd()
    os.chdir(d)
    try:
        for _ in range(warmup):
            dal.compare_self(A, spec, 1, name, dev=dev, cA=cA)
        dal.flush_writes()
        dal.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            stats = dal.compare_self(A, spec, 1, name, dev=dev, cA=cA, wait=False)
        dal.flush_writes()
        dal.synchronize()
        return (time.perf_counter() - t0) / steps, stats
    finally:
        os.chdir(cwd)


def gated_leg(steps):
    """The repeat-rich workload (first 200 Mbp block of the 50 Mbp genome with
    200 repeat families and 1 % tandems, 30x at 12 % error) against itself:
    the plain GPU path (checked against the reference's .las) and the same
    steps with the learned scorer dropping candidate pairs before alignment."""
    from marvel_amd import dbio, sim
    from marvel_amd import scorer as S
    from marvel_amd.daligner import Daligner
    if not os.path.exists(S.DEFAULT_WEIGHTS):
        return None
    W = S.ScorerWeights.load()
    d = workdir()
    try:
        t0 = time.time()
        reads = sim.config_repeats_50mbp(4).reads
        dbio.write_db(os.path.join(d, "R"), reads, size_mbp=200)
        t_db = time.time() - t0
        dal = Daligner()
        dal.configure(nthreads=8, mem_limit_gb=32, minover=2000, symmetric=1)
        dal.set_filter_params(14, 6, 0, 35)
        A = dal.open_block(os.path.join(d, "R.1"))
        spec = dal.new_align_spec(0.70, 100, list(A.freq), nthreads=8)
        dev = dal.block_device(A)
        cA = dal.complement_header(A)
        las = os.path.join(d, "001_00001", "R.1.R.1.las")
        t_plain, s_plain = run_steps(dal, A, spec, dev, cA, d, "R.1", steps, 1)
        n_plain, bases_plain, sha = las_summary(las)
        ref = os.path.join(d, "exact.las")
        shutil.copyfile(las, ref)
        sc = S.GpuScorer(dal.h)
        # the most any pair-level pruning can save at full recall: the aligner
        # work of the candidate pairs whose exact run writes no record
        sc.mode(1)
        try:
            run_steps(dal, A, spec, dev, cA, d, "R.1", 1, 0)
        finally:
            sc.mode(0)
        _, nrec, work = sc.take()
        work = work.astype("float64")
        share = float(work[nrec == 0].sum() / max(work.sum(), 1.0))
        sc.load(W)
        sc.mode(2)
        try:
            t_gate, s_gate = run_steps(dal, A, spec, dev, cA, d, "R.1", steps, 1)
        finally:
            sc.mode(0)
        n_gate, bases_gate, _ = las_summary(las)
        rec, hit, tot = S.record_recall(ref, las)
        return {"workload": "50 Mbp genome + 200 repeat families x 40 copies (6 kb, 10 %) + 1 % tandems, 30x at "
                            "12 % error; first 200 Mbp block vs itself, both strands",
                "reads": len(reads), "bases": sum(len(r) for r in reads), "daligner": " ".join(GATED_FLAGS),
                "plain": {"value": round(n_plain / t_plain, 1), "unit": "overlaps/s",
                          "ms_per_step": round(t_plain * 1e3, 2), "records": n_plain,
                          "aligned_gbp_per_s": round(bases_plain / t_plain / 1e9, 4),
                          "parity": "bit-exact vs reference daligner" if sha == GATED_SHA else f"MISMATCH {sha}",
                          "report_kernel_ms": round(sum(x["t_report_kernel"] for x in s_plain) * 1e3, 2)},
                "gated": {"value": round(n_gate / t_gate, 1), "unit": "overlaps/s",
                          "ms_per_step": round(t_gate * 1e3, 2), "records": n_gate,
                          "report_kernel_ms": round(sum(x["t_report_kernel"] for x in s_gate) * 1e3, 2),
                          "pairs_scored": sum(x["cand_scored"] for x in s_gate),
                          "pairs_dropped": sum(x["cand_skipped"] for x in s_gate)},
                "speedup": round(t_plain / t_gate, 4),
                "aligner_work_without_records": round(share, 5),
                "pruning_ceiling": round(1.0 / max(1e-9, 1.0 - share), 4),
                "candidate_pairs": int(len(nrec)), "candidate_pairs_without_records": int((nrec == 0).sum()),
                "record_recall": round(rec, 6), "records_exact": tot, "records_of_exact_kept": hit,
                "threshold": W.threshold, "trained_on": W.meta.get("configs"), "setup_s": round(t_db, 2)}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def batched_leg(n=2_000_000, seed=3):
    """n Local_Alignment tasks generated in HBM and aligned from one device
    queue (mi_la_batch_device); a warm-up batch, then one timed batch."""
    import numpy as np
    from marvel_amd.align import gen_tasks_device, la_batch_device
    t0 = time.perf_counter()
    dt = gen_tasks_device(n, seed)
    t_gen = time.perf_counter() - t0
    la_batch_device(dt)   # warm-up: grows the workspace, loads the kernels
    t0 = time.perf_counter()
    r = la_batch_device(dt)
    t = time.perf_counter() - t0
    res = r.fields()
    ok = res["status"] == 0
    w_alg = int(res["cells"][ok].sum() + res["snake"][ok].sum())   # the final run of every aligned task
    kern_s = sum(r.work[2:6]) / 1e6
    ach = w_alg / kern_s / 1e9 if kern_s > 0 else 0.0
    st = np.bincount(res["status"], minlength=8)
    return {"workload": f"{n} tasks: read pairs log-uniform 1-60 kb, pair error U[5 %, 20 %] (ins:del:sub 9:4:2), "
                        "seed up to 200 bp off the true diagonal, COMP on half; generated in HBM (seed 3)",
            "value": round(n / t, 1), "unit": "tasks/s", "s_per_batch": round(t, 3),
            "aligned": int(ok.sum()), "undefined_state": int(st[3]), "trace_values": r.used,
            "roofline": {"bound": "valu", "kernel": "la_batch_kernel (one wave64 per task, persistent queue)",
                         "achieved": round(ach, 2), "peak": VALU_PEAK_GOPS, "unit": "Gop/s (W_align lane-ops)",
                         "frac": round(ach / VALU_PEAK_GOPS, 5), "w_align": w_alg,
                         "kernel_s": round(kern_s, 4), "kernel_s_per_level": [x / 1e6 for x in r.work[2:6]],
                         "executed_cells_snake": r.work[:2], "task_reruns": r.work[6]},
            "generate_s": round(t_gen, 3)}


def c5_leg(with_cpu=True):
    """The nanopore-like block pair: the 100 Mbp genome with the repeat model,
    reads at 30x (lognormal 25 kb / 20 kb up to 200 kb templates, 8 % error,
    indels 3x as likely in homopolymer runs >= 4), block 1 (400 Mbp) against
    block 2 (400 Mbp), -e.75: the plain GPU path (checked against the
    reference's two .las), the scorer-gated GPU path and the reference
    daligner with every host core (-j the largest power of two)."""
    from marvel_amd import dbio, sim
    from marvel_amd import scorer as S
    from marvel_amd.daligner import Daligner
    flags = ["-k14", "-w6", "-h35", "-e.75", "-l1000", "-s100", "-j8", "-M32"]
    gold = {"001_00001/N.1.N.2.las": "b6d4104fa39cac81bbeeb064827b70a4f71f6c43aef630e8f2668bf820d60f9a",
            "001_00002/N.2.N.1.las": "b2b0dfc72f1df3d58cd12eaa6cb5fd52d2b4058d4e641b504f279b9269170da3"}
    d = workdir()
    out = {"workload": "100 Mbp genome + 200 repeat families x 40 copies (6 kb, 10 %) + 1 % tandems; 30x nanopore-"
                       "like reads (lognormal 25 / 20 kb, templates <= 200 kb, 8 % error, indels 3x in homopolymers "
                       ">= 4); block 1 (400 Mbp) vs block 2 (400 Mbp), both strands", "daligner": " ".join(flags)}
    try:
        t0 = time.time()
        reads = sim.config_nanopore_100mbp(5).reads
        dbio.write_db(os.path.join(d, "N"), reads, size_mbp=400)
        print(f"[c5] simulated and written in {time.time() - t0:.1f} s", flush=True)
        exe = os.path.join(ROOT, "marvel_amd", "bin", "daligner")

        def run_gpu(extra_env, tag):
            env = dict(os.environ, **extra_env)
            for k in gold:
                p = os.path.join(d, k)
                if os.path.exists(p):
                    os.remove(p)
            t = time.time()
            r = subprocess.run([exe] + flags + ["N.1", "N.2"], cwd=d, capture_output=True, text=True, timeout=900,
                               env=env)
            dt = time.time() - t
            if r.returncode != 0:
                raise RuntimeError(r.stderr[-2000:])
            shas = {k: hashlib.sha256(open(os.path.join(d, k), "rb").read()).hexdigest() for k in gold}
            n = struct.unpack("<q", open(os.path.join(d, "001_00001/N.1.N.2.las"), "rb").read(8))[0]
            print(f"[c5] {tag}: {dt:.1f} s, {n} records", flush=True)
            return dt, n, shas

        t_plain, n_plain, shas = run_gpu({}, "plain GPU")
        out["plain"] = {"value": round(2 * n_plain / t_plain, 1), "unit": "overlaps/s", "s": round(t_plain, 2),
                        "records": 2 * n_plain,
                        "parity": "bit-exact vs reference daligner" if shas == gold else f"MISMATCH {shas}",
                        "note": "whole command: DB load, both strands, both .las files"}
        ref = {k: os.path.join(d, k + ".exact") for k in gold}
        for k in gold:
            shutil.copyfile(os.path.join(d, k), ref[k])
        if os.path.exists(S.DEFAULT_WEIGHTS):
            t_g, n_g, _ = run_gpu({"MI_SCORER_WEIGHTS": S.DEFAULT_WEIGHTS}, "gated GPU")
            rec, hit, tot = S.record_recall(ref["001_00001/N.1.N.2.las"], os.path.join(d, "001_00001/N.1.N.2.las"))
            out["gated"] = {"value": round(2 * n_g / t_g, 1), "unit": "overlaps/s", "s": round(t_g, 2),
                            "records": 2 * n_g, "record_recall": round(rec, 6), "speedup": round(t_plain / t_g, 4)}
        if with_cpu:
            refexe = os.path.join(ROOT, "oracle", "_ref", "daligner")
            ncpu = len(os.sched_getaffinity(0))
            j = 1
            while j * 2 <= min(ncpu, 64):
                j *= 2
            cf = [f if not f.startswith("-j") else f"-j{j}" for f in flags]
            sd = os.path.join(d, "cpu")
            os.makedirs(sd)
            for f in os.listdir(d):
                if f.endswith((".db", ".idx", ".bps")):
                    os.symlink(os.path.join(d, f), os.path.join(sd, f))
            print(f"[c5] reference daligner {' '.join(cf)} on {ncpu} CPUs", flush=True)
            t = time.time()
            # -v: the reference reports each index build and comparison as it goes (to stderr here)
            subprocess.run([refexe, "-v"] + cf + ["N.1", "N.2"], cwd=sd, check=True, stdout=sys.stderr,
                           stderr=sys.stderr, timeout=1000)
            dt = time.time() - t
            n = struct.unpack("<q", open(os.path.join(sd, "001_00001/N.1.N.2.las"), "rb").read(8))[0]
            out["cpu"] = {"value": round(2 * n / dt, 1), "unit": "overlaps/s", "s": round(dt, 2), "cores": j,
                          "host_cpus": ncpu, "kind": "reference", "records": 2 * n}
            print(f"[c5] reference: {dt:.1f} s", flush=True)
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def workdir():
    base = "/dev/shm" if os.path.isdir("/dev/shm") and os.access("/dev/shm", os.W_OK) else None
    return tempfile.mkdtemp(prefix="mibench_", dir=base)


def cpu_baseline(d):
    """Reference daligner (compiled from source into oracle/_ref; the repo's C
    restatement when that is absent) on the same block, with -j the largest
    power of two not above the host's CPUs."""
    exe = os.path.join(ROOT, "oracle", "_ref", "daligner")
    kind = "reference"
    if not os.path.exists(exe):
        exe = os.path.join(ROOT, "oracle", "_bin", "daligner_oracle")
        kind = "port"
    if not os.path.exists(exe):
        return None
    ncpu = len(os.sched_getaffinity(0))
    j = 1
    while j * 2 <= min(ncpu, 64):
        j *= 2
    threads = j if kind == "reference" else 1
    flags = [f if not f.startswith("-j") else f"-j{threads}" for f in FLAGS]
    sd = os.path.join(d, "cpu")
    os.makedirs(sd, exist_ok=True)
    for f in os.listdir(d):
        if f.endswith((".db", ".idx", ".bps")):
            os.symlink(os.path.join(d, f), os.path.join(sd, f))
    t0 = time.time()
    subprocess.run([exe] + flags + ["C2.1", "C2.1"], cwd=sd, check=True, capture_output=True, timeout=900)
    dt = time.time() - t0
    n, bases, _ = las_summary(os.path.join(sd, "001_00001", "C2.1.C2.1.las"))
    return {"value": round(n / dt, 1), "unit": "overlaps/s", "cores": threads, "kind": kind, "host_cpus": ncpu,
            "sample": f"the whole benchmark block (C2.1 x C2.1, both strands), daligner {' '.join(flags)}, "
                      f"{dt:.2f} s wall, {n} records; -j is a power of two, so {threads} of the {ncpu} host CPUs",
            "aligned_gbp_per_s": round(bases / dt / 1e9, 6)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-gated", action="store_true", help="skip the scorer-gated leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-aligner leg")
    ap.add_argument("--c5", action="store_true",
                    help="only the 400 Mbp x 400 Mbp nanopore-like block pair: plain GPU, gated GPU, reference CPU")
    ap.add_argument("--c5-no-cpu", action="store_true", help="with --c5: skip the reference CPU run")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe << 20
        x = torch.ones(n, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "read": n, "written": n}), flush=True)
        return
    if args.c5:
        print(json.dumps({"c5": c5_leg(with_cpu=not args.c5_no_cpu)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}); "
              "block pairs shard across GPUs without communication (one block pair per GPU)", file=sys.stderr)
        sys.exit(2)

    from marvel_amd import dbio, sim
    from marvel_amd.daligner import Daligner

    d = workdir()
    try:
        t0 = time.time()
        reads = sim.config_5mbp(2).reads
        dbio.write_db(os.path.join(d, "C2"), reads, size_mbp=400)
        t_db = time.time() - t0

        dal = Daligner()
        dal.configure(nthreads=8, mem_limit_gb=32, minover=2000, symmetric=1)
        dal.set_filter_params(14, 6, 0, 35)
        t0 = time.time()
        A = dal.open_block(os.path.join(d, "C2.1"))
        spec = dal.new_align_spec(0.70, 100, list(A.freq), nthreads=8)
        dev = dal.block_device(A)        # reads packed and resident in HBM
        dal.synchronize()
        t_load = time.time() - t0
        cA = dal.complement_header(A)
        os.makedirs(os.path.join(d, "001_00001"), exist_ok=True)
        cwd = os.getcwd()
        os.chdir(d)
        try:
            # launch shapes: the cached choice for this GPU, else calibrated on this workload
            from marvel_amd import tune
            from marvel_amd.align import gen_tasks_device, la_batch_device
            cal = {}

            def batch_run():
                if "dt" not in cal:
                    cal["dt"] = gen_tasks_device(200_000, 11)
                return la_batch_device(cal["dt"]).work

            tuning = tune.ensure(report_step=lambda: dal.compare_self(A, spec, 1, "C2.1", dev=dev, cA=cA),
                                 batch_run=batch_run)
            cal.clear()
            for _ in range(args.warmup):
                dal.compare_self(A, spec, 1, "C2.1", dev=dev, cA=cA)
            # the timed region: K block-pair steps back to back, each step's
            # .las write overlapping the next step's GPU work, closed by the
            # wait for the last write and a device synchronize
            dal.flush_writes()
            dal.synchronize()
            times, stats, phases = [], [], []
            t_all = time.perf_counter()
            for _ in range(args.steps):
                t = time.perf_counter()
                s1, s2 = dal.compare_self(A, spec, 1, "C2.1", dev=dev, cA=cA, wait=False)
                dal.synchronize()
                times.append(time.perf_counter() - t)
                stats.append((s1, s2))
                phases.append(dal.last_phases)
            dal.flush_writes()
            dal.synchronize()
            total = time.perf_counter() - t_all
        finally:
            os.chdir(cwd)
        las = os.path.join(d, "001_00001", "C2.1.C2.1.las")
        n, bases, sha = las_summary(las)
        step = total / len(times)
        # dominant kernel: the report/alignment kernel, W_align lane-ops against the int32 VALU peak
        k_report = sum(a["t_report_kernel"] + b["t_report_kernel"] for a, b in stats) / len(stats)
        launches = sum(a["report_launches"] + b["report_launches"] for a, b in stats) / len(stats)
        w_align = stats[-1][0]["cells"] + stats[-1][0]["snake"] + stats[-1][1]["cells"] + stats[-1][1]["snake"]
        ach = w_align / k_report / 1e9 if k_report > 0 else 0.0
        # the ceiling that binds it: VALU instruction issue (instructions per wave step from the counter summary,
        # wave steps and kernel time of this run; the time includes the few wider-window launches)
        vps = profiled_valu_per_wave_step()
        wsteps = sum(x["wave_steps"] for x in stats[-1])
        issue = None
        if vps and k_report > 0:
            ia = vps * wsteps / k_report / 1e9
            issue = {"bound": "VALU instruction issue (one wave64 instruction per CU per cycle)",
                     "achieved": round(ia, 1), "peak": VALU_ISSUE_PEAK_G, "unit": "G wave64 VALU instructions/s",
                     "frac": round(ia / VALU_ISSUE_PEAK_G, 3), "valu_per_wave_step": vps,
                     "wave_steps_per_step": wsteps,
                     "source": "SQ_INSTS_VALU per wave step from profiles/k_report_pmc_5mbp.md"}
        # seed radix sort: 32 bytes per seed per 8-bit pass (read + write of a 16-byte seed)
        t_sort = sum(a["t_seed_sort"] + b["t_seed_sort"] for a, b in stats) / len(stats)
        sort_bytes = sum(x["nhits"] * 32 * x["seed_sort_passes"] for x in stats[-1])
        sort_gbs = sort_bytes / t_sort / 1e9 if t_sort > 0 else 0.0
        # HBM bytes of one seed-sort pass from the committed counter summary:
        # the digit-histogram read plus the scatter's reads and writes
        hist_b, scat_b = profiled_traffic("rs_hist<Seed>"), profiled_traffic("rs_scatter<Seed>")
        sort_traffic = hist_b + scat_b if hist_b and scat_b else None
        # the filter stage (both strands): k-mer index builds, merge-join and
        # seed sort, against implementation-independent bytes: per k-mer
        # indexed its 8-byte key written and sorted once (read + write), per
        # index entry merged 8 bytes read, per seed 16 bytes written and
        # sorted once (read + write)
        last = stats[-1]
        filt_bytes = sum(24 * x["nkmers_b"] + 8 * (x["nkmers_a"] + x["nkmers_b"]) + 48 * x["nhits"] for x in last)
        t_filt = phases[-1]["index"] + phases[-1]["complement_index"] + sum(x["t_filter"] for x in last)
        filt_gbs = filt_bytes / t_filt / 1e9 if t_filt > 0 else 0.0
        # one step with the block uploaded from host memory (PCIe included)
        cwd = os.getcwd()
        os.chdir(d)
        try:
            dal.flush_writes()
            dal.synchronize()
            t = time.perf_counter()
            dal.compare_self(A, spec, 1, "C2.1")
            dal.flush_writes()
            dal.synchronize()
            t_pcie = time.perf_counter() - t
        finally:
            os.chdir(cwd)
        ok = sha == GOLD_SHA
        cpu = None if args.no_cpu else cpu_baseline(d)
        batch = None if args.no_batch else batched_leg()
        gated = None if args.no_gated else gated_leg(2)
        line = {
            "metric": METRIC,
            "value": round(n / step, 1) if ok else None,
            "unit": "overlaps/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": round(step * 1e3, 2),
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "int32",
            "data": "synthetic (seeded simulator: 5 Mbp iid genome, 30x, 9/4/2% indel/sub errors)",
            "config": {"workload": "5 Mbp / 30x reads, one ~158 Mbp block vs itself, both strands",
                       "reads": len(reads), "bases": sum(len(r) for r in reads), "daligner": " ".join(FLAGS),
                       "parallelism": "1 GPU"},
            "aligned_gbp_per_s": round(bases / step / 1e9, 4),
            "records": n,
            "parity": "bit-exact vs reference daligner" if ok else f"MISMATCH sha256 {sha}",
            "pcie_inclusive": {"value": round(n / t_pcie, 1), "unit": "overlaps/s", "ms_per_step": round(t_pcie * 1e3, 2),
                               "note": "one step whose index build uploads the block's reads from host memory"},
            "roofline": {"bound": "valu", "kernel": "k_report (one wave64 per read pair: gating + Local_Alignment)",
                         "achieved": round(ach, 2), "peak": VALU_PEAK_GOPS, "unit": "Gop/s (W_align lane-ops)",
                         "frac": round(ach / VALU_PEAK_GOPS, 5), "traffic": profiled_traffic("k_report<1, false>"),
                         "traffic_unit": "HBM bytes per k_report<1, false> launch (rocprofv3 --pmc, calibrated)",
                         "kernel_ms_per_step": round(k_report * 1e3, 2), "launches_per_step": launches,
                         "w_align_per_step": w_align, "issue": issue},
            "roofline_hbm": {"bound": "hbm", "kernel": "seed LSD radix sort (rs_hist + rs_scatter)",
                             "achieved": round(sort_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(sort_gbs / HBM_PEAK_GBS, 4), "traffic": sort_traffic,
                             "traffic_unit": "HBM bytes per seed-sort pass, rs_hist + rs_scatter (rocprofv3 --pmc, "
                                             "calibrated); algorithmic: 32 B per seed",
                             "ms_per_step": round(t_sort * 1e3, 2)},
            "roofline_filter": {"bound": "hbm", "stage": "k-mer index builds, merge-join, seed sort; both strands",
                                "achieved": round(filt_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(filt_gbs / HBM_PEAK_GBS, 4), "bytes_per_step": int(filt_bytes),
                                "ms_per_step": round(t_filt * 1e3, 2),
                                "bytes_model": "24 B per k-mer indexed, 8 B per index entry merged, "
                                               "48 B per seed (written, then sorted once)"},
            "cpu_baseline": cpu,
            "tuning": tuning,
            "graphs": {"replayed_per_step": stats[-1][0]["graphs_replayed"] + stats[-1][1]["graphs_replayed"],
                       "captured_last_step": stats[-1][0]["graphs_captured"] + stats[-1][1]["graphs_captured"],
                       "kernel_launches_replayed_per_step": stats[-1][0]["launches_replayed"] +
                       stats[-1][1]["launches_replayed"],
                       "note": "radix sorts and scans of a pass recorded once as HIP graphs and replayed"},
            "batched_aligner": batch,
            "scorer_gated": gated,
            "times_s": [round(x, 4) for x in times],
            "phases_s": phases[-1],
            "setup_s": {"db_write": round(t_db, 2), "load_and_upload": round(t_load, 2)},
            "pass_stats": {k: [stats[-1][0][k], stats[-1][1][k]] for k in
                           ("nhits", "npairs", "njobs", "njobs_kept", "naligned", "nrecords", "retries", "edge_events",
                            "cells", "snake", "wave_steps",
                            "t_filter", "t_report", "t_copy", "t_total")},
        }
        print(json.dumps(line), flush=True)
        if not ok:
            sys.exit(3)
    finally:
        shutil.rmtree(d, ignore_errors=True)


if __name__ == "__main__":
    main()
