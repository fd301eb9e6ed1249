#!/usr/bin/env python3
"""Throughput of the MI355X distance engine on one GPU, with roofline and CPU baseline.

A step is one pass of the hot path over the resident input: the packed alignment is in HBM, the
step computes every pair's ts/tv/valid counts, the correction, the running maximum and the
saturation patch (2 * max) into a float32 upper triangle that stays in HBM (the reference's
computeDistances region, KimuraDistance.cpp:75-99, minus parsing). Default workload: 20000 taxa x
5000 sites, 5% gaps, K2P. Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--workload large|plain|small_jc|bootstrap|protein]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G site-pairs/s on distance matrix vs CPU; p50/p99 job latency served"
HBM_PEAK_GBS = 8000.0
# VALU instructions per 32 sites and pair of the count loop, by class: (bitwise, v_bcnt).
# K2P: v_and, 2 v_bitop3 (same-class sites, then transitions among them) and 3 v_bcnt;
# JC: v_and, 2 v_bitop3 (same class, then identical bases) and 2 v_bcnt.
OP_CLASSES = {"k2p": (3, 3), "jc": (3, 2)}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpus() -> dict:
    """The host CPUs this process may run on: the affinity mask, capped by a cgroup CPU quota
    (cpu.max) and by the CPU share the environment grants (OMP_NUM_THREADS) when they are set, and
    the CPU model from /proc/cpuinfo."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    usable = aff if quota is None else max(1, min(aff, int(quota)))
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        usable = min(usable, int(share))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"affinity": aff, "cgroup_quota": quota, "omp_num_threads": share or None,
            "usable": usable, "model": model}


def cpu_cores() -> int:
    """Every host CPU this process may use (worker pools and the CPU baseline)."""
    return host_cpus()["usable"]


def cpu_baseline(aln, model: str, sample_n: int, threads: int, protein: bool = False) -> dict:
    """Reference CPU engine (compiled from its sources into oracle/_ref) on the first sample_n
    taxa of the same alignment; falls back to the oracle's restatement when the harness is absent."""
    from oracle import oracle as O
    from rapidnj_amd.workloads import write_fasta
    sub = aln[:sample_n]
    n, L = sub.shape
    site_pairs = n * (n - 1) // 2 * L
    if O.ref_harness_available():
        with tempfile.TemporaryDirectory() as tmp:
            fa = os.path.join(tmp, "sample.fa")
            write_fasta(fa, sub)
            out = subprocess.run([O.REF_HARNESS, "ptime" if protein else "time", fa,
                                  "kim" if model == "k2p" else "jc", str(threads), "1"],
                                 check=True, capture_output=True, text=True)
            secs = json.loads(out.stdout.strip().splitlines()[-1])["best_s"]
        kind = "reference"
    else:
        t0 = time.perf_counter()
        if protein:
            O.protein_matrix(sub, model)
            threads = 1
        else:
            O.distance_matrix(sub, model, threads=threads)
        secs = time.perf_counter() - t0
        kind = "port"
    label = ("Kimura protein" if model == "k2p" else "JC alpha=20") if protein else model.upper()
    return {"value": round(site_pairs / secs / 1e9, 3), "unit": "G site-pairs/s",
            "cores": threads, "kind": kind,
            "sample": f"first {n} taxa x {L} sites of the same alignment, {label}, "
                      f"compute+post-process {secs:.2f} s"}


def _pmc_pass(counter: str, workload: str, tile_config: int) -> dict | None:
    """One rocprofv3 counter pass over a child `bench.py --probe-traffic` run: per-dispatch values
    of `counter` (KiB) for the last tile-kernel dispatch and for the 1 GiB elementwise kernel
    that ends the probe. None when rocprofv3 is missing or the pass fails."""
    import csv
    import shutil
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        cmd = [prof, "--pmc", counter, "-d", tmp, "-o", "run", "--output-format", "csv", "--",
               sys.executable, os.path.abspath(__file__), "--probe-traffic", "--workload", workload,
               "--tile-config", str(tile_config)]
        env = dict(os.environ, TMPDIR="/tmp", PYTHONPATH=ROOT)
        try:
            subprocess.run(cmd, cwd="/tmp", env=env, check=True, timeout=240,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        except (subprocess.SubprocessError, OSError) as e:
            log(f"rocprofv3 --pmc {counter} failed: {e}")
            return None
        files = []
        for root, _, names in os.walk(tmp):
            files += [os.path.join(root, f) for f in names if f.endswith("counter_collection.csv")]
        if not files:
            return None
        per_dispatch: dict[int, list] = {}
        with open(files[0]) as fh:
            for r in csv.DictReader(fh):
                d = int(r["Dispatch_Id"])
                ent = per_dispatch.setdefault(d, [r["Kernel_Name"], 0.0])
                ent[1] += float(r["Counter_Value"])
    order = sorted(per_dispatch)
    tiles = [d for d in order if "tile_kernel" in per_dispatch[d][0]]
    if not tiles:
        return None
    return {"tile_kib": per_dispatch[tiles[-1]][1], "copy_kib": per_dispatch[order[-1]][1],
            "tile_kernel": per_dispatch[tiles[-1]][0].split("(")[0]}


def measure_traffic(workload: str, tile_config: int) -> dict | None:
    """HBM bytes of one tile-kernel launch: FETCH_SIZE and WRITE_SIZE (two counter passes), each
    scaled by what the counter reads for a 1 GiB -> 1 GiB elementwise kernel in the same pass."""
    f = _pmc_pass("FETCH_SIZE", workload, tile_config)
    w = _pmc_pass("WRITE_SIZE", workload, tile_config)
    if not f or not w or f["copy_kib"] <= 0 or w["copy_kib"] <= 0:
        return None
    gib_kib = float(1 << 20)
    fetch_scale, write_scale = gib_kib / f["copy_kib"], gib_kib / w["copy_kib"]
    fetch_b = f["tile_kib"] * fetch_scale * 1024
    write_b = w["tile_kib"] * write_scale * 1024
    return {"bytes": fetch_b + write_b, "fetch_bytes": fetch_b, "write_bytes": write_b,
            "fetch_scale": round(fetch_scale, 3), "write_scale": round(write_scale, 3),
            "counters": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE (KiB per dispatch), each "
                        "scaled by its reading for a 1 GiB -> 1 GiB elementwise kernel in the "
                        "same run"}


def probe_traffic(args) -> int:
    """Child of measure_traffic: two raw launches of the workload, then one elementwise kernel that
    reads 1 GiB and writes 1 GiB (ordinary vector loads and stores, like the tile kernel's)."""
    import torch
    from rapidnj_amd import native as N
    from rapidnj_amd.workloads import WORKLOADS, make_alignment
    wl = WORKLOADS[args.workload]
    aln = make_alignment(wl)
    ctx = N.Context(0)
    if args.tile_config >= 0:
        ctx.set_tile_config(args.tile_config)
    A = ctx.alignment(aln)
    pairs = wl.n * (wl.n - 1) // 2
    out = torch.empty(pairs, dtype=torch.float32, device="cuda:0")
    for _ in range(2):
        ctx.distances(A, wl.model, out=out, raw=True)
    torch.cuda.synchronize()
    src = torch.ones((1 << 30) // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.mul(src, 2.0, out=dst)
    torch.cuda.synchronize()
    A.close()
    ctx.close()
    return 0


def ingest_compare(ctx, aln) -> dict:
    """FASTA and Stockholm -> packed alignment in HBM for the benchmark alignment: from host bytes
    (upload + device parse + pack), from a device-resident file image (parse + pack only), and
    from a file (pinned read + upload + parse + pack), next to the reference's readers
    (dataloaderFasta / dataloaderStockholm, readData with fastdist) on the same file. The FASTA
    text has 60-column lines; the Stockholm text has one line per sequence (the reference reads
    single-line Stockholm), a header and the closing "//"."""
    import torch
    from oracle import oracle as O
    lines = []
    for i, row in enumerate(aln):
        lines.append(b">t%d" % i)
        b = row.tobytes()
        lines += [b[k:k + 60] for k in range(0, len(b), 60)]
    fasta = b"\n".join(lines) + b"\n"
    sth = (b"# STOCKHOLM 1.0\n" + b"".join(b"t%d %s\n" % (i, row.tobytes())
                                          for i, row in enumerate(aln)) + b"//\n")

    def best(fn, reps=3):
        t = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            a = fn()
            torch.cuda.synchronize()
            t.append(time.perf_counter() - t0)
            a.close()
        return min(t)

    def leg(data, parse, parse_file, suffix, ref_mode):
        host = best(lambda: parse(data))
        dev = torch.frombuffer(bytearray(data), dtype=torch.uint8).to("cuda:0")
        device = best(lambda: parse(dev))
        del dev
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "bench" + suffix)
            with open(path, "wb") as fh:
                fh.write(data)
            from_file = best(lambda: parse_file(path))
            ref = None
            if O.ref_harness_available():
                r = subprocess.run([O.REF_HARNESS, ref_mode, path, "1"], check=True,
                                   capture_output=True, text=True)
                ref = json.loads(r.stdout.strip().splitlines()[-1])["best_s"]
        return {"bytes": len(data), "host_bytes_ms": round(host * 1e3, 2),
                "device_image_ms": round(device * 1e3, 2), "file_ms": round(from_file * 1e3, 2),
                "device_image_GBps": round(len(data) / device / 1e9, 1),
                "cpu_reference_reader_ms": None if ref is None else round(ref * 1e3, 1),
                "speedup_file_vs_reference": None if ref is None else round(ref / from_file, 1)}

    out = {"taxa": int(aln.shape[0]), "sites": int(aln.shape[1])}
    out.update(leg(fasta, ctx.alignment_from_fasta, ctx.alignment_from_fasta_file, ".fa", "load"))
    out["stockholm"] = leg(sth, ctx.alignment_from_stockholm, ctx.alignment_from_stockholm_file,
                           ".sth", "sload")
    return out


def handoff_compare(ctx, A) -> dict:
    """The in-memory NJ hand-off of the benchmark matrix: distances computed in HBM, then
    post-processed rows in host memory, as full rows (rnj_distances_rows) and as the lower
    triangle rapidNJMem keeps (rnj_distances_rows_lower)."""
    import ctypes
    import numpy as np
    from rapidnj_amd import native as N
    n = A.n
    full = np.empty((n, n), dtype=np.float32)
    lower = [np.empty(i + 1, dtype=np.float32) for i in range(n)]
    pf = (ctypes.c_void_p * n)(*[full.ctypes.data + i * full.strides[0] for i in range(n)])
    pl = (ctypes.c_void_p * n)(*[r.ctypes.data for r in lower])
    mx = ctypes.c_float()
    out = {"taxa": n}
    for name, fn, rows, nbytes in (("full_rows", N.lib().rnj_distances_rows, pf, n * n * 4),
                                   ("lower_rows", N.lib().rnj_distances_rows_lower, pl,
                                    n * (n + 1) // 2 * 4)):
        best = 1e30
        for _ in range(3):
            t0 = time.perf_counter()
            N.check(fn(ctx.handle, A.handle, N.K2P, ctypes.cast(rows, ctypes.c_void_p),
                       ctypes.byref(mx)))
            best = min(best, time.perf_counter() - t0)
        out[name + "_ms"] = round(best * 1e3, 1)
        out[name + "_host_GBps"] = round(nbytes / best / 1e9, 1)
    return out


def op_classes(A, model: str) -> tuple[int, int]:
    """(bitwise, v_bcnt) VALU instructions per 32 sites and pair of the tile kernel that counts
    alignment A."""
    if A.is_protein:
        return A.info.nplanes, 2  # B code planes: v_xor, B - 1 v_bitop3, v_and; 2 v_bcnt
    return OP_CLASSES[model]


def ops_per_word_pair(A, model: str) -> int:
    return sum(op_classes(A, model))


def valu_peak(ctx, A, model: str, naive: bool = False) -> dict:
    """Int32 VALU ceiling of the count loop that alignment A runs under `model`: the measured
    rate at which the device executes that loop's own per-pair instruction sequence
    (rnj_valu_mix_rate: from registers, 8 waves per SIMD, no LDS, staging or epilogue). Also
    reported: the per-class issue rates (rnj_valu_issue_rates: bitwise, v_bcnt, each alone) and
    the ceiling they would give if the classes combined independently, (b + c) / (b / bitwise +
    c / bcnt) for b bitwise and c v_bcnt instructions per 32 site-pairs -- mixed, they do not (the
    bitwise instructions lose most of their speed-up between v_bcnt), so that figure is not
    reachable and is kept for reference only."""
    if A.is_protein:
        mix, planes = "protein", A.info.nplanes - 1
    else:
        mix, planes = ("naive" if naive else model), 0
    rate = ctx.valu_mix_rate(mix, planes)
    b, c = (6, 3) if naive else op_classes(A, model)
    bw, bc = ctx.valu_issue_rates()
    harmonic = (b + c) / (b / bw + c / bc)
    return {"peak_tops": rate / 1e12,
            "ceiling": f"measured rate of the {mix} count loop's instruction sequence "
                       "(rnj_valu_mix_rate)",
            "class_model": {"bitwise_tops": round(bw / 1e12, 2), "bcnt_tops": round(bc / 1e12, 2),
                            "classes": {"bitwise": b, "bcnt": c},
                            "independent_classes_tops": round(harmonic / 1e12, 2)}}


def peak_fields(pk: dict, achieved: float) -> dict:
    cm = dict(pk["class_model"])
    cm["frac_vs_independent_classes"] = round(achieved / cm["independent_classes_tops"], 4)
    return {"peak": round(pk["peak_tops"], 2), "frac": round(achieved / pk["peak_tops"], 4),
            "ceiling": pk["ceiling"], "class_model": cm}


def model_side(ctx, A, model: str, naive: bool, wl, out, stream) -> dict:
    """Alignment A (the workload's taxa and sites) under `model`: step time, and the raw launch
    (reset + tile kernel + job maximum) against the ceiling of its count loop's instruction
    mix."""
    import torch
    pairs = wl.n * (wl.n - 1) // 2
    for _ in range(3):
        ctx.distances(A, model, out=out, stream=stream)
    torch.cuda.synchronize()
    k = 10
    t0 = time.perf_counter()
    for _ in range(k):
        ctx.distances(A, model, out=out, stream=stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / k * 1e3
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(k):
        ctx.distances(A, model, out=out, raw=True, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    kern_ms = e0.elapsed_time(e1) / k
    pk = valu_peak(ctx, A, model, naive=naive)
    ops = 9 if naive else ops_per_word_pair(A, model)
    achieved = ops * pairs * (wl.L / 32.0) / (kern_ms / 1e3) / 1e12
    cfg = ctx.tile_config_for(A, model)
    return {"model": model, "ms_per_step": round(ms, 4),
            "value": round(pairs * wl.L / (ms / 1e3) / 1e9, 2),
            "unit": "G site-pairs/s", "kernel_ms": round(kern_ms, 4),
            "valu": {"achieved": round(achieved, 3), "unit": "TOP/s",
                     **peak_fields(pk, achieved), "ops_per_32_site_pairs": ops},
            "tile_config": tile_name(cfg)}


def naive_side(ctx, aln, wl, out, stream) -> dict:
    """The workload's alignment under the naive character rules (rnj_set_char_rules: a fourth
    plane, 6 bitwise + 3 v_bcnt per 32 site-pairs)."""
    ctx.set_char_rules(True)
    try:
        An = ctx.alignment(aln)
    finally:
        ctx.set_char_rules(False)
    try:
        return model_side(ctx, An, wl.model, True, wl, out, stream)
    finally:
        An.close()


def tile_name(cfg: int) -> str:
    from rapidnj_amd import native as N
    return N.tile_configs()[cfg]


def protein_side(ctx, stream, cpu: bool) -> dict:
    """Kimura-protein distances of the 20000 x 1250 protein workload (resident, float32 upper
    triangle in HBM): throughput, the protein tile kernel's VALU roofline, and the reference's
    CPU protein engine (bitDistanceProtein + KimuraDistance) on the first 10000 taxa."""
    import torch
    from rapidnj_amd import native as N
    from rapidnj_amd.workloads import WORKLOADS, make_alignment
    wl = WORKLOADS["protein"]
    aln = make_alignment(wl)
    A = ctx.alignment(aln)
    n, L = wl.n, wl.L
    pairs = n * (n - 1) // 2
    out = torch.empty(pairs, dtype=torch.float32, device="cuda:0")
    for _ in range(3):  # the first launch tunes the tile configuration
        ctx.distances(A, wl.model, out=out, stream=stream)
    torch.cuda.synchronize()
    k = 10
    t0 = time.perf_counter()
    for _ in range(k):
        ctx.distances(A, wl.model, out=out, stream=stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / k * 1e3
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(k):
        ctx.distances(A, wl.model, out=out, raw=True, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    kern_ms = e0.elapsed_time(e1) / k
    ops = ops_per_word_pair(A, wl.model) * pairs * (L / 32.0)
    tops = ops / (kern_ms / 1e3) / 1e12
    pk = valu_peak(ctx, A, wl.model)
    res = {"workload": wl.name, "code_planes": A.info.nplanes - 1,
           "ms_per_step": round(ms, 3), "value": round(pairs * L / (ms / 1e3) / 1e9, 2),
           "unit": "G site-pairs/s", "kernel_ms": round(kern_ms, 3),
           "valu": {"achieved": round(tops, 3), "unit": "TOP/s", **peak_fields(pk, tops),
                    "ops_per_32_site_pairs": ops_per_word_pair(A, wl.model)},
           "tile_config": N.tile_configs()[ctx.tile_config_for(A, wl.model)]}
    if cpu:
        res["cpu_reference"] = cpu_baseline(aln, wl.model, 10000, cpu_cores(), protein=True)
        res["speedup_vs_cpu"] = round(res["value"] / res["cpu_reference"]["value"], 1)
    A.close()
    return res


def nj_side(ctx, A, cpu: bool, sample_n: int = 5000) -> dict:
    """Neighbour joining on the device over the benchmark's 20000 x 5000 K2P matrix kept in HBM
    (rnj_nj_device), and on its leading sample_n x sample_n block next to the reference's own
    rapidNJ (compiled from its sources, ``ref_harness nj``) on the same block; the two Newick
    texts are compared byte for byte."""
    import numpy as np
    import torch
    from oracle import oracle as O
    n = A.n
    sample_n = min(sample_n, n)
    d = torch.empty((n, n), dtype=torch.float32, device="cuda:0")
    ctx.distances(A, "k2p", layout="square", out=d, exact=True)
    torch.cuda.synchronize()
    out = {"taxa": n}
    t0 = time.perf_counter()
    tree = ctx.nj(d)
    out["device_s"] = round(time.perf_counter() - t0, 3)
    out["host_searches"] = tree.host_searches  # iterations whose search the host settled
    tree.close()
    sub = d[:sample_n, :sample_n].contiguous()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree = ctx.nj(sub)
    out[f"device_{sample_n}_s"] = round(time.perf_counter() - t0, 3)
    text = tree.newick()
    tree.close()
    if cpu and O.ref_harness_available():
        with tempfile.TemporaryDirectory() as tmp:
            mf, nwk = os.path.join(tmp, "m.bin"), os.path.join(tmp, "t.nwk")
            m = sub.cpu().numpy()
            with open(mf, "wb") as fh:
                np.array([sample_n], dtype=np.uint32).tofile(fh)
                m.tofile(fh)
            t0 = time.perf_counter()
            subprocess.run([O.REF_HARNESS, "nj", mf, nwk], check=True, capture_output=True)
            out[f"cpu_reference_{sample_n}_s"] = round(time.perf_counter() - t0, 3)
            with open(nwk, "rb") as fh:
                out[f"same_tree_as_reference_{sample_n}"] = fh.read() == text
        out[f"speedup_{sample_n}"] = round(out[f"cpu_reference_{sample_n}_s"] / out[f"device_{sample_n}_s"], 1)
    del d, sub
    return out


def side_main(args) -> int:
    """--side: one side measurement alone (for profiling it)."""
    from rapidnj_amd import native as N
    from rapidnj_amd.workloads import WORKLOADS, make_alignment
    if args.side == "valu":
        # the VALU ceilings alone: both instruction classes' issue rates and every count loop's
        # instruction-mix rate (lane-ops/s), three measurements each
        ctx = N.Context(0)
        res = {"issue_rates_tops": [], "mix_rates_tops": {}}
        for _ in range(3):
            bw, bc = ctx.valu_issue_rates()
            res["issue_rates_tops"].append({"bitwise": round(bw / 1e12, 2),
                                            "bcnt": round(bc / 1e12, 2)})
        for mix, planes in (("k2p", 0), ("jc", 0), ("naive", 0), ("protein", 5), ("protein", 6),
                            ("protein", 8)):
            key = mix if not planes else f"{mix}_b{planes}"
            res["mix_rates_tops"][key] = [round(ctx.valu_mix_rate(mix, planes) / 1e12, 2)
                                          for _ in range(3)]
        print(json.dumps({"side": "valu", **res}), flush=True)
        ctx.close()
        return 0
    wl = WORKLOADS[args.workload]
    aln = make_alignment(wl)
    ctx = N.Context(0)
    if args.side == "ingest":
        res = ingest_compare(ctx, aln)
    elif args.side == "nj":
        A = ctx.alignment(aln)
        res = nj_side(ctx, A, not args.no_cpu_baseline, args.nj_reference_taxa)
        A.close()
    elif args.side in ("jc", "naive"):
        import torch
        out = torch.empty(wl.n * (wl.n - 1) // 2, dtype=torch.float32, device="cuda:0")
        stream = torch.cuda.current_stream()
        if args.side == "naive":
            res = naive_side(ctx, aln, wl, out, stream)
        else:
            A = ctx.alignment(aln)
            res = model_side(ctx, A, "jc", False, wl, out, stream)
            A.close()
        del out
    else:
        A = ctx.alignment(aln)
        res = handoff_compare(ctx, A)
        A.close()
    print(json.dumps({"side": args.side, "workload": wl.name, **res}), flush=True)
    ctx.close()
    return 0


def _gen_job(job):
    from rapidnj_amd.workloads import simulate_alignment
    return simulate_alignment(job.n, job.L, job.gap, job.seed)


def generate_stream(jobs, procs: int):
    """Alignments of the served stream, generated in a process pool (before any GPU use)."""
    import multiprocessing as mp
    with mp.get_context("fork").Pool(procs) as pool:
        return pool.map(_gen_job, jobs, chunksize=1)


def served_run(ctx, jobs, alns, open_loop: bool, slots: int = 4, max_batch: int = 64) -> dict:
    """Submit the jobs (at their Poisson arrival times, or all at once) and wait for each in order.
    Latency = submit -> post-processed triangle on the host."""
    import threading
    import numpy as np
    srv = ctx.server(slots=slots, max_batch_jobs=max_batch)
    tickets = [None] * len(jobs)
    t0 = time.perf_counter()

    def submitter():
        for k, j in enumerate(jobs):
            if open_loop:
                delay = t0 + j.arrival_s - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            tickets[k] = srv.submit(alns[k], j.model)

    th = threading.Thread(target=submitter)
    th.start()
    lat, batch_jobs, stages = [], [], []
    for k, j in enumerate(jobs):
        while tickets[k] is None:
            time.sleep(0.0002)
        _, st = srv.wait(tickets[k], j.n)
        lat.append(st.total_s)
        batch_jobs.append(st.batch_jobs)
        stages.append((st.queued_s, st.upload_s, st.launch_s, st.run_s, st.download_s))
    makespan = time.perf_counter() - t0
    th.join()
    srv.close()
    sp = sum(j.n * (j.n - 1) // 2 * j.L for j in jobs)
    lat_ms = np.array(lat) * 1e3
    return {"jobs": len(jobs), "makespan_s": round(makespan, 3),
            "value": round(sp / makespan / 1e9, 2), "unit": "G site-pairs/s",
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 2),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 2),
            "mean_batch_jobs": round(float(np.mean(batch_jobs)), 2),
            "stage_p50_ms": dict(zip(("queued", "upload", "launch", "run", "download"),
                                     (round(float(v), 3) for v in
                                      np.percentile(np.array(stages) * 1e3, 50, axis=0))))}


def served_resident(ctx, jobs, alns, slots: int = 4, max_batch: int = 64) -> dict:
    """The stream with resident I/O: every alignment already packed in HBM and every triangle
    written to a device buffer (rnj_server_submit_resident), all jobs submitted at once and waited
    for in order. Latency = submit -> triangle complete in HBM."""
    import numpy as np
    import torch
    A = [ctx.alignment(a) for a in alns]
    outs = [torch.empty((max(j.n * (j.n - 1) // 2, 1),), dtype=torch.float32, device="cuda:0")
            for j in jobs]
    torch.cuda.synchronize()
    srv = ctx.server(slots=slots, max_batch_jobs=max_batch)
    t0 = time.perf_counter()
    tickets = [srv.submit_resident(a, j.model, out=o) for a, j, o in zip(A, jobs, outs)]
    lat, batch_jobs = [], []
    for j, t in zip(jobs, tickets):
        _, st = srv.wait(t, j.n, copy=False)
        lat.append(st.total_s)
        batch_jobs.append(st.batch_jobs)
    makespan = time.perf_counter() - t0
    srv.close()
    for a in A:
        a.close()
    del outs
    sp = sum(j.n * (j.n - 1) // 2 * j.L for j in jobs)
    lat_ms = np.array(lat) * 1e3
    return {"jobs": len(jobs), "makespan_s": round(makespan, 3),
            "value": round(sp / makespan / 1e9, 2), "unit": "G site-pairs/s",
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 2),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 2),
            "mean_batch_jobs": round(float(np.mean(batch_jobs)), 2)}


def one_job_per_launch(ctx, jobs, alns) -> dict:
    """The same jobs one per launch: the same asynchronous server (stream-ordered uploads through
    pinned staging, two batches in flight, results copied as they finish) with batches of one job,
    end to end and with resident I/O."""
    e2e = served_run(ctx, jobs, alns, open_loop=False, max_batch=1)
    res = served_resident(ctx, jobs, alns, max_batch=1)
    return {"end_to_end": {k: e2e[k] for k in ("makespan_s", "value", "unit", "p50_ms", "p99_ms")},
            "resident": {k: res[k] for k in ("makespan_s", "value", "unit", "p50_ms", "p99_ms")}}


def uneven_batch_compare(ctx, jobs, alns, count: int = 64) -> dict:
    """The first `count` alignments of a job list (uneven taxa and site counts, JC and K2P),
    resident in HBM: one captured multi-alignment launch (rnj_batch_plan) against one distance
    launch per alignment, outputs in HBM either way."""
    import torch
    jobs, alns = jobs[:count], alns[:count]
    A = [ctx.alignment(a) for a in alns]
    models = [j.model for j in jobs]
    outs = [torch.empty(max(1, j.n * (j.n - 1) // 2), dtype=torch.float32, device="cuda:0")
            for j in jobs]
    sp = sum(j.n * (j.n - 1) // 2 * j.L for j in jobs)
    k = 10
    for _ in range(2):
        for a, m, o in zip(A, models, outs):
            ctx.distances(a, m, out=o)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        for a, m, o in zip(A, models, outs):
            ctx.distances(a, m, out=o)
    torch.cuda.synchronize()
    single = (time.perf_counter() - t0) / k
    ref = [o.clone() for o in outs]
    with ctx.batch_plan(A, models, outs=outs) as plan:
        plan.run()
        torch.cuda.synchronize()
        same = all(bool(torch.equal(o, r)) for o, r in zip(outs, ref))
        plan.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            plan.run()
        torch.cuda.synchronize()
        graphed = (time.perf_counter() - t0) / k
    for a in A:
        a.close()
    return {"alignments": len(jobs), "site_pairs": sp, "graph_ms": round(graphed * 1e3, 3),
            "graph_value": round(sp / graphed / 1e9, 2),
            "one_launch_per_alignment_ms": round(single * 1e3, 3),
            "one_launch_per_alignment_value": round(sp / single / 1e9, 2),
            "speedup": round(single / graphed, 2), "identical_results": same,
            "unit": "G site-pairs/s"}


def small_alignments(count: int = 256, seed: int = 17):
    """`count` small alignments (20-300 taxa, 200-2000 sites, gaps 0-8 %, JC/K2P alternating),
    as jobs of the served-stream shape, with their character matrices."""
    import numpy as np
    from types import SimpleNamespace
    from rapidnj_amd.workloads import simulate_alignment
    rng = np.random.default_rng(seed)
    jobs, alns = [], []
    for k in range(count):
        n, L = int(rng.integers(20, 301)), int(rng.integers(200, 2001))
        jobs.append(SimpleNamespace(n=n, L=L, model="jc" if k % 2 else "k2p"))
        alns.append(simulate_alignment(n, L, float(rng.uniform(0, 0.08)), 9000 + k))
    return jobs, alns


def bootstrap_compare(ctx, dev) -> dict:
    """1000 x 5000 K2P, 100 column-weight replicates: one launch vs one launch per replicate."""
    import torch
    from rapidnj_amd.workloads import WORKLOADS, bootstrap_weights, simulate_alignment
    wl = WORKLOADS["bootstrap"]
    aln = simulate_alignment(wl.n, wl.L, wl.gap, wl.seed)
    A = ctx.alignment(aln)
    w = bootstrap_weights(wl.L, wl.replicates, seed0=0)
    pairs = wl.n * (wl.n - 1) // 2
    out = torch.empty((wl.replicates, pairs), dtype=torch.float32, device=dev)
    for _ in range(3):
        ctx.bootstrap(A, w, wl.model, out=out)
    torch.cuda.synchronize()
    k = 20
    t0 = time.perf_counter()
    for _ in range(k):
        ctx.bootstrap(A, w, wl.model, out=out)
    torch.cuda.synchronize()
    batched = (time.perf_counter() - t0) / k
    ref = out.clone()
    for r in range(wl.replicates):
        ctx.bootstrap(A, w[r:r + 1], wl.model, out=out[r:r + 1])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        for r in range(wl.replicates):
            ctx.bootstrap(A, w[r:r + 1], wl.model, out=out[r:r + 1])
    torch.cuda.synchronize()
    single = (time.perf_counter() - t0) / 2
    same = bool(torch.equal(out, ref))
    # the same launch captured into a HIP graph, replayed with fresh weights per call (the plan
    # alternates two graphs, one per pinned weight buffer: both replay once before the timing)
    sets = [bootstrap_weights(wl.L, wl.replicates, seed0=1000 * (i + 1)) for i in range(4)]
    with ctx.bootstrap_plan(A, wl.replicates, wl.model, out=out) as plan:
        plan.run(w)
        torch.cuda.synchronize()
        same = same and bool(torch.equal(out, ref))
        for ws_ in sets[:2]:
            plan.run(ws_)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(k):
            plan.run(sets[i % 4])
        torch.cuda.synchronize()
        graphed = (time.perf_counter() - t0) / k
    sp = pairs * wl.L * wl.replicates
    A.close()
    return {"workload": wl.name, "graph_ms": round(graphed * 1e3, 3),
            "graph_value": round(sp / graphed / 1e9, 2),
            "batched_ms": round(batched * 1e3, 3),
            "batched_value": round(sp / batched / 1e9, 2),
            "one_per_launch_ms": round(single * 1e3, 3),
            "one_per_launch_value": round(sp / single / 1e9, 2),
            "speedup": round(single / batched, 2), "identical_results": same,
            "unit": "G site-pairs/s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="large",
                    choices=["large", "plain", "small_jc", "bootstrap", "served", "protein"])
    ap.add_argument("--served-jobs", type=int, default=512,
                    help="jobs of the served stream (the first this many of the 512-job stream)")
    ap.add_argument("--tile-config", type=int, default=-1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the 2000 x 10000 side run")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (a known-traffic kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure HBM bytes per launch")
    ap.add_argument("--probe-traffic", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--side", choices=["ingest", "handoff", "nj", "valu", "jc", "naive"],
                    help="run only this side measurement on the workload and print its JSON")
    ap.add_argument("--nj-reference-taxa", type=int, default=5000,
                    help="with --side nj: the leading block (taxa) on which the reference's "
                         "rapidNJ is timed and its tree compared (default 5000; the whole "
                         "matrix takes the reference about a minute)")
    args = ap.parse_args()
    if args.probe_traffic:
        return probe_traffic(args)
    if args.side:
        return side_main(args)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: rapidnj_amd has no multi-GPU path (the distance "
              "launch runs on one MI355X); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np  # noqa: F401
    import torch

    from rapidnj_amd import native as N
    from rapidnj_amd.workloads import (WORKLOADS, bootstrap_weights, make_alignment, served_stream,
                                       simulate_alignment)

    served_jobs, served_alns = [], []
    if args.workload == "served" or not args.no_extra:
        count = max(1, min(512, args.served_jobs))
        served_jobs = served_stream()[:count]
        log(f"served stream: generating {count} jobs in a process pool")
        served_alns = generate_stream(served_jobs, cpu_cores())

    if args.workload == "served":
        return served_main(args, served_jobs, served_alns)

    wl = WORKLOADS[args.workload]
    log(f"workload {wl.name}: generating {wl.n} x {wl.L} (seed {wl.seed})")
    aln = make_alignment(wl)
    ctx = N.Context(0)
    if args.tile_config >= 0:
        ctx.set_tile_config(args.tile_config)
    A = ctx.alignment(aln)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    n, L = wl.n, wl.L
    pairs = n * (n - 1) // 2
    R = max(1, wl.replicates)
    site_pairs = pairs * L * R

    if wl.replicates:
        weights = bootstrap_weights(L, wl.replicates, seed0=0)
        out = torch.empty((R, pairs), dtype=torch.float32, device=dev)

        def step(raw=False):
            ctx.bootstrap(A, weights, wl.model, out=out, raw=raw, stream=stream)
    else:
        out = torch.empty(pairs, dtype=torch.float32, device=dev)

        def step(raw=False):
            ctx.distances(A, wl.model, out=out, raw=raw, stream=stream)

    log(f"warmup {args.warmup} (first launch per shape tunes the tile configuration)")
    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    ms = secs / args.steps * 1e3
    value = site_pairs / (ms / 1e3) / 1e9
    log(f"{ms:.3f} ms/step = {value:.1f} G site-pairs/s")

    # kernel duration: raw launches (statistics reset + tile kernel + job maximum, no patch)
    # between HIP
    # events recorded on the stream the library launches on
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(3, args.steps)
    step(raw=True)
    e0.record(stream)
    for _ in range(reps):
        step(raw=True)
    e1.record(stream)
    torch.cuda.synchronize()
    kern_ms = e0.elapsed_time(e1) / reps
    if wl.replicates:  # the bootstrap step also gathers replicate columns; report the whole launch
        kern_label = "bootstrap launch (gather + tile kernel)"
    else:
        kern_label = "tile_kernel (+ stats reset, job_max_kernel)"
    planes = A.info.nplanes
    ops = ops_per_word_pair(A, wl.model) * pairs * (L / 32.0) * R
    achieved = ops / (kern_ms / 1e3) / 1e12
    algo_bytes = pairs * 4 * R + n * planes * L / 8 * R
    cfg = ctx.tile_config_for(A, wl.model)
    pk = valu_peak(ctx, A, wl.model)
    pf = peak_fields(pk, achieved)
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": pf["peak"],
                "unit": "TOP/s", "frac": pf["frac"], "traffic": None,
                "ceiling": pf["ceiling"], "class_model": pf["class_model"],
                "kernel": kern_label, "kernel_ms": round(kern_ms, 4),
                "ops_per_launch": ops, "tile_config": N.tile_configs()[cfg],
                "hbm": {"achieved": round(algo_bytes / (kern_ms / 1e3) / 1e9, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "algo_bytes": algo_bytes}}

    if args.calibrate_copy_mib:
        nbytes = args.calibrate_copy_mib << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy_mib} MiB read + {args.calibrate_copy_mib} MiB "
            "written by one copy kernel")
        del src, dst

    extra = {}
    if not args.no_extra and args.workload == "large":
        pw = WORKLOADS["plain"]
        log(f"side run {pw.name}")
        p_aln = simulate_alignment(pw.n, pw.L, pw.gap, pw.seed)
        PA = ctx.alignment(p_aln)
        p_pairs = pw.n * (pw.n - 1) // 2
        p_out = torch.empty(p_pairs, dtype=torch.float32, device=dev)
        for _ in range(3):
            ctx.distances(PA, pw.model, out=p_out, stream=stream)
        torch.cuda.synchronize()
        k = 20
        t1 = time.perf_counter()
        for _ in range(k):
            ctx.distances(PA, pw.model, out=p_out, stream=stream)
        torch.cuda.synchronize()
        p_ms = (time.perf_counter() - t1) / k * 1e3
        extra["plain_2000x10000_k2p"] = {"ms_per_step": round(p_ms, 4),
                                         "value": round(p_pairs * pw.L / (p_ms / 1e3) / 1e9, 2),
                                         "unit": "G site-pairs/s"}
        PA.close()
        del p_out
        if not wl.protein and not wl.replicates:
            log(f"side run: naive character rules (the reference's -f counter), {wl.name}")
            extra["naive_rules"] = naive_side(ctx, aln, wl, out, stream)
            if wl.model != "jc":
                log(f"side run: JC on the same alignment (5 ops per 32 site-pairs), {wl.name}")
                extra["jc"] = model_side(ctx, A, "jc", False, wl, out, stream)
        log("side run: host hand-off of the 20000 x 5000 matrix (full rows, lower triangle)")
        extra["handoff"] = handoff_compare(ctx, A)
        log("side run: FASTA and Stockholm ingest (device readers) vs the reference readers")
        extra["ingest"] = ingest_compare(ctx, aln)
        log("side run: batched bootstrap vs one launch per replicate")
        extra["bootstrap"] = bootstrap_compare(ctx, dev)
        log("side run: neighbour joining on the device (20000 taxa; 5000-taxa block vs the reference)")
        extra["nj"] = nj_side(ctx, A, not args.no_cpu_baseline)
        log("side run: protein distances (Kimura protein, 20000 x 1250)")
        extra["protein"] = protein_side(ctx, stream, not args.no_cpu_baseline)
        log(f"side run: served stream, {len(served_jobs)} jobs")
        extra["served"] = {
            "stream": f"{len(served_jobs)} jobs of the 512-job stream (Poisson 20/s; "
                      "taxa log-uniform 50-5000, sites 500-30000, gaps 0-8%, JC/K2P 50/50)",
            "open_loop": served_run(ctx, served_jobs, served_alns, open_loop=True),
            "closed_loop": served_run(ctx, served_jobs, served_alns, open_loop=False),
            "resident": served_resident(ctx, served_jobs, served_alns),
            "one_job_per_launch": one_job_per_launch(ctx, served_jobs, served_alns)}
        log("side run: 64 uneven alignments in one captured launch vs one launch each")
        extra["uneven_batch"] = uneven_batch_compare(ctx, served_jobs, served_alns)
        log("side run: 256 small alignments in one captured launch vs one launch each")
        sj, sa = small_alignments()
        extra["uneven_batch_small"] = uneven_batch_compare(ctx, sj, sa, count=len(sj))

    if not args.no_traffic and not wl.replicates:
        log("HBM traffic: two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE)")
        t = measure_traffic(args.workload, cfg)
        if t is not None:
            roofline["traffic"] = round(t["bytes"])
            roofline["traffic_detail"] = {k: (round(v) if isinstance(v, float) and v > 1e3 else v)
                                          for k, v in t.items() if k != "bytes"}
            roofline["traffic_detail"]["algo_bytes"] = algo_bytes

    base = None
    if not args.no_cpu_baseline:
        cpus = host_cpus()
        threads = cpus["usable"]
        # every usable core on the whole workload (a few seconds on 16 cores), then one thread on
        # the first half of the taxa (a quarter of the pairs: 10-15 s)
        sample_n = min(n, 20000)
        log(f"CPU baseline on {sample_n} taxa with {threads} threads ({cpus['model']})")
        base = cpu_baseline(aln, wl.model, sample_n, threads, protein=wl.protein)
        one_n = min(n, 10000)
        log(f"CPU baseline on {one_n} taxa with 1 thread")
        one = cpu_baseline(aln, wl.model, one_n, 1, protein=wl.protein)
        base.update({"cpu_model": cpus["model"], "cores_affinity": cpus["affinity"],
                     "cgroup_cpu_quota": cpus["cgroup_quota"],
                     "omp_num_threads": cpus["omp_num_threads"],
                     "single_thread": {k: one[k] for k in ("value", "cores", "sample")},
                     "gpu_speedup": round(value / base["value"], 1),
                     "gpu_speedup_vs_single_thread": round(value / one["value"], 1)})

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "G site-pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded %s-evolved alignment on a random-joining tree)"
                % ("20-state" if wl.protein else "K2P"),
        "config": {"workload": wl.name, "taxa": n, "sites": L, "gap": wl.gap, "model": wl.model,
                   "replicates": wl.replicates, "output": "float32 upper triangle in HBM",
                   "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": base,
    }
    if extra:
        line["extra"] = extra
    print(json.dumps(line), flush=True)
    A.close()
    ctx.close()
    return 0


def served_main(args, jobs, alns) -> int:
    """--workload served: the served stream as the measured workload."""
    from rapidnj_amd import native as N
    ctx = N.Context(0)
    warm = served_run(ctx, jobs[:8], alns[:8], open_loop=False)  # tunes shape classes once
    log(f"warm: {warm}")
    opened = served_run(ctx, jobs, alns, open_loop=True)
    closed = served_run(ctx, jobs, alns, open_loop=False)
    resident = served_resident(ctx, jobs, alns)
    single = one_job_per_launch(ctx, jobs, alns)
    line = {
        "metric": METRIC, "value": closed["value"], "unit": "G site-pairs/s", "n_gpus": 1,
        "steps": len(jobs), "warmup": 8, "ms_per_step": round(closed["makespan_s"] * 1e3 / len(jobs), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded generator per job)",
        "config": {"workload": f"served stream, {len(jobs)} jobs, Poisson 20/s",
                   "parallelism": "single GPU"},
        "served": {"open_loop": opened, "closed_loop": closed, "resident": resident,
                   "one_job_per_launch": single},
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
