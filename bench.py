#!/usr/bin/env python3
"""Benchmark of the MI355X SBWT-Search index search (one GPU).

One step = one pass of the fused index search (2-bit packing, k-mer offsets,
SBWT search with Poppy-style rank, invalid/not-found classification) over one
batch of reads already resident in HBM, through the C ABI
(sbwt_searcher_search_reads_device). Default workload: the plain index search
at scale (~95 M-node union SBWT, 10 M reads x 150 bp); see
sbwt-search_amd/bindings/configs.py.

Prints one JSON line with the throughput, the roofline of the search kernel
(exact algorithmic bytes from a census run / its hipEvent-timed duration) and
a CPU baseline (the C/OpenMP restatement of the reference algorithm on the
host cores, on a bounded sample of the same reads).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "sbwt-search_amd", "bindings"))

METRIC = "k-mers/s (index) + reads/s (color) on 1 MI355X; batched vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0
LINE = 64  # bytes per rank-block line / presearch line


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2"])
    ap.add_argument("--reads", type=int, default=0, help="override the number of reads")
    ap.add_argument("--cpu-sample-reads", type=int, default=200_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--presearch-letters", type=int, default=12)
    ap.add_argument("--algorithm", default="streaming", choices=["streaming", "per-kmer"])
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run an elementwise copy of this many MiB (a kernel of known "
                         "HBM traffic, to scale rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def traffic_from_profile(name):
    """HBM bytes per search launch from the committed rocprofv3 --pmc summary."""
    path = os.path.join(REPO, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get("hbm_bytes_per_launch")
    except (ValueError, OSError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the search path runs on a single MI355X here and "
              f"--gpus {args.gpus} is not supported (see DESIGN.md for how it would shard).",
              file=sys.stderr)
        return 2
    import configs
    import sbwt_amd as S
    import torch
    import workloads as W

    t0 = time.time()
    if args.config == "c1":
        wl = configs.c1(**({"n_reads": args.reads} if args.reads else {}))
    else:
        wl = configs.c2(**({"n_reads": args.reads} if args.reads else {}))
    log(f"workload {wl.name}: {wl.n_reads} reads, {wl.n_chars} chars ({time.time() - t0:.1f} s)")
    t0 = time.time()
    built = W.BuiltSbwt(wl.genomes, wl.genome_ends, configs.K)
    log(f"SBWT: {built.n} nodes, {built.n_kmers} k-mers ({time.time() - t0:.1f} s)")
    bits = [built.bits(c) for c in range(4)]
    t0 = time.time()
    idx = S.Index.create(bits, built.n, configs.K, presearch_letters=args.presearch_letters,
                         suffix_group_starts=built.bits(4))
    log(f"device index: {idx.device_bytes / 2**20:.1f} MiB, wide={idx.wide} "
        f"({time.time() - t0:.1f} s)")

    dev = torch.device("cuda:0")
    d_chars = torch.from_numpy(wl.chars).to(dev)
    d_ends = torch.from_numpy(wl.read_ends.view(np.int64)).to(dev)
    d_out = torch.empty(wl.n_chars, dtype=torch.int64, device=dev)
    n_kmers = wl.n_kmers()
    s = idx.searcher(wl.n_chars, wl.n_reads)
    s.set_algorithm(S.ALGO_STREAMING if args.algorithm == "streaming" else S.ALGO_PER_KMER)

    def step():
        s.search_reads_device(d_chars.data_ptr(), wl.n_chars, d_ends.data_ptr(), wl.n_reads,
                              d_out.data_ptr())

    # census (untimed): exact work of this batch for the algorithmic byte count
    s.set_stats(True)
    step()
    s.sync()
    stats = s.last_stats()
    s.set_stats(False)
    resolved = (stats["full_searches"] + stats["streamed"] + stats["invalid"]
                - stats["stream_fallbacks"])
    assert resolved == n_kmers, stats
    # one 64-byte line per presearch lookup, SBWT step and streamed step, two per
    # key-walk step (mark + rank block); u64 output; packed input and invalid mask
    lines = (stats["full_searches"] + stats["steps"] + stats["streamed"]
             + stats["stream_extra_lines"] + 2 * stats["walk_steps"])
    algo_bytes = LINE * lines + 8 * n_kmers + 0.375 * wl.n_chars

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    s.sync()
    search_ms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step()
        search_ms.append(s.last_timings()["search_ms"])
    s.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1000.0 * elapsed / args.steps
    value = n_kmers / (elapsed / args.steps)
    kern_ms = float(np.mean(search_ms))
    achieved = algo_bytes / (kern_ms / 1000.0) / 1e9
    res = d_out[:n_kmers].cpu().numpy().view(np.uint64)
    found = int(np.count_nonzero(res < np.uint64(S.INVALID)))
    log(f"{ms_per_step:.2f} ms/step, search kernel {kern_ms:.2f} ms, {value / 1e9:.3f} G k-mers/s,"
        f" found {found / n_kmers:.3f}")

    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib << 20
        src = torch.ones(n // 8, dtype=torch.int64, device=dev)
        dst = torch.empty_like(src)
        torch.add(src, 0, out=dst)  # reads n bytes, writes n bytes
        torch.cuda.synchronize()
        log(f"calibration copy: {n} bytes read, {n} bytes written")
        del src, dst

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, wl, bits, built, res)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": wl.name + " " + wl.description,
            "sbwt_nodes": int(built.n),
            "kmers_per_step": n_kmers,
            "reads_per_step": wl.n_reads,
            "chars_per_step": wl.n_chars,
            "found_fraction": found / n_kmers,
            "presearch_letters": idx.presearch_letters,
            "algorithm": args.algorithm,
            "search_kernel_ms": kern_ms,
            "reads_per_s": wl.n_reads / (elapsed / args.steps),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profile(f"search_reads_{wl.name.lower()}_pmc.json"),
            "algorithmic_bytes_per_launch": algo_bytes,
            "census": stats,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(args, wl, bits, built, gpu_res):
    """The oracle (C/OpenMP restatement of the reference) on a bounded sample."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_py as O
    n_sample = min(args.cpu_sample_reads, wl.n_reads)
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    end_char = int(wl.read_ends[n_sample - 1])
    chars = wl.chars[:end_char].tobytes()
    ends = wl.read_ends[:n_sample].copy()
    orc = O.OracleSbwt(acgt=bits, num_bits=built.n, k=configs_k())
    orc.presearch(12)
    t0 = time.perf_counter()
    ref = orc.index_reads(chars, ends, kp=12, threads=threads)
    dt = time.perf_counter() - t0
    if not np.array_equal(ref, gpu_res[:len(ref)]):
        raise SystemExit("GPU results differ from the CPU oracle on the baseline sample")
    log(f"cpu baseline: {len(ref)} k-mers in {dt:.2f} s on {threads} threads")
    return {"value": len(ref) / dt, "unit": "k-mers/s", "cores": threads, "kind": "port",
            "sample": f"first {n_sample} reads ({len(ref)} k-mers) of the same batch; "
                      "pack + positions + presearch-12 search + classify, results checked "
                      "equal to the GPU's"}


def configs_k():
    import configs
    return configs.K


if __name__ == "__main__":
    sys.exit(main())
