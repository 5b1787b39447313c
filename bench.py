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
LINE = 128  # bytes per random access: one MI355X L2 line (node line or presearch line)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c5"],
                    help="index mode workload: c2 (default, the headline), c1, or c5 (the "
                         "serving index, HBM-resident node lines)")
    ap.add_argument("--mode", default="index",
                    choices=["index", "batched", "serve", "colors", "cli", "cli-colors",
                             "gather-probe"],
                    help="index: the headline line (default); batched: C4 plain vs batched "
                         "form; serve: Poisson request stream, served vs plain sequential; "
                         "colors: the color stage on the C3 recipe; cli: the sbwt_search "
                         "command end to end on C2 written as FASTA and FASTQ; cli-colors: "
                         "`sbwt_search colors` on C3 index-result files (ascii and binary); "
                         "gather-probe: the random 128-byte line ceiling over working-set sizes")
    # the C3 recipe: 1000 colors of 2-5 Mbp genomes from 20 roots (4.09 G nodes, built on the GPU)
    ap.add_argument("--color-genomes", type=int, default=1000)
    ap.add_argument("--color-roots", type=int, default=20)
    ap.add_argument("--color-min-len", type=int, default=2_000_000)
    ap.add_argument("--color-max-len", type=int, default=5_000_000)
    ap.add_argument("--color-reads", type=int, default=5_000_000)
    ap.add_argument("--batch-chars", type=int, default=600_000_000)
    ap.add_argument("--long-reads", type=int, default=1_000_000)
    ap.add_argument("--short-reads", type=int, default=4_000_000)
    ap.add_argument("--serve-genomes", type=int, default=200)
    ap.add_argument("--serve-roots", type=int, default=10)
    ap.add_argument("--serve-length", type=int, default=5_000_000)
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rates", default="50,1000",
                    help="Poisson arrival rates (requests/s) to run, comma separated")
    ap.add_argument("--serve-max-reads", type=int, default=100_000)
    ap.add_argument("--reads", type=int, default=0, help="override the number of reads")
    ap.add_argument("--cli-runs", default="",
                    help="cli mode: comma list of input:output:parsing runs (e.g. "
                         "fastq:bool:device); default all")
    ap.add_argument("--cpu-sample-reads", type=int, default=1_500_000,
                    help="reads of the batch timed on the CPU oracle (about 10 s of host work)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--leg-timeout", type=int, default=300)
    ap.add_argument("--legs", default="colors,batched,serve",
                    help="index mode: the other legs of the metric (colors: C3 reads/s; batched: "
                         "C4 batched vs plain; serve: C5 p50/p99), each run first as a child "
                         "process and summarized under \"legs\" in the one JSON line; empty: none")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kmer-table", action="store_true",
                    help="search without the index's k-mer table (fresh searches from the "
                         "presearch table and SBWT steps)")
    ap.add_argument("--presearch-letters", type=int, default=0,
                    help="presearch prefix letters (0: the library's choice for the index size; "
                         "the reference fixes 12)")
    ap.add_argument("--algorithm", default="streaming", choices=["streaming", "per-kmer", "auto"],
                    help="auto: the library's default (per k-mer below 2^22 chars, streaming "
                         "with work items sized to the batch above)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape instead of autotuning it on the batch")
    ap.add_argument("--launch", default="",
                    help="CHUNK,BLOCKS: fixed launch shape of the streaming search (no autotune)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run an elementwise copy of this many MiB (a kernel of known "
                         "HBM traffic, to scale rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def traffic_from_profile(config, algorithm, presearch_letters, kmer_table):
    """HBM bytes per search launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a
    copy of known size) from the committed summary of the same workload, or None."""
    path = os.path.join(REPO, "profiles", f"search_{algorithm}_{config}_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
    except (ValueError, OSError):
        return None
    if (d.get("config"), d.get("algorithm"), d.get("presearch_letters"),
            d.get("kmer_table", False)) != (config, algorithm, presearch_letters, kmer_table):
        return None
    return d.get("hbm_bytes_per_launch")


def main():
    global _ARGS
    args = parse()
    _ARGS = args
    if args.gpus != 1:
        print("bench.py measures one GPU: the search path runs on a single MI355X here and "
              f"--gpus {args.gpus} is not supported (the path would shard by replicating the "
              "index on every GPU and splitting the reads).",
              file=sys.stderr)
        return 2
    import configs
    import sbwt_amd as S
    import torch
    import workloads as W

    if args.mode == "batched":
        return mode_batched(args, configs, S, torch, W)
    if args.mode == "serve":
        return mode_serve(args, configs, S, torch, W)
    if args.mode == "cli":
        return mode_cli(args, configs, S, torch, W)
    if args.mode == "colors":
        return mode_colors(args, configs, S, torch, W)
    if args.mode == "cli-colors":
        return mode_cli_colors(args, configs, S, torch, W)
    if args.mode == "gather-probe":
        return mode_gather_probe(args, S)
    legs = run_legs(args) if args.legs else None
    t0 = time.time()
    if args.config == "c1":
        wl = configs.c1(**({"n_reads": args.reads} if args.reads else {}))
    elif args.config == "c5":
        wl = configs.c5_index(**({"n_reads": args.reads} if args.reads else {}))
    else:
        wl = configs.c2(**({"n_reads": args.reads} if args.reads else {}))
    log(f"workload {wl.name}: {wl.n_reads} reads, {wl.n_chars} chars ({time.time() - t0:.1f} s)")
    built, _ = _gpu_built(W, wl.genomes, wl.genome_ends, configs.K)
    bits = [built.bits(c) for c in range(4)]
    t0 = time.time()
    idx = S.Index.create(bits, built.n, configs.K, presearch_letters=args.presearch_letters,
                         suffix_group_starts=built.bits(4))
    log(f"device index: {idx.device_bytes / 2**20:.1f} MiB, wide={idx.wide}, "
        f"presearch {idx.presearch_letters} letters "
        f"({time.time() - t0:.1f} s)")
    table = _kmer_table(args, idx)

    dev = torch.device("cuda:0")
    d_chars = torch.from_numpy(wl.chars).to(dev)
    d_ends = torch.from_numpy(wl.read_ends.view(np.int64)).to(dev)
    d_out = torch.empty(wl.n_chars, dtype=torch.int64, device=dev)
    n_kmers = wl.n_kmers()
    s = idx.searcher(wl.n_chars, wl.n_reads)
    s.set_algorithm({"streaming": S.ALGO_STREAMING, "per-kmer": S.ALGO_PER_KMER,
                     "auto": S.ALGO_AUTO}[args.algorithm])

    def step():
        s.search_reads_device(d_chars.data_ptr(), wl.n_chars, d_ends.data_ptr(), wl.n_reads,
                              d_out.data_ptr())

    launch = None
    if args.algorithm == "streaming" and args.launch:
        ch, bpc = (int(x) for x in args.launch.split(","))
        s.set_launch(ch, bpc)
        launch = {"chunk_kmers": ch, "blocks_per_cu": bpc}
    elif args.algorithm == "streaming" and not args.no_autotune:
        ch, bpc, ms = s.autotune(d_chars.data_ptr(), wl.n_chars, d_ends.data_ptr(), wl.n_reads,
                                 d_out.data_ptr())
        launch = {"chunk_kmers": ch, "blocks_per_cu": bpc}
        log(f"autotune: {ch} k-mers per work item, {bpc} workgroups per CU ({ms:.2f} ms)")
    elif args.algorithm == "streaming":
        ch, bpc = s.launch()
        launch = {"chunk_kmers": ch, "blocks_per_cu": bpc}

    # census (untimed): exact work of this batch for the algorithmic byte count
    s.set_stats(True)
    step()
    s.sync()
    stats = s.last_stats()
    s.set_stats(False)
    assert S.Searcher.resolved(stats) == n_kmers, stats
    # one 128-byte line per presearch lookup (searches and absence checks), SBWT
    # step, streamed step, extra group-start line and key-walk step; u64 output;
    # packed input and invalid mask
    lines = S.Searcher.lines(stats)
    algo_bytes = LINE * lines + 8 * n_kmers + 0.375 * wl.n_chars

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    s.sync()
    search_ms, pack_ms, offsets_ms = [], [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step()
        tm = s.last_timings()
        search_ms.append(tm["search_ms"])
        pack_ms.append(tm["pack_ms"])
        offsets_ms.append(tm["offsets_ms"])
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
    # the random-line ceiling: at the node lines' working set (what most of the
    # search's lines hit) and at an HBM-resident size
    node_line_bytes = (int(built.n) // 96 + 1) * 128
    probes = {}
    for name, ws in (("node_lines", node_line_bytes), ("hbm_32GiB", 32 << 30)):
        best = max((S.random_line_probe(ws, bpc) for bpc in (4, 8)), key=lambda r: r["gbs"])
        probes[name] = {"working_set_bytes": best["table_bytes"], "gbs": best["gbs"],
                        "lines_per_s": best["lines_per_s"]}
        log(f"random-line probe {name}: {best['gbs']:.0f} GB/s "
            f"({best['lines_per_s'] / 1e9:.1f} G lines/s over {best['table_bytes'] >> 20} MiB)")

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
            "kmer_table": table,
            "algorithm": args.algorithm,
            "launch": launch,
            "search_kernel_ms": kern_ms,
            "pack_ms": float(np.mean(pack_ms)),
            "offsets_and_work_items_ms": float(np.mean(offsets_ms)),
            "reads_per_s": wl.n_reads / (elapsed / args.steps),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            # the lines the search reads are mostly node lines (a 128 B line per
            # random access) of a structure that fits the 256 MB Infinity Cache:
            # the measured ceiling of random lines at that size is the relevant one
            "regime": ("cache-assisted: node lines of %d MiB fit the 256 MB Infinity Cache"
                       if node_line_bytes <= (256 << 20) else
                       "HBM: node lines of %d MiB exceed the 256 MB Infinity Cache")
                      % (node_line_bytes >> 20),
            "random_line_ceiling": probes,
            "frac_of_random_line_ceiling": achieved / probes["node_lines"]["gbs"],
            "frac_of_hbm_random_line_ceiling": achieved / probes["hbm_32GiB"]["gbs"],
            "traffic": traffic_from_profile(wl.name.lower(), args.algorithm.replace("-", "_"),
                                            idx.presearch_letters, table["built"]),
            "algorithmic_bytes_per_launch": algo_bytes,
            # the reference algorithm's cost model (one 64-byte line per presearch and
            # per SBWT step of every valid k-mer, u64 out): what an independent search
            # of each k-mer would read
            "reference_model_bytes_per_launch": (64 * (configs.K - 12) + 64 + 8)
            * (n_kmers - stats["invalid"]) + 8 * stats["invalid"],
            "census": stats,
        },
        "cpu_baseline": cpu,
    }
    if legs is not None:
        out["legs"] = legs
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(args, wl, bits, built, gpu_res):
    """The oracle (C/OpenMP restatement of the reference) on a bounded sample."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_py as O
    n_sample = min(args.cpu_sample_reads, wl.n_reads)
    threads = args.cpu_threads or host_threads()
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
    # one thread on a tenth of the sample
    n1 = max(1, n_sample // 10)
    e1 = int(wl.read_ends[n1 - 1])
    t0 = time.perf_counter()
    ref1 = orc.index_reads(wl.chars[:e1].tobytes(), wl.read_ends[:n1].copy(), kp=12, threads=1)
    dt1 = time.perf_counter() - t0
    log(f"cpu baseline: {len(ref1)} k-mers in {dt1:.2f} s on 1 thread")
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": len(ref) / dt, "unit": "k-mers/s", "cores": threads, "kind": "port",
            "sample": f"first {n_sample} reads ({len(ref)} k-mers) of the same batch; "
                      "pack + positions + presearch-12 search + classify, results checked "
                      "equal to the GPU's",
            "value_1_thread": len(ref1) / dt1,
            "host": f"{model}, {os.cpu_count()} logical cpus"}


_ARGS = None


def host_threads():
    """Every CPU this process may run on (all the host cores it is given)."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def _leg_summary(leg, r):
    """The figures of a leg's JSON line that the headline line carries."""
    if leg == "colors":
        return {"metric": "reads/s (color)", "value": r["value"], "unit": "reads/s",
                "config": r["config"]["workload"], "sbwt_nodes": r["config"]["sbwt_nodes"],
                "form": r["form"], "color_stage_reads_per_s": r["color_stage_reads_per_s"],
                "index_stage_kmers_per_s": r["index_stage_kmers_per_s"],
                "key_node_form_reads_per_s": r["key_node_form"]["reads_per_s"],
                "outputs_identical": r["outputs_identical"],
                "index_stage_roofline_frac": r["index_roofline"]["frac"],
                "index_stage_frac_of_random_line_ceiling":
                    r["index_roofline"]["frac_of_random_line_ceiling"],
                "color_kernel_roofline_frac": r["color_roofline"]["color_table"]["frac"],
                "color_kernel_frac_of_random_line_ceiling":
                    r["color_roofline"]["color_table"]["frac_of_random_line_ceiling"],
                "color_kernel_ms": r["color_kernel_ms"],
                "key_runs_roofline_frac": r["key_runs_roofline"]["frac"],
                "seconds_per_step": r["seconds"]}
    if leg == "batched":
        keep = ("reads", "batches", "plain_s", "batched_unfused_s", "batched_s", "speedup",
                "speedup_unfused", "plain_reads_per_s", "batched_reads_per_s")
        return {"metric": "reads/s (index + color), batched vs plain",
                "config": r["config"]["workload"],
                **{k: r[k] for k in keep if k in r}}
    if leg == "serve":
        runs = []
        for run in r["runs"]:
            runs.append({"rate_per_s": run["rate_per_s"], "returns": run["returns"],
                         "requests": run["requests"], "reads": run["reads"],
                         **{form: {k: run[form][k] for k in ("p50_ms", "p99_ms", "reads_per_s")}
                            for form in ("served", "served_no_graphs", "plain")},
                         "outputs_equal": run["outputs_equal"]})
        return {"metric": "p50/p99 request latency + reads/s, served vs plain",
                "config": r["config"], "latency_from": "scheduled arrival", "runs": runs}
    return r


def run_legs(args):
    """The metric's other legs (C3 color search, C4 batched vs plain, C5 serving),
    each as a child process on the same GPU before the headline measurement;
    their stderr streams through, their JSON lines are summarized."""
    import subprocess
    out = {}
    for leg in [x for x in args.legs.split(",") if x]:
        cmd = [sys.executable, "-u", os.path.abspath(__file__), "--mode", leg, "--steps", "3",
               "--warmup", "1"]
        log(f"leg {leg}: {' '.join(cmd[2:])}")
        t0 = time.time()
        try:
            p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True,
                               timeout=args.leg_timeout)
            lines = [x for x in p.stdout.splitlines() if x.startswith("{")]
            if p.returncode != 0 or not lines:
                out[leg] = {"error": f"exit status {p.returncode}"}
            else:
                out[leg] = _leg_summary(leg, json.loads(lines[-1]))
        except subprocess.TimeoutExpired:
            out[leg] = {"error": f"no result within {args.leg_timeout} s"}
        out[leg]["wall_s"] = time.time() - t0
        log(f"leg {leg}: done in {out[leg]['wall_s']:.0f} s")
    return out


def _kmer_table(args, idx):
    """Builds the index's k-mer table (once per index, before any timing) unless
    --no-kmer-table; returns {"built": bool, "build_ms": float}."""
    if args is not None and args.no_kmer_table:
        return {"built": False, "build_ms": 0.0}
    t0 = time.perf_counter()
    built = idx.build_kmer_table()
    ms = (time.perf_counter() - t0) * 1e3
    log(f"k-mer table: {'built' if built else 'not built'} ({ms:.0f} ms)")
    return {"built": built, "build_ms": ms}


def _gpu_built(W, codes, ends, k, color_of=None, d=20):
    """SBWT (and color index) of the genomes from the GPU builder (identical to
    the host builder's); construction is set-up, outside every timed region."""
    t0 = time.time()
    built, col, st = W.gpu_build(codes, ends, k, True,
                                 None if color_of is None else np.asarray(color_of, np.uint32), d)
    log(f"GPU build: {st['nodes']} nodes, {st['kmers']} k-mers, {st['sets']} color sets, "
        f"{st['ms_total']:.0f} ms on the device, peak {st['peak_device_bytes'] / 2**30:.1f} GiB "
        f"({time.time() - t0:.1f} s)")
    return built, col


def _colored_index(S, built, col, tag):
    """Device index + color index of a built SBWT and color index (the colors
    written to a temporary .tcolors and loaded)."""
    import tempfile
    t0 = time.time()
    tmp = tempfile.mkdtemp(prefix=f"sbwt_{tag}_")
    path = os.path.join(tmp, "index.tcolors")
    col.write(path)
    km = col.key_marks()
    idx = S.Index.create([built.bits(c) for c in range(4)], built.n, 31, key_marks=km,
                         suffix_group_starts=built.bits(4))
    _kmer_table(_ARGS, idx)
    cidx = S.ColorIndex(path)
    log(f"colors: {col.num_colors} colors, {col.num_sets} sets ({col.num_dense} dense), "
        f"{cidx.device_bytes / 2**30:.1f} GiB on the device, {time.time() - t0:.1f} s")
    os.remove(path)
    os.rmdir(tmp)
    return idx, cidx


def mode_batched(args, configs, S, torch, W):
    """C4: plain per-batch launches vs the batched persistent form (reads/s)."""
    import modes
    t0 = time.time()
    wl = configs.c4(args.long_reads, args.short_reads)
    log(f"workload {wl.name}: {wl.n_reads} reads, {wl.n_chars} chars ({time.time() - t0:.1f} s)")
    built, col = _gpu_built(W, wl.genomes, wl.genome_ends, configs.K,
                            np.arange(len(wl.genome_ends)), 20)
    idx, cidx = _colored_index(S, built, col, "c4")
    t0 = time.perf_counter()
    kind = idx.build_color_table(cidx)
    log(f"color table: {'masks' if kind == 1 else 'key ranks'} "
        f"({1000 * (time.perf_counter() - t0):.0f} ms)")
    r = modes.run_batched(S, torch, wl, idx, cidx, args.batch_chars, log=log)
    print(json.dumps({"metric": "reads/s (index + color), batched vs plain", "mode": "batched",
                      "config": {"workload": wl.name + " " + wl.description,
                                 "sbwt_nodes": int(built.n)}, **r}), flush=True)
    return 0


def mode_serve(args, configs, S, torch, W):
    """C5-style serving: Poisson requests, served (continuous batching) vs plain sequential."""
    import modes
    t0 = time.time()
    codes, ends = configs.c5_genomes(args.serve_genomes, args.serve_roots, args.serve_length)
    built, col = _gpu_built(W, codes, ends, configs.K, np.arange(len(ends)), 20)
    log(f"serve index: {built.n} nodes ({time.time() - t0:.1f} s)")
    idx, cidx = _colored_index(S, built, col, "serve")

    def reads(seed, n):
        lens = W.uniform_lengths(seed, n, 75, 250)
        c, e, _ = W.reads(seed, codes, ends, lens, random_frac=0.0, err_rate=0.01)
        return c, e

    runs = []
    for rate in (float(x) for x in args.serve_rates.split(",")):
        for kr in (False, True):
            runs.append(modes.run_serve(S, idx, cidx, reads, rate, args.serve_requests,
                                        args.serve_max_reads, 5, S.color_params(0.7), log=log,
                                        kmer_results=kr))
    print(json.dumps({"metric": "p50/p99 request latency + reads/s, served vs plain",
                      "mode": "serve", "config": {
                          "index": f"{args.serve_genomes} genomes x {args.serve_length} bp from "
                                   f"{args.serve_roots} roots (3% substitutions), "
                                   f"{built.n} nodes, {args.serve_genomes} colors",
                          "requests": "Poisson, N log-uniform in [1, max] reads of U[75,250] bp"},
                      "runs": runs}), flush=True)
    return 0


def mode_colors(args, configs, S, torch, W):
    """C3 recipe: index search with the key-k-mer move, then the color stage (reads/s)."""
    t0 = time.time()
    wl = configs.c3(args.color_reads, args.color_genomes, args.color_roots, args.color_min_len,
                    args.color_max_len)
    log(f"workload {wl.name}: {wl.n_reads} reads ({time.time() - t0:.1f} s)")
    built, col = _gpu_built(W, wl.genomes, wl.genome_ends, configs.K,
                            np.arange(len(wl.genome_ends)), 20)
    idx, cidx = _colored_index(S, built, col, "c3")
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()
    s = idx.searcher(wl.n_chars, wl.n_reads, move_to_key_kmer=True, stream=stream.cuda_stream)
    cs = cidx.searcher(wl.n_chars, wl.n_reads, stream=stream.cuda_stream)
    d_chars = torch.from_numpy(wl.chars).to(dev)
    d_ends = torch.from_numpy(wl.read_ends.view(np.int64)).to(dev)
    d_res = torch.empty(wl.n_chars, dtype=torch.int64, device=dev)
    d_off = torch.empty(wl.n_reads + 1, dtype=torch.int64, device=dev)
    cap = min(wl.n_reads * cidx.num_colors, wl.n_chars)
    d_ids = torch.empty(cap, dtype=torch.int32, device=dev)
    params = S.color_params(0.7)

    def run(fused):
        """W + K passes of both stages; the key-node form (fused False) or the
        color-table form (the index stage writes key ranks or masks)."""
        s.set_color_output(fused)
        cs.set_input(idx if fused else None)
        times = {"index": [], "color": [], "total": []}
        n_ids = 0
        for it in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s.search_reads_device(d_chars.data_ptr(), wl.n_chars, d_ends.data_ptr(), wl.n_reads,
                                  d_res.data_ptr())
            stream.synchronize()
            t1 = time.perf_counter()
            n_ids = cs.pseudoalign_device(d_res.data_ptr(), s.kmer_ends_device(), wl.n_reads,
                                          params, d_off.data_ptr(), d_ids.data_ptr(), cap)
            stream.synchronize()
            t2 = time.perf_counter()
            if it >= args.warmup:
                times["index"].append(t1 - t0)
                times["color"].append(t2 - t1)
                times["total"].append(t2 - t0)
        med = {k: float(np.median(v)) for k, v in times.items()}
        lists = (d_off.cpu().numpy().copy(), d_ids[:n_ids].cpu().numpy().copy())
        tm = s.last_timings()
        return med, n_ids, lists, tm, cs.last_time()

    t0 = time.perf_counter()
    kind = idx.build_color_table(cidx)
    table_ms = 1000 * (time.perf_counter() - t0)
    log(f"color table: {'masks' if kind == 1 else 'key ranks'} ({table_ms:.0f} ms)")
    # launch shape of the streaming search with the key move and color output
    s.set_color_output(True)
    launch = None
    if not args.no_autotune:
        ch, bpc, tms = s.autotune(d_chars.data_ptr(), wl.n_chars, d_ends.data_ptr(), wl.n_reads,
                                  d_res.data_ptr())
        launch = {"chunk_kmers": ch, "blocks_per_cu": bpc}
        log(f"autotune: {ch} k-mers per work item, {bpc} workgroups per CU ({tms:.2f} ms)")
        # work-item shape of the color stage, timed on the color-table form's input
        cs.set_input(idx)
        s.search_reads_device(d_chars.data_ptr(), wl.n_chars, d_ends.data_ptr(), wl.n_reads,
                              d_res.data_ptr())
        s.sync()
        items, ims = cs.autotune(d_res.data_ptr(), s.kmer_ends_device(), wl.n_reads, params,
                                 d_off.data_ptr(), d_ids.data_ptr(), cap)
        launch["color_item_kmers"] = items
        log(f"color autotune: {items} k-mers per work item ({ims:.2f} ms)")
    med_f, n_ids_f, lists_f, tm_f, color_ms_f = run(True)
    med, n_ids, lists, tm, color_ms = run(False)
    same = n_ids == n_ids_f and all(np.array_equal(a, b) for a, b in zip(lists, lists_f))
    log(f"key nodes: {wl.n_reads / med['total'] / 1e6:.1f} M reads/s; color table: "
        f"{wl.n_reads / med_f['total'] / 1e6:.1f} M reads/s; lists identical: {same}")
    if not same:
        raise SystemExit("the color-table form's color lists differ from the key-node form's")
    # census of the index stage in the color-table form (untimed): lines read by
    # the search and the key move
    search_ms = tm_f["search_ms"]
    s.set_color_output(True)
    s.set_stats(True)
    s.search_reads_device(d_chars.data_ptr(), wl.n_chars, d_ends.data_ptr(), wl.n_reads,
                          d_res.data_ptr())
    s.sync()
    stats = s.last_stats()
    s.set_stats(False)
    n_kmers = wl.n_kmers()
    algo_bytes = LINE * S.Searcher.lines(stats) + 8 * n_kmers + 0.375 * wl.n_chars
    node_line_bytes = (int(built.n) // 96 + 1) * 128
    probe = max((S.random_line_probe(min(node_line_bytes, 32 << 30), bpc) for bpc in (4, 8)),
                key=lambda r: r["gbs"])
    achieved = algo_bytes / (search_ms / 1e3) / 1e9
    log(f"index stage census: {S.Searcher.lines(stats) / n_kmers:.2f} lines per k-mer, "
        f"{achieved:.0f} GB/s over {search_ms:.2f} ms; random-line ceiling at "
        f"{probe['table_bytes'] >> 20} MiB: {probe['gbs']:.0f} GB/s")
    # the key-runs pass on its own: each k-mer's result read and rewritten (16 B)
    # and one color-table line per chain end (the census' walk steps)
    kr_ms = tm_f["key_runs_ms"]
    kr_bytes = 16 * n_kmers + LINE * stats["walk_steps"]
    kr_gbs = kr_bytes / (kr_ms / 1e3) / 1e9 if kr_ms > 0 else None
    if kr_gbs:
        log(f"key runs: {kr_ms:.2f} ms, {stats['walk_steps'] / n_kmers:.2f} chain-end lines per "
            f"k-mer, {kr_gbs:.0f} GB/s by census")
    # census of the color stage in both forms (untimed): key-set, key-mark and
    # payload lines of the color kernel, against its hipEvent time
    words = (int(cidx.num_colors) + 63) // 64
    color_roofline = {}
    for fused, kernel_ms in ((True, color_ms_f), (False, color_ms)):
        s.set_color_output(fused)
        cs.set_input(idx if fused else None)
        s.search_reads_device(d_chars.data_ptr(), wl.n_chars, d_ends.data_ptr(), wl.n_reads,
                              d_res.data_ptr())
        s.sync()
        cs.set_census(True)
        cs.pseudoalign_device(d_res.data_ptr(), s.kmer_ends_device(), wl.n_reads, params,
                              d_off.data_ptr(), d_ids.data_ptr(), cap)
        cen = cs.last_census()
        cs.set_census(False)
        c_bytes = (LINE * (cen["set_lines"] + cen["mark_lines"] + cen["payload_lines"])
                   + 8 * cen["kmers"] + wl.n_reads * (8 + 8 + 8 * words))
        c_gbs = c_bytes / (kernel_ms / 1e3) / 1e9
        color_roofline["color_table" if fused else "key_nodes"] = {
            "bound": "hbm", "achieved": c_gbs, "peak": 8000.0, "unit": "GB/s",
            "frac": c_gbs / 8000.0, "kernel_ms": kernel_ms, "algorithmic_bytes": c_bytes,
            "frac_of_random_line_ceiling": c_gbs / probe["gbs"], "census": cen}
        log(f"color stage census ({'color table' if fused else 'key nodes'}): "
            f"{cen['set_lines'] / wl.n_reads:.2f} set + {cen['mark_lines'] / wl.n_reads:.2f} mark + "
            f"{cen['payload_lines'] / wl.n_reads:.2f} payload lines per read, {c_gbs:.0f} GB/s "
            f"over {kernel_ms:.2f} ms")
    out = {"metric": "reads/s (color)", "mode": "colors",
           "value": wl.n_reads / med_f["total"], "unit": "reads/s",
           "form": "color table (the index stage writes each found k-mer's key rank or mask)",
           "color_stage_reads_per_s": wl.n_reads / med_f["color"],
           "index_stage_kmers_per_s": n_kmers / med_f["index"],
           "color_table": {"kind": "masks" if kind == 1 else "key ranks", "build_ms": table_ms},
           "key_node_form": {"reads_per_s": wl.n_reads / med["total"],
                             "color_stage_reads_per_s": wl.n_reads / med["color"],
                             "index_stage_kmers_per_s": n_kmers / med["index"],
                             "color_kernel_ms": color_ms, "seconds": med},
           "outputs_identical": same,
           "index_roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0,
                              "unit": "GB/s", "frac": achieved / 8000.0,
                              "search_ms": search_ms, "algorithmic_bytes": algo_bytes,
                              "node_line_bytes": node_line_bytes,
                              "random_line_ceiling_gbs": probe["gbs"],
                              "frac_of_random_line_ceiling": achieved / probe["gbs"],
                              "census": stats},
           "key_runs_roofline": {"bound": "hbm", "achieved": kr_gbs, "peak": 8000.0,
                                 "unit": "GB/s", "frac": kr_gbs / 8000.0 if kr_gbs else None,
                                 "kernel_ms": kr_ms, "algorithmic_bytes": kr_bytes},
           "color_kernel_ms": color_ms_f, "color_roofline": color_roofline,
           "ids_per_step": n_ids,
           "config": {"workload": wl.name + " " + wl.description, "sbwt_nodes": int(built.n),
                      "launch": launch, "num_colors": int(cidx.num_colors),
                      "color_sets": int(cidx.num_color_sets), "threshold": 0.7},
           "seconds": med_f}
    print(json.dumps(out), flush=True)
    return 0


def mode_cli(args, configs, S, torch, W):
    """C2 reads as a FASTA and as a four-line FASTQ file through `sbwt_search
    index` (file to file), with device and with host (--host-parse) parsing."""
    import subprocess
    import tempfile
    wl = configs.c2(args.reads or 10_000_000)
    lens = np.diff(np.concatenate([[0], wl.read_ends.astype(np.int64)]))
    if not np.all(lens == lens[0]):
        raise SystemExit("cli mode writes fixed-length reads")
    built, _ = _gpu_built(W, wl.genomes, wl.genome_ends, configs.K)
    tmp = tempfile.mkdtemp(prefix="sbwt_cli_")
    sbwt_path = os.path.join(tmp, "index.sbwt")
    built.write(sbwt_path)
    n, L = wl.n_reads, int(lens[0])
    inputs = {}
    rec = np.empty((n, L + 4), np.uint8)
    rec[:, 0], rec[:, 1], rec[:, 2] = ord(">"), ord("r"), ord("\n")
    rec[:, 3:3 + L] = wl.chars.reshape(n, L)
    rec[:, 3 + L] = ord("\n")
    inputs["fasta"] = (os.path.join(tmp, "reads.fna"), rec.nbytes)
    rec.tofile(inputs["fasta"][0])
    rec = np.empty((n, 2 * L + 7), np.uint8)
    rec[:, 0], rec[:, 1], rec[:, 2] = ord("@"), ord("r"), ord("\n")
    rec[:, 3:3 + L] = wl.chars.reshape(n, L)
    rec[:, 3 + L], rec[:, 4 + L], rec[:, 5 + L] = ord("\n"), ord("+"), ord("\n")
    rec[:, 6 + L:6 + 2 * L] = ord("I")
    rec[:, 6 + 2 * L] = ord("\n")
    inputs["fastq"] = (os.path.join(tmp, "reads.fq"), rec.nbytes)
    rec.tofile(inputs["fastq"][0])
    del rec
    exe = os.path.join(REPO, "sbwt-search_amd", "host", "sbwt_search")
    out = {"metric": "sbwt_search index end to end (sequence file to results file)", "mode": "cli",
           "config": {"workload": wl.name + " " + wl.description,
                      "fasta_bytes": inputs["fasta"][1], "fastq_bytes": inputs["fastq"][1],
                      "tmpdir": tmp}}
    ext = {"ascii": ".txt", "binary": ".bin", "bool": ".bool"}
    # one untimed run first: the first large output file of a fresh box pays for
    # growing the page cache
    warm = os.path.join(tmp, "warm")
    subprocess.run([exe, "index", "-q", inputs["fasta"][0], "-i", sbwt_path, "-o", warm],
                   check=True, capture_output=True)
    os.remove(warm + ".txt")
    runs = [(kind, fmt, parse)
            for kind, fmt in (("fasta", "ascii"), ("fasta", "binary"), ("fastq", "ascii"),
                              ("fastq", "bool"))
            for parse in ("device", "host")]
    if args.cli_runs:
        runs = [tuple(r.split(":")) for r in args.cli_runs.split(",")]
    for kind, fmt, parse in runs:
        path, in_bytes = inputs[kind]
        prefix = os.path.join(tmp, "out_" + fmt)
        env = dict(os.environ, SBWT_SEARCH_TIMING="1")
        extra = ["--host-parse"] if parse == "host" else []
        t0 = time.perf_counter()
        r = subprocess.run([exe, "index", "-q", path, "-i", sbwt_path, "-o", prefix, "-p", fmt]
                           + extra, env=env, capture_output=True, text=True)
        dt = time.perf_counter() - t0
        if r.returncode != 0:
            raise SystemExit(f"sbwt_search failed: {r.stderr}")
        out_file = prefix + ext[fmt]
        size = os.path.getsize(out_file)
        log(f"cli {kind} -> {fmt} ({parse} parsing): {dt:.2f} s, {r.stderr.strip()}")
        out[f"{kind}_{fmt}_{parse}_parse"] = {
            "seconds": dt, "reads_per_s": n / dt, "kmers_per_s": wl.n_kmers() / dt,
            "input_MB_per_s": in_bytes / dt / 1e6, "output_bytes": size,
            "stages": r.stderr.strip()}
        os.remove(out_file)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cli_cpu_baseline(args, wl, exe, sbwt_path, inputs["fasta"][0], L, tmp)
    for path, _ in inputs.values():
        os.remove(path)
    os.remove(sbwt_path)
    os.rmdir(tmp)
    print(json.dumps(out), flush=True)
    return 0


def cli_cpu_baseline(args, wl, exe, sbwt_path, fasta, L, tmp):
    """The oracle's FASTA -> ascii file (C/OpenMP restatement of the reference's
    index command: read, presearch-12 search, classify, print) on the first
    reads of the same FASTA file; the command line runs on the same sample and
    its file must be byte-identical."""
    import subprocess
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_py as O
    n_sample = min(args.cpu_sample_reads, wl.n_reads)
    threads = args.cpu_threads or host_threads()
    sample = os.path.join(tmp, "sample.fna")
    with open(fasta, "rb") as f, open(sample, "wb") as g:
        g.write(f.read(n_sample * (L + 4)))
    orc = O.OracleSbwt(sbwt_path)
    t0 = time.perf_counter()
    orc.index_file(sample, os.path.join(tmp, "cpu.txt"), 0, True, threads=threads)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    subprocess.run([exe, "index", "-q", sample, "-i", sbwt_path, "-o", os.path.join(tmp, "gpu")],
                   check=True, capture_output=True)
    dt_gpu = time.perf_counter() - t0
    same = open(os.path.join(tmp, "cpu.txt"), "rb").read() == open(os.path.join(tmp, "gpu.txt"), "rb").read()
    for name in ("cpu.txt", "gpu.txt", "sample.fna"):
        os.remove(os.path.join(tmp, name))
    if not same:
        raise SystemExit("sbwt_search output differs from the CPU oracle's file on the sample")
    log(f"cli cpu baseline: {n_sample} reads FASTA -> ascii in {dt:.2f} s on {threads} threads "
        f"(sbwt_search on the same sample: {dt_gpu:.2f} s, byte-identical)")
    return {"value": n_sample / dt, "unit": "reads/s", "cores": threads, "kind": "port",
            "sample": f"first {n_sample} reads of the FASTA file, FASTA -> ascii file, "
                      "including the oracle's index load and presearch",
            "gpu_same_sample_seconds": dt_gpu}


def mode_cli_colors(args, configs, S, torch, W):
    """C3 recipe through `sbwt_search colors`: index-result files (written by the
    index stage with the key-k-mer move, ascii and binary) to color files."""
    import subprocess
    import tempfile
    t0 = time.time()
    wl = configs.c3(args.color_reads, args.color_genomes, args.color_roots, args.color_min_len,
                    args.color_max_len)
    built, col = _gpu_built(W, wl.genomes, wl.genome_ends, configs.K,
                            np.arange(len(wl.genome_ends)), 20)
    log(f"workload {wl.name}: {wl.n_reads} reads, SBWT {built.n} nodes ({time.time() - t0:.1f} s)")
    tmp = tempfile.mkdtemp(prefix="sbwt_clicol_")
    colors_path = os.path.join(tmp, "index.tcolors")
    col.write(colors_path)
    idx = S.Index.create([built.bits(c) for c in range(4)], built.n, 31, key_marks=col.key_marks(),
                         suffix_group_starts=built.bits(4))
    dev = torch.device("cuda:0")
    s = idx.searcher(wl.n_chars, wl.n_reads, move_to_key_kmer=True)
    d_chars = torch.from_numpy(wl.chars).to(dev)
    d_ends = torch.from_numpy(wl.read_ends.view(np.int64)).to(dev)
    d_res = torch.empty(wl.n_chars, dtype=torch.int64, device=dev)
    s.search_reads_device(d_chars.data_ptr(), wl.n_chars, d_ends.data_ptr(), wl.n_reads,
                          d_res.data_ptr())
    s.sync()
    lens = np.diff(np.concatenate([[0], wl.read_ends.astype(np.int64)]))
    kends = np.cumsum(np.maximum(lens - configs.K + 1, 0)).astype(np.uint64)
    n_kmers = int(kends[-1])
    res = d_res[:n_kmers].cpu().numpy().view(np.uint64)
    del d_chars, d_ends, d_res
    exe = os.path.join(REPO, "sbwt-search_amd", "host", "sbwt_search")
    out = {"metric": "sbwt_search colors end to end (index-result file to colors file)",
           "mode": "cli-colors",
           "config": {"workload": wl.name + " " + wl.description, "sbwt_nodes": int(built.n),
                      "num_colors": int(col.num_colors), "kmers": n_kmers, "threshold": 0.7}}
    fmt = S.Formatter(n_kmers, wl.n_reads)
    closes = np.ones(wl.n_reads, np.uint8)
    for in_mode, mode_id, ext in (("ascii", S.FORMAT_ASCII, ".txt"),
                                  ("binary", S.FORMAT_BINARY, ".bin")):
        body = fmt.format(mode_id, res, kends, closes)
        in_path = os.path.join(tmp, "idx" + ext)
        with open(in_path, "wb") as f:
            name = in_mode.encode()
            f.write(len(name).to_bytes(8, "little") + name + (4).to_bytes(8, "little") + b"v1.0")
            f.write(body)
        in_bytes = len(body) + 8 + len(in_mode) + 12
        del body
        for out_mode, out_ext in (("ascii", ".txt"), ("binary", ".bin"), ("csv", ".csv")):
            if out_mode != "csv" and out_mode != in_mode:
                continue
            prefix = os.path.join(tmp, "col")
            env = dict(os.environ, SBWT_SEARCH_TIMING="1")
            t0 = time.perf_counter()
            r = subprocess.run([exe, "colors", "-q", in_path, "-k", colors_path, "-o", prefix,
                                "-p", out_mode, "-t", "0.7"], env=env, capture_output=True,
                               text=True)
            dt = time.perf_counter() - t0
            if r.returncode != 0:
                raise SystemExit(f"sbwt_search colors failed: {r.stderr}")
            size = os.path.getsize(prefix + out_ext)
            key = f"{in_mode}_to_{out_mode}"
            log(f"cli colors {key}: {dt:.2f} s, {r.stderr.strip()}")
            out[key] = {"seconds": dt, "reads_per_s": wl.n_reads / dt,
                        "kmers_per_s": n_kmers / dt, "input_bytes": in_bytes,
                        "input_MB_per_s": in_bytes / dt / 1e6, "output_bytes": size,
                        "stages": r.stderr.strip()}
            os.remove(prefix + out_ext)
        os.remove(in_path)
    os.remove(colors_path)
    os.rmdir(tmp)
    print(json.dumps(out), flush=True)
    return 0


def mode_gather_probe(args, S):
    """The random-line probe over working-set sizes (the ceiling of dependent random
    128-byte line reads, cache-resident to HBM-resident)."""
    rows = []
    for mib in (16, 64, 128, 256, 512, 1024, 4096, 16384, 65536, 131072):
        for bpc in (2, 4, 8):
            r = S.random_line_probe(mib << 20, bpc)
            rows.append({"working_set_mib": r["table_bytes"] >> 20, "blocks_per_cu": bpc,
                         "lines_per_s": r["lines_per_s"], "gbs": r["gbs"], "ms": r["ms"]})
            log(f"{mib} MiB x {bpc}/CU: {r['lines_per_s'] / 1e9:.1f} G lines/s, {r['gbs']:.0f} GB/s")
    print(json.dumps({"metric": "random 128-byte line reads/s by working set", "mode":
                      "gather-probe", "rows": rows}), flush=True)
    return 0


def configs_k():
    import configs
    return configs.K


if __name__ == "__main__":
    sys.exit(main())
