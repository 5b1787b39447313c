"""Measurement modes beyond the headline index search (used by bench.py --mode ...).

batched: the uneven-length workload (C4) through the plain form -- one launch of the
  streaming search and of the color stage per batch, batch after batch on one stream --
  and the batched form -- the same kernels with batches alternating over two streams so
  one batch's tail overlaps the next one's start. Outputs are compared in full.
colors: the color stage alone on a colored workload (C3 recipe), reads/s.
serve: a Poisson stream of requests (N log-uniform in [1, max] reads of U[75, 250] bp)
  replayed by the library's C++ load generator through the serving loop and through
  plain sequential processing; p50/p99 latency from the scheduled arrival, reads/s.
"""
import time

import numpy as np


def _split_batches(read_ends, max_chars):
    """[(read0, read1)] of whole reads with at most max_chars chars per batch."""
    out = []
    n = len(read_ends)
    r0 = 0
    ends = read_ends.astype(np.int64)
    while r0 < n:
        c0 = int(ends[r0 - 1]) if r0 else 0
        r1 = int(np.searchsorted(ends, c0 + max_chars, side="right"))
        r1 = max(r1, r0 + 1)
        out.append((r0, r1))
        r0 = r1
    return out


def run_batched(S, torch, wl, idx, col, batch_chars, tau=0.7, log=print, collect=False,
                item_kmers=1024):
    """C4: the same streaming search, key-k-mer chains and color decode in three forms,
    on the same batches.

    plain: one launch of each stage per batch on one stream, batch after batch,
      each batch's color ids copied back before the next starts; the index stage
      writes key nodes and the color stage runs one wave per read, ranking each
      key and reading its set (the reference's two-stage, one-launch-per-batch
      form, with this library's kernels).
    batched_unfused: the color stage over a device queue of (read, k-mer range)
      work items of item_kmers k-mers, the batches spread over two streams so one
      batch's tail overlaps the next one's start.
    batched: as batched_unfused, with index and color search fused through the
      index's color table (idx.build_color_table(col) must have run): the search
      leaves its key chains (COLOR_OUTPUT_CHAINS) and the color stage resolves
      them to each k-mer's color entry (a mask, or a key rank) while it counts,
      masks by ballots (ranks: one key-set lookup per run): no key pass.

    Outputs (per batch: CSR offsets and color ids) are compared in full with the
    plain form's. With collect, the outputs are returned too."""
    dev = torch.device("cuda:0")
    d_chars = torch.from_numpy(wl.chars).to(dev)
    ends64 = wl.read_ends.astype(np.int64)
    batches = _split_batches(wl.read_ends, batch_chars)
    # per-batch relative read ends, resident on the device like the reads
    rel = []
    for r0, r1 in batches:
        c0 = int(ends64[r0 - 1]) if r0 else 0
        rel.append((c0, int(ends64[r1 - 1]) - c0, torch.from_numpy(ends64[r0:r1] - c0).to(dev)))
    max_chars = max(n for _, n, _ in rel)
    max_reads = max(r1 - r0 for r0, r1 in batches)
    nc = col.num_colors
    params = S.color_params(tau)

    def make_lane(items, fused):
        stream = torch.cuda.Stream()
        s = idx.searcher(max_chars, max_reads, move_to_key_kmer=True, stream=stream.cuda_stream)
        s.set_algorithm(S.ALGO_STREAMING)
        cs = col.searcher(max_chars, max_reads, stream=stream.cuda_stream)
        cs.set_work_items(items)
        if fused:
            s.set_color_output(S.COLOR_OUTPUT_CHAINS)
            cs.set_input(idx, chains=True)
        res = torch.empty(max_chars, dtype=torch.int64, device=dev)
        off = torch.empty(max_reads + 1, dtype=torch.int64, device=dev)
        cap = max_reads * nc
        ids = torch.empty(cap, dtype=torch.int32, device=dev)
        # page-locked landing buffers for the copies back (a pageable copy is
        # staged through the runtime's bounce buffer in pieces)
        h_off = torch.empty(max_reads + 1, dtype=torch.int64, pin_memory=True)
        h_ids = torch.empty(cap, dtype=torch.int32, pin_memory=True)
        return {"stream": stream, "s": s, "cs": cs, "res": res, "off": off, "ids": ids,
                "cap": cap, "h_off": h_off, "h_ids": h_ids}

    def run_batch(lane, b):
        c0, n, d_ends = rel[b]
        r0, r1 = batches[b]
        lane["s"].search_reads_device(d_chars.data_ptr() + c0, n, d_ends.data_ptr(), r1 - r0,
                                      lane["res"].data_ptr())
        n_ids = lane["cs"].pseudoalign_device(lane["res"].data_ptr(),
                                              lane["s"].kmer_ends_device(), r1 - r0, params,
                                              lane["off"].data_ptr(), lane["ids"].data_ptr(),
                                              lane["cap"])
        # the batch's output back on the host (the same copies in every form),
        # then out of the lane's landing buffers before its next batch
        nr = r1 - r0
        with torch.cuda.stream(lane["stream"]):
            lane["h_off"][:nr + 1].copy_(lane["off"][:nr + 1], non_blocking=True)
            lane["h_ids"][:n_ids].copy_(lane["ids"][:n_ids], non_blocking=True)
        lane["stream"].synchronize()
        return (lane["h_off"][:nr + 1].numpy().view(np.uint64).copy(),
                lane["h_ids"][:n_ids].numpy().view(np.uint32).copy())

    import threading

    def run_form(n_lanes, items, fused):
        lanes = [make_lane(items, fused) for _ in range(n_lanes)]
        out = [None] * len(batches)

        def worker(li):
            for b in range(li, len(batches), n_lanes):
                out[b] = run_batch(lanes[li], b)

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if n_lanes == 1:
            worker(0)
        else:
            th = [threading.Thread(target=worker, args=(i,)) for i in range(n_lanes)]
            for t in th:
                t.start()
            for t in th:
                t.join()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out

    t_plain, out_plain = run_form(1, 0, False)
    log(f"plain: {len(batches)} batches in {t_plain:.3f} s")
    t_unfused, out_unfused = run_form(2, item_kmers, False)
    log(f"batched without fusion: {len(batches)} batches in {t_unfused:.3f} s")
    t_batched, out_b = run_form(2, item_kmers, True)
    log(f"batched: {len(batches)} batches in {t_batched:.3f} s")
    for name, other in (("batched", out_b), ("batched without fusion", out_unfused)):
        for b in range(len(batches)):
            if not (np.array_equal(out_plain[b][0], other[b][0])
                    and np.array_equal(out_plain[b][1], other[b][1])):
                raise SystemExit(f"{name} and plain color outputs differ in batch {b}")
    n_ids = sum(len(o[1]) for o in out_plain)
    r = {"reads": wl.n_reads, "batches": len(batches), "batch_chars": batch_chars,
         "algorithm": "streaming search with key-k-mer chains (every form)",
         "color_ids": n_ids, "outputs_equal": True,
         "forms": {"plain": "one stream; key nodes; color stage one wave per read",
                   "batched_unfused": f"two streams; key nodes; {item_kmers}-k-mer work items",
                   "batched": f"two streams; key chains resolved to color-table entries by the "
                              f"color stage (fused); {item_kmers}-k-mer work items"},
         "plain_s": t_plain, "batched_unfused_s": t_unfused, "batched_s": t_batched,
         "plain_reads_per_s": wl.n_reads / t_plain,
         "batched_unfused_reads_per_s": wl.n_reads / t_unfused,
         "batched_reads_per_s": wl.n_reads / t_batched,
         "speedup": t_plain / t_batched, "speedup_unfused": t_plain / t_unfused}
    if collect:
        r["outputs"] = {"batches": batches, "plain": out_plain, "batched": out_b,
                        "batched_unfused": out_unfused}
    return r


def run_serve(S, idx, col, wl_reads, rate_per_s, n_requests, max_reads, seed, params,
              slots=4, log=print, kmer_results=True):
    """wl_reads(seed, n_reads) -> (chars uint8 array, ends). Poisson arrivals at
    rate_per_s, N log-uniform in [1, max_reads] reads per request. The same request
    stream is replayed by the library's load generator (sbwt_server_replay: a C++
    submitter on the schedule and a C++ collector, no Python in the loop) through
      served: the serving loop with `slots` slots and HIP graphs,
      served_no_graphs: the same without graphs,
      plain: one slot with eager launches, i.e. one request at a time in arrival
        order, each one search + color stage (the plain sequential form).
    With kmer_results off (and a color index) requests return their color lists
    only. Latency runs from each request's scheduled arrival to its end on the server's
    clock, so time spent behind a late submitter counts. Every request's outputs
    (k-mer results and color lists) are compared across the three forms by digest."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(0, np.log(max_reads), n_requests)).astype(np.int64)
    sizes = np.maximum(sizes, 1)
    arrival = np.cumsum(rng.exponential(1.0 / rate_per_s, n_requests))
    arrival -= arrival[0]
    parts, ends, first = [], [], [0]
    total_chars = 0
    for i in range(n_requests):
        c, e = wl_reads(seed * 100_000 + i, int(sizes[i]))
        parts.append(np.frombuffer(c, np.uint8) if isinstance(c, bytes) else c)
        ends.append(e.astype(np.uint64) + np.uint64(total_chars))
        total_chars += len(parts[-1])
        first.append(first[-1] + len(e))
    chars = np.concatenate(parts)
    read_ends = np.concatenate(ends)
    first = np.array(first, np.uint64)
    total_reads = int(sizes.sum())
    req_chars = np.diff(np.concatenate([[0], read_ends[first[1:] - 1].astype(np.int64)]))
    slot_chars = max(int(req_chars.max()), 1 << 20)
    out = {}
    digests = {}

    def replay(name, n_slots, graphs):
        srv = S.Server(idx, col, slots=n_slots, max_chars_per_slot=slot_chars,
                       max_reads_per_slot=max(max_reads, 1), move_to_key_kmer=col is not None,
                       params=params, graphs=graphs, kmer_results=kmer_results)
        # one untimed pass so that first-touch costs (graph capture, page faults)
        # stay out of the measured replay
        srv.replay(chars, read_ends, first, np.zeros(n_requests))
        r = srv.replay(chars, read_ends, first, arrival)
        lat, lag = r["latency_s"], r["lag_s"]
        digests[name] = r["digest"]
        res = {"slots": srv.slots, "peak_in_flight": srv.peak_in_flight(),
               "p50_ms": 1000 * float(np.percentile(lat, 50)),
               "p99_ms": 1000 * float(np.percentile(lat, 99)),
               "max_ms": 1000 * float(lat.max()),
               "submit_lag_p99_ms": 1000 * float(np.percentile(lag, 99)),
               "submit_lag_max_ms": 1000 * float(lag.max()),
               "reads_per_s": total_reads / r["wall_s"], "wall_s": r["wall_s"],
               "schedule_s": float(arrival[-1])}
        srv.close()
        log(f"{name}: {res}")
        return res

    out["served"] = replay("served", slots, True)
    out["served_no_graphs"] = replay("served_no_graphs", slots, False)
    out["plain"] = replay("plain", 1, False)
    for name in ("served", "served_no_graphs"):
        if not np.array_equal(digests[name], digests["plain"]):
            raise SystemExit(f"{name} outputs differ from the plain form's")
    out["outputs_equal"] = True
    out["requests"] = n_requests
    out["reads"] = total_reads
    out["rate_per_s"] = rate_per_s
    out["latency_from"] = "scheduled arrival"
    out["returns"] = "k-mer results and color lists" if kmer_results or col is None else \
        "color lists"
    return out
