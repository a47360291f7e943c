"""bench.py -- XTNeighbor-streaming nearest-neighbour search on one MI355X.

Workload (default): 1,000,000 unique synthetic CDR3 strings (lengths
~N(15, 2) clipped to [8, 22], 'C' first, 'F'/'W' last, 25% planted 1-2 edit
relatives, seed 2), Levenshtein distance 2. One step is one full pass of the
device pipeline over the resident batch: packing, deletion-variant
generation, key sort, run detection, candidate pairs, pair sort/unique,
verification and the compacted sorted (i, j, distance) output in HBM.

The inputs are uploaded once before the timed region (xtn_set_sequences);
the host hand-off (D2H of the output) is not part of a step. Synchronisation
is hipDeviceSynchronize from the HIP runtime the engine itself uses: torch is
not imported, because its bundled HIP runtime has the same soname as the
system ROCm's and would replace the one libxtn.so was built against.

Prints one JSON line. Also reports verified pairs/s, p50/p99 step latency,
the roofline of the dominant kernel (HIP events around every launch on the
engine's stream), a plain copy kernel's GB/s, and a CPU baseline: the
oracle's restatement of the reference algorithm on the host cores (OpenMP)
over the same workload, plus its single-thread form on a bounded prefix.
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "xt-neighbor_amd"))

import numpy as np  # noqa: E402

import xtn  # noqa: E402

METRIC = "seqs/s + verified pairs/s at 1M CDR3 d=2; batched vs plain; p50/p99 latency"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (n, mu, sigma, lmin, lmax, planted, max_edits, distance, measure, seed)
    "nn-1M-cdr3-lev-d2": (1_000_000, 15.0, 2.0, 8, 22, 0.25, 2, 2, 0, 2),
    "nn-10k-cdr3-lev-d1": (10_000, 15.0, 2.0, 8, 22, 0.20, 1, 1, 0, 1),
    "nn-5M-cdr3-lev-d1": (5_000_000, 15.0, 2.5, 6, 30, 0.25, 1, 1, 0, 3),
}


def hip_sync():
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipDeviceSynchronize.restype = ctypes.c_int
    rc = lib.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")


def cpu_baseline(chars, offsets, d, measure, sample, threads):
    """The oracle's restatement of the reference algorithm on the host cores:
    the OpenMP port on `threads` threads over the whole workload (median of
    5 runs), and the single-thread port on a bounded prefix."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    n = len(offsets) - 1
    # the process (and so the port's OpenMP threads) pinned to `threads` of
    # the CPUs it may use, for the duration of the timing
    allowed = sorted(os.sched_getaffinity(0))
    pinned = allowed[:threads]
    os.sched_setaffinity(0, pinned)
    try:
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            tri, st = oracle.nn_port_mt(chars, offsets, d, measure, threads)
            times.append(time.perf_counter() - t0)
        dt = statistics.median(times)
        ns = min(sample, n)
        sub_off = offsets[: ns + 1].copy()
        sub_chars = chars[: int(sub_off[-1])].copy()
        t0 = time.perf_counter()
        _, st1 = oracle.nn_port(sub_chars, sub_off, d, measure)
        dt1 = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, allowed)
    return {"value": n / dt, "unit": "seqs/s", "cores": threads, "kind": "port",
            "sample": f"the whole workload ({n} sequences, d={d}), {threads} OpenMP threads, median of 5 "
                      f"runs {dt:.2f} s, {st['n_output']} output pairs, {st['n_unique']} unique candidates",
            "verified_pairs_per_s": st["n_unique"] / dt,
            "one_thread": {"value": ns / dt1, "unit": "seqs/s", "cores": 1,
                           "sample": f"first {ns} sequences, {dt1:.1f} s, {st1['n_output']} output pairs"},
            "host": host_cpu(),
            "pinned_cpus": f"{len(pinned)} of the {len(allowed)} CPUs allowed to the process "
                           f"(ids {pinned[0]}..{pinned[-1]})"}


def host_cpu() -> str:
    """model name and core count of the host"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs visible"


def batched_vs_plain(eng, offsets, n_sets, seed):
    """the same rows as n_sets query sets of log-uniform size: one pipeline
    pass for all sets (batched) against one pass per set (plain)"""
    n = len(offsets) - 1
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(100), np.log(10000), n_sets))
    cuts = np.concatenate([[0], np.cumsum(sizes)])
    so = np.round(cuts / cuts[-1] * n).astype(np.uint32)
    eng.kernel_profile_reset()
    eng.configure(profile=False)
    out = {}
    for name, batched in (("plain", False), ("batched", True)):
        eng.nn_sets(so, batched=batched, collect=False)  # warm-up
        hip_sync()
        t0 = time.perf_counter()
        eng.nn_sets(so, batched=batched, collect=False)
        hip_sync()
        dt = time.perf_counter() - t0
        out[f"{name}_ms"] = dt * 1e3
        out[f"{name}_seqs_per_s"] = n / dt
        bst = eng.stats()
        out[f"{name}_output_pairs"] = bst["n_output"]
        out[f"{name}_launches"] = bst["n_launches"]
        out[f"{name}_syncs"] = bst["n_syncs"]
        out[f"{name}_graph_replays"] = bst["graph_replays"]
    out["sets"] = n_sets
    out["rows"] = n
    out["speedup"] = out["plain_ms"] / out["batched_ms"]
    return out


def host_boundary(eng, chars, offsets, runs=3):
    """the drop-in boundary as the reference's callers use it: host chars and
    offsets in (xtn_set_sequences), xtn_perform, the sorted triples handed to
    the callback in host memory; both PCIe transfers are inside the time"""
    n = len(offsets) - 1
    times = []
    for _ in range(runs + 1):  # the first run is a warm-up
        hip_sync()
        t0 = time.perf_counter()
        eng.set_sequences(chars=chars, offsets=offsets)
        pairs, _, calls = eng.nn()
        times.append(time.perf_counter() - t0)
    best = min(times[1:])
    return {"ms": best * 1e3, "seqs_per_s": n / best, "output_pairs": len(pairs), "callbacks": calls,
            "runs": runs}


def served(n_index, n_queries, seed, plain_queries=4):
    """Query sets (log-uniform 1k..100k rows, 30% planted relatives of index
    rows) against a resident index of n_index rows, Levenshtein d=1; latency
    is submit to completion. Capacity: every query submitted at once, rows
    per second of wall time until the last completes. Then Poisson arrivals
    at 50, 80 and 100% of that rate (queries/s = capacity / mean rows):
    p50/p99 latency and the rows/s sustained. Beside it, the same query sets
    one at a time through the plain pipeline, which has no resident index:
    one NN run over index + query (device time, input already resident),
    rows/s = query rows / run time."""
    import threading

    ichars, ioffs = xtn.synth_nn(seed, n_index, 15.0, 2.0, 8, 22, 0.25, 1)
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(100000), n_queries)).astype(int)
    queries = [xtn.synth_queries(seed * 1000 + k, ichars, ioffs, int(sz), planted=0.3)
               for k, sz in enumerate(sizes)]
    nq_rows = int(sizes.sum())
    t_build = time.perf_counter()
    srv = xtn.Server(ichars, ioffs, distance=1, slots=4, max_query_rows=1 << 17)
    build_ms = (time.perf_counter() - t_build) * 1e3
    for t in [srv.submit(*queries[k % n_queries]) for k in range(8)]:  # warm-up, every slot
        srv.wait(t)

    def run(gaps):
        tickets = [None] * n_queries
        lat = [0.0] * n_queries
        done_at = [0.0] * n_queries
        ready = threading.Semaphore(0)

        def waiter():
            for k in range(n_queries):
                ready.acquire()
                _, _, l = srv.wait(tickets[k])
                lat[k], done_at[k] = l, time.perf_counter()

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        next_t = t0
        for k in range(n_queries):
            if gaps is not None:
                next_t += gaps[k]
                delay = next_t - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            tickets[k] = srv.submit(*queries[k])
            ready.release()
        th.join()
        wall = max(done_at) - t0
        return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "seqs_per_s": nq_rows / wall, "wall_s": wall}

    # capacity: one untimed burst, then the median of five (a burst is ~10 ms
    # of wall time, so single bursts vary with host scheduling)
    run(None)
    bursts = sorted((run(None) for _ in range(5)), key=lambda r: r["seqs_per_s"])
    burst = bursts[2]
    capacity = burst["seqs_per_s"]
    loads = {}
    for frac in (0.5, 0.8, 1.0):
        rate = capacity * frac / (nq_rows / n_queries)  # queries per second
        gaps = rng.exponential(1.0 / rate, n_queries)
        r = run(gaps)
        r["arrival_rate_hz"] = rate
        loads[f"{int(frac * 100)}pct"] = r
    srv.close()
    # the plain pipeline: index + query as one input, one NN run per query
    plain = []
    with xtn.Engine(distance=1) as eng:
        for k in range(min(plain_queries, n_queries)):
            qc, qo = queries[k]
            chars = np.concatenate([ichars, qc])
            offs = np.concatenate([ioffs, qo[1:].astype(np.int64) + int(ioffs[-1])]).astype(np.uint32)
            eng.set_sequences(chars=chars, offsets=offs)
            eng.nn_resident()  # warm-up
            hip_sync()
            t0 = time.perf_counter()
            eng.nn_resident()
            hip_sync()
            plain.append((len(qo) - 1, time.perf_counter() - t0))
    plain_rows = sum(r for r, _ in plain)
    plain_s = sum(t for _, t in plain)
    return {"index_seqs": n_index, "queries": n_queries, "query_seqs": nq_rows, "slots": 4,
            "index_build_ms": build_ms, "capacity_seqs_per_s": capacity, "burst": burst,
            "bursts_seqs_per_s": [r["seqs_per_s"] for r in bursts], "load": loads,
            "p50_ms": loads["80pct"]["p50_ms"], "p99_ms": loads["80pct"]["p99_ms"],
            "plain_per_query": {"queries": len(plain), "query_seqs": plain_rows,
                                "ms_per_query": plain_s / max(len(plain), 1) * 1e3,
                                "seqs_per_s": plain_rows / plain_s if plain_s > 0 else None,
                                "note": "one NN run over index + query per query set (no resident index)"},
            "capacity_vs_plain": capacity / (plain_rows / plain_s) if plain_s > 0 else None}


def end_to_end(chars, offsets, d, measure):
    """the reference's published measurement: the CLI on a CSV of the same 1M
    sequences, process wall time without -o (02D_symdel_large_scale.ipynb
    cell 23-25) and with -o"""
    import subprocess
    import tempfile

    n = len(offsets) - 1
    tmp = tempfile.mkdtemp(prefix="xtn_bench_")
    path = os.path.join(tmp, "input.csv")
    body = chars.tobytes()
    with open(path, "wb") as f:
        f.write(b"cdr3\n")
        f.write(b"\n".join(body[offsets[i]:offsets[i + 1]] for i in range(n)))
        f.write(b"\n")
    cmd = [xtn.CLI_PATH, "-i", path, "-n", str(n), "-d", str(d), "-m", "hamming" if measure else "leven"]
    out = {}
    env = dict(os.environ, XTN_CLI_TIMING="1")
    for label, extra in (("no_output", []), ("with_output", ["-o", os.path.join(tmp, "out.txt")])):
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            r = subprocess.run(cmd + extra, capture_output=True, env=env)
            wall = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError(f"CLI exit status {r.returncode} ({label}): stderr {r.stderr.decode()[-2000:]!r} "
                                   f"stdout {r.stdout.decode()[-500:]!r}")
            # the CLI's own stage times (ms): parse (the device opens on a
            # second thread meanwhile), the rest of that open, upload, run and
            # hand-off, close
            stages = {}
            for line in r.stderr.decode().splitlines():
                if line.startswith("xtn_cli_times"):
                    for kv in line.split()[1:]:
                        k, v = kv.split("=")
                        stages[k] = float(v)
            runs.append((wall, stages))
        wall, stages = min(runs, key=lambda x: x[0])
        out[f"{label}_s"] = wall
        out[f"{label}_stages_ms"] = stages
    out["seqs_per_s"] = n / out["no_output_s"]
    out["published_reference_s"] = 5.449
    out["speedup_vs_published"] = 5.449 / out["no_output_s"]
    for fn in os.listdir(tmp):
        os.remove(os.path.join(tmp, fn))
    os.rmdir(tmp)
    return out


def overlap_leg(seed: int, n_reps: int, size_min: int, size_max: int, steps: int, e2e: bool,
                tune: bool = True) -> dict:
    """Repertoire overlap on the batched-repertoire workload: n_reps
    repertoires of log-uniform size, Zipf(2) frequencies capped at 1e5, 10%
    of sequences shared or 1-edit-shared across repertoires, Levenshtein d=1.
    A run is xtn_perform in overlap mode over all repertoires at once (host
    SeqInfo in, R x R sums out); with e2e, the CLI on the same rows too."""
    import subprocess
    import tempfile

    chars, offsets, freq, sizes = xtn.synth_overlap(seed, n_reps, size_min, size_max, shared=0.1)
    rep = xtn.reps_from_sizes(sizes)
    n = len(freq)
    out = {"repertoires": n_reps, "rows": n, "sizes": [int(size_min), int(size_max)], "distance": 1,
           "measure": "leven"}
    with xtn.Engine(distance=1) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        out["tuning"] = autotuned(eng, tune)
        res = eng.overlap(freq, rep)
        times = []
        for _ in range(steps):
            t0 = time.perf_counter()
            res = eng.overlap(freq, rep)
            times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    out.update({"ms": t * 1e3, "seqs_per_s": n / t, "output_rows": int(len(res))})
    if e2e:
        tmp = tempfile.mkdtemp(prefix="xtn_bench_ov_")
        seqs, info = os.path.join(tmp, "seqs.csv"), os.path.join(tmp, "info.csv")
        xtn.write_csv(seqs, chars, offsets, freq)
        with open(info, "w") as f:
            f.write("size\n" + "".join(f"{int(x)}\n" for x in sizes))
        cmd = [xtn.CLI_PATH, "-i", seqs, "-n", str(n), "-I", info, "-N", str(n_reps)]
        wall = []
        for _ in range(2):
            t0 = time.perf_counter()
            r = subprocess.run(cmd, capture_output=True)
            wall.append(time.perf_counter() - t0)
            if r.returncode != 0:
                raise RuntimeError(f"CLI exit status {r.returncode} (overlap): stderr {r.stderr.decode()[-2000:]!r} "
                                   f"stdout {r.stdout.decode()[-500:]!r}")
        out["cli_no_output_s"] = min(wall)
        out["cli_seqs_per_s"] = n / min(wall)
        for fn in os.listdir(tmp):
            os.remove(os.path.join(tmp, fn))
        os.rmdir(tmp)
    out["reference_published"] = {"seqs_per_s": 11968694 / 12.589,
                                  "note": "64 Emerson repertoires (11,968,694 seqs), Levenshtein d=1, "
                                          "12.589 s end to end, GPU unrecorded (data/airr_overlap.csv:41)"}
    return out


TUNE_CACHE = None


def autotuned(eng, enabled: bool) -> dict:
    """the engine's autotuner run on its current workload (cache in a fresh
    temporary file, so the choice is measured in this run): chosen parameters,
    their run time and the defaults' run time"""
    global TUNE_CACHE
    if not enabled:
        return {"enabled": False, **eng.tuning()}
    if TUNE_CACHE is None:
        import tempfile
        TUNE_CACHE = os.path.join(tempfile.mkdtemp(prefix="xtn_tune_"), "tuning.json")
    t = eng.autotune(TUNE_CACHE, runs=5)
    return {"enabled": True, "bucket_capacity": t["bucket_capacity"], "bucket_target": t["bucket_target"],
            "sort_threads": t["sort_threads"], "sort_tile": t["sort_tile"], "select_threads": t["select_threads"],
            "sort_low_bits": t["sort_low_bits"], "coarse_separate": t["coarse_separate"], "ms_tuned": t["ms"],
            "ms_default": t["ms_default"], "candidates": t["candidates"]}


def timed_runs(eng, steps):
    """ms per nn_resident run (median of `steps` after one warm-up) and the stats"""
    eng.nn_resident()
    times = []
    for _ in range(steps):
        hip_sync()
        t0 = time.perf_counter()
        eng.nn_resident()
        hip_sync()
        times.append(time.perf_counter() - t0)
    return statistics.median(times) * 1e3, eng.stats()


def chunked_leg(steps: int = 5, tune: bool = True) -> dict:
    """5M synthetic CDR3 (lengths ~N(15,2.5) in [6,30], 25% planted 1-edit
    relatives), d=1, Levenshtein and Hamming: one key-space chunk against the
    key space forced into chunks of at most 20M records (each chunk
    regenerates its part of the records, so the chunked run pays the
    generator once per chunk plus a counting pass)"""
    n = 5_000_000
    chars, offsets = xtn.synth_nn(3, n, 15.0, 2.5, 6, 30, 0.25, 1)
    out = {"n_seqs": n, "distance": 1}
    with xtn.Engine(distance=1) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        for measure in ("leven", "hamming"):
            eng.configure(measure=measure, max_records_per_chunk=0)
            out[f"{measure}_tuning"] = autotuned(eng, tune)
            for label, chunk in (("one_chunk", 0), ("chunks", 20_000_000)):
                eng.configure(measure=measure, max_records_per_chunk=chunk)
                ms, st = timed_runs(eng, steps)
                out[f"{measure}_{label}"] = {"ms": ms, "seqs_per_s": n / (ms / 1e3), "chunks": st["n_chunks"],
                                             "variants": st["n_variants"], "output_pairs": st["n_output"],
                                             "syncs": st["n_syncs"], "launches": st["n_launches"]}
    return out


def distance_leg(steps: int = 5) -> dict:
    """100k synthetic CDR3 at d = 1..5 (the reference publishes 100k at
    d = 1..5, data/gpu_dist_benchmark.csv); d >= 4 takes the generator's
    colex unranking"""
    n = 100_000
    chars, offsets = xtn.synth_nn(7, n, planted=0.25, max_edits=2)
    out = {}
    with xtn.Engine(distance=1) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        for d in range(1, 6):
            eng.configure(distance=d)
            ms, st = timed_runs(eng, steps)
            out[f"d{d}"] = {"ms": ms, "seqs_per_s": n / (ms / 1e3), "variants": st["n_variants"],
                            "unique_pairs": st["n_unique"], "output_pairs": st["n_output"]}
    out["reference_published_s"] = {"d1": 0.436, "d2": 0.836, "d3": 2.676, "d4": 11.03, "d5": 49.10,
                                    "note": "XTNeighbor-streaming (V2) at 100k, end to end incl. process "
                                            "start, data/gpu_dist_benchmark.csv:2-13; not the same "
                                            "measurement"}
    return out


def copy_probe(nbytes: int = 1 << 30) -> dict:
    """a plain copy kernel of nbytes on the same GPU: the practical HBM ceiling"""
    probe = ctypes.CDLL(os.path.join(REPO, "xt-neighbor_amd", "lib", "libxtn_probe.so"))
    probe.xtn_probe_copy.restype = ctypes.c_double
    probe.xtn_probe_copy.argtypes = [ctypes.c_int64, ctypes.c_int32]
    ms = probe.xtn_probe_copy(nbytes, 5)
    return {"bytes": nbytes, "ms": ms, "GBps": 2 * nbytes / (ms / 1e3) / 1e9 if ms > 0 else None}


def progress(msg: str) -> None:
    """one line per finished leg on stderr (stdout carries only the JSON line)"""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def coarse_ab(eng, steps):
    """ms per step of the same workload with the records' first ordering
    pass in the generator and as a pass of its own, back to back"""
    out = {}
    for sep in (0, 1):
        eng.configure(coarse_separate=sep)
        for _ in range(3):
            eng.nn_resident()
        hip_sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.nn_resident()
        hip_sync()
        key = "separate" if sep else "in_generator"
        out[key + "_ms"] = (time.perf_counter() - t0) / steps * 1e3
        out[key + "_took_generator_path"] = bool(eng.stats()["coarse_in_generator"])
    eng.configure(coarse_separate=0)
    return out


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs summarised in profiles/ (calibrated on a copy kernel), or None"""
    path = os.path.join(REPO, "profiles", f"{workload}_hbm_traffic.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return k["traffic_bytes"] if k else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="nn-1M-cdr3-lev-d2", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample", type=int, default=300_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="OpenMP threads of the CPU baseline (one GPU's share of the host: 16)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-sets", type=int, default=1000)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-overlap", action="store_true")
    ap.add_argument("--no-chunked", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default engine parameters")
    ap.add_argument("--no-distances", action="store_true")
    ap.add_argument("--no-e2e", action="store_true")
    ap.add_argument("--no-coarse-ab", action="store_true")
    ap.add_argument("--coarse-separate", action="store_true",
                    help="run the records' first ordering pass as a pass of its own (not in the generator)")
    ap.add_argument("--served-index", type=int, default=10_000_000)
    ap.add_argument("--served-queries", type=int, default=64)
    ap.add_argument("--probe-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain copy kernel of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.probe_copy:
        probe = ctypes.CDLL(os.path.join(REPO, "xt-neighbor_amd", "lib", "libxtn_probe.so"))
        probe.xtn_probe_copy.restype = ctypes.c_double
        probe.xtn_probe_copy.argtypes = [ctypes.c_int64, ctypes.c_int32]
        ms = probe.xtn_probe_copy(args.probe_copy, 5)
        print(json.dumps({"probe_copy_bytes": args.probe_copy, "ms": ms,
                          "GBps": 2 * args.probe_copy / (ms / 1e3) / 1e9 if ms > 0 else None}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build runs the search on a single MI355X "
              "(the multi-GPU sharding is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)

    n, mu, sigma, lmin, lmax, planted, edits, d, measure, seed = WORKLOADS[args.workload]
    chars, offsets = xtn.synth_nn(seed, n, mu, sigma, lmin, lmax, planted, edits)

    eng = xtn.Engine(distance=d, measure=measure, profile=False)
    eng.configure(coarse_separate=1 if args.coarse_separate else 0)
    eng.set_sequences(chars=chars, offsets=offsets)
    tuning = autotuned(eng, not args.no_autotune)
    for _ in range(args.warmup):
        eng.nn_resident()

    # timed region: the plain pipeline, no per-launch events
    lat = []
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        out_len = eng.nn_resident()
        lat.append((time.perf_counter() - ts) * 1e3)
    hip_sync()
    t1 = time.perf_counter()
    st = eng.stats()
    step_s = (t1 - t0) / args.steps

    # the same steps again with a HIP event pair around every launch, on the
    # engine's stream: per-kernel durations for the roofline
    eng.configure(profile=True)
    eng.nn_resident()
    eng.kernel_profile_reset()
    hip_sync()
    tp0 = time.perf_counter()
    for _ in range(args.steps):
        eng.nn_resident()
    hip_sync()
    profiled_ms = (time.perf_counter() - tp0) / args.steps * 1e3
    progress(f"{args.workload}: {step_s * 1e3:.3f} ms per step")
    kernels = ["gen", "gen_coarse", "sort_rec_hist", "sort_rec_scatter", "region", "bucket", "runs", "pairs_small",
               "pairs_big",
               "sort_seg_hist", "sort_seg_scatter", "tail_hist", "sort_pairs", "select", "pack", "scan",
               "gen_count"]
    prof = {k: eng.kernel_profile(k) for k in kernels}
    # the library prices the record-path launches at the V records of the
    # variant count; the generator's coarse pass writes only n_records of them
    # (repeated variants of a sequence left out), and the passes after it
    # move those: their algorithmic bytes per launch, from the run's counts
    rs = eng.stats()
    per_launch = {"gen_coarse": 8.0 * rs["n_records"] + 24.0 * rs["n_sequences"],
                  "region": 16.0 * rs["n_records"], "bucket": 8.0 * rs["n_records"]}
    for k, b in per_launch.items():
        if prof[k][0] and rs["n_chunks"] == 1:
            prof[k] = (prof[k][0], prof[k][1], b * prof[k][0])
    dom = max(prof, key=lambda k: prof[k][1])
    launches, ms, byts = prof[dom]
    avg_s = ms / 1e3 / max(launches, 1)
    achieved = (byts / max(launches, 1)) / avg_s / 1e9 if avg_s > 0 else 0.0
    per_kernel = {k: {"launches": v[0], "ms_avg": (v[1] / v[0] if v[0] else 0.0),
                      "GBps": ((v[2] / v[0]) / (v[1] / v[0] / 1e3) / 1e9 if v[0] and v[1] else 0.0)}
                  for k, v in prof.items() if v[0]}

    result = {
        "metric": METRIC,
        "value": n / step_s,
        "unit": "seqs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": args.workload, "n_seqs": n, "distance": d,
                   "measure": "hamming" if measure else "leven", "planted": planted,
                   "max_edits": edits, "seed": seed},
        "verified_pairs_per_s": st["n_unique"] / step_s,
        "launches_per_step": st["n_launches"],
        "syncs_per_step": st["n_syncs"],
        "graph_replay": bool(st["graph_replays"]),
        "tail_sort_passes": st["tail_passes"],
        "coarse_in_generator": bool(st["coarse_in_generator"]),
        "output_pairs": out_len,
        "variants": st["n_variants"],
        "records": st["n_records"],
        "candidate_pairs": st["n_candidates"],
        "unique_pairs": st["n_unique"],
        "ms_per_step_profiled": profiled_ms,
        "p50_ms": statistics.median(lat),
        "p99_ms": float(np.percentile(lat, 99)),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.workload, dom)},
        "kernels": per_kernel,
        "tuning": tuning,
        "reference_published": {"seqs_per_s": 183500, "note": "1M Emerson CDR3, d=2, 5.449 s end to "
                                "end incl. process start and parsing, GPU unrecorded "
                                "(data/gpu_benchmark.csv:42); not the same measurement"},
    }
    eng.configure(profile=False)

    def leg(name, done, fn):
        """a secondary leg: its result under `name`, or the error it raised
        (recorded in the line, so the main figures are still printed)"""
        try:
            result[name] = fn()
            progress(done)
        except Exception as e:  # noqa: BLE001 - reported, not hidden
            result[name] = {"error": f"{type(e).__name__}: {e}"}
            progress(f"{name} FAILED: {type(e).__name__}: {e}")

    if not args.no_coarse_ab:
        leg("coarse_pass", "coarse pass A/B done", lambda: coarse_ab(eng, args.steps))
    leg("host_boundary", "host boundary done", lambda: host_boundary(eng, chars, offsets))
    if not args.no_batched:
        leg("batched", "batched vs plain done", lambda: batched_vs_plain(eng, offsets, args.batched_sets, seed))
    eng.close()
    if not args.no_e2e:
        leg("end_to_end_cli", "CLI end to end done", lambda: end_to_end(chars, offsets, d, measure))
    if not args.no_served:
        leg("served", "served done", lambda: served(args.served_index, args.served_queries, 5))
    if not args.no_chunked:
        leg("chunked_5M_d1", "chunked 5M done", lambda: chunked_leg(tune=not args.no_autotune))
    if not args.no_distances:
        leg("distances_100k", "distance scaling done", distance_leg)
    if not args.no_overlap:
        leg("overlap", "overlap done",
            lambda: overlap_leg(4, 64, 10_000, 500_000, 3, not args.no_e2e, not args.no_autotune))
    leg("copy_probe", "copy probe done", copy_probe)
    if not args.no_cpu_baseline:
        leg("cpu_baseline", "CPU baseline done",
            lambda: cpu_baseline(chars, offsets, d, measure, args.cpu_sample, args.cpu_threads))
    print(json.dumps(result))


if __name__ == "__main__":
    main()
