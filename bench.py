"""Benchmark of the exact WFA hot path on one MI355X.

Workload (one "step"): the short-read batch -- 1,000,000 gap-affine
(match 0, mismatch 4, open 6, extend 2) end-to-end pairs, text length uniform
in [100,300] bp, error rate uniform in [0,10%] per pair (substitutions and
1-base indels), score + CIGAR, synthetic seeded data.  The pairs are packed
and resident in HBM before timing; one step aligns all of them (forward
kernel + backtrace kernel, CIGARs left in HBM).

Prints one JSON line (rank 0) with pairs/s as `value`, equivalent GCUPS,
the roofline of the dominant (forward) kernel, the plain one-pair-per-launch
GPU form against the batched form on the same pairs, the 1 kbp / 5% case, the
serving loop's p50/p99 latency at a fixed Poisson arrival rate, and the
reference's OpenMP CPU build timed on this host.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wfa2-lib_amd"))

HBM_PEAK_GBS = 8000.0                 # MI355X HBM3E
VALU_INT32_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12   # 256 CUs x 64 lanes x 2.4 GHz
BYTES_PER_DSTEP = 44                  # gap-affine: 4 int32 in, 3 int32 out, 2 x 8 B sequence window
OPS_PER_DSTEP = 20
# per metric: (algorithmic bytes, integer ops) per diagonal-step
DSTEP_COST = {"affine": (44, 20), "affine2p": (64, 30)}


def host_cpu():
    """The host's CPU: model name, logical CPUs, physical cores (distinct
    (package, core) ids), the CPUs this process may run on, the cgroup CPU
    quota, and the threads the CPU baselines use: the job's CPU share (the
    affinity set, the quota and OMP_NUM_THREADS, whichever is smallest)."""
    model, cores = None, set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    logical = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = logical
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    share = affinity
    if quota:
        share = min(share, max(1, int(quota)))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    return {"model": model, "logical_cpus": logical, "physical_cores": len(cores) or None,
            "affinity_cpus": affinity, "cgroup_cpu_quota": quota, "threads_used": share}


_HOST = None


def cpu_threads():
    global _HOST
    if _HOST is None:
        _HOST = host_cpu()
    return _HOST["threads_used"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per forward launch measured with rocprofv3 PMC counters."""
    path = os.path.join(ROOT, "profiles", "forward_short_reads_counters.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(ps, threads, res=None):
    """The reference (WFA2-lib compiled from source, OpenMP) on this host.
    Also compares the GPU results of a seeded sample with its outputs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle as O
    except Exception as exc:  # pragma: no cover
        log("cpu baseline unavailable: %s" % exc)
        return None
    if not O.ref_available():
        log("cpu baseline unavailable: oracle/_ref not built")
        return None
    pairs = O.Pairs(ps.patterns(), ps.texts())
    conf = O.Conf("affine", (0, 4, 6, 2), "high")
    secs, _ = O.ref_time(conf, pairs, threads)
    checked = bad = 0
    if res is not None:
        import numpy as np
        idx = np.sort(np.random.default_rng(0).choice(ps.n, size=min(2000, ps.n), replace=False))
        sub = O.Pairs([ps.pattern(i) for i in idx], [ps.text(i) for i in idx])
        ref = O.ref_align(conf, sub, threads)
        for j, i in enumerate(idx):
            if res.status[i] != ref.status[j] or res.score[i] != ref.score[j] or res.ops(i) != ref.ops(j):
                bad += 1
        checked = len(idx)
        log("parity spot check against the reference build: %d pairs, %d differ" % (checked, bad))
    return {"value": pairs.n / secs, "unit": "pairs/s", "cores": threads, "kind": "reference",
            "parity_spot_check": {"pairs": int(checked), "differ": int(bad)},
            "sample": "all %d pairs of the workload, wavefront_align of WFA2-lib v2.3 built -O3 "
                      "(%s) with one aligner per OpenMP thread, memory-high, score+CIGAR; "
                      "%.2f s wall" % (pairs.n, os.path.basename(O.ref_lib_path()), secs),
            "gcups": ps.cells / secs / 1e9, "host": _HOST}


# Configurations beyond the headline one (indel / edit / gap-linear, ends-free,
# extension, the wf-adaptive heuristic): (name, workload, metric, penalties,
# aligner keywords, reference-build ISA for the spot check).  The ends-free
# forms are checked against the reference's scalar build, whose diagonal
# order the device follows.
VARIANTS = [
    ("edit end-to-end", "short", "edit", (0, 1, 1), {}, None),
    ("indel end-to-end", "short", "indel", (0, 1, 1), {}, None),
    ("gap-linear (0,4,2) end-to-end", "short", "linear", (0, 4, 2), {}, None),
    ("gap-affine (0,4,6,2) wf-adaptive (10,50,1), the default attributes", "short", "affine", (0, 4, 6, 2),
     {"heuristic": (4, 10, 50, 1)}, None),
    ("gap-affine (0,4,6,2) ends-free, text ends free up to 30", "semiglobal", "affine", (0, 4, 6, 2),
     {"ends_free": (0, 0, 30, 30)}, "v2"),
    ("gap-affine (0,4,6,2) extension", "semiglobal", "affine", (0, 4, 6, 2), {"extension": True}, "v2"),
    ("gap-affine-2p (0,4,6,2,24,1) end-to-end", "short", "affine2p", (0, 4, 6, 2, 24, 1), {}, None),
]


def variant_measurements(n, cpu_sample, threads, no_cpu):
    """pairs/s of each configuration of VARIANTS over n resident pairs (short
    reads, or short reads with random flanks of up to 30 bases for the
    ends-free forms), beside the reference build on a sample of the same pairs
    (one aligner per thread), with a parity spot check of that sample."""
    import numpy as np
    import wfa_amd
    from wfa_amd import datagen
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    O = None
    if not no_cpu:
        try:
            import oracle as O
        except Exception as exc:  # pragma: no cover
            log("reference build unavailable: %s" % exc)
    sets = {"short": datagen.generate(1, 7, n), "semiglobal": datagen.semiglobal(1, 7, n)}
    out = []
    for name, wl, metric, pen, kw, isa in VARIANTS:
        ps = sets[wl]
        al = wfa_amd.Aligner(metric=metric, penalties=pen, **kw)
        rate, dt, st = batch_rate(al, ps)
        row = {"config": name, "pairs": ps.n, "value": rate, "unit": "pairs/s",
               "gcups": ps.cells / dt / 1e9, "diagonal_steps": int(st.diagonal_steps)}
        if O is not None and O.ref_available(isa):
            k = min(cpu_sample, ps.n)
            sub = ps.subset(list(range(k)))
            pairs = O.Pairs(sub.patterns(), sub.texts())
            conf = O.Conf(metric, pen, "high", **kw)
            ref = O.ref_align(conf, pairs, threads, isa=isa)
            got = al.align_batch(sub)
            differ = sum(1 for i in range(k) if got.status[i] != ref.status[i] or got.score[i] != ref.score[i]
                         or got.ops(i) != ref.ops(i))
            row["cpu_reference"] = {"value": k / ref.seconds, "unit": "pairs/s", "cores": threads,
                                    "sample": "first %d pairs, %s" % (k, os.path.basename(O.ref_lib_path(isa)))}
            row["parity_spot_check"] = {"pairs": k, "differ": differ}
            if differ:
                raise RuntimeError("%s: GPU results differ from the reference build" % name)
        log("%s: %.3e pairs/s%s" % (name, rate, (" (reference %.3e pairs/s on %d threads)"
                                                  % (row["cpu_reference"]["value"], threads))
                                    if "cpu_reference" in row else ""))
        out.append(row)
    return out


def seq_file_measurement(n):
    """End-to-end .seq -> .alg throughput of wfa_align_seq_file (parse, device
    batches, format and write, overlapped) on n short-read pairs written to a
    local temporary file first (outside the timing)."""
    import shutil
    import tempfile
    import wfa_amd
    from wfa_amd import datagen
    ps = datagen.generate(1, 9, n)
    tmp = tempfile.mkdtemp(prefix="wfa_seq_")
    try:
        src, dst = os.path.join(tmp, "in.seq"), os.path.join(tmp, "out.alg")
        with open(src, "wb") as f:
            for i in range(ps.n):
                f.write(b">" + ps.pattern(i) + b"\n<" + ps.text(i) + b"\n")
        al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
        # The first file of a new aligner also pays its set-up (stream,
        # device and pinned buffers, the launch-shape tuner's trials); the
        # second, the same file again, is the rate of a process that streams
        # file after file.  Both are reported; value is the second.
        runs = []
        for _ in range(2):
            rc, st = al.align_seq_file(src, dst)
            if rc != 0 or st.pairs != ps.n:
                raise RuntimeError("seq file run failed (%d, %d pairs)" % (rc, st.pairs))
            runs.append(st)
        cold, st = runs
        out = {"pairs": int(st.pairs), "value": st.pairs / st.seconds, "unit": "pairs/s",
               "input_mb_s": st.bytes_in / st.seconds / 1e6, "seconds": st.seconds,
               "read_s": st.read_seconds, "align_s": st.align_seconds, "write_s": st.write_seconds,
               "first_file": {"value": cold.pairs / cold.seconds, "seconds": cold.seconds,
                              "read_s": cold.read_seconds, "align_s": cold.align_seconds,
                              "write_s": cold.write_seconds},
               "sample": "%d short-read pairs (gap-affine 0,4,6,2, score+CIGAR) from a local .seq file "
                         "of %.0f MB to a .alg file of %.0f MB, the second file of the aligner (first_file: "
                         "the first, with the aligner's set-up)" % (ps.n, st.bytes_in / 1e6, st.bytes_out / 1e6)}
        log(".seq -> .alg: %.3e pairs/s (%.0f MB/s in; read %.2f s, align %.2f s, write %.2f s, total %.2f s; "
            "first file %.2f s)" % (out["value"], out["input_mb_s"], st.read_seconds, st.align_seconds,
                                    st.write_seconds, st.seconds, cold.seconds))
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def serve_measurement(duration_s, seed, log, lanes=(1, 1, 1)):
    """Serving: single-pair requests from the mixture of 60% 1 kbp / 5% pairs,
    35% 1-20 kbp / 1-15% pairs and 5% 100 kbp / 15% pairs (BiWFA), gap-affine
    (0,4,6,2), score + CIGAR.  The capacity of the request loop is measured
    closed loop (a burst of requests at once); then requests arrive as a
    Poisson process at 70% of that rate for duration_s, and the latency of
    each (submit to result ready) is recorded."""
    import numpy as np
    import wfa_amd
    from wfa_amd import datagen

    cfg = dict(short_max_length=2048, long_min_length=32768, short_lanes=lanes[0], medium_lanes=lanes[1],
               long_lanes=lanes[2], max_batch_pairs=4096)

    srv = wfa_amd.Server(**cfg)

    def run(ps, arrivals):
        t0 = srv.clock()
        base = srv.submitted if hasattr(srv, "submitted") else 0
        srv.replay(ps, arrivals)
        got = {}
        while len(got) < ps.n:
            for r in srv.collect(max_results=4096, timeout_ms=500):
                got[r.id - base] = r
            if srv.clock() - t0 > 600:
                raise RuntimeError("serving run did not finish")
        srv.submitted = base + ps.n
        res = [got[i] for i in range(ps.n)]
        if any(r.status != 0 for r in res):
            raise RuntimeError("a served request failed")
        return res, t0

    # The burst also warms the lanes (device buffers, record arenas).
    burst = datagen.generate(4, seed, 1200)
    res, t0 = run(burst, np.zeros(burst.n))
    makespan = max(r.finish for r in res) - min(r.submit for r in res)
    capacity = burst.n / makespan
    rate = 0.7 * capacity
    n = max(200, int(rate * duration_s))
    log("serving: capacity %.1f requests/s (burst of %d), open loop at %.1f requests/s for %d requests"
        % (capacity, burst.n, rate, n))
    ps = datagen.generate(4, seed + 1, n)
    arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, n))
    res, t0 = run(ps, arrivals)
    srv.close()
    lat = np.array([r.latency for r in res]) * 1e3
    L = np.maximum(ps.plen, ps.tlen)
    cls = np.where(L >= cfg["long_min_length"], "long", np.where(L <= cfg["short_max_length"], "short", "medium"))
    span = max(r.finish for r in res) - min(r.submit for r in res)
    queue_ms = np.array([r.start - r.submit for r in res]) * 1e3   # waiting for a free lane
    run_ms = np.array([r.finish - r.start for r in res]) * 1e3      # its device batch
    by = {}
    for c in ("short", "medium", "long"):
        m = cls == c
        if m.any():
            slow = m & (lat >= np.percentile(lat[m], 99))
            by[c] = {"requests": int(m.sum()), "p50_ms": float(np.percentile(lat[m], 50)),
                     "p99_ms": float(np.percentile(lat[m], 99)),
                     "p99_tail_queue_ms": float(queue_ms[slow].mean()),
                     "p99_tail_batch_ms": float(run_ms[slow].mean())}
    return {
        "workload": "Poisson arrivals of single-pair requests: 60% 1 kbp/5% (1-base indels), 35% "
                    "1-20 kbp log-uniform/1-15% (geometric indels, mean 3), 5% 100 kbp/15% "
                    "(geometric indels, mean 2, BiWFA); gap-affine (0,4,6,2), score+CIGAR",
        "requests": int(n), "rate_req_s": rate, "capacity_req_s": capacity,
        "throughput_req_s": n / span,
        "p50_ms": float(np.percentile(lat, 50)), "p90_ms": float(np.percentile(lat, 90)),
        "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
        "by_class": by,
        "lanes": {"short": cfg["short_lanes"], "medium": cfg["medium_lanes"], "long": cfg["long_lanes"]},
    }


def batch_rate(al, ps, steps=3):
    """pairs/s of wfa_device_batch_run over a resident batch (after one warm-up)."""
    db = al.upload(ps)
    if db.run() != 0:
        raise RuntimeError("batch run failed")
    t = time.perf_counter()
    for _ in range(steps):
        if db.run() != 0:
            raise RuntimeError("batch run failed")
    dt = (time.perf_counter() - t) / steps
    st = db.stats()
    db.free()
    return ps.n / dt, dt, st


LONG_WORKLOADS = {
    # name: (datagen shape, metric, penalties, GPU memory mode, default pairs, CPU sample pairs,
    #        CPU memory mode, description)
    # The reference's memory mode med gives the CIGARs of memory high (tests/wfa.utest.sh:67-78) with
    # O(s) memory per thread, so its CPU sample of the 1-20 kbp pairs runs in med on every thread.
    "affine2p-1-20kbp": (2, "affine2p", (0, 4, 6, 2, 24, 1), "high", 4000, 1000, "med",
                         "dual-cost gap-affine (0,4,6,2,24,1) end-to-end pairs, L log-uniform in [1,20] kbp, "
                         "error ~U[1%,15%], geometric indels (mean 3), score+CIGAR, memory-high"),
    "biwfa-100kbp": (3, "affine", (0, 4, 6, 2), "ultralow", 256, 32, "ultralow",
                     "gap-affine (0,4,6,2) end-to-end pairs of 100 kbp, 15% errors, geometric indels "
                     "(mean 2), score+CIGAR, BiWFA (memory ultralow)"),
}


def long_workload(name, n, no_cpu):
    """One run over the first n pairs of a long-pair workload (a run takes
    seconds, so one untimed warm-up on a small slice only).  Reports pairs/s,
    equivalent GCUPS, the HBM and VALU fractions of the forward launches (the
    contract's bytes and ops per diagonal-step over their HIP-event time),
    and the reference on this host's CPU share over the first pairs of the
    same set, whose results must equal the GPU's."""
    import numpy as np
    import wfa_amd
    from wfa_amd import datagen
    shape, metric, pen, mem, n_def, n_cpu, cpu_mem, desc = LONG_WORKLOADS[name]
    n = n or n_def
    ps = datagen.generate(shape, 7, n)
    al = wfa_amd.Aligner(metric=metric, penalties=pen, memory=mem)
    al.align_batch(ps.subset(list(range(min(4, n)))))
    if mem != "ultralow" and os.environ.get("WFA_AUTOTUNE", "1") != "0":
        # the launch-shape tuner measures its candidates on pairs of the same
        # workload (another seed), outside the timed run
        t = time.perf_counter()
        al.align_batch(datagen.generate(shape, 8, min(n, 800)))
        log("%s: tuning run on %d pairs of seed 8 in %.2f s" % (name, min(n, 800), time.perf_counter() - t))
    db = al.upload(ps)
    t = time.perf_counter()
    rc = db.run()
    dt = time.perf_counter() - t
    st = db.stats()
    if rc != 0:
        raise RuntimeError("run returned %d" % rc)
    res = db.download()
    db.free()
    bytes_ds, ops_ds = DSTEP_COST[metric]
    W = int(st.diagonal_steps)
    fwd_s = st.forward_ms * 1e-3
    log("%s: %d pairs in %.2f s (%.1f pairs/s, %.3e diagonal-steps/s; forward %.2f s, backtrace %.2f s)"
        % (name, n, dt, n / dt, W / dt, fwd_s, st.backtrace_ms * 1e-3))
    cpu = None
    spot = None
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        if O.ref_available():
            k = min(n_cpu, n)
            idx = list(range(k))
            sub = O.Pairs([ps.pattern(i) for i in idx], [ps.text(i) for i in idx])
            threads = cpu_threads()
            ref = O.ref_align(O.Conf(metric, pen, cpu_mem), sub, threads)
            secs = ref.seconds
            bad = sum(1 for j, i in enumerate(idx) if res.status[i] != ref.status[j] or res.score[i] != ref.score[j]
                      or res.ops(i) != ref.ops(j))
            spot = {"pairs": k, "differ": int(bad)}
            if bad:
                raise RuntimeError("%s: GPU results differ from the reference build on %d pairs" % (name, bad))
            cells = float(sum(int(ps.plen[i]) * int(ps.tlen[i]) for i in idx))
            cpu = {"value": k / secs, "unit": "pairs/s", "cores": threads, "kind": "reference",
                   "sample": "first %d pairs of the set, wavefront_align of the reference built -O3 with OpenMP "
                             "(%s), one aligner per thread, memory mode %s; %.1f s wall; all %d results identical "
                             "to the GPU's" % (k, os.path.basename(O.ref_lib_path()), cpu_mem, secs, k),
                   "gcups": cells / secs / 1e9, "host": _HOST}
            log("%s reference: %.2f pairs/s on %d threads; %d pairs identical" % (name, k / secs, threads, k))
    return {"workload": "%s: first %d pairs, %s" % (name, n, desc), "pairs": n, "value": n / dt,
            "unit": "pairs/s", "seconds": dt, "sum_nm_cells": ps.cells, "gcups": ps.cells / dt / 1e9,
            "diagonal_steps": W, "diagonal_steps_per_s": W / dt,
            "forward_ms": st.forward_ms, "backtrace_ms": st.backtrace_ms, "launches": st.forward_launches,
            "roofline": {"bytes_per_diagonal_step": bytes_ds, "ops_per_diagonal_step": ops_ds,
                         "hbm_achieved_gbs": W * bytes_ds / fwd_s / 1e9 if fwd_s else None,
                         "hbm_frac": W * bytes_ds / fwd_s / (HBM_PEAK_GBS * 1e9) if fwd_s else None,
                         "valu_int32_frac": W * ops_ds / fwd_s / (VALU_INT32_PEAK_TOPS * 1e12) if fwd_s else None,
                         "over": "forward launches (HIP events on the aligner's stream)"},
            "parity_spot_check": spot, "cpu_baseline": cpu}


def long_line(name, n, no_cpu):
    """A long workload as a bench line of its own (--workload)."""
    r = long_workload(name, n, no_cpu)
    return {"metric": "pairs/s & equiv GCUPS, batched, exact CIGAR", "value": r["value"], "unit": "pairs/s",
            "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": r["seconds"] * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic (seeded SplitMix64 generator)",
            "config": {"workload": r["workload"], "pairs": r["pairs"], "sum_nm_cells": r["sum_nm_cells"],
                       "diagonal_steps": r["diagonal_steps"]},
            "detail": r, "cpu_baseline": r["cpu_baseline"]}


def copy_calibration(mb):
    """A plain device copy of known size: reads and writes mb MiB once each.
    rocprofv3 FETCH_SIZE / WRITE_SIZE of this kernel calibrate the counters."""
    import torch
    n = mb * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    torch.add(x, 0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=1000000)
    ap.add_argument("--plain-sample", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="open-loop serving run length (0: skip)")
    ap.add_argument("--serve-lanes", default="2,2,1", help="short,medium,long lanes of the request loop")
    ap.add_argument("--extras", type=int, default=1,
                    help="0: only the short-read batch (plain form, 1 kbp case and serving skipped)")
    ap.add_argument("--seq-pairs", type=int, default=500000, help="pairs of the .seq -> .alg run (0: skip)")
    ap.add_argument("--variant-pairs", type=int, default=100000,
                    help="pairs per configuration of the variant table (0: skip)")
    ap.add_argument("--workload", choices=("short-reads",) + tuple(LONG_WORKLOADS), default="short-reads",
                    help="short-reads: the headline batch; the others print their own line")
    ap.add_argument("--long-pairs", type=int, default=0, help="pairs of a long workload (0: its default)")
    ap.add_argument("--long-sets", type=int, default=1,
                    help="1: the default run also measures the 1-20 kbp and 100 kbp workloads on subsets")
    ap.add_argument("--mode", choices=("align", "copy"), default="align",
                    help="copy: one device copy of --copy-mb MiB (HBM counter calibration)")
    ap.add_argument("--copy-mb", type=int, default=1024)
    args = ap.parse_args()
    if args.mode == "copy":
        copy_calibration(args.copy_mb)
        return
    if args.workload != "short-reads":
        if args.gpus != 1:
            print("bench.py measures one GPU; run with --gpus 1.", file=sys.stderr)
            sys.exit(2)
        print(json.dumps(long_line(args.workload, args.long_pairs, args.no_cpu)))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the pairs are independent, so N GPUs would each run "
              "their own share with no exchange; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import wfa_amd
    from wfa_amd import datagen

    t0 = time.time()
    ps = datagen.generate(1, 1, args.pairs)
    log("generated %d pairs in %.2f s (sum n*m = %.3e)" % (ps.n, time.time() - t0, ps.cells))
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
    db = al.upload(ps)
    for i in range(args.warmup):
        rc = db.run()
        if rc != 0:
            raise RuntimeError("warmup run returned %d" % rc)
    times, fwd, bt, launches = [], [], [], []
    for i in range(args.steps):
        t = time.perf_counter()
        rc = db.run()
        times.append(time.perf_counter() - t)
        st = db.stats()
        fwd.append(st.forward_ms)
        bt.append(st.backtrace_ms)
        launches.append(st.forward_launches)
        if rc != 0:
            raise RuntimeError("timed run returned %d" % rc)
        log("step %d: %.2f ms (forward %.2f ms in %d launch(es), backtrace %.2f ms)"
            % (i, times[-1] * 1e3, st.forward_ms, st.forward_launches, st.backtrace_ms))
    dsteps = db.stats().diagonal_steps
    total = sum(times)
    ms_per_step = total / args.steps * 1e3
    pairs_per_s = ps.n * args.steps / total
    gcups = ps.cells * args.steps / total / 1e9

    # Results -> host (outside the timed region), for the PCIe-inclusive
    # rates: run-length CIGARs made on the device into a reused buffer
    # (wfa_device_batch_download_runs), and every operation byte
    # (wfa_device_batch_download).
    runs_buf = None
    d2h_runs = []
    for _ in range(3):
        t = time.perf_counter()
        rres = db.download_runs(runs_buf)
        d2h_runs.append(time.perf_counter() - t)
        runs_buf = rres.runs.base if rres.runs.base is not None else rres.runs
        db.run()  # the next download makes its runs again
    d2h_runs_s = min(d2h_runs)
    t = time.perf_counter()
    res = db.download()
    d2h_s = time.perf_counter() - t
    if (res.status != 0).any():
        raise RuntimeError("some pairs did not complete")
    import numpy as _np
    for i in _np.random.default_rng(1).choice(ps.n, size=2000, replace=False):
        if rres.ops(int(i)) != res.ops(int(i)) or rres.score[i] != res.score[i]:
            raise RuntimeError("run-length CIGAR of pair %d differs from its operations" % i)
    log("results to host: run-length CIGARs %.1f ms (%d runs), operation bytes %.1f ms"
        % (d2h_runs_s * 1e3, len(rres.runs), d2h_s * 1e3))

    # Plain form: the same kernels, one pair per launch (wavefront_align),
    # against the batched form on 10^5 mixed-length pairs (the first 10^5
    # short-read pairs, 100-300 bp), timed on a sample of the plain calls.
    plain = None
    k = min(args.plain_sample, ps.n) if args.extras else 0
    if k:
        mixed = ps.subset(list(range(min(100000, ps.n))))
        al_plain = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
        al_plain.align(ps.pattern(0), ps.text(0))
        t = time.perf_counter()
        for i in range(k):
            al_plain.align(mixed.pattern(i), mixed.text(i))
        plain_pairs_s = k / (time.perf_counter() - t)
        mixed_rate, _, _ = batch_rate(al, mixed)
        plain = {"value": plain_pairs_s, "unit": "pairs/s",
                 "sample": "first %d of %d mixed-length (100-300 bp) pairs, one wavefront_align call "
                           "(forward + backtrace launches) per pair" % (k, mixed.n),
                 "batched_same_pairs": mixed_rate, "batched_over_plain": mixed_rate / plain_pairs_s}
        log("plain form %.0f pairs/s; batched on the same %d pairs %.3e pairs/s (x%.0f)"
            % (plain_pairs_s, mixed.n, mixed_rate, mixed_rate / plain_pairs_s))

    # 1 kbp / 5% pairs (10,000, seed 0), the reference's own CPU-runnable case.
    c0_out = None
    if args.extras:
        c0 = datagen.generate(0, 0, 10000)
        c0_rate, c0_dt, _ = batch_rate(al, c0)
        log("1 kbp/5%%: %d pairs in %.2f ms (%.3e pairs/s)" % (c0.n, c0_dt * 1e3, c0_rate))
        c0_out = {"pairs": c0.n, "value": c0_rate, "unit": "pairs/s", "gcups": c0.cells * c0_rate / c0.n / 1e9}
        if not args.no_cpu:
            # every pair against the reference build (10,000 pairs take it
            # well under a second on the host's cores)
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            try:
                import oracle as O
            except Exception as exc:  # pragma: no cover
                O = None
                log("reference build unavailable: %s" % exc)
            if O is not None and O.ref_available():
                threads = cpu_threads()
                ref = O.ref_align(O.Conf("affine", (0, 4, 6, 2), "high"), O.Pairs(c0.patterns(), c0.texts()),
                                  threads)
                got = al.align_batch(c0)
                differ = sum(1 for i in range(c0.n) if got.status[i] != ref.status[i]
                             or got.score[i] != ref.score[i] or got.ops(i) != ref.ops(i))
                c0_out["cpu_reference"] = {"value": c0.n / ref.seconds, "unit": "pairs/s", "cores": threads,
                                           "sample": "all %d pairs, %s" % (c0.n, os.path.basename(O.ref_lib_path()))}
                c0_out["parity_check"] = {"pairs": c0.n, "differ": differ}
                log("1 kbp/5%%: every pair against the reference build: %d differ (reference %.3e pairs/s on %d "
                    "threads)" % (differ, c0.n / ref.seconds, threads))
                if differ:
                    raise RuntimeError("1 kbp/5%: GPU results differ from the reference build")

    variants = None
    if args.extras and args.variant_pairs > 0:
        variants = variant_measurements(args.variant_pairs, 4000, cpu_threads(), args.no_cpu)

    seq_file = seq_file_measurement(args.seq_pairs) if args.extras and args.seq_pairs > 0 else None

    serve = None
    if args.extras and args.serve_seconds > 0:
        serve = serve_measurement(args.serve_seconds, 41, log, tuple(int(x) for x in args.serve_lanes.split(",")))
        log("serving: p50 %.2f ms, p99 %.2f ms at %.1f requests/s"
            % (serve["p50_ms"], serve["p99_ms"], serve["rate_req_s"]))

    # The 1-20 kbp dual-cost and the 100 kbp BiWFA workloads, over seeded
    # subsets (sizes in each sub-line), with the reference on this host.
    longs = {}
    if args.extras and args.long_sets:
        for key, name in (("affine2p_1_20kbp", "affine2p-1-20kbp"), ("biwfa_100kbp", "biwfa-100kbp")):
            longs[key] = long_workload(name, 0, args.no_cpu)

    fwd_ms = float(np.mean(fwd))
    launches_per_step = float(np.mean(launches))
    alg_bytes = BYTES_PER_DSTEP * dsteps
    achieved = alg_bytes / (fwd_ms / launches_per_step * launches_per_step) / 1e6  # GB/s over the launches
    traffic = traffic_from_profiles()

    cpu = None
    if not args.no_cpu:
        threads = cpu_threads()
        cpu = cpu_baseline(ps, threads, res)
        if cpu and cpu["parity_spot_check"]["differ"]:
            raise RuntimeError("GPU results differ from the reference build")

    out = {
        "metric": "pairs/s & equiv GCUPS, batched vs plain GPU, exact CIGAR; serve p50/p99 ms",
        "value": pairs_per_s,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded SplitMix64 generator, uniform ACGT, substitution/1-base indel errors)",
        "config": {
            "workload": "short-read batch: %d gap-affine (0,4,6,2) end-to-end pairs, L~U[100,300] bp, "
                        "error ~U[0,10%%], score+CIGAR, memory-high semantics" % ps.n,
            "pairs": ps.n,
            "sum_nm_cells": ps.cells,
            "diagonal_steps": int(dsteps),
        },
        "gcups": gcups,
        "forward_ms_per_step": fwd_ms,
        "backtrace_ms_per_step": float(np.mean(bt)),
        "forward_launches_per_step": launches_per_step,
        "pcie_inclusive_pairs_per_s": ps.n / (total / args.steps + d2h_runs_s),
        "pcie_inclusive": {"runs_pairs_per_s": ps.n / (total / args.steps + d2h_runs_s),
                           "ops_pairs_per_s": ps.n / (total / args.steps + d2h_s),
                           "runs_download_ms": d2h_runs_s * 1e3, "ops_download_ms": d2h_s * 1e3,
                           "runs": int(len(rres.runs)),
                           "what": "one step plus the copy of every pair's status, score and CIGAR to host "
                                   "memory: as run-length words made on the device (the reported rate), or as "
                                   "one byte per operation"},
        "plain_gpu": plain,
        "kbp1_5pct": c0_out,
        "variants": variants,
        "seq_file": seq_file,
        "serve": serve,
        "long_workloads": longs or None,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": alg_bytes / launches_per_step,
            "valu_int32_frac": OPS_PER_DSTEP * dsteps / (fwd_ms * 1e-3) / (VALU_INT32_PEAK_TOPS * 1e12),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
