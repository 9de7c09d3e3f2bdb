"""Benchmark of the exact WFA hot path on one MI355X.

Workload (one "step"): the short-read batch -- 1,000,000 gap-affine
(match 0, mismatch 4, open 6, extend 2) end-to-end pairs, text length uniform
in [100,300] bp, error rate uniform in [0,10%] per pair (substitutions and
1-base indels), score + CIGAR, synthetic seeded data.  The pairs are packed
and resident in HBM before timing; one step aligns all of them (forward
kernel + backtrace kernel, CIGARs left in HBM).

Prints one JSON line (rank 0) with pairs/s as `value`, equivalent GCUPS,
the roofline of the dominant (forward) kernel, the plain one-pair-per-launch
GPU form and the reference's OpenMP CPU build timed on this host.
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
            "gcups": ps.cells / secs / 1e9}


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
    ap.add_argument("--mode", choices=("align", "copy"), default="align",
                    help="copy: one device copy of --copy-mb MiB (HBM counter calibration)")
    ap.add_argument("--copy-mb", type=int, default=1024)
    args = ap.parse_args()
    if args.mode == "copy":
        copy_calibration(args.copy_mb)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the pairs are independent, so N GPUs would each run "
              "their own share with no exchange (see DESIGN.md); run with --gpus 1.", file=sys.stderr)
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

    # Results -> host (outside the timed region), for the PCIe-inclusive rate.
    t = time.perf_counter()
    res = db.download()
    d2h_s = time.perf_counter() - t
    if (res.status != 0).any():
        raise RuntimeError("some pairs did not complete")

    # Plain form: the same kernels, one pair per launch (wavefront_align).
    al_plain = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
    k = min(args.plain_sample, ps.n)
    if k:
        al_plain.align(ps.pattern(0), ps.text(0))
    t = time.perf_counter()
    for i in range(k):
        al_plain.align(ps.pattern(i), ps.text(i))
    plain_s = time.perf_counter() - t
    plain_pairs_s = k / plain_s if k else None

    fwd_ms = float(np.mean(fwd))
    launches_per_step = float(np.mean(launches))
    alg_bytes = BYTES_PER_DSTEP * dsteps
    achieved = alg_bytes / (fwd_ms / launches_per_step * launches_per_step) / 1e6  # GB/s over the launches
    traffic = traffic_from_profiles()

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
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
        "pcie_inclusive_pairs_per_s": ps.n / (total / args.steps + d2h_s),
        "plain_gpu": {"value": plain_pairs_s, "unit": "pairs/s", "sample": "first %d pairs, one "
                      "wavefront_align call (one forward + one backtrace launch) per pair" % k,
                      "batched_over_plain": (pairs_per_s / plain_pairs_s) if plain_pairs_s else None},
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
