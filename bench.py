"""Benchmark: block-sparse contraction GFLOP/s on one MI355X.

Default workload: C[i,j,m] = sum_{k,l} A[i,j,k,l] B[k,l,m], every mode 600
long, split into uneven blocks drawn from {5,7,9,11,13,15,21,28} (seeded),
blocks present with p = 0.1, values u(-1,1) * exp(-0.3*|b0-b1|), FP64.  A
step is one execution of the contraction plan (all block products of the
contraction, one grouped launch per kernel class) with A, B resident in HBM.

Prints one JSON line.  Beside the FP64 drop-in (`value`) it reports the other
precision forms, p50/p99 step latency, the dominant kernel's roofline from HIP
events, parity against the CPU oracle on a sampled set of C blocks, and the
CPU baseline (the oracle's AVX2 fma-chain restatement with a work-stealing
pool) timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X dense peaks per useful flop: FP64 (vector = matrix) 78.6 TFLOP/s,
# FP32 MFMA 157.3; split-BF16 runs six bf16 products per useful product
# (2.5 PFLOP/s / 6).
PEAK_BY_PREC = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15 / 6}


def peak_of(form):
    return PEAK_BY_PREC[form.split("_")[0]]


HBM_PEAK = 8.0e12
# kernel class reported by sht_plan_execute_timed -> substring of the kernel's
# symbol in rocprofv3's output
KERNEL_SYMBOL = {
    "as_mfma64": "k_asp<2, 8, 0, 1, true>", "as_mfma64_g": "k_asl<0, 4, true, 4>",
    "as_mfma32": "k_asp<2, 8, 1, 1, false>", "as_mfma32_g": "k_asl<1, 4, false, 4>", "as_bf16x3": "k_asp<2, 8, 2, 1, false>",
    "as_bf16x3_g": "k_asb", "as_mfma64_n4": "k_asp<4, 4, 0, 1, true>",
    "as_mfma32_n4": "k_asp<4, 4, 1, 1, false>",
    "as_exact": "k_as_exact", "exact_n8": "k_exact<8>", "exact_n16": "k_exact<16>", "exact_n32": "k_exact<32>",
    "mfma64_n16": "k_mfma64<1, 4>", "mfma64_n32": "k_mfma64<2, 4>",
    "mfma64_n16_r32": "k_mfma64<1, 2>", "mfma32_n16_r32": "k_mfma32<1, 2>",
    "mfma64_n64w": "k_mfmaw<0, 64, 64>", "mfma32_n64w": "k_mfmaw<1, 64, 64>", "bf16x3_n64w": "k_bf3w<64, 64>",
    "mfma64_n32s": "k_mfmaw<0, 32, 32>", "mfma32_n32s": "k_mfmaw<1, 32, 32>",
    "mfma64_n16w": "k_mfmaw<0, 64, 16>", "mfma32_n16w": "k_mfmaw<1, 64, 16>",
    "mfma64_n32w": "k_mfmaw<0, 64, 32>", "mfma32_n32w": "k_mfmaw<1, 64, 32>",
    "mfma32_n16w_r128": "k_mfmaw<1, 128, 16>", "mfma32_n32w_r128": "k_mfmaw<1, 128, 32>", "bf16x3_n32s": "k_bf3w<32, 32>",
    "bf16x3_n16w": "k_bf3w<64, 16>",
    "mfma64_n16_r128": "k_mfma64<1, 8>", "mfma64_n32_r128": "k_mfma64<2, 8>",
}
CHOICES = [5, 7, 9, 11, 13, 15, 21, 28]

CONFIGS = {
    # name: (expr, length, block sizes or None (uneven), p, lambda, seed)
    "c0": ("ijk,kl->ijl", 256, [16], 0.30, 0.0, 1000),
    "c1": ("ijkl,klm->ijm", 600, None, 0.10, 0.3, 1001),
    "c2": ("ijk,km->ijm", 1200, None, 0.10, 0.3, 1002),
    "c3_32": ("ik,kj->ij", 4096, [32], 0.25, 0.1, 1003),
    "c3_64": ("ik,kj->ij", 4096, [64], 0.25, 0.1, 1003),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def axis_sizes(cfg, length=None):
    expr, L, fixed, p, lam, seed = CONFIGS[cfg]
    L = length or L
    letters = []
    for ch in expr.replace("->", ",").replace(",", ""):
        if ch not in letters:
            letters.append(ch)
    out = {}
    for ax, ch in enumerate(letters):
        if fixed:
            b = fixed[0]
            out[ch] = [b] * (L // b) + ([L % b] if L % b else [])
        else:
            import shtensor_amd as sht
            out[ch] = sht.gen_block_sizes(seed, ax, L, CHOICES)
    return out


def build_tensors(sht, ctx, cfg, length=None):
    expr, L, fixed, p, lam, seed = CONFIGS[cfg]
    sizes = axis_sizes(cfg, length)
    ea, rest = expr.split(",")
    eb, ec = rest.split("->")
    A = sht.Tensor(ctx, [sizes[x] for x in ea])
    B = sht.Tensor(ctx, [sizes[x] for x in eb])
    A.reserve_random(seed, 0, p)
    B.reserve_random(seed, 1, p)
    A.fill_random(seed, 0, lam)
    B.fill_random(seed, 1, lam)
    Cc = sht.Tensor(ctx, [sizes[x] for x in ec])
    return expr, sizes, (ea, eb, ec), A, B, Cc


def pick_sample(cfg, sizes, eabc, orc, target_flops):
    """CPU sample: A restricted to a seeded set of A row groups (the
    combined A free-index block coordinates), B whole.  The C blocks of those
    row groups are exact copies of the full contraction's blocks (C blocks are
    independent), so parity on them is exact."""
    expr, L, fixed, p, lam, seed = CONFIGS[cfg]
    ea, eb, ec = eabc
    bsA = [sizes[x] for x in ea]
    bsB = [sizes[x] for x in eb]
    bsC = [sizes[x] for x in ec]
    nA = np.array([len(b) for b in bsA], dtype=np.int64)
    nB = np.array([len(b) for b in bsB], dtype=np.int64)
    idsA = orc.gen_presence(seed, 0, nA, p)
    idsB = orc.gen_presence(seed, 1, nB, p)
    cA = [(idsA // s) % n for s, n in zip(np.cumprod(np.r_[1, nA[:-1]]), nA)]
    cB = [(idsB // s) % n for s, n in zip(np.cumprod(np.r_[1, nB[:-1]]), nB)]
    con = [x for x in ea if x in eb]
    kA = np.zeros(len(idsA), np.int64)
    kB = np.zeros(len(idsB), np.int64)
    mul = 1
    for x in con:
        kA += cA[ea.index(x)] * mul
        kB += cB[eb.index(x)] * mul
        mul *= len(sizes[x])
    volA = np.ones(len(idsA), np.float64)
    for d, x in enumerate(ea):
        volA *= np.asarray(sizes[x], np.float64)[cA[d]]
    nfreeB = np.ones(len(idsB), np.float64)
    for d, x in enumerate(eb):
        if x not in ea:
            nfreeB *= np.asarray(sizes[x], np.float64)[cB[d]]
    S = np.bincount(kB, weights=nfreeB, minlength=mul)
    flopsA = 2.0 * volA * S[kA]
    row = np.zeros(len(idsA), np.int64)
    mul = 1
    for d, x in enumerate(ea):
        if x not in eb:
            row += cA[d] * mul
            mul *= len(sizes[x])
    rowflops = np.bincount(row, weights=flopsA, minlength=mul)
    order = np.random.default_rng(seed).permutation(mul)
    cum = np.cumsum(rowflops[order])
    nrows = int(min(len(order), np.searchsorted(cum, target_flops) + 1))
    chosen = order[:nrows]
    As = orc.BlockSparse(bsA).reserve_ids(idsA[np.isin(row, chosen)])
    orc.gen_values(seed, 0, As, lam)
    B = orc.BlockSparse(bsB).reserve_ids(idsB)
    orc.gen_values(seed, 1, B, lam)
    Cs = orc.BlockSparse(bsC)
    _, _, cids = orc.tasks(expr, As, B, Cs)
    Cs.reserve_ids(cids)
    return As, B, Cs, nrows, float(cum[nrows - 1])


def measure_traffic(args, form, schedule, kclass, c_nelem):
    """HBM bytes per launch of the dominant kernel: two child runs of this
    script under `rocprofv3 --pmc` (FETCH_SIZE, then WRITE_SIZE; KiB per
    dispatch), each scaled by what the same counter reads for a plain
    read+write pass of known size (C = 1*C over C's elements, the --calib
    pass).  Runs after this process has released its device memory."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    rp = shutil.which("rocprofv3")
    sym = KERNEL_SYMBOL.get(kclass)
    if not rp or not sym:
        return None
    res = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="sht_pmc_", dir="/tmp")
        cmd = [rp, "--pmc", ctr, "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable,
               os.path.join(ROOT, "bench.py"), "--config", args.config, "--forms", form, "--steps", "1",
               "--warmup", "0", "--no-cpu", "--no-traffic", "--calib", "--schedule", str(schedule)]
        if args.length:
            cmd += ["--length", str(args.length)]
        log("traffic: %s" % " ".join(cmd))
        try:
            subprocess.run(cmd, cwd=d, env=env, timeout=300, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        except subprocess.TimeoutExpired:
            log("traffic: %s run timed out" % ctr)
            return None
        files = glob.glob(os.path.join(d, "**", "run_counter_collection.csv"), recursive=True)
        if not files:
            log("traffic: no counter output for %s" % ctr)
            return None
        kern, calib = [], []
        with open(files[0]) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != ctr:
                    continue
                v = float(row["Counter_Value"]) * 1024.0
                if sym in row["Kernel_Name"]:
                    kern.append(v)
                elif "k_scale" in row["Kernel_Name"]:
                    calib.append(v)
        shutil.rmtree(d, ignore_errors=True)
        if not kern or not calib or max(calib) <= 0:
            log("traffic: kernel or calibration dispatch missing for %s" % ctr)
            return None
        factor = 8.0 * c_nelem / max(calib)
        res[ctr] = {"measured": float(np.mean(kern)), "factor": factor, "bytes": float(np.mean(kern)) * factor}
    res["bytes"] = res["FETCH_SIZE"]["bytes"] + res["WRITE_SIZE"]["bytes"]
    return res


SERVE_EXPR = {2: "ik,kl->il", 3: "ijk,kl->ijl", 4: "ijmk,kl->ijml"}


def serve_bench(args):
    """Mixed request stream through the serving loop (continuous batching,
    admission by footprint, merged launches): requests of order 2-4 (the last
    mode of A contracted with B[k,l]), mode length L in [128, 1024], uniform
    block size b in {8,16,24,32}, occupancy p ~ U(0.05, 0.5), FP64 or FP32,
    arriving as a Poisson process of rate lambda.  A request whose expected
    footprint (A + B + C bytes) exceeds the budget is rejected by the same rule
    the server applies at submit, before its operands would be generated.
    Admitted requests are generated once (seeded, on the GPU) and replayed for
    every rate; latency = submit to completion."""
    import shtensor_amd as sht
    rng = np.random.default_rng(1004)
    budget = int(float(os.environ.get("SHT_SERVE_BUDGET_GB", "200")) * (1 << 30))
    gen_cap = budget // 2  # operands generated for the stream stay within half the budget
    ctx = sht.Context(budget + (16 << 30))
    reqs, rejected, gen_bytes = [], 0, 0
    for r in range(args.serve_requests):
        d = int(rng.choice([2, 3, 4]))
        b = int(rng.choice([8, 16, 24, 32]))
        L = int(rng.integers(128, 1025)) // b * b
        p = float(rng.uniform(0.05, 0.5))
        prec = sht.PREC_FP64 if rng.random() < 0.5 else sht.PREC_FP32
        nb = L // b
        a_el = p * nb ** d * b ** d
        b_el = p * nb ** 2 * b ** 2
        c_el = min(1.0, p * nb * p) * nb ** d * b ** d  # C block present if any k matches
        foot = 8.0 * (a_el + b_el + c_el)
        if foot > budget:
            rejected += 1
            continue
        if gen_bytes + foot > gen_cap:
            continue
        gen_bytes += foot
        expr = SERVE_EXPR[d]
        ea, rest = expr.split(",")
        eb, ec = rest.split("->")
        bs = {ch: [b] * nb for ch in set(expr) if ch.isalpha()}
        A = sht.Tensor(ctx, [bs[x] for x in ea])
        B = sht.Tensor(ctx, [bs[x] for x in eb])
        A.reserve_random(2000 + r, 0, p)
        B.reserve_random(2000 + r, 1, p)
        A.fill_random(2000 + r, 0, 0.0)
        B.fill_random(2000 + r, 1, 0.0)
        reqs.append({"expr": expr, "A": A, "B": B, "cbs": [bs[x] for x in ec], "prec": prec, "d": d, "L": L,
                     "b": b, "p": p})
    sht.call("sht_device_synchronize")
    log("serve: %d requests admitted and generated (%.1f GB), %d rejected by footprint" % (
        len(reqs), gen_bytes / 1e9, rejected))
    # useful flops of each request (one solo run, also a warm-up of every shape's path)
    for q in reqs:
        Ct = sht.Tensor(ctx, q["cbs"])
        pl = sht.Plan(ctx, q["expr"], q["A"], q["B"], Ct, q["prec"])
        q["flops"] = pl.stats()["flops"]
        pl.execute(1.0, 0.0)
        sht.call("sht_device_synchronize")
        pl.close()
        Ct.close()
    total_flops = sum(q["flops"] for q in reqs)
    rates = [float(x) for x in args.serve_rates.split(",")]
    runs = []
    for lam in rates:
        srv = sht.Server(ctx, budget, max_batch=64)
        arr = np.cumsum(np.random.default_rng(7).exponential(1.0 / lam, len(reqs)))
        outs, ids = [], []
        t0 = time.perf_counter()
        for q, ta in zip(reqs, arr):
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            Ct = sht.Tensor(ctx, q["cbs"])
            outs.append(Ct)
            ids.append(srv.submit(q["expr"], 1.0, q["A"], q["B"], 0.0, Ct, q["prec"]))
        lat = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        for Ct in outs:
            Ct.close()
        lat = np.array(lat)
        runs.append({"lambda_per_s": lam, "requests": len(reqs), "wall_s": round(wall, 3),
                     "gflops": round(total_flops / wall / 1e9, 3),
                     "p50_ms": round(float(np.percentile(lat, 50)), 3),
                     "p99_ms": round(float(np.percentile(lat, 99)), 3),
                     "batches": st["batches"], "max_merged": st["max_merged"], "launches": st["launches"]})
        log("serve lambda=%g/s: %s" % (lam, runs[-1]))
    best = max(runs, key=lambda x: x["gflops"])
    out = {
        "metric": "contraction GFLOP/s FP64 vs batched/FP32 at 1e-6 rel err; p50/p99 latency",
        "value": best["gflops"], "unit": "GFLOP/s", "n_gpus": 1, "steps": len(reqs), "warmup": 1,
        "ms_per_step": round(1e3 * best["wall_s"] / max(1, len(reqs)), 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic (seeded SplitMix64)",
        "config": {"workload": "c4 mixed request stream (orders 2-4, L 128-1024, b 8-32, p 0.05-0.5, FP64/FP32)",
                   "admitted": len(reqs), "rejected_by_footprint": rejected, "budget_bytes": budget,
                   "useful_flops": total_flops},
        "latency_ms": {"p50": best["p50_ms"], "p99": best["p99_ms"]},
        "serve": runs, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    for q in reqs:
        q["A"].close()
        q["B"].close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c1", choices=sorted(CONFIGS) + ["c4"])
    ap.add_argument("--serve-requests", type=int, default=200, help="c4: requests drawn for the stream")
    ap.add_argument("--serve-rates", default="50,200,1000", help="c4: Poisson arrival rates (requests/s)")
    ap.add_argument("--length", type=int, default=None, help="override every mode length (smaller runs)")
    ap.add_argument("--forms", default="fp64,fp64_exact,fp32,bf16x3",
                    help="fp64 (autotuned schedule), fp64_cs, fp64_as1, fp64_as2, fp64_as3, fp64_exact, fp32, "
                         "bf16x3; _cs = C-stationary (schedule 0), _asN = A-stationary schedule N (1: one-wave "
                         "units, 2: four units of a block row per workgroup, 3: narrow sweeps)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU baseline work")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-max-flops", type=float, default=1.5e11, help="cap on the CPU sample's work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--schedule", type=int, default=-1,
                    help="force the execution schedule (0, 1, 2) instead of the form's default/autotuned one")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE runs that measure the dominant kernel's HBM bytes")
    ap.add_argument("--graph", action="store_true", help="time hipGraph replays instead of eager steps")
    ap.add_argument("--calib", action="store_true",
                    help="also run C = 1*C (alpha=0) through the A-stationary path: a plain read+write of C "
                         "(counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d refused)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.config == "c4":
        serve_bench(args)
        return
    import shtensor_amd as sht

    t0 = time.time()
    base = {"fp64": sht.PREC_FP64, "fp64_exact": sht.PREC_FP64_EXACT, "fp32": sht.PREC_FP32, "bf16x3": sht.PREC_BF16X3}
    precs = dict(base)
    for b, pr in base.items():
        for suf in ("_cs", "_as1", "_as2", "_as3"):
            precs[b + suf] = pr
    forms = [f for f in args.forms.split(",") if f]
    arena = int(float(os.environ.get("SHT_BENCH_ARENA_GB", "200")) * (1 << 30))
    ctx = sht.Context(arena)
    expr, sizes, eabc, A, B, Cc = build_tensors(sht, ctx, args.config, args.length)
    log("built tensors in %.1fs: A %d blocks (%d elems), B %d blocks" % (
        time.time() - t0, A.get_nb_nzblocks_local(), A.nelem(), B.get_nb_nzblocks_local()))
    stream = sht.Stream()
    results = {}
    ref_form = None
    Ctens = {}
    for f in forms:
        Cf = sht.Tensor(ctx, Cc.block_sizes)
        tp = time.time()
        plan = sht.Plan(ctx, expr, A, B, Cf, precs[f])
        tuned = None
        if args.schedule >= 0:
            plan.set_schedule(args.schedule)
        elif f in ("fp64", "fp64_exact", "fp32", "bf16x3"):
            tuned = plan.autotune(1.0, 0.0, reps=2, stream=stream)  # picks the schedule
        elif "_as" in f:
            plan.set_schedule(int(f[-1]))
        st = plan.stats()
        st["schedule"] = plan.get_schedule()
        st["autotune_ms"] = tuned
        log("%s plan: %d tasks, %d C blocks, %.3e flop, %.2fs" % (f, st["ntasks"], st["ncblocks"], st["flops"],
                                                                  time.time() - tp))
        for _ in range(args.warmup):
            plan.execute(1.0, 0.0, stream)
        stream.synchronize()
        times = []
        if args.graph:
            plan.graph_replay(1.0, 0.0, 1, stream)
            stream.synchronize()
        for _ in range(args.steps):
            ts = time.perf_counter()
            if args.graph:
                plan.graph_replay(1.0, 0.0, 1, stream)
            else:
                plan.execute(1.0, 0.0, stream)
            stream.synchronize()
            times.append(time.perf_counter() - ts)
        timed = plan.execute_timed(1.0, 0.0, stream)
        if args.calib:
            sched = plan.get_schedule()
            if sched == 0:
                plan.set_schedule(1)  # the A-stationary path starts with C = beta*C (k_scale)
            plan.execute(0.0, 1.0, stream)  # k_scale only: reads and writes every C element once
            stream.synchronize()
            plan.set_schedule(sched)
            log("calibration: C = 1*C over %d elements (%.3e bytes read, same written)" % (Cf.nelem(),
                                                                                         8.0 * Cf.nelem()))
        times = np.array(times)
        results[f] = {"stats": st, "times": times, "launches": timed, "plan": plan}
        Ctens[f] = Cf
        log("%s: median %.3f ms -> %.1f GFLOP/s; launches %s" % (
            f, 1e3 * np.median(times), st["flops"] / np.median(times) / 1e9, timed))

    # parity + CPU baseline on a sample of C blocks
    parity = {}
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
        tgt = min(args.cpu_seconds * nthreads * 3e9, args.cpu_max_flops)
        ts = time.time()
        As, Bs, Cs, nrows, sflops = pick_sample(args.config, sizes, eabc, orc, tgt)
        log("CPU sample: %d A row groups, %d A blocks, %d C blocks, %.3e flop (%.1fs to build)" % (
            nrows, len(As.indices), len(Cs.indices), sflops, time.time() - ts))
        reps = []
        for _ in range(3):
            tc = time.perf_counter()
            ntask = orc.contract(expr, 1.0, As, Bs, 0.0, Cs, mode=0, nthreads=nthreads)
            reps.append(time.perf_counter() - tc)
        cpu_s = float(np.median(reps))
        cpu = {"value": round(sflops / cpu_s / 1e9, 3), "unit": "GFLOP/s", "cores": nthreads, "kind": "port",
               "sample": "%d of %d C blocks (every C block of %d seeded A row groups), %d tasks, %.3e flop, "
                         "median of 3 runs %.2f s" % (len(Cs.indices), results[forms[0]]["stats"]["ncblocks"], nrows,
                                                      ntask, sflops, cpu_s)}
        log("CPU: %s" % cpu)
        ref = Cs.data
        ends = np.r_[Cs.offsets[1:], len(ref)]
        for f in forms:
            Cf = Ctens[f]
            info = Cf.get_sparse_info()
            data = Cf.download()
            pos = np.searchsorted(info["indices"], Cs.indices)
            assert np.array_equal(info["indices"][pos], Cs.indices), "C pattern differs on the sample"
            got = np.empty_like(ref)
            mx = 0.0
            for q, pp in enumerate(pos):
                o0, o1 = Cs.offsets[q], ends[q]
                g0 = info["offsets"][pp]
                got[o0:o1] = data[g0:g0 + (o1 - o0)]
                nb = np.linalg.norm(ref[o0:o1])
                if nb > 0:
                    mx = max(mx, float(np.linalg.norm(got[o0:o1] - ref[o0:o1]) / nb))
            err = float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))
            parity[f] = {"normwise_rel": err, "max_blockwise_rel": mx, "bitwise": bool(np.array_equal(got, ref)),
                         "cblocks": int(len(pos))}
            log("parity %s: %s" % (f, parity[f]))

    main_f = forms[0]
    r = results[main_f]
    med = float(np.median(r["times"]))
    value = r["stats"]["flops"] / med / 1e9
    # dominant kernel roofline (HIP events around each grouped launch)
    dom = max(r["launches"], key=lambda x: x[1])
    tot_ms = sum(ms for _, ms in r["launches"])
    # The bound follows the workload's algorithmic intensity: flops per
    # algorithmic byte below the ridge (peak / 8 TB/s) is HBM-bound.  The
    # dominant launch is credited with its share (by time) of the step's work.
    share = dom[1] / tot_ms
    alg_bytes = r["stats"].get("bytes") or 0.0
    ai = r["stats"]["flops"] / alg_bytes if alg_bytes else float("inf")
    if ai < peak_of(main_f) / HBM_PEAK:
        roof = {"bound": "hbm", "kernel": dom[0],
                "achieved": round(alg_bytes * share / (dom[1] * 1e-3) / 1e9, 1),
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "traffic": None}
    else:
        roof = {"bound": "mfma", "kernel": dom[0],
                "achieved": round(r["stats"]["flops"] * share / (dom[1] * 1e-3) / 1e12, 3),
                "peak": round(peak_of(main_f) / 1e12, 1), "unit": "TFLOP/s", "traffic": None}
    roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    roof["intensity_flop_per_byte"] = round(ai, 2)
    if roof["bound"] == "mfma" and precs[main_f] in (sht.PREC_FP64, sht.PREC_FP64_EXACT):
        # the FP64 MFMA rate a register-resident loop sustains on this device
        loop = sht.diag_mfma_f64_rate(stream)
        roof["mfma_loop_tflops"] = round(loop, 2)
        roof["frac_of_mfma_loop"] = round(roof["achieved"] / loop, 4)
        log("FP64 MFMA loop on this device: %.2f TFLOP/s" % loop)
    out = {
        "metric": "contraction GFLOP/s FP64 vs batched/FP32 at 1e-6 rel err; p50/p99 latency",
        "value": round(value, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * med, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64, generated on the GPU)",
        "config": {"workload": "%s %s L=%d" % (args.config, expr, args.length or CONFIGS[args.config][1]),
                   "ntasks": r["stats"]["ntasks"], "ncblocks": r["stats"]["ncblocks"],
                   "flops_per_step": r["stats"]["flops"], "graph": bool(args.graph)},
        "latency_ms": {"p50": round(1e3 * float(np.percentile(r["times"], 50)), 4),
                       "p99": round(1e3 * float(np.percentile(r["times"], 99)), 4)},
        "forms": {f: {"gflops": round(results[f]["stats"]["flops"] / float(np.median(results[f]["times"])) / 1e9, 3),
                      "ms_p50": round(1e3 * float(np.percentile(results[f]["times"], 50)), 4),
                      "ms_p99": round(1e3 * float(np.percentile(results[f]["times"], 99)), 4),
                      "launches": results[f]["launches"],
                      "schedule": results[f]["stats"].get("schedule"),
                      "autotune_ms": results[f]["stats"].get("autotune_ms"),
                      "parity": parity.get(f)} for f in forms},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    c_nelem = Ctens[main_f].nelem()
    for f in forms:
        results[f]["plan"].close()
    ctx.close()
    if not args.no_traffic:
        tr = measure_traffic(args, main_f, r["stats"]["schedule"], dom[0], c_nelem)
        if tr:
            roof["traffic"] = round(tr["bytes"], 0)
            roof["traffic_detail"] = {k: {"measured_bytes": round(v["measured"], 0), "calibration": round(v["factor"], 4)}
                                      for k, v in tr.items() if k != "bytes"}
            roof["algorithmic_bytes"] = r["stats"].get("bytes")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
