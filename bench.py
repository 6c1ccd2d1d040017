#!/usr/bin/env python3
"""Benchmark of the RNS multi-limb negacyclic NTT/INTT on one MI355X.

Workload: Montgomery variant, N = 2^16, 16 limbs (the 16 largest primes
q < 2^30 with q = 1 mod 2^17), a batch of B operands (default 256 = 1 GiB,
four times the 256 MB last-level cache) resident in HBM.  One step is a
forward NTT then an INTT of every operand in the batch (the batched path: one
launch per pass over the whole batch).  Operands are uniform residues
generated on the device; after the timed region the batch must equal its
initial contents (the roundtrip check).

Prints ONE JSON line on stdout (diagnostics go to stderr):
  value      total butterflies (forward + inverse) per second, in G/s
  roofline   the slowest kernel of the engine in use: algorithmic bytes per
             launch (8 B per word: read once, written once) / its mean
             duration, measured with HIP events on the launch stream; traffic
             = counter-measured HBM bytes of that kernel (profiles/)
  compute_roof  the same kernel's butterflies/s against the butterfly roof
             measured live (the kernels' butterfly code on registers only)
  cpu_baseline  the oracle's butterfly loops (src/v04_montgomery.cu:742-784
             restated in C) on one host core, table generation excluded
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "NTT/INTT Gbutterfly/s & %HBM peak, logN=16×16 limbs; served p50/p99 ms"
HBM_PEAK_GBS = 8000.0
RTX4090_V04_NTT_GBFLY = 487.3     # reference README.md:209, 0.017 ms incl. restore copy


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--logN", type=int, default=16)
    ap.add_argument("--limbs", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the per-transform and single-transform legs")
    ap.add_argument("--no-served", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--served-requests", type=int, default=20000)
    ap.add_argument("--no-mixed", action="store_true", help="skip the batched uneven (256-job) leg")
    ap.add_argument("--no-tune", action="store_true", help="use the default phase config instead of the autotuner's")
    ap.add_argument("--engine", choices=["auto", "fused", "two_pass"], default="auto",
                    help="force a kernel family for the batched leg")
    ap.add_argument("--phase-config", default=None,
                    help="fixed two-pass config 'R,S,W;R,S,W;R,S,W;R,S,W' (ntt1;ntt2;intt1;intt2), for profiling")
    ap.add_argument("--fused-shape", type=int, choices=[0, 6, 7], default=0,
                    help="force the fused workgroup shape (0: the tuned one)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain copy of the batch (known bytes, for counter calibration)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "hbm_traffic.txt"),
                    help="HBM bytes per dispatch measured with rocprofv3 counters (see that file)")
    return ap.parse_args()


def hbm_traffic(path: str, kernel: str, workload: str):
    """HBM bytes per dispatch of `kernel` on `workload` from a counter summary:
    lines `kernel=<name>;workload=<tag>;fetch_bytes=<n>;write_bytes=<n>` (already
    scaled by the copy-kernel calibration).  None when not measured."""
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        if ln.startswith("#") or "=" not in ln:
            continue
        kv = dict(f.split("=", 1) for f in ln.strip().split(";") if "=" in f)
        if kv.get("kernel") == kernel and kv.get("workload") == workload:
            return int(kv["fetch_bytes"]) + int(kv["write_bytes"])
    return None


def bench_word64(ntt_amd, dev, stream, logN: int, limbs: int, pool_bytes: int = 1 << 30, reps: int = 5):
    """64-bit words (ntt_plan64): forward + inverse over a 1 GiB pool of
    operands of `limbs` 62-bit limbs, against the live 64-bit butterfly roof."""
    import numpy as np
    moduli = ntt_amd.generate_moduli64(logN, limbs)
    N = 1 << logN
    count = max(1, pool_bytes // (8 * limbs * N))
    rng = np.random.default_rng(0)
    one = np.stack([rng.integers(0, int(m.q), N, dtype=np.uint64) for m in moduli])
    pool = dev.DeviceBuffer(8 * limbs * N * count)
    for k in range(count):
        pool.upload(one, offset_bytes=8 * limbs * N * k)
    ctx = ntt_amd.Context(0)
    plan = ctx.plan64(logN, moduli)
    e0, e1 = dev.Event(), dev.Event()
    plan.forward(pool, count, stream)
    plan.inverse(pool, count, stream)
    stream.synchronize()
    out = {}
    for name, fn in (("forward", lambda: plan.forward(pool, count, stream)),
                     ("inverse", lambda: plan.inverse(pool, count, stream))):
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        out[name + "_ms"] = round(e0.elapsed_ms(e1) / reps, 4)
    # reps forwards then reps inverses: the pool must be back to its contents
    roundtrip = bool(np.array_equal(pool.numpy(np.uint64, nbytes=8 * limbs * N).reshape(one.shape), one))
    bfly = count * limbs * (N // 2) * logN
    step_ms = out["forward_ms"] + out["inverse_ms"]
    roof = dev.butterfly_roof(dev.BUTTERFLY_CT64, stream)
    plan.close()
    return {"workload": f"N=2^{logN}, {limbs} limbs of 62-bit primes, {count} operands ({8 * limbs * N * count >> 20} MiB)",
            **out, "gbfly_s": round(2 * bfly / (step_ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(2 * 2 * 8 * limbs * N * count / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "roof_ct64_gbfly_s": round(roof / 1e9, 1),
            "forward_vs_roof": round(bfly / (out["forward_ms"] * 1e-3) / roof, 4), "roundtrip_ok": roundtrip}


def bench_mixed(ntt_amd, dev, stream, reps: int = 10, tune: bool = True, seed: int = 0):
    """The batched uneven workload: 256 jobs, logN ~ U{12..17}, L ~ U{1..48},
    direction ~ Bernoulli(1/2), Montgomery domain, seed 0, HBM-resident.
    Times the captured graph, the eager batch, and one plan launch per job."""
    import numpy as np
    from ntt_amd import workloads

    wl = workloads.batched_uneven(seed)
    ctx = ntt_amd.Context(0)
    t0 = time.perf_counter()
    if tune:
        # tune each size at the launch it will see: one launch per
        # (logN, direction) group over all its jobs' limbs
        for n in sorted({j.logN for j in wl}):
            tot = max(sum(j.limbs for j in wl if j.logN == n and j.direction == d) for d in (0, 1))
            limbs = min(tot, 48)
            ctx.autotune(n, workloads.moduli_for(n, limbs), count=max(1, tot // limbs), reps=2)
    tune_s = time.perf_counter() - t0
    bufs, spec = [], []
    qcache = {}
    for k, j in enumerate(wl):
        m = workloads.moduli_for(j.logN, j.limbs)
        buf = dev.DeviceBuffer(4 * j.words)
        qd = qcache.get((j.logN, j.limbs))
        if qd is None:
            qd = qcache[(j.logN, j.limbs)] = dev.DeviceBuffer.from_numpy(np.array([q for q, _ in m], np.uint32))
        dev.fill_uniform(buf, qd, j.limbs, j.logN, 1, seed=k)
        bufs.append(buf)
        spec.append((buf, j.logN, m, j.direction, ntt_amd.MONTGOMERY))
    dev.synchronize()
    total_bfly = sum(j.butterflies for j in wl)
    total_bytes = sum(8 * j.words for j in wl)
    batch = ctx.batch(spec)
    batch.capture(stream)
    e0, e1 = dev.Event(), dev.Event()

    def timed(fn, n):
        fn()
        stream.synchronize()
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_ms(e1) / n

    graph_ms = timed(lambda: batch.replay(stream), reps)
    eager_ms = timed(lambda: batch.run(stream), reps)
    plans = {}
    for (buf, n, m, d, dom), j in zip(spec, wl):
        if (n, j.limbs) not in plans:
            plans[(n, j.limbs)] = ctx.plan(n, m, dom)

    def per_transform():
        for (buf, n, m, d, dom), j in zip(spec, wl):
            p = plans[(n, j.limbs)]
            (p.inverse if d else p.forward)(buf, 1, stream)

    per_ms = timed(per_transform, max(2, reps // 3))
    info = batch.info()
    out = {
        "workload": "256 jobs, logN ~ U{12..17}, limbs ~ U{1..48}, direction ~ Bernoulli(1/2), "
                    "Montgomery, moduli = first L primes < 2^30 with q = 1 mod 2N, seed 0",
        "operand_bytes": total_bytes // 2, "butterflies": total_bfly,
        "graph_ms": round(graph_ms, 4), "graph_gbfly_s": round(total_bfly / (graph_ms * 1e-3) / 1e9, 1),
        "eager_ms": round(eager_ms, 4), "eager_gbfly_s": round(total_bfly / (eager_ms * 1e-3) / 1e9, 1),
        "per_transform_ms": round(per_ms, 4),
        "per_transform_gbfly_s": round(total_bfly / (per_ms * 1e-3) / 1e9, 1),
        "graph_vs_per_transform": round(per_ms / graph_ms, 3),
        "hbm_frac": round(total_bytes / (graph_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "fused_groups": info["fused_groups"], "two_pass_jobs": info["two_pass_jobs"],
        "streams": info["streams"], "twiddle_tables": ctx.twiddle_entries,
        "twiddle_bytes": ctx.twiddle_bytes, "tune_s": round(tune_s, 2),
    }
    batch.close()
    for b in bufs:
        b.free()
    return out


def bench_served(ntt_amd, requests: int = 20000, fractions=(0.25, 0.5, 0.75, 0.9, 1.0, 1.1),
                 headline: float = 0.75, seed: int = 0, budget_gb: float = 64.0, tune: bool = True):
    """The served stream: `requests` requests, logN ~ U{13..17}, L log-uniform
    on [4, 64], direction ~ Bernoulli(1/2), Montgomery, seeded operands
    generated on the device at admission, 64 GB operand budget.  First the
    saturation throughput (every request arriving at t = 0), then Poisson
    arrivals at the given fractions of it."""
    import numpy as np
    from ntt_amd import workloads

    ctx = ntt_amd.Context(0)
    if tune:
        # launches near saturation hold a few hundred limbs per (logN, direction)
        for n in range(13, 18):
            ctx.autotune(n, workloads.moduli_for(n, 48), count=8, reps=2)
    srv = ctx.server(hbm_budget_bytes=int(budget_gb * (1 << 30)), slots=4, max_jobs_per_launch=256)
    base = workloads.serving_stream(seed, requests, rate_per_s=1.0)
    burst = [workloads.Request(0.0, r.job, r.seed) for r in base]
    bfly = sum(r.job.butterflies for r in base)
    srv.run_trace(burst[:2000])                      # warm-up: plans, tables, kernels
    lat, wall = srv.run_trace(burst)
    sat_rps = requests / wall
    out = {"workload": f"{requests} requests, logN ~ U{{13..17}}, limbs log-uniform on [4, 64] (rounded), "
                       "direction ~ Bernoulli(1/2), Montgomery, Poisson arrivals, operands generated on the "
                       "device at admission, 64 GB operand budget, 4 launch slots",
           "saturation_req_s": round(sat_rps, 1), "saturation_gbfly_s": round(bfly / wall / 1e9, 1),
           "saturation_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
           "saturation_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3), "sweep": []}
    for f in fractions:
        rate = f * sat_rps
        reqs = workloads.serving_stream(seed + 1, requests, rate_per_s=rate)
        st0 = srv.stats()
        lat, wall = srv.run_trace(reqs)
        st1 = srv.stats()
        launches = st1["launches"] - st0["launches"]
        row = {"fraction": f, "rate_req_s": round(rate, 1), "throughput_req_s": round(requests / wall, 1),
               "gbfly_s": round(sum(r.job.butterflies for r in reqs) / wall / 1e9, 1),
               "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
               "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
               "jobs_per_launch": round(requests / max(1, launches), 2)}
        out["sweep"].append(row)
        log(f"served {row}")
        if f == headline:
            out.update({"p50_ms": row["p50_ms"], "p99_ms": row["p99_ms"], "at_fraction": f,
                        "throughput_req_s": row["throughput_req_s"], "gbfly_s": row["gbfly_s"]})
    st = srv.stats()
    out["arena_peak_bytes"] = st["arena_peak_bytes"]
    out["max_slots_busy"] = st["max_slots_busy"]
    srv.close()
    return out


def fused_kernel_name(logN: int, inverse: bool, eb: int, lazy: bool, logw: int, ot: bool) -> str:
    """Symbol of the fused kernel a plan launch runs, with its template
    arguments as the kernel trace prints them."""
    tf = lambda b: "true" if b else "false"
    if logN == 17:   # one limb per workgroup as two 2^16-word halves: width 6, OT inverse only
        return f"fused_split_kernel<{tf(inverse)}, {tf(lazy)}, {tf(ot and inverse)}>"
    return f"fused_kernel<{logN}, {tf(inverse)}, {tf(lazy)}, {eb}, false, {tf(ot)}, {logw}>"


def _timed(dev, stream, fn, reps):
    e0, e1 = dev.Event(), dev.Event()
    fn()
    stream.synchronize()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_ms(e1) / reps


def phase_kernel_name(logN, phase, inv, col, mode):
    """Symbol of the two-pass kernel launch_pass picks for one pass (R, S, W):
    the radix fixed at compile time for row passes with R 8-10 and S >= 4
    (not the table-converting mode 1), the whole geometry fixed for lazy
    logN-17 column passes with R = 8 or 9, W = 5 or 6 (at most 1024 threads)."""
    r, sm, w = phase
    if mode == 2 and sm >= 4 and col and logN == 17 and r in (8, 9) and w in (5, 6) and r - sm + w <= 10:
        rf, nf, wf = r, 17, w
    elif mode != 1 and sm >= 4 and not col and r in (8, 9, 10):
        rf, nf, wf = r, 0, 0
    else:
        rf, nf, wf = 0, 0, 0
    b = lambda v: "true" if v else "false"
    return f"phase_kernel<{sm}, {b(inv)}, {b(col)}, {mode}, {rf}, {nf}, {wf}>"


def _pass_kernel_ms(dev, stream, plan, pool, count, cfg, lazy, twconv=False, steps=5):
    """Per-pass kernel times of a two-pass launch (HIP events on the launch
    stream), keyed by kernel symbol."""
    mode = 1 if twconv else (2 if lazy else 0)
    c = cfg.astuple()

    def pk(ph, inv, col):
        return phase_kernel_name(plan.logN, c[ph], inv == "true", col == "true", mode)

    kern = {pk(0, "false", "true"): lambda: plan.run_pass(pool, count, 0, 0, stream),
            pk(1, "false", "false"): lambda: plan.run_pass(pool, count, 0, 1, stream),
            pk(2, "true", "false"): lambda: plan.run_pass(pool, count, 1, 0, stream),
            pk(3, "true", "true"): lambda: plan.run_pass(pool, count, 1, 1, stream)}
    names = list(kern)
    ev = [dev.Event() for _ in range(len(names) + 1)]
    per = {n: [] for n in names}
    for _ in range(steps):
        ev[0].record(stream)
        for i, n in enumerate(names):
            kern[n]()
            ev[i + 1].record(stream)
        ev[-1].synchronize()
        for i, n in enumerate(names):
            per[n].append(ev[i].elapsed_ms(ev[i + 1]))
    return {n: statistics.mean(v) for n, v in per.items()}


def _roofline(kernel_ms: dict, bytes_per_launch, workload_tag: str, traffic_path: str, units: str):
    """bytes_per_launch: one figure, or per kernel (a dict keyed like kernel_ms)
    of (bytes, operands per launch); the workload tag and units then take the
    dominant kernel's operand count ({ops}) and limb count ({limbs})."""
    dom = max(kernel_ms, key=kernel_ms.get)
    if isinstance(bytes_per_launch, dict):
        bytes_per_launch, ops = bytes_per_launch[dom]
        units = units.format(limbs=bytes_per_launch // (8 << 17))
        workload_tag = workload_tag.format(ops=ops)
    achieved = bytes_per_launch / (kernel_ms[dom] * 1e-3) / 1e9
    traffic = hbm_traffic(traffic_path, dom, workload_tag)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_ratio": round(traffic / bytes_per_launch, 3) if traffic else None,
            "kernel": dom, "kernel_ms": round(kernel_ms[dom], 4), "bytes_per_launch": bytes_per_launch,
            "units_per_launch": units}


def bench_c3(ntt_amd, dev, stream, traffic_path: str, tune: bool = True, count: int = 43, reps: int = 10):
    """The larger single-transform shape: N = 2^17, 48 limbs (the 48 largest
    primes q < 2^30 with q = 1 mod 2^18), Montgomery, a pool of `count`
    operands (43 x 24 MiB = 1.03 GiB, beyond the last-level cache), forward
    then inverse, roundtrip checked."""
    import numpy as np
    logN, limbs = 17, 48
    N = 1 << logN
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    ctx = ntt_amd.Context(0)
    cfg = None
    if tune:
        cfg = ntt_amd.TransformConfig.make(*ctx.autotune(logN, moduli, count=count, reps=2)["cfg"])
    plan = ctx.plan(logN, moduli, ntt_amd.MONTGOMERY)
    cfg = cfg or plan.config
    words = limbs * N
    pool = dev.DeviceBuffer(4 * words * count)
    qd = dev.DeviceBuffer.from_numpy(np.array([q for q, _ in moduli], np.uint32))
    dev.fill_uniform(pool, qd, limbs, logN, count, seed=0)
    dev.synchronize()
    before = pool.numpy(np.uint32, nbytes=4 * words)
    fwd = _timed(dev, stream, lambda: plan.forward(pool, count, stream), reps)
    inv = _timed(dev, stream, lambda: plan.inverse(pool, count, stream), reps)
    ok = bool(np.array_equal(before, pool.numpy(np.uint32, nbytes=4 * words)))
    bfly = 0.5 * N * logN * limbs * count
    engs = [{1: "two_pass", 2: "fused"}.get(plan.engine(count, d), "?") for d in (ntt_amd.FORWARD, ntt_amd.INVERSE)]
    eng = engs[0] if engs[0] == engs[1] else f"forward {engs[0]}, inverse {engs[1]}"
    # kernel times per direction: one kernel per plan launch on the fused
    # engine (the split kernels, one limb per workgroup), the two passes on
    # the two-pass engine
    tp = _pass_kernel_ms(dev, stream, plan, pool, count, cfg, all(q < (1 << 30) for q, _ in moduli))
    # a split launch whose last round would be nearly empty runs the whole
    # operands of one round fewer on the split kernels and hands the rest to
    # the two-pass kernels (plan_launch_fused); the split kernel alone is the
    # launch over those head operands
    cus = dev.device_info(0)["compute_units"]
    rounds = -(-limbs * count // cus)
    k = (rounds - 1) * cus // limbs if rounds > 1 else 0
    head = k if 0 < k < count and (count - k) * limbs <= cus // 2 else count
    kms, kbytes = {}, {}
    for d, t in ((0, fwd), (1, inv)):
        if engs[d] == "fused":
            eb, lz, lw, ot = plan.fused_variant_full(count, d)
            name = fused_kernel_name(logN, d == 1, eb, lz, lw, ot)
            if head < count:
                run = plan.forward if d == 0 else plan.inverse
                t = _timed(dev, stream, lambda: run(pool, head, stream), reps)
            kms[name] = t
            kbytes[name] = (8 * limbs * N * head, head)
        else:
            for kn, v in tp.items():
                if (kn.split(", ")[1] == "true") == (d == 1):
                    kms[kn] = v
                    kbytes[kn] = (8 * words * count, count)
    step = fwd + inv
    out = {"workload": f"N=2^17, 48 limbs (48 largest primes < 2^30, q = 1 mod 2^18), Montgomery, {count} operands "
                       f"({4 * words * count / 2**30:.2f} GiB), forward then inverse, seed 0",
           "engine": eng, "forward_ms": round(fwd, 4), "inverse_ms": round(inv, 4),
           "gbfly_s": round(2 * bfly / (step * 1e-3) / 1e9, 1),
           "hbm_frac": round(2 * 8 * words * count / (step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "kernel_ms": {k: round(v, 4) for k, v in kms.items()},
           "split_head_operands": head if "fused" in eng else None,
           "roofline": _roofline(kms, kbytes, "logN17x48_batch{ops}", traffic_path,
                                 "{limbs} limbs of 2^17 words, 8 B per word per transform"),
           "phase_config": cfg.astuple(), "roundtrip_ok": ok}
    plan.close()
    pool.free()
    return out


def bench_c1_plain(ntt_amd, dev, stream, traffic_path: str, logN: int = 16, limbs: int = 16, count: int = 256,
                   reps: int = 5):
    """The plain (normal-domain) variant at N = 2^16 x 16 limbs through the
    drop-in launchers (src/v03_template_config.cu:549-603): normal-domain
    twiddle tables converted on load.  Batched: the 256-operand pool as one
    launch of 4096 limbs with the tables of each limb (as the reference's
    launcher takes them); single transform: reference-style (restore copy +
    transform) and kernel only, forward and inverse."""
    import numpy as np
    N = 1 << logN
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    t = ntt_amd.make_tables(logN, moduli, ntt_amd.PLAIN)
    L = limbs * count
    words = limbs * N
    tw = dev.DeviceBuffer(4 * words * count)
    itw = dev.DeviceBuffer(4 * words * count)
    for k in range(count):
        tw.upload(t.twiddles, offset_bytes=4 * words * k)
        itw.upload(t.inv_twiddles, offset_bytes=4 * words * k)
    q = dev.DeviceBuffer.from_numpy(np.tile(t.moduli, count))
    dinv = dev.DeviceBuffer.from_numpy(np.tile(t.degree_inv, count))
    pool = dev.DeviceBuffer(4 * words * count)
    qd = dev.DeviceBuffer.from_numpy(np.array([m for m, _ in moduli], np.uint32))
    dev.fill_uniform(pool, qd, limbs, logN, count, seed=0)
    dev.synchronize()
    before = pool.numpy(np.uint32, nbytes=4 * words)
    f = lambda: ntt_amd.launch_ntt_plain(pool, tw, q, L, logN, None, stream)
    i = lambda: ntt_amd.launch_intt_plain(pool, itw, q, dinv, L, logN, None, stream)
    fwd = _timed(dev, stream, f, reps)
    inv = _timed(dev, stream, i, reps)
    ok = bool(np.array_equal(before, pool.numpy(np.uint32, nbytes=4 * words)))
    bfly = 0.5 * N * logN * L
    cfg = ntt_amd.default_launch_config(logN).astuple()
    # single transform, reference-style: restore copy + transform (src/v03_template_config.cu timing loop)
    one = dev.DeviceBuffer(4 * words)
    src = dev.DeviceBuffer(4 * words)
    src.copy_from(pool, 4 * words, stream)
    single = {}
    for name, fn in (("ntt", lambda: ntt_amd.launch_ntt_plain(one, tw, q, limbs, logN, None, stream)),
                     ("intt", lambda: ntt_amd.launch_intt_plain(one, itw, q, dinv, limbs, logN, None, stream))):
        single[f"single_{name}_ms_reference_style"] = round(
            _timed(dev, stream, lambda: (one.copy_from(src, stream=stream), fn()), 30), 4)
        single[f"single_{name}_ms_kernel_only"] = round(_timed(dev, stream, fn, 30), 4)
    # the same pool through a plain-domain plan (the library's own entry
    # point: cached Montgomery-form tables, fused kernels where they apply)
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(logN, moduli, ntt_amd.PLAIN)
    pf = _timed(dev, stream, lambda: plan.forward(pool, count, stream), reps)
    pi = _timed(dev, stream, lambda: plan.inverse(pool, count, stream), reps)
    pok = bool(np.array_equal(before, pool.numpy(np.uint32, nbytes=4 * words)))
    pv = [plan.fused_variant_full(count, d) for d in (ntt_amd.FORWARD, ntt_amd.INVERSE)]
    peng = {1: "two_pass", 2: "fused"}.get(plan.engine(count), "?")
    plan_rec = {"engine": peng, "forward_ms": round(pf, 4), "inverse_ms": round(pi, 4),
                "gbfly_s": round(2 * bfly / ((pf + pi) * 1e-3) / 1e9, 1),
                "hbm_frac": round(2 * 8 * words * count / ((pf + pi) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "kernels": {"forward": fused_kernel_name(logN, False, *pv[0]) if peng == "fused" else None,
                            "inverse": fused_kernel_name(logN, True, *pv[1]) if peng == "fused" else None},
                "roundtrip_ok": pok}
    plan.close()
    step = fwd + inv
    out = {"workload": f"plain variant, N=2^{logN}, {limbs} limbs, {count} operands as one drop-in launch of {L} limbs "
                       f"with per-limb normal-domain tables ({4 * words * count / 2**30:.2f} GiB operands), seed 0",
           "forward_ms": round(fwd, 4), "inverse_ms": round(inv, 4),
           "gbfly_s": round(2 * bfly / (step * 1e-3) / 1e9, 1),
           "hbm_frac": round(2 * 8 * words * count / (step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "phase_config": cfg,
           "kernels": {"forward": phase_kernel_name(logN, cfg[0], False, True, 1) + " + " +
                                  phase_kernel_name(logN, cfg[1], False, False, 1),
                       "inverse": phase_kernel_name(logN, cfg[2], True, False, 1) + " + " +
                                  phase_kernel_name(logN, cfg[3], True, True, 1)},
           "roofline": {"bound": "hbm", "achieved": round(8 * words * count / (max(fwd, inv) * 1e-3) / 1e9, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(8 * words * count / (max(fwd, inv) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "scope": "whole two-pass transform (both kernels) per direction",
                        "bytes_per_launch": 8 * words * count},
           **single, "roundtrip_ok": ok, "plan": plan_rec}
    for b in (tw, itw, pool, one, src):
        b.free()
    return out


def _cpu_model():
    try:
        return next(ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name"))
    except (OSError, StopIteration):
        return ""


def cpu_jobs(jobs, seconds: float, threads: int):
    """Oracle butterfly loops (src/v04_montgomery.cu:742-784 restated in C) over
    the limbs of `jobs` = [(logN, moduli, direction, seed)], in order, limbs
    of a job spread over `threads` host threads, until all are done or
    `seconds` of compute have passed.  Input and table generation are not
    timed.  Returns (limbs done, butterflies done, seconds, limbs total)."""
    import ctypes
    import concurrent.futures as cf

    from oracle import oracle as orc

    L = orc.lib()
    P = ctypes.POINTER(ctypes.c_uint32)
    tabs = {}
    total = sum(len(m) for _, m, _, _ in jobs)

    def one(w):
        logN, q, key, direction, row = w
        tw, itw, dinv = tabs[key]
        a = row.copy()
        if direction == 0:
            L.oracle_cpu_ntt_butterfly(a.ctypes.data_as(P), logN, q, tw.ctypes.data_as(P))
        else:
            L.oracle_cpu_intt_butterfly(a.ctypes.data_as(P), logN, q, itw.ctypes.data_as(P), dinv)
        return 0.5 * (1 << logN) * logN

    done, bfly, busy = 0, 0.0, 0.0
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        for logN, moduli, direction, seed in jobs:
            if busy >= seconds:
                break
            x = orc.fill_random_limbs(logN, moduli, seed)
            work = []
            for l, (q, g) in enumerate(moduli):
                key = (logN, q, g)
                if key not in tabs:
                    rc, rt = orc.make_modulus_runtime(logN, q, g)
                    tabs[key] = (orc.twiddle_table(logN, q, rt["psi"]), orc.twiddle_table(logN, q, rt["psi_inv"]),
                                 rt["degree_inv"])
                work.append((logN, q, key, direction, x[l]))
            t0 = time.perf_counter()
            bfly += sum(ex.map(one, work))
            busy += time.perf_counter() - t0
            done += len(work)
    return done, bfly, busy, total


def cpu_baselines_more(ntt_amd, seconds: float, threads: int):
    """CPU baselines of the other workloads, each on a bounded sample:
    the 17 x 48 shape, the uneven 256-job batch (its jobs in order) and a
    500-request prefix of the served stream; full-workload times are
    extrapolated from the sample rate."""
    from ntt_amd import workloads
    model = _cpu_model()
    out = {}
    m48 = [tuple(m) for m in ntt_amd.generate_moduli(17, 48)]
    c3_jobs = [(17, m48, d, k) for d in (0, 1) for k in range(43)]
    d, b, t, n = cpu_jobs(c3_jobs, seconds, threads)
    out["c3"] = {"value": round(b / t / 1e9, 3), "unit": "Gbutterfly/s", "cores": threads, "kind": "port",
                 "extrapolated_ms": round(2 * 43 * 48 * 0.5 * (1 << 17) * 17 / (b / t) * 1e3, 1),
                 "sample": f"{d} of {n} limbs of the 17 x 48 workload (43 operands forward, then 43 inverse, "
                           f"seed k per operand), {t:.1f} s, {threads} thread(s) on {model}"}
    wl = workloads.batched_uneven(0)
    jobs = [(j.logN, workloads.moduli_for(j.logN, j.limbs), j.direction, k) for k, j in enumerate(wl)]
    d, b, t, n = cpu_jobs(jobs, seconds, threads)
    total = sum(j.butterflies for j in wl)
    out["mixed"] = {"value": round(b / t / 1e9, 3), "unit": "Gbutterfly/s", "cores": threads, "kind": "port",
                    "extrapolated_ms": round(total / (b / t) * 1e3, 1),
                    "sample": f"first {d} of {n} limbs of the 256-job uneven workload (seed 0), {t:.1f} s, "
                              f"{threads} thread(s) on {model}; full-workload time extrapolated from the rate"}
    reqs = workloads.serving_stream(0, 500, rate_per_s=1.0)
    jobs = [(r.job.logN, workloads.moduli_for(r.job.logN, r.job.limbs), r.job.direction, k)
            for k, r in enumerate(reqs)]
    d, b, t, n = cpu_jobs(jobs, 4 * seconds, threads)
    rate = b / t
    req_s = 500 / (t * n / max(d, 1))
    out["served"] = {"value": round(rate / 1e9, 3), "unit": "Gbutterfly/s", "cores": threads, "kind": "port",
                     "requests_per_s": round(req_s, 1),
                     "extrapolated_20k_s": round(20000 / req_s, 1),
                     "sample": f"{d} of {n} limbs of a 500-request prefix of the served stream (seed 0), "
                               f"{t:.1f} s, {threads} thread(s) on {model}; 20k-request time extrapolated"}
    return out


def cpu_baseline(logN: int, moduli, seconds: float, threads: int = 1):
    """Oracle butterfly loops on host cores over a bounded sample: forward +
    inverse of whole limbs of the workload (seed 0 inputs), until `seconds`."""
    import ctypes
    import concurrent.futures as cf

    import numpy as np

    from oracle import oracle as orc

    L = orc.lib()
    tabs = []
    for q, g in moduli:
        rc, rt = orc.make_modulus_runtime(logN, q, g)
        tabs.append((q, orc.twiddle_table(logN, q, rt["psi"]), orc.twiddle_table(logN, q, rt["psi_inv"]),
                     rt["degree_inv"]))
    x = orc.fill_random_limbs(logN, moduli, 0)
    P = ctypes.POINTER(ctypes.c_uint32)

    def one_limb(i):
        q, tw, itw, dinv = tabs[i % len(tabs)]
        a = x[i % len(tabs)].copy()
        L.oracle_cpu_ntt_butterfly(a.ctypes.data_as(P), logN, q, tw.ctypes.data_as(P))
        L.oracle_cpu_intt_butterfly(a.ctypes.data_as(P), logN, q, itw.ctypes.data_as(P), dinv)
        return a

    limbs_done = 0
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        while time.perf_counter() - t0 < seconds:
            list(ex.map(one_limb, range(limbs_done, limbs_done + threads)))
            limbs_done += threads
    dt = time.perf_counter() - t0
    bfly = 2 * 0.5 * (1 << logN) * logN * limbs_done
    model = ""
    try:
        model = next(ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": bfly / dt / 1e9, "unit": "Gbutterfly/s", "cores": threads, "kind": "port",
            "sample": f"{limbs_done} limbs of N=2^{logN} (forward + inverse each, seed-0 inputs, "
                      f"tables prebuilt), {dt:.1f} s, {threads} thread(s) on {model}"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library runs on a single MI355X (asked for --gpus {args.gpus}). "
              "Limbs and batch operands shard across GPUs with no exchange.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import ntt_amd
    from ntt_amd import device as dev

    if not dev.is_available():
        log("no GPU visible")
        sys.exit(1)
    info = dev.device_info(0)
    log(f"device {info['name']} {info['arch']} CUs {info['compute_units']}")

    logN, limbs, B = args.logN, args.limbs, args.batch
    N = 1 << logN
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    ctx = ntt_amd.Context(0)
    tuning = {}
    if not args.no_tune:
        t0 = time.perf_counter()
        tuning["batch"] = ctx.autotune(logN, moduli, count=B, reps=3)
        tuning["single"] = ctx.autotune(logN, moduli, count=1, reps=5)
        tuning["seconds"] = round(time.perf_counter() - t0, 2)
        log(f"autotune: {json.dumps(tuning)}")
    fixed_cfg = None
    if args.phase_config:
        fixed_cfg = ntt_amd.TransformConfig.make(*[tuple(int(v) for v in ph.split(",")) for ph in
                                                   args.phase_config.split(";")])
    plan = ctx.plan(logN, moduli, ntt_amd.MONTGOMERY, fixed_cfg)
    if args.engine != "auto":
        plan.set_engine({"fused": ntt_amd.ENGINE_FUSED, "two_pass": ntt_amd.ENGINE_TWO_PASS}[args.engine])
    if args.fused_shape:
        plan.set_fused_shape(args.fused_shape)
    engine_used = {1: "two_pass", 2: "fused"}.get(plan.engine(B), "?")
    cfg = plan.config if fixed_cfg or not tuning else ntt_amd.TransformConfig.make(*tuning["batch"]["cfg"])
    words_per_op = limbs * N
    pool = dev.DeviceBuffer(4 * words_per_op * B)
    qdev = dev.DeviceBuffer.from_numpy(np.array([m[0] for m in moduli], np.uint32))
    dev.fill_uniform(pool, qdev, limbs, logN, B, seed=0)
    stream = dev.Stream()
    dev.synchronize()
    probe = [0, B - 1]
    before = {k: pool.numpy(np.uint32, offset_bytes=4 * words_per_op * k, nbytes=4 * words_per_op) for k in probe}

    if args.copy_calibration:
        scratch = dev.DeviceBuffer(4 * words_per_op * B)
        dev.stream_copy(scratch, pool, 4 * words_per_op * B, stream)
        stream.synchronize()
        log(f"copy calibration: {4 * words_per_op * B} bytes read and written")
        del scratch

    bfly_per_transform = ntt_amd.butterflies(logN, limbs)
    bytes_per_transform = ntt_amd.algorithmic_bytes(logN, limbs)

    def step():
        plan.forward(pool, B, stream)
        plan.inverse(pool, B, stream)

    for _ in range(args.warmup):
        step()
    stream.synchronize()
    e0, e1 = dev.Event(), dev.Event()
    dev.synchronize()
    e0.record(stream)
    t_wall = time.perf_counter()
    for _ in range(args.steps):
        step()
    e1.record(stream)
    e1.synchronize()
    t_wall = time.perf_counter() - t_wall
    dev.synchronize()
    total_ms = e0.elapsed_ms(e1)
    ms_per_step = total_ms / args.steps
    value = 2 * B * bfly_per_transform / (ms_per_step * 1e-3) / 1e9
    hbm_frac_transform = 2 * B * bytes_per_transform / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS
    log(f"batched: {ms_per_step:.3f} ms/step ({B} x fwd+inv), {value:.1f} Gbfly/s, wall {t_wall:.3f} s")

    ok = all(np.array_equal(before[k], pool.numpy(np.uint32, offset_bytes=4 * words_per_op * k,
                                                   nbytes=4 * words_per_op)) for k in probe)
    if not ok:
        log("ROUNDTRIP MISMATCH after the timed region")

    # ---- kernel durations of the engine in use (HIP events on the launch stream) ----
    ebf, lzf, lwf, otf = plan.fused_variant_full(B, ntt_amd.FORWARD)
    ebi, lzi, lwi, oti = plan.fused_variant_full(B, ntt_amd.INVERSE)
    lazy = all(q < (1 << 30) for q, _ in moduli)   # two-pass kernels then use lazy reduction
    if engine_used == "fused":
        kern = {fused_kernel_name(logN, False, ebf, lzf, lwf, otf): lambda: plan.forward(pool, B, stream),
                fused_kernel_name(logN, True, ebi, lzi, lwi, oti): lambda: plan.inverse(pool, B, stream)}
    else:
        mode = 2 if lazy else 0
        c = cfg.astuple()
        def pk(ph, inv, col):
            return phase_kernel_name(logN, c[ph], inv == "true", col == "true", mode)
        kern = {pk(0, "false", "true"): lambda: plan.run_pass(pool, B, 0, 0, stream),
                pk(1, "false", "false"): lambda: plan.run_pass(pool, B, 0, 1, stream),
                pk(2, "true", "false"): lambda: plan.run_pass(pool, B, 1, 0, stream),
                pk(3, "true", "true"): lambda: plan.run_pass(pool, B, 1, 1, stream)}
    names = list(kern)
    ev = [dev.Event() for _ in range(len(names) + 1)]
    per = {n: [] for n in names}
    for _ in range(max(5, args.steps // 2)):
        ev[0].record(stream)
        for i, n in enumerate(names):
            kern[n]()
            ev[i + 1].record(stream)
        ev[-1].synchronize()
        for i, n in enumerate(names):
            per[n].append(ev[i].elapsed_ms(ev[i + 1]))
    kernel_ms = {n: statistics.mean(v) for n, v in per.items()}
    dom = max(kernel_ms, key=kernel_ms.get)
    bytes_per_launch = B * bytes_per_transform
    achieved = bytes_per_launch / (kernel_ms[dom] * 1e-3) / 1e9
    workload_tag = f"logN{logN}x{limbs}_batch{B}"
    traffic = hbm_traffic(args.traffic, dom, workload_tag)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "kernel_ms": round(kernel_ms[dom], 4), "bytes_per_launch": bytes_per_launch,
                "units_per_launch": f"{B * limbs} limbs of 2^{logN} words, 8 B per word (read once, written once)"}
    log("kernel ms: " + ", ".join(f"{n} {v:.4f}" for n, v in kernel_ms.items()))

    # the binding compute roof: the same butterfly code on register-resident operands
    # INV: the first template argument of the split kernels, the second of the others
    targs = dom.split("<", 1)[1].split(",")
    inverse_dom = targs[0 if dom.startswith("fused_split_kernel") else 1].strip().rstrip(">") == "true"
    lazy_dom = (lazy and engine_used != "fused") or (engine_used == "fused" and (lzi if inverse_dom else lzf))
    variant = (dev.BUTTERFLY_GS if inverse_dom else dev.BUTTERFLY_CT) + (1 if lazy_dom else 0)
    roof_bfly = dev.butterfly_roof(variant, stream)
    bfly_dom = B * bfly_per_transform / (kernel_ms[dom] * 1e-3)
    compute_roof = {"bound": "valu", "unit": "Gbutterfly/s", "achieved": round(bfly_dom / 1e9, 1),
                    "peak": round(roof_bfly / 1e9, 1), "frac": round(bfly_dom / roof_bfly, 4),
                    "butterfly": ("GS" if inverse_dom else "CT") + (" lazy" if variant & 1 else " canonical"),
                    "kernel": dom}
    log(f"compute roof: {compute_roof}")

    extras = {}
    if not args.no_extras:
        # plain per-transform launch: one plan call per operand
        ev_a, ev_b = dev.Event(), dev.Event()
        reps = 3
        ev_a.record(stream)
        for _ in range(reps):
            for k in range(B):
                off = pool.ptr + 4 * words_per_op * k
                plan.forward(off, 1, stream)
                plan.inverse(off, 1, stream)
        ev_b.record(stream)
        ev_b.synchronize()
        ms = ev_a.elapsed_ms(ev_b) / reps
        extras["per_transform_launch_gbfly_s"] = round(2 * B * bfly_per_transform / (ms * 1e-3) / 1e9, 1)
        # single transform, reference-style (restore D2D copy + transform, src/v04_montgomery.cu:980-988)
        one = dev.DeviceBuffer(4 * words_per_op)
        src = dev.DeviceBuffer(4 * words_per_op)
        src.copy_from(pool, 4 * words_per_op, stream)
        for _ in range(10):
            one.copy_from(src, stream=stream)
            plan.forward(one, 1, stream)
        iters = 30
        ev_a.record(stream)
        for _ in range(iters):
            one.copy_from(src, stream=stream)
            plan.forward(one, 1, stream)
        ev_b.record(stream)
        ev_b.synchronize()
        ref_ms = ev_a.elapsed_ms(ev_b) / iters
        ev_a.record(stream)
        for _ in range(iters):
            plan.forward(one, 1, stream)
        ev_b.record(stream)
        ev_b.synchronize()
        k_ms = ev_a.elapsed_ms(ev_b) / iters
        ev_a.record(stream)
        for _ in range(iters):
            plan.inverse(one, 1, stream)
        ev_b.record(stream)
        ev_b.synchronize()
        ki_ms = ev_a.elapsed_ms(ev_b) / iters
        # reference-style INTT: restore copy + inverse (src/v04_montgomery.cu:1037-1062)
        ev_a.record(stream)
        for _ in range(iters):
            one.copy_from(src, stream=stream)
            plan.inverse(one, 1, stream)
        ev_b.record(stream)
        ev_b.synchronize()
        refi_ms = ev_a.elapsed_ms(ev_b) / iters
        extras.update({
            "single_ntt_ms_reference_style": round(ref_ms, 4),
            "single_ntt_gbfly_s_reference_style": round(bfly_per_transform / (ref_ms * 1e-3) / 1e9, 1),
            "single_ntt_ms_kernel_only": round(k_ms, 4),
            "single_intt_ms_kernel_only": round(ki_ms, 4),
            "single_intt_ms_reference_style": round(refi_ms, 4),
            "vs_rtx4090_v04_ntt_reference_style": round(bfly_per_transform / (ref_ms * 1e-3) / 1e9
                                                        / RTX4090_V04_NTT_GBFLY, 3),
        })
        # fused mod-down form over the whole batch: dst = (sub - NTT(src)) * f, in place
        # on the pool; algorithmic bytes 12 per word (src and sub read once, dst written once)
        sub = dev.DeviceBuffer(4 * words_per_op * B)
        dev.fill_uniform(sub, qdev, limbs, logN, B, seed=1)
        plan.set_mod_down_factors([3 + l for l in range(limbs)])
        md = {}
        for name, fn in (("forward", lambda: plan.forward(pool, B, stream)),
                         ("mod_down", lambda: plan.forward_mod_down(pool, pool, sub, B, stream))):
            fn()
            ev_a.record(stream)
            for _ in range(5):
                fn()
            ev_b.record(stream)
            ev_b.synchronize()
            md[name] = ev_a.elapsed_ms(ev_b) / 5
        plan.set_mod_down_factors(None)
        del sub
        extras["mod_down"] = {
            "ms": round(md["mod_down"], 4), "forward_ms": round(md["forward"], 4),
            "gbfly_s": round(B * bfly_per_transform / (md["mod_down"] * 1e-3) / 1e9, 1),
            "algorithmic_gb_s": round(12 * B * words_per_op / (md["mod_down"] * 1e-3) / 1e9, 1),
            "cost_vs_forward": round(md["mod_down"] / md["forward"], 3)}
        # fused engine with on-the-fly twiddles in the low-bit round, against the table
        if engine_used == "fused":
            ot = {}
            for mode, tag in ((ntt_amd.TWIDDLE_TABLE, "table"), (ntt_amd.TWIDDLE_OT, "ot")):
                plan.set_twiddle_mode(mode)
                for name, fn in (("forward", lambda: plan.forward(pool, B, stream)),
                                 ("inverse", lambda: plan.inverse(pool, B, stream))):
                    fn()
                    ev_a.record(stream)
                    for _ in range(5):
                        fn()
                    ev_b.record(stream)
                    ev_b.synchronize()
                    ot[f"{tag}_{name}_ms"] = round(ev_a.elapsed_ms(ev_b) / 5, 4)
            plan.set_twiddle_mode(ntt_amd.TWIDDLE_AUTO)
            extras["twiddle_ot"] = ot
        extras["word64"] = bench_word64(ntt_amd, dev, stream, logN, limbs)
        log("extras: " + json.dumps(extras))

    c3 = c1 = None
    if not args.no_extras:
        c3 = bench_c3(ntt_amd, dev, stream, args.traffic, tune=not args.no_tune)
        log("c3: " + json.dumps(c3))
        c1 = bench_c1_plain(ntt_amd, dev, stream, args.traffic)
        log("c1_plain: " + json.dumps(c1))

    mixed = None
    if not args.no_mixed:
        mixed = bench_mixed(ntt_amd, dev, stream, tune=not args.no_tune)
        log("mixed: " + json.dumps(mixed))

    served = None
    if not args.no_served:
        served = bench_served(ntt_amd, requests=args.served_requests, tune=not args.no_tune)
        log("served: " + json.dumps({k: v for k, v in served.items() if k != "sweep"}))

    cpu = cpu1 = cpu_more = None
    if not args.no_cpu:
        # limbs are independent: spread over the host threads this job may use
        # (at most 16, the CPU share of one GPU on the box)
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        cpu = cpu_baseline(logN, moduli, args.cpu_seconds, threads=threads)
        log(f"cpu baseline ({threads} threads): {cpu}")
        cpu1 = cpu_baseline(logN, moduli, args.cpu_seconds / 2, threads=1)
        log(f"cpu baseline (1 thread): {cpu1}")
        cpu_more = cpu_baselines_more(ntt_amd, args.cpu_seconds / 2, threads)
        log(f"cpu baselines (other workloads): {cpu_more}")

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Gbutterfly/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"Montgomery NTT then INTT, N=2^{logN}, {limbs} limbs (largest primes < 2^30, "
                        f"q = 1 mod 2^{logN + 1}), batch of {B} operands "
                        f"({4 * words_per_op * B / 2**30:.2f} GiB, HBM-resident), batched launch",
            "logN": logN, "limbs": limbs, "batch": B,
            "engine": engine_used,
            "fused_variant": {"forward": {"eb": ebf, "lazy": lzf, "log2_words": lwf, "on_the_fly_twiddles": otf},
                              "inverse": {"eb": ebi, "lazy": lzi, "log2_words": lwi, "on_the_fly_twiddles": oti}}
                             if engine_used == "fused" else None,
            "phase_config": cfg.astuple(),
            "chunk_operands": tuning["batch"]["chunk"] if tuning else 0,
            "parallelism": "single GPU",
        },
        "hbm_frac_transform": round(hbm_frac_transform, 4),
        "served_p50_ms": served["p50_ms"] if served else None,
        "served_p99_ms": served["p99_ms"] if served else None,
        "roundtrip_ok": ok,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_1core": cpu1,
        "cpu_baseline_other": cpu_more,
        "c3": c3,
        "c1_plain": c1,
        "mixed": mixed,
        "served": served,
        "compute_roof": compute_roof,
        "kernel_ms": {k: round(v, 4) for k, v in kernel_ms.items()},
        "autotune": tuning or None,
        **extras,
    }
    print(json.dumps(out), flush=True)
    if not ok:
        sys.exit(3)


if __name__ == "__main__":
    main()
