"""Benchmark of the MI355X branch evaluator.

Headline (the JSON line's `value`, Gray-code branches/s at T=36): the
plain-GPU-at-scale workload, one step = the 16 amplitudes <0|C|y> of the n=40
circuit "H on all qubits; 36 T + 300 CX/CZ + 100 S interleaved (SplitMix64
seed 5); H on all qubits" (t = 30 terms after T merging, branch dimension
d = 30: 16 x 2^30 branches), all 16 evaluated in one launch of the per-branch
walk (`k_z8_walk`): every one of the 2^30 branches of each amplitude is
visited on the GPU, one FP64 FMA per branch (the sign flipped once per 16
branches), with integer Z8 phases and FP64 accumulation.  Engine images are
resident in HBM before
timing starts; the host build of the engine (compile, normal form, branch
tables) is timed apart and reported in `config`.  Reported beside it:
  * the default served form (segment walk: per-engine FP64 Walsh-Hadamard
    tables over 14-16 inner coordinates, outer Gray walk in double-float
    lanes) on the same 16 amplitudes: amplitudes/s, its kernel's roofline on
    the outer points it walks, and a saturating launch (4096 amplitudes),
  * the plain FP64 form (one query per launch, FP64 port of the reference's
    Gray walk) on the same 16 amplitudes, and the per-amplitude error of the
    other forms against it,
  * the CPU oracle (a C++ restatement of the reference's Rayon engine, OpenMP
    over 2^14-branch Gray chunks) on the host cores, on a bounded sample, with
    the same 2^d branches per amplitude (Cached), and with the oracle's own
    best strategy (Auto: pair decomposition into 2^p smaller engines, built
    and evaluated in parallel) for an amplitudes/s comparison; engine build
    times are reported apart from evaluation,
  * a VALU issue-rate probe (pq_valu_probe) run first on the same device:
    the roofline peaks are its measured lanes per clock per CU x CUs x the
    2.4 GHz peak engine clock.
Further sections (each bounded; --sections picks them):
  precision  n=48, 40 Rz(theta) (seed 4, d = 30): 1024 amplitudes in the
             segment (z8) and branch-walk forms against the FP64 GPU form, the
             FP64 form against the CPU oracle on 16 of them; per-amplitude errors, with
             |a| < 2^-20 reported separately (--precision-out writes them all)
  batch      256 circuits (n ~ U[20,56], T ~ U[8,32]) x 16 (x, y) pairs: host
             builds, segment-split autotuner, one persistent launch, the same
             batch captured in a hipGraph and replayed
  serve      Poisson arrivals (200 req/s, 60 s): 50% new circuits of the
             batch generator, 50% repeats; 70% amplitude, 20% probability,
             10% 4-qubit marginal; 64 GB HBM budget; p50/p99 latency,
             throughput, and results re-evaluated alone for comparison
  serve_sweep the serve mix at Poisson rates 200, 400, ... req/s (each up to
             --sweep-seconds) until p99 bends: p50/p99, throughput, largest
             batch and overlapped batches per rate

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-plain]
                       [--sections headline,saturated,precision,batch,serve,serve_sweep] [--serve-seconds S]
Prints one JSON line.
"""
import argparse
import json
import os
import random
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "polyq-v1.0_amd"))

METRIC = "Gray-code branches/sec at T=36; amplitudes/sec + p50/p99 latency under load"
SEED = 5
SMALL = 2.0 ** -20


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    return O


def _err_summary(got, ref):
    import numpy as np
    got = np.asarray(got)
    ref = np.asarray(ref)
    ab = np.abs(got - ref)
    mag = np.abs(ref)
    small = mag < SMALL
    big = ~small
    rel = np.where(mag > 0, ab / np.where(mag > 0, mag, 1), ab)
    out = {"n": int(len(ref)), "max_abs": float(ab.max()) if len(ab) else 0.0}
    if big.any():
        out["large"] = {"n": int(big.sum()), "max_rel": float(rel[big].max()), "median_rel": float(np.median(rel[big]))}
    if small.any():
        out["below_2^-20"] = {"n": int(small.sum()), "max_abs": float(ab[small].max()),
                              "max_rel": float(rel[small].max()), "median_rel": float(np.median(rel[small])),
                              "median_abs_amp": float(np.median(mag[small]))}
    return out


SEG_OPS_PER_POINT = 30.0
# Algorithmic lane-ops per outer point of k_seg_walk: the integer walk step
# (phase update with the cross-term parity, sign-mask flip, table index: 6)
# and the double-float arithmetic (walk magnitude times the complex table
# entry: 2 x 4; compensated complex accumulation: 2 x 8).  The lane seed, the
# lane factor and the reductions are overhead, not counted.
# HBM bytes per walk launch, keyed by launch shape, table bits, lane walk
# length and waves per workgroup (FETCH_SIZE x 2 + WRITE_SIZE, copy-calibrated;
# profiles/k_seg_walk_config2.md): the table once per XCD L2 and the item
# partials.
SEG_TRAFFIC = {("headline", 16, 2, 4): 1.84e6, ("saturated", 16, 6, 4): 5.37e6}
# HBM bytes per branch-walk launch on the headline step, keyed by the mid
# coordinates M a lane walks (FETCH_SIZE x 2 + WRITE_SIZE, copy-calibrated;
# profiles/k_z8_walk_config2.md): the item partials (4 doubles per 256-lane
# item, 2^(12 - M) items per amplitude with 10 inner slots) and the tables
# once per XCD.
BRANCH_TRAFFIC = {3: 0.93e6, 4: 0.67e6, 5: 0.54e6, 6: 0.47e6, 7: 0.44e6}
PEAK_CLOCK_HZ = 2.4e9       # MI355X peak engine clock


def _valu_probe(pq):
    """Measured VALU issue rates (lanes per clock per CU) of this device."""
    try:
        return pq.valu_probe(blocks_per_cu=8, iters=16384)
    except Exception as e:      # the probe is a measurement aid; the bench still runs
        return {"error": str(e)}


def _peak(probe, op, fallback_lanes):
    """Peak lane-ops/s for an instruction mix: the probe's lanes per clock per
    CU for that mix x CUs x the peak engine clock (fallback_lanes per clock
    when the probe did not run)."""
    ops = probe.get("ops", {})
    cus = probe.get("cus", 256)
    lanes = ops[op]["lanes_per_clk_per_cu"] if op in ops else fallback_lanes
    return cus * lanes * PEAK_CLOCK_HZ, lanes, op if op in ops else "assumed"


def _seg_roofline(probe, points, kernel_ms, extra=None, traffic=None, kernel="k_seg_walk"):
    # the segment walk issues int32 and FP32 ops: price it at the faster of
    # the two measured rates
    p_int, l_int, _ = _peak(probe, "v_xor_b32", 64)
    p_f32, l_f32, _ = _peak(probe, "v_fma_f32", 64)
    peak = max(p_int, p_f32)
    achieved = points * SEG_OPS_PER_POINT / (kernel_ms * 1e-3)
    r = {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "Tlane-op/s", "frac": achieved / peak,
         "traffic": traffic, "kernel": kernel, "kernel_avg_ms": kernel_ms, "ops_per_outer_point": SEG_OPS_PER_POINT,
         "outer_points_per_launch": points, "peak_lanes_per_clk_per_cu": max(l_int, l_f32)}
    if extra:
        r.update(extra)
    return r


def _s8d_ops(st):
    """Lane-ops per branch of the reference's per-branch algorithm (popcount
    phase8 on the packed shift row, term update, combine): 4 W_g + 4 W_o +
    4 W_t + 14 with W = 32-bit words of planes, lines and terms."""
    w = lambda n: (n + 31) // 32
    return 4 * w(st[1]) + 4 * w(st[2]) + 4 * w(st[4]) + 14


def _apply_tuning(pq, text):
    """Apply "offset,max_mid[,waves]" to the segment walk; returns (offset, max_mid)."""
    v = [int(t) for t in text.split(",")]
    pq.set_walk_tuning(v[0], v[1])
    if len(v) > 2:
        pq.set_walk_shape(v[2])
    return v[0], v[1]


SETTLE_S = 0.06


def _settle(fn):
    """Run fn back to back for SETTLE_S seconds (untimed; the steps it took)."""
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < SETTLE_S:
        fn()
        n += 1
    return n


def _timed_kernel_ms(pq, np, fn, reps):
    ms = []
    with pq.timed():
        for _ in range(reps):
            out = fn()
            ms.append(pq.last_timing()["kernel_ms"])
    return float(np.mean(ms)), out


def headline(args, pq, W, np, probe):
    dev = pq.device_info()
    c, xs, ys = W.config2(SEED)
    tb = time.perf_counter()
    sim = pq.Simulator(c, strategy="cached", form="branch_walk")
    build_s = time.perf_counter() - tb          # compile + normal form + elimination + walk images + segment tables
    st = sim.stats()
    info = sim.z8_info()
    d = st[5]
    nq = len(ys)
    branches_per_step = nq * (1 << d)
    # the basis indices as uint64 arrays, the form the C ABI takes them in
    xs = np.asarray(xs, dtype=np.uint64)
    ys = np.asarray(ys, dtype=np.uint64)
    tu = time.perf_counter()
    sim.amplitudes(xs[:1], ys[:1])
    upload_s = time.perf_counter() - tu          # first device use: engine images to HBM (one amplitude included)

    # lane walk length of the branch walk, autotuned for this batch shape
    if args.headline_tuning:
        tuned = tuple(int(v) for v in args.headline_tuning.split(","))
        pq.set_walk_tuning(*tuned, form="branch_walk")
    else:
        tuned = pq.autotune([sim] * nq, xs, ys, form="branch_walk", repeats=8, rounds=4, cache=False)[:2]
    # The GPU clock rises over the first 20-40 ms of load after any pause
    # (0.60 -> 0.53 ms per step here), and the tuning trials end in host-side
    # work: the chosen split first runs for SETTLE_S, then the W warm-up steps
    settle = _settle(lambda: sim.amplitudes(xs, ys))
    for _ in range(args.warmup):
        sim.amplitudes(xs, ys)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        got = sim.amplitudes(xs, ys)     # synchronous: returns after the device results are on the host
    elapsed = time.perf_counter() - t0
    # Kernel durations: the same steps again with HIP timing events on the
    # launch stream (the events are queue barriers, so the throughput loop
    # above runs without them).
    kernel_avg_ms, again = _timed_kernel_ms(pq, np, lambda: sim.amplitudes(xs, ys), min(args.steps, 50))
    timing = pq.last_timing()
    assert np.array_equal(again, got)
    ops = info["ops_per_branch"]
    mid_bits = pq.branch_walk_mid_bits(d, tuned[0], tuned[1], info["inner_slots"])
    # without inner pivots the inner step is one FP64 FMA per branch, with the
    # table entry broadcast by DPP (the sign XOR comes once per 16 branches):
    # priced at the faster of the measured v_fma_f64 and v_fmac_f64_dpp rates;
    # with pivots at the xor/FMA mix
    if ops == 1:
        peak, lanes, peak_op = max(_peak(probe, "v_fma_f64", 64), _peak(probe, "v_fmac_f64_dpp row_newbcast", 64))
    else:
        peak, lanes, peak_op = _peak(probe, "v_xor_b32+v_fma_f64", 64)
    achieved = branches_per_step * ops / (kernel_avg_ms * 1e-3)
    s8d = _s8d_ops(st)
    kern = "k_z8_walk<%s, %s>" % ("true" if info["inner_pivots"] else "false", "true" if info["imag_slot0"] else "false")
    res = {
        "sim": sim, "c": c, "xs": xs, "ys": ys, "st": st, "info": info, "d": d, "nq": nq, "got": got,
        "branches_per_step": branches_per_step, "elapsed": elapsed, "kernel_avg_ms": kernel_avg_ms,
        "timing": timing, "dev": dev, "tuning": list(tuned), "build_s": build_s, "upload_s": upload_s, "settle_steps": settle,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "Tlane-op/s",
                     "frac": achieved / peak, "traffic": BRANCH_TRAFFIC.get(mid_bits), "kernel": kern, "mid_bits": mid_bits,
                     "kernel_avg_ms": kernel_avg_ms, "ops_per_branch": ops, "branches_per_launch": branches_per_step,
                     "peak_mix": peak_op, "peak_lanes_per_clk_per_cu": lanes,
                     "kernel_branches_per_s": branches_per_step / (kernel_avg_ms * 1e-3),
                     # the same branch rate priced at the reference algorithm's per-branch work
                     "ref_algorithm_ops_per_branch": s8d,
                     "ref_algorithm_equiv_frac": branches_per_step * s8d / (kernel_avg_ms * 1e-3) / peak},
    }

    # The default served form (segment walk) on the same 16 amplitudes.
    seg = sim.seg_info()
    points_per_step = nq * (1 << seg["outer"])
    if args.segment_tuning:
        stune = _apply_tuning(pq, args.segment_tuning)
    else:
        stune = pq.autotune([sim] * nq, xs, ys, repeats=7, rounds=5, cache=False)[:2]
    swaves = pq.get_walk_shape()
    _settle(lambda: sim.amplitudes(xs, ys, form="z8"))
    for _ in range(args.warmup):
        sim.amplitudes(xs, ys, form="z8")
    ts = time.perf_counter()
    for _ in range(args.steps):
        zgot = sim.amplitudes(xs, ys, form="z8")
    z_elapsed = time.perf_counter() - ts
    z_ms, zagain = _timed_kernel_ms(pq, np, lambda: sim.amplitudes(xs, ys, form="z8"), min(args.steps, 50))
    assert np.array_equal(zagain, zgot)
    seg = sim.seg_info()            # with the tuned lane walk length
    zkern = "k_seg_walk_args<%s, %d>" % ("true" if seg["mid_bits"] <= 4 else "false", swaves)
    res["segment"] = {
        "form": "z8 (default: per-engine FP64 Walsh-Hadamard tables, outer Gray walk, double-float lanes)",
        "amplitudes_per_s": nq * args.steps / z_elapsed, "ms_per_step": 1e3 * z_elapsed / args.steps,
        "walk_tuning": list(stune), "workgroup_waves": swaves, "segment": seg, "got": zgot,
        "roofline": _seg_roofline(probe, points_per_step, z_ms, kernel=zkern,
                                  traffic=SEG_TRAFFIC.get(("headline", seg["inner"], seg["mid_bits"], swaves))),
    }

    if not args.no_plain:
        psim = pq.Simulator(c, strategy="cached", form="plain_fp64")
        psim.amplitudes(xs[:1], ys[:1])
        tp0 = time.perf_counter()
        pa = psim.amplitudes(xs, ys)
        tp1 = time.perf_counter()
        p_int, l_int, op_int = _peak(probe, "v_xor_b32", 64)
        s8d = _s8d_ops(st)
        res["plain"] = {
            "amplitudes_per_s": nq / (tp1 - tp0),
            "branches_per_s": branches_per_step / (tp1 - tp0),
            "branch_walk_vs_fp64": _err_summary(got, pa),
            "z8_vs_fp64": _err_summary(zgot, pa),
            # k_plain priced at the reference algorithm's per-branch integer
            # work (popcount phase8, term update, combine) against the int32
            # VALU peak; its FP64 complex accumulation per branch comes on top
            "roofline_ref_ops": {"kernel": "k_plain", "ops_per_branch": s8d, "peak": p_int / 1e12, "unit": "Tlane-op/s",
                                 "peak_op": op_int, "achieved": branches_per_step * s8d / (tp1 - tp0) / 1e12,
                                 "frac": branches_per_step * s8d / (tp1 - tp0) / p_int},
        }

    if not args.no_cpu:
        O = _oracle()
        # a one-GPU job on the GPU host gets a share of 16 CPU threads, whatever
        # os.cpu_count() reports there (the count is recorded as host_nproc)
        threads = max(1, min(16, os.cpu_count() or 1))
        oc = O.OCircuit.from_circuit(c)
        osim = O.OSim(oc, "cached", threads=threads)
        nhi = 1 << (d - min(d, 14))
        done_br, done_sec, amps = 0, 0.0, 0
        cpu_err = 0.0
        for i in range(nq):
            a, sec, br = osim.chunk_range_timed(int(xs[i]), int(ys[i]), 0, nhi, threads)
            done_br += br; done_sec += sec; amps += 1
            cpu_err = max(cpu_err, abs(a - got[i]) / max(abs(a), 1e-300))
            if done_sec > args.cpu_seconds:
                break
        res["cpu"] = {
            "value": done_br / done_sec, "unit": "branches/s", "cores": threads, "kind": "port",
            "sample": "%d full amplitudes (%d x 2^%d branches, 2^14-branch Gray chunks, OpenMP dynamic) of the same workload"
                      % (amps, amps, d),
            "amplitudes_per_s": amps / done_sec, "max_rel_err_branch_walk_vs_cpu": cpu_err,
            "engine_new_s": osim.build_seconds, "host_nproc": os.cpu_count(), "host_cpu": _cpu_model(),
        }
        # the oracle's own best strategy: Auto (pair decomposition by the
        # cost model, branch engines built and evaluated in parallel); eval
        # timed alone, median of 5 passes over the 16 amplitudes after one
        auto = O.OSim(oc, "auto", threads=threads)
        ast = auto.stats()
        auto.amplitudes_timed(xs[:1], ys[:1], threads)
        secs, auto_err = [], 0.0
        for _ in range(5):
            aa, sec = auto.amplitudes_timed(xs, ys, threads)
            secs.append(sec)
            auto_err = max(auto_err, float(np.max(np.abs(aa - zgot) / np.maximum(np.abs(aa), 1e-300))))
        res["cpu_auto"] = {
            "strategy": "auto", "pairs": ast[6], "engines": 1 << ast[6], "engine_branch_dim": ast[5],
            "branches_per_amplitude": (1 << ast[6]) << ast[5], "engine_new_s": auto.build_seconds,
            "amplitudes_per_s": nq / float(np.median(secs)), "cores": threads,
            "sample": "5 passes over the %d amplitudes of one step after one warm-up amplitude; median pass" % nq,
            "max_rel_err_z8_vs_cpu_auto": auto_err,
        }
    return res


def saturated(args, pq, W, np, probe, sim=None):
    """k_seg_walk at full occupancy: 4096 amplitudes <0|C|y> (random y) of the
    n=40 workload circuit in one launch, with the lane walk length autotuned
    for this batch (then restored)."""
    if sim is None:
        c, _, _ = W.config2(SEED)
        sim = pq.Simulator(c, strategy="cached", form="z8")
    seg = sim.seg_info()
    d = seg["d"]
    rng = W.Rng(99)
    big_x = [rng.below(1 << 40) for _ in range(4096)]
    big_y = [0] * 4096
    default, default_waves = pq.get_walk_tuning(), pq.get_walk_shape()
    S = [sim] * 4096
    if args.saturated_tuning:
        off, mx = _apply_tuning(pq, args.saturated_tuning)
    else:
        off, mx, table = pq.autotune(S, big_x, big_y, repeats=2, cache=False, objective="kernel")
    waves = pq.get_walk_shape()
    _settle(lambda: sim.amplitudes(big_x, big_y, form="z8"))
    bms, bwall = [], []
    for _ in range(args.saturated_reps):
        tb = time.perf_counter()
        sim.amplitudes(big_x, big_y, form="z8")
        bwall.append(time.perf_counter() - tb)
    with pq.timed():
        for _ in range(args.saturated_reps):
            sim.amplitudes(big_x, big_y, form="z8")
            bms.append(pq.last_timing()["kernel_ms"])
    tuned = sim.seg_info()["mid_bits"]
    pq.set_walk_tuning(*default)
    pq.set_walk_shape(default_waves)
    return _seg_roofline(probe, 4096 * (1 << seg["outer"]), float(np.median(bms)),
                         {"queries": 4096, "walk_tuning": [off, mx], "workgroup_waves": waves, "mid_bits": tuned,
                          "amplitudes_per_s": 4096 / float(np.median(bwall)),
                          "branches_per_s": 4096 * (1 << d) / float(np.median(bwall))},
                         traffic=SEG_TRAFFIC.get(("saturated", seg["inner"], tuned, waves)),
                         kernel="k_seg_walk<%s, %d>" % ("true" if tuned <= 4 else "false", waves))


def precision(args, pq, W, np):
    """n=48 with 40 Rz(theta): every amplitude's error against FP64."""
    c, xs, ys = W.config3(4, n_amps=1024)
    z = pq.Simulator(c, strategy="cached", form="z8")
    d = z.stats()[5]
    t0 = time.perf_counter(); az = z.amplitudes(xs, ys); t1 = time.perf_counter()
    aa = z.amplitudes(xs, ys, form="branch_walk"); t2 = time.perf_counter()
    p = pq.Simulator(c, strategy="cached", form="plain_fp64")
    t3 = time.perf_counter(); ap = p.amplitudes(xs, ys); t4 = time.perf_counter()
    out = {"workload": "n=48: H^48; 40 Rz(theta ~ U[0,2pi)) + 300 CX/CZ + 100 S (seed 4); H^48; 1024 amplitudes <0|C|y>",
           "branch_dim": d, "z8_vs_fp64": _err_summary(az, ap), "branch_walk_vs_fp64": _err_summary(aa, ap),
           "z8_s": t1 - t0, "branch_walk_s": t2 - t1, "plain_fp64_s": t4 - t3}
    if not args.no_cpu:
        O = _oracle()
        osim = O.OSim(O.OCircuit.from_circuit(c), "cached")
        osim.set_threads(max(1, min(16, os.cpu_count() or 1)))
        idx = list(range(0, 1024, 64))
        ao = np.array([osim.amplitude(xs[i], ys[i]) for i in idx])
        out["fp64_gpu_vs_cpu_oracle"] = _err_summary(ap[idx], ao)
        out["z8_vs_cpu_oracle"] = _err_summary(az[idx], ao)
    if args.precision_out:
        with open(args.precision_out, "w") as f:
            json.dump({"y": [int(v) for v in ys], "fp64": [[float(a.real), float(a.imag)] for a in ap],
                       "z8": [[float(a.real), float(a.imag)] for a in az],
                       "branch_walk": [[float(a.real), float(a.imag)] for a in aa]}, f)
    return out


def batch(args, pq, W, np):
    """256 circuits x 16 queries: one persistent launch, autotuned, graph-replayed."""
    tg = time.perf_counter()
    work = W.config4(1, n_circuits=args.batch_circuits)
    tb = time.perf_counter()
    sims = [pq.Simulator(c, strategy="cached") for c, _ in work]
    te = time.perf_counter()
    S, X, Y = [], [], []
    for s, (c, pairs) in zip(sims, work):
        for x, y in pairs:
            S.append(s); X.append(x); Y.append(y)
    ds = [s.stats()[5] for s in sims]
    branches = sum(len(p) * (1 << dd) for dd, (_, p) in zip(ds, work))
    first = pq.batch_amplitudes(S, X, Y)          # uploads the engine images
    default = pq.get_walk_tuning()
    # the replayed graph is timed on the device, so the split is chosen on
    # the walk kernel's time (the eager wall time also counts the prologues)
    off, mx, table = pq.autotune(S, X, Y, repeats=3, cache=False, objective="kernel")
    eager_ms = []
    for _ in range(3):
        t0 = time.perf_counter(); eager = pq.batch_amplitudes(S, X, Y); eager_ms.append(1e3 * (time.perf_counter() - t0))
    with pq.timed():
        pq.batch_amplitudes(S, X, Y)
        tm = pq.last_timing()
    g = pq.Graph(S, X, Y)
    reps, gms, wall = 10, [], []
    same = True
    for _ in range(reps):
        t0 = time.perf_counter()
        got, ms = g.run()
        wall.append(1e3 * (time.perf_counter() - t0))
        gms.append(ms)
        same = same and bool(np.array_equal(got, eager))
    out = {
        "workload": "%d circuits (n ~ U[20,56], T ~ U[8,32], 7.5n CX/CZ, 2.5n S, seed 1) x 16 (x, y)" % len(work),
        "queries": len(S), "branch_dim_min": min(ds), "branch_dim_max": max(ds), "branches": branches,
        "generate_s": tb - tg, "host_build_s": te - tb, "host_build_ms_per_circuit": 1e3 * (te - tb) / len(work),
        "tuning_default": list(default), "tuning_chosen": [off, mx], "workgroup_waves": pq.get_walk_shape(),
        "autotune_kernel_ms": table,
        "eager_ms_median": float(np.median(eager_ms)), "eager_kernel_ms": tm["kernel_ms"], "items": tm["items"],
        "graph_ms_median": float(np.median(gms)), "graph_wall_ms_median": float(np.median(wall)),
        "graph_replays_equal_eager": same,
        "amplitudes_per_s": len(S) / (1e-3 * float(np.median(wall))),
        "branches_per_s": branches / (1e-3 * float(np.median(wall))),
    }
    if not args.no_plain:
        P = {}
        for s, (c, _) in zip(sims, work):
            P[id(s)] = pq.Simulator(c, strategy="cached", form="plain_fp64")
        t0 = time.perf_counter()
        ref = np.array([P[id(s)].amplitudes([x], [y])[0] for s, x, y in zip(S, X, Y)])
        t1 = time.perf_counter()
        out["plain_fp64_amplitudes_per_s"] = len(S) / (t1 - t0)
        out["z8_vs_fp64"] = _err_summary(eager, ref)
        out["speedup_vs_plain_amps"] = out["amplitudes_per_s"] / out["plain_fp64_amplitudes_per_s"]
    del first
    return out


def _serve_schedule(W, rate, secs, rseed=2024, cseed=77, max_requests=None):
    """Poisson arrivals at `rate` for `secs` seconds: 50% new circuits of the
    batch generator, 50% repeats; 70% amplitude, 20% probability, 10% 4-qubit
    marginal."""
    rnd = random.Random(rseed)
    rng = W.Rng(cseed)
    t, sched, circuits, fresh = 0.0, [], [], []
    while max_requests is None or len(sched) < max_requests:
        t += rnd.expovariate(rate)
        if t >= secs:
            break
        if not circuits or rnd.random() < 0.5:
            c = W.config4_circuit(rng)
            circuits.append(c)
            fresh.append(True)
        else:
            c = circuits[rnd.randrange(len(circuits))]
            fresh.append(False)
        u = rnd.random()
        n = c.num_qubits
        x, y = rnd.getrandbits(n), rnd.getrandbits(n)
        if u < 0.7:
            sched.append((t, c, "amplitude", x, y, ()))
        elif u < 0.9:
            sched.append((t, c, "probability", x, y, ()))
        else:
            sched.append((t, c, "marginal", x, 0, tuple((q, rnd.random() < 0.5) for q in rnd.sample(range(n), 4))))
    for c in circuits:
        c.arrays()                  # gate arrays built before the clock starts
    return sched, circuits, fresh


def _serve_run(pq, np, sched, fresh):
    srv = pq.Server(hbm_budget_bytes=64 << 30, workers=8)
    tickets = []
    t0 = time.perf_counter()
    for (ta, c, kind, x, y, meas) in sched:
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        tickets.append(srv.submit(c, kind, x, y, meas))
    t_sub = time.perf_counter() - t0
    lat, vals, failed, errors, lat_new, lat_rep = [], [], 0, [], [], []
    for i, tk in enumerate(tickets):
        try:
            v, l = srv.wait(tk, timeout=120.0)
            lat.append(l); vals.append(v)
            (lat_new if fresh[i] else lat_rep).append(l)
        except pq.PolyqError as e:
            failed += 1; vals.append(None)
            if len(errors) < 5:
                _, c, kind, x, y, meas = sched[i]
                errors.append({"error": str(e), "kind": kind, "n": c.num_qubits, "measured": [list(m) for m in meas]})
    t_end = time.perf_counter() - t0
    st = srv.stats()
    hs = srv.host_stats()
    phases = srv.phase_stats()
    srv.close()
    lat = np.array(lat)
    pct = lambda a, q: float(np.percentile(a, q)) if len(a) else None
    out = {
        "requests": len(sched), "failed": failed, "errors": errors,
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "max_ms": float(lat.max()) if len(lat) else None,
        # first request of a circuit (host build and upload included) / repeats of resident circuits
        "new_circuit_p50_ms": pct(lat_new, 50), "new_circuit_p99_ms": pct(lat_new, 99),
        "repeat_p50_ms": pct(lat_rep, 50), "repeat_p99_ms": pct(lat_rep, 99),
        "throughput_req_per_s": len(lat) / t_end, "submit_span_s": t_sub, "drain_s": t_end - t_sub,
        "batches": st["batches"], "overlapped_batches": st["overlapped_batches"], "max_batch": st["max_batch"],
        "builds": st["builds"], "evictions": st["evictions"], "resident_bytes": st["resident_bytes"],
        "host_bytes": hs["host_bytes"], "host_drops": hs["host_drops"], "walk_kernel_ms_total": st["kernel_ms"],
        # mean ms per request phase, over all requests and over the slowest 1%
        "phases_ms": phases,
    }
    return out, vals


def serve(args, pq, W, np):
    """Open-loop Poisson load against the request loop."""
    rate, secs = args.serve_rate, args.serve_seconds
    sched, circuits, fresh = _serve_schedule(W, rate, secs)
    out, vals = _serve_run(pq, np, sched, fresh)
    # Alone vs in the mix: re-evaluate a sample of served requests one by one.
    rnd = random.Random(7)
    sample = rnd.sample(range(len(sched)), min(96, len(sched)))
    equal, maxdiff, cache = 0, 0.0, {}
    for i in sample:
        _, c, kind, x, y, meas = sched[i]
        if vals[i] is None:
            continue
        s = cache.get(id(c))
        if s is None:
            s = cache[id(c)] = pq.Simulator(c, strategy="cached")
        if kind == "amplitude":
            want = complex(*s.amplitude(x, y))
        elif kind == "probability":
            want = complex(s.probability(x, y), 0.0)
        else:
            want = complex(s.marginal(x, meas), 0.0)
        equal += int(want == vals[i])
        maxdiff = max(maxdiff, abs(want - vals[i]))
    kinds = {k: sum(1 for r in sched if r[2] == k) for k in ("amplitude", "probability", "marginal")}
    out.update({
        "workload": "Poisson %.0f req/s for %.0f s; 50%% new circuits (n ~ U[20,56], T ~ U[8,32]), 50%% repeats; "
                    "70/20/10 amplitude/probability/4-qubit marginal; 64 GB HBM budget" % (rate, secs),
        "new_circuits": len(circuits), "kinds": kinds,
        "alone_vs_mix": {"checked": len(sample), "bitwise_equal": equal, "max_abs_diff": maxdiff},
    })
    return out


def serve_sweep(args, pq, W, np):
    """The same request mix at rising Poisson rates (each for up to
    --sweep-seconds, at most 40000 requests) until the p99 latency bends:
    the served capacity of one GPU, with the largest batch and the number of
    batches that overlapped another on the device."""
    rows = []
    base_p99 = None
    for rate in (200, 400, 800, 1600, 3200, 6400, 12800, 25600):
        secs = min(args.sweep_seconds, 40000.0 / rate)
        sched, circuits, fresh = _serve_schedule(W, rate, secs, rseed=rate, cseed=rate + 1)
        out, _ = _serve_run(pq, np, sched, fresh)
        row = {"rate_req_per_s": rate, "seconds": secs}
        row.update({k: out[k] for k in ("requests", "failed", "p50_ms", "p99_ms", "max_ms", "throughput_req_per_s", "max_batch",
                                        "overlapped_batches", "batches", "builds", "host_drops")})
        row["slowest_1pct_phases_ms"] = out["phases_ms"]["slowest_1pct"]
        rows.append(row)
        print("bench: serve_sweep rate %d p50 %.3f p99 %.3f thr %.0f" % (rate, out["p50_ms"], out["p99_ms"], out["throughput_req_per_s"]),
              file=sys.stderr, flush=True)
        base_p99 = base_p99 or out["p99_ms"]
        if out["p99_ms"] > 10 * base_p99 or out["throughput_req_per_s"] < 0.8 * rate or out["failed"]:
            break
    return {"workload": "the serve mix (50% new config-4 circuits, 70/20/10 amplitude/probability/4-qubit marginal, 64 GB HBM, "
                        "8 host workers) at rising Poisson rates; stops once p99 exceeds 10x its value at 200 req/s or "
                        "throughput falls below 80% of the offered rate", "rows": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--sections", default="headline,saturated,precision,batch,serve,serve_sweep")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--sweep-seconds", type=float, default=5.0)
    ap.add_argument("--batch-circuits", type=int, default=256)
    ap.add_argument("--saturated-reps", type=int, default=5)
    ap.add_argument("--saturated-tuning", default=None,
                    help="offset,max_mid[,waves] for the saturating launch instead of autotuning")
    ap.add_argument("--precision-out", default=None)
    ap.add_argument("--headline-tuning", default="",
                    help="offset,max_mid of the headline (branch walk) lane walk instead of autotuning it (profiling runs)")
    ap.add_argument("--segment-tuning", default="",
                    help="offset,max_mid[,waves] of the segment walk on the headline step instead of autotuning it")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 256 MiB device copies (reference traffic for FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 64 << 20
        a = torch.ones(n, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_kernel": n * 4, "kernels": 3}))
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported (the 2^d branch range would shard by "
              "coset index across GPUs with one 64-byte reduction per query)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import polyq_amd as pq
    from polyq_amd import workloads as W

    sections = set(args.sections.split(","))
    probe = _valu_probe(pq)
    if "headline" not in sections:
        # profiling aid: only the listed sections (e.g. the saturating launches)
        out = {"valu_probe": probe}
        if "saturated" in sections:
            out["roofline_saturated"] = saturated(args, pq, W, np, probe)
        for name, fn in (("precision", precision), ("batch", batch), ("serve", serve), ("serve_sweep", serve_sweep)):
            if name in sections:
                out[name] = fn(args, pq, W, np)
        print(json.dumps(out))
        return
    h = headline(args, pq, W, np, probe)
    print("bench: headline done", file=sys.stderr, flush=True)
    h["segment"]["saturated"] = saturated(args, pq, W, np, probe, h["sim"]) if "saturated" in sections else None
    st, d, nq, elapsed = h["st"], h["d"], h["nq"], h["elapsed"]
    value = h["branches_per_step"] * args.steps / elapsed
    bw_amps = nq * args.steps / elapsed
    extra = {}
    for name, fn in (("precision", precision), ("batch", batch), ("serve", serve), ("serve_sweep", serve_sweep)):
        if name in sections:
            print("bench: section %s" % name, file=sys.stderr, flush=True)
            extra[name] = fn(args, pq, W, np)
    plain, cpu = h.get("plain"), h.get("cpu")
    segd = dict(h["segment"])
    segd.pop("got", None)
    seg_amps = segd["amplitudes_per_s"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "branches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (FP64 FMA accumulation per branch; int32 Z8 phases)",
        "data": "synthetic (seeded SplitMix64 circuits)",
        "config": {
            "workload": "n=40: H^40; 36 T + 300 CX/CZ + 100 S interleaved (seed %d); H^40; 16 amplitudes <0|C|y>, "
                        "every branch visited (per-branch walk)" % SEED,
            "terms": st[4], "branch_dim": d, "free_vars": st[0], "planes": st[1], "lines": st[2], "radical": st[3],
            "branches_per_step": h["branches_per_step"], "branch_walk": h["info"],
            "items": h["timing"]["items"], "grid": h["timing"]["grid"], "device": h["dev"],
            "walk_tuning": h["tuning"],
            "engine_build_ms": 1e3 * h["build_s"], "engine_upload_ms": 1e3 * h["upload_s"],
            # untimed steps of the chosen split run for SETTLE_S before the warm-up steps
            "settle": {"seconds": SETTLE_S, "steps": h["settle_steps"]},
        },
        "amplitudes_per_s_branch_walk": bw_amps,
        "amplitudes_per_s": seg_amps,
        "segment_walk": segd,
        "plain_fp64": plain,
        "speedup_vs_plain_amps": seg_amps / plain["amplitudes_per_s"] if plain else None,
        "speedup_vs_plain_branches": value / plain["branches_per_s"] if plain else None,
        "roofline": h["roofline"],
        "valu_probe": probe,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": value / cpu["value"] if cpu else None,
        "cpu_auto": h.get("cpu_auto"),
        "speedup_vs_cpu_auto_amps": seg_amps / h["cpu_auto"]["amplitudes_per_s"] if h.get("cpu_auto") else None,
    }
    line.update(extra)
    if "serve" in extra:
        line["p50_ms"] = extra["serve"]["p50_ms"]
        line["p99_ms"] = extra["serve"]["p99_ms"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
