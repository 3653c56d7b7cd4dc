"""Benchmark: logical Jaro-Winkler pair comparisons per second of the linkage
hot path on one MI355X.

Workload (default): plain linkage of 1,000,000 x 1,000,000 synthetic records,
three string variables (first_name, last_name, street) plus an exact-match
int32 house number. One step = the device stage of Comparison.fit from
encoded, HBM-resident columns: all-pairs JW levels of the three string
variables (unique values, expansion to record pairs, sort), the exact-match
pairs, and the pattern merge into the 54-pattern histogram. value =
n_A * n_B * 3 / step time.

Prints one JSON line (rank 0) with the roofline of the comparison kernel and a
CPU baseline (the oracle's C restatement on a bounded row sample).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]
"""
import argparse
import contextlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fast-er_amd"))

# MI355X peaks used for the roofline fractions.
# VALU: the published FP32 vector rate, 157.3 TFLOP/s = 7.86e13 FMA lane-ops/s
# (256 CUs x 4 SIMD x 32 lanes x 2.4 GHz: each SIMD can issue two VALU
# instructions per quad-cycle, counted by SQ_ACTIVE_INST_VALU2). A probe of
# the match recurrence's own ops (bitop3, add) sustains 6.4e13 on the device.
VALU_LANE_OPS_PEAK = 256 * 4 * 32 * 2.4e9
HBM_PEAK_GBS = 8000.0
FROZEN_PEAK = 256 * 64 * 2.4e9  # the peak the fixed work model was quoted against


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def comparison_roofline(stats_all, kernel_name, traffic):
    """Roofline of the comparison kernel from the per-call work accounting:
    algorithmic VALU lane-ops over the HIP-event time of the launches."""
    import numpy as np

    kern_ms = float(np.mean([sum(s["pass_ms"] for s in st) for st in stats_all]))
    lane = float(np.mean([sum(s["alg_lane_ops"] for s in st) for st in stats_all]))
    # The prefix form skips the recurrence steps of shared prefixes: its
    # algorithmic work is the per-pair count less those steps.
    skipped = float(np.mean([sum(s.get("prefix_skipped_ops", 0.0) for s in st) for st in stats_all]))
    executed = lane - skipped
    achieved = executed / (kern_ms / 1e3)
    frozen = float(np.mean([sum(s.get("frozen_model_ops", 0.0) for s in st) for st in stats_all]))
    return {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_LANE_OPS_PEAK / 1e12,
            "unit": "Tlane-op/s", "frac": achieved / VALU_LANE_OPS_PEAK, "traffic": traffic,
            "kernel": kernel_name, "kernel_ms": kern_ms, "alg_ops": executed,
            # Lane-ops of the same pairs when every character step is
            # computed (the per-lane form's work), at this kernel time.
            "per_pair_form_ops": lane, "per_pair_form_frac": lane / (kern_ms / 1e3) / VALU_LANE_OPS_PEAK,
            # The work model fixed before the kernel was written (8 lane-ops per
            # string-1 character + 24 per evaluated pair) against 3.93e13
            # lane-ops/s (the FP32 vector rate read as packed): a model of a
            # non-packed recurrence, so it reads above 1 for the packed one.
            "frozen_model": {"ops": frozen, "peak": FROZEN_PEAK / 1e12, "unit": "Tlane-op/s",
                             "achieved": frozen / (kern_ms / 1e3) / 1e12,
                             "frac": frozen / (kern_ms / 1e3) / FROZEN_PEAK}}


KERNEL_NAMES = {"tiles": "k_jw_tiles (bit-parallel, one string-1 value per lane)",
                "prefix": "k_jw_prefix (bit-parallel, string-1 values walked per wave, shared prefixes reused)"}


def roofline_steps(step, method, traffic, n_vars, repeats=2):
    """The comparison kernel's roofline from `repeats` extra steps run with
    the variables one after another (faster.concurrent.SERIAL): in the timed
    steps the variables' kernels share the GPU, so their HIP-event times
    overlap and would not add up to kernel time. Returns (roofline, stats of
    the last step)."""
    import torch
    from faster import concurrent

    stats_all = []
    concurrent.SERIAL = True
    try:
        for _ in range(repeats):
            st = []
            step(st)
            torch.cuda.synchronize()
            stats_all.append(st)
    finally:
        concurrent.SERIAL = False
    return comparison_roofline(stats_all, f"{KERNEL_NAMES[method]}, {n_vars} vars", traffic), stats_all[-1]


def valu_probe(dev):
    """Times fh_probe_valu (the match recurrence's three VALU ops on register
    operands, 2048 x 256 threads) with HIP events on its launch stream and
    prints the sustained lane-op rate against VALU_LANE_OPS_PEAK."""
    import ctypes

    import torch
    from faster import _lib

    L = _lib.load()
    blocks, iters = 2048, 20000
    out = torch.empty(blocks * 256, dtype=torch.int32, device=dev)
    ops = ctypes.c_double(0.0)
    rates = []
    for _ in range(4):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        _lib.check(L.fh_probe_valu(_lib.ptr(out), blocks, iters, ctypes.byref(ops), _lib.stream()), "fh_probe_valu")
        e1.record()
        torch.cuda.synchronize()
        rates.append(ops.value / (e0.elapsed_time(e1) / 1e3))
    best = max(rates[1:])
    print(json.dumps({"valu_probe_lane_ops_per_s": best, "peak": VALU_LANE_OPS_PEAK, "frac": best / VALU_LANE_OPS_PEAK,
                      "lane_ops_per_launch": ops.value, "rates": rates}), flush=True)


def measured_traffic(workload, n):
    """HBM bytes moved by the comparison-kernel dispatches of one step, from
    the rocprofv3 FETCH_SIZE / WRITE_SIZE runs recorded in
    profiles/traffic.json (each counter scaled by what it reads for a plain
    copy of known size). None when no record matches this workload size."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    if not rec or int(rec.get("n", -1)) != int(n):
        return None
    return int(rec["bytes"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["link", "dedup", "blocked", "realdata"], default="link",
                    help="link: 1M x 1M plain linkage (default); dedup: 2M-row deduplication; "
                         "blocked: 8M x 8M in Zipf blocks, batched launch vs per-block loops; "
                         "realdata: the reference's 100k x 100k example voter records (tests/golden fixture)")
    ap.add_argument("--blocks", type=int, default=4000, help="blocks of the blocked workload")
    ap.add_argument("--no-loops", action="store_true", help="blocked: skip the per-block loop baselines")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain device copy of 256 MiB three times (calibrates PMC byte counters)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="run only the VALU throughput probe of the match recurrence (checks the VALU peak)")
    ap.add_argument("--n", type=int, default=None, help="records per side (default 1M link, 2M dedup)")
    ap.add_argument("--method", choices=["prefix", "tiles"], default="prefix",
                    help="all-pairs comparison: the bit-parallel match with string-1 values walked per wave and "
                         "shared prefixes reused (prefix, default), or the same match with one string-1 value per lane "
                         "(tiles)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-blocked", action="store_true",
                    help="link workload: skip the batched-block section (8M x 8M in Zipf blocks)")
    ap.add_argument("--blocked-n", type=int, default=8_000_000, help="records per side of the batched-block section")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline duration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (it would shard by contiguous "
              "A-row ranges); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import faster
    from faster import _ops
    from faster.comparison import exact_keys_link, fuzzy_keys_link
    from faster.encode import group_exact_link_device, group_packed
    from faster.synthetic import STRING_VARS, linkage_tables

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    from faster import _lib

    _lib.set_jw_method(args.method)
    if args.copy_probe:
        src = torch.empty(256 << 20, dtype=torch.uint8, device=dev).random_(0, 255)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_per_launch": 256 << 20, "launches": 3}), flush=True)
        return
    if args.valu_probe:
        return valu_probe(dev)
    if args.workload == "dedup":
        return bench_dedup(args)
    if args.workload == "blocked":
        return bench_blocked(args)
    if args.workload == "realdata":
        return bench_realdata(args)
    n = args.n or 1_000_000
    log(f"generating {n} x {n} synthetic records")
    t0 = time.time()
    dfa, dfb, _ = linkage_tables(n, n, seed=args.seed)
    t_gen = time.time() - t0
    vf = list(STRING_VARS)
    comp0 = faster.Comparison(dfa, dfb, vf, vf, ["house_number"], ["house_number"])
    log(f"generated in {t_gen:.1f} s; packing columns on the host and copying them to the device")
    t0 = time.time()
    inputs = comp0.device_inputs()
    torch.cuda.synchronize()
    t_prep = time.time() - t0
    log(f"host packing + copy {t_prep:.2f} s")

    def step(stats_out=None):
        c = faster.Comparison(dfa, dfb, vf, vf, ["house_number"], ["house_number"])
        c.fit_device(inputs, 0.1, 0.88, 0.94, Exact_Mode="list")
        if stats_out is not None:
            stats_out.extend(c.Stats)
        return c

    for w in range(args.warmup):
        log(f"warmup step {w + 1}/{args.warmup}")
        step()
        torch.cuda.synchronize()
    times = []
    stats_all = []
    comp = None
    for k in range(args.steps):
        st = []
        del comp
        comp = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        comp = step(st)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        times.append(dt)
        stats_all.append(st)
        log(f"step {k + 1}/{args.steps}: {dt * 1e3:.1f} ms (Comparison.fit device stage incl. grouping)")
    ms = 1e3 * sum(times) / len(times)
    comps = float(n) * float(n) * 3
    value = comps / (ms / 1e3)

    roof, st0 = roofline_steps(step, args.method, measured_traffic(f"link_{args.method}", n), 3)
    kern_ms = roof["kernel_ms"]
    counts = comp.Counts
    level_sizes = [int(k.numel()) for k in comp._var_keys]
    n_union = int(comp._pattern_ids.numel())
    del comp

    # The step from columns already grouped on the device (grouping outside).
    dgroups = [(group_packed(a), group_packed(b)) for a, b in inputs.fuzzy]
    ex = group_exact_link_device(inputs.exact[0], n)
    torch.cuda.synchronize()

    def step_grouped():
        keys = [fuzzy_keys_link(a, b, 0.1, 0.88, 0.94) for a, b in dgroups]
        keys.append(exact_keys_link(ex.group_a, ex.starts_b, ex.rows_b, ex.n_groups, n))
        return _ops.pattern_merge(keys, [3, 3, 3, 2])

    step_grouped()
    g_times = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, offs_g, _ = step_grouped()
        torch.cuda.synchronize()
        g_times.append(time.perf_counter() - t0)
    counts_g = np.concatenate([[n * n - int(offs_g[-1])], np.diff(np.asarray(offs_g))[1:]]).astype(np.int64)
    log(f"step from grouped columns: {1e3 * np.mean(g_times):.1f} ms; counts equal: {np.array_equal(counts_g, counts)}")
    del dgroups, ex

    # EM on the histogram (reported, not part of value).
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    np.random.seed(0)
    est = faster.Estimation(3, 1, counts)
    with contextlib.redirect_stdout(sys.stderr):  # stdout carries only the JSON line
        est.fit(Tolerance=1e-6)
    t_em = time.perf_counter() - t0

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(dfa, dfb, vf, ["house_number"], args.cpu_seconds)
    del inputs
    torch.cuda.empty_cache()

    blocked = None
    if not args.no_blocked:
        blocked = blocked_section(args)

    line = {
        "metric": "JW pair-comparisons/s at 1M x 1M, 3 string vars (plain linkage); batched-block vs plain speedup",
        "value": value,
        "unit": "comparisons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded generator, lognormal lengths, 30% perturbed copies)",
        "config": {"workload": f"plain linkage {n} x {n} records, 3 string vars + int32 house number, thresholds "
                               "[0.88, 0.94], p = 0.1; step = Comparison.fit from device-resident packed columns "
                               "(device grouping of all 8 columns, all-pairs JW, expansion, sort, exact pairs, pattern "
                               "merge)", "n_a": n, "n_b": n, "method": args.method,
                   "string_vars": list(STRING_VARS), "exact_vars": ["house_number"], "patterns": 54},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extras": {
            "pairs_kernel_per_var": [s["pairs_kernel"] for s in st0],
            "pairs_pruned_per_var": [s["pairs_pruned"] for s in st0],
            "candidates_per_var": [s["pairs_candidates"] for s in st0],
            "comparison_stage_ms_per_var": [s["kernel_ms"] for s in st0],
            "alphabet": [s["alphabet"] for s in st0],
            "kernel_comparisons_per_s": comps / (kern_ms / 1e3),
            "level_list_sizes": level_sizes,
            "union_pairs": n_union,
            "counts_head": counts[:6].tolist(),
            "em_ms": t_em * 1e3,
            "em_iterations": est.Iterations,
            "host_pack_and_copy_s": t_prep,
            "value_incl_host_pack": comps / (ms / 1e3 + t_prep),
            "step_from_grouped_ms": 1e3 * float(np.mean(g_times)),
            "step_from_grouped_counts_equal": bool(np.array_equal(counts_g, counts)),
            "step_ms_all": [t * 1e3 for t in times],
            "blocked": blocked,
        },
    }
    print(json.dumps(line), flush=True)


def bench_dedup(args):
    """2M-row deduplication: 3 string vars (upper triangle) + birth year kept
    as row ids (its ~2.6e10 equal pairs are counted in closed form)."""
    import numpy as np
    import torch

    import faster
    from faster import _lib, _ops, patterns
    from faster.deduplication import fuzzy_keys_dedup
    from faster.encode import group_exact_dedup, group_strings_device
    from faster.synthetic import STRING_VARS, dedup_table

    n = args.n or 2_000_000
    log(f"generating {n} records for deduplication")
    t0 = time.time()
    df, _ = dedup_table(n, seed=args.seed)
    log(f"generated in {time.time() - t0:.1f} s; encoding")
    t0 = time.time()
    dgroups = [group_strings_device(df[v].to_numpy()) for v in STRING_VARS]
    by = group_exact_dedup(df["birth_year"].to_numpy())
    by_dev = by.to_device()
    torch.cuda.synchronize()
    t_enc = time.time() - t0
    log(f"encoded in {t_enc:.2f} s; unique: " + ", ".join(str(g.n_unique) for g in dgroups))

    def step(stats=None):
        keys = [fuzzy_keys_dedup(g, 0.1, 0.88, 0.94, stats) for g in dgroups]
        ids, offs, idx, counts = patterns.merge_with_rowids(
            keys, [3, 3, 3], [(by_dev[0], by_dev[0])], [(by.group, by.group)], n, True, n * (n + 1) // 2,
            lambda *a: None)
        return keys, counts

    for w in range(args.warmup):
        log(f"warmup step {w + 1}/{args.warmup}")
        step()
        torch.cuda.synchronize()
    times, stats_all = [], []
    for k in range(args.steps):
        st = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        keys, counts = step(st)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        stats_all.append(st)
        log(f"step {k + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms")
    ms = 1e3 * sum(times) / len(times)
    comps = float(n) * float(n - 1) / 2 * 3
    roof = comparison_roofline(stats_all, f"{KERNEL_NAMES[args.method]} (triangle), 3 vars",
                               measured_traffic(f"dedup_{args.method}", n))
    np.random.seed(0)
    est = faster.Estimation(3, 1, counts)
    with contextlib.redirect_stdout(sys.stderr):  # stdout carries only the JSON line
        est.fit(Tolerance=1e-6)
    line = {
        "metric": "JW pair-comparisons/s, 2M-row deduplication (upper triangle), 3 string vars",
        "value": comps / (ms / 1e3), "unit": "comparisons/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic (20% perturbed duplicates in clusters of 2-5)",
        "config": {"workload": f"deduplication of {n} records, 3 string vars + int16 birth year (row ids)",
                   "n": n, "patterns": 54, "method": args.method},
        "roofline": roof,
        "cpu_baseline": None,
        "extras": {"level_list_sizes": [int(k.numel()) for k in keys], "counts_head": counts[:4].tolist(),
                   "exact_equal_pairs": int(patterns.exact_pattern_totals([by.group], [by.group], True)[1]),
                   "em_iterations": est.Iterations, "encode_s": t_enc,
                   "step_ms_all": [t * 1e3 for t in times]},
    }
    print(json.dumps(line), flush=True)


def blocked_section(args, recipe_loop=True, autotune_space=None):
    """Batched blocked linkage (8M x 8M records in Zipf(1.2) blocks): the
    batched form (one persistent launch per run, hipGraph replay, autotuned)
    against the plain per-block launch loop over the same plan (one launch set
    per block, stream-ordered, counts read back once) and against the
    reference's recipe (one Comparison.fit per block), with per-block counts
    compared for equality."""
    import numpy as np
    import torch

    from faster.blocking import BlockedLinkage, per_block_loop
    from faster.synthetic import STRING_VARS, blocked_tables

    n = args.blocked_n if getattr(args, "blocked_n", None) else 8_000_000
    if getattr(args, "workload", "link") == "blocked" and args.n:
        n = args.n
    log(f"blocked: generating {n} x {n} records in {args.blocks} Zipf(1.2) blocks")
    t0 = time.time()
    dfa, dfb, ba, bb, sizes = blocked_tables(n, n_blocks=args.blocks, seed=args.seed)
    comps = float(np.dot(sizes, sizes)) * 3
    log(f"blocked: generated in {time.time() - t0:.1f} s; sum of block products {comps / 3:.3e}; "
        f"largest block {sizes.max()}")
    vf = list(STRING_VARS)
    t0 = time.time()
    bl = BlockedLinkage(dfa, dfb, ba, bb, vf, vf, ["house_number"], ["house_number"])
    t_enc = time.time() - t0
    log(f"blocked: encoded in {t_enc:.1f} s; autotuning")
    t0 = time.time()
    space = autotune_space or dict(tile_width=[256], tiles_per_item=[4, 8], waves_per_item=[8], wgs_per_cu=[2, 3])
    best, tune_res = bl.autotune(space=space)
    t_tune = time.time() - t0
    log(f"blocked: autotune {t_tune:.1f} s: best {best}")
    for _ in range(2):
        bl.fit(use_graph=True)
    walls, devs = [], []
    for k in range(5):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bl.fit(use_graph=True)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
        devs.append(bl.last_ms)
    batched = bl.Block_Counts.copy()
    info = bl.info()
    ms = 1e3 * float(np.mean(walls))
    log(f"blocked: batched {ms:.1f} ms wall, {np.mean(devs):.1f} ms device")
    # Plain per-block launch loop over the same plan.
    bl.run_per_block()
    bl.collect()
    lw, ld = [], []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bl.run_per_block()
        loop_counts = bl.collect().copy()
        torch.cuda.synchronize()
        lw.append(time.perf_counter() - t0)
        ld.append(bl.last_ms)
    loop_ms = 1e3 * float(np.mean(lw))
    loop_equal = bool(np.array_equal(loop_counts, batched))
    log(f"blocked: per-block launch loop {loop_ms:.1f} ms wall, {np.mean(ld):.1f} ms device, counts equal: {loop_equal}")
    achieved = info["alg_lane_ops"] / (float(np.mean(devs)) / 1e3)
    out = {"workload": f"blocked linkage {n} x {n} records, {args.blocks} Zipf(1.2) blocks (sizes 10-60000 before "
                       "rescaling), 3 string vars + house number",
           "batched_wall_ms": ms, "batched_device_ms": float(np.mean(devs)),
           "per_block_launch_loop_wall_ms": loop_ms, "per_block_launch_loop_device_ms": float(np.mean(ld)),
           "per_block_launch_loop_counts_equal": loop_equal,
           "speedup_vs_per_block_launch_loop": loop_ms / ms,
           "comparisons_per_s": comps / (ms / 1e3), "largest_block": int(sizes.max()), "blocks": int(args.blocks),
           "union_pairs": int(bl.n_union), "counts_head": bl.Counts[:6].tolist(), "tuning": best,
           "autotune_s": t_tune, "host_encode_s": t_enc, "plan": info,
           "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_LANE_OPS_PEAK / 1e12,
                        "unit": "Tlane-op/s", "frac": achieved / VALU_LANE_OPS_PEAK,
                        "kernel": "k_jw_blocked (persistent, all blocks)"}}
    if recipe_loop:
        log("blocked: the reference's recipe, one Comparison.fit per block")
        ref, t_ref = per_block_loop(dfa, dfb, ba, bb, vf, ["house_number"],
                                    progress=lambda k: log(f"  recipe loop: {k} blocks done"))
        out["recipe_loop_s"] = t_ref
        out["recipe_counts_equal"] = bool(np.array_equal(ref, batched))
        out["speedup_vs_recipe_loop"] = t_ref / (ms / 1e3)
        log(f"blocked: recipe loop {t_ref:.1f} s, equal counts: {out['recipe_counts_equal']}")
    del bl
    torch.cuda.empty_cache()
    return out


def bench_blocked(args):
    """The batched-block section alone, as its own line (value = speedup of the
    batched form over the plain per-block launch loop)."""
    out = blocked_section(args, recipe_loop=not args.no_loops,
                          autotune_space=dict(tile_width=[128, 256], tiles_per_item=[2, 8], waves_per_item=[4, 8],
                                              wgs_per_cu=[2, 3]))
    line = {
        "metric": "batched-block vs plain per-block speedup (8M x 8M, Zipf blocks, 3 string vars)",
        "value": out["speedup_vs_per_block_launch_loop"], "unit": "x", "n_gpus": 1, "steps": 5,
        "warmup": 2, "ms_per_step": out["batched_wall_ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic (Zipf(1.2) block sizes, 30% perturbed copies)",
        "config": {"workload": out["workload"]},
        "roofline": dict(out["roofline"], traffic=measured_traffic("blocked", args.n or 8_000_000)),
        "cpu_baseline": None,
        "extras": out,
    }
    print(json.dumps(line), flush=True)


def bench_realdata(args):
    """The reference's example data (100,000 x 100,000 voter records, linked as
    its notebook does: last_name, first_name, house_number, street_name fuzzy,
    birth_year exact). One step = Comparison.fit_device from HBM-resident
    packed columns (device grouping included)."""
    import numpy as np
    import torch

    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_realdata

    dfa, dfb, _ = make_realdata.load()
    n_a, n_b = len(dfa), len(dfb)
    vf = ["last_name", "first_name", "house_number", "street_name"]
    import faster

    inputs = faster.Comparison(dfa, dfb, vf, vf, ["birth_year"], ["birth_year"]).device_inputs()
    torch.cuda.synchronize()

    def step(stats_out=None):
        c = faster.Comparison(dfa, dfb, vf, vf, ["birth_year"], ["birth_year"])
        c.fit_device(inputs, 0.1, 0.88, 0.94, Exact_Mode="list")
        if stats_out is not None:
            stats_out.extend(c.Stats)
        return c

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    times, stats_all = [], []
    comp = None
    for k in range(args.steps):
        st = []
        comp = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        comp = step(st)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        stats_all.append(st)
        log(f"step {k + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms")
    ms = 1e3 * float(np.mean(times))
    comps = float(n_a) * float(n_b) * len(vf)
    roof, st0 = roofline_steps(step, args.method, None, 4)
    keys = comp._var_keys
    n_union = int(comp._pattern_ids.numel())
    e2e = end_to_end_realdata(dfa, dfb, vf)
    line = {
        "metric": "JW pair-comparisons/s, reference example data 100k x 100k, 4 string vars",
        "value": comps / (ms / 1e3), "unit": "comparisons/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "reference example voter records (example/df_{A,B}_100000.Rdata)",
        "config": {"workload": "linkage of the reference's 100,000 x 100,000 example records as its notebook does: "
                               "4 fuzzy vars + birth_year exact; step = Comparison.fit from device-resident packed "
                               "columns", "method": args.method},
        "roofline": roof, "cpu_baseline": None,
        "extras": {"unique_pairs_per_var": [s["pairs_logical"] for s in st0],
                   "comparison_stage_ms_per_var": [s["kernel_ms"] for s in st0],
                   "pass_ms_per_var": [s["pass_ms"] for s in st0],
                   "scoring_ms_per_var": [s["eval_ms"] for s in st0],
                   "candidates_per_var": [s["pairs_candidates"] for s in st0],
                   "level_list_sizes": [int(k.numel()) for k in keys], "union_pairs": int(n_union),
                   "step_ms_all": [t * 1e3 for t in times], "end_to_end": e2e},
    }
    print(json.dumps(line), flush=True)


def end_to_end_realdata(dfa, dfb, vf):
    """The reference notebook's pipeline through the public classes, from the
    DataFrames (host encoding, device grouping, comparison, merge, EM,
    Linkage.transform with its pandas merge): Comparison(...).fit(),
    Estimation(4, 1, Counts).fit(), Linkage(...).transform(0.85), as in
    example/Example.ipynb cell 5 (seeded EM). One untimed run, then one timed.
    The notebook reports 28.4 s wall for these calls on an NVIDIA T4."""
    out = {}
    for rep in range(2):
        with contextlib.redirect_stdout(sys.stderr):  # Estimation.fit prints; stdout carries the JSON line only
            t = _end_to_end_once(dfa, dfb, vf)
        log(f"end to end run {rep + 1}/2: {t['total_s']:.2f} s ({t})")
        out = t
    out["reference_notebook_s"] = 28.4
    out["reference_notebook_hardware"] = "NVIDIA T4 (example/Example.ipynb cell 5)"
    return out


def _end_to_end_once(dfa, dfb, vf):
    import numpy as np
    import torch

    import faster

    t = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    comp = faster.Comparison(dfa, dfb, vf, vf, ["birth_year"], ["birth_year"])
    comp.fit()
    torch.cuda.synchronize()
    t["comparison_fit_s"] = time.perf_counter() - t0
    t1 = time.perf_counter()
    np.random.seed(0)
    est = faster.Estimation(len(vf), 1, comp.Counts)
    est.fit()
    t["estimation_fit_s"] = time.perf_counter() - t1
    t2 = time.perf_counter()
    linked = faster.Linkage(dfa, dfb, comp.Indices, est.Ksi).transform(0.85)
    torch.cuda.synchronize()
    t["linkage_transform_s"] = time.perf_counter() - t2
    t["total_s"] = time.perf_counter() - t0
    t["linked_pairs"] = int(len(linked))
    return t


def cpu_baseline(dfa, dfb, vf, ex, seconds):
    """The oracle's whole Comparison.fit pipeline on the host cores (C
    Jaro-Winkler over a thread pool, NumPy for unique values, expansion, exact
    pairs and the pattern merge), run on two contiguous samples of A rows
    against every B row. The time is fitted as fixed + per-row * rows (the
    fixed part, mostly the B side's unique values, is paid once per job) and
    extrapolated to all A rows: work per A row is uniform under the
    generator."""
    import time as _t

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import faster_oracle as oracle

    threads = oracle.n_threads()
    n_a, n_b = len(dfa), len(dfb)

    def run(rows):
        sub = dfa.iloc[:rows]
        t0 = _t.perf_counter()
        oracle.comparison_fit([sub[v].to_numpy() for v in vf], [dfb[v].to_numpy() for v in vf],
                              [sub[e].to_numpy() for e in ex], [dfb[e].to_numpy() for e in ex], threads=threads)
        return _t.perf_counter() - t0

    # Per-row cost estimate from the comparison scan alone (B packed
    # beforehand), to size the second sample.
    probe = 16
    packed_b = [oracle.pack_strings(list(dfb[v].to_numpy())) for v in vf]
    t0 = _t.perf_counter()
    for v, pb in zip(vf, packed_b):
        oracle.jaro_winkler_all_pairs(list(dfa[v].to_numpy()[:probe]), pb, threads=threads)
    per_row = (_t.perf_counter() - t0) / probe
    del packed_b
    r1 = 32
    r2 = r1 + max(64, int(seconds * 0.5 / max(per_row, 1e-6)))
    t1 = run(r1)
    t2 = run(r2)
    c = max(t2 - t1, 1e-9) / (r2 - r1)
    fixed = max(t1 - r1 * c, 0.0)
    full = fixed + c * n_a
    log(f"CPU baseline: {threads} threads; {r1} rows {t1:.1f} s, {r2} rows {t2:.1f} s -> fixed {fixed:.1f} s + "
        f"{c * 1e3:.2f} ms per A row -> {full:.0f} s for all {n_a} rows")
    return {"value": float(n_a) * n_b * len(vf) / full, "unit": "comparisons/s", "cores": threads, "kind": "port",
            "sample": f"the oracle's Comparison.fit (unique values, all-pairs JW, expansion, exact pairs, pattern merge, "
                      f"{len(vf)} string vars + {len(ex)} exact) on the first {r1} and {r2} A rows x all {n_b} B rows "
                      f"({t1:.1f} s and {t2:.1f} s), fitted as {fixed:.1f} s + {c * 1e3:.2f} ms per A row and "
                      f"extrapolated to {n_a} rows ({full:.0f} s)",
            "extrapolated_s": full}


if __name__ == "__main__":
    main()
