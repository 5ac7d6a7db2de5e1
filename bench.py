"""Benchmark of the SQD hot path on one MI355X.

A step is one matvec of the FP32-compensated bucket-join kernel over the
whole subspace of the workload (default: the 28-qubit Heisenberg chain plus
disorder, 2^22 states). ``value`` is the algorithmic HBM rate of the matvec,
B_mv / t_mv with

    B_mv = N (8 + 2 s_v) + J_s N 8 + H s_v

(own key + v read + Hv write per row, one 8-byte key per searched
(row, off-diagonal group), one gathered element per hit; s_v = 8 bytes for
complex64; J_s = off-diagonal groups searched after parity pruning; H = hits).
The same JSON line carries the roofline object, the CPU baseline (the oracle's
reference dataflow on a bounded group sample, extrapolated) and the time to a
1e-6 relative residual of the fused FP32/FP64 LOBPCG next to the plain FP64
reference-structured LOBPCG, both from the same seeded Gaussian start.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "LOBPCG time to 1e-6 residual + matvec GB/s (% HBM peak), vs plain FP64"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def count_hits(prob) -> int:
    """Total (row, off-diagonal group) pairs whose target is in the subspace."""
    import torch
    total = 0
    for x in prob.op.host.xmask:
        total += int(torch.count_nonzero(prob.space.xsource(int(x)) >= 0).item())
    return total


def read_traffic(path):
    """HBM bytes per launch from a committed rocprof summary (or None)."""
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(w, j_s, n_groups_sample, b_mv):
    """Oracle (numpy port of the reference dataflow) on a group sample."""
    from oracle import sqd_oracle as so
    from rqutils_amd.operator import keys_to_states
    x, z, c, _ = so.from_paulisum(w.labels, w.coeffs, force_real=True, add_padding=True)
    packed = so.pack_states(keys_to_states(w.keys, w.num_qubits))
    rng = np.random.default_rng(0)
    v = (rng.normal(size=w.n) + 1j * rng.normal(size=w.n)).astype(np.complex64)
    groups = [g for g in range(1, x.shape[0])][: n_groups_sample]
    t0 = time.perf_counter()
    out = np.zeros(w.n, np.complex128)
    for g in groups:
        xs = so.get_xsource(x[g], packed)
        out += so.apply_xgrp(xs, so.get_diagonal(z[g], c[g], packed), v)
    dt = time.perf_counter() - t0
    t_full = dt * j_s / max(1, len(groups))
    return {"value": round(b_mv / t_full / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"oracle sort-join matvec over {len(groups)} of {j_s} searched X groups, "
                      f"all {w.n} rows, {dt:.1f} s, extrapolated x{j_s / max(1, len(groups)):.1f}",
            "seconds_per_matvec_extrapolated": round(t_full, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="heisenberg28",
                    choices=["heisenberg28", "random40", "small16", "batch64", "netstart", "nextrows",
                             "svsim"])
    ap.add_argument("--instances", type=int, default=64, help="instances of batch64")
    ap.add_argument("--no-solve", action="store_true", help="skip the time-to-solution block")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-groups", type=int, default=3)
    ap.add_argument("--rows-per-thread", type=int, default=0,
                    help="fix rows per thread (0: autotune the launch)")
    ap.add_argument("--group-split", type=int, default=1)
    ap.add_argument("--stage-v", type=int, default=0, help="stage v slices in LDS (0/1)")
    ap.add_argument("--compact", type=int, default=0,
                    help="with an explicit launch: apply hits through the per-wave LDS queue (0/1)")
    ap.add_argument("--bucket-cap", type=int, default=None, help="max keys per bucket")
    ap.add_argument("--dense-bits", type=int, default=0,
                    help="dense lookup at this bucket width (0: autotune or search)")
    ap.add_argument("--no-net", action="store_true", help="skip the network-start block")
    ap.add_argument("--no-intake", action="store_true", help="skip the sample-intake block")
    ap.add_argument("--no-cache", action="store_true", help="skip the cached hit table")
    ap.add_argument("--no-svsim", action="store_true", help="skip the state-vector block")
    ap.add_argument("--net-train", type=int, default=100,
                    help="training Hamiltonians of the network-start block")
    ap.add_argument("--net-test", type=int, default=25, help="held-out Hamiltonians")
    ap.add_argument("--net-steps", type=int, default=3000, help="Adam steps")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 1 GiB device copy first (scale for rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from rqutils_amd import workloads as W
    from rqutils_amd.sqd import SQDProblem

    dev = torch.device("cuda", 0)
    if args.workload == "batch64":
        bench_batch(args, dev)
        return
    if args.workload == "svsim":
        return bench_svsim(args, dev)
    if args.workload == "netstart":
        bench_netstart(args, dev)
        return
    if args.workload == "nextrows":
        bench_next_rows(args, dev)
        return
    if args.calibrate_copy:
        # A plain copy of known size: 1 GiB read + 1 GiB written per launch.
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("calibration copy: 3 x 1 GiB device copies done")
    t = time.time()
    w = W.make(args.workload)
    log(f"workload {w.name}: {w.n} states, {len(w.labels)} terms ({time.time() - t:.1f} s)")
    t = time.time()
    if args.bucket_cap or args.rows_per_thread or args.dense_bits:
        space_kw = {"rows_per_thread": args.rows_per_thread or 4, "dense": False}
        if args.bucket_cap:
            space_kw["bucket_cap"] = args.bucket_cap
        if args.dense_bits:
            space_kw.update(dense=True, bucket_bits=args.dense_bits)
        prob = SQDProblem(w.hamiltonian, w.keys, w.num_qubits, device=dev,
                          group_split=args.group_split, compact=args.compact, **space_kw)
    else:
        prob = SQDProblem(w.hamiltonian, w.keys, w.num_qubits, device=dev, tune="auto")
        log(f"autotune: {[t for t in prob.tuned.table]} -> cap={prob.tuned.bucket_cap} "
            f"dense_bits={prob.tuned.dense_bits} "
            f"rows={prob.tuned.rows_per_thread} split={prob.tuned.group_split} "
            f"compact={prob.tuned.compact} ({time.time() - t:.1f} s incl. tuning)")
    prob.op.tune.stage_v = args.stage_v
    torch.cuda.synchronize()
    sp, op = prob.space, prob.op
    log(f"space: bucket_bits={sp.bucket_bits} fine_bits={sp.fine_bits} buckets={sp.n_buckets} "
        f"max_bucket={sp.max_bucket} items={sp.n_items}; groups searched={op.n_groups} "
        f"(pruned {op.host.pruned_groups}), classes={op.n_classes}, diag terms={op.diag_terms} "
        f"({time.time() - t:.1f} s)")

    n = w.n
    s_v = 8
    j_s = op.n_groups
    hits = count_hits(prob)
    b_mv = n * (8 + 2 * s_v) + j_s * n * 8 + hits * s_v
    log(f"hits={hits} ({hits / n:.2f}/row), B_mv={b_mv / 1e9:.3f} GB")

    g = torch.Generator(device="cpu").manual_seed(1234)
    v = torch.complex(torch.randn(n, generator=g), torch.randn(n, generator=g)).to(
        torch.complex64).to(dev)
    out = torch.empty_like(v)
    for _ in range(args.warmup):
        prob.matvec(v, out)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        prob.matvec(v, out)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    achieved = b_mv / (ms * 1e-3) / 1e9
    log(f"matvec c64: {ms:.3f} ms, {achieved:.1f} GB/s algorithmic ({achieved / HBM_PEAK_GBS:.1%})")

    # complex128 matvec for comparison
    v128 = v.to(torch.complex128)
    o128 = torch.empty_like(v128)
    prob.matvec(v128, o128)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(max(3, args.steps // 4)):
        prob.matvec(v128, o128)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms128 = ev0.elapsed_time(ev1) / max(3, args.steps // 4)
    b128 = n * (8 + 2 * 16) + j_s * n * 8 + hits * 16
    diff = (out.to(torch.complex128) - o128).norm().item() / o128.norm().item()
    log(f"matvec c128: {ms128:.3f} ms; c64 vs c128 rel diff {diff:.2e}")

    record = {
        "metric": METRIC,
        "value": round(achieved, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 two-sum accumulate)",
        "data": "synthetic (seeded generator, rqutils_amd.workloads)",
        "config": {"workload": w.name, "num_qubits": w.num_qubits, "states": n,
                   "terms": len(w.labels), "groups_searched": j_s,
                   "groups_pruned_odd": op.host.pruned_groups, "classes": op.n_classes,
                   "hits": hits, "bucket_bits": sp.bucket_bits, "fine_bits": sp.fine_bits,
                   "rows_per_thread": sp.rows_per_thread, "group_split": op.tune.group_split,
                   "lookup": "dense bitmap" if sp.dense else "search",
                   "compact_hits": bool(op.tune.compact),
                   "stage_v": args.stage_v, "autotuned": prob.tuned is not None,
                   "parallelism": "single GPU"},
        "matvecs_per_s": round(1e3 / ms, 2),
        "matvec_c128_ms": round(ms128, 4),
        "matvec_c128_gbs": round(b128 / (ms128 * 1e-3) / 1e9, 2),
        "c64_vs_c128_rel_diff": diff,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic(os.path.join(
                         ROOT, "profiles", f"bucket_join_kernel_{w.name}.md")),
                     "kernel": f"bucket_join_kernel<float2,uint32_t,{sp.rows_per_thread},"
                               f"{str(bool(args.stage_v)).lower()},{str(sp.dense).lower()},"
                               f"{str(bool(op.tune.compact)).lower()}>",
                     "bytes_per_launch": b_mv},
        "wall_s_timed": round(wall, 3),
    }

    if not args.no_cache:
        record["hit_table"] = hit_table_block(prob, v, out, o128, hits, args.steps, dev)
    if not args.no_solve:
        record["time_to_solution"] = time_to_solution(prob, w, dev, cached=not args.no_cache)
    if not args.no_intake:
        record["intake"] = intake_block(w, dev)
    if not args.no_svsim and w.name == "heisenberg28":
        record["svsim"] = svsim_block(dev, cpu=not args.no_cpu)
    if not args.no_net and w.name == "heisenberg28":
        record["network_start"] = {
            "heisenberg28_2^20": network_start(dev, "heisenberg28_2^20", args.net_train,
                                               args.net_test, args.net_steps),
            "heisenberg16_sector": network_start(dev, "heisenberg16_sector", 6, 2,
                                                 args.net_steps)}
    if not args.no_cpu:
        log("cpu baseline ...")
        record["cpu_baseline"] = cpu_baseline(w, j_s, args.cpu_groups, b_mv)
    print(json.dumps(record), flush=True)


def hit_table_block(prob, v, out_join, o128_join, hits, steps, dev):
    """The cached hit table: build time and size in each precision, then the
    product timed with HIP events like the bucket join above, and its
    difference from the join's product. Algorithmic bytes per product: the
    hits' entries (4 + s_c each; the padding is not read), the diagonal, v,
    out and the row length (3 s_v + 4 per row), one s_v gather per hit and 16
    per slice."""
    import torch
    from rqutils_amd import _lib
    op = prob.op
    op.use_cache = True
    stream = torch.cuda.current_stream(dev)
    res = {}
    for prec, vec, ref in ((_lib.C64, v, out_join), (_lib.C128, v.to(torch.complex128), o128_join)):
        name = "c64" if prec == _lib.C64 else "c128"
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.build_cache(prec)
        torch.cuda.synchronize()
        build_s = time.perf_counter() - t0
        info = op.cache_info[prec]
        o = torch.empty_like(vec)
        prob.matvec(vec, o)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            prob.matvec(vec, o)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        s_c = 16 if prec == _lib.C128 else 8
        algo = hits * (4 + s_c) + prob.n * (3 * s_c + 4) + hits * s_c + info["slices"] * 16
        diff = (o.to(torch.complex128) - ref.to(torch.complex128)).norm().item() / \
            ref.to(torch.complex128).norm().item()
        gbs = algo / (ms * 1e-3) / 1e9
        res[name] = {"build_s": round(build_s, 4), "entries": info["entries"],
                     "padding": round(info["entries"] / max(hits, 1), 3),
                     "table_bytes": info["bytes"], "ms": round(ms, 4),
                     "matvecs_per_s": round(1e3 / ms, 1), "algorithmic_bytes": algo,
                     "achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                     "rel_diff_vs_bucket_join": diff}
        log(f"hit table {name}: built in {build_s:.3f} s, {info['entries']} entries "
            f"({info['bytes'] / 1e9:.2f} GB, {info['entries'] / max(hits, 1):.2f}x the hits); "
            f"product {ms:.3f} ms ({gbs:.0f} GB/s algorithmic), rel diff vs join {diff:.1e}")
    op.use_cache = False
    return res


def time_to_solution(prob, w, dev, cached=False):
    import torch
    n = w.n
    g = torch.Generator(device="cpu").manual_seed(77)
    x0 = torch.complex(torch.randn(n, generator=g, dtype=torch.float64),
                       torch.randn(n, generator=g, dtype=torch.float64)).to(dev)
    res = {}
    runs = [("fused_c64", dict(mode="fused", precision="c64")),
            ("fused_c128", dict(mode="fused", precision="c128")),
            ("fused_c64_diag_precond", dict(mode="fused", precision="c64", precond="diag")),
            ("reference_c128", dict(mode="reference", precision="c128"))]
    if cached:
        runs += [("cached_fused_c64", dict(mode="fused", precision="c64")),
                 ("cached_fused_c64_diag_precond",
                  dict(mode="fused", precision="c64", precond="diag")),
                 ("cached_reference_c128", dict(mode="reference", precision="c128"))]
    for name, kw in runs:
        prob.op.use_cache = name.startswith("cached_")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vec, info = prob.solve(x0, criterion="relative", tol=1e-6, maxiter=3000, **kw)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[name] = {"seconds": round(dt, 3), "iterations": info.niter, "matvecs": info.matvecs,
                     "energy": info.eigval, "converged": info.converged,
                     "resid_rel": info.resid_norm / abs(info.eigval)}
        log(f"{name}: {dt:.2f} s, {info.niter} it, {info.matvecs} matvecs, E={info.eigval:.12f}, "
            f"r={info.resid_norm / abs(info.eigval):.2e}")
    prob.op.use_cache = False
    # the reference's default start: one-hot at the smallest diagonal entry
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vec, info = prob.solve(prob.vinit_min_diag(), mode="fused", precision="c64",
                           criterion="relative", tol=1e-6, maxiter=3000)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res["fused_c64_min_diag_start"] = {"seconds": round(dt, 3), "iterations": info.niter,
                                       "matvecs": info.matvecs, "energy": info.eigval,
                                       "converged": info.converged,
                                       "resid_rel": info.resid_norm / abs(info.eigval)}
    log(f"fused_c64 from the min-diagonal one-hot: {dt:.2f} s, {info.niter} it")
    # FP64 energy at a tight residual as the reference value
    vec, info = prob.solve(x0, mode="fused", precision="c128", criterion="relative", tol=1e-10,
                           maxiter=5000)
    e_ref = info.eigval
    for name in list(res):
        res[name]["energy_rel_err_vs_fp64_1e-10"] = abs(res[name]["energy"] - e_ref) / abs(e_ref)
    res["fp64_reference_energy"] = e_ref
    base = res["reference_c128"]["seconds"]
    res["speedup_fused_c64_vs_plain_fp64"] = round(base / res["fused_c64"]["seconds"], 2)
    res["speedup_fused_c64_diag_precond_vs_plain_fp64"] = round(
        base / res["fused_c64_diag_precond"]["seconds"], 2)
    if cached:
        # (the tables were built in the hit_table block, which reports their build time)
        cb = res["cached_reference_c128"]["seconds"]
        res["speedup_cached_fused_c64_diag_precond_vs_cached_plain_fp64"] = round(
            cb / res["cached_fused_c64_diag_precond"]["seconds"], 2)
        res["speedup_cached_fused_c64_diag_precond_vs_plain_fp64"] = round(
            base / res["cached_fused_c64_diag_precond"]["seconds"], 2)
    prob.op.use_cache = False
    return res


def intake_block(w, dev, repeats: int = 4, reps: int = 3):
    """Sample intake on the device: ``repeats * N`` state rows (every key of
    the workload at least once, the rest drawn with replacement, shuffled)
    resident in HBM as uint8 [rows, n]; timed: pack into keys and unique
    (occupancy bitmap for keys of at most 28 bits, radix sort above). The
    result must be the workload's keys bit for bit. CPU baseline: numpy
    (pack + np.unique) on a 2^22-row sample, scaled by the row count.
    Algorithmic bytes: the rows read and their keys written (n + 8 per row),
    then for the bitmap the grouping by slice (24 per row: histogram read,
    scatter read and write; none up to 20 bits), the grouped keys read (8 per
    row), the 2^n / 8-byte bitmap written, counted and read, and 8 per
    distinct key written; for the radix sort 24 per row per 8-bit pass and 24
    for the unique."""
    import torch
    from rqutils_amd.intake import pack_states, unique_keys, unpack_keys
    from rqutils_amd.operator import states_to_keys
    n, nq = w.n, w.num_qubits
    rng = np.random.default_rng(5)
    idx = np.concatenate([np.arange(n), rng.integers(0, n, size=(repeats - 1) * n)])
    rng.shuffle(idx)
    kd = torch.from_numpy(w.keys.view(np.int64)).to(dev)[torch.from_numpy(idx).to(dev)]
    states = unpack_keys(kd, nq)
    del kd
    rows = states.shape[0]
    keys = pack_states(states, dev)
    u = unique_keys(keys, nq)                        # warm-up
    ok = bool(np.array_equal(u.cpu().numpy().view(np.uint64), w.keys))
    stream = torch.cuda.current_stream(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        keys = pack_states(states, dev)
        u = unique_keys(keys, nq)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    sample = states[: 1 << 22].cpu().numpy()
    t0 = time.perf_counter()
    np.unique(states_to_keys(sample))
    cpu_s = (time.perf_counter() - t0) * rows / sample.shape[0]
    bitmap = nq <= 28 and ((1 << nq) // 8 <= 4 << 20 or (1 << nq) // 8 <= 32 * rows)
    if bitmap:
        algo = (rows * (nq + 8 + (24 if nq > 20 else 0) + 8) + 3 * max((1 << nq) // 8, 4)
                + 8 * u.numel())
    else:
        algo = rows * (nq + 8 + 24 * -(-nq // 8) + 24)
    gbs = algo / (ms * 1e-3) / 1e9
    log(f"intake: {rows} rows -> {u.numel()} distinct in {ms:.2f} ms on the GPU "
        f"({rows / (ms * 1e-3) / 1e9:.2f} G rows/s, {'bitmap' if bitmap else 'radix sort'}, "
        f"{gbs:.0f} GB/s algorithmic), bit-exact={ok}; numpy {cpu_s:.2f} s (scaled)")
    return {"rows": rows, "distinct": int(u.numel()), "ms": round(ms, 3),
            "rows_per_s": round(rows / (ms * 1e-3), 1), "bit_exact_vs_workload_keys": ok,
            "method": "bitmap" if bitmap else "radix_sort",
            "algorithmic_bytes": int(algo), "achieved_gbs": round(gbs, 1),
            "hbm_frac": round(gbs / 8000.0, 3),
            "state_bytes": rows * nq,
            "cpu_numpy_s_scaled": round(cpu_s, 3),
            "cpu_sample": f"pack + np.unique of {sample.shape[0]} rows, x{rows / sample.shape[0]:.1f}"}


NET_STARTS = ("gaussian", "min_diag", "network_bits", "network_bits_mix", "network_bits_diag",
              "network_bits_diag_mix", "gaussian_precond", "network_bits_diag_precond")
NET_MIX = 0.3   # norm of the Gaussian part of the *_mix network starts


def network_start(dev, family, n_train, n_test, steps):
    """LOBPCG (fused, FP32 vectors, 1e-6 relative residual) on held-out
    Hamiltonians of a family from six starts: a seeded Gaussian, the one-hot
    at the smallest diagonal entry, the start network on bits, and the start
    network on bits plus two diagonal-energy features, each network alone and
    with a Gaussian part of norm NET_MIX; then the Gaussian and the feature
    network's starts again with the diagonal preconditioner (``*_precond``).
    Both networks are trained on the FP64 ground vectors of ``n_train`` other
    members of the family."""
    import functools
    import multiprocessing as mp
    import torch
    from rqutils_amd import workloads as W
    from rqutils_amd.sqd import SQDProblem
    from rqutils_amd.startnet import StartNet, energy_features, gaussian_start, ground_targets

    if family == "heisenberg28_2^20":
        nq, make = 28, functools.partial(W.heisenberg28, n_states=1 << 20)
    elif family == "heisenberg16_sector":
        nq, make = 16, functools.partial(W.heisenberg28, n_states=None, n_random=200,
                                         num_qubits=16)
    else:
        raise ValueError(family)
    seeds = list(range(1000, 1000 + n_train)) + list(range(2000, 2000 + n_test))
    t = time.time()
    with mp.get_context("spawn").Pool(min(8, len(seeds))) as pool:
        ws = pool.map(make, seeds)
    gen_s = time.time() - t
    log(f"[{family}] {len(ws)} instances generated ({gen_s:.1f} s)")
    keys, tgts, feats = [], [], []
    torch.cuda.synchronize()
    t = time.time()
    for j, w in enumerate(ws[:n_train]):
        p = SQDProblem(w.hamiltonian, w.keys, nq, device=dev)
        tg, _, info = ground_targets(p)
        keys.append(p.space.keys)
        tgts.append(tg)
        feats.append(energy_features(p))
        if (j + 1) % 25 == 0:
            log(f"[{family}] {j + 1}/{n_train} ground vectors")
    torch.cuda.synchronize()
    ground_s = time.time() - t
    log(f"[{family}] {n_train} FP64 ground vectors ({ground_s:.1f} s)")
    nets, train_s, final_loss = {}, {}, {}
    for ne in (0, 2):
        net = StartNet(nq, device=dev, seed=0, n_extra=ne)
        t = time.time()
        losses = net.train(keys, tgts, feats if ne else None, steps=steps, lr=3e-3,
                           log_every=max(1, steps // 3))
        torch.cuda.synchronize()
        name = "network_bits_diag" if ne else "network_bits"
        nets[name], train_s[name], final_loss[name] = net, time.time() - t, losses[-1]
        log(f"[{family}] {name}: {steps} Adam steps ({train_s[name]:.1f} s), loss {losses[-1]:.4f}")
    del keys, tgts, feats
    per = {k: {"iterations": [], "seconds": [], "start_seconds": [], "overlap": [],
               "energy_rel_err": [], "converged": 0} for k in NET_STARTS}
    for w in ws[n_train:]:
        p = SQDProblem(w.hamiltonian, w.keys, nq, device=dev)
        _, psi, info = ground_targets(p)
        e_ref = info.eigval
        for name in NET_STARTS:
            base = name.removesuffix("_precond")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if base == "gaussian":
                x0 = gaussian_start(p.n, dev, w.seed)
            elif base == "min_diag":
                x0 = p.vinit_min_diag()
            else:
                net_name = base.removesuffix("_mix")
                mix = NET_MIX if base.endswith("_mix") else 0.0
                extra = energy_features(p) if net_name == "network_bits_diag" else None
                x0 = nets[net_name].start_vector(p.space.keys, extra, mix=mix, seed=w.seed)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            _, inf = p.solve(x0, mode="fused", precision="c64", criterion="relative", tol=1e-6,
                             maxiter=3000, precond="diag" if base != name else None)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            r = per[name]
            r["iterations"].append(inf.niter)
            r["seconds"].append(t2 - t0)
            r["start_seconds"].append(t1 - t0)
            r["overlap"].append(abs(torch.vdot(psi, x0.to(psi.dtype)).item())
                                / (psi.norm().item() * x0.norm().item()))
            r["energy_rel_err"].append(abs(inf.eigval - e_ref) / abs(e_ref))
            r["converged"] += int(inf.converged)
        log(f"[{family}] seed {w.seed}: " + ", ".join(
            f"{k} {per[k]['iterations'][-1]} it" for k in NET_STARTS))
    out = {"family": family, "train": n_train, "held_out": n_test, "states": ws[-1].n,
           "generate_s": round(gen_s, 2), "ground_vectors_s": round(ground_s, 2),
           "train_s": {k: round(v, 2) for k, v in train_s.items()},
           "final_loss": {k: round(v, 5) for k, v in final_loss.items()}}
    for k, r in per.items():
        out[k] = {"iterations_mean": round(float(np.mean(r["iterations"])), 2),
                  "iterations_median": float(np.median(r["iterations"])),
                  "seconds_mean": round(float(np.mean(r["seconds"])), 4),
                  "start_seconds_mean": round(float(np.mean(r["start_seconds"])), 4),
                  "overlap_mean": round(float(np.mean(r["overlap"])), 4),
                  "energy_rel_err_max": float(np.max(r["energy_rel_err"])),
                  "converged": r["converged"]}
    g = out["gaussian"]["seconds_mean"]
    for k in NET_STARTS[1:]:
        out[k]["speedup_vs_gaussian"] = round(g / out[k]["seconds_mean"], 2)
    out["network_bits_diag_precond"]["speedup_vs_gaussian_precond"] = round(
        out["gaussian_precond"]["seconds_mean"] / out["network_bits_diag_precond"]["seconds_mean"], 2)
    return out


def svsim_block(dev, n=30, layers=10, cpu=True, cpu_qubits=18):
    """State-vector simulation (rqutils/svsim.py) of the reference example's
    GHZ circuit and of seeded random layers on n qubits in complex128, timed
    with HIP events on the launch stream; the state (2^n amplitudes) is
    resident before the timed region. HBM bytes per pass: one read and one
    write of the state (32 B per amplitude)."""
    import torch
    from rqutils_amd import workloads as W
    from rqutils_amd.svsim import apply, to_circuitxz
    state = torch.zeros(1 << n, dtype=torch.complex128, device=dev)
    stream = torch.cuda.current_stream(dev)
    out = {"num_qubits": n, "precision": "c128", "state_bytes": 16 << n}
    for name, gates in (("ghz", W.ghz_circuit(n)), (f"layers{layers}", W.layered_circuit(n, layers))):
        c = to_circuitxz(gates)
        c.num_qubits = n
        state.zero_()
        state[0] = 1.0
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        runs = apply(state, c)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        traffic = runs * 32 * (1 << n)
        rec = {"gates": len(gates), "passes": runs, "ms": round(ms, 3),
               "gates_per_s": round(len(gates) / (ms * 1e-3), 1),
               "amplitude_updates_per_s": float(len(gates) * (1 << n) / (ms * 1e-3)),
               "hbm_gbs": round(traffic / (ms * 1e-3) / 1e9, 1),
               "hbm_frac": round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
               "per_gate_pass_bytes": 32 * (1 << n) * len(gates)}
        if name == "ghz":
            a0 = float(state[0].abs().item())
            a1 = float(state[-1].abs().item())
            rest = float(torch.linalg.vector_norm(state[1:-1]).item())
            rec["check"] = {"abs_first": a0, "abs_last": a1, "norm_rest": rest,
                            "ok": abs(a0 - 2 ** -0.5) < 1e-12 and abs(a1 - 2 ** -0.5) < 1e-12
                            and rest < 1e-12}
        log(f"svsim {name} {n}q: {len(gates)} gates in {runs} passes, {ms:.2f} ms "
            f"({rec['hbm_gbs']} GB/s of state traffic)")
        out[name] = rec
    # The practical read-write streaming rate on this state for comparison: a
    # plain in-place multiply by 1.0 (no change to the values), best of three.
    best = None
    for _ in range(3):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        state.mul_(1.0)
        e1.record(stream)
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1)
        best = t if best is None else min(best, t)
    out["inplace_stream_reference"] = {
        "op": "torch in-place multiply of the state by 1.0", "ms": round(best, 3),
        "gbs": round(32 * (1 << n) / (best * 1e-3) / 1e9, 1)}
    log(f"in-place torch stream over the same state: {best:.2f} ms "
        f"({out['inplace_stream_reference']['gbs']} GB/s)")
    del state
    if cpu:
        from oracle import svsim_oracle as sv
        g = W.ghz_circuit(cpu_qubits)
        t0 = time.perf_counter()
        sv.svsim(*sv.to_circuitxz(g, cpu_qubits))
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"amplitude_updates_per_s": len(g) * (1 << cpu_qubits) / dt,
                               "cores": 1, "kind": "port",
                               "sample": f"oracle GHZ circuit on {cpu_qubits} qubits "
                                         f"({len(g)} gates, {dt:.1f} s)"}
    return out


def bench_svsim(args, dev):
    """State-vector simulation alone: ``svsim_block`` at 30 qubits."""
    res = svsim_block(dev, cpu=not args.no_cpu)
    lay = res["layers10"]
    record = {"metric": "svsim amplitude updates/s (gate x amplitude), complex128",
              "value": lay["amplitude_updates_per_s"], "unit": "amplitude updates/s",
              "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": lay["ms"],
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
              "dtype": "c128", "data": "synthetic (GHZ and seeded random layers)",
              "config": {"workload": "svsim", "num_qubits": res["num_qubits"],
                         "parallelism": "single GPU"},
              "roofline": {"bound": "hbm", "achieved": lay["hbm_gbs"], "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": lay["hbm_frac"], "traffic": None,
                           "kernel": "svsim_reg_kernel<double2,11,3>"},
              "svsim": res}
    if "cpu_baseline" in res:
        record["cpu_baseline"] = {**res["cpu_baseline"], "value": res["cpu_baseline"][
            "amplitude_updates_per_s"], "unit": "amplitude updates/s"}
    print(json.dumps(record), flush=True)


def bench_netstart(args, dev):
    """The network-start protocol alone: ``--net-train`` training and
    ``--net-test`` held-out 28-qubit Hamiltonians at 2^20 states."""
    res = network_start(dev, "heisenberg28_2^20", args.net_train, args.net_test,
                        args.net_steps)
    best = res["network_bits_diag"]
    record = {"metric": METRIC, "value": best["seconds_mean"], "unit": "s",
              "n_gpus": 1, "steps": args.net_steps, "warmup": 0,
              "ms_per_step": None, "higher_is_better": False, "scaling": "weak",
              "vs_baseline": None, "dtype": "bf16 network, c64 LOBPCG (fp64 Gram/RR)",
              "data": "synthetic (seeded generator, rqutils_amd.workloads.heisenberg28)",
              "config": {"workload": "netstart", "family": res["family"],
                         "train": args.net_train, "held_out": args.net_test,
                         "parallelism": "single GPU"},
              "network_start": res}
    print(json.dumps(record), flush=True)


def bench_next_rows(args, dev):
    """The rows next to the hot path, timed against their CPU ports: the
    full-space product PauliSumXZ.matmul (symplectic.py:131-150) on the whole
    2^22 space of a 22-qubit chain with 1,000 disorder strings (complex128;
    reported as products/s and gathers/s, beside the least HBM traffic of one
    fused pass and the bytes the reference's per-group passes would move),
    and hproj (sqd.py:251-296) on 2^18 sampled
    states of a 24-qubit chain (one hit-table build, unpacking and the scipy
    CSR assembly). CPU: the oracle's NumPy restatements on a sample of 3
    groups, scaled by the group count."""
    import torch
    from oracle import sqd_oracle as so
    from rqutils_amd import workloads as W
    from rqutils_amd.operator import keys_to_states
    from rqutils_amd.paulis.symplectic import PauliSumXZ
    from rqutils_amd.sqd import hproj
    rec = {}
    nq = 22
    w = W.heisenberg28(seed=2, n_states=1 << 10, n_random=1000, num_qubits=nq)
    ham = PauliSumXZ.from_paulisum(w.hamiltonian, force_real=True, add_padding=True)
    g = torch.Generator(device=dev).manual_seed(3)
    v = torch.complex(torch.randn(1 << nq, generator=g, device=dev, dtype=torch.float64),
                      torch.randn(1 << nq, generator=g, device=dev, dtype=torch.float64))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = ham.matmul(v)
    torch.cuda.synchronize()
    first_s = time.perf_counter() - t0
    stream = torch.cuda.current_stream(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        out = ham.matmul(v)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    groups = int(ham.x.shape[0])
    # The reference's per-group passes move 48 B per amplitude and group
    # through memory; the fused kernel keeps each output in registers, so its
    # least HBM traffic is one read of the right-hand side and one write of
    # the output (32 B per amplitude), and the group gathers hit the cache.
    algo = 48 * (1 << nq) * groups
    min_bytes = 32 * (1 << nq)
    gathers = (1 << nq) * groups
    vh = v.cpu().numpy()
    t0 = time.perf_counter()
    part = so.fullspace_matmul(ham.x, ham.z, ham.c, nq, vh, groups=[0, 1, 2])
    cpu_s = (time.perf_counter() - t0) * groups / 3
    got = ham.matmul(v).cpu().numpy()
    # the sampled groups' share of the product, computed on the device
    ham3 = PauliSumXZ(ham.x[:3], ham.z[:3], ham.c[:3], nq, ham.padding)
    share = ham3.matmul(v).cpu().numpy()
    diff = float(np.linalg.norm(share - part) / np.linalg.norm(part))
    rec["fullspace_matmul"] = {
        "num_qubits": nq, "groups": groups, "terms": len(w.labels), "precision": "c128",
        "first_call_s": round(first_s, 3), "ms": round(ms, 3),
        "products_per_s": round(1e3 / ms, 2),
        "reference_pass_bytes": algo,
        "reference_pass_equivalent_gbs": round(algo / (ms * 1e-3) / 1e9, 1),
        "min_hbm_bytes": min_bytes,
        "min_hbm_gbs": round(min_bytes / (ms * 1e-3) / 1e9, 1),
        "gathers_per_s": round(gathers / (ms * 1e-3), 1),
        "cpu_port_s_extrapolated": round(cpu_s, 2),
        "cpu_sample": f"oracle fullspace_matmul over 3 of {groups} groups, x{groups / 3:.1f}",
        "rel_diff_vs_oracle_on_3_groups": diff, "norm": float(np.linalg.norm(got))}
    log(f"full-space matmul 2^{nq} x {groups} groups: {ms:.2f} ms per product (first call "
        f"{first_s:.2f} s with setup), {gathers / (ms * 1e-3) / 1e9:.0f} G gathers/s "
        f"(the reference's passes would move {algo / (ms * 1e-3) / 1e12:.1f} TB/s); CPU port "
        f"{cpu_s:.1f} s (extrapolated); 3-group share vs oracle {diff:.1e}")
    w2 = W.heisenberg28(seed=2, n_states=1 << 18, n_random=1000, num_qubits=24)
    states = keys_to_states(w2.keys, 24)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m = hproj(w2.hamiltonian, states, unique_states=True)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    xs, zs, cs, _ = so.from_paulisum(w2.labels, w2.coeffs, force_real=True, add_padding=True)
    packed = so.pack_states(states)
    t0 = time.perf_counter()
    r, c, _ = so.hproj_coo(xs[:3], zs[:3], cs[:3], packed)
    cpu_h = (time.perf_counter() - t0) * xs.shape[0] / 3
    rec["hproj"] = {"num_qubits": 24, "states": int(w2.n), "groups": int(xs.shape[0]),
                    "nnz": int(m.nnz), "seconds": round(gpu_s, 3),
                    "cpu_port_s_extrapolated": round(cpu_h, 2),
                    "cpu_sample": f"oracle hproj_coo over 3 of {xs.shape[0]} groups, "
                                  f"x{xs.shape[0] / 3:.1f}"}
    log(f"hproj 2^18 states x {xs.shape[0]} groups: {gpu_s:.2f} s ({m.nnz} nonzeros, scipy CSR "
        f"included); CPU port {cpu_h:.1f} s (extrapolated)")
    print(json.dumps({"metric": METRIC, "value": rec["fullspace_matmul"]["ms"], "unit": "ms",
                      "n_gpus": 1, "steps": args.steps, "warmup": 1,
                      "ms_per_step": rec["fullspace_matmul"]["ms"], "higher_is_better": False,
                      "scaling": "weak", "vs_baseline": None, "dtype": "c128",
                      "data": "synthetic (seeded generator, rqutils_amd.workloads.heisenberg28)",
                      "config": {"workload": "nextrows", "parallelism": "single GPU"},
                      "next_rows": rec}), flush=True)


def bench_batch(args, dev):
    """Batched workload: many independent instances, one launch per step."""
    import torch
    from rqutils_amd import workloads as W
    from rqutils_amd.batch import SQDBatch, autotune_batch
    from rqutils_amd.sqd import SQDProblem

    t = time.time()
    ws = [W.batch_instance(seed) for seed in range(args.instances)]
    insts = [(w.hamiltonian, w.keys, w.num_qubits) for w in ws]
    log(f"batch of {len(ws)} instances, {sum(w.n for w in ws)} rows, qubits "
        f"{min(w.num_qubits for w in ws)}-{max(w.num_qubits for w in ws)} ({time.time() - t:.1f} s)")
    t = time.time()
    best, table = autotune_batch(insts, device=dev)
    log(f"batch autotune {table} -> cap={best[0]} rows={best[1]} split={best[2]} "
        f"compact={best[4]} dense={best[5]} ({time.time() - t:.1f} s)")
    b = SQDBatch(insts, device=dev, bucket_cap=best[0], rows_per_thread=best[1],
                 group_split=best[2], compact=best[4], dense=best[5])
    # algorithmic bytes, summed over instances
    s_v = 8
    b_mv = 0
    hits_total = 0
    for sp, op in zip(b.spaces, b.ops):
        hits = 0
        for x in op.host.xmask:
            hits += int(torch.count_nonzero(sp.xsource(int(x)) >= 0).item())
        hits_total += hits
        b_mv += sp.n * (8 + 2 * s_v) + op.n_groups * sp.n * 8 + hits * s_v
    g = torch.Generator(device="cpu").manual_seed(1234)
    v = torch.complex(torch.randn(b.n, generator=g), torch.randn(b.n, generator=g)).to(
        torch.complex64).to(dev)
    out = torch.empty_like(v)
    for _ in range(args.warmup):
        b.matvec(v, out)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        b.matvec(v, out)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    achieved = b_mv / (ms * 1e-3) / 1e9
    log(f"batched matvec: {ms:.3f} ms, {achieved:.1f} GB/s algorithmic")

    x0 = torch.complex(torch.randn(b.n, generator=g, dtype=torch.float64),
                       torch.randn(b.n, generator=g, dtype=torch.float64)).to(dev)
    res64 = b.solve(x0, precision="c64", tol=1e-6, maxiter=2000, use_graph=True)
    log(f"batched fused c64 (graph): {res64.seconds:.3f} s, "
        f"{len(ws) / res64.seconds:.1f} solves/s")
    res_nog = b.solve(x0, precision="c64", tol=1e-6, maxiter=2000, use_graph=False)
    log(f"batched fused c64 (no graph): {res_nog.seconds:.3f} s")
    ref = b.solve(x0, precision="c128", tol=1e-10, maxiter=4000, use_graph=True)
    errs = [abs(a.eigval - r.eigval) / abs(r.eigval) for a, r in zip(res64.infos, ref.infos)]
    # sequential baseline: the same fused solver, one instance per call
    probs = [SQDProblem(w.hamiltonian, w.keys, w.num_qubits, device=dev, bucket_cap=best[0],
                        rows_per_thread=best[1], group_split=best[2], compact=best[4],
                        dense=best[5])
             for w in ws]
    starts = [b.segment(x0, i).contiguous() for i in range(len(ws))]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for prob, xs in zip(probs, starts):
        prob.solve(xs, mode="fused", precision="c64", criterion="relative", tol=1e-6,
                   maxiter=2000)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    log(f"sequential per-instance fused c64 solves: {t_seq:.3f} s")
    # the batch's hit table: build once, then the same product and solve through it
    from rqutils_amd import _lib
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cinfo = b.build_cache(_lib.C64)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    b.use_cache = True
    b.matvec(v, out)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        b.matvec(v, out)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_c = ev0.elapsed_time(ev1) / args.steps
    res_c = b.solve(x0, precision="c64", tol=1e-6, maxiter=2000, use_graph=True)
    b.use_cache = False
    errs_c = [abs(a.eigval - r.eigval) / abs(r.eigval) for a, r in zip(res_c.infos, ref.infos)]
    log(f"batch hit table: built in {build_s:.3f} s ({cinfo['entries']} entries); product "
        f"{ms_c:.3f} ms; solves {res_c.seconds:.3f} s ({len(ws) / res_c.seconds:.1f} solves/s), "
        f"max energy error {max(errs_c):.1e}")
    record = {
        "metric": METRIC,
        "value": round(achieved, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 two-sum accumulate)",
        "data": "synthetic (seeded generator, rqutils_amd.workloads.batch_instance)",
        "config": {"workload": f"batch{len(ws)}", "instances": len(ws), "states": b.n,
                   "hits": hits_total, "bucket_cap": best[0], "rows_per_thread": best[1],
                   "group_split": best[2], "compact_hits": bool(best[4]),
                   "dense_lookup_instances": b.n_dense, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "bucket_join_kernel (batched)",
                     "bytes_per_launch": b_mv},
        "batch": {"solves_per_s": round(len(ws) / res64.seconds, 2),
                  "seconds_graph": round(res64.seconds, 3),
                  "seconds_no_graph": round(res_nog.seconds, 3),
                  "sequential_seconds": round(t_seq, 3),
                  "converged": sum(i.converged for i in res64.infos),
                  "iterations_median": float(np.median([i.niter for i in res64.infos])),
                  "energy_rel_err_vs_fp64_1e-10": {"max": max(errs), "median": float(np.median(errs))}},
        "batch_hit_table": {"build_s": round(build_s, 4), "entries": cinfo["entries"],
                            "table_bytes": cinfo["bytes"], "product_ms": round(ms_c, 4),
                            "solves_per_s": round(len(ws) / res_c.seconds, 2),
                            "seconds_graph": round(res_c.seconds, 3),
                            "converged": sum(i.converged for i in res_c.infos),
                            "energy_rel_err_vs_fp64_1e-10": {
                                "max": max(errs_c), "median": float(np.median(errs_c))}},
    }
    print(json.dumps(record), flush=True)


if __name__ == "__main__":
    main()
