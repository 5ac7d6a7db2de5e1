"""Benchmark of the SQD hot path on one MI355X.

Headline (default, ``--workload random40``): the 40-qubit instance (20,000
random Pauli strings of weight 2-4 over 2^24 uniform 20-ones-sector states).
A step is one matvec of the FP32-compensated bucket-join kernel over the
whole subspace; ``value`` is its algorithmic HBM rate B_mv / t_mv with

    B_mv = N (8 + 2 s_v) + J_s N 8 + H s_v

(own key + v read + Hv write per row, one 8-byte key per searched
(row, off-diagonal group), one gathered element per hit; s_v = 8 bytes for
complex64; J_s = off-diagonal groups searched after parity pruning; H = hits).
The same JSON line carries the roofline object, the time to a 1e-6 relative
residual of every solver variant next to the plain FP64 reference-structured
LOBPCG (setup included in ``seconds``), the CPU baseline (the oracle's C
restatement of the reference's sort-join dataflow on all host cores, over a
seeded group sample, extrapolated), and the other workloads as extra keys:
``c1_small16`` (FP64 time to 1e-8), ``c2_heisenberg28``, ``c4_batch64``,
``c5_network_start`` (200 training / 50 held-out Hamiltonians, 28-input
network), ``intake`` and ``svsim``.

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
# Host threads of the CPU baseline: the box's CPU share for one GPU is 16,
# whatever the affinity mask shows.
CPU_THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sync():
    import torch
    torch.cuda.synchronize()


def count_hits(prob) -> int:
    """Total (row, off-diagonal group) pairs whose target is in the subspace."""
    import torch
    total = 0
    for x in prob.op.host.xmask:
        total += int(torch.count_nonzero(prob.space.xsource(int(x)) >= 0).item())
    return total




def read_limiter(path):
    """The measured limiter (resource, fraction used) from a committed rocprof
    summary's ``limiter:`` line, or None."""
    if not os.path.exists(path):
        return None
    m = re.search(r"limiter:\s*(\S+)\s+([0-9.eE+]+)", open(path).read())
    return {"resource": m.group(1), "frac": float(m.group(2)),
            "source": os.path.relpath(path, ROOT)} if m else None


def read_traffic(path):
    """HBM bytes per launch from a committed rocprof summary (or None)."""
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None




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
        op.drop_cache()             # each precision's build timed from scratch
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




def build_problem(w, dev, tune=True, launch=None):
    """SQDProblem of a workload whose keys are already resident on the device.

    The autotuner (one-time per problem, memoised) picks the launch first;
    ``setup_s`` then times what every new problem pays with that launch:
    grouping the Pauli sum on the host and building the bucket directories and
    operator tables on the device. ``launch`` (a dict of ``SQDProblem``
    keywords) fixes the launch instead of the autotuner."""
    import torch
    from rqutils_amd.sqd import SQDProblem
    keys = torch.from_numpy(np.ascontiguousarray(w.keys).view(np.int64)).to(dev)
    sync()
    tuned = None
    if tune and not launch:
        t = time.time()
        tuned = SQDProblem(w.hamiltonian, keys, w.num_qubits, device=dev, tune="auto").tuned
        log(f"[{w.name}] autotune -> cap={tuned.bucket_cap} dense_bits={tuned.dense_bits} "
            f"rows={tuned.rows_per_thread} split={tuned.group_split} compact={tuned.compact} "
            f"fine_shift={tuned.fine_shift} "
            f"({time.time() - t:.1f} s)")
    if tuned is None:
        # first use of the kernels and the host code (the autotuner's runs do it otherwise)
        SQDProblem(w.hamiltonian, keys, w.num_qubits, device=dev, **(launch or {}))
    sync()
    t0 = time.perf_counter()
    if launch:
        prob = SQDProblem(w.hamiltonian, keys, w.num_qubits, device=dev, **launch)
    elif tuned is not None:
        prob = SQDProblem(w.hamiltonian, keys, w.num_qubits, device=dev,
                          group_split=tuned.group_split, compact=tuned.compact,
                          **tuned.space_kw())
    else:
        prob = SQDProblem(w.hamiltonian, keys, w.num_qubits, device=dev)
    sync()
    setup_s = time.perf_counter() - t0
    prob.tuned = tuned
    return prob, setup_s


def matvec_block(prob, w, args, dev):
    """The bucket-join matvec in complex64 (the step) and complex128, timed with
    HIP events on the launch stream over resident vectors."""
    import torch
    sp, op = prob.space, prob.op
    n = w.n
    j_s = op.n_groups
    hits = count_hits(prob)
    b_mv = n * (8 + 2 * 8) + j_s * n * 8 + hits * 8
    log(f"[{w.name}] space: bucket_bits={sp.bucket_bits} fine_bits={sp.fine_bits} "
        f"items={sp.n_items}; groups searched={j_s} (pruned {op.host.pruned_groups}), "
        f"classes={op.n_classes}; hits={hits} ({hits / n:.2f}/row), B_mv={b_mv / 1e9:.3f} GB")
    g = torch.Generator(device="cpu").manual_seed(1234)
    v = torch.complex(torch.randn(n, generator=g), torch.randn(n, generator=g)).to(
        torch.complex64).to(dev)
    out = torch.empty_like(v)
    for _ in range(args.warmup):
        prob.matvec(v, out)
    sync()
    stream = torch.cuda.current_stream(dev)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        prob.matvec(v, out)
    ev1.record(stream)
    sync()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    achieved = b_mv / (ms * 1e-3) / 1e9
    log(f"[{w.name}] matvec c64: {ms:.3f} ms, {achieved:.1f} GB/s algorithmic "
        f"({achieved / HBM_PEAK_GBS:.1%})")
    v128 = v.to(torch.complex128)
    o128 = torch.empty_like(v128)
    prob.matvec(v128, o128)
    sync()
    reps = max(3, args.steps // 4)
    ev0.record(stream)
    for _ in range(reps):
        prob.matvec(v128, o128)
    ev1.record(stream)
    sync()
    ms128 = ev0.elapsed_time(ev1) / reps
    b128 = n * (8 + 2 * 16) + j_s * n * 8 + hits * 16
    diff = (out.to(torch.complex128) - o128).norm().item() / o128.norm().item()
    log(f"[{w.name}] matvec c128: {ms128:.3f} ms; c64 vs c128 rel diff {diff:.2e}")
    res = {"ms": round(ms, 4), "achieved_gbs": round(achieved, 2),
           "hbm_frac": round(achieved / HBM_PEAK_GBS, 4), "matvecs_per_s": round(1e3 / ms, 2),
           "bytes_per_launch": b_mv, "hits": hits, "groups_searched": j_s,
           "groups_pruned_odd": op.host.pruned_groups, "classes": op.n_classes,
           "matvec_c128_ms": round(ms128, 4),
           "matvec_c128_gbs": round(b128 / (ms128 * 1e-3) / 1e9, 2),
           "c64_vs_c128_rel_diff": diff, "wall_s_timed": round(wall, 3),
           "launch": {"bucket_bits": sp.bucket_bits, "fine_bits": sp.fine_bits,
                      "rows_per_thread": sp.rows_per_thread, "group_split": op.tune.group_split,
                      "lookup": "dense bitmap" if sp.dense else "search",
                      "compact_hits": bool(op.tune.compact)},
           "kernel": f"bucket_join_kernel<float2,{'uint32_t' if sp.bucket_bits < 32 else 'uint64_t'},"
                     f"{sp.rows_per_thread},false,{str(sp.dense).lower()},"
                     f"{str(bool(op.tune.compact)).lower()}>"}
    return res, (v, out, o128, hits)


TTS_RUNS = {
    # name: (mode, precision, precond, cached)
    "reference_c128": ("reference", "c128", None, False),      # the plain FP64 baseline
    "fused_c128": ("fused", "c128", None, False),
    "fused_c64": ("fused", "c64", None, False),
    "fused_c64_diag_precond": ("fused", "c64", "diag", False),
    "cached_reference_c128": ("reference", "c128", None, True),
    "cached_fused_c64": ("fused", "c64", None, True),
    "cached_fused_c64_diag_precond": ("fused", "c64", "diag", True),
}


GRAPH_RUNS = ("fused_c64", "cached_fused_c64_diag_precond")


def graph_block(prob, x0, runs, tol, reps=3):
    """The fused solve launched directly and with each refresh period replayed
    as one captured HIP graph (use_graph): best of ``reps`` wall times, and
    whether the two give the same eigenvalue bits and iteration count."""
    out = {}
    for name in runs:
        mode, prec, precond, cached = TTS_RUNS[name]
        prob.op.use_cache = cached
        rec = {}
        for graph in (False, True):
            best, info = None, None
            for _ in range(reps):
                sync()
                t0 = time.perf_counter()
                _, info = prob.solve(x0, mode=mode, precision=prec, precond=precond,
                                     criterion="relative", tol=tol, maxiter=3000,
                                     use_graph=graph)
                sync()
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            rec["graph" if graph else "direct"] = {"solve_s": round(best, 5),
                                                   "iterations": info.niter,
                                                   "energy": info.eigval}
        rec["identical"] = (rec["graph"]["energy"] == rec["direct"]["energy"] and
                            rec["graph"]["iterations"] == rec["direct"]["iterations"])
        rec["graph_speedup"] = round(rec["direct"]["solve_s"] / rec["graph"]["solve_s"], 3)
        out[name] = rec
    return out


def time_to_solution(prob, w, dev, setup_s, runs=None, tol=1e-6):
    """Time to a ``tol`` relative residual from one seeded Gaussian start.

    ``seconds`` = ``setup_s`` (the problem built from resident keys) + the
    hit-table build for the cached runs (``table_build_s``, built once per
    precision and charged to every run that reads it) + ``solve_s``. The
    reference energy is the FP64 fused solve at a 1e-10 residual."""
    import torch
    from rqutils_amd import _lib
    from rqutils_amd.startnet import gaussian_start
    runs = list(TTS_RUNS) if runs is None else runs
    x0 = gaussian_start(w.n, dev, 77)
    op = prob.op
    op.drop_cache()
    build = {}
    precs = sorted({TTS_RUNS[k][1] for k in runs if TTS_RUNS[k][3]})
    for prec in precs:
        # each precision's table timed on its own (a second precision reuses
        # the first one's layout), then all of them built for the runs
        op.drop_cache()
        sync()
        t0 = time.perf_counter()
        op.build_cache(_lib.C128 if prec == "c128" else _lib.C64)
        sync()
        build[prec] = time.perf_counter() - t0
    for prec in precs:
        op.build_cache(_lib.C128 if prec == "c128" else _lib.C64)
    res = {"setup_s": round(setup_s, 4), "table_build_s": {k: round(v, 4) for k, v in build.items()},
           "start": "seeded complex Gaussian (seed 77)", "tol": tol}
    for name in runs:
        mode, prec, precond, cached = TTS_RUNS[name]
        op.use_cache = cached
        sync()
        t0 = time.perf_counter()
        _, info = prob.solve(x0, mode=mode, precision=prec, precond=precond,
                             criterion="relative", tol=tol, maxiter=3000)
        sync()
        dt = time.perf_counter() - t0
        total = setup_s + dt + (build.get(prec, 0.0) if cached else 0.0)
        res[name] = {"seconds": round(total, 4), "solve_s": round(dt, 4),
                     "iterations": info.niter, "matvecs": info.matvecs, "energy": info.eigval,
                     "converged": info.converged, "status": info.status,
                     "resid_rel": info.resid_norm / abs(info.eigval)}
        log(f"[{w.name}] {name}: {total:.3f} s ({dt:.3f} s solve), {info.niter} it, "
            f"{info.matvecs} matvecs, E={info.eigval:.12f}")
    if all(k in runs for k in GRAPH_RUNS):
        res["graph_replay"] = graph_block(prob, x0, GRAPH_RUNS, tol)
        log(f"[{w.name}] graph replay speedup: " + ", ".join(
            f"{k} {v['graph_speedup']}" for k, v in res["graph_replay"].items()))
    op.use_cache = True
    if _lib.C128 not in op._caches:
        op.build_cache(_lib.C128)
    _, ref = prob.solve(x0, mode="fused", precision="c128", criterion="relative", tol=1e-10,
                        maxiter=5000)
    op.use_cache = False
    e_ref = ref.eigval
    res["fp64_reference_energy_1e-10"] = e_ref
    for name in runs:
        res[name]["energy_rel_err_vs_fp64_1e-10"] = abs(res[name]["energy"] - e_ref) / abs(e_ref)
    if "reference_c128" in runs:
        base = res["reference_c128"]["seconds"]
        res["speedup_vs_plain_fp64"] = {k: round(base / res[k]["seconds"], 2) for k in runs
                                        if k != "reference_c128"}
        best = min((k for k in runs if k != "reference_c128"), key=lambda k: res[k]["seconds"])
        res["best"] = {"run": best, "seconds": res[best]["seconds"],
                       "speedup_vs_plain_fp64": round(base / res[best]["seconds"], 2),
                       "energy_rel_err_vs_fp64_1e-10": res[best]["energy_rel_err_vs_fp64_1e-10"]}
    return res


def cpu_baseline(w, b_mv, n_groups, threads):
    """The oracle's C restatement of the reference's matvec dataflow (per
    group a stable sort-join of [S; S ^ x] for the source indices, the
    per-term diagonal, the gather: cache_level (0, 0), since the source-index
    cache of the default (1, 0) takes 4 J N bytes, 200 GB here) on ``threads``
    host threads over a seeded sample of the reference's X groups, all rows,
    extrapolated by the reference's group count."""
    from oracle import ref_dataflow as rd
    tables = rd.RefTables(w.labels, w.coeffs, w.num_qubits)
    j = tables.n_groups
    rng = np.random.default_rng(0)
    groups = np.sort(rng.choice(j, size=min(n_groups, j), replace=False)).astype(np.int32)
    v = np.random.default_rng(1).normal(size=w.n) + 0j
    t0 = time.perf_counter()
    _, used = rd.matvec(tables, w.keys, v, groups=groups, nthreads=threads)
    dt = time.perf_counter() - t0
    t_full = dt * j / groups.size
    log(f"cpu baseline: {groups.size} of {j} groups on {used} threads in {dt:.1f} s "
        f"-> {t_full:.1f} s per matvec (extrapolated)")
    return {"value": round(b_mv / t_full / 1e9, 3), "unit": "GB/s", "cores": used,
            "kind": "port",
            "sample": f"C/OpenMP restatement of the reference's sort-join matvec (cache_level "
                      f"(0,0)) over {groups.size} of the reference's {j} X groups, all {w.n} "
                      f"rows, {dt:.1f} s on {used} threads, extrapolated x{j / groups.size:.1f}",
            "seconds_per_matvec_extrapolated": round(t_full, 2),
            "matvecs_per_s": round(1.0 / t_full, 6)}


def c1_block(dev):
    """16-qubit, 200-term, 4,096-state workload: FP64 time to a 1e-8 relative
    residual from a seeded Gaussian start, GPU (plain reference iteration and
    fused) against the NumPy oracle's reference iteration (cached source
    indices, one core); energies against dense eigh."""
    import torch
    from oracle import locg_oracle as lo, sqd_oracle as so
    from rqutils_amd import workloads as W
    from rqutils_amd.operator import keys_to_states
    from rqutils_amd.startnet import gaussian_start
    w = W.small16()
    prob, setup_s = build_problem(w, dev, tune=False)
    x0 = gaussian_start(w.n, dev, 7)
    res = {"states": w.n, "terms": len(w.labels), "tol": 1e-8, "setup_s": round(setup_s, 4)}
    for name, kw in (("reference_c128", dict(mode="reference")), ("fused_c128", dict(mode="fused"))):
        for _ in range(2):              # the first call warms the kernels up
            sync()
            t0 = time.perf_counter()
            _, info = prob.solve(x0, precision="c128", criterion="relative", tol=1e-8,
                                 maxiter=3000, **kw)
            sync()
            dt = time.perf_counter() - t0
        res[name] = {"seconds": round(setup_s + dt, 5), "solve_s": round(dt, 5),
                     "iterations": info.niter, "energy": info.eigval,
                     "converged": info.converged}
    from rqutils_amd import _lib
    prob.op.build_cache(_lib.C64)
    res["graph_replay"] = graph_block(prob, x0, GRAPH_RUNS, 1e-6)
    prob.op.drop_cache()
    x, z, c, _ = so.from_paulisum(w.labels, w.coeffs, force_real=True, add_padding=True)
    packed = so.pack_states(keys_to_states(w.keys, w.num_qubits))
    t0 = time.perf_counter()
    xs = np.stack([so.get_xsource(x[g], packed) for g in range(x.shape[0])])
    mv = lambda u: so.apply_h_x_cached(u, xs, z, c, packed)  # noqa: E731
    th, _, it = lo.ground_locg(mv, x0.cpu().numpy(), maxiter=3000, rel_tol=1e-8)
    cpu_s = time.perf_counter() - t0
    h = np.zeros((w.n, w.n), complex)
    rows, cols, vals = so.hproj_coo(x, z, c, packed)
    np.add.at(h, (rows, cols), vals)
    e0 = float(np.linalg.eigvalsh(h)[0])
    res["cpu_oracle"] = {"seconds": round(cpu_s, 3), "iterations": it, "energy": th,
                         "cores": 1, "kind": "port",
                         "sample": "NumPy reference iteration, source indices cached (the "
                                   "reference's default cache_level (1,0)), setup included"}
    res["dense_eigh_energy"] = e0
    for k in ("reference_c128", "fused_c128", "cpu_oracle"):
        res[k]["energy_rel_err_vs_eigh"] = abs(res[k]["energy"] - e0) / abs(e0)
    log(f"[small16] FP64 to 1e-8: GPU reference {res['reference_c128']['seconds']:.4f} s, "
        f"fused {res['fused_c128']['seconds']:.4f} s; CPU oracle {cpu_s:.2f} s")
    del prob
    torch.cuda.empty_cache()
    return res


def c2_block(args, dev):
    """28-qubit chain + 2,000 disorder strings, 2^22 domain-wall states."""
    import torch
    from rqutils_amd import workloads as W
    t = time.time()
    w = W.heisenberg28()
    log(f"[heisenberg28] generated ({time.time() - t:.1f} s)")
    prob, setup_s = build_problem(w, dev)
    mv, (v, out, o128, hits) = matvec_block(prob, w, args, dev)
    res = {"states": w.n, "terms": len(w.labels), "matvec": mv}
    res["hit_table"] = hit_table_block(prob, v, out, o128, hits, args.steps, dev)
    prob.op.drop_cache()
    res["time_to_solution"] = time_to_solution(
        prob, w, dev, setup_s, runs=["reference_c128", "fused_c64", "fused_c64_diag_precond",
                                     "cached_fused_c64", "cached_fused_c64_diag_precond"])
    prob.op.drop_cache()
    if not args.no_intake:
        res["intake"] = intake_block(w, dev)
    del prob
    torch.cuda.empty_cache()
    return res


def c4_block(args, dev):
    """64 instances (20-32 qubits, 500-5,000 terms, 2^12-2^20 states) in one
    batch: the batched matvec, the batched fused FP32/FP64 solves (HIP graph
    and direct launches, joined and through the batch's hit table) against
    the same fused solver called once per instance; energies against the
    batched FP64 solve at 1e-10."""
    import torch
    from rqutils_amd import _lib, workloads as W
    from rqutils_amd.batch import SQDBatch, autotune_batch
    from rqutils_amd.sqd import SQDProblem
    t = time.time()
    ws = [W.batch_instance(seed) for seed in range(args.instances)]
    insts = [(w.hamiltonian, w.keys, w.num_qubits) for w in ws]
    log(f"[batch] {len(ws)} instances, {sum(w.n for w in ws)} rows ({time.time() - t:.1f} s)")
    t = time.time()
    best, table = autotune_batch(insts, device=dev, caps=(2048, 8192), splits=(1, 2),
                                 dense=(False,))
    log(f"[batch] autotune -> cap={best[0]} rows={best[1]} split={best[2]} compact={best[4]} "
        f"({time.time() - t:.1f} s)")
    sync()
    t0 = time.perf_counter()
    b = SQDBatch(insts, device=dev, bucket_cap=best[0], rows_per_thread=best[1],
                 group_split=best[2], compact=best[4], dense=best[5])
    sync()
    setup_s = time.perf_counter() - t0
    b_mv = 0
    hits_total = 0
    for sp, op in zip(b.spaces, b.ops):
        hits = sum(int(torch.count_nonzero(sp.xsource(int(x)) >= 0).item())
                   for x in op.host.xmask)
        hits_total += hits
        b_mv += sp.n * (8 + 2 * 8) + op.n_groups * sp.n * 8 + hits * 8
    g = torch.Generator(device="cpu").manual_seed(1234)
    v = torch.complex(torch.randn(b.n, generator=g), torch.randn(b.n, generator=g)).to(
        torch.complex64).to(dev)
    out = torch.empty_like(v)
    for _ in range(args.warmup):
        b.matvec(v, out)
    sync()
    stream = torch.cuda.current_stream(dev)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        b.matvec(v, out)
    ev1.record(stream)
    sync()
    ms = ev0.elapsed_time(ev1) / args.steps
    achieved = b_mv / (ms * 1e-3) / 1e9
    x0 = b.start_vector("gaussian", seed=5)
    ref = b.solve(x0, precision="c128", tol=1e-10, maxiter=4000)
    res = {"instances": len(ws), "rows": b.n, "hits": hits_total, "setup_s": round(setup_s, 3),
           "launch": {"bucket_cap": best[0], "rows_per_thread": best[1], "group_split": best[2],
                      "compact_hits": bool(best[4]), "autotune_table": table},
           "matvec": {"ms": round(ms, 4), "achieved_gbs": round(achieved, 2),
                      "hbm_frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_launch": b_mv}}

    def record(name, r, extra_s=0.0):
        errs = [abs(a.eigval - q.eigval) / abs(q.eigval) for a, q in zip(r.infos, ref.infos)]
        total = setup_s + extra_s + r.seconds
        res[name] = {"seconds": round(total, 4), "solve_s": round(r.seconds, 4),
                     "solves_per_s": round(len(ws) / total, 2),
                     "converged": sum(i.converged for i in r.infos),
                     "iterations_median": float(np.median([i.niter for i in r.infos])),
                     "energy_rel_err_vs_fp64_1e-10": {"max": max(errs),
                                                      "median": float(np.median(errs))}}
        log(f"[batch] {name}: {total:.3f} s ({len(ws) / total:.1f} solves/s), max energy "
            f"error {max(errs):.1e}")

    record("fused_c64_graph", b.solve(x0, precision="c64", tol=1e-6, maxiter=3000, use_graph=True))
    record("fused_c64_direct", b.solve(x0, precision="c64", tol=1e-6, maxiter=3000,
                                       use_graph=False))
    sync()
    t0 = time.perf_counter()
    cinfo = b.build_cache(_lib.C64)
    sync()
    build_s = time.perf_counter() - t0
    b.use_cache = True
    b.matvec(v, out)
    sync()
    ev0.record(stream)
    for _ in range(args.steps):
        b.matvec(v, out)
    ev1.record(stream)
    sync()
    res["hit_table"] = {"build_s": round(build_s, 4), "entries": cinfo["entries"],
                        "table_bytes": cinfo["bytes"],
                        "product_ms": round(ev0.elapsed_time(ev1) / args.steps, 4)}
    record("cached_fused_c64_graph", b.solve(x0, precision="c64", tol=1e-6, maxiter=3000,
                                             use_graph=True), build_s)
    record("cached_fused_c64_direct", b.solve(x0, precision="c64", tol=1e-6, maxiter=3000,
                                              use_graph=False), build_s)
    b.use_cache = False
    # the same fused solver called once per instance (problems built beforehand)
    probs = [SQDProblem(w.hamiltonian, w.keys, w.num_qubits, device=dev, bucket_cap=best[0],
                        rows_per_thread=best[1], group_split=best[2], compact=best[4],
                        dense=best[5]) for w in ws]
    sync()
    t0 = time.perf_counter()
    for i, prob in enumerate(probs):
        prob.solve(b.segment(x0, i).contiguous(), mode="fused", precision="c64",
                   criterion="relative", tol=1e-6, maxiter=3000)
    sync()
    res["sequential_fused_c64_s"] = round(time.perf_counter() - t0, 3)
    best_name = min((k for k in res if k.startswith(("fused", "cached"))),
                    key=lambda k: res[k]["seconds"])
    res["best"] = {"run": best_name, "solves_per_s": res[best_name]["solves_per_s"],
                   "speedup_vs_sequential": round(res["sequential_fused_c64_s"] /
                                                  res[best_name]["solve_s"], 2)}
    log(f"[batch] sequential per-instance solves {res['sequential_fused_c64_s']:.3f} s")
    del probs, b
    torch.cuda.empty_cache()
    return res


def _make_family_instance(seed):
    from rqutils_amd import workloads as W
    return W.heisenberg28(seed=seed, n_states=1 << 20)


C5_STARTS = {
    # name: (start, mix, precond); every one of these is a product start
    "gaussian": ("gaussian", None, None),
    "gaussian_precond": ("gaussian", None, "diag"),
    "min_diag_mix": ("min_diag", 0.1, None),
    "min_diag_mix_precond": ("min_diag", 0.1, "diag"),
    "network": ("net", None, None),
    "network_precond": ("net", None, "diag"),
    # 30 inputs: the bits plus two diagonal-energy features (not the 28-input
    # configuration; shown beside it)
    "network_features": ("net_features", None, None),
    "network_features_precond": ("net_features", None, "diag"),
}
C5_UNGUARDED = {
    # for comparison only: starts without a Gaussian part
    "reference_min_diag_one_hot": ("min_diag", 0.0, None),
    "network_without_gaussian_part": ("net", 0.0, None),
}


GEMM_SHAPES = (
    # (name, M, N, K, trans_a, trans_b, ksplit): the start network's products at
    # a training batch of 4096, and one large square product
    ("hidden_forward", 4096, 256, 256, 0, 0, 1),
    ("hidden_input_grad", 4096, 256, 256, 0, 1, 1),
    ("hidden_weight_grad", 256, 256, 4096, 1, 0, 8),
    ("square_8192", 8192, 8192, 8192, 0, 0, 1),
)


def gemm_block(dev, reps=20):
    """BF16 GEMM (sqdk_gemm_bf16, FP32 accumulation) timed with HIP events on
    the stream it is launched on; TFLOP/s against the 2.5 PFLOP/s dense BF16
    peak."""
    import ctypes as C
    import torch
    from rqutils_amd import _lib
    L = _lib.lib()
    stream = torch.cuda.current_stream(dev)
    g = torch.Generator(device=dev).manual_seed(5)
    out = {}
    for name, M, N, K, ta, tb, ks in GEMM_SHAPES:
        a = torch.randn((K, M) if ta else (M, K), device=dev, generator=g).to(torch.bfloat16)
        b = torch.randn((N, K) if tb else (K, N), device=dev, generator=g).to(torch.bfloat16)
        c = torch.empty((M, N), dtype=torch.float32, device=dev)
        nb = C.c_size_t(0)
        _lib.check(L.sqdk_gemm_bf16_workspace(M, N, ks, C.byref(nb)), "gemm workspace")
        work = torch.empty(max(nb.value, 4), dtype=torch.uint8, device=dev)

        def run():
            _lib.check(L.sqdk_gemm_bf16(M, N, K, a.data_ptr(), a.shape[1], ta, b.data_ptr(),
                                        b.shape[1], tb, c.data_ptr(), N, None, ks,
                                        work.data_ptr() if ks > 1 else None, nb.value,
                                        stream.cuda_stream), "sqdk_gemm_bf16")
        run()
        ref = (a.float().t() if ta else a.float()) @ (b.float().t() if tb else b.float())
        err = ((c - ref).norm() / ref.norm()).item()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        n = reps if M * N * K < (1 << 34) else 3
        e0.record(stream)
        for _ in range(n):
            run()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / n
        tf = 2.0 * M * N * K / (ms * 1e-3) / 1e12
        out[name] = {"shape": [M, N, K], "trans": [ta, tb], "ksplit": ks, "ms": round(ms, 5),
                     "tflops": round(tf, 2), "frac_of_2500": round(tf / 2500.0, 4),
                     "rel_err_vs_fp32": float(f"{err:.2e}")}
        del a, b, c, ref, work
    return out


def c5_block(dev, n_train, n_test, steps, features=True):
    """Network-started LOBPCG on the 28-qubit chain family at 2^20 states: a
    28-input BF16 MLP (bits as +-1) trained with the library's MFMA kernels on
    the FP64 ground vectors of ``n_train`` Hamiltonians, then used as the start
    on ``n_test`` held-out ones; with ``features`` also a 30-input network
    (bits plus two diagonal-energy inputs) beside it. Every start is judged
    against the FP64 energy of plain LOBPCG from a Gaussian start (fused
    complex128, 1e-10)."""
    import multiprocessing as mp
    import torch
    from rqutils_amd.sqd import SQDProblem, start_vector
    from rqutils_amd.startnet import StartNet, energy_features, ground_targets
    seeds = list(range(1000, 1000 + n_train)) + list(range(2000, 2000 + n_test))
    t = time.time()
    with mp.get_context("spawn").Pool(min(CPU_THREADS, len(seeds))) as pool:
        ws = pool.map(_make_family_instance, seeds, chunksize=4)
    gen_s = time.time() - t
    log(f"[netstart] {len(ws)} instances generated ({gen_s:.1f} s)")
    keys, tgts, feats = [], [], []
    sync()
    t = time.time()
    for j, w in enumerate(ws[:n_train]):
        p = SQDProblem(w.hamiltonian, w.keys, 28, device=dev)
        tg, _, info = ground_targets(p, precond="diag")
        keys.append(p.space.keys)
        tgts.append(tg)
        if features:
            feats.append(energy_features(p))
        if (j + 1) % 50 == 0:
            log(f"[netstart] {j + 1}/{n_train} ground vectors")
    sync()
    ground_s = time.time() - t
    net = StartNet(28, device=dev, seed=0)
    sync()
    t = time.time()
    losses = net.train(keys, tgts, steps=steps, lr=3e-3)
    sync()
    train_s = time.time() - t
    net_f = None
    if features:
        net_f = StartNet(28, device=dev, seed=0, n_extra=2)
        losses_f = net_f.train(keys, tgts, feats, steps=steps, lr=3e-3)
        sync()
    flop_per_sample = 2 * 3 * (net.kpad * 256 + 256 * 256 + 256 * 2)
    log(f"[netstart] {n_train} ground vectors ({ground_s:.1f} s); {steps} Adam steps "
        f"({train_s:.1f} s), loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    del keys, tgts, feats
    starts = {k: v for k, v in {**C5_STARTS, **C5_UNGUARDED}.items()
              if features or v[0] != "net_features"}
    per = {k: {"iterations": [], "seconds": [], "energy_rel_err": [], "converged": 0,
               "status4": 0, "overlap": []} for k in starts}
    nets = {"net": net, "net_features": net_f}
    for w in ws[n_train:]:
        p = SQDProblem(w.hamiltonian, w.keys, 28, device=dev)
        _, psi, ref = ground_targets(p, seed=w.seed)
        e_ref = ref.eigval
        for name, (st, mix, precond) in starts.items():
            sync()
            t0 = time.perf_counter()
            x0 = start_vector(p.op, nets.get(st, st), mix=mix, seed=w.seed)
            _, inf = p.solve(x0, mode="fused", precision="c64", criterion="relative", tol=1e-6,
                             maxiter=3000, precond=precond)
            sync()
            r = per[name]
            r["iterations"].append(inf.niter)
            r["seconds"].append(time.perf_counter() - t0)
            r["energy_rel_err"].append(abs(inf.eigval - e_ref) / abs(e_ref))
            r["converged"] += int(inf.converged)
            r["status4"] += int(inf.status == 4)
            r["overlap"].append(abs(torch.vdot(psi, x0.to(psi.dtype)).item())
                                / (psi.norm().item() * x0.norm().item()))
        log(f"[netstart] seed {w.seed}: " + ", ".join(
            f"{k} {per[k]['iterations'][-1]} it/{per[k]['energy_rel_err'][-1]:.0e}"
            for k in starts))
    out = {"family": "28-qubit chain + 2,000 disorder strings, 2^20 domain-wall states",
           "inputs": net.n_in, "train": n_train, "held_out": n_test, "states": ws[-1].n,
           "generate_s": round(gen_s, 2), "ground_vectors_s": round(ground_s, 2),
           "train_s": round(train_s, 2), "adam_steps": steps,
           "train_tflops": round(flop_per_sample * 4096 * steps / train_s / 1e12, 3),
           "train_ms_per_step": round(1e3 * train_s / max(steps, 1), 3),
           "gemm": gemm_block(dev),
           "loss_first_last": [round(losses[0], 5), round(losses[-1], 5)],
           "features_network_loss_first_last": None if net_f is None else
           [round(losses_f[0], 5), round(losses_f[-1], 5)],
           "judge": "FP64 fused LOBPCG from a Gaussian start to 1e-10",
           "wrong_state_threshold": 1e-6, "guarded_starts": {}, "unguarded_for_comparison": {}}
    for k, r in per.items():
        errs = np.array(r["energy_rel_err"])
        rec = {"iterations_mean": round(float(np.mean(r["iterations"])), 2),
               "iterations_median": float(np.median(r["iterations"])),
               "seconds_mean": round(float(np.mean(r["seconds"])), 4),
               "energy_rel_err_max": float(errs.max()),
               "wrong_state": int(np.count_nonzero(errs > 1e-6)),
               "converged": r["converged"], "invariant_start_status": r["status4"],
               "overlap_mean": round(float(np.mean(r["overlap"])), 4)}
        (out["guarded_starts"] if k in C5_STARTS else out["unguarded_for_comparison"])[k] = rec
    gs = out["guarded_starts"]
    for k, rec in list(gs.items()) + list(out["unguarded_for_comparison"].items()):
        base = "gaussian_precond" if k.endswith("precond") else "gaussian"
        rec["speedup_vs_" + base] = round(gs[base]["seconds_mean"] / rec["seconds_mean"], 2)
    ref_md = out["unguarded_for_comparison"]["reference_min_diag_one_hot"]
    for k in ("network", "network_precond"):
        gs[k]["speedup_vs_reference_min_diag_one_hot"] = round(
            ref_md["seconds_mean"] / gs[k]["seconds_mean"], 2)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="random40",
                    choices=["random40", "heisenberg28", "small16", "batch64", "netstart",
                             "nextrows", "svsim"])
    ap.add_argument("--instances", type=int, default=64, help="instances of batch64")
    ap.add_argument("--no-solve", action="store_true", help="skip the time-to-solution block")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-groups", type=int, default=16,
                    help="X groups in the CPU baseline's sample")
    ap.add_argument("--no-extra", action="store_true",
                    help="headline workload only (no c1/c2/c4/c5/svsim blocks)")
    ap.add_argument("--no-net", action="store_true", help="skip the network-start block")
    ap.add_argument("--no-intake", action="store_true", help="skip the sample-intake block")
    ap.add_argument("--no-cache", action="store_true", help="skip the hit-table block")
    ap.add_argument("--no-svsim", action="store_true", help="skip the state-vector block")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-instance block")
    ap.add_argument("--net-train", type=int, default=200, help="training Hamiltonians")
    ap.add_argument("--net-test", type=int, default=50, help="held-out Hamiltonians")
    ap.add_argument("--net-steps", type=int, default=3000, help="Adam steps")
    ap.add_argument("--no-net-features", action="store_true",
                    help="skip the 30-input network (bits plus two diagonal-energy features)")
    ap.add_argument("--bucket-cap", type=int, default=None,
                    help="fix the launch: largest bucket (search lookup), no autotuning")
    ap.add_argument("--dense-bits", type=int, default=0,
                    help="fix the launch: dense lookup at this bucket width")
    ap.add_argument("--rows-per-thread", type=int, default=4, help="with a fixed launch")
    ap.add_argument("--group-split", type=int, default=1, help="with a fixed launch")
    ap.add_argument("--compact", type=int, default=0, help="with a fixed launch (0/1)")
    ap.add_argument("--fine-shift", type=int, default=-2,
                    help="with a fixed search launch: fine bits = log2(mean bucket) + this")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 1 GiB device copy first (scale for rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from rqutils_amd import workloads as W

    dev = torch.device("cuda", 0)
    if args.workload == "svsim":
        return bench_svsim(args, dev)
    if args.workload == "nextrows":
        return bench_next_rows(args, dev)
    if args.workload in ("batch64", "netstart", "small16"):
        if args.workload == "batch64":
            res, key = c4_block(args, dev), "c4_batch64"
        elif args.workload == "netstart":
            res, key = c5_block(dev, args.net_train, args.net_test, args.net_steps,
                                not args.no_net_features), "c5_network_start"
        else:
            res, key = c1_block(dev), "c1_small16"
        print(json.dumps({"metric": METRIC, "value": None, "unit": None, "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup,
                          "config": {"workload": args.workload, "parallelism": "single GPU"},
                          key: res}), flush=True)
        return
    if args.calibrate_copy:
        # A plain copy of known size: 1 GiB read + 1 GiB written per launch.
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        sync()
        del src, dst
        log("calibration copy: 3 x 1 GiB device copies done")
    t = time.time()
    w = W.make(args.workload)
    log(f"workload {w.name}: {w.n} states, {len(w.labels)} terms ({time.time() - t:.1f} s)")
    launch = None
    if args.bucket_cap or args.dense_bits:
        launch = {"rows_per_thread": args.rows_per_thread, "group_split": args.group_split,
                  "compact": args.compact}
        launch.update({"dense": True, "bucket_bits": args.dense_bits} if args.dense_bits
                      else {"dense": False, "bucket_cap": args.bucket_cap,
                            "fine_shift": args.fine_shift})
    prob, setup_s = build_problem(w, dev, launch=launch)
    mv, (v, out, o128, hits) = matvec_block(prob, w, args, dev)
    record = {
        "metric": METRIC,
        "value": mv["achieved_gbs"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mv["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 two-sum accumulate)",
        "data": "synthetic (seeded generator, rqutils_amd.workloads)",
        "config": {"workload": w.name, "num_qubits": w.num_qubits, "states": w.n,
                   "terms": len(w.labels), **mv["launch"], "groups_searched": mv["groups_searched"],
                   "groups_pruned_odd": mv["groups_pruned_odd"], "classes": mv["classes"],
                   "hits": hits, "autotuned": prob.tuned is not None,
                   "parallelism": "single GPU"},
        "matvec": mv,
        "roofline": {"bound": "hbm", "achieved": mv["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": mv["hbm_frac"],
                     "traffic": read_traffic(os.path.join(
                         ROOT, "profiles", f"bucket_join_kernel_{w.name}.md")),
                     "kernel": mv["kernel"], "bytes_per_launch": mv["bytes_per_launch"],
                     "measured_limiter": read_limiter(os.path.join(
                         ROOT, "profiles", f"bucket_join_kernel_{w.name}.md"))},
    }
    if not args.no_cache:
        record["hit_table"] = hit_table_block(prob, v, out, o128, hits, args.steps, dev)
        prob.op.drop_cache()
    del v, out, o128
    if not args.no_solve:
        record["time_to_solution"] = time_to_solution(prob, w, dev, setup_s)
        prob.op.drop_cache()
    if not args.no_intake:
        record["intake"] = intake_block(w, dev)
    del prob
    torch.cuda.empty_cache()
    if not args.no_cpu:
        record["cpu_baseline"] = cpu_baseline(w, mv["bytes_per_launch"], args.cpu_groups,
                                              CPU_THREADS)
    if w.name == "random40" and not args.no_extra:
        record["c1_small16"] = c1_block(dev)
        record["c2_heisenberg28"] = c2_block(args, dev)
        if not args.no_batch:
            record["c4_batch64"] = c4_block(args, dev)
        if not args.no_net:
            record["c5_network_start"] = c5_block(dev, args.net_train, args.net_test,
                                                  args.net_steps, not args.no_net_features)
        if not args.no_svsim:
            record["svsim"] = svsim_block(dev, cpu=not args.no_cpu)
    print(json.dumps(record), flush=True)


if __name__ == "__main__":
    main()
