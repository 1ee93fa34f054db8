#!/usr/bin/env python3
"""Benchmark of the drop-in block-sparse multiply and the purification loops on one MI355X.

Default run (no --workload): the headline line is the plain FP64 GPU multiply at
DZP scale — 4096 atoms at 0.05 atoms/A^3 (seeded, minimum separation 2 A), 13
support functions per atom (13x13 blocks), result range 16 A = (A range 8 A) .
(B range 16 A)^T in CONQUEST's reduction form (m_kern_min), entries
N(0,1)*exp(-r/2 A), FP64 — with, in `config`, the other workloads of the metric:
  * multiply_4x4: 512 atoms, 4x4 blocks, C(20 A) = A(8 A).B(12 A), full product
    (m_kern_max), FP64, against the HBM roofline (8 (lenA + lenB + lenC) bytes);
  * multiply_mixed_species: 8192 atoms, 50% 4 / 25% 9 / 25% 13 SF, C(16 A) =
    A(8 A).B(12 A), FP64 and FP32 accumulation with their relative errors;
  * purification: McWeeny (Hotelling S^-1 -> InitMcW -> McWMin) of 2048 atoms x
    13 SF at S/H/L 8/8/12 A, one all-FP64 solve and one mixed-precision solve
    (FP32 MFMA products with FP64 block sums, FP64 traces/mu/Ne and final steps):
    solves/s, band-energy error per atom and electron-count error against FP64.

A multiply "step" is one cqmm_mat_mult over device-resident operands (plan and
HBM copies built before the timed region). value = useful GFLOP/s = (sum over
products of 2*n_i*n_k*n_j) / wall time per step. Each roofline object uses the
mean kernel time from HIP events on the library stream; cpu_baseline times the
oracle's restatement of multiply_kernel_ompGemm on a bounded sample of halo
partitions on the host cores. --workload cfg5 runs the served request stream.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload all|cfg1|cfg2|cfg3|cfg4|cfg5]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "conquest-release_amd")]

import numpy as np  # noqa: E402

import cqmm  # noqa: E402

PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 matrix
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E

WORKLOADS = {
    # name: natoms, nsf, mult_type, (r_a, r_b, r_c) as mat_mult's a, b, c
    "cfg1": dict(natoms=512, nsf=4, mult_type=1, ra=8.0, rb=12.0, rc=20.0, bound="hbm",
                 desc="plain FP64, 512 atoms, 4x4 blocks, C(20 A) = A(8 A).B(12 A), full product (m_kern_max)"),
    "cfg2": dict(natoms=4096, nsf=13, mult_type=2, ra=16.0, rb=16.0, rc=8.0, bound="mfma",
                 desc="plain FP64 drop-in at DZP scale, 4096 atoms, 13x13 blocks, "
                      "X(16 A) = A(8 A).B(16 A)^T, C truncated to 16 A (m_kern_min)"),
    "cfg3": dict(natoms=8192, nsf=((4, .5), (9, .25), (13, .25)), mult_type=1, ra=8.0, rb=12.0, rc=16.0,
                 bound="mfma",
                 desc="batched uneven blocks, 8192 atoms, species 50% 4 / 25% 9 / 25% 13 SF (rectangular "
                      "blocks, one launch), C(16 A) = A(8 A).B(12 A), FP64 and FP32-accumulate"),
}


def species_nsf(natoms, mix, seed):
    """Support functions per atom: an int, or a seeded species draw ((nsf, fraction), ...)."""
    if isinstance(mix, int):
        return mix
    rng = np.random.default_rng(seed + 7919)
    vals = np.array([m[0] for m in mix], np.int32)
    probs = np.array([m[1] for m in mix], np.float64)
    return vals[rng.choice(len(vals), size=natoms, p=probs / probs.sum())]


TILE_KERNELS = {1: "k_tile_dense4_f64", 2: "k_tile_dense_mx_f64", 3: "k_tile_mixed_f64", 5: "k_tile_quad_f64"}


def kernel_name(info, nsf):
    """Name of the FP64 product kernel a plan runs (as rocprofv3 lists it, shortened)."""
    mixed = not isinstance(nsf, int)
    if info.kernel == cqmm.KERNEL_TILE:
        return TILE_KERNELS.get(info.tile_family, "k_tile_products_f64")
    return "k_block_products_f64" if mixed else "k_block_products_f64_u"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_for(workload, kernel):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(f"{workload}:{kernel}", {}).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(desc, a, b, c, target_s):
    """Oracle ompGemm restatement on a bounded sample of halo partitions."""
    import oracle
    provider = oracle.use_numpy_openblas()
    nparts = desc.ahalo.np_in_halo
    # Calibrate on one partition, then size the sample to ~target_s seconds.
    t = oracle.mat_mult_ompgemm(desc, a, b, c, 1, 2)
    f = oracle.flops(desc, 1, 2)
    rate = f / max(t, 1e-6)
    total = oracle.flops(desc)
    per_part = total / nparts
    n = int(max(1, min(nparts, target_s * rate / max(per_part, 1.0))))
    log(f"cpu baseline: {rate / 1e9:.2f} GFLOP/s on calibration, sampling {n}/{nparts} halo partitions")
    t = oracle.mat_mult_ompgemm(desc, a, b, c, 1, 1 + n)
    f = oracle.flops(desc, 1, 1 + n)
    return {
        "value": round(f / t / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": oracle.max_threads(),
        "kind": "port",
        "sample": f"multiply_kernel_ompGemm restatement (OpenMP schedule(runtime), {provider}); "
                  f"halo partitions 1..{n} of {nparts} ({f / 1e9:.1f} GFLOP, {t:.1f} s, partition loop only)",
    }


def multiply_workload(name, args, fp32_modes=True, cpu=False):
    """Build, tune and time one multiply workload; returns a result dict (and the
    host operands when the CPU baseline is wanted)."""
    w = WORKLOADS[name]
    t_setup = time.perf_counter()
    xyz, edge = cqmm.random_structure(w["natoms"], 0.05, 2.0, args.seed)
    ra, rb, rc = w["ra"], w["rb"], w["rc"]
    rcover = max(ra + rb, rc)
    sysm = cqmm.System(xyz, edge, species_nsf(w["natoms"], w["nsf"], args.seed), rcover,
                       cqmm.default_grid(edge, 4.0))
    ids = {}
    for r in sorted({ra, rb, rc}):
        ids[r] = sysm.add_range(r)
    log(f"{name}: {w['natoms']} atoms, box {edge:.2f} A, ranges {sorted(ids)} built in "
        f"{time.perf_counter() - t_setup:.1f} s")
    desc = sysm.mult_desc(w["mult_type"], ids[ra], ids[rb], ids[rc])
    # Operands as mat_mult receives them: type 1 reads a, b -> c; type 2 reads c, b -> a.
    b = sysm.fill(ids[rb], 0, args.seed, 2, (2.0,))
    if w["mult_type"] == 1:
        a = sysm.fill(ids[ra], 0, args.seed, 1, (2.0,))
        c = np.zeros(sysm.length(ids[rc]))
    else:
        c = sysm.fill(ids[rc], 0, args.seed, 3, (2.0,))
        a = np.zeros(sysm.length(ids[ra]))
    log(f"{name}: operands lena={a.size} lenb={b.size} lenc={c.size} doubles")
    cqmm.init(0)
    t0 = time.perf_counter()
    plan = cqmm.Plan(desc)
    t_plan = time.perf_counter() - t0
    plan.set_kernel({"auto": cqmm.KERNEL_AUTO, "triplet": cqmm.KERNEL_TRIPLET, "tile": cqmm.KERNEL_TILE}[args.kernel])
    info = plan.info()
    flops = info.flops
    kname = kernel_name(info, w["nsf"])
    log(f"{name}: {info.n_triplets} products, {flops / 1e12:.4f} TFLOP/step, plan built in {t_plan:.2f} s; "
        f"{info.n_tiles} tiles, {info.n_entries} entries, "
        f"{info.n_loads / max(info.n_triplets, 1):.3f} operand blocks per product")
    da, db, dc = (cqmm.DeviceArray.from_host(x) for x in (a, b, c))
    setup_s = time.perf_counter() - t_setup
    if args.calibrate_copy_mib > 0:
        n = args.calibrate_copy_mib * (1 << 20) // 8
        src, dst = cqmm.DeviceArray(n), cqmm.DeviceArray(n)
        cqmm.device_copy(dst, src)
        cqmm.synchronize()
        log(f"calibration copy: {n * 8} bytes read, {n * 8} bytes written")
        src.free()
        dst.free()
    tune = None
    if args.kernel == "auto":
        # Kernel autotuning on these operands: the faster of the tile and per-block kernels stays.
        ms_tile, ms_trip = plan.tune_dev(da.ptr, da.n, db.ptr, db.n, dc.ptr, dc.n, reps=2)
        info = plan.info()
        kname = kernel_name(info, w["nsf"])
        tune = {"tile_ms": None if ms_tile is None else round(ms_tile, 3), "per_block_ms": round(ms_trip, 3),
                "chosen": kname, "persistent_workgroups": int(info.tile_wgs) or None}
        log(f"{name}: autotune tile {ms_tile} ms, per-block {ms_trip:.3f} ms -> {kname}")
    flags = cqmm.CQMM_DEVICE_PTRS | cqmm.CQMM_NO_SYNC
    for _ in range(args.warmup):
        plan.mat_mult_dev(da.ptr, da.n, db.ptr, db.n, dc.ptr, dc.n, flags)
    cqmm.synchronize()
    t0 = time.perf_counter()
    cqmm.timer_start()
    for _ in range(args.steps):
        plan.mat_mult_dev(da.ptr, da.n, db.ptr, db.n, dc.ptr, dc.n, flags)
    kern_ms = cqmm.timer_stop()  # waits for the stream
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    kern_avg_ms = kern_ms / args.steps
    value = flops / (ms_step / 1e3) / 1e9
    out = (dc if w["mult_type"] == 1 else da).to_host()
    log(f"{name}: step {ms_step:.3f} ms (kernel {kern_avg_ms:.3f} ms), {value:.1f} GFLOP/s; "
        f"checksum {float(np.abs(out).sum()):.6e}")
    algo_bytes = 8.0 * (a.size + b.size + c.size)
    if w["bound"] == "hbm":
        achieved = algo_bytes / (kern_avg_ms / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4)}
    else:
        achieved = flops / (kern_avg_ms / 1e3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4)}
    roof.update(traffic=traffic_for(name, kname), kernel=kname, kernel_avg_ms=round(kern_avg_ms, 4),
                algorithmic_bytes=algo_bytes, tflops_fp64=round(flops / (kern_avg_ms / 1e3) / 1e12, 3))
    # The same products with FP32 MFMA on FP32 copies of the operands (per call:
    # FP64->FP32 operand copies, products, FP32->FP64 result): FP32 accumulation
    # (CQMM_FP32), each product's FP32 result summed in FP64 (CQMM_FP32_ACC64), and
    # a block's products summed in FP32 within a chunk, chunks in FP64
    # (CQMM_FP32_RUN64); errors against the FP64 result.
    fp32 = {}
    modes = (("fp32_accumulate", 0), ("fp32_acc64", cqmm.CQMM_FP32_ACC64), ("fp32_run64", cqmm.CQMM_FP32_RUN64))
    if info.kernel == cqmm.KERNEL_TILE and info.tile_family in (cqmm.TILES_CLASS, cqmm.TILES_UNIFORM + 9,
                                                                cqmm.TILES_UNIFORM + 13, cqmm.TILES_UNIFORM + 16):
        # split-BF16 x3 (hi.hi + hi.lo + lo.hi BF16 MFMA, FP64 sums): class tiles and 9-16 blocks
        modes += (("bf16x3", cqmm.CQMM_BF16X3),)
    for key, extra in modes if fp32_modes else ():
        f32 = cqmm.CQMM_DEVICE_PTRS | cqmm.CQMM_FP32 | cqmm.CQMM_NO_SYNC | extra
        plan.mat_mult_dev(da.ptr, da.n, db.ptr, db.n, dc.ptr, dc.n, f32)
        cqmm.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.mat_mult_dev(da.ptr, da.n, db.ptr, db.n, dc.ptr, dc.n, f32)
        cqmm.synchronize()
        ms32 = (time.perf_counter() - t0) / args.steps * 1e3
        out32 = (dc if w["mult_type"] == 1 else da).to_host()
        d = out32 - out
        fp32[key] = {"ms_per_step": round(ms32, 4), "gflops": round(flops / (ms32 / 1e3) / 1e9, 3),
                     "speedup_vs_fp64": round(ms_step / ms32, 3),
                     "max_rel_err": float(np.abs(d).max() / np.abs(out).max()),
                     "frobenius_rel_err": float(np.linalg.norm(d) / np.linalg.norm(out))}
        log(f"{name} {key}: {ms32:.3f} ms, {fp32[key]['gflops']:.1f} GFLOP/s, max rel err "
            f"{fp32[key]['max_rel_err']:.3e}, Frobenius rel err {fp32[key]['frobenius_rel_err']:.3e}")
    res = {
        "workload": w["desc"], "natoms": w["natoms"],
        "nsf": w["nsf"] if isinstance(w["nsf"], int) else "4/9/13 mix 0.5/0.25/0.25",
        "mult_type": w["mult_type"], "ranges_a_b_c": [ra, rb, rc], "products": int(info.n_triplets),
        "flop_per_step": flops, "output_blocks": int(info.n_out_total), "plan_build_s": round(t_plan, 3),
        "setup_s": round(setup_s, 2), "kernel": kname, "tiles": int(info.n_tiles), "autotune": tune,
        "ms_per_step": round(ms_step, 4), "gflops": round(value, 3), "roofline": roof,
        "operand_blocks_per_product": round(info.n_loads / max(info.n_triplets, 1), 4)
        if info.kernel == cqmm.KERNEL_TILE else 2.0,
    }
    if fp32:
        res["fp32"] = fp32
    for x in (da, db, dc):
        x.free()
    plan.close()
    # The descriptor's arrays live in the system, which it keeps alive.
    host = (desc, a, b, c) if cpu else None
    return res, host


def purification_pair(args, natoms=2048, nsf=13):
    """The mixed-precision McWeeny workload: one all-FP64 solve and one solve per
    reduced-precision mode of the same structure (precision 1: FP32 MFMA with
    FP64 block sums; 4: split-BF16 x3 MFMA with FP64 block sums), each with FP64
    traces, c, N_e and final steps; band-energy and electron-count errors
    against the FP64 solve."""
    from cqmm import serve as sv
    cqmm.init(0)
    out = {}
    precs = [int(p) for p in args.precisions.split(",")]
    if 0 not in precs:
        precs = [0] + precs
    # Symmetric-result plans for LSL and LSLSL (half the blocks computed, the
    # rest mirrored) in the listed precisions.
    sym = [int(p) for p in args.symmetric_precisions.split(",")] if args.symmetric_precisions else []
    runs = [(prec, 0) for prec in precs] + [(prec, 1) for prec in sym]
    for prec, symm in runs:
        r = sv.purify(natoms, nsf, args.seed, sv.default_params(precision=prec, symmetric=symm))
        if r["status"] != 0:
            raise SystemExit(f"purify failed: {r}")
        out[(prec, symm)] = r
        log(f"purification precision {prec}{' symmetric' if symm else ''}: solve {r['solve_s']:.2f} s "
            f"(McWeeny {r['mcw_s']:.2f} s), {r['iterations']} it ({r['fp32_iterations']} reduced precision), "
            f"E_band {r['band_energy']:.10f} Ha, Ne error {r['ne_error']:.2e}")
    e64 = out[(0, 0)]
    names = {0: "fp64", 1: "fp32_mfma_fp64_sums", 2: "fp32_accumulate", 3: "fp32_chunk_runs", 4: "split_bf16x3",
             5: "split_bf16x3_then_fp32"}
    modes = {}
    for prec, symm in runs:
        if (prec, symm) == (0, 0):
            continue
        e = out[(prec, symm)]
        modes[("symmetric_" if symm else "") + names[prec]] = {
            "precision": prec, "symmetric": symm, "solves_per_s": round(1.0 / e["solve_s"], 5),
            "speedup_vs_fp64": round(e64["solve_s"] / e["solve_s"], 3), "solve_s": round(e["solve_s"], 3),
            "mcw_s": round(e["mcw_s"], 3), "iterations": e["iterations"],
            "reduced_precision_iterations": e["fp32_iterations"], "band_energy": e["band_energy"],
            "band_energy_err_per_atom": abs(e["band_energy"] - e64["band_energy"]) / natoms,
            "n_electrons_err_vs_fp64": abs(e["n_electrons"] - e64["n_electrons"]), "ne_error": e["ne_error"],
        }
    best = max(modes.values(), key=lambda m: m["solves_per_s"] if m["band_energy_err_per_atom"] <= 1e-6 else 0.0)
    res = {
        "workload": f"mixed-precision McWeeny, {natoms} atoms x {nsf} SF, S/H/L ranges 8/8/12 A (Hotelling -> "
                    "InitMcW -> McWMin; S = I + 0.1 N(0,1) exp(-r/1.5), H onsite U(-0.5,0.5) split -/+1.5 Ha, "
                    "off-diagonal -0.2 N(0,1) exp(-r/2), + 3 S)",
        "fp64_solves_per_s": round(1.0 / e64["solve_s"], 5), "solve_s_fp64": round(e64["solve_s"], 3),
        "mcw_s_fp64": round(e64["mcw_s"], 3), "iterations_fp64": e64["iterations"],
        "band_energy_fp64": e64["band_energy"], "ne_error_fp64": e64["ne_error"],
        "setup_s": round(e64["setup_s"], 2), "modes": modes,
        "mixed_solves_per_s": best["solves_per_s"], "speedup_vs_fp64": best["speedup_vs_fp64"],
        "roofline": {
            "bound": "mfma", "unit": "TFLOP/s",
            "fp64": {"achieved": round(e64["mcw_flops"] / e64["mcw_s"] / 1e12, 3), "peak": PEAK_FP64_TFLOPS,
                     "frac": round(e64["mcw_flops"] / e64["mcw_s"] / 1e12 / PEAK_FP64_TFLOPS, 4)},
            "note": "useful FLOP of McWMin's products / McWMin wall time",
        },
    }
    return res


def served_stream(args):
    """The served request stream: plain FP64 (requests one at a time, no
    batching) against batched FP64 and batched mixed precision."""
    from cqmm import serve as sv
    cqmm.init(0)
    reqs = sv.poisson_stream(args.requests, args.rate, args.seed)
    if args.request_range:
        # A contiguous part of the stream (requests a..b-1, arrivals shifted to
        # start at 0): the full stream can be served in parts and the parts'
        # per-request results combined.
        a, b = (int(x) for x in args.request_range.split(":"))
        t0 = reqs[a]["arrival_s"]
        reqs = [dict(q, arrival_s=q["arrival_s"] - t0, index=a + k) for k, q in enumerate(reqs[a:b])]
    n = len(reqs)
    runs = {
        "plain_fp64": dict(precision=0, batch=0, slots=1),
        "batched_fp64": dict(precision=0, batch=1, slots=args.slots),
        "batched_mixed": dict(precision=args.mixed_precision, batch=1, slots=args.slots),
    }
    only = args.serve_modes.split(",") if args.serve_modes else list(runs)
    lines = {}
    energies = {}
    for key in only:
        m = runs[key]
        params = sv.default_params(precision=m["precision"], symmetric=args.serve_symmetric if m["batch"] else 0)
        log(f"serving {n} requests at {args.rate}/s: {key} ({m['slots']} slots, batch {m['batch']})")
        res, wall = sv.serve(reqs, slots=m["slots"], hbm_budget=int(args.hbm_gb * (1 << 30)), params=params,
                             log_progress=True, batch=m["batch"], max_batch=args.max_batch)
        ok = [r for r in res if r["status"] == 0]
        lat = np.array([r["latency_s"] for r in ok])
        svc = np.array([r["done_s"] - r["admit_s"] for r in ok])
        lines[key] = dict(solves_per_s=len(ok) / wall, wall_s=wall, failed=n - len(ok),
                          p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                          p50_service_s=float(np.percentile(svc, 50)), p99_service_s=float(np.percentile(svc, 99)),
                          max_ne_error=max(r["ne_error"] for r in ok))
        energies[key] = np.array([r["band_energy"] for r in res])
        if args.serve_out:
            rows = [dict(index=q.get("index", k), natoms=q["natoms"], nsf=q["nsf"], seed=q["seed"],
                         arrival_s=q["arrival_s"], **{f: r[f] for f in ("status", "band_energy", "n_electrons",
                                                                        "ne_error", "iterations", "latency_s",
                                                                        "admit_s", "done_s")})
                    for k, (q, r) in enumerate(zip(reqs, res))]
            with open(f"{args.serve_out}.{key}.json", "w") as f:
                json.dump({"mode": key, "wall_s": wall, "precision": m["precision"], "requests": rows}, f)
        log(f"{key}: {len(ok)}/{n} solved in {wall:.1f} s ({lines[key]['solves_per_s']:.4f}/s), "
            f"latency p50 {lines[key]['p50_latency_s']:.2f} s p99 {lines[key]['p99_latency_s']:.2f} s")
    na = np.array([q["natoms"] for q in reqs])
    part = f" (requests {args.request_range} of {args.requests})" if args.request_range else ""
    cfg = {"workload": f"served stream of {n} purification requests{part} (atoms log-uniform [64, 2048], 4 or 13 SF, "
                       f"S/H/L 8/8/12 A), Poisson {args.rate}/s, HBM budget {args.hbm_gb} GiB, "
                       f"max batch {args.max_batch}",
           "symmetric_plans_in_batched_modes": bool(args.serve_symmetric),
           **lines}
    if "batched_mixed" in energies and "plain_fp64" in energies:
        cfg["max_band_energy_err_per_atom"] = float(np.max(np.abs(energies["batched_mixed"] -
                                                                   energies["plain_fp64"]) / na))
        cfg["speedup_vs_plain_fp64"] = round(lines["batched_mixed"]["solves_per_s"] /
                                             lines["plain_fp64"]["solves_per_s"], 3)
    if "batched_fp64" in energies and "plain_fp64" in energies:
        cfg["batched_fp64_equals_plain"] = bool(np.array_equal(energies["batched_fp64"], energies["plain_fp64"]))
    head = lines.get("batched_mixed", next(iter(lines.values())))
    line = {
        "metric": "purification solves/s", "value": round(head["solves_per_s"], 5), "unit": "solves/s",
        "n_gpus": 1, "steps": n, "warmup": 0, "ms_per_step": round(1e3 / max(head["solves_per_s"], 1e-9), 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded Poisson stream; seeded random periodic structures; S/H recipe of the workload)",
        "config": cfg, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all", choices=["all"] + sorted(WORKLOADS) + ["cfg4", "cfg5"])
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--request-range", default="", help="serve only requests a:b of the stream (arrivals shifted)")
    ap.add_argument("--serve-out", default="", help="write per-request results to <prefix>.<mode>.json")
    ap.add_argument("--rate", type=float, default=0.05)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--max-batch", type=int, default=16)
    ap.add_argument("--serve-modes", default="", help="comma list of plain_fp64,batched_fp64,batched_mixed")
    ap.add_argument("--hbm-gb", type=float, default=240.0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-purification", action="store_true", help="default run: skip the purification pair")
    ap.add_argument("--mixed-precision", type=int, default=1,
                    help="precision of the served stream's mixed solves (1 FP32 MFMA + FP64 sums, 4 split-BF16 x3, "
                         "5 split-BF16 x3 then FP32 MFMA + FP64 sums)")
    ap.add_argument("--precisions", default="0,1,4",
                    help="purification solves: 0 all FP64, 1 FP32 MFMA + FP64 sums, 4 split-BF16 x3, 5 split-BF16 x3 "
                         "then FP32 MFMA + FP64 sums (comma list)")
    ap.add_argument("--symmetric-precisions", default="0,4,5",
                    help="purification: also solve with symmetric-result plans for LSL and LSLSL in these precisions")
    ap.add_argument("--serve-symmetric", type=int, default=0,
                    help="served stream: batched modes use symmetric-result plans for LSL and LSLSL")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run the streaming copy kernel once over this many MiB before the timed region "
                         "(reference kernel for scaling HBM traffic counters)")
    ap.add_argument("--kernel", default="auto", choices=["auto", "triplet", "tile"],
                    help="FP64 kernel: auto (tile schedule when present), triplet, tile")
    ap.add_argument("--tiles", default="default", choices=["default", "class"],
                    help="tile schedule: default, or class tiles for every block size above 4 (needed by the "
                         "split-BF16 x3 mode)")
    args = ap.parse_args()
    if args.tiles == "class":
        os.environ["CQMM_UNIFORM_TILES"] = "class"
        os.environ["CQMM_MIXED_TILES"] = "class"
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU design is described in DESIGN.md and is not built", file=sys.stderr)
        sys.exit(2)
    if args.workload == "cfg5":
        served_stream(args)
        return
    if args.workload == "cfg4":
        res = purification_pair(args)
        line = {"metric": "purification solves/s", "value": res["mixed_solves_per_s"], "unit": "solves/s",
                "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(1e3 / res["mixed_solves_per_s"], 2),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
                "data": "synthetic (seeded random periodic structure; S/H recipe of the workload)",
                "config": res, "roofline": None, "cpu_baseline": None}
        print(json.dumps(line), flush=True)
        return
    head = "cfg2" if args.workload == "all" else args.workload
    res, host = multiply_workload(head, args, fp32_modes=True, cpu=not args.no_cpu)
    cpu = None
    if host is not None:
        desc, a, b, c = host
        cpu = cpu_baseline(desc, a.copy() if WORKLOADS[head]["mult_type"] == 2 else a, b,
                           c if WORKLOADS[head]["mult_type"] == 2 else c.copy(), args.cpu_seconds)
        del desc, a, b, c, host
    cfg = dict(res)
    roof = cfg.pop("roofline")
    cfg.pop("gflops")
    cfg.pop("ms_per_step")
    if args.workload == "all":
        sub1, _ = multiply_workload("cfg1", args, fp32_modes=False)
        sub3, _ = multiply_workload("cfg3", args, fp32_modes=True)
        cfg["multiply_4x4"] = sub1
        cfg["multiply_mixed_species"] = sub3
        if not args.no_purification:
            cfg["purification"] = purification_pair(args)
    line = {
        "metric": "block-sparse mult GFLOP/s",
        "value": res["gflops"],
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random periodic structures; N(0,1)*exp(-r/2) block entries; purification S/H "
                "recipe)",
        "config": cfg,
        "roofline": {k: v for k, v in roof.items() if k not in ("tflops_fp64",)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
