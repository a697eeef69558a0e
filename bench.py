#!/usr/bin/env python3
"""Benchmark of the plain H·v (mltply) on one MI355X.

Workload: spin-1/2 J1-J2 Heisenberg model on the 8x4 periodic square lattice,
J1 = 1, J2 = 0.5, canonical Sz = 0 sector (dim C(32,16) = 601,080,390,
9.6 GB per complex128 vector), synthetic input vector (SplitMix64, seed 2).
One step = one H·v over the whole sector (v0 += H v1), tables and diagonal
already resident in HBM, vectors resident in HBM.

value = algorithmic bytes of one H·v / time per H·v, where
  B_alg = 16 NNZ_off + dim (16 [own v1] + 32 [v0 read+write] + 8 [list_Diagonal] + 8 [list_1])
and NNZ_off = 128 bonds x 2 C(30,15) = 39,710,085,120 off-diagonal elements.

cpu_baseline: the CPU oracle (a C + OpenMP restatement of the reference's
per-term scatter sweep) timed on a bounded, contiguous slice of source
states of the same workload (all tables and full-length vectors), its
algorithmic bytes counted the same way.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H·v GB/s (% HBM peak); solves/s + p50/p99 latency vs plain, |ΔE|≤1e-10"
HBM_PEAK_GBS = 8000.0
WORKLOAD = "J1-J2 Heisenberg 8x4 PBC, J1=1, J2=0.5, Sz=0 (dim 601,080,390), plain H·v"


def b_alg(dim: int, nnz: int, basis_bytes: int = 8, v0_bytes: int = 32) -> int:
    return 16 * nnz + dim * (16 + v0_bytes + 8 + basis_bytes)


def plain_nnz(spec) -> int:
    """Off-diagonal elements of one H·v of a canonical sector: an exchange bond acts
    where its two spins differ, C(N-2, Nup-1) states each way; a directed hop
    c+_{i s} c_{j s} where orbital (j, s) is filled and (i, s) empty,
    C(N-2, n_s - 1) C(N, n_other) states."""
    from hphi_amd import lattice as L

    n = spec.nsite
    if spec.model == L.SPIN:
        return len(spec.exchange) * 2 * math.comb(n - 2, spec.nup - 1)
    nnz = 0
    for (i, si, j, sj, _, _) in spec.trans:
        ns, no = (spec.nup, spec.ndown) if si == 0 else (spec.ndown, spec.nup)
        nnz += math.comb(n - 2, ns - 1) * math.comb(n, no)
    return nnz


# per workload of the HBM-traffic runs: profile name, title, seed of the probe vector
PMC_WORKLOADS = {3: ("hv_j1j2_8x4", "J1-J2 8x4 Sz=0, dim 601,080,390", 2),
                 2: ("hv_hubbard_4x4", "Hubbard 4x4 Nup=Ndown=8, dim 165,636,900", 1)}


def traffic_from_profiles(name: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    copy-kernel calibrated), recorded under profiles/ by the profiling runs."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def gpu_leg(args):
    from hphi_amd import capi, lattice as L

    spec = L.baseline_config(3)
    t0 = time.time()
    eng = capi.Engine.from_spec(spec)
    eng.sz(copy=False)
    eng.diagonalcalc(copy=False)
    t_setup = time.time() - t0
    dim = eng.dim
    nnz = plain_nnz(spec)
    balg = b_alg(dim, nnz)
    v1 = capi.DeviceArray(dim)
    v0 = capi.DeviceArray(dim)
    eng.fill_random(v1, 1, 2)
    v0.zero()
    # warmup, then K timed H·v bracketed by device synchronisation
    if args.warmup:
        eng.time_mltply(1, v0, v1, 0, args.warmup)
    capi.lib().hphi_synchronize()
    t = time.time()
    ms = eng.time_mltply(1, v0, v1, 0, args.steps)  # HIP events on the engine's stream
    capi.lib().hphi_synchronize()
    wall = time.time() - t
    # spread of single launches (each bracketed by its own pair of HIP events)
    import numpy as np
    single = np.array([eng.time_mltply(1, v0, v1, 0, 1) for _ in range(max(args.steps, 5))])
    spread = {q: round(float(np.percentile(single, p)), 3) for q, p in (("p10", 10), ("p50", 50), ("p90", 90))}
    # Lanczos step (H·v + update + beta + tridiagonal), device-resident driver
    lz = eng.lanczos(n_eigen=1, max_iter=args.lanczos_steps, eps=1e-300, target=2, seed=2)
    v0.free()
    v1.free()
    blk = None
    if args.block_k > 1:  # block H·v: the same sector for k columns in one launch
        k = args.block_k
        V1 = capi.DeviceArray(dim * k)
        V0 = capi.DeviceArray(dim * k)
        eng.fill_random(V1, k, 2)
        V0.zero()
        blk = eng.time_mltply(k, V0, V1, 1, 2) / k
        V0.free()
        V1.free()
    device_gb = (eng.device_bytes + 2 * 16 * dim) / 1e9  # the device-resident path, before any staging buffer
    host = None
    if args.host_reps > 0:  # compatibility mode: the reference's 1-indexed host arrays, staged per call
        import numpy as np

        h1 = np.full(dim + 1, 1.0 / math.sqrt(dim), np.complex128)
        h1[0] = 0.0
        h0 = np.zeros(dim + 1, np.complex128)
        eng.mltply(h0, h1)  # the first call allocates the staging buffers
        t = time.time()
        for _ in range(args.host_reps):
            eng.mltply(h0, h1)
        hs = (time.time() - t) / args.host_reps
        host = {"ms_per_call": round(1e3 * hs, 1), "gbs_algorithmic": round(balg / hs / 1e9, 1),
                "pcie_bytes_per_call": 2 * 16 * dim, "reps": args.host_reps,
                "note": "hphi_mltply on host arrays: v1 in and H v1 out through pinned bounce buffers on "
                        "worker threads, v0 added on the host, on every call; a labelled compatibility "
                        "rate, never the value"}
        del h0, h1
    res = {
        "dim": dim, "nnz": nnz, "balg": balg, "ms": ms, "wall_s": wall, "setup_s": t_setup, "host": host,
        "spread": spread,
        "lanczos_step_ms": 1e3 * lz.seconds / max(lz.iterations, 1),
        "device_gb": device_gb,
        "block_ms_per_column": blk,
    }
    eng.close()
    return res


def cfg4_leg(args):
    """Workload 4: triangular 13x2 XXZ in a transverse field (SpinGC, dim 2^26).
    mTPQ with k = 16 complex Gaussian samples: per-step time of the block
    layout (one block H·v for all samples, hipGraph-captured step) and of the
    per-sample layout (k single-vector H·v); canonical TPQ on the same 16
    samples (10 block H·v per step); then LOBPCG with a block of 8."""
    from hphi_amd import capi, lattice as L

    spec = L.baseline_config(4)
    eng = capi.Engine.from_spec(spec)
    eng.sz(copy=False)
    eng.diagonalcalc(copy=False)
    dim = eng.dim
    k = 16
    nnz_state = None
    out = {"dim": dim, "k": k}
    for layout, name in ((1, "block"), (0, "per_sample")):
        t = []
        for steps in (args.tpq_steps // 4, args.tpq_steps):
            _, r = eng.tpq(k, steps, seed=0, init_kind=-1, layout=layout, use_graph=True)
            t.append(r.seconds)
        ms = 1e3 * (t[1] - t[0]) / (args.tpq_steps - args.tpq_steps // 4)
        out[f"tpq_{name}_ms_per_step"] = round(ms, 3)
        out[f"tpq_{name}_ms_per_sample_step"] = round(ms / k, 3)
    # off-diagonal elements per H·v: each distinct exchange pair acts where the two
    # spins differ (half of the states; the torus lists 78 entries for 65 pairs, and
    # the engine merges identical terms), 52 transverse entries act where their
    # spin matches (half)
    pairs = len({(min(i, j), max(i, j)) for i, j, _ in spec.exchange})
    out["exchange_pairs"] = pairs
    nnz = pairs * dim // 2 + 52 * dim // 2
    nnz_state = nnz / dim
    balg = k * (16 * nnz + 16 * dim) + 8 * dim + 48 * dim * k  # overwrite H·v + diag + fused step pass
    out["tpq_block_gbs"] = round(balg / (out["tpq_block_ms_per_step"] / 1e3) / 1e9, 1)
    out["offdiag_per_state"] = nnz_state
    # canonical TPQ: ExpandCoef (10) block H·v per step, the same 16 samples
    t = []
    for steps in (2, 5):
        _, r = eng.ctpq(k, steps, seed=0, init_kind=-1)
        t.append(r.seconds)
    ms = 1e3 * (t[1] - t[0]) / 3
    out["ctpq_ms_per_step"] = round(ms, 2)
    out["ctpq_ms_per_block_hv"] = round(ms / 10, 2)
    r, hist = eng.lobpcg(8, max_iter=args.lobpcg_iters, eps=1e-7, seed=0)
    out["lobpcg_k"] = 8
    out["lobpcg_iterations"] = r.iterations
    out["lobpcg_converged"] = bool(r.converged)
    out["lobpcg_ms_per_iteration"] = round(1e3 * r.seconds / max(r.iterations, 1), 2)
    out["lobpcg_hv_fraction"] = round(r.hv_seconds / max(r.seconds, 1e-9), 3)
    out["lobpcg_eigenvalues"] = [round(x, 10) for x in r.eigenvalues[:8]]
    out["lobpcg_max_residual"] = r.max_residual
    eng.close()
    return out


def cfg2_leg(args):
    """Workload 2: Hubbard 4x4 PBC, t=1, U=4, Nup=Ndown=8 (dim 165,636,900):
    plain H·v (HIP events over K launches, seed 1) and the Lanczos step."""
    from hphi_amd import capi, lattice as L

    spec = L.baseline_config(2)
    eng = capi.Engine.from_spec(spec)
    eng.sz(copy=False)
    eng.diagonalcalc(copy=False)
    dim = eng.dim
    nnz = plain_nnz(spec)
    balg = b_alg(dim, nnz)
    v1 = capi.DeviceArray(dim)
    v0 = capi.DeviceArray(dim)
    eng.fill_random(v1, 1, 1)
    v0.zero()
    if args.warmup:
        eng.time_mltply(1, v0, v1, 0, args.warmup)
    ms = eng.time_mltply(1, v0, v1, 0, args.steps)
    v0.free()
    v1.free()
    lz = eng.lanczos(n_eigen=1, max_iter=args.lanczos_steps, eps=1e-300, target=2, seed=1)
    out = {"dim": dim, "nnz_offdiag": nnz, "bytes_alg_per_hv": balg, "hv_ms": round(ms, 3),
           "hv_gbs": round(balg / (ms / 1e3) / 1e9, 1), "hbm_frac": round(balg / (ms / 1e3) / 8e12, 4),
           "lanczos_step_ms": round(1e3 * lz.seconds / max(lz.iterations, 1), 3),
           "traffic_bytes_per_hv": traffic_from_profiles("hv_hubbard_4x4")}
    eng.close()
    return out


def pmc_probe(args):
    """A few plain H·v launches on the 8x4 J1-J2 sector, then the same number of
    launches of a streaming kernel of known traffic (y = a x + b y over two
    601,080,390-element complex vectors: 32 B read and 16 B written per
    element) that calibrates the FETCH_SIZE / WRITE_SIZE counters."""
    from hphi_amd import capi, lattice as L
    import numpy as np

    name, _, seed = PMC_WORKLOADS[args.pmc_cfg]
    eng = capi.Engine.from_spec(L.baseline_config(args.pmc_cfg))
    eng.sz(copy=False)
    eng.diagonalcalc(copy=False)
    dim = eng.dim
    v1 = capi.DeviceArray(dim)
    v0 = capi.DeviceArray(dim)
    eng.fill_random(v1, 1, seed)
    v0.zero()
    ms = eng.time_mltply(1, v0, v1, 0, args.steps)
    ab = np.array([0.5, 0.0, 0.25, 0.0])
    for _ in range(args.steps):
        capi.check(capi.lib().hphi_vec_axpby(eng.h, v1.ptr, v0.ptr, 1, ab.ctypes.data_as(C.POINTER(C.c_double))))
    print(json.dumps({"probe": name, "dim": dim, "hv_ms": ms, "steps": args.steps,
                      "calib_read_bytes": 32 * dim, "calib_write_bytes": 16 * dim}), flush=True)
    v0.free()
    v1.free()
    eng.close()


# kernels of one plain H·v: k_hv_canon, and for large canonical Spin sectors
# also the tile-major copy of v1 and the high-half pass (one dispatch each)
HV_KERNELS = ("k_hv_canon", "k_hi_copy", "k_hv_high")


def _pmc_means(directory: str, counter: str):
    """Mean counter value per dispatch of each H·v kernel and of the
    calibration kernel from rocprofv3 --output-format csv; "k_hv" sums the H·v
    kernels (the counter per H·v)."""
    import csv
    import glob

    sums = {}
    for f in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True):
        for row in csv.DictReader(open(f)):
            if row.get("Counter_Name") != counter:
                continue
            name = row["Kernel_Name"]
            key = next((k for k in HV_KERNELS + ("k_axpby",) if k in name), None)
            if key:
                a = sums.setdefault(key, [0.0, 0])
                a[0] += float(row["Counter_Value"])
                a[1] += 1
    means = {k: v[0] / v[1] for k, v in sums.items() if v[1]}
    means["k_hv"] = sum(means.get(k, 0.0) for k in HV_KERNELS)
    return means


def summarize_stats(csv_path: str, out: str, title: str, cmd: str, kernel: str, balg: float):
    """Short text summary of a rocprofv3 --stats kernel_stats.csv: per kernel
    calls and average / min / max duration.  `kernel` names the kernel (or a
    comma-separated list of the kernels launched once each per operation,
    whose averages add up) that gets the algorithmic rate balg / average."""
    import csv
    rows = list(csv.DictReader(open(csv_path)))
    lines = [f"# rocprofv3 kernel statistics: {title}", f"# command: {cmd}"]
    names = [k for k in kernel.split(",") if k] if kernel else []
    hit = [(k, r) for k in names for r in rows if k in r["Name"]]
    if hit:
        avg = sum(float(r["AverageNs"]) for _, r in hit) * 1e-9
        label = " + ".join(r["Name"] for _, r in hit)
        if balg > 0:
            lines.append(f"# {label}: average {avg * 1e3:.3f} ms -> {balg / avg / 1e12:.3f} TB/s algorithmic "
                         f"({balg:.6e} B per launch) = {100 * balg / avg / (HBM_PEAK_GBS * 1e9):.1f} % of 8 TB/s")
        else:  # launches of varying size (packed served slabs): the average only
            lines.append(f"# {label}: average {avg * 1e3:.3f} ms over launches of varying size")
    lines.append("")
    lines.append(f"{'kernel':<72} {'calls':>6} {'avg_ms':>10} {'min_ms':>10} {'max_ms':>10} {'pct':>6}")
    for r in rows:
        lines.append(f"{r['Name'][:72]:<72} {int(r['Calls']):>6} {float(r['AverageNs']) * 1e-6:>10.3f} "
                     f"{float(r['MinNs']) * 1e-6:>10.3f} {float(r['MaxNs']) * 1e-6:>10.3f} "
                     f"{float(r['Percentage']):>6.2f}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines[:4]))


def reduce_pmc(fetch_dir: str, write_dir: str, out: str, cmd: str, cfg: int = 3):
    """HBM bytes per H·v launch from a FETCH_SIZE and a WRITE_SIZE run of
    --pmc-probe, each scaled by what the counter reads for the calibration
    kernel of known traffic."""
    from hphi_amd import lattice as L

    spec = L.baseline_config(cfg)
    dim = spec.dim
    f, w = _pmc_means(fetch_dir, "FETCH_SIZE"), _pmc_means(write_dir, "WRITE_SIZE")
    ff = (32 * dim / 1024) / f["k_axpby"]
    fw = (16 * dim / 1024) / w["k_axpby"]
    fetch_b = f["k_hv"] * 1024 * ff
    write_b = w["k_hv"] * 1024 * fw
    nnz = plain_nnz(spec)
    used = [k for k in HV_KERNELS if k in f]
    lines = [
        f"# HBM traffic of one plain H·v ({' + '.join(used)}), {PMC_WORKLOADS[cfg][1]}",
        f"# command: {cmd}",
        *[f"FETCH_SIZE {k} (KiB/dispatch, raw): {f[k]:.1f}" for k in used],
        f"FETCH_SIZE k_axpby calibration (KiB/dispatch, raw): {f['k_axpby']:.1f}  known: {32 * dim / 1024:.1f}"
        f"  factor: {ff:.4f}",
        *[f"WRITE_SIZE {k} (KiB/dispatch, raw): {w[k]:.1f}" for k in used if k in w],
        f"WRITE_SIZE k_axpby calibration (KiB/dispatch, raw): {w['k_axpby']:.1f}  known: {16 * dim / 1024:.1f}"
        f"  factor: {fw:.4f}",
        f"fetch_bytes_per_launch: {fetch_b:.6e}",
        f"write_bytes_per_launch: {write_b:.6e}",
        f"algorithmic_bytes_per_launch: {b_alg(dim, nnz):.6e}",
        f"traffic_bytes_per_launch: {fetch_b + write_b:.6e}",
    ]
    open(out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


ISSUE_COUNTERS = ("SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_INSTS_SMEM",
                  "SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES")


def reduce_issue(directory: str, out: str, cmd: str, cfg: int = 2):
    """Issue-side counters of the plain H·v (one rocprofv3 --pmc pass of
    --pmc-probe with the SQ counters of ISSUE_COUNTERS): mean per dispatch, per
    64-state tile (one wave's worth of states), and the VALU issue time a tile
    needs (4 cycles per wave64 instruction on a 16-lane SIMD) against the
    cycles a tile may take at the measured launch time."""
    from hphi_amd import lattice as L

    spec = L.baseline_config(cfg)
    dim = spec.dim
    vals = {c: _pmc_means(directory, c).get("k_hv_canon") for c in ISSUE_COUNTERS}
    tiles = dim / 64.0
    lines = [f"# issue counters of one plain H·v launch (k_hv_canon), {PMC_WORKLOADS[cfg][1]}",
             f"# command: {cmd}"]
    for c in ISSUE_COUNTERS:
        v = vals[c]
        lines.append(f"{c}: {v:.6e} per dispatch" + (f", {v / tiles:.1f} per 64-state tile" if v else ""))
    hv_ms = _probe_hv_ms(directory)
    if vals["SQ_INSTS_VALU"] and hv_ms:
        simds = 1024
        cyc_tile = hv_ms * 1e-3 * 2.4e9 * simds / tiles
        lines.append(f"H·v {hv_ms:.3f} ms -> {cyc_tile:.0f} SIMD cycles per tile (1024 SIMDs at 2.4 GHz)")
        lines.append(f"VALU issue: {4 * vals['SQ_INSTS_VALU'] / tiles:.0f} cycles per tile "
                     f"({100 * 4 * vals['SQ_INSTS_VALU'] / tiles / cyc_tile:.1f} % of the tile's cycles)")
    open(out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def _probe_hv_ms(directory: str):
    """hv_ms the probe printed into the run directory's stdout copy, if kept."""
    p = os.path.join(directory, "probe.json")
    if os.path.exists(p):
        for ln in open(p):
            if ln.startswith("{"):
                return float(json.loads(ln)["hv_ms"])
    return None


# The port against the reference's own mltply (oracle/_ref, compiled from the
# reference's sources), both on the whole 8x4 J1-J2 sector with the same 8
# threads of the build container (Intel Xeon @ 2.70GHz), medians of 3 calls:
# reference 90.71 s (7.43 GB/s algorithmic), port 54.12 s (12.45 GB/s)
PORT_OVER_REFERENCE = {"speed_ratio": 1.676, "reference_gbs": 7.43, "port_gbs": 12.45, "threads": 8,
                       "cpu": "Intel(R) Xeon(R) Processor @ 2.70GHz (build container)",
                       "workload": "whole J1-J2 8x4 Sz=0 sector, one mltply, median of 3"}


def cpu_leg(args):
    import numpy as np

    from hphi_amd import lattice as L
    from oracle import oracle as O

    spec = L.baseline_config(3)
    orc = O.Oracle(spec, L.processed_terms(spec))
    orc.sz()
    dim = orc.dim
    S = min(int(args.cpu_sample), dim)
    jlo = 1 + dim // 3
    jhi = jlo + S - 1
    orc.diag = np.zeros(dim + 1)
    O.lib().or_diagonalcalc(C.byref(orc.m), orc.list_1[jlo - 1:].ctypes.data_as(O._u64p), S,
                            orc.diag[jlo - 1:].ctypes.data_as(O._f64p))
    v1 = np.empty(dim + 1, np.complex128)
    v0 = np.empty(dim + 1, np.complex128)
    O.lib().or_fill_vector(2, dim, v1.ctypes.data_as(O._f64p))
    O.lib().or_fill_vector(3, dim, v0.ctypes.data_as(O._f64p))  # touches every page of v0
    nnz = int(O.lib().or_count_nnz(C.byref(orc.m), orc.list_1.ctypes.data_as(O._u64p), jlo, jhi))
    times = []
    for r in range(args.cpu_reps):
        t = time.time()
        orc.mltply(v0, v1, jlo, jhi)
        times.append(time.time() - t)
    t = statistics.median(times)
    cpu = "unknown CPU"
    try:
        cpu = next(ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": b_alg(S, nnz) / t / 1e9, "unit": "GB/s", "cores": O.lib().or_num_threads(),
            "kind": "port",
            "sample": f"{S} contiguous source states (j={jlo}..{jhi}) of the same 8x4 J1-J2 sector, "
                      f"all 128 Exchange sweeps + diagonal, full-length tables and vectors; median of "
                      f"{args.cpu_reps} runs, {t:.2f} s each ({', '.join(f'{x:.2f}' for x in times)}); CPU "
                      f"oracle (C+OpenMP restatement of the reference's per-term scatter) on "
                      f"{O.lib().or_num_threads()} threads of {cpu}",
            "port_over_reference": PORT_OVER_REFERENCE,
            "seconds": t}


def stream_classes(reqs):
    """Indices of the stream's requests per structure class (the lattice
    name), classes in order of first appearance."""
    from collections import defaultdict
    by = defaultdict(list)
    for i, r in enumerate(reqs):
        by[r.name].append(i)
    return dict(by)


def stratified_seconds_per_request(classes, class_seconds):
    """Mean seconds per request of the whole stream from per-class samples:
    classes = {name: request indices}, class_seconds = {name: mean seconds of
    that class's sampled solves}; each class weighs by its request count."""
    n = sum(len(ids) for ids in classes.values())
    return sum(class_seconds[name] * len(ids) for name, ids in classes.items()) / n


def serve_leg(args):
    """Served stream (workload 5): Poisson arrivals of random small clusters,
    lowest 4 distinct eigenvalues each, through the continuous-batching
    service; the same requests solved one at a time by the plain path give
    the sequential baseline."""
    import numpy as np

    from hphi_amd import capi, lattice as L

    rng = np.random.default_rng(3)
    reqs = [L.stream_request(rng) for _ in range(args.requests)]
    arr_rng = np.random.default_rng(3)
    out = {"requests": args.requests, "dims": [int(min(r.dim for r in reqs)), int(max(r.dim for r in reqs))]}
    # sequential plain solves, stratified by structure class: the first
    # args.seq_per_class requests of every class of the stream (all of them
    # with 0), each class's mean weighted by its count in the stream
    by = stream_classes(reqs)
    warm = capi.Engine.from_spec(reqs[0])  # first-call set-up kept out of the timing
    warm.lanczos(n_eigen=4, max_iter=2000, eps=1e-12, check_every=5, mode=1, seed=0)
    warm.close()
    class_s, n_seq = {}, 0
    for name, ids in by.items():
        take = ids if args.seq_per_class <= 0 else ids[:args.seq_per_class]
        t = time.time()
        for i in take:
            eng = capi.Engine.from_spec(reqs[i])
            eng.lanczos(n_eigen=4, max_iter=2000, eps=1e-12, check_every=5, mode=1, seed=int(i))
            eng.close()
        class_s[name] = (time.time() - t) / len(take)
        n_seq += len(take)
        print(f"[serve] sequential {name}: {len(take)} solves, {1e3 * class_s[name]:.2f} ms each",
              file=sys.stderr, flush=True)
    seq_s = stratified_seconds_per_request(by, class_s)
    out["sequential_solves_per_s"] = 1.0 / seq_s
    print(f"[serve] sequential: {n_seq} solves over {len(by)} classes, {1.0 / seq_s:.1f}/s", file=sys.stderr,
          flush=True)
    out["sequential_sample"] = int(n_seq)
    out["sequential_method"] = (f"every request" if args.seq_per_class <= 0 else
                                f"first {args.seq_per_class} of each of the {len(by)} structure classes, "
                                f"weighted by the class counts")
    rates = [float(x) for x in args.rates.split(",")]
    points = []
    for rate in rates:
        # latency points (rate > 0) may run on the first args.latency_requests requests
        nr = len(reqs) if rate <= 0 or args.latency_requests <= 0 else min(len(reqs), args.latency_requests)
        gaps = arr_rng.exponential(1.0 / rate, size=len(reqs)) if rate > 0 else np.zeros(len(reqs))
        arrivals = np.cumsum(gaps) - gaps[0]
        srv = capi.Server(hbm_budget=int(args.hbm_gb * (1 << 30)), max_jobs=args.max_jobs, n_eigen=4,
                          eps=1e-12, check_every=5)
        start = srv.now() + 0.05
        for i, r in enumerate(reqs[:nr]):
            srv.submit(r, i, start + float(arrivals[i]))
        res = []
        ticks = 0
        while srv.now() < start:
            time.sleep(0.001)
        last = time.time()
        while len(res) < nr:
            srv.tick()
            ticks += 1
            res += srv.results()
            if time.time() - last > 10.0:
                last = time.time()
                print(f"[serve] rate {rate}: {len(res)}/{nr} done", file=sys.stderr, flush=True)
        st = srv.stats()
        srv.close()
        lat = np.array([r.t_done - r.t_arrival for r in res])
        makespan = max(r.t_done for r in res) - start
        conv = sum(r.converged for r in res)
        unconv = [{"id": int(r.id), "dim": int(r.dim), "iterations": int(r.iterations),
                   "model": reqs[r.id].name} for r in res if not r.converged]
        hv = {"hv_gbs": round(st["hv_bytes"] / st["hv_s"] / 1e9, 1) if st["hv_s"] > 0 else None,
              "hv_frac": round(st["hv_bytes"] / st["hv_s"] / 1e9 / HBM_PEAK_GBS, 4) if st["hv_s"] > 0 else None}
        points.append({"rate": rate, "requests": nr, "solves_per_s": len(res) / makespan,
                       "p50_s": float(np.percentile(lat, 50)),
                       **hv,
                       "unconverged": unconv[:8],
                       "p99_s": float(np.percentile(lat, 99)), "makespan_s": makespan,
                       "converged": int(conv), "mean_iterations": float(np.mean([r.iterations for r in res])),
                       **{k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}})
    out["points"] = points
    best = max(p["solves_per_s"] for p in points)
    out["speedup_vs_sequential"] = best / out["sequential_solves_per_s"]
    return out


PER_SIZE = (("chain16", lambda j2: ("heisenberg_chain", (16, 1.0, j2))),
            ("ladder20", lambda j2: ("j1j2_ladder", (10, 1.0, j2))),
            ("j1j2_6x4", lambda j2: ("j1j2_square", (6, 4, 1.0, j2))))


def serve_per_size(args, n_jobs: int = 32, n_plain: int = 32):
    """Served vs plain per job size: n_jobs requests of one lattice (random
    J2, seed 5) submitted at once to the service, and the first n_plain of them
    (by default the same requests) solved one at a time by the plain path (one
    engine each), both with the
    stream's settings (lowest 4 distinct eigenvalues, eps 1e-12, check every
    5 steps).  ms per solve each way and their ratio."""
    import numpy as np

    from hphi_amd import capi, lattice as L

    rng = np.random.default_rng(5)
    out = {}
    for name, make in PER_SIZE:
        specs = []
        for _ in range(n_jobs):
            fn, a = make(float(rng.uniform(0.0, 1.0)))
            specs.append(getattr(L, fn)(*a))
        t = time.time()
        for i in range(n_plain):
            eng = capi.Engine.from_spec(specs[i])
            eng.lanczos(n_eigen=4, max_iter=2000, eps=1e-12, check_every=5, mode=1, seed=i)
            eng.close()
        plain_ms = 1e3 * (time.time() - t) / n_plain
        srv = capi.Server(hbm_budget=int(args.hbm_gb * (1 << 30)), max_jobs=args.max_jobs, n_eigen=4, eps=1e-12,
                          check_every=5)
        t0 = time.time()
        for i, sp in enumerate(specs):
            srv.submit(sp, i, 0.0)
        done = 0
        while done < n_jobs:
            srv.tick()
            done += len(srv.results())
        served_ms = 1e3 * (time.time() - t0) / n_jobs
        st = srv.stats()
        srv.close()
        out[name] = {"dim": int(specs[0].dim), "plain_ms_per_solve": round(plain_ms, 2),
                     "served_ms_per_solve": round(served_ms, 2), "ratio": round(plain_ms / served_ms, 2),
                     "slab_hv_gbs": round(st["hv_bytes"] / st["hv_s"] / 1e9, 1) if st["hv_s"] > 0 else None}
        print(f"[serve] {name}: plain {plain_ms:.2f} ms, served {served_ms:.2f} ms per solve", file=sys.stderr,
              flush=True)
    return out


def serve_classes(args, n_jobs: int = 24, n_plain: int = 24):
    """Served vs plain per structure class of the stream: for each of the
    stream's structures (seed 3, args.requests requests) the first n_jobs
    requests of that structure go to the service at once and the first
    n_plain of them (by default the same requests) are solved one at a time
    by the plain path (one engine each), with the stream's settings.  Weighting each class's ms per solve by
    its share of the stream gives the whole-stream ratio and shows which
    classes cap it."""
    import numpy as np

    from hphi_amd import capi, lattice as L

    rng = np.random.default_rng(3)
    reqs = [L.stream_request(rng) for _ in range(args.requests)]
    by = stream_classes(reqs)
    warm = capi.Engine.from_spec(reqs[0])  # first-call set-up kept out of the first class's time
    warm.lanczos(n_eigen=4, max_iter=2000, eps=1e-12, check_every=5, mode=1, seed=0)
    warm.close()
    rows = {}
    for name, ids in sorted(by.items(), key=lambda kv: -reqs[kv[1][0]].dim):
        t = time.time()
        its = []
        for i in ids[:n_plain]:
            eng = capi.Engine.from_spec(reqs[i])
            its.append(eng.lanczos(n_eigen=4, max_iter=2000, eps=1e-12, check_every=5, mode=1, seed=int(i)).iterations)
            eng.close()
        plain_ms = 1e3 * (time.time() - t) / len(ids[:n_plain])
        take = ids[:n_jobs]
        srv = capi.Server(hbm_budget=int(args.hbm_gb * (1 << 30)), max_jobs=args.max_jobs, n_eigen=4, eps=1e-12,
                          check_every=5)
        t0 = time.time()
        for i in take:
            srv.submit(reqs[i], int(i), 0.0)
        done = []
        while len(done) < len(take):
            srv.tick()
            done += srv.results()
        served_ms = 1e3 * (time.time() - t0) / len(take)
        st = srv.stats()
        srv.close()
        rows[name] = {"dim": int(reqs[ids[0]].dim), "count": len(ids), "plain_ms": plain_ms, "served_ms": served_ms,
                      "plain_iters": float(np.mean(its)), "served_iters": float(np.mean([r.iterations for r in done])),
                      "cols_per_slab_step": st["job_steps"] / max(st["slab_steps"], 1)}
        print(f"[classes] {name}: plain {plain_ms:.2f} ms, served {served_ms:.2f} ms per solve", file=sys.stderr,
              flush=True)
    tp = sum(v["plain_ms"] * v["count"] for v in rows.values())
    ts = sum(v["served_ms"] * v["count"] for v in rows.values())
    for v in rows.values():
        v["plain_share"] = v["plain_ms"] * v["count"] / tp
        v["served_share"] = v["served_ms"] * v["count"] / ts
    return {"requests": args.requests, "plain_ms_per_request": tp / args.requests,
            "served_ms_per_request": ts / args.requests, "ratio": tp / ts, "classes": rows}


def write_classes(res: dict, out: str, cmd: str):
    """Text summary of serve_classes: one line per class, largest served share first."""
    lines = [f"# served vs plain per structure class of the {res['requests']:,}-request stream (seed 3)",
             f"# command: {cmd}",
             "# plain: the class's first requests solved one at a time (engine set-up included);",
             "# served: the same requests submitted at once to the service;",
             "# shares weight ms per solve by the class's count in the stream",
             f"# stream-weighted: plain {res['plain_ms_per_request']:.2f} ms, served "
             f"{res['served_ms_per_request']:.2f} ms per request -> ratio {res['ratio']:.2f}",
             f"{'class':18s} {'dim':>9s} {'count':>6s} {'plain ms':>9s} {'served ms':>9s} {'ratio':>6s} "
             f"{'served share':>12s} {'plain share':>11s} {'iters p/s':>11s} {'cols/step':>9s}"]
    cum = 0.0
    for n, v in sorted(res["classes"].items(), key=lambda kv: -kv[1]["served_share"]):
        cum += v["served_share"]
        lines.append(f"{n:18s} {v['dim']:9d} {v['count']:6d} {v['plain_ms']:9.2f} {v['served_ms']:9.2f} "
                     f"{v['plain_ms'] / v['served_ms']:6.2f} {v['served_share']:12.3f} {v['plain_share']:11.3f} "
                     f"{v['plain_iters']:5.0f}/{v['served_iters']:<5.0f} {v['cols_per_slab_step']:9.2f}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--lanczos-steps", type=int, default=5)
    ap.add_argument("--cpu-sample", type=float, default=4e8)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--block-k", type=int, default=4, help="also time the block H·v with this many columns")
    ap.add_argument("--pmc-probe", action="store_true", help="H·v + calibration launches for counter runs")
    ap.add_argument("--cfg2", action="store_true", help="plain H·v and Lanczos step on the Hubbard 4x4 workload")
    ap.add_argument("--cfg4", action="store_true", help="TPQ (k=16) and LOBPCG (block 8) on the 26-site SpinGC workload")
    ap.add_argument("--tpq-steps", type=int, default=40)
    ap.add_argument("--lobpcg-iters", type=int, default=60)
    ap.add_argument("--reduce-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/traffic_hv_j1j2_8x4.txt from two rocprofv3 --pmc runs of --pmc-probe")
    ap.add_argument("--pmc-cmd", type=str, default="", help="command line recorded in the traffic file")
    ap.add_argument("--reduce-issue", type=str, metavar="DIR",
                    help="write profiles/issue_<workload>.txt from a rocprofv3 --pmc run of --pmc-probe with the "
                         "SQ issue counters (DIR/probe.json: the probe's stdout)")
    ap.add_argument("--pmc-cfg", type=int, default=3, choices=(2, 3), help="workload of --pmc-probe / --reduce-pmc")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-sample", type=int, default=10000,
                    help="requests of the served stream timed with the default run (0: skip)")
    ap.add_argument("--latency-requests", type=int, default=0,
                    help="requests of the stream served at the arrival rates > 0 (0: all); the default run uses 2000")
    ap.add_argument("--host-reps", type=int, default=2, help="timed host-pointer (compatibility mode) calls; 0: skip")
    ap.add_argument("--summarize-stats", nargs=2, metavar=("KERNEL_STATS_CSV", "OUT"),
                    help="write a text summary of a rocprofv3 --stats run (with --stats-cmd, --stats-title)")
    ap.add_argument("--stats-cmd", type=str, default="")
    ap.add_argument("--stats-title", type=str, default="plain H·v on the J1-J2 8x4 Sz=0 sector (dim 601,080,390)")
    ap.add_argument("--stats-kernel", type=str, default=",".join(HV_KERNELS))
    ap.add_argument("--stats-balg", type=float, default=float(b_alg(601080390, 39710085120)),
                    help="algorithmic bytes per launch of --stats-kernel (0: report the average only)")
    ap.add_argument("--serve", action="store_true", help="served-stream benchmark instead of plain H·v")
    ap.add_argument("--serve-classes", type=str, default="",
                    help="served vs plain per structure class of the stream; writes the summary to this file")
    ap.add_argument("--requests", type=int, default=10000)
    ap.add_argument("--rates", type=str, default="0", help="arrival rates /s, comma separated; 0 = all at once")
    ap.add_argument("--seq-per-class", type=int, default=16,
                    help="sequential plain solves per structure class of the stream (0: every request)")
    ap.add_argument("--max-jobs", type=int, default=512)
    ap.add_argument("--hbm-gb", type=float, default=96.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(jobs, TPQ samples and basis shards would each map to one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.summarize_stats:
        summarize_stats(args.summarize_stats[0], args.summarize_stats[1], args.stats_title, args.stats_cmd,
                        args.stats_kernel, args.stats_balg)
        return
    if args.reduce_pmc:
        reduce_pmc(args.reduce_pmc[0], args.reduce_pmc[1],
                   os.path.join(ROOT, "profiles", f"traffic_{PMC_WORKLOADS[args.pmc_cfg][0]}.txt"), args.pmc_cmd,
                   args.pmc_cfg)
        return
    if args.reduce_issue:
        reduce_issue(args.reduce_issue, os.path.join(ROOT, "profiles", f"issue_{PMC_WORKLOADS[args.pmc_cfg][0]}.txt"),
                     args.pmc_cmd, args.pmc_cfg)
        return
    if args.pmc_probe:
        pmc_probe(args)
        return
    if args.cfg2:
        c2 = cfg2_leg(args)
        print(json.dumps({"metric": "H·v GB/s (% HBM peak)", "value": c2["hv_gbs"], "unit": "GB/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": c2["hv_ms"],
                          "higher_is_better": True, "dtype": "c128", "data": "synthetic",
                          "config": {"workload": "Hubbard 4x4 PBC t=1 U=4 Nup=Ndown=8, plain H·v", **c2}}),
              flush=True)
        return
    if args.cfg4:
        c4 = cfg4_leg(args)
        print(json.dumps({"metric": "mTPQ ms per step (k=16) and LOBPCG ms per iteration (block 8)",
                          "value": c4["tpq_block_ms_per_step"], "unit": "ms/step", "n_gpus": 1,
                          "higher_is_better": False, "dtype": "c128", "data": "synthetic",
                          "config": {"workload": "triangular 13x2 XXZ Jz=1 Jxy=0.8 hx=0.1, SpinGC dim 2^26", **c4}}),
              flush=True)
        return
    if args.serve_classes:
        res = serve_classes(args)
        write_classes(res, args.serve_classes,
                      f"python bench.py --serve-classes {args.serve_classes} --requests {args.requests}")
        print(json.dumps({"metric": "served vs plain per structure class", "value": round(res["ratio"], 3),
                          "unit": "x", "n_gpus": 1, "higher_is_better": True, "dtype": "c128",
                          "data": "synthetic", "config": {"workload": "served stream classes", **res}}), flush=True)
        return
    if args.serve:
        sv = serve_leg(args)
        best = max(sv["points"], key=lambda p: p["solves_per_s"])
        print(json.dumps({
            "metric": METRIC, "value": round(best["solves_per_s"], 2), "unit": "solves/s", "n_gpus": 1,
            "steps": args.requests, "warmup": 0, "ms_per_step": round(1e3 / best["solves_per_s"], 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
            "data": "synthetic",
            "config": {"workload": "served stream: random chains / 2-leg ladders / square J1-J2 (N 12..24, Sz=0 "
                                   "or 1/2) and half-filled Hubbard chains (N 8..12), lowest 4 eigenvalues each, "
                                   "Poisson arrivals (seed 3)", **sv},
        }), flush=True)
        return

    g = gpu_leg(args)
    s_per = g["ms"] / 1e3
    value = g["balg"] / s_per / 1e9
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(g["ms"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "dim": g["dim"], "nnz_offdiag": g["nnz"],
                   "bytes_alg_per_hv": g["balg"], "hbm_frac": round(value / HBM_PEAK_GBS, 4),
                   "hv_ms_single_launch": g["spread"],
                   "lanczos_step_ms": round(g["lanczos_step_ms"], 3),
                   "block_k": args.block_k if g["block_ms_per_column"] else None,
                   "block_ms_per_column": round(g["block_ms_per_column"], 3) if g["block_ms_per_column"] else None,
                   "block_hbm_frac_per_column": round(g["balg"] / (g["block_ms_per_column"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                   if g["block_ms_per_column"] else None,
                   "device_gb": round(g["device_gb"], 2), "setup_s": round(g["setup_s"], 2),
                   "host_pointer_mode": g["host"]},
        "roofline": {"bound": "hbm", "achieved": round(value, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(value / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles("hv_j1j2_8x4")},
    }
    if args.serve_sample > 0:  # the metric's served half: the stream back-to-back, its first 2,000 at 120 req/s
        sa = argparse.Namespace(**vars(args))
        sa.requests, sa.rates = args.serve_sample, "0,120"
        sa.latency_requests = args.latency_requests if args.latency_requests > 0 else 2000
        sv = serve_leg(sa)
        p0 = next(p for p in sv["points"] if p["rate"] == 0.0)
        p1 = next(p for p in sv["points"] if p["rate"] == 120.0)
        line["config"]["served_sample"] = {
            "requests": sa.requests, "solves_per_s": round(p0["solves_per_s"], 2),
            "sequential_solves_per_s": round(sv["sequential_solves_per_s"], 2),
            "sequential_method": sv["sequential_method"], "latency_requests": p1["requests"],
            "speedup_vs_sequential": round(p0["solves_per_s"] / sv["sequential_solves_per_s"], 3),
            "rate_per_s": 120.0, "p50_s": round(p1["p50_s"], 3), "p99_s": round(p1["p99_s"], 3),
            "converged": [p0["converged"], p1["converged"]],
            "slab_hv": {"kernel": "k_sl_hv", "achieved_gbs": p0["hv_gbs"], "frac": p0["hv_frac"],
                        "launches": p0["hv_launches"],
                        "note": "algorithmic bytes (per slab 8 dim, per running column 56 dim + 16 per "
                                "off-diagonal element) over HIP-event time of the slab H·v launches"},
            "per_size": serve_per_size(args),
            "note": "the seed-3 stream back-to-back; p50/p99 from its first latency_requests requests at 120 "
                    "req/s (bench.py --serve --rates 0,120,150 runs every rate on the whole stream)"}
    if not args.no_cpu:
        c = cpu_leg(args)
        line["cpu_baseline"] = {k: c[k] for k in ("value", "unit", "cores", "kind", "sample", "port_over_reference")}
        line["cpu_baseline"]["value"] = round(c["value"], 3)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
