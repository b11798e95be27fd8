#!/usr/bin/env python3
"""Benchmark of the H2 J-build matvec on one MI355X.

Workload (default, --config 2): linear alkane C100H202, cc-pVDZ (2510
Cartesian basis functions), H2 relative tolerance 1e-8, the largest
representation of the benchmark configurations on one GPU (6.7 GB of FP64
matrix data).  The representation is constructed on the host and uploaded
once.  A step is one H2 matvec y = A_H2 x on the uncontracted density (seeded
damped density, 401 occupied columns), replayed from a captured HIP graph,
with x and y resident in HBM.  Timed in the same run:

* one density per pass: FP64 storage (the drop-in) and FP32-stored U/B (FP64
  near field and accumulation); the executor, tile size and near-field
  schedule are the project's autotuner's choice for the FP32-stored form;
* the batching workload on the same representation: 16 seeded densities in
  one graph replay (one pass over the matrix data, the tiles as FP64 MFMA
  products), FP64 and FP32-stored, with the autotuner's choice for 16
  densities per pass.

`value` is the FP32-stored single-density form's algorithmic HBM throughput
(bytes of the representation and vectors one matvec must move, divided by
the graph replay's time).  Beside it: the batched forms, the precision error
(FP32-stored J against FP64 J), the FP64 J against the oracle, the
compression error of both against directly evaluated ERI rows (256 sampled
bra shell pairs), the roofline of the dominant kernel from the union of its
launch intervals, per-phase times, and the oracle's matvec on the host cores
(median of 10 after a warm-up, all cores and one core).

--config 1 and --config 4 run the same measurement on C30H62 (reltol 1e-6)
and (H2O)64 (reltol 1e-6, with the exchange build K beside J); --config 3
runs the molecule group of the batching workload (C20..C160 in one
persistent launch).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "h2p-eri_amd"))

METRIC = "J-build H2 matvec ms & HBM GB/s, FP32-stored vs FP64 on 1 GPU, rel.err ≤1e-7"
HBM_PEAK_GBS = 8000.0
MFMA_FP64_PEAK_TFLOPS = 78.6
MALL_BYTES = 256e6
EPS_PREC_GATE = 1e-7


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# timing helpers
# ---------------------------------------------------------------------------
def pass_widths(n):
    """Vectors per pass of a batched call of n vectors (16, 8, 4, 2, 1 as the
    library splits it)."""
    out = []
    while n > 0:
        w = 16 if n >= 16 else 8 if n >= 8 else 4 if n >= 4 else 2 if n >= 2 else 1
        out.append(w)
        n -= w
    return out


def time_graph(h, torch, stream, steps, warmup):
    """ms per replay of the captured graph, HIP events on the replay stream."""
    for _ in range(warmup):
        h.graph_replay(stream.cuda_stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        h.graph_replay(stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def interval_union(recs):
    """Length of the union of the launch intervals [start, start + ms]."""
    iv = sorted((r["start"], r["start"] + r["ms"]) for r in recs)
    tot, cur0, cur1 = 0.0, None, None
    for a, b in iv:
        if cur1 is None or a > cur1:
            if cur1 is not None:
                tot += cur1 - cur0
            cur0, cur1 = a, b
        else:
            cur1 = max(cur1, b)
    if cur1 is not None:
        tot += cur1 - cur0
    return tot


def launch_profile(h, torch, x_ptr, y_ptr, stream, reps):
    """Per-launch durations of one eager single-density matvec (HIP events
    around each launch, on the stream the launch runs on, plus a base event
    before the pass), averaged over `reps` matvecs: the records, the mean
    union of the launch intervals and the mean sum of the launch durations."""
    h.set_phase_timing(True)
    acc, unions, sums = None, [], []
    for _ in range(reps):
        h.matvec_dev(x_ptr, y_ptr, stream.cuda_stream)
        torch.cuda.synchronize()
        st = h.launch_stats()
        unions.append(interval_union(st))
        sums.append(sum(r["ms"] for r in st))
        if acc is None:
            acc = [dict(r) for r in st]
        else:
            for a, r in zip(acc, st):
                a["ms"] += r["ms"]
    h.set_phase_timing(False)
    for a in acc:
        a["ms"] /= reps
    return acc, sum(unions) / reps, sum(sums) / reps


def uncontract(tr, D, n):
    """x of src/H2ERI_build_Coulomb.c:26-62 (host copy for the timed input)."""
    import numpy as np
    x = np.empty(n)
    sp, sb, bf = tr["sp_shell_idx"], tr["sp_bfp_sidx"], tr["shell_bf_sidx"]
    npair = tr["num_sp"]
    for i in range(npair):
        s0, s1 = sp[i], sp[i + npair]
        r0, nr = bf[s0], bf[s0 + 1] - bf[s0]
        c0, nc = bf[s1], bf[s1 + 1] - bf[s1]
        blk = D[c0:c0 + nc, r0:r0 + nr] * (1.0 if s0 == s1 else 2.0)
        x[sb[i]:sb[i + 1]] = blk.reshape(-1)
    return x


def build_molecule(h2eri, M, atoms, reltol, label):
    """Host construction with progress lines from the construction itself."""
    t0 = time.perf_counter()
    h = h2eri.H2ERI(1e-10, 1e-10, reltol)
    h.read_mol_text(M.mol_text(atoms))
    h.process_shells()
    h.partition()
    log(f"[bench] {label}: {h.nbf} basis functions, {h.s.num_sp} shell pairs, N={h.n}; building H2 ...")
    h.s.print_dbginfo = 1
    h.build_h2()
    h.s.print_dbginfo = 0
    log(f"[bench] {label}: built in {time.perf_counter() - t0:.1f} s")
    return h


# ---------------------------------------------------------------------------
# CPU baseline: the oracle's OpenMP matvec in a child process, so that the
# OpenMP binding variables take effect before its runtime starts
# ---------------------------------------------------------------------------
def host_info():
    info = {}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)", "NUMA node(s)") \
                    or (k.startswith("NUMA node") and "CPU(s)" in k):
                info[k] = v
    except OSError:
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return info


def cpu_threads():
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


def cpu_child(rep_path, x_path, threads, reps):
    """Runs in the child: one warm-up, then `reps` timed oracle matvecs."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind
    lib = oracle_bind.load()
    lib.orc_set_threads(threads)
    o = oracle_bind.Oracle(rep_path)
    x = np.load(x_path)
    o.matvec(x)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        o.matvec(x)
        ts.append((time.perf_counter() - t0) * 1e3)
    o.close()
    ts.sort()
    print(json.dumps({"median_ms": ts[len(ts) // 2] if len(ts) % 2 else 0.5 * (ts[len(ts) // 2 - 1] + ts[len(ts) // 2]),
                      "ms": ts, "threads": threads}), flush=True)


def cpu_baseline(rep_path, x, reps=10, one_core=True):
    """Median ms of the oracle matvec on all host cores given to this process
    and on one core, each in a child process with OMP_PROC_BIND=close and
    OMP_PLACES=cores."""
    import numpy as np
    res = {}
    with tempfile.TemporaryDirectory() as d:
        xp = os.path.join(d, "x.npy")
        np.save(xp, x)
        for label, nthr in (("all", cpu_threads()), ("one", 1)):
            if label == "one" and not one_core:
                continue
            env = dict(os.environ, OMP_NUM_THREADS=str(nthr), OMP_PROC_BIND="close", OMP_PLACES="cores")
            t0 = time.perf_counter()
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", rep_path, xp, str(nthr), str(reps)],
                               stdout=subprocess.PIPE, env=env, text=True, check=True)
            res[label] = json.loads(r.stdout.strip().splitlines()[-1])
            log(f"[bench] CPU oracle matvec, {nthr} thread(s): median {res[label]['median_ms']:.2f} ms "
                f"({time.perf_counter() - t0:.1f} s)")
    return res


def direct_rows_error(h, x, ys, npair=256, seed=7):
    """Relative error of each y in `ys` against directly evaluated rows of the
    ERI matrix times x, over the rows of `npair` seeded random shell pairs."""
    import numpy as np
    pairs = np.sort(np.random.default_rng(seed).choice(h.s.num_sp, min(npair, h.s.num_sp), replace=False))
    t0 = time.perf_counter()
    rows, yd = h.direct_matvec_rows(pairs, x)
    log(f"[bench] direct ERI rows of {len(pairs)} shell pairs ({len(rows)} rows) in {time.perf_counter() - t0:.1f} s")
    nrm = float(np.linalg.norm(yd))
    return {k: float(np.linalg.norm(np.asarray(y)[rows] - yd) / nrm) for k, y in ys.items()}, len(pairs), len(rows)


# ---------------------------------------------------------------------------
# one representation: single density and the 16-density batch
# ---------------------------------------------------------------------------
def single_workload(args):
    import numpy as np
    import torch

    import h2eri
    from h2eri import molecules as M

    cfg = args.config
    atoms, _, D, reltol = M.workload(cfg)
    wname = M.workload_name(cfg)
    nocc = M.CONFIGS[cfg]["nocc"]
    t0 = time.perf_counter()
    h = build_molecule(h2eri, M, atoms, reltol, f"{wname} reltol {reltol:g}")
    t_build = time.perf_counter() - t0
    sizes = {k: round(h.s.mat_size[i] * 8 / 1e6, 2) for i, k in enumerate(["U_MB", "B_MB", "D_MB"])}

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    tr = h.tree()
    x = uncontract(tr, D, h.n)
    Ds = M.batch_densities(atoms, nocc, args.batch)
    X = np.stack([uncontract(tr, Dk, h.n) for Dk in Ds])
    x_dev = torch.from_numpy(x).to(dev)
    y_dev = torch.empty_like(x_dev)
    X_dev = torch.from_numpy(X).to(dev)
    Y_dev = torch.empty_like(X_dev)
    D_dev = torch.from_numpy(D).to(dev)
    J_dev = torch.empty_like(D_dev)

    if args.calibrate_copy_mb > 0:
        n = args.calibrate_copy_mb * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {n * 8} bytes read and written")
        del a, b

    # executor, tile size and near-field schedule: the project's autotuner on
    # the headline form (FP32-stored U/B, one density per pass)
    h.set_precision(32)
    h.upload()
    if args.tuned:
        vals = [int(v) for v in args.tuned.split(",")]
        ex, tb, sc = vals[:3]
        wb = vals[3] if len(vals) > 3 else 32768
        ng = vals[4] if len(vals) > 4 else 0
        h.set_executor(ex)
        h.set_tile_bytes(tb)
        h.set_schedule(sc)
        h.set_wave_tile_bytes(wb)
        h.set_near_field_grid(ng)
        tune = dict(ms=0.0, tile_bytes=tb, schedule=sc, executor=ex, wave_tile_bytes=wb, near_field_grid=ng)
        log(f"[bench] configuration given: {tune}")
    else:
        h.s.print_dbginfo = 1   # the autotuner prints a line per configuration it times
        tune = h.autotune(nv=1, reps=10)
        log(f"[bench] autotune (FP32-stored, 1 density per pass): {tune}")
    kernel_name = "queue_kernel<1>" if tune["executor"] == 1 else "stage_kernel<1>"

    def config_record(c):
        return {"executor": ["per stage", "persistent"][c["executor"]], "tile_bytes": int(c["tile_bytes"]),
                "schedule": c["schedule"], "wave_tile_bytes": int(c["wave_tile_bytes"]),
                "near_field_grid": int(c.get("near_field_grid", 0))}

    if args.passes_only > 0:
        # profiling runs: only the headline form (or the batch of
        # --passes-width densities), a fixed number of eager passes after two
        # warm-ups, so its dispatches are easy to pick out
        h.set_precision(32)
        h.upload()
        pw = args.passes_width

        def one_pass():
            if pw == 1:
                h.matvec_dev(x_dev.data_ptr(), y_dev.data_ptr(), stream.cuda_stream)
            else:
                h.matvec_multi_dev(pw, X_dev.data_ptr(), Y_dev.data_ptr(), stream.cuda_stream)
        for _ in range(2):
            one_pass()
        torch.cuda.synchronize()
        log(f"[bench] profiling: {args.passes_only} FP32-stored passes of {pw} densities follow")
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.passes_only):
            one_pass()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.passes_only
        nb = h.matvec_bytes()
        print(json.dumps({"workload": wname, "passes": args.passes_only, "ms_per_pass": round(ms, 5),
                          "alg_MB": round((nb["U"] + nb["B"] + nb["D"] + nb["vec"]) / 1e6, 2),
                          "tuned": config_record(tune)}), flush=True)
        return

    res = {}
    for bits in (64, 32):
        h.set_precision(bits)
        h.upload()
        nb = h.matvec_bytes()
        mat = nb["U"] + nb["B"] + nb["D"]
        r = {"nb": nb, "mat": mat}
        # one density per pass
        h.graph_capture("matvec", 1, x_dev.data_ptr(), y_dev.data_ptr())
        ms1 = time_graph(h, torch, stream, args.steps, args.warmup)
        r["ms1"], r["alg1"] = ms1, mat + nb["vec"]
        r["y"] = y_dev.cpu().numpy()
        log(f"[bench] FP{bits}-stored, 1 density: {ms1:.4f} ms, {r['alg1'] / ms1 / 1e6:.0f} GB/s algorithmic")
        r["prof"], r["union_ms"], r["sum_ms"] = launch_profile(h, torch, x_dev.data_ptr(), y_dev.data_ptr(), stream, 10)
        h.build_coulomb_dev(D_dev.data_ptr(), J_dev.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        r["J"] = J_dev.cpu().numpy()
        # the batch: the autotuner's choice for the width of its first pass
        widths = pass_widths(args.batch)
        if args.tuned:
            tuneb = tune
        else:
            tuneb = h.autotune(nv=widths[0], reps=5)
            log(f"[bench] autotune (FP{bits}-stored, {widths[0]} densities per pass): {tuneb}")
        h.graph_capture("matvec", args.batch, X_dev.data_ptr(), Y_dev.data_ptr())
        msb = time_graph(h, torch, stream, max(5, args.steps // 2), max(1, args.warmup // 2))
        passes = len(widths)
        algb = passes * mat + args.batch * nb["vec"]
        Yb = Y_dev.cpu().numpy()
        dev_single = 0.0
        for k in sorted({0, args.batch - 1}):
            h.matvec_dev(X_dev[k].data_ptr(), y_dev.data_ptr(), stream.cuda_stream)
            torch.cuda.synchronize()
            y1 = y_dev.cpu().numpy()
            dev_single = max(dev_single, float(np.linalg.norm(Yb[k] - y1) / np.linalg.norm(y1)))
        # per-launch record of one eager batched call (the last pass of it)
        h.set_phase_timing(True)
        h.matvec_multi_dev(args.batch, X_dev.data_ptr(), Y_dev.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        bprof = h.launch_stats()
        h.set_phase_timing(False)
        bphase = {}
        for p in ("up", "B", "y1", "down", "D", "gather", "pass"):
            q = [a for a in bprof if a["phase"] == p]
            if q:
                bphase[p] = {"ms_sum": round(sum(a["ms"] for a in q), 5), "launches": len(q),
                             "matrix_GB": round(sum(a["mat_bytes"] for a in q) / 1e9, 4)}
        bphase["union_ms"] = round(interval_union(bprof), 5)
        r["batch"] = {"densities": args.batch, "passes": widths, "ms": msb, "ms_per_density": msb / args.batch,
                      "phases_last_pass": bphase,
                      "alg": algb, "GBs": algb / msb / 1e6, "config_first_pass": config_record(tuneb),
                      "max_rel_dev_vs_single": dev_single}
        log(f"[bench] FP{bits}-stored, {args.batch} densities: {msb:.4f} ms per replay, {msb / args.batch:.4f} ms/density, "
            f"{algb / msb / 1e6:.0f} GB/s algorithmic")
        # the single-density configuration back for the next precision
        h.set_executor(tune["executor"])
        h.set_tile_bytes(tune["tile_bytes"])
        h.set_schedule(tune["schedule"])
        h.set_wave_tile_bytes(tune["wave_tile_bytes"])
        h.set_near_field_grid(tune["near_field_grid"])
        res[bits] = r
    r64, r32 = res[64], res[32]
    eps_prec = float(np.linalg.norm(r32["J"] - r64["J"]) / np.linalg.norm(r64["J"]))
    log(f"[bench] J FP32-stored vs FP64: {eps_prec:.3e}")

    # compression error against directly evaluated ERI rows
    eps_direct, npair_s, nrow_s = direct_rows_error(h, x, {"fp64": r64["y"], "fp32": r32["y"]}, npair=args.direct_pairs)

    # exchange build (configuration 4)
    exch = None
    if args.exchange:
        exch = exchange_measure(h, torch, D, D_dev, stream)
    jit = None if args.no_jit else jit_measure(torch, stream)

    # dominant kernel of the headline form
    prof = r32["prof"]
    alg1 = r32["alg1"]
    dom_gbs = alg1 / r32["union_ms"] / 1e6
    phases = {}
    for p in ("up", "B", "y1", "down", "D", "gather", "pass"):
        q = [a for a in prof if a["phase"] == p]
        if q:
            ms = sum(a["ms"] for a in q)
            mb = sum(a["mat_bytes"] for a in q)
            phases[p] = {"ms_sum": round(ms, 5), "launches": len(q), "matrix_GB": round(mb / 1e9, 4),
                         "GBs_matrix": round(mb / ms / 1e6, 1) if ms > 0 else None}

    h.s.print_dbginfo = 0
    cpu = None
    eps_oracle = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_bind
        with tempfile.TemporaryDirectory() as d:
            rp = os.path.join(d, "rep.h2")
            h.save(rp)
            c = cpu_baseline(rp, x, reps=10, one_core=not args.no_one_core)
            o = oracle_bind.Oracle(rp)
            J_cpu = o.coulomb(D)
            o.close()
        eps_oracle = {"fp64": float(np.linalg.norm(r64["J"] - J_cpu) / np.linalg.norm(J_cpu)),
                      "fp32": float(np.linalg.norm(r32["J"] - J_cpu) / np.linalg.norm(J_cpu))}
        alg64 = r64["alg1"]
        cpu = {"value": round(alg64 / c["all"]["median_ms"] / 1e6, 2), "unit": "GB/s", "cores": c["all"]["threads"],
               "kind": "port",
               "sample": f"oracle FP64 matvec of the same {wname} representation (C restatement of "
                         f"src/H2ERI_matvec.c), 1 warm-up + 10 timed, median {c['all']['median_ms']:.2f} ms; "
                         f"OMP_PROC_BIND=close, OMP_PLACES=cores",
               "median_ms": round(c["all"]["median_ms"], 3),
               "host": host_info()}
        if "one" in c:
            cpu["one_core"] = {"median_ms": round(c["one"]["median_ms"], 3),
                               "value": round(alg64 / c["one"]["median_ms"] / 1e6, 2)}
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{wname.lower().replace('(', '').replace(')', '')}_j_matvec.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            rec = json.load(f)
        # a recording of another launch structure does not describe this plan
        if rec.get("launches_per_matvec") == len(prof) and rec.get("kernel") == kernel_name:
            traffic = rec.get("traffic_bytes_per_matvec")
    b64, b32 = r64["batch"], r32["batch"]
    fp64_plain = r64["ms1"]
    out = {
        "metric": METRIC,
        "value": round(alg1 / r32["ms1"] / 1e6, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms1"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-stored U/B, f64 D and accumulation",
        "data": f"synthetic (seeded {wname} geometry and densities)",
        "config": {"workload": f"{wname} cc-pVDZ J matvec, reltol {reltol:g}, 1 density per step; "
                               f"{args.batch} densities per graph replay beside it",
                   "num_bf": h.nbf, "num_shell_pairs": h.s.num_sp, "krnl_mat_size": h.n, "H2_MB": sizes,
                   "alg_MB_fp32": round(alg1 / 1e6, 2), "alg_MB_fp64": round(r64["alg1"] / 1e6, 2),
                   "cache_resident": alg1 < MALL_BYTES, "timing": "hipGraph replays, HIP events",
                   "autotune": config_record(tune)},
        "fp64": {"ms": round(fp64_plain, 5), "GBs": round(r64["alg1"] / fp64_plain / 1e6, 2)},
        "fp32_speedup_vs_fp64": round(fp64_plain / r32["ms1"], 4),
        "batched": {f"fp{bits}": {k: (float(f"{v:.5g}") if isinstance(v, float) else v)
                                  for k, v in res[bits]["batch"].items() if k != "alg"} for bits in (64, 32)},
        "fp32_batched": {"densities": b32["densities"], "ms_per_density": round(b32["ms_per_density"], 5),
                         "speedup_vs_fp64_plain": round(fp64_plain / b32["ms_per_density"], 3),
                         "GBs_algorithmic": round(b32["GBs"], 2),
                         "roofline_frac": round(b32["GBs"] / HBM_PEAK_GBS, 4)},
        "rel_err_J_fp32_vs_fp64": eps_prec,
        "precision_gate_met": eps_prec <= EPS_PREC_GATE,
        "rel_err_J_vs_oracle": eps_oracle,
        "rel_err_vs_direct_eri": {"fp64": eps_direct["fp64"], "fp32": eps_direct["fp32"],
                                  "sample": f"rows of {npair_s} seeded random bra shell pairs ({nrow_s} rows of y)"},
        "build_s": round(t_build, 2),
        "exchange_fp64": exch,
        "bd_jit_fp64": jit,
        "roofline": {"bound": "hbm", "achieved": round(dom_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kernel_name, "launches_per_matvec": len(prof),
                     "method": "algorithmic bytes of one FP32-stored matvec / union of its launch intervals "
                               "(HIP events around every launch on its own stream, both streams)",
                     "alg_bytes_per_matvec": round(alg1, 1),
                     "kernel_ms_union": round(r32["union_ms"], 6),
                     "kernel_ms_sum": round(r32["sum_ms"], 6),
                     "avg_launch_ms": round(r32["sum_ms"] / len(prof), 6),
                     "streamed_matrix_bytes_per_matvec": sum(p["mat_bytes"] for p in prof),
                     # the bytes the plan actually streams (a low-rank block's smaller half twice)
                     # over the same union: how close the launches come to the HBM peak
                     "streamed_frac": round(sum(p["mat_bytes"] for p in prof) / (r32["union_ms"] * 1e6)
                                            / HBM_PEAK_GBS, 4)},
        "phases_fp32": phases,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def jit_measure(torch, stream, cfg=1, reps=3, host_compare=True):
    """The on-the-fly form (BD_JIT = 1: B and D evaluated from the shell
    pairs in every matvec) on C30H62 at reltol 1e-6: ms per FP64 matvec
    (HIP events, device vectors), with the integral work of one matvec."""
    import numpy as np

    import h2eri
    from h2eri import molecules as M
    atoms, txt, _, reltol = M.workload(cfg)
    t0 = time.perf_counter()
    hj = h2eri.H2ERI(1e-10, 1e-10, reltol)
    hj.read_mol_text(txt).process_shells().partition().build_h2(1)
    t_build = time.perf_counter() - t0
    st = hj.jit_stats()
    x = torch.from_numpy(np.random.default_rng(17).standard_normal(hj.n)).to("cuda:0")
    y = torch.empty_like(x)
    hj.upload()
    hj.matvec_dev(x.data_ptr(), y.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        hj.matvec_dev(x.data_ptr(), y.data_ptr(), stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    best = min(ms)
    log(f"[bench] on-the-fly B/D ({M.workload_name(cfg)}): {best:.1f} ms per matvec, "
        f"{st['eri_values'] / best / 1e6:.2f} G integrals/s")
    rec = {"workload": f"{M.workload_name(cfg)} cc-pVDZ, reltol {reltol:g}", "krnl_mat_size": hj.n,
           "ms_per_matvec": round(best, 3), "ms_all": [round(m, 3) for m in ms],
           "integrals_per_matvec": st["eri_values"], "shell_quartets_per_matvec": st["quartets"],
           "primitive_quartets_per_matvec": st["prim_quartets"], "row_tasks": st["tasks"],
           "G_integrals_per_s": round(st["eri_values"] / best / 1e6, 3),
           "resident_MB": round(hj.upload() / 1e6, 1), "build_s": round(t_build, 2)}
    hj.close()
    if host_compare:
        rec["host_engine_same_blocks"] = jit_host_compare(torch, stream)
    return rec


def jit_host_compare(torch, stream, cfg=0, reps=3):
    """The reference's CPU on-the-fly path needs SIMINT, absent here; the
    like-for-like CPU figure is this project's host integral engine
    evaluating the same (dense) B and D blocks, which a dense build times
    (the B and D construction timers).  Beside it, the GPU on-the-fly matvec
    on the same molecule (C10H22)."""
    import numpy as np

    import h2eri
    from h2eri import molecules as M
    atoms, txt, _, reltol = M.workload(cfg)
    hd = h2eri.H2ERI(1e-10, 1e-10, reltol)
    hd.set_block_compression(0)
    hd.build_from_mol(text=txt)
    cpu_s = float(hd.s.timers[2] + hd.s.timers[3])   # B_BUILD_TIMER_IDX, D_BUILD_TIMER_IDX
    hd.close()
    hj = h2eri.H2ERI(1e-10, 1e-10, reltol)
    hj.read_mol_text(txt).process_shells().partition().build_h2(1)
    x = torch.from_numpy(np.random.default_rng(18).standard_normal(hj.n)).to("cuda:0")
    y = torch.empty_like(x)
    hj.matvec_dev(x.data_ptr(), y.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        hj.matvec_dev(x.data_ptr(), y.data_ptr(), stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    hj.close()
    threads = cpu_threads()
    log(f"[bench] {M.workload_name(cfg)} B and D blocks: host engine {cpu_s:.2f} s ({threads} threads), "
        f"GPU on-the-fly matvec {min(ms):.1f} ms")
    return {"workload": f"{M.workload_name(cfg)} cc-pVDZ, reltol {reltol:g}", "cpu_s": round(cpu_s, 3),
            "cpu_threads": threads, "gpu_ms_per_matvec": round(min(ms), 3),
            "ratio": round(cpu_s * 1e3 / min(ms), 1),
            "note": "CPU: the host integral engine evaluating the dense B and D blocks (B and D construction "
                    "timers of a dense build); GPU: one on-the-fly matvec, which evaluates the same blocks "
                    "from both sides and applies them"}


def exchange_measure(h, torch, D, D_dev, stream):
    """K build (H2ERI_build_exchange_dev) on the representation, FP64: ms per
    build (plan uploaded once), the MFMA rate of the whole build, and K
    against the oracle."""
    import numpy as np
    t0 = time.perf_counter()
    kst = h.exchange_stats()   # plans and uploads the exchange data, runs the basis products
    k_upload_s = time.perf_counter() - t0
    K_dev = torch.empty_like(D_dev)
    h.build_exchange_dev(D_dev.data_ptr(), K_dev.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(3):
        h.build_exchange_dev(D_dev.data_ptr(), K_dev.data_ptr(), stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1) / 3
    Kh = K_dev.cpu().numpy()
    exch = {"ms": round(k_ms, 3), "upload_s": round(k_upload_s, 2), "resident_GB": round(kst["resident_bytes"] / 1e9, 2),
            "basis_gemm_gflop": round(kst["gemm_flops"] / 1e9, 2), "terms": int(kst["terms"]),
            "rel_asym": float(np.linalg.norm(Kh - Kh.T) / np.linalg.norm(Kh))}
    log(f"[bench] exchange K: {k_ms:.2f} ms per build (plan + basis products {k_upload_s:.2f} s once)")
    h.exchange_workbuf_free()
    return exch


# ---------------------------------------------------------------------------
# molecule group of the batching workload
# ---------------------------------------------------------------------------
def group_workload(args):
    """The batching workload's uneven molecules: 8 alkanes C20..C160 (reltol
    1e-6, FP32-stored U/B) as one molecule group, one persistent launch over
    all their tasks per step, against the same matvecs one after another."""
    import numpy as np
    import torch

    import h2eri
    from h2eri import molecules as M

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    hs, X, Y, ns, xs = [], [], [], [], []
    t0 = time.perf_counter()
    for n in range(20, 161, 20):
        r = M.rng_for(300 + n)
        at = M.alkane(n, r)
        D = M.density(at, 4 * n + 1, r)
        hh = build_molecule(h2eri, M, at, 1e-6, f"C{n}H{2 * n + 2}")
        hh.set_precision(32)
        hh.upload()
        x = uncontract(hh.tree(), D, hh.n)
        hs.append(hh)
        xs.append(x)
        X.append(torch.from_numpy(x).to(dev))
        Y.append(torch.empty_like(X[-1]))
        ns.append(hh.nbf)
    t_build = time.perf_counter() - t0
    g = h2eri.Group(hs, [x.data_ptr() for x in X], [y.data_ptr() for y in Y])
    # the group's two modes (one grouped persistent launch; each molecule's
    # own pass in turn), timed by the library, the faster kept
    gt = g.autotune(reps=5)
    log(f"[bench] group modes: one grouped launch {gt['mode0_ms']:.3f} ms, molecules' own passes "
        f"{gt['mode1_ms']:.3f} ms; keeping mode {gt['mode']}")
    for _ in range(args.warmup):
        g.matvec(stream.cuda_stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        g.matvec(stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_group = e0.elapsed_time(e1) / args.steps
    g.status()   # raises if a dependency wait of the persistent launch exceeded its bound
    Yg = [y.cpu().numpy() for y in Y]
    e0.record(stream)
    for _ in range(args.steps):
        for hh, x, y in zip(hs, X, Y):
            hh.matvec_dev(x.data_ptr(), y.data_ptr(), stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_seq = e0.elapsed_time(e1) / args.steps
    dev_group = max(float(np.linalg.norm(a - y.cpu().numpy()) / np.linalg.norm(a)) for a, y in zip(Yg, Y))
    alg_group = sum(hh.matvec_bytes()["alg"] for hh in hs)
    log(f"[bench] 8 molecules: group {ms_group:.3f} ms vs one after another {ms_seq:.3f} ms")
    g.close()
    cpu = None
    if not args.no_cpu:
        tot, per = 0.0, []
        with tempfile.TemporaryDirectory() as d:
            for k, hh in enumerate(hs):
                rp = os.path.join(d, f"m{k}.h2")
                hh.save(rp)
                c = cpu_baseline(rp, xs[k], reps=10, one_core=False)
                per.append(round(c["all"]["median_ms"], 3))
                tot += c["all"]["median_ms"]
        alg64 = 0.0
        for hh in hs:
            hh.set_precision(64)
            alg64 += hh.matvec_bytes()["alg"]
        cpu = {"value": round(alg64 / tot / 1e6, 2), "unit": "GB/s", "cores": cpu_threads(), "kind": "port",
               "sample": "oracle FP64 matvec of each of the 8 molecules, 1 warm-up + 10 timed each, medians summed "
                         f"({tot:.2f} ms); OMP_PROC_BIND=close, OMP_PLACES=cores",
               "median_ms_per_molecule": per, "host": host_info()}
    out = {"metric": METRIC, "value": round(alg_group / ms_group / 1e6, 2), "unit": "GB/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_group, 5), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32-stored U/B, f64 D and accumulation",
           "data": "synthetic (seeded alkane geometries and densities)",
           "config": {"workload": "8 alkanes C20H42..C160H322 (reltol 1e-6) as one molecule group, one matvec each "
                                  "per step", "num_bf": ns, "timing": "back-to-back group launches, HIP events"},
           "ms_sequential_matvecs": round(ms_seq, 4), "rel_dev_group_vs_single": dev_group,
           "group_mode": {"kept": gt["mode"], "grouped_launch_ms": round(gt["mode0_ms"], 4),
                          "own_passes_ms": round(gt["mode1_ms"], 4)},
           "build_s": round(t_build, 1),
           "roofline": {"bound": "hbm", "achieved": round(alg_group / ms_group / 1e6, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(alg_group / ms_group / 1e6 / HBM_PEAK_GBS, 4), "traffic": None,
                        "kernel": "queue_kernel<1, group>" if gt["mode"] == 0 else "stage_kernel<1>",
                        "launches_per_matvec": 1 if gt["mode"] == 0 else None,
                        "method": "algorithmic bytes of the 8 matvecs / time per group matvec"},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4])
    ap.add_argument("--batch", type=int, default=16, help="densities per graph replay of the batched form")
    ap.add_argument("--direct-pairs", type=int, default=256, help="shell pairs whose rows are evaluated directly")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-one-core", action="store_true")
    ap.add_argument("--exchange", action="store_true", help="time the exchange build too (default with --config 4)")
    ap.add_argument("--no-jit", action="store_true", help="skip the on-the-fly (BD_JIT) matvec record")
    ap.add_argument("--jit-only", type=int, default=0, metavar="REPS",
                    help="profiling: only the on-the-fly (BD_JIT) record, REPS timed matvecs")
    ap.add_argument("--no-jit-cpu", action="store_true",
                    help="skip the host-engine comparison of the on-the-fly record")
    ap.add_argument("--tuned", default="", help="EXECUTOR,TILE_BYTES,SCHEDULE[,WAVE_TILE_BYTES[,NEAR_FIELD_GRID]] instead of running the "
                    "autotuner (profiling runs, where the profiler distorts the tuner's timings)")
    ap.add_argument("--passes-only", type=int, default=0,
                    help="profiling: run only this many eager FP32-stored single-density passes (needs --tuned)")
    ap.add_argument("--passes-width", type=int, default=1, choices=[1, 16],
                    help="profiling: densities per pass of the --passes-only passes")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="run a plain device copy of this many MB first (HBM counter calibration)")
    ap.add_argument("--cpu-child", nargs=4, metavar=("REP", "X", "THREADS", "REPS"), help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:
        rep, xp, thr, reps = args.cpu_child
        cpu_child(rep, xp, int(thr), int(reps))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (got --gpus {args.gpus}); the single-GPU "
              "matvec is the unit of work and no multi-GPU path is provided.", file=sys.stderr)
        sys.exit(2)
    if args.config == 4:
        args.exchange = True
    if args.jit_only:
        import torch
        print(json.dumps(jit_measure(torch, torch.cuda.current_stream(), reps=args.jit_only,
                                     host_compare=not args.no_jit_cpu)), flush=True)
        return
    if args.config == 3:
        group_workload(args)
    else:
        single_workload(args)


if __name__ == "__main__":
    main()
