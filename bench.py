#!/usr/bin/env python3
"""Benchmark of the H2 J-build matvec on one MI355X.

Workload (default): linear alkane C30H62, cc-pVDZ (760 Cartesian basis
functions), H2 relative tolerance 1e-6, one seeded damped density with 121
occupied columns.  The H2 representation is constructed on the host, uploaded
once, and each step is one H2 matvec y = A_H2 x on the uncontracted density,
with x and y resident in HBM.  Both the FP64 form (the drop-in) and the
FP32-stored U/B form (FP64 near field and accumulation) are timed; `value` is
the FP32-stored form's algorithmic HBM bandwidth.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "h2p-eri_amd"))

METRIC = "J-build H2 matvec ms & HBM GB/s, FP32-stored vs FP64 on 1 GPU, rel.err ≤1e-7"
HBM_PEAK_GBS = 8000.0
MALL_BYTES = 256e6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(h, x, y, stream, steps, warmup, torch):
    for _ in range(warmup):
        h.matvec_dev(x.data_ptr(), y.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        h.matvec_dev(x.data_ptr(), y.data_ptr(), stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def launch_profile(h, x, y, stream, steps, torch):
    """Per-launch HIP-event times averaged over `steps` timed matvecs."""
    h.set_phase_timing(True)
    acc = None
    for _ in range(steps):
        h.matvec_dev(x.data_ptr(), y.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        st = h.launch_stats()
        if acc is None:
            acc = [dict(r) for r in st]
        else:
            for a, r in zip(acc, st):
                a["ms"] += r["ms"]
    h.set_phase_timing(False)
    for a in acc:
        a["ms"] /= steps
    return acc


def cpu_baseline(rep_path, x, seconds):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind
    nthr = os.cpu_count() or 1
    try:
        nthr = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    nthr = min(nthr, int(os.environ.get("OMP_NUM_THREADS", nthr)))
    lib = oracle_bind.load()
    lib.orc_set_threads(nthr)
    o = oracle_bind.Oracle(rep_path)
    y = o.matvec(x)   # warm-up (first touch)
    t0 = time.perf_counter()
    n = 0
    while n < 10 or (time.perf_counter() - t0 < seconds and n < 200):
        y = o.matvec(x)
        n += 1
    ms = (time.perf_counter() - t0) / n * 1e3
    return ms, nthr, n, y, o


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="run a plain device copy of this many MB first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (got --gpus {args.gpus}); the single-GPU "
              "kernel is the unit of work and no multi-GPU path is provided.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import h2eri
    from h2eri import molecules as M

    t0 = time.perf_counter()
    atoms, mol, D, reltol = M.workload(args.config)
    h = h2eri.H2ERI(1e-10, 1e-10, reltol)
    h.read_mol_text(mol)
    h.process_shells()
    h.partition()
    log(f"[bench] config {args.config}: {len(atoms)} atoms, {h.nbf} basis functions, "
        f"{h.s.num_sp} shell pairs, N={h.n}; building H2 ...")
    h.build_h2()
    t_build = time.perf_counter() - t0
    sizes = {k: h.s.mat_size[i] * 8 / 1e6 for i, k in enumerate(["U_MB", "B_MB", "D_MB"])}
    log(f"[bench] H2 built in {t_build:.1f} s: U {sizes['U_MB']:.1f} MB, B {sizes['B_MB']:.1f} MB, "
        f"D {sizes['D_MB']:.1f} MB")

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    D_dev = torch.from_numpy(D).to(dev)
    J_dev = torch.empty_like(D_dev)
    tr = h.tree()
    # uncontracted density x (src/H2ERI_build_Coulomb.c:26-62 layout)
    x = np.empty(h.n)
    sp, sb, bf = tr["sp_shell_idx"], tr["sp_bfp_sidx"], tr["shell_bf_sidx"]
    npair = tr["num_sp"]
    for i in range(npair):
        s0, s1 = sp[i], sp[i + npair]
        r0, nr = bf[s0], bf[s0 + 1] - bf[s0]
        c0, nc = bf[s1], bf[s1 + 1] - bf[s1]
        blk = D[c0:c0 + nc, r0:r0 + nr] * (1.0 if s0 == s1 else 2.0)
        x[sb[i]:sb[i + 1]] = blk.reshape(-1)
    x_dev = torch.from_numpy(x).to(dev)
    y_dev = torch.empty_like(x_dev)

    if args.calibrate_copy_mb > 0:
        n = args.calibrate_copy_mb * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {n * 8} bytes read and written")
        del a, b

    res = {}
    for bits in (64, 32):
        h.set_precision(bits)
        resident = h.upload()
        nb = h.matvec_bytes()
        ms = time_steps(h, x_dev, y_dev, stream, args.steps, args.warmup, torch)
        prof = launch_profile(h, x_dev, y_dev, stream, max(3, min(args.steps, 20)), torch)
        h.build_coulomb_dev(D_dev.data_ptr(), J_dev.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        res[bits] = dict(ms=ms, alg=nb["alg"], ref=nb["ref_accounting"], resident=resident, nb=nb, prof=prof,
                         J=J_dev.cpu().numpy(), y=y_dev.cpu().numpy())
        log(f"[bench] FP{bits}-stored U/B: {ms:.4f} ms/matvec, {nb['alg'] / ms / 1e6:.0f} GB/s algorithmic "
            f"({nb['alg'] / 1e6:.1f} MB/matvec)")
    r64, r32 = res[64], res[32]
    eps_prec = float(np.linalg.norm(r32["J"] - r64["J"]) / np.linalg.norm(r64["J"]))

    # dominant kernel: the tile kernel (bgemv_kernel), all its launches of the
    # FP32-stored form; algorithmic matrix bytes (U twice, B and D once)
    tiles = [p for p in r32["prof"] if p["kind"] == "tiles"]
    tile_ms = sum(p["ms"] for p in tiles)
    alg_mat = r32["nb"]["U"] + r32["nb"]["B"] + r32["nb"]["D"]
    dom_gbs = alg_mat / tile_ms / 1e6
    cpu = None
    eps_parity = None
    if not args.no_cpu:
        with tempfile.TemporaryDirectory() as d:
            rp = os.path.join(d, "rep.h2")
            h.save(rp)
            cpu_ms, nthr, nrep, y_cpu, o = cpu_baseline(rp, x, args.cpu_seconds)
            J_cpu = o.coulomb(D)
            o.close()
        eps_parity = float(np.linalg.norm(r64["J"] - J_cpu) / np.linalg.norm(J_cpu))
        cpu = {"value": round(r64["alg"] / cpu_ms / 1e6, 2), "unit": "GB/s", "cores": nthr, "kind": "port",
               "sample": f"{nrep} oracle FP64 matvecs on the same C30H62 representation, "
                         f"{cpu_ms:.2f} ms each"}
    out = {
        "metric": METRIC,
        "value": round(r32["alg"] / r32["ms"] / 1e6, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-stored U/B, f64 D and accumulation",
        "data": "synthetic (seeded C30H62 geometry and density)",
        "config": {"workload": f"C30H62 cc-pVDZ J matvec, reltol {reltol:g}, 1 density",
                   "num_bf": h.nbf, "krnl_mat_size": h.n, "H2_MB": sizes,
                   "alg_MB_fp32": round(r32["alg"] / 1e6, 2), "alg_MB_fp64": round(r64["alg"] / 1e6, 2),
                   "cache_resident": r32["alg"] < MALL_BYTES},
        "fp64": {"ms": round(r64["ms"], 5), "GBs": round(r64["alg"] / r64["ms"] / 1e6, 2)},
        "fp32_speedup_vs_fp64": round(r64["ms"] / r32["ms"], 4),
        "rel_err_J_fp32_vs_fp64": eps_prec,
        "rel_err_J_fp64_vs_oracle": eps_parity,
        "build_s": round(t_build, 2),
        "roofline": {"bound": "hbm", "achieved": round(dom_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "bgemv_kernel", "launches": len(tiles),
                     "kernel_ms_per_matvec": round(tile_ms, 5),
                     "alg_bytes_per_launch": round(alg_mat / len(tiles), 1),
                     "avg_launch_ms": round(tile_ms / len(tiles), 6),
                     "streamed_bytes_per_matvec": sum(p["mat_bytes"] for p in tiles)},
        "phases_ms_fp32": {p: round(sum(q["ms"] for q in r32["prof"] if q["phase"] == p), 5)
                           for p in ("up", "B", "y1", "down", "D", "gather")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
