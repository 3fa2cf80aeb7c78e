#!/usr/bin/env python3
"""Plain FP64 CP-ALS sweep throughput on one MI355X.

Workload: an order-3 Float64 512x512x512 tensor (1 GiB, column-major) equal to a
rank-64 sum of N(0,1) factor outer products plus N(0, 1e-2) noise, decomposed at
rank 64 by normal-equation ALS (KRPFreeNormal semantics) from a random_CPD init,
with FitCheck evaluating the fit after every sweep.  One step = one ALS sweep
(per mode: MTTKRP, Gram Hadamard, pivoted-Cholesky solve, row_norm, Gram update;
then the fit).  The tensor is resident in HBM before the timed region starts.

Prints one JSON line: sweeps/s, the roofline of the dominant MTTKRP GEMM kernel
(HIP events on the plan's stream over the timed region), and the CPU oracle's
throughput on a bounded sample of the same workload on this host's cores.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
METRIC = ("CP-ALS sweeps/s + time-to-fit within 1e-6 of FP64: mixed/warm vs plain FP64")
FP64_MFMA_PEAK_TFLOPS = 78.6
TRAFFIC_FILE = ROOT / "profiles" / "mttkrp_gemm_512cubed_r64_hbm_traffic.txt"


def synthetic_tensor(I: int, R: int, seed: int) -> np.ndarray:
    """T(i,j,k) = sum_r A(i,r) B(j,r) C(k,r) + 1e-2 N(0,1), returned as an (I,I,I)
    Fortran-ordered view (i fastest, Julia's layout)."""
    rng = np.random.default_rng(seed)
    A, B, Cf = (rng.standard_normal((I, R)) for _ in range(3))
    KR = (Cf[:, None, :] * B[None, :, :]).reshape(I * I, R)     # row = (k, j)
    Tc = KR @ A.T                                                # Tc[(k,j), i]
    del KR
    step = 16384
    for r0 in range(0, I * I, step):                              # noise in slabs
        Tc[r0:r0 + step] += 1e-2 * rng.standard_normal((min(step, I * I - r0), I))
    return Tc.reshape(I, I, I).transpose(2, 1, 0)


def sweep_flops(dims, R) -> float:
    """Algorithmic flops of one dimension-tree sweep (order 3): the two MTTKRP
    GEMMs (2 * prod(I) * R each), the Hadamard reductions of the partials, the
    Gram updates and the R x R solves."""
    I, J, K = dims
    gemm = 4.0 * I * J * K * R
    ttv = 2.0 * (I * J * R + I * J * R)        # left partial -> modes 0, 1 (s = 2)
    gram = sum(2.0 * d * R * R for d in dims)
    solve = sum(R ** 3 / 3.0 + 2.0 * d * R * R for d in dims)
    return gemm + ttv + gram + solve


def read_traffic():
    """HBM bytes per launch of the left/right GEMM, measured with rocprofv3 PMC
    counters (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated); None if absent."""
    if not TRAFFIC_FILE.exists():
        return None
    out = {}
    for line in TRAFFIC_FILE.read_text().splitlines():
        m = re.match(r"\s*traffic_bytes_per_launch_(left|right)\s*[:=]\s*([0-9.eE+]+)", line)
        if m:
            out[m.group(1)] = float(m.group(2))
    return out or None


def cpu_baseline(T: np.ndarray, R: int, seed: int):
    """The CPU oracle on this host: full sweeps of the two reference MTTKRP
    forms on the same tensor, each repeated until a few seconds have passed."""
    sys.path.insert(0, str(ROOT))
    import oracle as O
    try:
        from threadpoolctl import threadpool_info
        cores = max([d.get("num_threads", 1) for d in threadpool_info()
                     if d.get("user_api") == "blas"] or [1])
    except Exception:  # threadpoolctl missing: the env setting is what BLAS uses
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    init = O.random_CPD(T, R, np.random.default_rng(seed))
    res = {}
    for name, form, budget in (("KRPFreeNormal (rank-slice GEMVs)", "slices", 6.0),
                               ("KRPNormal (explicit KRP + GEMM)", "gemm", 4.0)):
        cp, n, t0 = init, 0, time.perf_counter()
        while True:
            cp = O.als_optimize(T, cp, check=O.NoCheck(1), form=form)
            n += 1
            el = time.perf_counter() - t0
            if el >= budget:
                break
        res[name] = (n, el)
    best = max(res, key=lambda k: res[k][0] / res[k][1])
    n, el = res[best]
    return {
        "value": n / el, "unit": "sweeps/s", "cores": cores, "kind": "port",
        "sample": (f"{best}: {n} full sweep(s) of the {'x'.join(map(str, T.shape))} rank-{R} "
                   f"workload in {el:.2f} s "
                   f"(numpy/OpenBLAS oracle, {cores} BLAS threads)"),
        "variants": {k: {"sweeps": v[0], "seconds": round(v[1], 3),
                         "sweeps_per_s": v[0] / v[1]} for k, v in res.items()},
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed ALS sweeps")
    ap.add_argument("--warmup", type=int, default=5, help="untimed ALS sweeps")
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--rank", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus}: this engine runs one CP-ALS problem per GPU and "
              "only the one-GPU measurement is defined (no multi-GPU path is built)",
              file=sys.stderr)
        return 2

    sys.path.insert(0, str(ROOT))
    import __graft_entry__ as G
    E = G.load_package()
    L = E._lib

    I, R = args.size, args.rank
    t_gen = time.perf_counter()
    T = synthetic_tensor(I, R, 2024)
    print(f"[bench] tensor {I}^3 generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr,
          flush=True)

    with E.ALSPlan(T, R) as plan:
        plan.random_cpd(7)                                   # random_CPD on the device
        nrm = plan.tensor_norm()
        if args.warmup > 0:
            plan.run(E.NoCheck(args.warmup))
        plan.synchronize()
        plan.set_profiling(True)
        check = E.FitCheck(0.0, args.steps, nrm)             # tol 0: exactly `steps` sweeps
        plan.synchronize()
        t0 = time.perf_counter()
        info = plan.run(check)                               # returns after a stream sync
        t1 = time.perf_counter()
        stats = {k: plan.kernel_stats(code) for k, code in
                 (("left", L.ITCPD_KERNEL_GEMM_LEFT), ("right", L.ITCPD_KERNEL_GEMM_RIGHT))}
        plan.set_profiling(False)
    if info["sweeps"] != args.steps:
        raise RuntimeError(f"ran {info['sweeps']} sweeps, expected {args.steps}")
    secs = t1 - t0
    ms_per_step = 1e3 * secs / args.steps
    print(f"[bench] {args.steps} sweeps in {secs * 1e3:.1f} ms; final fit {check.final_fit:.12f}",
          file=sys.stderr, flush=True)

    kern = {}
    for side, st in stats.items():
        if st["launches"]:
            avg_ms = st["total_ms"] / st["launches"]
            fl = st["flops"] / st["launches"]
            kern[side] = {"launches": st["launches"], "avg_ms": avg_ms, "flops_per_launch": fl,
                          "tflops": fl / (avg_ms * 1e-3) / 1e12,
                          "T_bytes_per_launch": st["bytes"] / st["launches"]}
    dom = max(kern, key=lambda k: kern[k]["avg_ms"] * kern[k]["launches"])
    traffic = read_traffic()
    achieved = kern[dom]["tflops"]
    sweep_tf = sweep_flops((I, I, I), R) / (ms_per_step * 1e-3) / 1e12
    roofline = {
        "bound": "mfma", "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": achieved / FP64_MFMA_PEAK_TFLOPS,
        "traffic": traffic.get(dom) if traffic else None,
        "kernel": f"MTTKRP {dom} GEMM (2*I*J*K*R = {kern[dom]['flops_per_launch']:.4g} flop/launch)",
        "kernels": kern,
        "sweep_achieved": sweep_tf, "sweep_frac": sweep_tf / FP64_MFMA_PEAK_TFLOPS,
    }

    out = {
        "metric": METRIC, "value": args.steps / secs, "unit": "sweeps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded rank-64 N(0,1) factor sum + N(0,1e-2) noise)",
        "config": {"workload": f"plain FP64 CP-ALS, {I}x{I}x{I} Float64, rank {R}, random_CPD "
                               f"init, FitCheck fit every sweep, {args.steps} sweeps",
                   "dims": [I, I, I], "rank": R, "precision": "fp64"},
        "final_fit": check.final_fit,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        print("[bench] CPU oracle baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(T, R, 7)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
