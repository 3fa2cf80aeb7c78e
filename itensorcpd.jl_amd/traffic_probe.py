"""HBM-traffic probe for the MTTKRP GEMM kernels, meant to run under
``rocprofv3 --pmc FETCH_SIZE`` and ``rocprofv3 --pmc WRITE_SIZE``.

It launches, in this order:
  1. one device-to-device copy of a 1 GiB float64 buffer (torch ``clone``): the
     calibration kernel, 1 GiB read + 1 GiB written;
  2. ``norm(T)`` of the 512^3 Float64 tensor (reads 1 GiB, writes bytes);
  3. ``--sweeps`` plain FP64 CP-ALS sweeps at rank 64 on that tensor.

With ``--workload precision`` steps 2-3 use the 224^4 FP32-stored rank-128 tensor
of the precision workload instead, and run ``--sweeps`` sweeps in each of FP64,
FP32 and BF16x3 arithmetic (the LDS ring GEMMs of the reduced forms).

Counter values per dispatch divided by the calibration kernel's value per byte
give the HBM bytes each kernel really moved.  torch is imported first: it
carries its own HIP runtime, which must be initialised before libitcpd_hip.so.
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

import numpy as np


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--rank", type=int, default=64)
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--workload", choices=["plain", "precision"], default="plain")
    args = ap.parse_args()

    import torch
    x = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0")
    x.fill_(1.0)
    y = x.clone()                         # the 1 GiB copy kernel
    torch.cuda.synchronize()
    del x, y
    torch.cuda.empty_cache()
    print("[probe] copy calibration done", flush=True)

    root = Path(__file__).resolve().parent.parent
    sys.path.insert(0, str(root))
    import __graft_entry__ as G
    import bench
    E = G.load_package()
    if args.workload == "precision":
        T = bench.c3_tensor(224, 128, 2025)
        with E.ALSPlan(T, 128, prec="fp64") as plan:
            del T
            plan.random_cpd(11)
            nrm = plan.tensor_norm()
            for prec in ("fp64", "fp32", "bf16x3"):
                plan.set_precision(prec)
                plan.run(E.NoCheck(args.sweeps))
                plan.synchronize()
                print(f"[probe] {prec}: {args.sweeps} sweeps done", flush=True)
        print(f"[probe] norm {nrm:.6e}", flush=True)
        return 0
    T = bench.synthetic_tensor(args.size, args.rank, 2024)
    with E.ALSPlan(T, args.rank) as plan:
        plan.random_cpd(7)
        nrm = plan.tensor_norm()
        plan.run(E.NoCheck(args.sweeps))
        plan.synchronize()
    print(f"[probe] norm {nrm:.6e}, {args.sweeps} sweeps done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
