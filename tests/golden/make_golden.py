"""Writes the small golden fixtures in this directory from the CPU oracle.

Each fixture holds the inputs (tensor, initial CPD, check settings) and the
oracle's outputs (factors, lambda, fit trace, check state).  Run from the repo
root:  python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent


def lowrank(dims, R, noise, seed):
    rng = np.random.default_rng(seed)
    true = [rng.standard_normal((d, R)) for d in dims]
    letters = "abcdefgh"[:len(dims)]
    T = np.einsum(",".join(l + "z" for l in letters) + "->" + letters, *true)
    return T + noise * rng.standard_normal(dims)


def write(name, T, init, check, dup=False):
    tr = {}
    out = O.als_optimize(T, init, check=check, trace=tr)
    np.savez_compressed(
        OUT / f"{name}.npz", T=T, rank=init.rank,
        **{f"init_{n}": f for n, f in enumerate(init.factors)}, init_lam=init.lam,
        **{f"out_{n}": f for n, f in enumerate(out.factors)}, out_lam=out.lam,
        check_kind=1 if isinstance(check, O.FitCheck) else 0,
        tol=getattr(check, "tolerance", 0.0), max_counter=check.max_counter,
        ref_norm=getattr(check, "ref_norm", 0.0), fits=np.array(tr["fits"]),
        total_iter=getattr(check, "total_iter", 0), final_fit=getattr(check, "final_fit", 0.0),
        fallbacks=tr["fallbacks"])
    print(name, T.shape, "rank", init.rank, "fallbacks", tr["fallbacks"], "fits", len(tr["fits"]))


def main():
    T = lowrank((16, 12, 10), 4, 1e-2, 11)
    init = O.random_CPD(T, 4, np.random.default_rng(12))
    write("order3_fitcheck", T, init, O.FitCheck(1e-10, 12, float(np.linalg.norm(T))))

    T = lowrank((14, 13, 11), 3, 1e-3, 21)
    init = O.random_CPD(T, 3, np.random.default_rng(22))
    write("order3_converge", T, init, O.FitCheck(1e-6, 100, float(np.linalg.norm(T))))

    T = lowrank((8, 7, 6, 5), 3, 1e-2, 13)
    init = O.random_CPD(T, 3, np.random.default_rng(14))
    write("order4_nocheck", T, init, O.NoCheck(8))

    T = lowrank((12, 9), 4, 1e-2, 15)
    init = O.random_CPD(T, 4, np.random.default_rng(16))
    write("order2_nocheck", T, init, O.NoCheck(5))

    # duplicated component: every V is singular, so every solve takes the QR path
    T = np.random.default_rng(17).standard_normal((10, 9, 8))
    init = O.random_CPD(T, 5, np.random.default_rng(18))
    for f in init.factors:
        f[:, 1] = f[:, 0]
    write("dup_fallback", T, init, O.NoCheck(4))


if __name__ == "__main__":
    main()
