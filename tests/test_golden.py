"""The oracle reproduces the committed golden fixtures (a regression check on the
oracle itself; the GPU suite compares the engine against the same files)."""
from pathlib import Path

import numpy as np
import pytest

import oracle as O

GOLD = Path(__file__).resolve().parent / "golden"
CASES = sorted(p.stem for p in GOLD.glob("*.npz"))


def load_case(name):
    z = np.load(GOLD / f"{name}.npz")
    N = z["T"].ndim
    init = O.CPD([z[f"init_{n}"] for n in range(N)], z["init_lam"])
    if int(z["check_kind"]) == 1:
        check = O.FitCheck(float(z["tol"]), int(z["max_counter"]), float(z["ref_norm"]))
    else:
        check = O.NoCheck(int(z["max_counter"]))
    return z, init, check


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_golden(name):
    z, init, check = load_case(name)
    tr = {}
    out = O.als_optimize(z["T"], init, check=check, trace=tr)
    for n, f in enumerate(out.factors):
        np.testing.assert_allclose(f, z[f"out_{n}"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(out.lam, z["out_lam"], rtol=1e-10)
    np.testing.assert_allclose(np.array(tr["fits"]), z["fits"], rtol=0, atol=1e-12)
    assert tr["fallbacks"] == int(z["fallbacks"])
    if isinstance(check, O.FitCheck):
        assert check.total_iter == int(z["total_iter"])
        assert abs(check.final_fit - float(z["final_fit"])) < 1e-12


def test_golden_set_covers_the_paths():
    assert {"order3_fitcheck", "order3_converge", "order4_nocheck", "order2_nocheck",
            "dup_fallback"} <= set(CASES)
    z, _, _ = load_case("dup_fallback")
    assert int(z["fallbacks"]) > 0
    z, _, _ = load_case("order3_converge")
    assert int(z["total_iter"]) < int(z["max_counter"])
