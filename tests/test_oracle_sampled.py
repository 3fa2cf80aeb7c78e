"""Pins of the leverage-score-sampled ALS oracle (oracle/sampled.py) against what the
mathematics fixes and the reference's own test (test/rand_cp_als.jl:104-131).
CPU only."""
import itertools

import numpy as np
import pytest

import oracle as O
import oracle.sampled as S


def test_leverage_scores_are_the_hat_diagonal():
    """probability.jl:3-10: for a full-column-rank A the row norms of Q are the
    diagonal of A (A^T A)^-1 A^T; divided by R they sum to 1."""
    rng = np.random.default_rng(0)
    for m, n in ((30, 6), (7, 7), (50, 1)):
        A = rng.standard_normal((m, n))
        p = S.leverage_score_probability(A)
        hat = np.einsum("ij,jk,ik->i", A, np.linalg.inv(A.T @ A), A) / n
        assert np.abs(p - hat).max() < 1e-13
        assert abs(p.sum() - 1.0) < 1e-13
    # fewer rows than columns: Q is square, every row has norm 1, p = 1 / m
    p = S.leverage_score_probability(rng.standard_normal((3, 5)))
    assert np.abs(p - 1.0 / 3).max() < 1e-13


def test_pivot_hadamard_rows_of_the_full_krp():
    """had_contract.jl:277-295: row s is the Khatri-Rao row at the sampled tuple."""
    rng = np.random.default_rng(1)
    B, C, D = rng.standard_normal((4, 3)), rng.standard_normal((5, 3)), rng.standard_normal((2, 3))
    smp = np.array([[3, 0, 1], [0, 4, 0], [2, 2, 1], [3, 0, 1]], dtype=np.int32)
    K = S.pivot_hadamard([B, C, D], smp)
    for s, (j, k, l) in enumerate(smp):
        for r in range(3):
            assert K[s, r] == pytest.approx(B[j, r] * C[k, r] * D[l, r], rel=1e-15)


def test_fused_flatten_sample_is_fiber_indexing():
    """pivot_mapping.jl:90-118 against explicit loops, every mode of an order-4 tensor."""
    rng = np.random.default_rng(2)
    T = rng.standard_normal((3, 4, 5, 2))
    for k in range(4):
        others = [m for m in range(4) if m != k]
        smp = np.stack([rng.integers(0, T.shape[m], 6) for m in others], axis=1).astype(np.int32)
        Ts = S.fused_flatten_sample(T, k, smp)
        assert Ts.shape == (T.shape[k], 6)
        for s in range(6):
            for i in range(T.shape[k]):
                idx = [0] * 4
                for q, m in enumerate(others):
                    idx[m] = smp[s, q]
                idx[k] = i
                assert Ts[i, s] == T[tuple(idx)]


@pytest.mark.parametrize("normal", [True, False])
def test_exhaustive_samples_give_the_plain_als_step(normal):
    """Every index tuple of the other modes sampled once: K_s^T K_s is the Hadamard of
    the Grams and T_s K_s the MTTKRP, so the projected step IS the normal-equation step
    (the normal form by construction, the QR form as the same least-squares problem)."""
    rng = np.random.default_rng(3)
    dims, R = (5, 4, 3), 3
    T = rng.standard_normal(dims)
    cp = O.random_CPD(T, R, np.random.default_rng(4))
    for n in range(3):
        others = [m for m in range(3) if m != n]
        grid = np.array(list(itertools.product(*[range(dims[m]) for m in others])), dtype=np.int32)
        An, lam = S.sampled_solve(T, cp.factors, n, grid, normal)
        X, _ = O.solve_ls_problem(O.compute_krp(O.grams(cp.factors), n), O.mttkrp(T, cp.factors, n))
        A2, lam2 = O.row_norm(X)
        assert np.abs(An - A2).max() < 1e-12
        assert np.abs(lam - lam2).max() < 1e-12 * np.abs(lam2).max()


def test_qr_form_is_the_minimum_norm_least_squares():
    """qr(K, ColumnNorm()) \\ B (dgelsy) against numpy's SVD lstsq: tall, and wide
    (fewer samples than the rank: the minimum-norm solution)."""
    rng = np.random.default_rng(5)
    for m, n in ((40, 6), (2, 5), (5, 5)):
        K, B = rng.standard_normal((m, n)), rng.standard_normal((m, 4))
        x = O.qr_colnorm_solve(K, B)
        assert np.abs(x - np.linalg.lstsq(K, B, rcond=None)[0]).max() < 1e-12


def test_sampler_draws_from_the_weights():
    """probability.jl:12-33: one column per other mode, indices in range, and the
    empirical frequencies follow the weights."""
    rng = np.random.default_rng(6)
    probs = [np.array([0.5, 0.25, 0.25]), np.array([0.0, 1.0]), np.array([0.1, 0.2, 0.3, 0.4])]
    smp = S.sample_factor_matrices(20000, 0, probs, rng)
    assert smp.shape == (20000, 2)
    assert (smp[:, 0] == 1).all()
    freq = np.bincount(smp[:, 1], minlength=4) / 20000
    assert np.abs(freq - probs[2]).max() < 0.02


@pytest.mark.parametrize("normal", [False, True])
def test_reference_bar_rand_cp_als(normal):
    """rand_cp_als.jl:104-131: LevScoreSampled(100), CPAngleCheck(1e-5, 10) on a 12x13x3
    exact rank-5 target, the best of 3 tries within relative error 0.1 (the bar depends
    on the draws; these seeds meet it)."""
    cpT = O.random_CPD(np.zeros((12, 13, 3)), 5, np.random.default_rng(101))
    T = O.reconstruct(cpT.factors, cpT.lam)
    cp0 = O.random_CPD(T, 5, np.random.default_rng(201))
    errs = []
    for tr in range(3):
        out = S.optimize_sampled(T, cp0, 100, check=O.CPAngleCheck(1e-5, 10), normal=normal, rng=301 + tr)
        errs.append(np.linalg.norm(O.reconstruct(out.factors, out.lam) - T) / np.linalg.norm(T))
    assert min(errs) < 0.1


def test_stop_resample_zero_keeps_the_first_draws():
    """stop_resample = 0 (krp_lev_score_sampled.jl:24-26): no draw after compute_als, so
    a run equals a loop over the initial samples."""
    rng = np.random.default_rng(7)
    T = rng.standard_normal((6, 5, 4))
    cp0 = O.random_CPD(T, 3, np.random.default_rng(8))
    out = S.optimize_sampled(T, cp0, 30, check=O.NoCheck(3), normal=True, stop_resample=0, rng=9)
    r = np.random.default_rng(9)
    w = [S.leverage_score_probability(A) for A in cp0.factors]
    smp = [S.sample_factor_matrices(30, m, w, r) for m in range(3)]
    f = [A.copy() for A in cp0.factors]
    for _ in range(3):
        for n in range(3):
            f[n], lam = S.sampled_solve(T, f, n, smp[n], True)
    assert max(np.abs(a - b).max() for a, b in zip(out.factors, f)) == 0.0


def test_block_sampler_structure():
    """probability.jl:66-107: full rounds fill runs of consecutive indices of one block
    of the first non-skipped factor with the other modes fixed; leftovers are single
    draws; the blocks cover the factor (the first size % block_size one longer)."""
    rng = np.random.default_rng(11)
    probs = [np.full(7, 1 / 7), np.arange(1, 11) / 55.0, np.full(4, 0.25)]
    smp = S.block_sample_factor_matrices(17, probs, 3, 0, rng)   # blocked: mode 1 (10 = 4+3+3)
    assert smp.shape == (17, 2)
    starts = [0, 4, 7, 10]
    row = 0
    for _ in range(17 // 3):
        b = [k for k in range(3) if starts[k] == smp[row, 0]]
        assert len(b) == 1
        ln = starts[b[0] + 1] - starts[b[0]]
        run = smp[row:row + ln]
        assert (run[:, 0] == np.arange(starts[b[0]], starts[b[0]] + len(run))).all()
        assert (run[:, 1] == run[0, 1]).all()
        row += len(run)
        if row >= 17:
            break
    assert ((smp[:, 0] >= 0) & (smp[:, 0] < 10)).all() and ((smp[:, 1] >= 0) & (smp[:, 1] < 4)).all()
    with pytest.raises(ValueError):
        S.block_sample_factor_matrices(5, probs, 11, 0, rng)


@pytest.mark.parametrize("blocks", [1, 12])
def test_block_reference_bar(blocks):
    """rand_cp_als.jl:133-149: BlockLevScoreSampled((50, 50, 500), blocks), best of 3 tries
    within relative error 0.1 (these seeds meet it)."""
    cpT = O.random_CPD(np.zeros((12, 13, 3)), 5, np.random.default_rng(101))
    T = O.reconstruct(cpT.factors, cpT.lam)
    cp0 = O.random_CPD(T, 5, np.random.default_rng(201))
    errs = []
    for tr in range(3):
        out = S.optimize_sampled(T, cp0, (50, 50, 500), check=O.CPAngleCheck(1e-5, 10), rng=301 + tr,
                                 blocks=blocks)
        errs.append(np.linalg.norm(O.reconstruct(out.factors, out.lam) - T) / np.linalg.norm(T))
    assert min(errs) < 0.1


def test_product_samplers_draw_what_the_oracle_draws():
    """The engine's host-side samplers (the reference draws on the host too) consume the
    generator exactly as the oracle's restatements do, so GPU runs and oracle runs with
    one seed see the same samples."""
    import __graft_entry__ as G
    E = G.load_package()
    probs = [np.random.default_rng(1).random(n) for n in (9, 13, 5)]
    for skip in range(3):
        a = E.LevScoreSampled(40).draw(skip, probs, np.random.default_rng(7))
        b = S.sample_factor_matrices(40, skip, probs, np.random.default_rng(7))
        assert (a == b).all()
        a = E.BlockLevScoreSampled(40, 4).draw(skip, probs, np.random.default_rng(8))
        b = S.block_sample_factor_matrices(40, probs, 4, skip, np.random.default_rng(8))
        assert (a == b).all()
