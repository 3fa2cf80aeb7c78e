"""numpy restatement of ITensorCPD.jl's leverage-score-sampled ALS (test infrastructure).

Only tests use this module, as the checker; the product path (``itensorcpd.jl_amd``)
never imports it.  Modes and sample indices are 0-based here (1-based in Julia).

What pins each function (checks in ``tests/test_oracle_sampled.py``):
  leverage_score_probability  the hat-matrix diagonal diag(A (A^T A)^-1 A^T) / R of a
                              full-column-rank A; probabilities sum to 1
  pivot_hadamard              rows of the full Khatri-Rao product at the sampled tuples
  fused_flatten_sample        plain fancy indexing of T
  sampled_solve               with every index tuple sampled once, the normal form IS the
                              plain ALS step (V = Hadamard of Grams, M = MTTKRP); the QR
                              form agrees with numpy's minimum-norm lstsq, also on
                              wide (nsamps < R) samples
  optimize_sampled            the reference's own bar (rand_cp_als.jl:106-124): a
                              12x13x3 rank-5 target recovered to relative error < 0.1
"""
from __future__ import annotations

import numpy as np

from .als import CPD, FitCheck, NoCheck, check_converge, grams, ldiv_solve, qr_colnorm_solve, row_norm


def leverage_score_probability(A: np.ndarray) -> np.ndarray:
    """probability.jl:3-10: q = Q of qr(A) (thin), sum_j q_ij^2 / min(size(A))."""
    q, _ = np.linalg.qr(A, mode="reduced")
    return (q * q).sum(axis=1) / min(A.shape)


def samples_from_probability_vector(pw: np.ndarray, nsamps: int, rng) -> np.ndarray:
    """probability.jl:12-14: ``sample(1:n, Weights(abs.(PW)))`` nsamps times (with
    replacement); here numpy's weighted choice, 0-based."""
    w = np.abs(pw)
    return rng.choice(len(w), size=nsamps, replace=True, p=w / w.sum()).astype(np.int32)


def sample_factor_matrices(nsamps: int, skip: int, probs, rng) -> np.ndarray:
    """probability.jl:22-33: nsamps independent draws from each factor's weights except
    ``skip``'s, in mode order -> (nsamps, N-1) int array."""
    cols = [samples_from_probability_vector(p, nsamps, rng) for m, p in enumerate(probs) if m != skip]
    return np.stack(cols, axis=1)


def sample_single_col_from_factors(skip: int, probs, rng) -> list:
    """probability.jl:16-18: one draw from each factor's weights except ``skip``'s."""
    return [int(samples_from_probability_vector(p, 1, rng)[0]) for m, p in enumerate(probs) if m != skip]


def block_sample_factor_matrices(nsamps: int, probs, block_size: int, skip: int, rng) -> np.ndarray:
    """probability.jl:66-107 (BlockLevScoreSampled): the first factor other than
    ``skip`` is cut into size_fast // block_size blocks (the first size_fast % block_size
    one index longer) weighted by their summed probabilities; each of nsamps //
    block_size rounds draws one index of every other factor, then a block, and fills
    consecutive rows with the block's indices (the other modes fixed to the round's
    draws); rows left over take single draws.  0-based, -> (nsamps, N-1)."""
    blocked = 1 if skip == 0 else 0
    bf = np.asarray(probs[blocked])
    size_fast = len(bf)
    nblocks, resid = size_fast // block_size, size_fast % block_size
    if nblocks == 0:
        raise ValueError("block size larger than the blocked factor")
    starts = [0]
    bprob = np.empty(nblocks)
    m = 0
    for i in range(nblocks):
        ln = block_size + (1 if i < resid else 0)
        bprob[i] = bf[m:m + ln].sum()
        m += ln
        starts.append(m)
    out = np.empty((nsamps, len(probs) - 1), dtype=np.int32)
    row = 0
    for _ in range(nsamps // block_size):
        other = sample_single_col_from_factors(skip, probs, rng)
        b = int(rng.choice(nblocks, p=bprob / bprob.sum()))
        for idx in range(starts[b], starts[b + 1]):
            if row >= nsamps:
                break
            out[row, 0] = idx
            out[row, 1:] = other[1:]
            row += 1
    for r in range(row, nsamps):
        out[r, :] = sample_single_col_from_factors(skip, probs, rng)
    return out


def pivot_hadamard(others, samples: np.ndarray) -> np.ndarray:
    """had_contract.jl:277-295: prod over the other factors of their sampled rows."""
    out = np.ones((samples.shape[0], others[0].shape[1]))
    for A, col in zip(others, samples.T):
        out *= A[col, :]
    return out


def fused_flatten_sample(T: np.ndarray, k: int, samples: np.ndarray) -> np.ndarray:
    """pivot_mapping.jl:90-118: column s = the mode-k fiber of T at the sampled indices
    of the other modes -> (I_k, nsamps)."""
    others = [m for m in range(T.ndim) if m != k]
    ns = samples.shape[0]
    idx = [None] * T.ndim
    for q, m in enumerate(others):
        idx[m] = samples[None, :, q]
    idx[k] = np.arange(T.shape[k])[:, None]
    return np.ascontiguousarray(T[tuple(np.broadcast_arrays(*idx))].reshape(T.shape[k], ns))


def sampled_solve(T: np.ndarray, factors, n: int, samples: np.ndarray, normal: bool):
    """One step of optimize.jl:17-25 for a ProjectionAlgorithm: compute_krp ->
    project_krp (pivot_hadamard), matricize_tensor (fused_flatten_sample),
    solve_ls_problem (ProjectionAlgorithm.jl:54-67), row_norm.  Returns (A_n, lambda)."""
    others = [A for m, A in enumerate(factors) if m != n]
    K = pivot_hadamard(others, samples)
    Ts = fused_flatten_sample(T, n, samples)
    if normal:
        X, _ = ldiv_solve(K.T @ K, np.ascontiguousarray((Ts @ K).T))
    else:
        X = qr_colnorm_solve(K, np.ascontiguousarray(Ts.T))
    return row_norm(np.ascontiguousarray(X.T))


def _projection_check(check, factors, lam) -> bool:
    """ProjectionAlgorithm.jl:20-53: FitCheck only counts sweeps (never converges);
    every other check runs its own check_converge."""
    if isinstance(check, FitCheck):
        check.iter += 1
        if check.iter >= check.max_counter:
            check.iter = 0
        return False
    return check_converge(check, factors, lam, grams(factors))


def optimize_sampled(T: np.ndarray, cp: CPD, nsamples, check=None, normal: bool = False,
                     stop_resample: int = -1, rng=None, blocks=None) -> CPD:
    """compute_als(::LevScoreSampled) (optimizers/.../krp_lev_score_sampled.jl:1-37) and
    optimize (optimize.jl:6-32) with project_krp / post_solve of
    krp_lev_score_sampled.jl:18-57.  nsamples: int or one per mode.  With ``blocks``
    (int or one per mode) the BlockLevScoreSampled draws (krp_lev_score_sampled.jl:
    66-103, optimizers/.../krp_lev_score_sampled.jl:39-83)."""
    rng = np.random.default_rng(rng)
    N = T.ndim
    ns = [nsamples] * N if np.isscalar(nsamples) else list(nsamples)
    if len(ns) == 1:
        ns = ns * N
    bs = None
    if blocks is not None:
        bs = [blocks] * N if np.isscalar(blocks) else list(blocks)
        if len(bs) == 1:
            bs = bs * N

    def draw(n, weights):
        if bs is None:
            return sample_factor_matrices(ns[n], n, weights, rng)
        return block_sample_factor_matrices(ns[n], weights, bs[n], n, rng)
    check = check if check is not None else NoCheck(100)
    factors = [np.array(A, dtype=np.float64) for A in cp.factors]
    lam = np.array(cp.lam, dtype=np.float64)
    weights = [leverage_score_probability(A) for A in factors]
    samples = [draw(m, weights) for m in range(N)]
    it = check.iter
    while it < check.max_counter:
        for n in range(N):
            if stop_resample < 0 or stop_resample > check.iter:
                samples[n] = draw(n, weights)
            factors[n], lam = sampled_solve(T, factors, n, samples[n], normal)
            weights[n] = leverage_score_probability(factors[n])
        if _projection_check(check, factors, lam):
            break
        it += 1
    return CPD(factors, lam)
