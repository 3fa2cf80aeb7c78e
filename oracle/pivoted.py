"""numpy restatement of ITensorCPD.jl's pivot-based projected ALS (test infrastructure).

Only tests use this module, as the checker; the product path (``itensorcpd.jl_amd``)
never imports it.  Modes, columns and sample indices are 0-based here (1-based in Julia).

The path (src/optimizers/als_optimizers/randomized/qr_lev_score_sampled.jl,
src/algorithms/als_algorithms/randomized/qr_lev_score_sampled.jl, src/algebra/SEQRCS.jl):
for each mode n the columns of the unfolding T_(n) are ordered by a column-pivoted QR
(Julia ``qr(A, ColumnNorm())`` = LAPACK ``dgeqp3``), the pivots beyond the effective
rank are optionally shuffled, a contiguous range of them becomes the fixed sample set of
mode n, and every sweep solves the projected least-squares problems of
``oracle.sampled.sampled_solve`` on those samples.  SE-QRCS replaces the QR of T_(n) by
a QR of the sketch T_(n) Omega^T (Omega a sparse sign embedding), maps the sketch pivots
back to the columns of T_(n) they touch and orders those by a QR of that column subset.

What pins each function (checks in ``tests/test_oracle_pivoted.py``):
  qrcp                     LAPACK dgeqp3 through scipy (the routine Julia calls): the same
                           pivots, |diag R| to rounding, on tall, wide, square and
                           rank-deficient matrices with min(m, n) <= 128 (dgeqp3's
                           unblocked range)
  column_to_multi_coords   numpy's unravel_index in Fortran order, and its inverse
  sketched_matricization   the dense product T_(n) Omega^T
  seqrcs                   an embedding that keeps every column (one sketch row per
                           column) reduces it to the plain column-pivoted QR of T_(n);
                           the map-back order against a brute-force scan
  sparse_sign_structure    the structural invariants of sparse_sign.c / sparsestack.c
                           (values +-1/sqrt(zeta), distinct rows per column, injective
                           blocks); the draws themselves come from libc rand() in the
                           reference and are not reproducible (parity unpinned for the
                           draws: the engine's generator is checked by the same invariants)
  krp_columns /            pivot_hadamard rows at every multi-index (the explicit
  omega_hadamard           Khatri-Rao product) and the dense product Omega KRP
  seqrcs_krp               the map-back order against a brute-force scan
  optimize_projected       the reference's own bars (rand_cp_als.jl:40-48, 78-83, 166-178)
"""
from __future__ import annotations

import math

import numpy as np

from .als import CPD, NoCheck
from .sampled import _projection_check, sampled_solve

# sqrt(dlamch('E')) with dlamch('E') = 2^-53: the threshold below which dlaqp2
# recomputes a downdated column norm.
TOL3Z = math.sqrt(2.0 ** -53)


def qrcp(A: np.ndarray):
    """Column-pivoted Householder QR, LAPACK dgeqp3 with free columns -> xLAQP2
    (unblocked, the path dgeqp3 takes for min(m, n) <= 128).  Returns (jpvt, rdiag):
    the column order (0-based) and the signed diagonal of R.

    Per step j: pivot = first position of the largest partial norm vn1 among j..n-1
    (IDAMAX), swap it to position j (columns, jpvt, vn1/vn2), Householder reflector of
    rows j..m-1 of that column (DLARFG: beta = -sign(alpha) hypot(alpha, ||x||)),
    applied to the remaining columns (DLARF), then the norm downdate
    vn1 *= sqrt(max(0, 1 - (|a_jl| / vn1)^2)) with recomputation when
    (1 - (|a_jl| / vn1)^2) (vn1 / vn2)^2 <= sqrt(eps)."""
    W = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = W.shape
    mn = min(m, n)
    jpvt = np.arange(n)
    vn1 = np.sqrt((W * W).sum(axis=0))
    vn2 = vn1.copy()
    rdiag = np.zeros(mn)
    for j in range(mn):
        pvt = j + int(np.argmax(vn1[j:]))          # first index of the maximum
        if pvt != j:
            W[:, [j, pvt]] = W[:, [pvt, j]]
            jpvt[[j, pvt]] = jpvt[[pvt, j]]
            vn1[pvt] = vn1[j]
            vn2[pvt] = vn2[j]
        alpha = W[j, j]
        x = W[j + 1:, j]
        xnorm = math.sqrt(float(x @ x)) if x.size else 0.0
        if xnorm == 0.0:
            tau, beta = 0.0, alpha
        else:
            beta = -math.copysign(math.hypot(alpha, xnorm), alpha)
            tau = (beta - alpha) / beta
            W[j + 1:, j] = x / (alpha - beta)
        rdiag[j] = beta
        if j + 1 < n and tau != 0.0:
            v = np.concatenate(([1.0], W[j + 1:, j]))
            C = W[j:, j + 1:]
            w = v @ C
            C -= tau * np.outer(v, w)
        W[j, j] = beta
        # partial norms of the remaining columns
        rest = np.arange(j + 1, n)
        live = rest[vn1[rest] != 0.0]
        if live.size:
            temp = 1.0 - (np.abs(W[j, live]) / vn1[live]) ** 2
            temp = np.maximum(temp, 0.0)
            temp2 = temp * (vn1[live] / vn2[live]) ** 2
            redo = temp2 <= TOL3Z
            keep = live[~redo]
            vn1[keep] = vn1[keep] * np.sqrt(temp[~redo])
            for q in live[redo]:
                if j + 1 < m:
                    y = W[j + 1:, q]
                    vn1[q] = math.sqrt(float(y @ y))
                    vn2[q] = vn1[q]
                else:
                    vn1[q] = 0.0
                    vn2[q] = 0.0
    return jpvt, rdiag


def unfold(T: np.ndarray, n: int) -> np.ndarray:
    """reshape(array(T, (i_n, others...)), (I_n, prod others)): columns enumerate the
    other modes with the first one fastest (Julia's column-major reshape)."""
    others = [m for m in range(T.ndim) if m != n]
    return np.moveaxis(T, n, 0).reshape(T.shape[n], -1, order="F") if others else T[:, None]


def column_to_multi_coords(cols, dims) -> np.ndarray:
    """pivot_mapping.jl:17-35: column index -> one index per other mode (first fastest)."""
    cols = np.asarray(cols, dtype=np.int64)
    out = np.empty((cols.size, len(dims)), dtype=np.int32)
    rem = cols.copy()
    for q, d in enumerate(dims):
        out[:, q] = rem % d
        rem //= d
    return out


def effective_rank(rdiag: np.ndarray, trunc_tol: float) -> int:
    """qr_lev_score_sampled.jl:29: sum(|r_jj| / max|r_jj| > trunc_tol)."""
    a = np.abs(rdiag)
    return int((a / a.max() > trunc_tol).sum())


def sparse_sign_structure_ok(d: int, m: int, zeta: int, vals, rows, injective: bool) -> bool:
    """Invariants of sparse_sign.c:25-69 / sparsestack.c:24-78 (0-based rows): zeta =
    min(zeta, d) entries per column of value +-1/sqrt(zeta); rows distinct within a
    column; in the injective form entry q lies in block q of the near-equal split of
    0..d-1 (the first d % zeta blocks one longer)."""
    z = min(zeta, d)
    vals = np.asarray(vals).reshape(m, z)
    rows = np.asarray(rows).reshape(m, z)
    if not np.all(np.abs(np.abs(vals) - 1.0 / math.sqrt(z)) <= 1e-15):
        return False
    if rows.min() < 0 or rows.max() >= d:
        return False
    if any(len(set(r)) != z for r in rows.tolist()):
        return False
    if injective:
        q, r = divmod(d, z)
        start = np.array([j * (q + 1) if j < r else r * (q + 1) + (j - r) * q for j in range(z)])
        size = np.array([q + 1 if j < r else q for j in range(z)])
        if not (np.all(rows >= start[None, :]) and np.all(rows < (start + size)[None, :])):
            return False
    return True


def sketched_matricization(T: np.ndarray, k: int, l: int, rows, vals, s: int) -> np.ndarray:
    """pivot_mapping.jl:103-139 (the form without the sparse matrix): A_sk(:, j) = sum
    over the columns c of Omega with a nonzero in row j, in increasing c, of
    Omega(j, c) T_(k)(:, c) -> I_k x l.  rows / vals: s entries per column of Omega."""
    Tk = unfold(T, k)
    n = Tk.shape[1]
    rows = np.asarray(rows).reshape(n, s)
    vals = np.asarray(vals).reshape(n, s)
    out = np.zeros((Tk.shape[0], l))
    for j in range(l):
        cs, qs = np.nonzero(rows == j)              # row-major: increasing c
        if cs.size:
            out[:, j] = Tk[:, cs] @ vals[cs, qs]
    return out


def seqrcs(T: np.ndarray, k: int, l: int, s: int, t: int, rows, vals):
    """SEQRCS(::Val{false}, ...) (SEQRCS.jl:127-175) for a given embedding (rows / vals,
    0-based rows, s per column): QR-CP of the sketch, its first t pivots, the columns of
    T_(k) that touch those sketch rows (pivot order, then increasing column, first
    occurrence kept), QR-CP of that column subset; pivots = subset in that order, then the
    untouched columns increasing.  Returns (p, rdiag of the subset QR)."""
    Tk = unfold(T, k)
    n = Tk.shape[1]
    A_sk = sketched_matricization(T, k, l, rows, vals, s)
    p_sk, _ = qrcp(A_sk)
    rmat = np.asarray(rows).reshape(n, s)
    seen = np.zeros(n, dtype=bool)
    indices = []
    for p in p_sk[:t]:
        for c in np.nonzero((rmat == p).any(axis=1))[0]:
            if not seen[c]:
                seen[c] = True
                indices.append(int(c))
    indices = np.array(indices, dtype=np.int64)
    p_sub, rdiag = qrcp(Tk[:, indices])
    rem = np.nonzero(~seen)[0]
    return np.concatenate([indices[p_sub], rem]), rdiag


def krp_columns(factors, n: int, cols=None) -> np.ndarray:
    """The transposed Khatri-Rao product of the factors other than n (R x prod others,
    column c = the KRP row at c's multi-index, first other mode fastest), restricted to
    `cols` (qr_lev_score_sampled.jl:259-261, SEQRCS.jl:217)."""
    from .sampled import pivot_hadamard
    others = [np.asarray(A) for m, A in enumerate(factors) if m != n]
    odims = [A.shape[0] for A in others]
    if cols is None:
        cols = np.arange(int(np.prod(odims)))
    return pivot_hadamard(others, column_to_multi_coords(cols, odims)).T.copy()


def omega_hadamard(factors, n: int, l: int, rows, vals, s: int) -> np.ndarray:
    """had_contract.jl:300-329: row j = sum over the nonzeros of row j of Omega (increasing
    column c) of Omega(j, c) KRP(c, :) -> l x R."""
    K = krp_columns(factors, n).T
    ncol = K.shape[0]
    rows = np.asarray(rows).reshape(ncol, s)
    vals = np.asarray(vals).reshape(ncol, s)
    out = np.zeros((l, K.shape[1]))
    for j in range(l):
        cs, qs = np.nonzero(rows == j)
        if cs.size:
            out[j] = vals[cs, qs] @ K[cs]
    return out


def seqrcs_krp(factors, n: int, l: int, s: int, t: int, rows, vals):
    """SEQRCS(krp::Vector{ITensor}, ...) (SEQRCS.jl:178-240) for a given embedding:
    QR-CP of (Omega KRP)^T, the columns of Omega with a nonzero in its first t pivot rows
    (increasing), QR-CP of KRP^T on them; pivots = that subset in QR order, then the
    rest increasing.  Returns (p, rdiag of the subset QR)."""
    A_sk = omega_hadamard(factors, n, l, rows, vals, s)
    p_sk, _ = qrcp(A_sk.T)
    ncol = int(np.prod([np.asarray(A).shape[0] for m, A in enumerate(factors) if m != n]))
    touched = np.isin(np.asarray(rows).reshape(ncol, s), p_sk[:t]).any(axis=1)
    indices = np.nonzero(touched)[0]
    p_sub, rdiag = qrcp(krp_columns(factors, n, indices))
    return np.concatenate([indices[p_sub], np.nonzero(~touched)[0]]), rdiag


def _range(v, n):
    return v if np.isscalar(v) else v[n]


def projection_samples(T: np.ndarray, start, end, trunc_tol=0.01, shuffle_pivots=True,
                       rng=None, random_modes=(), rank_vect=None, embeddings=None, krp_factors=None):
    """compute_als(::QRPivProjected / ::SEQRCSPivProjected / ::KSEQRCSPivProjected)
    (optimizers/.../qr_lev_score_sampled.jl:1-281): per mode the pivot order, the
    effective rank, the shuffled tail and the sample rows start..end (1-based,
    inclusive; end 0 = all).  ``embeddings[n] = (rows, vals)`` supplies the sparse sign
    embedding of a random mode (the reference draws it from libc rand()).  With
    ``krp_factors`` (KSEQRCS: the preliminary CPD's factors) the pivots come from their
    Khatri-Rao product instead of T.  Returns (samples, pivots, meffs)."""
    rng = np.random.default_rng(rng)
    samples, pivots, meffs = [], [], []
    for n in range(T.ndim):
        others = [m for m in range(T.ndim) if m != n]
        odims = [T.shape[m] for m in others]
        dR = int(np.prod(odims))
        e = _range(end, n)
        e = dR if e == 0 else min(e, dR)
        b = _range(start, n)
        if not (0 < b <= e):
            raise ValueError("sample range start must lie in 1..end")
        m = T.shape[n]
        if n in random_modes:
            k_sk = e if rank_vect is None else rank_vect[n]
            l = int(round(3 * m * math.log(m)))
            s = int(round(math.log(m)))
            rows, vals = embeddings[n]
            if krp_factors is not None:
                p, rdiag = seqrcs_krp(krp_factors, n, l, s, k_sk, rows, vals)
            else:
                p, rdiag = seqrcs(T, n, l, s, k_sk, rows, vals)
        elif krp_factors is not None:
            p, rdiag = qrcp(krp_columns(krp_factors, n))
        else:
            p, rdiag = qrcp(unfold(T, n))
        meff = effective_rank(rdiag, trunc_tol)
        tail = p[meff:]
        if shuffle_pivots:
            tail = tail[rng.permutation(tail.size)]
        p = np.concatenate([p[:meff], tail])
        coords = column_to_multi_coords(p, odims)
        pivots.append(coords)
        meffs.append(meff)
        samples.append(coords[b - 1:e])
    return samples, pivots, meffs


def optimize_projected(T: np.ndarray, cp: CPD, samples, check=None, normal=True) -> CPD:
    """optimize (optimize.jl:6-32) with the pivot-based project_krp / matricize_tensor
    (algorithms/.../qr_lev_score_sampled.jl:150-169): fixed samples, no post_solve."""
    check = check if check is not None else NoCheck(100)
    factors = [np.array(A, dtype=np.float64) for A in cp.factors]
    lam = np.array(cp.lam, dtype=np.float64)
    it = check.iter
    while it < check.max_counter:
        for n in range(T.ndim):
            factors[n], lam = sampled_solve(T, factors, n, samples[n], normal)
        if _projection_check(check, factors, lam):
            break
        it += 1
    return CPD(factors, lam)
