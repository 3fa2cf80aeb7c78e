"""Pins of the pivot-based projection oracle (oracle/pivoted.py) against LAPACK's
dgeqp3 (the routine behind Julia's qr(A, ColumnNorm())), what the mathematics fixes,
and the reference's own bars (test/rand_cp_als.jl).  CPU only."""
import math

import numpy as np
import pytest
from scipy.linalg import lapack

import oracle as O
import oracle.pivoted as P


def dgeqp3(A):
    """LAPACK dgeqp3 with every column free: (0-based pivots, diag R)."""
    m, n = A.shape
    qr, jpvt, _, _, info = lapack.dgeqp3(np.array(A, order="F"), lwork=max(1, 3 * n + 1) * 64)
    assert info == 0
    return jpvt - 1, np.diag(qr)[:min(m, n)].copy()


SHAPES = [(20, 1200), (30, 800), (40, 600), (13, 36), (36, 13), (64, 64), (5, 1), (1, 7), (128, 300)]


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{m}x{n}" for m, n in SHAPES])
def test_qrcp_is_dgeqp3(shape):
    rng = np.random.default_rng(sum(shape))
    A = rng.standard_normal(shape)
    p, r = P.qrcp(A)
    p_ref, r_ref = dgeqp3(A)
    assert (p == p_ref).all()
    assert np.abs(np.abs(r) - np.abs(r_ref)).max() <= 1e-12 * np.abs(r_ref).max()


def test_qrcp_rank_deficient_and_zero_columns():
    """A rank-6 matrix with repeated and zero columns: the pivots that matter (the first
    six) match dgeqp3, and |r_jj| collapses after them on both sides."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((25, 6)) @ rng.standard_normal((6, 90))
    A[:, 10] = A[:, 3]
    A[:, 20] = 0.0
    p, r = P.qrcp(A)
    p_ref, r_ref = dgeqp3(A)
    assert (p[:6] == p_ref[:6]).all()
    assert np.abs(r[6:]).max() < 1e-12 * np.abs(r).max()
    assert np.abs(np.abs(r[:6]) - np.abs(r_ref[:6])).max() <= 1e-12 * np.abs(r_ref).max()
    assert P.effective_rank(r, 0.01) == 6


def test_qrcp_factorisation_holds():
    """R's diagonal and the pivots are those of an exact factorisation A P = Q R: the
    column norms of R's leading triangle reproduce the pivoted columns' norms."""
    rng = np.random.default_rng(4)
    A = rng.standard_normal((12, 40))
    p, r = P.qrcp(A)
    Rfull = np.linalg.qr(A[:, p], mode="r")
    assert np.abs(np.abs(np.diag(Rfull)) - np.abs(r)).max() < 1e-12 * np.abs(r).max()
    # each pivot has the largest residual norm among the columns left at its step
    Q, _ = np.linalg.qr(A[:, p[:5]])
    res = A - Q @ (Q.T @ A)
    left = p[5:]
    assert np.argmax(np.linalg.norm(res[:, left], axis=0)) == 0


def test_unfold_and_multi_coords():
    """reshape(array(T, (i, others...)), (I, :)) and column_to_multi_coords
    (pivot_mapping.jl:17-35) agree with numpy's Fortran-order index maps."""
    rng = np.random.default_rng(5)
    T = rng.standard_normal((3, 4, 5, 2))
    for n in range(4):
        others = [m for m in range(4) if m != n]
        odims = [T.shape[m] for m in others]
        U = P.unfold(T, n)
        cols = rng.integers(0, U.shape[1], 20)
        C = P.column_to_multi_coords(cols, odims)
        assert (C == np.stack(np.unravel_index(cols, odims, order="F"), axis=1)).all()
        for c, co in zip(cols, C):
            idx = [0] * 4
            for q, m in enumerate(others):
                idx[m] = co[q]
            for i in range(T.shape[n]):
                idx[n] = i
                assert U[i, c] == T[tuple(idx)]


def _embedding(d, m, zeta, rng, injective=False):
    z = min(zeta, d)
    rows = np.empty((m, z), dtype=np.int32)
    q, r = divmod(d, z)
    for c in range(m):
        if injective:
            for j in range(z):
                st = j * (q + 1) if j < r else r * (q + 1) + (j - r) * q
                rows[c, j] = st + rng.integers(0, q + 1 if j < r else q)
        else:
            rows[c] = rng.choice(d, z, replace=False)
    vals = rng.choice([-1.0, 1.0], size=(m, z)) / math.sqrt(z)
    return rows.ravel(), vals.ravel()


def test_sparse_sign_invariants_catch_mistakes():
    rng = np.random.default_rng(6)
    rows, vals = _embedding(50, 30, 4, rng)
    assert P.sparse_sign_structure_ok(50, 30, 4, vals, rows, False)
    bad = rows.copy()
    bad[1] = bad[0]                                   # a repeated row in column 0
    assert not P.sparse_sign_structure_ok(50, 30, 4, vals, bad, False)
    rows, vals = _embedding(50, 30, 4, rng, injective=True)
    assert P.sparse_sign_structure_ok(50, 30, 4, vals, rows, True)
    rows2, _ = _embedding(50, 30, 4, rng)
    assert not P.sparse_sign_structure_ok(50, 30, 4, vals, rows2, True)


def test_sketched_matricization_is_the_dense_product():
    rng = np.random.default_rng(7)
    T = rng.standard_normal((6, 5, 4))
    for k in range(3):
        n = T.size // T.shape[k]
        l, s = 9, 3
        rows, vals = _embedding(l, n, s, rng)
        Om = np.zeros((l, n))
        Om[rows.reshape(n, s), np.arange(n)[:, None]] = vals.reshape(n, s)
        assert np.abs(P.sketched_matricization(T, k, l, rows, vals, s) - P.unfold(T, k) @ Om.T).max() < 1e-13


def test_seqrcs_with_an_identity_embedding_is_plain_qrcp():
    """One sketch row per column (Omega = I, s = 1, t = n): every column is touched, in
    the sketch's pivot order, and the subset QR-CP of all columns is the plain QR-CP."""
    rng = np.random.default_rng(8)
    T = rng.standard_normal((7, 6, 5))
    n = 7 * 5
    p, r = P.seqrcs(T, 1, n, 1, n, np.arange(n), np.ones(n))
    p0, r0 = P.qrcp(P.unfold(T, 1))
    assert (p == p0).all() and np.abs(r - r0).max() == 0.0


def test_seqrcs_map_back_order():
    """The subset: columns touching the first t sketch pivots, pivot by pivot, in
    increasing column, first occurrence kept; untouched columns follow in order."""
    rng = np.random.default_rng(9)
    T = rng.standard_normal((8, 6, 5))
    k, n, l, s, t = 0, 30, 12, 2, 4
    rows, vals = _embedding(l, n, s, rng)
    p, _ = P.seqrcs(T, k, l, s, t, rows, vals)
    p_sk, _ = P.qrcp(P.sketched_matricization(T, k, l, rows, vals, s))
    want = []
    for q in p_sk[:t]:
        for c in range(n):
            if q in rows[c * s:(c + 1) * s] and c not in want:
                want.append(c)
    assert sorted(p[:len(want)].tolist()) == sorted(want)
    assert p[len(want):].tolist() == sorted(set(range(n)) - set(want))
    assert sorted(p.tolist()) == list(range(n))


def test_projection_samples_ranges_and_shuffle():
    """start..end rows of the (shuffled-tail) pivot list; the first meff pivots never
    move; end 0 means every column (qr_lev_score_sampled.jl:33-56)."""
    rng = np.random.default_rng(10)
    T = rng.standard_normal((6, 5, 4))
    smp, piv, meff = P.projection_samples(T, 1, 0, trunc_tol=0.01, shuffle_pivots=False)
    for n in range(3):
        odims = [T.shape[m] for m in range(3) if m != n]
        p, _ = P.qrcp(P.unfold(T, n))
        assert (piv[n] == P.column_to_multi_coords(p, odims)).all()
        assert smp[n].shape[0] == int(np.prod(odims))
    smp2, piv2, meff2 = P.projection_samples(T, 2, 7, trunc_tol=0.5, shuffle_pivots=True, rng=1)
    for n in range(3):
        assert (piv2[n][:meff2[n]] == piv[n][:meff2[n]]).all()
        assert (smp2[n] == piv2[n][1:7]).all()
    with pytest.raises(ValueError):
        P.projection_samples(T, 9, 7)


def _low_rank_target(dims, R, seed):
    cp = O.random_CPD(np.zeros(dims), R, np.random.default_rng(seed))
    return O.reconstruct(cp.factors, cp.lam)


def test_reference_bar_qrpiv_full_rank():
    """The bar of rand_cp_als.jl:160-173 (QRPivProjected() keeps every column, so the
    run must give the plain ALS answer within 1e-2), on a smaller exact-rank target."""
    T = _low_rank_target((6, 7, 8), 12, 11)
    cp0 = O.random_CPD(T, 12, np.random.default_rng(12))
    chk = O.FitCheck(1e-6, 30, 1.0)
    ref = O.als_optimize(T, cp0, check=chk)
    smp, _, _ = P.projection_samples(T, 1, 0, rng=1)
    got = P.optimize_projected(T, cp0, smp, check=O.FitCheck(1e-6, 30, 1.0))
    X = O.reconstruct(ref.factors, ref.lam)
    assert np.linalg.norm(O.reconstruct(got.factors, got.lam) - X) / np.linalg.norm(X) < 1e-2


@pytest.mark.parametrize("normal", [True, False])
def test_reference_bar_qrpiv_interpolates_low_rank(normal):
    """rand_cp_als.jl:73-83: a rank-20 target on 20x30x40, rank-10 runs with
    CPDiffCheck(1e-5, 100); QRPivProjected(1200) reaches the plain ALS error within 10%."""
    T = _low_rank_target((20, 30, 40), 20, 21)
    cp0 = O.random_CPD(T, 10, np.random.default_rng(22))
    ref = O.als_optimize(T, cp0, check=O.CPDiffCheck(1e-5, 100))
    exact = np.linalg.norm(T - O.reconstruct(ref.factors, ref.lam)) / np.linalg.norm(T)
    smp, _, _ = P.projection_samples(T, 1, 1200, rng=2)
    got = P.optimize_projected(T, cp0, smp, check=O.CPDiffCheck(1e-5, 100), normal=normal)
    err = np.linalg.norm(T - O.reconstruct(got.factors, got.lam)) / np.linalg.norm(T)
    assert abs(exact - err) / exact < 0.1


def test_krp_columns_and_omega_hadamard():
    """KRP^T columns are the Khatri-Rao rows at each multi-index (first other mode
    fastest), and omega_hadamard (had_contract.jl:300-329) is the dense product."""
    rng = np.random.default_rng(12)
    fac = [rng.standard_normal((d, 3)) for d in (4, 5, 6)]
    for n in range(3):
        others = [f for m, f in enumerate(fac) if m != n]
        K = P.krp_columns(fac, n)
        a, b = others
        for c in range(K.shape[1]):
            i, j = c % a.shape[0], c // a.shape[0]
            assert np.abs(K[:, c] - a[i] * b[j]).max() < 1e-15
        ncol = K.shape[1]
        l, s = 7, 2
        rows, vals = _embedding(l, ncol, s, rng)
        Om = np.zeros((l, ncol))
        Om[rows.reshape(ncol, s), np.arange(ncol)[:, None]] = vals.reshape(ncol, s)
        assert np.abs(P.omega_hadamard(fac, n, l, rows, vals, s) - Om @ K.T).max() < 1e-13


def test_seqrcs_krp_map_back():
    """SEQRCS on the Khatri-Rao product (SEQRCS.jl:178-240): the subset is every column
    of Omega touching the first t sketch pivots, in increasing order; the rest follow."""
    rng = np.random.default_rng(13)
    fac = [rng.standard_normal((d, 4)) for d in (6, 5, 7)]
    n, l, s, t = 0, 11, 2, 3
    ncol = 35
    rows, vals = _embedding(l, ncol, s, rng)
    p, r = P.seqrcs_krp(fac, n, l, s, t, rows, vals)
    p_sk, _ = P.qrcp(P.omega_hadamard(fac, n, l, rows, vals, s).T)
    want = sorted({c for c in range(ncol) for q in p_sk[:t] if q in rows[c * s:(c + 1) * s]})
    sub, r0 = P.qrcp(P.krp_columns(fac, n, want))
    assert p[:len(want)].tolist() == [want[k] for k in sub]
    assert p[len(want):].tolist() == sorted(set(range(ncol)) - set(want))
    assert np.abs(r - r0).max() == 0.0
