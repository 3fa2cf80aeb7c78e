"""MI355X engine for ITensorCPD.jl's dense normal-equation CP-ALS path.

The public names mirror the reference's Julia API on this path so a caller (or
a test written like the reference's ``test/cp_als.jl``) reads the same:

    decompose(A, rank; rng, alg, check, maxiter, verbose)      decompose.jl:13-30
    decompose(A, epsilon, max_rank; ...)  -> decompose_adaptive decompose.jl:32-69
    random_CPD(A, rank; rng)                                   cpd.jl:63-82
    als_optimize(A, cp; alg, check, maxiter, verbose)          als_optimizer.jl:15-25
    FitCheck(tol, max, norm), NoCheck(max), CPDiffCheck(tol, max),
    CPAngleCheck(tol, max), CPDFit(check)                      converge_checks/*.jl
    KRPFreeNormal(), KRPNormal()                               algorithms/.../standard/tensor.jl
    LevScoreSampled(n); als_optimize(...; normal, stop_resample) randomized/krp_lev_score_sampled.jl
    QRPivProjected(n), SEQRCSPivProjected(s, e, modes, ranks), KSEQRCSPivProjected(...);
    als_optimize(...; shuffle_pivots, trunc_tol, injective),
    update_samples                                          randomized/qr_lev_score_sampled.jl
    CPD (factors, lam), cp_rank                                cpd.jl:7-46

All arithmetic runs in libitcpd_hip.so on the GPU; there is no CPU fallback.
"""
from .api import (  # noqa: F401
    CPD, ALSPlan, BatchALS, DeviceTensor, MLP, reconstruct, CPDFit, FitCheck, CPDiffCheck, CPAngleCheck, KRPFreeNormal, KRPNormal, NoCheck, CPDOptimizer,
    LevScoreSampled, BlockLevScoreSampled, ProjectionAlgorithm, QRPivProjected, SEQRCSPivProjected,
    KSEQRCSPivProjected, PivotSetup, compute_projection, optimize_projected, update_samples, sparse_sign,
    als_optimize, cp_rank, decompose, decompose_adaptive, decompose_batch, increase_cpd_rank,
    random_CPD, tensor_norm,
)
from ._lib import ITCPDError  # noqa: F401

__all__ = [
    "CPD", "ALSPlan", "BatchALS", "decompose_batch", "DeviceTensor", "MLP", "reconstruct", "CPDFit", "FitCheck", "CPDiffCheck", "CPAngleCheck", "KRPFreeNormal", "KRPNormal", "NoCheck",
    "LevScoreSampled", "BlockLevScoreSampled", "ProjectionAlgorithm", "QRPivProjected",
    "SEQRCSPivProjected", "KSEQRCSPivProjected", "PivotSetup", "compute_projection", "optimize_projected", "update_samples",
    "sparse_sign", "CPDOptimizer", "als_optimize", "cp_rank", "decompose", "decompose_adaptive",
    "increase_cpd_rank", "random_CPD", "tensor_norm", "ITCPDError",
]
