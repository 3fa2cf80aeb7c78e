// Batched CP-ALS: thousands of small independent order-3 problems, each run for a
// fixed number of sweeps (NoCheck semantics, no_check.jl:1-20) inside ONE
// persistent kernel.  A workgroup takes the next problem from a device queue,
// keeps its three factors, Grams and the R x R work in LDS, streams its tensor from
// global memory (L2 / infinity cache) three times per sweep and writes the result.
//
// Per sweep and mode n the steps are the reference's (optimize.jl:19-30):
//   V = Hadamard of the other Grams (MttkrpAlgorithm.jl:18-31),
//   M = MTTKRP (standard/tensor.jl:32-44),
//   X = V \ M^T by pivoted Cholesky (tol 1e-6) or the column-pivoted QR fallback
//       (ldiv_solve.jl:13-29), applied as an explicit R x R solution operator
//       built by solving the R unit right-hand sides,
//   (A_n, lambda) = row_norm(X) (row_norm.jl:4-24), G_n = A_n^T A_n.
// After the last sweep the FP64 fit (fit_check.jl:24-32) is stored per problem.
#include "dense_la.h"
#include <hip/hip_runtime.h>

namespace itcpd {

namespace {

constexpr int LDT = BATCH_MAXR + 2;

// LDS layout (in doubles) shared by every problem a workgroup runs.
struct BatchLds {
    double* A[3];      // factors, row-major [i][r], leading dimension RB (zero padded)
    double* G[3];      // Grams, R x R
    double* M;         // MTTKRP of the current mode, [i][r], ld RB
    double* V;         // V, then U (pivoted, row-major) or the QR data W
    double* Z;         // solution operator: Z[k * R + c] = (V^-1 e_k)_c
    double* UbT;       // left-looking Cholesky columns, [index][step], ld LDT
    double* qtau;
    double* ztau;
    double* vwork;     // 2 R
    double* diagU;
    double* lam;
    double* redv;      // NT
    int* piv;
    int* jpvt;
    int* finfo;        // rank, fallback, qr rank
    int* redi;         // NT
};

template <int NT>
__host__ __device__ constexpr size_t batch_lds_doubles() {
    return 3 * (size_t)BATCH_MAXD * BATCH_MAXR + 3 * (size_t)BATCH_MAXR * BATCH_MAXR +
           (size_t)BATCH_MAXD * BATCH_MAXR + 2 * (size_t)BATCH_MAXR * BATCH_MAXR +
           (size_t)BATCH_MAXR * LDT + 7 * (size_t)BATCH_MAXR + NT +
           (3 * BATCH_MAXR + 4 + NT + 1) / 2;
}

template <int NT>
__device__ BatchLds carve(double* base) {
    BatchLds L;
    double* p = base;
    for (int n = 0; n < 3; ++n) { L.A[n] = p; p += BATCH_MAXD * BATCH_MAXR; }
    for (int n = 0; n < 3; ++n) { L.G[n] = p; p += BATCH_MAXR * BATCH_MAXR; }
    L.M = p; p += BATCH_MAXD * BATCH_MAXR;
    L.V = p; p += BATCH_MAXR * BATCH_MAXR;
    L.Z = p; p += BATCH_MAXR * BATCH_MAXR;
    L.UbT = p; p += BATCH_MAXR * LDT;
    L.qtau = p; p += BATCH_MAXR;
    L.ztau = p; p += BATCH_MAXR;
    L.vwork = p; p += 2 * BATCH_MAXR;
    L.diagU = p; p += BATCH_MAXR;
    L.lam = p; p += 2 * BATCH_MAXR;
    L.redv = p; p += NT;
    int* q = reinterpret_cast<int*>(p);
    L.piv = q; q += BATCH_MAXR;
    L.jpvt = q; q += BATCH_MAXR;
    L.finfo = q; q += 4;
    L.redi = q;
    return L;
}

// M(idx, r) for mode n of one problem.  Threads are (outer group, rank chunk of
// 8, output index): lanes run over the output index, each thread keeps the 8
// ranks of its chunk, and the "outer" other index is split over up to 8 groups;
// the "inner" one is looped over with the factor rows read from LDS.  The outer
// factor multiplies each inner partial sum (KRP(i, j, k) = A(j) A(k) is never
// formed).  Groups add into M one after another in a fixed order.
// acc(8 ranks of chunk rc0) = sum over outer o = o0, o0 + ostep, ... of
// A_out(o, .) * sum_q T(element) A_in(q, .) for output index idx of mode n.
template <int RB>
__device__ __forceinline__ void mttkrp_task(const double* __restrict__ T, int I, int J, int n,
                                            int idx, int rc0, int o0, int ostep, int outerN,
                                            int innerN, const BatchLds& L, double (&acc)[8]) {
    constexpr int RC = 8;
    const double* Ain = (n == 0 ? L.A[1] : L.A[0]) + rc0;
    const double* Aout = (n == 2 ? L.A[1] : L.A[2]) + rc0;
    const int64_t IJ = (int64_t)I * J;
#pragma unroll
    for (int r = 0; r < RC; ++r) acc[r] = 0.0;
    for (int o = o0; o < outerN; o += ostep) {
        double s[RC];
#pragma unroll
        for (int r = 0; r < RC; ++r) s[r] = 0.0;
        // element (idx, inner q, outer o) of T for this mode
        int64_t base, step;
        if (n == 0) { base = idx + IJ * o; step = I; }                      // T(idx, q, o)
        else if (n == 1) { base = (int64_t)I * idx + IJ * o; step = 1; }    // T(q, idx, o)
        else { base = (int64_t)I * o + IJ * idx; step = 1; }                // T(q, o, idx)
        // eight tensor loads issued before their use (the loop is latency bound); the
        // products still add up in q order
        int q = 0;
        for (; q + 8 <= innerN; q += 8) {
            double tq[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) tq[u] = T[base + step * (q + u)];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const dv2* a = reinterpret_cast<const dv2*>(Ain + (q + u) * RB);
#pragma unroll
                for (int r2 = 0; r2 < RC / 2; ++r2) {
                    const dv2 x = a[r2];
                    s[2 * r2] += tq[u] * x.x;
                    s[2 * r2 + 1] += tq[u] * x.y;
                }
            }
        }
        for (; q < innerN; ++q) {
            const double t0 = T[base + step * q];
            const dv2* a0 = reinterpret_cast<const dv2*>(Ain + q * RB);
#pragma unroll
            for (int r2 = 0; r2 < RC / 2; ++r2) {
                const dv2 x0 = a0[r2];
                s[2 * r2] += t0 * x0.x;
                s[2 * r2 + 1] += t0 * x0.y;
            }
        }
        const double* w = Aout + o * RB;
#pragma unroll
        for (int r = 0; r < RC; ++r) acc[r] += s[r] * w[r];
    }
}

// M(idx, r) for mode n of one problem.  A task is (output index, chunk of 8 ranks):
// lanes run over the output index and the "inner" other index is looped over
// with the factor rows read from LDS; the outer factor multiplies each inner
// partial sum (KRP(i, j, k) = A(j) A(k) is never formed).  When the tasks leave
// threads idle, the "outer" other index is split over P <= 8 groups of tasks,
// which add into M one after another in a fixed order; otherwise (P = 1) each
// thread runs tasks t, t + NT, ... and writes them.
template <int NT, int RB>
__device__ void batch_mttkrp(const double* __restrict__ T, int I, int J, int K, int n,
                             const BatchLds& L) {
    constexpr int RC = 8, NGR = RB / RC;
    const int t = threadIdx.x;
    const int D = n == 0 ? I : (n == 1 ? J : K);
    const int outerN = n == 2 ? J : K;
    const int innerN = n == 0 ? J : I;
    const int nslot = D * NGR;
    int P = NT / nslot;
    if (P > 8) P = 8;
    if (P > outerN) P = outerN;
    if (P <= 1) {
        for (int task = t; task < nslot; task += NT) {
            const int idx = task % D, rc0 = RC * (task / D);
            double acc[RC];
            mttkrp_task<RB>(T, I, J, n, idx, rc0, 0, 1, outerN, innerN, L, acc);
            double* Mrow = L.M + idx * RB + rc0;
#pragma unroll
            for (int r = 0; r < RC; ++r) Mrow[r] = acc[r];
        }
        __syncthreads();
        return;
    }
    // here nslot * P <= NT: one task per thread
    const int idx = t % D;
    const int rgrp = (t / D) % NGR;
    const int grp = t / nslot;
    const bool active = grp < P;
    const int rc0 = RC * rgrp;
    double acc[RC];
    if (active) mttkrp_task<RB>(T, I, J, n, idx, rc0, grp, P, outerN, innerN, L, acc);
    double* Mrow = L.M + idx * RB + rc0;
    for (int g = 0; g < P; ++g) {
        if (active && grp == g) {
#pragma unroll
            for (int r = 0; r < RC; ++r) Mrow[r] = (g == 0 ? 0.0 : Mrow[r]) + acc[r];
        }
        __syncthreads();
    }
}

// Pivoted Cholesky of V (R <= 32, LAPACK xPSTF2 semantics as in the per-plan
// factorisation): wave 0, lane l owns index l; left-looking columns in UbT.
// Full rank: V is overwritten by U (row-major, pivoted order), diagU / piv set.
__device__ void batch_cholesky(int R, double tol, const BatchLds& L) {
    const int t = threadIdx.x;
    // zero UbT: entries t >= j of row l are read as zeros by the padded dot products
    for (int x = t; x < BATCH_MAXR * LDT; x += blockDim.x) L.UbT[x] = 0.0;
    __syncthreads();
    if (t < 64) {
        const int l = t;
        bool rem = l < R;
        const double vdiag = rem ? L.V[l * R + l] : 0.0;
        double sumsq = 0.0;
        int rank = R;
        for (int j = 0; j < R; ++j) {
            double best = rem ? vdiag - sumsq : -__builtin_inf();
            int bi = rem ? l : R;
            wave_argmax(best, bi);
            const int p = __builtin_amdgcn_readfirstlane(bi);
            const double d = best;
            const bool bad = (p >= R) || ((j == 0) ? !(d > 0.0) : !(d > tol));
            if (bad) { rank = j; break; }
            const double ujj = sqrt(d);
            const double inv = 1.0 / ujj;
            if (l == 0) { L.diagU[j] = ujj; L.piv[j] = p; }
            if (l == p) rem = false;
            double u = 0.0;
            if (rem) {
                const double* up = L.UbT + p * LDT;
                const double* ul = L.UbT + l * LDT;
                double acc[4] = {0.0, 0.0, 0.0, 0.0};
                for (int tt = 0; tt < j; tt += 8) {     // zero terms past j: whole batches
                    dv2 a[4], c[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        a[k] = *reinterpret_cast<const dv2*>(up + tt + 2 * k);
                        c[k] = *reinterpret_cast<const dv2*>(ul + tt + 2 * k);
                    }
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        acc[k] += a[k].x * c[k].x;
                        acc[k] += a[k].y * c[k].y;
                    }
                }
                u = (L.V[p * R + l] - ((acc[0] + acc[1]) + (acc[2] + acc[3]))) * inv;
                sumsq += u * u;
            }
            if (l < R) L.UbT[l * LDT + j] = u;
        }
        if (l == 0) L.finfo[0] = rank;
    }
    __syncthreads();
    const int rank = L.finfo[0];
    if (rank == R) {
        for (int x = t; x < R * R; x += blockDim.x) {
            const int j = x / R, tt = x % R;
            double v = 0.0;
            if (tt > j) v = L.UbT[L.piv[tt] * LDT + j];
            else if (tt == j) v = L.diagU[j];
            L.V[x] = v;
        }
        if (t == 0) { L.finfo[1] = 0; L.finfo[2] = R; }
    }
    __syncthreads();
}

// Z[k * R + c] = x_c with V x = e_k (x = P U^-1 U^-T P^T e_k, or the QR minimum-norm
// solution), one wave per unit right-hand side.
template <int NT>
__device__ void batch_solution_operator(int R, const BatchLds& L) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const bool fallback = L.finfo[1] != 0;
    const int rk = fallback ? L.finfo[2] : R;
    for (int k = wave; k < R; k += NT / 64) {
        double y[1];
        int pk = 0;
        if (!fallback) {
            pk = lane < R ? L.piv[lane] : 0;
            y[0] = (lane < R && pk == k) ? 1.0 : 0.0;
            for (int j = 0; j < R; ++j) {          // U^T z = y
                const double zj = bcast<1>(y, j) / L.diagU[j];
                if (lane == j) y[0] = zj;
                else if (lane > j && lane < R) y[0] -= L.V[j * R + lane] * zj;
            }
            for (int j = R - 1; j >= 0; --j) {     // U w = z
                const double wj = bcast<1>(y, j) / L.diagU[j];
                if (lane == j) y[0] = wj;
                else if (lane < j) y[0] -= L.V[lane * R + j] * wj;
            }
        } else {
            pk = lane < R ? L.jpvt[lane] : 0;
            y[0] = (lane == k) ? 1.0 : 0.0;
            fallback_apply<1>(y, lane, R, rk, L.V, L.qtau, L.ztau);
        }
        if (lane < R) L.Z[k * R + pk] = y[0];
    }
    __syncthreads();
}

// One mode update after the MTTKRP: V, factorisation, X = M Z^T, row_norm, Gram.
template <int NT, int RB>
__device__ void batch_update(int n, int D, int R, const BatchLds& L) {
    const int t = threadIdx.x;
    GramSet gs;
    gs.N = 3;
    for (int m = 0; m < 3; ++m) gs.G[m] = L.G[m];
    for (int m = 3; m < ITCPD_MAX_ORDER; ++m) gs.G[m] = nullptr;
    fill_hadamard<NT>(L.V, gs, n, R * R);
    __syncthreads();
    batch_cholesky(R, 1e-6, L);
    if (L.finfo[0] < R) {
        qr_fallback<NT>(gs, n, R, L.V, L.qtau, L.ztau, L.jpvt, L.vwork, L.finfo, L.redv, L.redi,
                        L.finfo[0]);
        __syncthreads();
    }
    batch_solution_operator<NT>(R, L);
    // X(i, c) = sum_k M(i, k) Z(k, c) into A_n (ld RB, pad columns stay zero)
    constexpr int PER = (BATCH_MAXD * BATCH_MAXR + NT - 1) / NT;
    double xv[PER];
    const int nx = D * R;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int x = t + NT * u;
        double acc = 0.0;
        if (x < nx) {
            const int i = x / R, c = x - (x / R) * R;
            for (int k = 0; k < R; ++k) acc += L.M[i * RB + k] * L.Z[k * R + c];
        }
        xv[u] = acc;
    }
    __syncthreads();
    double* A = L.A[n];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int x = t + NT * u;
        if (x < nx) A[(x / R) * RB + (x % R)] = xv[u];
    }
    __syncthreads();
    // row_norm: lambda_c = sqrt(sum_i X(i,c)^2), A = X / lambda
    for (int c = t; c < R; c += NT) {
        double ss = 0.0;
        for (int i = 0; i < D; ++i) ss += A[i * RB + c] * A[i * RB + c];
        L.lam[c] = sqrt(ss);
    }
    __syncthreads();
    for (int x = t; x < D * R; x += NT) {
        const int i = x / R, c = x % R;
        A[i * RB + c] /= L.lam[c];
    }
    __syncthreads();
    double* G = L.G[n];
    for (int x = t; x < R * R; x += NT) {
        const int c = x / R, c2 = x % R;
        double ss = 0.0;
        for (int i = 0; i < D; ++i) ss += A[i * RB + c] * A[i * RB + c2];
        G[x] = ss;
    }
    __syncthreads();
}

template <int NT, int RB>
__device__ __forceinline__ void batch_problem(const BatchProblem& pb, const double* __restrict__ Tall,
                              double* __restrict__ factors, double* __restrict__ lambdas,
                              double* __restrict__ fits, int sweeps, const BatchLds& L) {
    const int t = threadIdx.x;
    const int dims[3] = {pb.I, pb.J, pb.K};
    const int R = pb.R;
    const double* T = Tall + pb.toff;
    // factors (column-major per mode, concatenated) -> LDS rows, zero pad
    int64_t off = pb.foff;
    for (int n = 0; n < 3; ++n) {
        const int D = dims[n];
        for (int x = t; x < D * RB; x += NT) {
            const int i = x / RB, c = x % RB;
            L.A[n][x] = c < R ? factors[off + i + (int64_t)D * c] : 0.0;
        }
        off += (int64_t)D * R;
    }
    __syncthreads();
    for (int n = 0; n < 3; ++n) {
        const int D = dims[n];
        for (int x = t; x < R * R; x += NT) {
            const int c = x / R, c2 = x % R;
            double ss = 0.0;
            for (int i = 0; i < D; ++i) ss += L.A[n][i * RB + c] * L.A[n][i * RB + c2];
            L.G[n][x] = ss;
        }
    }
    // ||T||^2
    const int64_t nel = (int64_t)pb.I * pb.J * pb.K;
    double ss = 0.0;
    for (int64_t x = t; x < nel; x += NT) ss += T[x] * T[x];
    const double norm2 = block_sum<NT>(ss, L.redv);
    __syncthreads();

    for (int sw = 0; sw < sweeps; ++sw) {
        for (int n = 0; n < 3; ++n) {
            batch_mttkrp<NT, RB>(T, pb.I, pb.J, pb.K, n, L);
            batch_update<NT, RB>(n, dims[n], R, L);
        }
    }

    // fit from the last mode's MTTKRP (still in M) and the Grams
    double inner = 0.0, fs = 0.0;
    for (int x = t; x < pb.K * R; x += NT) {
        const int k = x / R, c = x % R;
        inner += L.M[k * RB + c] * L.A[2][k * RB + c] * L.lam[c];
    }
    for (int x = t; x < R * R; x += NT) {
        const int c = x / R, c2 = x % R;
        fs += L.lam[c] * L.lam[c2] * L.G[0][x] * L.G[1][x] * L.G[2][x];
    }
    inner = block_sum<NT>(inner, L.redv);
    fs = block_sum<NT>(fs, L.redv);
    if (t == 0) {
        const double ref = sqrt(norm2);
        fits[pb.index] = 1.0 - sqrt(fabs(norm2 + fs - 2.0 * fabs(inner))) / ref;
    }
    off = pb.foff;
    for (int n = 0; n < 3; ++n) {
        const int D = dims[n];
        for (int x = t; x < D * R; x += NT) {
            const int i = x % D, c = x / D;
            factors[off + x] = L.A[n][i * RB + c];
        }
        off += (int64_t)D * R;
    }
    for (int c = t; c < R; c += NT) lambdas[pb.loff + c] = L.lam[c];
    __syncthreads();
}

// Persistent: every workgroup pulls problem indices from `counter` (in the order
// of `order`) until the queue is empty; the exit test is uniform per workgroup.
template <int NT>
__global__ __launch_bounds__(NT) void k_batch_als(const BatchProblem* __restrict__ probs,
                                                  const int* __restrict__ order, int nprob,
                                                  const double* __restrict__ T,
                                                  double* __restrict__ factors,
                                                  double* __restrict__ lambdas,
                                                  double* __restrict__ fits, int sweeps,
                                                  int* __restrict__ counter) {
    extern __shared__ double lds[];
    __shared__ int s_next;
    const BatchLds L = carve<NT>(lds);
    for (;;) {
        if (threadIdx.x == 0) s_next = atomicAdd(counter, 1);
        __syncthreads();
        const int q = s_next;
        __syncthreads();
        if (q >= nprob) break;
        const BatchProblem pb = probs[order[q]];
        if (pb.R <= 8) batch_problem<NT, 8>(pb, T, factors, lambdas, fits, sweeps, L);
        else if (pb.R <= 16) batch_problem<NT, 16>(pb, T, factors, lambdas, fits, sweeps, L);
        else if (pb.R <= 24) batch_problem<NT, 24>(pb, T, factors, lambdas, fits, sweeps, L);
        else batch_problem<NT, 32>(pb, T, factors, lambdas, fits, sweeps, L);
    }
}

__global__ void k_batch_reset(int* __restrict__ counter) {
    if (threadIdx.x == 0) counter[0] = 0;
}

}  // namespace

size_t batch_lds_bytes(int nt) {
    return (nt == 256 ? batch_lds_doubles<256>() : batch_lds_doubles<512>()) * 8;
}

// Enqueues reset + persistent kernel on `st` (also used under stream capture).
// nt: threads per workgroup, 256 or 512 (1024 would spill the rank chunks).
int launch_batch(const BatchDev& b, int nt, int sweeps, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, b.counter);
    ITCPD_HIP(hipGetLastError());
    const size_t lds = batch_lds_bytes(nt);
    const unsigned grid = (unsigned)b.grid;
    switch (nt) {
        case 256:
            hipLaunchKernelGGL(k_batch_als<256>, dim3(grid), dim3(256), lds, st, b.probs, b.order,
                               b.nprob, b.T, b.factors, b.lambdas, b.fits, sweeps, b.counter);
            break;
        default:
            hipLaunchKernelGGL(k_batch_als<512>, dim3(grid), dim3(512), lds, st, b.probs, b.order,
                               b.nprob, b.T, b.factors, b.lambdas, b.fits, sweeps, b.counter);
            break;
    }
    ITCPD_HIP(hipGetLastError());
    return ITCPD_OK;
}

}  // namespace itcpd
