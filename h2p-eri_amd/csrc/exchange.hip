// Exchange matrix through the H2 representation on the MI355X:
//
//   K_mq = sum_{n,p} (mn|pq) D_np            (src/H2ERI_build_exchange.c:1356-1450)
//
// The reference evaluates K with partial multi-column H2 products, one per
// (N,Q) shell pair (:1021-1353), with gathered inputs (:609-688) and an atomic
// scatter of the outputs into K (:691-771).  This device version evaluates
// the same sum block by block of A_H2 instead:
//
//   A_H2 = sum_i D_ii + sum_{(i,j)} [F_ij + F_ij^T],   F_ij = L_ij R_ij^T
//
// - diagonal near-field blocks are applied once;
// - every reduced admissible (B) or inadmissible (D) pair is applied with its
//   transpose (src/H2ERI_matvec.c:299-407,572-660).
// L and R come from the expanded nested bases (P_n = U of a leaf, or
// blockdiag(P of children) U of an inner node) and the block's core (dense C,
// or U V^T for a low-rank block), as listed at build_terms.
//
// Row x of A_H2 is the basis-function pair (a_x, b_x) of a screened shell
// pair, stored with the smaller shell first.  When the two shells differ,
// the row stands for both (ab| and (ba|; a same-shell pair has one row per
// ordered (a,b).  For a term with rows in node i and columns in node j, the
// contraction factorises through the basis-function sets B(i), B(j) of the
// two nodes:
//
//   G[n][t][q] = sum_{c in j, (p,q) orientation of c} R[c][t] D[n][p]   (ket, n in B(i), q in B(j))
//   K[m][q]   += sum_{r in i, (m,n) orientation of r} sum_t L[r][t] G[n][t][q]   (bra)
//
// For big nodes this costs kk (|B(i)| m_j + m_i |B(j)|) rather than the
// m_i m_j kk of forming F.  Dense near-field blocks are contracted directly.
//
// Accumulation and symmetry:
// - For symmetric D the transpose terms give the transpose of the F terms:
//   K = K1 + K1^T + K_diag.
// - Contributions go into K1 and K_diag with FP64 atomic adds, as the
//   reference's scatter does (:767), so the last bits of K depend on the
//   order of the adds.
// - The L and R products of the cores, the ket and the bra are FP64 MFMA
//   tiles (v_mfma_f64_16x16x4f64) with FP64 accumulation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>

#include "h2eri_internal.h"

namespace h2e {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

enum KSrc { KS_PB = 0, KS_CB = 1, KS_SC = 2 };

struct KTerm {
    int direct;           // 1: dense core contracted directly (identity bases on both sides)
    int diag;             // 1: diagonal near-field block (into K_diag), 0: into K1
    int ni, nj;           // node-list ids of the row and column ranges
    int m0, m1, kk;
    int Lsrc, Lld;        // L[r][t] = src[Loff + r*Lld + t]   (direct: C[r][c] = src[Loff + r*Lld + c])
    long long Loff;
    int Rsrc, Rrs, Rcs;   // R[c][t] = src[Roff + c*Rrs + t*Rcs]
    long long Roff;
    long long Goff;       // G in its chunk's scratch: [q / 8][n][t][q % 8], gld(|B(j)|) / 8 blocks
    int pad_[2];
};

struct KGemm {           // out (m x n, ld ldo) = A (m x k, PB, ld lda) * X (k x n, CB, ld ldx)
    long long Aoff, Xoff, Ooff;
    int m, n, k, lda, ldx, ldo;
};

struct KTask {           // one workgroup: term and tile origin
    int term;
    int e0;               // ket: first local n; bra: first local q; direct: first (m,q) index
    int t0;               // ket: first t; bra: split of local m
    int pad_;             // ket: first local q
};

struct KGemmTask {
    int g, row0, col0, pad_;
};

struct KNode {           // per row/column range: basis-function set and the orientation lists
    int bf_off, nbf;      // B(node) = bfidx[bf_off .. bf_off + nbf)
    int q_ptr;            // byQ CSR rows: qptr[q_ptr .. q_ptr + nbf]   entries (column c, p)
    int m_ptr;            // byM CSR rows: mptr[m_ptr .. m_ptr + nbf]   entries (row r, local n)
    int s_ptr;            // byN split table: nspl[s_ptr + n * (nsplit + 1) + s] = first entry (row r, local m)
    int nsplit;           //   of local n with local m >= s * ssz (entries of each n sorted by m)
    int ssz;              // functions m per split: the nbf functions in nsplit = ceil(nbf / KSPLIT) even parts
};

static const int KWG = 256;
// G is stored in blocks of 8 functions q, [q / 8][n][t][q % 8]: a ket task
// (8 q) then writes whole 64-byte (n, t) records, two per 128-byte line, and
// four consecutive q form one aligned 32-byte store.  gld(|B(j)|) is the
// padded q count.
__host__ __device__ inline int gld(int nbf_j) { return (nbf_j + 7) & ~7; }
__device__ inline size_t gidx(int q, int n, int t, int nbf_i, int kk) {
    return (((size_t)(q >> 3) * nbf_i + n) * kk + t) * 8 + (q & 7);
}
// Workgroups are dispatched to the 8 XCDs round-robin by id, and each XCD has
// its own L2.  This bijection of [0, n) hands out runs of C consecutive tasks
// of the list, run j to XCD j mod 8 (the tail of fewer than 8 runs is split
// into 8 contiguous ranges), so tasks next to each other in the list, which
// share G columns, L rows or R rows, read them through one L2.
__device__ inline int xcd_task(int bid, int n, int C) {
    const int full = n / (8 * C);
    const int x = bid & 7, k = bid >> 3;
    if (bid < full * 8 * C) return ((k / C) * 8 + x) * C + k % C;
    const int n2 = n - full * 8 * C, k2 = (bid - full * 8 * C) >> 3, q = n2 >> 3, r = n2 & 7;
    return full * 8 * C + x * q + min(x, r) + k2;
}
// runs of consecutive tasks per XCD (measured on C30H62: K 16.3 ms with the
// plain round-robin order, 15.5 ms with these; one contiguous range per XCD
// cut the bra's HBM reads from 39 to 22 GB but unbalanced the XCDs)
static const int KRUN_KET = 64, KRUN_BRA = 8;
static const int KQC = 8;       // column functions q per ket task (one group of four per wave)
// most local m rows per bra task (LDS block KSPLIT x 32); on C30H62, with
// tasks of 16 q, K took 13.0 ms with 256, 12.5 ms with 192 and even splits,
// 12.9 ms with 160
static const int KSPLIT = 192;
static const int KEB = 4;       // entry blocks of 16 sharing one walk over t in the bra

// Ket on MFMA: G[n][t][q] = sum over the entries (c, p) of q of R[c][t] D[n][p].
// Per q this is a GEMM (rows t, columns n, inner dimension the entries of q).
// Task = (term, t tile of 16, 64 functions n, KQC functions q); the four
// waves split it 2 x 2 into 32 functions n (two MFMA tiles that share every
// gathered R operand and entry index) by KQC / 2 functions q.  A wave computes
// four consecutive q together, walking their entry lists in step, and stores
// the four q of each (n, t) record as one aligned 32-byte write (single
// 8-byte writes scattered over 16 rows cost four times their bytes in HBM
// write traffic).  Registers are capped for four waves per SIMD: uncapped, the
// eight accumulators take 200 and halve the waves that hide the gathers.
__global__ __launch_bounds__(KWG) __attribute__((amdgpu_waves_per_eu(4))) void kx_ket(const KTask *__restrict__ tasks, const KTerm *__restrict__ terms,
                                               const KNode *__restrict__ nodes, const int *__restrict__ bfidx,
                                               const int *__restrict__ qptr, const int2 *__restrict__ qent,
                                               const double *__restrict__ PB, const double *__restrict__ CB,
                                               const double *__restrict__ SC, double *__restrict__ GS,
                                               const double *__restrict__ D, int nbf) {
    const KTask tk = tasks[xcd_task(blockIdx.x, gridDim.x, KRUN_KET)];   // term, nl0, t0, ql0
    const KTerm t = terms[tk.term];
    const KNode ni = nodes[t.ni], nj = nodes[t.nj];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // wave w: functions n [nl0, nl0 + 32) as two MFMA tiles that share every
    // R operand, and half of the task's functions q
    const int nl0 = tk.e0 + 32 * (w & 1);
    static_assert(KQC % 8 == 0, "each wave stores whole groups of four q");
    const int qbeg = tk.pad_ + (KQC / 2) * (w >> 1);
    const int qend = min(qbeg + KQC / 2, nj.nbf);
    if (nl0 >= ni.nbf || qbeg >= qend) return;   // whole wave
    const double *R = (t.Rsrc == KS_PB ? PB : (t.Rsrc == KS_CB ? CB : SC)) + t.Roff;
    int nl[2], bfn[2];
    bool nv[2];
    #pragma unroll
    for (int h = 0; h < 2; h++) {
        nl[h] = nl0 + 16 * h + (lane & 15);
        nv[h] = nl[h] < ni.nbf;
        // D is symmetric (the entry point's contract), so D[n][p] is read as
        // D[p][n]: the 16 lanes of an entry (consecutive functions n of the
        // node) read along one row of D instead of down a column
        bfn[h] = nv[h] ? bfidx[ni.bf_off + nl[h]] : 0;
    }
    const int ta = tk.t0 + (lane & 15);   // A operand row (t)
    const bool tv = ta < t.kk;
    const int kl = lane >> 4;
    double *G = GS + t.Goff;
    for (int qg = qbeg; qg < qend; qg += 4) {
        // The entry lists of the four q are walked together (entries past a
        // list are zeros), so each iteration has the gathers of eight
        // independent accumulations in flight.
        int qa[4], qb[4], len = 0;
        #pragma unroll
        for (int j = 0; j < 4; j++) {
            const int ql = qg + j;
            qa[j] = ql < qend ? qptr[nj.q_ptr + ql] : 0;
            qb[j] = ql < qend ? qptr[nj.q_ptr + ql + 1] : 0;
            len = max(len, qb[j] - qa[j]);
        }
        d4 acc[2][4];
        #pragma unroll
        for (int h = 0; h < 2; h++)
            #pragma unroll
            for (int j = 0; j < 4; j++) acc[h][j] = d4{0.0, 0.0, 0.0, 0.0};
        // entry indices come in windows of 16 per q: lane l loads entry
        // (l & 15) of q (l >> 4) and the lanes of each step take theirs by
        // a cross-lane shuffle, which keeps them off the vector memory path
        const int jq = lane >> 4;
        const int wa = jq == 0 ? qa[0] : jq == 1 ? qa[1] : jq == 2 ? qa[2] : qa[3];
        const int wb = jq == 0 ? qb[0] : jq == 1 ? qb[1] : jq == 2 ? qb[2] : qb[3];
        for (int w0 = 0; w0 < len; w0 += 16) {
            const int ew = wa + w0 + (lane & 15);
            const int2 mine = ew < wb ? qent[ew] : make_int2(-1, 0);
            for (int s4 = 0; s4 < 16 && w0 + s4 < len; s4 += 4) {
                int2 cp[4];
                #pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int src = 16 * j + s4 + kl;
                    cp[j] = make_int2(__shfl(mine.x, src), __shfl(mine.y, src));
                }
                double a[4], b[2][4];
                #pragma unroll
                for (int j = 0; j < 4; j++) {
                    const bool ev = cp[j].x >= 0;
                    a[j] = (ev && tv) ? R[(long long)cp[j].x * t.Rrs + (long long)ta * t.Rcs] : 0.0;
                    #pragma unroll
                    for (int h = 0; h < 2; h++) b[h][j] = (ev && nv[h]) ? D[(size_t)cp[j].y * nbf + bfn[h]] : 0.0;
                }
                #pragma unroll
                for (int j = 0; j < 4; j++)
                    #pragma unroll
                    for (int h = 0; h < 2; h++) acc[h][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[j], b[h][j], acc[h][j], 0, 0, 0);
            }
        }
        #pragma unroll
        for (int h = 0; h < 2; h++)
            #pragma unroll
            for (int r = 0; r < 4; r++) {
                const int tt = tk.t0 + kl + 4 * r;
                if (tt < t.kk && nv[h])
                    *reinterpret_cast<d4 *>(G + gidx(qg, nl[h], tt, ni.nbf, t.kk)) =
                        d4{acc[h][0][r], acc[h][1][r], acc[h][2][r], acc[h][3][r]};
            }
    }
}

// Bra on MFMA: K[m][q] += sum_n sum_{(r,m) of n} sum_t L[r][t] G[n][t][q].
// For each n the rows r that pair with n form a GEMM (rows: those entries,
// columns q, inner dimension t).  Task = (term, 32 functions q as two MFMA
// column tiles that share every L operand, split of ssz <= KSPLIT functions
// m); wave w takes the functions n = w, w + 4, ... and adds its 16 x 16
// output tiles into an LDS block of the split with FP64 LDS atomics (tiles
// of different n may share an m); the block is added into K at the end.
// Tasks of 16 q read every L row twice as often: K 12.5 ms against 12.1.  The t loop issues the operands of four MFMA steps at once,
// and one walk over t serves up to KEB blocks of 16 entries of n, which share
// its G operands.
__global__ __launch_bounds__(KWG) void kx_bra(const KTask *__restrict__ tasks, const KTerm *__restrict__ terms,
                                               const KNode *__restrict__ nodes, const int *__restrict__ bfidx,
                                               const int *__restrict__ nspl, const int2 *__restrict__ nent,
                                               const double *__restrict__ PB, const double *__restrict__ CB,
                                               const double *__restrict__ SC, const double *__restrict__ GS,
                                               double *__restrict__ K1, double *__restrict__ Kd, int nbf) {
    __shared__ double Kb[KSPLIT * 32];
    const KTask tk = tasks[xcd_task(blockIdx.x, gridDim.x, KRUN_BRA)];   // term, ql0, split
    const KTerm t = terms[tk.term];
    const KNode ni = nodes[t.ni], nj = nodes[t.nj];
    const int split = tk.t0, ql0 = tk.e0;
    const int ml0 = split * ni.ssz;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, kl = lane >> 4;
    for (int i = threadIdx.x; i < KSPLIT * 32; i += KWG) Kb[i] = 0.0;
    __syncthreads();
    const double *L = (t.Lsrc == KS_PB ? PB : (t.Lsrc == KS_CB ? CB : SC)) + t.Loff;
    const double *G = GS + t.Goff;
    bool qv[2];
    #pragma unroll
    for (int h = 0; h < 2; h++) qv[h] = ql0 + 16 * h + (lane & 15) < nj.nbf;
    const int ns1 = ni.nsplit + 1;
    for (int nl = w; nl < ni.nbf; nl += KWG / 64) {
        const int eb = nspl[ni.s_ptr + nl * ns1 + split], ee = nspl[ni.s_ptr + nl * ns1 + split + 1];
        const double *Gn[2];
        #pragma unroll
        for (int h = 0; h < 2; h++) Gn[h] = G + gidx(min(ql0 + 16 * h + (lane & 15), gld(nj.nbf) - 1), nl, 0, ni.nbf, t.kk);
        // up to KEB blocks of 16 entries per walk over t, so that each G
        // operand load feeds KEB accumulations
        for (int e0 = eb; e0 < ee; e0 += 16 * KEB) {
            const double *Lr[KEB];
            bool evj[KEB];
            #pragma unroll
            for (int j = 0; j < KEB; j++) {
                const int e = e0 + 16 * j + (lane & 15);
                evj[j] = e < ee;
                Lr[j] = L + (long long)(evj[j] ? nent[e].x : 0) * t.Lld;
            }
            const int nblk = min(KEB, (ee - e0 + 15) / 16);   // blocks with entries (wave-uniform)
            d4 acc[KEB][2];
            #pragma unroll
            for (int j = 0; j < KEB; j++) acc[j][0] = acc[j][1] = d4{0.0, 0.0, 0.0, 0.0};
            for (int t0 = 0; t0 < t.kk; t0 += 16) {
                double b[2][4];
                #pragma unroll
                for (int h = 0; h < 2; h++)
                    #pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int tt = t0 + 4 * u + kl;
                        b[h][u] = (qv[h] && tt < t.kk) ? Gn[h][(size_t)tt * 8] : 0.0;
                    }
                #pragma unroll
                for (int j = 0; j < KEB; j++) {
                    if (j >= nblk) break;
                    double a[4];
                    #pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int tt = t0 + 4 * u + kl;
                        a[u] = (evj[j] && tt < t.kk) ? Lr[j][tt] : 0.0;
                    }
                    #pragma unroll
                    for (int u = 0; u < 4; u++)
                        #pragma unroll
                        for (int h = 0; h < 2; h++) acc[j][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[h][u], acc[j][h], 0, 0, 0);
                }
            }
            #pragma unroll
            for (int j = 0; j < KEB; j++) {
                if (j >= nblk) break;
                #pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int er = e0 + 16 * j + kl + 4 * r;
                    if (er < ee) {
                        const int mrow = nent[er].y - ml0;
                        #pragma unroll
                        for (int h = 0; h < 2; h++)
                            if (qv[h]) atomicAdd(&Kb[mrow * 32 + 16 * h + (lane & 15)], acc[j][h][r]);
                    }
                }
            }
        }
    }
    __syncthreads();
    double *Kout = t.diag ? Kd : K1;
    const int mlN = min(ni.ssz, ni.nbf - ml0);
    for (int i = threadIdx.x; i < mlN * 32; i += KWG) {
        const int qq = ql0 + (i & 31);
        const double v = Kb[i];
        if (qq < nj.nbf && v != 0.0)
            unsafeAtomicAdd(Kout + (size_t)bfidx[ni.bf_off + ml0 + (i >> 5)] * nbf + bfidx[nj.bf_off + qq], v);
    }
}

// Dense core with identity bases: K[m][q] += sum_{(r,n)} sum_{(c,p)} C[r][c] D[n][p].
__global__ __launch_bounds__(KWG) void kx_direct(const KTask *__restrict__ tasks, const KTerm *__restrict__ terms,
                                                  const KNode *__restrict__ nodes, const int *__restrict__ bfidx,
                                                  const int *__restrict__ qptr, const int2 *__restrict__ qent,
                                                  const int *__restrict__ mptr, const int2 *__restrict__ ment,
                                                  const double *__restrict__ CB, const double *__restrict__ D,
                                                  double *__restrict__ K1, double *__restrict__ Kd, int nbf) {
    const KTask tk = tasks[blockIdx.x];
    const KTerm t = terms[tk.term];
    const KNode ni = nodes[t.ni], nj = nodes[t.nj];
    const int e = tk.e0 + threadIdx.x;
    if (e >= ni.nbf * nj.nbf) return;
    const int ml = e / nj.nbf, ql = e - ml * nj.nbf;
    const double *C = CB + t.Loff;
    double acc = 0.0;
    const int q0 = qptr[nj.q_ptr + ql], q1 = qptr[nj.q_ptr + ql + 1];
    for (int j = mptr[ni.m_ptr + ml]; j < mptr[ni.m_ptr + ml + 1]; j++) {
        const int2 rn = ment[j];
        const double *Cr = C + (long long)rn.x * t.Lld;
        const double *Dn = D + (size_t)bfidx[ni.bf_off + rn.y] * nbf;
        double s = 0.0;
        for (int i = q0; i < q1; i++) {
            const int2 cp = qent[i];
            s = fma(Cr[cp.x], Dn[cp.y], s);
        }
        acc += s;
    }
    const int m = bfidx[ni.bf_off + ml], q = bfidx[nj.bf_off + ql];
    unsafeAtomicAdd((t.diag ? Kd : K1) + (size_t)m * nbf + q, acc);
}

// Grouped FP64 GEMM on MFMA: a 64 x 64 output tile per workgroup, wave w owns
// rows [16w, 16w+16) and four 16 x 16 column tiles.  Operand layout of
// v_mfma_f64_16x16x4f64 (checked on gfx950): lane l holds A[l%16][l/16] and
// B[l/16][l%16]; output register r of lane l is C[l/16 + 4r][l%16].
__global__ __launch_bounds__(KWG) void kx_gemm(const KGemmTask *__restrict__ tasks, const KGemm *__restrict__ gemms,
                                                const double *__restrict__ PB, const double *__restrict__ CB,
                                                double *__restrict__ SC) {
    const KGemmTask tk = tasks[blockIdx.x];
    const KGemm g = gemms[tk.g];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ar = tk.row0 + 16 * w + (lane & 15);
    const int kl = lane >> 4;
    const double *A = PB + g.Aoff;
    const double *X = CB + g.Xoff;
    d4 acc[4];
    #pragma unroll
    for (int c = 0; c < 4; c++) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < g.k; k0 += 4) {
        const int kk = k0 + kl;
        const double a = (ar < g.m && kk < g.k) ? A[(long long)ar * g.lda + kk] : 0.0;
        #pragma unroll
        for (int c = 0; c < 4; c++) {
            const int col = tk.col0 + 16 * c + (lane & 15);
            const double b = (kk < g.k && col < g.n) ? X[(long long)kk * g.ldx + col] : 0.0;
            acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
        }
    }
    #pragma unroll
    for (int c = 0; c < 4; c++) {
        const int col = tk.col0 + 16 * c + (lane & 15);
        #pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = tk.row0 + 16 * w + (lane >> 4) + 4 * r;
            if (row < g.m && col < g.n) SC[g.Ooff + (long long)row * g.ldo + col] = acc[c][r];
        }
    }
}

// Density screening (src/H2ERI_build_exchange.c:774-838): block (N,P) of D
// takes part in K only when some |D_np|, n in N and p in P, exceeds
// DTOL = 1e-10 (:790).  The reference restricts every P list to dlist[N]
// (:410-417,1397-1398), which sums over the kept blocks only; here every
// other block is written as zeros into Ds, the density the kernels read.
// One thread per block; the threads of a row of blocks read consecutive
// columns of the same rows of D.
__global__ void kx_dlist(const int *__restrict__ sbf, int nshell, const double *__restrict__ D, double *__restrict__ Ds,
                         int nbf) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nshell * nshell) return;
    const int n = (int)(idx / nshell), p = (int)(idx - (long long)n * nshell);
    const int r0 = sbf[n], r1 = sbf[n + 1], c0 = sbf[p], c1 = sbf[p + 1];
    bool keep = false;
    for (int i = r0; i < r1; i++)
        for (int j = c0; j < c1; j++) keep |= fabs(D[(size_t)i * nbf + j]) > 1e-10;
    for (int i = r0; i < r1; i++)
        for (int j = c0; j < c1; j++) Ds[(size_t)i * nbf + j] = keep ? D[(size_t)i * nbf + j] : 0.0;
}

__global__ void kx_zero(double *__restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t s = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = 0.0;
}

// K = K1 + K1^T + K_diag
__global__ void kx_final(const double *__restrict__ K1, const double *__restrict__ Kd, double *__restrict__ K, int nbf) {
    const size_t n2 = (size_t)nbf * nbf;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t s = (size_t)gridDim.x * blockDim.x;
    for (; i < n2; i += s) {
        const size_t r = i / nbf, c = i - r * nbf;
        K[i] = K1[i] + K1[c * nbf + r] + Kd[i];
    }
}

// ---------------------------------------------------------------------------
// Host plan
// ---------------------------------------------------------------------------
struct KPlan {
    std::vector<double> PB, CB;
    std::vector<KTerm> terms;
    std::vector<KGemm> gemms;
    std::vector<KNode> nodes;
    std::vector<int> bfidx, qptr, mptr, nspl;
    std::vector<int2> qent, ment, nent;
    std::vector<KTask> ket, bra, direct;
    std::vector<KGemmTask> gtasks;
    long long nSC = 0;      // L / R products (computed once)
    long long nG = 0;       // G scratch of the largest chunk of terms
    std::vector<int> ket_chunk, bra_chunk;   // first ket / bra task of every chunk, and the end
    double flops = 0;
};

// Default bound of the G scratch of one chunk (bytes): the terms run in
// chunks whose G fits it, each chunk's ket then its bra; with two or more
// chunks, chunk c + 1's ket runs beside chunk c's bra on a second stream, in
// a second region (the scratch is twice the bound).  Measured K per build
// against the bound (unbounded / 4 / 2 / 1 GiB / 256 MiB): C30H62 12.2 /
// 12.7 / 13.1 / 13.9 / 22.2 ms with 16.0 / 10.3 / 6.0 / 3.9 / 2.3 GB
// resident; (H2O)64 75.7 ms unbounded (76 GB), 80.1 ms at 4 and at 1 GiB
// (16.0 and 9.6 GB).  On one stream the chunks cost more (C30H62 16.7 ms at
// 1 GiB): a launch of a small chunk has too few tasks to fill the GPU.
static const double KG_BUDGET_DEFAULT = 2.0 * 1024.0 * 1024.0 * 1024.0;

struct KState {
    KPlan P;
    bool ready = false;
    double g_budget = KG_BUDGET_DEFAULT;   // bytes of the G scratch (H2ERI_set_exchange_scratch)
    double *PB = nullptr, *CB = nullptr, *SC = nullptr, *K1 = nullptr, *Kd = nullptr, *D = nullptr, *K = nullptr;
    double *GS = nullptr;   // G of two chunks of terms (chunk c in region c % 2)
    hipStream_t side = nullptr;                        // odd chunks, beside the caller's stream
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    double *Ds = nullptr;   // the density after screening (the kernels' D)
    int *sbf = nullptr;     // shell_bf_sidx
    KTerm *terms = nullptr;
    KGemm *gemms = nullptr;
    KNode *nodes = nullptr;
    int *bfidx = nullptr, *qptr = nullptr, *mptr = nullptr, *nspl = nullptr;
    int2 *qent = nullptr, *ment = nullptr, *nent = nullptr;
    KTask *ket = nullptr, *bra = nullptr, *direct = nullptr;
    KGemmTask *gtasks = nullptr;
    size_t resident = 0;
};

static void kfree(void *&p) {
    if (p) { (void)hipFree(p); p = nullptr; }
}

static void kstate_release(KState *k) {
    void **ps[] = {(void **)&k->PB, (void **)&k->CB, (void **)&k->SC, (void **)&k->GS, (void **)&k->K1, (void **)&k->Kd, (void **)&k->D,
                   (void **)&k->K, (void **)&k->terms, (void **)&k->gemms, (void **)&k->nodes, (void **)&k->bfidx,
                   (void **)&k->qptr, (void **)&k->mptr, (void **)&k->qent, (void **)&k->ment, (void **)&k->nspl,
                   (void **)&k->nent, (void **)&k->ket,
                   (void **)&k->bra, (void **)&k->direct, (void **)&k->gtasks, (void **)&k->Ds, (void **)&k->sbf};
    for (void **p : ps) kfree(*p);
    if (k->side) { (void)hipStreamDestroy(k->side); k->side = nullptr; }
    if (k->ev_fork) { (void)hipEventDestroy(k->ev_fork); k->ev_fork = nullptr; }
    if (k->ev_join) { (void)hipEventDestroy(k->ev_join); k->ev_join = nullptr; }
    k->ready = false;
    k->P = KPlan();
}

struct KBuilder {
    H2ERI_p h;
    KPlan &P;
    double g_budget;                 // bytes of G per chunk of terms
    std::vector<long long> gsize;    // G doubles of every term (0: direct)
    std::vector<int> ra, rb;
    std::vector<char> sw;
    std::vector<long long> pb_off;   // expanded basis of each tree node in PB (-1: none)
    std::vector<int> node_list;      // tree node -> KNode id (-1: not built)

    KBuilder(H2ERI_p h_, KPlan &p_, double g_budget_) : h(h_), P(p_), g_budget(g_budget_) {}

    int rank_of(int n) const { return h->U[n] ? h->U[n]->ncol : 0; }
    int rows_of(int n) const { return h->mat_cluster[2 * n + 1] - h->mat_cluster[2 * n] + 1; }

    bool pair_rows() {
        const int N = h->krnl_mat_size, np = h->num_sp;
        ra.assign(N, -1); rb.assign(N, -1); sw.assign(N, 0);
        for (int i = 0; i < np; i++) {
            const int s0 = h->sp_shell_idx[i], s1 = h->sp_shell_idx[i + np];
            const int f0 = h->shell_bf_sidx[s0], n0 = h->shell_bf_sidx[s0 + 1] - f0;
            const int f1 = h->shell_bf_sidx[s1], n1 = h->shell_bf_sidx[s1 + 1] - f1;
            for (int j = 0; j < n1; j++)
                for (int k = 0; k < n0; k++) {
                    const int x = h->sp_bfp_sidx[i] + j * n0 + k;
                    if (x < 0 || x >= N) { set_error(ERR_ARG, "exchange: shell-pair layout out of range"); return false; }
                    ra[x] = f0 + k; rb[x] = f1 + j; sw[x] = s0 != s1;
                }
        }
        for (int x = 0; x < N; x++)
            if (ra[x] < 0) { set_error(ERR_ARG, "exchange: row %d belongs to no shell pair", x); return false; }
        return true;
    }

    // expanded nested bases of every node with a basis, leaves first
    void expand_bases() {
        const int nn = h->n_node, nl = h->n_leaf_node;
        pb_off.assign(nn, -1);
        for (int lv = h->max_level; lv >= 0; lv--)
            for (int j = 0; j < h->level_n_node[lv]; j++) {
                const int node = h->level_nodes[(size_t)lv * nl + j];
                const int k = rank_of(node);
                if (k == 0 || h->U[node]->nrow == 0) continue;
                const int s = h->mat_cluster[2 * node], m = rows_of(node);
                const long long off = (long long)P.PB.size();
                P.PB.resize(P.PB.size() + (size_t)m * k, 0.0);
                double *Pn = P.PB.data() + off;
                const H2E_dense_mat_p U = h->U[node];
                if (h->n_child[node] == 0) {
                    if (U->nrow != m) continue;
                    for (int r = 0; r < m; r++) memcpy(Pn + (size_t)r * k, U->data + (size_t)r * U->ld, sizeof(double) * k);
                } else {
                    int uoff = 0;
                    for (int c = 0; c < h->n_child[node]; c++) {
                        const int ch = h->children[(size_t)node * h->max_child + c];
                        const int kc = rank_of(ch);
                        if (kc > 0 && pb_off[ch] >= 0) {
                            const double *Pc = P.PB.data() + pb_off[ch];
                            const int cs = h->mat_cluster[2 * ch] - s, cm = rows_of(ch);
                            for (int row = 0; row < cm; row++)
                                for (int l = 0; l < kc; l++) {
                                    const double a = Pc[(size_t)row * kc + l];
                                    const double *urow = U->data + (size_t)(uoff + l) * U->ld;
                                    double *o = Pn + (size_t)(cs + row) * k;
                                    for (int t = 0; t < k; t++) o[t] += a * urow[t];
                                }
                        }
                        uoff += kc;
                    }
                }
                pb_off[node] = off;
            }
    }

    // basis-function set and orientation lists of a tree node's row range
    int list_of(int node) {
        if (node_list[node] >= 0) return node_list[node];
        const int s = h->mat_cluster[2 * node], m = rows_of(node);
        std::vector<int> bfs;
        bfs.reserve(2 * m);
        for (int r = 0; r < m; r++) { bfs.push_back(ra[s + r]); bfs.push_back(rb[s + r]); }
        std::sort(bfs.begin(), bfs.end());
        bfs.erase(std::unique(bfs.begin(), bfs.end()), bfs.end());
        const int nb = (int)bfs.size();
        std::vector<int> loc(h->num_bf, -1);
        for (int i = 0; i < nb; i++) loc[bfs[i]] = i;
        KNode kn;
        kn.bf_off = (int)P.bfidx.size();
        kn.nbf = nb;
        P.bfidx.insert(P.bfidx.end(), bfs.begin(), bfs.end());
        // byQ: ket orientations (p,q) of each row c; byM: bra orientations (m,n) of each row r
        std::vector<std::vector<int2>> byq(nb), bym(nb);
        for (int r = 0; r < m; r++) {
            const int x = s + r, a = ra[x], b = rb[x];
            byq[loc[b]].push_back(make_int2(r, a));
            bym[loc[a]].push_back(make_int2(r, loc[b]));
            if (sw[x]) {
                byq[loc[a]].push_back(make_int2(r, b));
                bym[loc[b]].push_back(make_int2(r, loc[a]));
            }
        }
        kn.q_ptr = (int)P.qptr.size();
        for (int i = 0; i < nb; i++) { P.qptr.push_back((int)P.qent.size()); P.qent.insert(P.qent.end(), byq[i].begin(), byq[i].end()); }
        P.qptr.push_back((int)P.qent.size());
        kn.m_ptr = (int)P.mptr.size();
        for (int i = 0; i < nb; i++) { P.mptr.push_back((int)P.ment.size()); P.ment.insert(P.ment.end(), bym[i].begin(), bym[i].end()); }
        P.mptr.push_back((int)P.ment.size());
        // byN: for each local n the (row, local m) entries, sorted by m, with the
        // first entry of every split of ssz functions m (at most KSPLIT, the
        // node's functions in even parts so that no split task is a sliver)
        std::vector<std::vector<int2>> byn(nb);
        for (int ml = 0; ml < nb; ml++)
            for (const int2 &e : bym[ml]) byn[e.y].push_back(make_int2(e.x, ml));
        kn.nsplit = (nb + KSPLIT - 1) / KSPLIT;
        kn.ssz = kn.nsplit > 0 ? (nb + kn.nsplit - 1) / kn.nsplit : KSPLIT;
        kn.s_ptr = (int)P.nspl.size();
        for (int nl = 0; nl < nb; nl++) {
            const int base = (int)P.nent.size();
            int k = 0;
            for (int sp = 0; sp <= kn.nsplit; sp++) {
                while (k < (int)byn[nl].size() && byn[nl][k].y < sp * kn.ssz) k++;
                P.nspl.push_back(base + k);
            }
            P.nent.insert(P.nent.end(), byn[nl].begin(), byn[nl].end());
        }
        P.nodes.push_back(kn);
        node_list[node] = (int)P.nodes.size() - 1;
        return node_list[node];
    }

    struct Core { bool lr; long long C, U, VT; int nrow, ncol, k; };
    Core store_core(H2E_dense_mat_p b) {
        Core c{false, -1, -1, -1, b->nrow, b->ncol, 0};
        const long long off = (long long)P.CB.size();
        if (b->ld >= 0) {
            P.CB.insert(P.CB.end(), b->data, b->data + (size_t)b->nrow * b->ncol);
            c.C = off;
        } else {
            c.lr = true;
            c.k = -b->ld;
            P.CB.insert(P.CB.end(), b->data, b->data + (size_t)(b->nrow + b->ncol) * c.k);
            c.U = off;
            c.VT = off + (long long)b->nrow * c.k;
        }
        return c;
    }
    long long salloc(long long n) {
        const long long o = P.nSC;
        P.nSC += (n + 31) & ~31LL;
        return o;
    }
    long long gemm(long long Aoff, int lda, int m, int k, long long Xoff, int ldx, int n) {
        KGemm g;
        g.Aoff = Aoff; g.lda = lda; g.m = m; g.k = k; g.Xoff = Xoff; g.ldx = ldx; g.n = n;
        g.Ooff = salloc((long long)m * n);
        g.ldo = n;
        P.gemms.push_back(g);
        P.flops += 2.0 * m * n * k;
        return g.Ooff;
    }

    // one term: rows of node n0 (basis side if bas0) x columns of node n1 (bas1), core b
    void add_term(int n0, bool bas0, int n1, bool bas1, H2E_dense_mat_p b, int diag) {
        if (!b || b->nrow == 0 || b->ncol == 0) return;
        const int m0 = rows_of(n0), m1 = rows_of(n1);
        const int k0 = bas0 ? rank_of(n0) : m0, k1 = bas1 ? rank_of(n1) : m1;
        if (b->nrow != k0 || b->ncol != k1) return;   // not part of A_H2
        if ((bas0 && pb_off[n0] < 0) || (bas1 && pb_off[n1] < 0)) return;
        const Core c = store_core(b);
        KTerm t;
        memset(&t, 0, sizeof(t));
        t.diag = diag;
        t.ni = list_of(n0);
        t.nj = list_of(n1);
        t.m0 = m0; t.m1 = m1;
        if (!bas0 && !bas1 && !c.lr) {            // dense near-field block: direct contraction
            t.direct = 1;
            t.kk = m1;
            t.Lsrc = KS_CB; t.Loff = c.C; t.Lld = m1;
            P.terms.push_back(t);
            return;
        }
        if (!c.lr) {
            if (bas0 && bas1) {        // L = P0 C, R = P1
                t.kk = k1;
                t.Lsrc = KS_SC; t.Loff = gemm(pb_off[n0], k0, m0, k0, c.C, k1, k1); t.Lld = k1;
                t.Rsrc = KS_PB; t.Roff = pb_off[n1]; t.Rrs = k1; t.Rcs = 1;
            } else if (bas0) {         // L = P0, R = C^T
                t.kk = k0;
                t.Lsrc = KS_PB; t.Loff = pb_off[n0]; t.Lld = k0;
                t.Rsrc = KS_CB; t.Roff = c.C; t.Rrs = 1; t.Rcs = m1;
            } else {                   // L = C, R = P1
                t.kk = k1;
                t.Lsrc = KS_CB; t.Loff = c.C; t.Lld = k1;
                t.Rsrc = KS_PB; t.Roff = pb_off[n1]; t.Rrs = k1; t.Rcs = 1;
            }
        } else {
            const int k = c.k;
            t.kk = k;
            if (bas0) { t.Lsrc = KS_SC; t.Loff = gemm(pb_off[n0], k0, m0, k0, c.U, k, k); t.Lld = k; }
            else { t.Lsrc = KS_CB; t.Loff = c.U; t.Lld = k; }
            if (bas1) { t.Rsrc = KS_SC; t.Roff = gemm(pb_off[n1], k1, m1, k1, c.VT, k, k); t.Rrs = k; t.Rcs = 1; }
            else { t.Rsrc = KS_CB; t.Roff = c.VT; t.Rrs = k; t.Rcs = 1; }
        }
        t.Goff = -1;   // placed when the terms are cut into chunks
        P.terms.push_back(t);
    }

    bool build() {
        if (h->num_bf <= 0 || !h->sp_shell_idx || !h->shell_bf_sidx) { set_error(ERR_ARG, "exchange needs shell-pair metadata"); return false; }
        if (h->BD_JIT) { set_error(ERR_UNSUPPORTED, "H2ERI_build_exchange does not support BD_JIT representations"); return false; }
        if (!pair_rows()) return false;
        expand_bases();
        node_list.assign(h->n_node, -1);
        const int nl = h->n_leaf_node;
        for (int i = 0; i < nl; i++) add_term(h->height_nodes[i], false, h->height_nodes[i], false, h->c_D_blks[i], 1);
        for (int i = 0; i < h->n_r_inadm_pair; i++)
            add_term(h->r_inadm_pairs[2 * i], false, h->r_inadm_pairs[2 * i + 1], false, h->c_D_blks[nl + i], 0);
        for (int i = 0; i < h->n_r_adm_pair; i++) {
            const int n0 = h->r_adm_pairs[2 * i], n1 = h->r_adm_pairs[2 * i + 1];
            const int l0 = h->node_level[n0], l1 = h->node_level[n1];
            add_term(n0, l0 >= l1, n1, l1 >= l0, h->c_B_blks[i], 0);
        }
        // launch task lists
        for (int g = 0; g < (int)P.gemms.size(); g++)
            for (int r0 = 0; r0 < P.gemms[g].m; r0 += 64)
                for (int c0 = 0; c0 < P.gemms[g].n; c0 += 64) P.gtasks.push_back(KGemmTask{g, r0, c0, 0});
        // chunks of consecutive terms whose G fits the budget (a larger term
        // alone); G offsets restart in every chunk
        const long long budget = std::max(1LL, (long long)(g_budget / 8.0));
        long long used = 0;
        bool open = false;
        for (int ti = 0; ti < (int)P.terms.size(); ti++) {
            KTerm &t = P.terms[ti];
            const KNode &ni = P.nodes[t.ni], &nj = P.nodes[t.nj];
            if (t.direct) {
                const long long ne = (long long)ni.nbf * nj.nbf;
                for (long long e0 = 0; e0 < ne; e0 += KWG) P.direct.push_back(KTask{ti, (int)e0, 0, 0});
                continue;
            }
            const long long gs = ((long long)t.kk * ni.nbf * gld(nj.nbf) + 31) & ~31LL;
            if (!open || (used > 0 && used + gs > budget)) {
                P.ket_chunk.push_back((int)P.ket.size());
                P.bra_chunk.push_back((int)P.bra.size());
                used = 0;
                open = true;
            }
            t.Goff = used;
            used += gs;
            P.nG = std::max(P.nG, used);
            for (int t0 = 0; t0 < t.kk; t0 += 16)
                for (int nl0 = 0; nl0 < ni.nbf; nl0 += 64)
                    for (int ql0 = 0; ql0 < nj.nbf; ql0 += KQC) P.ket.push_back(KTask{ti, nl0, t0, ql0});
            for (int ql0 = 0; ql0 < nj.nbf; ql0 += 32)
                for (int sp = 0; sp < ni.nsplit; sp++) P.bra.push_back(KTask{ti, ql0, sp, 0});
        }
        P.ket_chunk.push_back((int)P.ket.size());
        P.bra_chunk.push_back((int)P.bra.size());
        return validate();
    }

    // every access of every kernel checked on the host before any launch
    bool validate() const {
        const long long nPB = (long long)P.PB.size(), nCB = (long long)P.CB.size(), nSC = P.nSC;
        const int nbf = h->num_bf;
        auto size_of = [&](int src) { return src == KS_PB ? nPB : (src == KS_CB ? nCB : nSC); };
        for (const KGemm &g : P.gemms) {
            const bool ok = g.m > 0 && g.n > 0 && g.k > 0 && g.Aoff >= 0 && g.Aoff + (long long)(g.m - 1) * g.lda + g.k <= nPB &&
                            g.Xoff >= 0 && g.Xoff + (long long)(g.k - 1) * g.ldx + g.n <= nCB && g.Ooff >= 0 &&
                            g.Ooff + (long long)(g.m - 1) * g.ldo + g.n <= nSC && g.lda >= g.k && g.ldx >= g.n && g.ldo >= g.n;
            if (!ok) { set_error(ERR_ARG, "exchange plan check failed: gemm"); return false; }
        }
        for (const KNode &n : P.nodes) {
            bool ok = n.bf_off >= 0 && n.bf_off + n.nbf <= (long long)P.bfidx.size() && n.q_ptr + n.nbf < (long long)P.qptr.size() &&
                      n.m_ptr + n.nbf < (long long)P.mptr.size();
            for (int i = 0; ok && i < n.nbf; i++) ok = P.bfidx[n.bf_off + i] >= 0 && P.bfidx[n.bf_off + i] < nbf;
            if (!ok) { set_error(ERR_ARG, "exchange plan check failed: node lists"); return false; }
        }
        for (const KTerm &t : P.terms) {
            const KNode &ni = P.nodes[t.ni], &nj = P.nodes[t.nj];
            bool ok = t.kk > 0 && (long long)ni.nbf * nj.nbf < (1LL << 31);
            // entries: columns < m1, p < nbf; rows < m0, local n < |B(i)|
            for (int j = P.qptr[nj.q_ptr]; ok && j < P.qptr[nj.q_ptr + nj.nbf]; j++)
                ok = P.qent[j].x >= 0 && P.qent[j].x < t.m1 && P.qent[j].y >= 0 && P.qent[j].y < nbf;
            for (int j = P.mptr[ni.m_ptr]; ok && j < P.mptr[ni.m_ptr + ni.nbf]; j++)
                ok = P.ment[j].x >= 0 && P.ment[j].x < t.m0 && P.ment[j].y >= 0 && P.ment[j].y < ni.nbf;
            // byN splits: increasing, inside nent, entries of a split within its m range
            const int ns1 = ni.nsplit + 1;
            // (the splits cover every local m: nsplit * ssz >= nbf)
            ok = ok && ni.nsplit >= 1 && ni.ssz >= 1 && ni.ssz <= KSPLIT && (long long)ni.nsplit * ni.ssz >= ni.nbf &&
                 ni.s_ptr >= 0 && ni.s_ptr + (long long)ni.nbf * ns1 <= (long long)P.nspl.size();
            for (int nl = 0; ok && nl < ni.nbf; nl++)
                for (int sp = 0; ok && sp < ni.nsplit; sp++) {
                    const int eb = P.nspl[ni.s_ptr + nl * ns1 + sp], ee = P.nspl[ni.s_ptr + nl * ns1 + sp + 1];
                    ok = eb >= 0 && eb <= ee && ee <= (int)P.nent.size();
                    for (int e = eb; ok && e < ee; e++)
                        ok = P.nent[e].x >= 0 && P.nent[e].x < t.m0 && P.nent[e].y >= sp * ni.ssz &&
                             P.nent[e].y < std::min((sp + 1) * ni.ssz, ni.nbf);
                }
            ok = ok && t.Loff >= 0 && t.Loff + (long long)(t.m0 - 1) * t.Lld + t.kk <= size_of(t.Lsrc);
            if (!t.direct) {
                const long long last = t.Roff + (long long)(t.m1 - 1) * t.Rrs + (long long)(t.kk - 1) * t.Rcs;
                ok = ok && t.Roff >= 0 && last < size_of(t.Rsrc) && t.Goff >= 0 &&
                     t.Goff % 4 == 0 && t.Goff + (long long)t.kk * ni.nbf * gld(nj.nbf) <= P.nG;
            }
            if (!ok) { set_error(ERR_ARG, "exchange plan check failed: term"); return false; }
        }
        // chunks: the task ranges cover both lists in order, and all tasks of
        // a term fall in one chunk (the chunks share one G region)
        const size_t nc = P.ket_chunk.size();
        bool ok = nc >= 1 && P.bra_chunk.size() == nc && P.ket_chunk.front() == 0 && P.bra_chunk.front() == 0 &&
                  P.ket_chunk.back() == (int)P.ket.size() && P.bra_chunk.back() == (int)P.bra.size();
        std::vector<int> tchunk(P.terms.size(), -1);
        for (size_t c = 0; ok && c + 1 < nc; c++) {
            ok = P.ket_chunk[c] <= P.ket_chunk[c + 1] && P.bra_chunk[c] <= P.bra_chunk[c + 1];
            for (int side = 0; ok && side < 2; side++) {
                const std::vector<KTask> &L = side == 0 ? P.ket : P.bra;
                const std::vector<int> &B = side == 0 ? P.ket_chunk : P.bra_chunk;
                for (int i = B[c]; ok && i < B[c + 1]; i++) {
                    const int ti = L[i].term;
                    ok = ti >= 0 && ti < (int)P.terms.size() && !P.terms[ti].direct &&
                         (tchunk[ti] < 0 || tchunk[ti] == (int)c);
                    if (ok) tchunk[ti] = (int)c;
                }
            }
        }
        if (!ok) { set_error(ERR_ARG, "exchange plan check failed: chunks"); return false; }
        return true;
    }
};

}  // namespace

static KState *kstate(H2ERI_p h) { return (KState *)h->thread_Kmat_workbuf; }

void exchange_release(H2ERI_p h) {
    KState *k = kstate(h);
    if (!k) return;
    device_release_graph_kind(h, 2);   // a captured exchange graph refers to the buffers freed here
    kstate_release(k);
    delete k;
    h->thread_Kmat_workbuf = nullptr;
}

#define KHIP(expr)                                                                                           \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) {                                                                              \
            set_error(ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return false;                                                                                    \
        }                                                                                                    \
    } while (0)

static bool k_upload(H2ERI_p h) {
    if (!h->U || !h->c_B_blks || !h->c_D_blks) { set_error(ERR_ARG, "no H2 representation built or loaded"); return false; }
    if (!h->shell_bf_sidx || h->nshell <= 0) { set_error(ERR_ARG, "the exchange build needs shell metadata"); return false; }
    if (!h->thread_Kmat_workbuf) h->thread_Kmat_workbuf = (void **)new KState();
    KState *k = kstate(h);
    if (k->ready) return true;
    const double budget = k->g_budget;
    kstate_release(k);
    KBuilder b(h, k->P, budget);
    if (!b.build()) return false;
    const KPlan &P = k->P;
    const int nbf = h->num_bf;
    auto up = [&](void **dst, const void *src, size_t bytes) -> bool {
        if (bytes == 0) bytes = 64;
        KHIP(hipMalloc(dst, bytes));
        if (src) KHIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        k->resident += bytes;
        return true;
    };
    const size_t nn2 = (size_t)nbf * nbf;
    if (!up((void **)&k->PB, P.PB.empty() ? nullptr : P.PB.data(), P.PB.size() * 8) ||
        !up((void **)&k->CB, P.CB.empty() ? nullptr : P.CB.data(), P.CB.size() * 8) ||
        !up((void **)&k->SC, nullptr, (size_t)P.nSC * 8) ||
        !up((void **)&k->GS, nullptr, (size_t)P.nG * 8 * (P.ket_chunk.size() > 2 ? 2 : 1)) ||
        !up((void **)&k->K1, nullptr, nn2 * 8) ||
        !up((void **)&k->Kd, nullptr, nn2 * 8) || !up((void **)&k->D, nullptr, nn2 * 8) ||
        !up((void **)&k->K, nullptr, nn2 * 8) || !up((void **)&k->Ds, nullptr, nn2 * 8) ||
        !up((void **)&k->sbf, h->shell_bf_sidx, (size_t)(h->nshell + 1) * 4) ||
        !up((void **)&k->terms, P.terms.empty() ? nullptr : P.terms.data(), P.terms.size() * sizeof(KTerm)) ||
        !up((void **)&k->gemms, P.gemms.empty() ? nullptr : P.gemms.data(), P.gemms.size() * sizeof(KGemm)) ||
        !up((void **)&k->nodes, P.nodes.empty() ? nullptr : P.nodes.data(), P.nodes.size() * sizeof(KNode)) ||
        !up((void **)&k->bfidx, P.bfidx.data(), P.bfidx.size() * 4) || !up((void **)&k->qptr, P.qptr.data(), P.qptr.size() * 4) ||
        !up((void **)&k->mptr, P.mptr.data(), P.mptr.size() * 4) ||
        !up((void **)&k->nspl, P.nspl.data(), P.nspl.size() * 4) ||
        !up((void **)&k->nent, P.nent.empty() ? nullptr : P.nent.data(), P.nent.size() * sizeof(int2)) ||
        !up((void **)&k->qent, P.qent.empty() ? nullptr : P.qent.data(), P.qent.size() * sizeof(int2)) ||
        !up((void **)&k->ment, P.ment.empty() ? nullptr : P.ment.data(), P.ment.size() * sizeof(int2)) ||
        !up((void **)&k->ket, P.ket.empty() ? nullptr : P.ket.data(), P.ket.size() * sizeof(KTask)) ||
        !up((void **)&k->bra, P.bra.empty() ? nullptr : P.bra.data(), P.bra.size() * sizeof(KTask)) ||
        !up((void **)&k->direct, P.direct.empty() ? nullptr : P.direct.data(), P.direct.size() * sizeof(KTask)) ||
        !up((void **)&k->gtasks, P.gtasks.empty() ? nullptr : P.gtasks.data(), P.gtasks.size() * sizeof(KGemmTask)))
        return false;
    if (P.ket_chunk.size() > 2) {   // two or more chunks: odd ones run on a second stream
        KHIP(hipStreamCreateWithFlags(&k->side, hipStreamNonBlocking));
        KHIP(hipEventCreateWithFlags(&k->ev_fork, hipEventDisableTiming));
        KHIP(hipEventCreateWithFlags(&k->ev_join, hipEventDisableTiming));
    }
    // L / R products do not depend on D: computed once at upload
    if (!P.gtasks.empty()) {
        hipLaunchKernelGGL(kx_gemm, dim3((unsigned)P.gtasks.size()), dim3(KWG), 0, 0, k->gtasks, k->gemms, k->PB, k->CB, k->SC);
        KHIP(hipGetLastError());
        KHIP(hipDeviceSynchronize());
    }
    k->ready = true;
    return true;
}

static bool k_run(H2ERI_p h, const double *D_dev, double *K_dev, hipStream_t s) {
    if (!k_upload(h)) return false;
    KState *k = kstate(h);
    const KPlan &P = k->P;
    const int nbf = h->num_bf;
    const size_t nn2 = (size_t)nbf * nbf;
    const unsigned zb = (unsigned)std::min<size_t>((nn2 + 255) / 256, 2048);
    const long long nblk = (long long)h->nshell * h->nshell;
    hipLaunchKernelGGL(kx_dlist, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, s, k->sbf, h->nshell, D_dev, k->Ds, nbf);
    D_dev = k->Ds;
    hipLaunchKernelGGL(kx_zero, dim3(zb), dim3(256), 0, s, k->K1, nn2);
    hipLaunchKernelGGL(kx_zero, dim3(zb), dim3(256), 0, s, k->Kd, nn2);
    KHIP(hipGetLastError());
    // every chunk of terms: its ket writes G, its bra reads it back.  Chunk c
    // uses G region c % 2 on stream c % 2 (the caller's, or the second stream
    // forked from it after the zeroing), so chunk c + 1's ket runs beside
    // chunk c's bra and a region is reused only after its previous bra; both
    // streams add into K1 / Kd with atomics and join before kx_final.
    const int nchunk = (int)P.ket_chunk.size() - 1;
    const bool two = nchunk > 1;
    if (two) {
        KHIP(hipEventRecord(k->ev_fork, s));
        KHIP(hipStreamWaitEvent(k->side, k->ev_fork, 0));
    }
    if (!P.direct.empty())
        hipLaunchKernelGGL(kx_direct, dim3((unsigned)P.direct.size()), dim3(KWG), 0, s, k->direct, k->terms, k->nodes,
                           k->bfidx, k->qptr, k->qent, k->mptr, k->ment, k->CB, D_dev, k->K1, k->Kd, nbf);
    for (int c = 0; c < nchunk; c++) {
        hipStream_t q = (two && (c & 1)) ? k->side : s;
        double *G = k->GS + (two && (c & 1) ? P.nG : 0);
        const int kb = P.ket_chunk[c], ke = P.ket_chunk[c + 1], bb = P.bra_chunk[c], be = P.bra_chunk[c + 1];
        if (ke > kb)
            hipLaunchKernelGGL(kx_ket, dim3((unsigned)(ke - kb)), dim3(KWG), 0, q, k->ket + kb, k->terms, k->nodes,
                               k->bfidx, k->qptr, k->qent, k->PB, k->CB, k->SC, G, D_dev, nbf);
        if (be > bb)
            hipLaunchKernelGGL(kx_bra, dim3((unsigned)(be - bb)), dim3(KWG), 0, q, k->bra + bb, k->terms, k->nodes,
                               k->bfidx, k->nspl, k->nent, k->PB, k->CB, k->SC, G, k->K1, k->Kd, nbf);
    }
    KHIP(hipGetLastError());
    if (two) {
        KHIP(hipEventRecord(k->ev_join, k->side));
        KHIP(hipStreamWaitEvent(s, k->ev_join, 0));
    }
    hipLaunchKernelGGL(kx_final, dim3(zb), dim3(256), 0, s, k->K1, k->Kd, K_dev, nbf);
    KHIP(hipGetLastError());
    return true;
}

bool exchange_prepare(H2ERI_p h) { return k_upload(h); }

bool exchange_enqueue(H2ERI_p h, const double *D_dev, double *K_dev, void *stream) {
    return k_run(h, D_dev, K_dev, (hipStream_t)stream);
}

}  // namespace h2e

using namespace h2e;

extern "C" {

// Density screening list on the host (src/H2ERI_build_exchange.c:774-838):
// dlist_sidx[N]..dlist_sidx[N+1] lists, in increasing order, the shells P
// whose block (N,P) of D has some |D_np| > 1e-10 (:790).  The device build
// applies the same rule to its copy of D.
void H2ERI_exchange_dlist(H2ERI_p h, const double *den_mat) {
    if (!h || !den_mat) { set_error(ERR_ARG, "H2ERI_exchange_dlist: null argument"); return; }
    if (!h->shell_bf_sidx || h->nshell <= 0) { set_error(ERR_ARG, "H2ERI_exchange_dlist: needs shell-pair metadata (a molecule)"); return; }
    const int ns = h->nshell, nbf = h->num_bf;
    const int *sbf = h->shell_bf_sidx;
    std::vector<std::vector<int>> rows(ns);
    #pragma omp parallel for schedule(dynamic)
    for (int n = 0; n < ns; n++)
        for (int p = 0; p < ns; p++) {
            bool keep = false;
            for (int i = sbf[n]; i < sbf[n + 1] && !keep; i++)
                for (int j = sbf[p]; j < sbf[p + 1]; j++)
                    if (std::fabs(den_mat[(size_t)i * nbf + j]) > 1e-10) { keep = true; break; }
            if (keep) rows[n].push_back(p);
        }
    free(h->dlist_sidx);
    free(h->dlist);
    h->dlist_sidx = (int *)malloc(sizeof(int) * (ns + 1));
    size_t len = 0;
    for (int n = 0; n < ns; n++) len += rows[n].size();
    h->dlist = (int *)malloc(sizeof(int) * (len > 0 ? len : 1));
    if (!h->dlist_sidx || !h->dlist) { set_error(ERR_ALLOC, "H2ERI_exchange_dlist: allocation failed"); return; }
    h->dlist_sidx[0] = 0;
    for (int n = 0; n < ns; n++) {
        std::copy(rows[n].begin(), rows[n].end(), h->dlist + h->dlist_sidx[n]);
        h->dlist_sidx[n + 1] = h->dlist_sidx[n] + (int)rows[n].size();
    }
}

void H2ERI_build_exchange(H2ERI_p h, const double *den_mat, double *K_mat) {
    if (!h || !den_mat || !K_mat) { set_error(ERR_ARG, "H2ERI_build_exchange: null argument"); return; }
    H2ERI_exchange_dlist(h, den_mat);   // the reference's side effect on the handle
    if (has_error()) return;
    if (!k_upload(h)) return;
    KState *k = kstate(h);
    const size_t nn2 = (size_t)h->num_bf * h->num_bf;
    if (hipMemcpy(k->D, den_mat, nn2 * 8, hipMemcpyHostToDevice) != hipSuccess) { set_error(ERR_DEVICE, "copy D"); return; }
    if (!k_run(h, k->D, k->K, nullptr)) return;
    if (hipMemcpy(K_mat, k->K, nn2 * 8, hipMemcpyDeviceToHost) != hipSuccess) { set_error(ERR_DEVICE, "copy K"); return; }
    h->n_matvec++;
}

void H2ERI_build_exchange_dev(H2ERI_p h, const double *den_dev, double *K_dev, void *stream) {
    if (!h || !den_dev || !K_dev) { set_error(ERR_ARG, "H2ERI_build_exchange_dev: null argument"); return; }
    k_run(h, den_dev, K_dev, (hipStream_t)stream);
}

// Builds the exchange plan on the device now (otherwise done by the first
// call); the reference allocates its per-thread K work buffers here.
void H2ERI_exchange_workbuf_init(H2ERI_p h) {
    if (h) k_upload(h);
}

void H2ERI_exchange_workbuf_free(H2ERI_p h) {
    if (h) exchange_release(h);
}

// Builds the exchange plan on the host only (no HIP call) and checks every
// access of its kernels; out4 as H2ERI_exchange_stats with out4[2] the bytes
// the plan would make resident.  Returns the number of terms, -1 on failure.
int H2ERI_exchange_plan_check(H2ERI_p h, double *out4) {
    if (!h || !h->U || !h->c_B_blks || !h->c_D_blks) { set_error(ERR_ARG, "exchange_plan_check: nothing built"); return -1; }
    KPlan P;
    KBuilder b(h, P, kstate(h) ? kstate(h)->g_budget : KG_BUDGET_DEFAULT);
    if (!b.build()) return -1;
    if (out4) {
        const double nn2 = (double)h->num_bf * h->num_bf;
        out4[0] = P.flops;
        out4[1] = (double)P.terms.size();
        const double nG = (double)P.nG * (P.ket_chunk.size() > 2 ? 2 : 1);   // two regions with two or more chunks
        out4[2] = 8.0 * ((double)P.PB.size() + (double)P.CB.size() + (double)P.nSC + nG + 5.0 * nn2) +
                  4.0 * ((double)P.bfidx.size() + P.qptr.size() + P.mptr.size()) + 8.0 * ((double)P.qent.size() + P.ment.size());
        out4[3] = (double)P.nSC + nG;
    }
    return (int)P.terms.size();
}

// Work figures of the exchange plan: out4 = {GEMM flops at upload, terms,
// device bytes resident, scratch doubles}.
void H2ERI_exchange_stats(H2ERI_p h, double *out4) {
    if (!h || !out4 || !k_upload(h)) return;
    KState *k = kstate(h);
    out4[0] = k->P.flops;
    out4[1] = (double)k->P.terms.size();
    out4[2] = (double)k->resident;
    out4[3] = (double)(k->P.nSC + k->P.nG * (k->P.ket_chunk.size() > 2 ? 2 : 1));
}

// Bound of the exchange build's G scratch in bytes (additive, tuning): the
// block terms run in chunks whose intermediate G fits it, each chunk's ket
// then its bra, in one scratch region reused by every chunk.  Releases the
// exchange plan (rebuilt by the next K build).
void H2ERI_set_exchange_scratch(H2ERI_p h, const double bytes) {
    if (!h || !(bytes >= 1024.0)) { set_error(ERR_ARG, "H2ERI_set_exchange_scratch: at least 1024 bytes"); return; }
    if (!h->thread_Kmat_workbuf) h->thread_Kmat_workbuf = (void **)new KState();
    KState *k = kstate(h);
    if (k->g_budget == bytes) return;
    device_release_graph_kind(h, 2);
    kstate_release(k);
    k->g_budget = bytes;
}

}  // extern "C"
