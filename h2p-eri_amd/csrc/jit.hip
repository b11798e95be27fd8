// On-the-fly B and D blocks (a BD_JIT = 1 handle) on the GPU.
//
// The reference's BD_JIT mode (src/H2ERI_matvec.c:411-568 intermediate
// products, :761-879 near field, :911 and :928 the switch) stores U only and
// evaluates every coupling block B_ij and near-field block D_ij from the
// shell pairs at each matvec, through H2ERI_calc_ERI_pairs_to_mat
// (src/CMS.c:387-451).  Here the same blocks are evaluated on the device:
//
//  * Row tasks.  Every output row has one writer: a task owns one shell pair
//    a of one node n on one side (the node's full leaf rows, or its skeleton
//    rows, J_row) and walks every partner block of that side, evaluating each
//    shell quartet (a|b) and contracting it with the partner's input at once.
//    Side rules of the reference's three coupling cases
//    (src/H2ERI_matvec.c:455-545): node n is on its skeleton side against an
//    admissible partner m when level(n) >= level(m), else on its full side
//    (n is then a leaf).  Full rows are x-indexed and receive the near field
//    (the leaf itself and its inadmissible partners, both orientations) and
//    the coupling products of case (2)/(3) leaves; skeleton rows are y1 of
//    the node.  A block (n, m) is so evaluated once from each side, against
//    the reference's once with a bi-GEMV: twice the integrals, in exchange
//    for no atomics and a fixed summation order (bit-reproducible results).
//  * The integrals: McMurchie-Davidson (Helgaker, Jorgensen & Olsen,
//    Molecular Electronic-Structure Theory, eqs. 9.5.x and 9.9.33),
//    (ab|cd) = sum_{ij} 2 pi^{5/2} / (p q sqrt(p+q)) K_i K_j
//              sum_{tuv} E^{ab}_{tuv} sum_{tau nu phi} (-1)^{tau+nu+phi}
//              E^{cd}_{tau nu phi} R_{t+tau, u+nu, v+phi}(alpha, PQ)
//    with Hermite coefficients E from their two-term recurrences, Hermite
//    Coulomb integrals R from R^{(n)}_{000} = (-2 alpha)^n F_n(alpha |PQ|^2)
//    and R^{(n)}_{t+1,u,v} = t R^{(n+1)}_{t-1,u,v} + X_PQ R^{(n+1)}_{t,u,v},
//    and the Boys function of the host engine (the same table).  One wave
//    evaluates one quartet, or a few kets of one class against its bra pair
//    (jit_task): lanes over primitive quartets for the E and R tables (in
//    LDS), then lanes over the Cartesian components, each summing its
//    component over the primitive quartets in a fixed order.  Shells up to
//    d (cc-pVDZ); the primitive pairs, centres and conventions are the host
//    engine's (first shell of a pair = N, row index n * nbf(M) + m,
//    lexicographic Cartesians).
//  * Sweeps: the up sweep (y0 = U^T x, children first) and the down sweep
//    (parents first) are one launch per tree level over the nodes of that
//    level, with U in FP64.
//
// Sequence of one matvec: zero y1 and the row outputs, up sweep, the row
// tasks (they need y0), down sweep (it needs y1), y = down-sweep rows + row
// task rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>
#include <thread>
#include <utility>
#include <vector>

#include "h2eri_internal.h"

namespace h2e {

static const int JIT_LMAX = 2;     // shell angular momentum (d)
static const int JIT_LDS = 2048;   // LDS doubles of a task with a bra or ket pair of la + lb >= 3
static const int JIT_LDS_S = 1280; // LDS doubles of a task whose pairs all have la + lb <= 2
static const int JIT_NAB = 36;     // Cartesian rows of a shell pair, at most
static const int JIT_KMAX = 4;     // kets per pass, at most (8 and 16: 1% slower)
static const int JIT_PASSQ = 64;   // primitive quartets of a multi-ket pass, at most

// Everything a row task reads or writes; the pointers are device or, in the
// emulator, host arrays.
struct JitView {
    int N;
    const int *pinfo;     // per shell pair: la, lb, primitive pairs, first primitive
    const double *pgeo;   // per shell pair: A (3), B (3)
    const double *prim;   // per primitive pair: p, P (3), K, 1/(2p)
    const double *boys;   // Boys table [grid][bcol]
    int bgrid, bcol;
    double bstep, btmax, inv_bstep;
    const int *kl_pair, *kl_moff;     // per partner-list entry: shell pair, its input map
    const int *kmap;                  // input index per column (-1: not used)
    const int *t_pair, *t_e0, *t_e1;  // per task: bra pair, its range of gent
    const int *gent;                  // partner list entries of the tasks
    double *part;                     // per task: the partial rows (JIT_NAB doubles)
    // output rows of one bra pair on one side = the sum of its tasks' partials
    const int *r_pair, *r_omap, *r_t0, *r_t1, *r_tasks;
    const int *omap;                  // output index per row (-1: not written)
    const double *X, *Y0;             // inputs: index < N in X, else Y0[index - N]
    double *YJ, *Y1;                  // outputs: index < N in YJ, else Y1[index - N]
};

__host__ __device__ inline void jit_sync() {
#ifdef __HIP_DEVICE_COMPILE__
    __syncthreads();
#endif
}

__host__ __device__ inline int jit_ncart(int L) { return (L + 1) * (L + 2) / 2; }
__host__ __device__ constexpr int jit_nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
// index of the Hermite (or Cartesian) triple (t, u, v) among all triples of
// total order <= t+u+v, lexicographic within an order
__host__ __device__ constexpr int jit_herm(int t, int u, int v) {
    const int k = t + u + v, a = k - t;
    return k * (k + 1) * (k + 2) / 6 + a * (a + 1) / 2 + v;
}
// exponents of the k-th lexicographic Cartesian component of order L
__host__ __device__ constexpr void jit_cart(int L, int k, int &lx, int &ly, int &lz) {
    int a = 0;
    while ((a + 1) * (a + 2) / 2 <= k) a++;
    lx = L - a;
    ly = a - (k - a * (a + 1) / 2);
    lz = L - lx - ly;
}

// Index tables of the Hermite recursion, fixed at compile time: for each
// Hermite triple h = 1..nherm(8)-1 in jit_herm order, the step that builds
// it (R^{(n)}_h = X_dir R^{(n+1)}_{h1} + c R^{(n+1)}_{h2}); jit_herm of every
// triple of order <= 8 (-1 above); 1/k! and 1/(2m-1).
// Also: for each pair class (la, lb) and Cartesian pair index ab, the
// offsets of its E rows in the pair's table (x, y, z) and its orders per
// direction (packed 4 bits each); and magic[d] = ceil(2^32 / d) (2^32 for
// d = 1, so 64 bits), so that x / d = (x magic[d]) >> 32 exactly for
// x d < 2^32.
struct JitTabs {
    int rstep[165][4];   // h1, h2 (-1: none), c, direction
    short herm[9][9][9];
    double inv_fact[8];
    double inv_odd[10];
    short eoff[100][4];  // pair classes in (la, lb) order, each nab rows
    short pcoff[9];      // first eoff row of pair class la * 3 + lb
    unsigned long long magic[1297];
    constexpr JitTabs() : rstep{}, herm{}, inv_fact{}, inv_odd{}, eoff{}, pcoff{}, magic{} {
        int row = 0;
        for (int la = 0; la <= 2; la++)
            for (int lb = 0; lb <= 2; lb++) {
                pcoff[la * 3 + lb] = (short)row;
                const int L = la + lb, esd = (la + 1) * (lb + 1) * (L + 1);
                const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
                for (int ia = 0; ia < na; ia++)
                    for (int ib = 0; ib < nb; ib++, row++) {
                        int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
                        jit_cart(la, ia, ax, ay, az);
                        jit_cart(lb, ib, bx, by, bz);
                        eoff[row][0] = (short)((ax * (lb + 1) + bx) * (L + 1));
                        eoff[row][1] = (short)((ay * (lb + 1) + by) * (L + 1) + esd);
                        eoff[row][2] = (short)((az * (lb + 1) + bz) * (L + 1) + 2 * esd);
                        eoff[row][3] = (short)((ax + bx) | ((ay + by) << 4) | ((az + bz) << 8));
                    }
            }
        for (int d = 1; d < 1297; d++) magic[d] = (0x100000000ULL + (unsigned long long)d - 1) / (unsigned long long)d;
        for (int t = 0; t < 9; t++)
            for (int u = 0; u < 9; u++)
                for (int w = 0; w < 9; w++) herm[t][u][w] = (short)(t + u + w <= 8 ? jit_herm(t, u, w) : -1);
        for (int k = 1; k <= 8; k++)
            for (int t = k; t >= 0; t--)
                for (int u = k - t; u >= 0; u--) {
                    const int w = k - t - u, h = jit_herm(t, u, w);
                    int h1 = 0, h2 = -1, c = 0, d = 0;
                    if (t > 0) { h1 = jit_herm(t - 1, u, w); c = t - 1; d = 0; if (c > 0) h2 = jit_herm(t - 2, u, w); }
                    else if (u > 0) { h1 = jit_herm(t, u - 1, w); c = u - 1; d = 1; if (c > 0) h2 = jit_herm(t, u - 2, w); }
                    else { h1 = jit_herm(t, u, w - 1); c = w - 1; d = 2; if (c > 0) h2 = jit_herm(t, u, w - 2); }
                    rstep[h][0] = h1; rstep[h][1] = h2; rstep[h][2] = c; rstep[h][3] = d;
                }
        double f = 1.0;
        for (int k = 0; k < 8; k++) { inv_fact[k] = 1.0 / f; f *= (double)(k + 1); }
        inv_odd[0] = 0.0;
        for (int m = 1; m < 10; m++) inv_odd[m] = 1.0 / (double)(2 * m - 1);
    }
};
__constant__ JitTabs jit_tabs_dev = JitTabs();
static constexpr JitTabs jit_tabs_host = JitTabs();
__host__ __device__ inline const JitTabs &jit_tabs() {
#ifdef __HIP_DEVICE_COMPILE__
    return jit_tabs_dev;
#else
    return jit_tabs_host;
#endif
}

// F_m(T) for m = 0..mmax into out[m * stride]: the host engine's method
// (Taylor expansion of the tabulated grid for T <= tmax, downward
// recursion; erf and upward recursion above).
__host__ __device__ inline void jit_boys(const JitView &v, int mmax, double T, double *out, int stride) {
    if (T <= v.btmax) {
        const int gi = (int)(T * v.inv_bstep + 0.5);
        const double dT = gi * v.bstep - T;
        const double *row = v.boys + (size_t)gi * v.bcol;
        const JitTabs &tb = jit_tabs();
        double f = 0.0, pw = 1.0;
        for (int k = 0; k < 8; k++) {
            f += row[mmax + k] * pw * tb.inv_fact[k];
            pw *= dT;
        }
        out[mmax * stride] = f;
        const double eT = exp(-T);
        for (int m = mmax; m > 0; m--) out[(m - 1) * stride] = (2.0 * T * out[m * stride] + eT) * tb.inv_odd[m];
    } else {
        const double eT = exp(-T), inv2T = 0.5 / T;
        out[0] = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
        for (int m = 0; m < mmax; m++) out[(m + 1) * stride] = ((2 * m + 1) * out[m * stride] - eT) * inv2T;
    }
}

// Components of the Hermite triple of index h (inverse of jit_herm).
__host__ __device__ constexpr int jit_hcomp(int h, int which) {
    for (int k = 0; k <= 8; k++)
        for (int t = k; t >= 0; t--)
            for (int u = k - t; u >= 0; u--)
                if (jit_herm(t, u, k - t - u) == h) return which == 0 ? t : (which == 1 ? u : k - t - u);
    return 0;
}
template <int H1, int H2>
struct JitSum {   // index of the triple sum and the sign of the ket triple
    static constexpr int t1 = jit_hcomp(H1, 0), u1 = jit_hcomp(H1, 1), w1 = jit_hcomp(H1, 2);
    static constexpr int t2 = jit_hcomp(H2, 0), u2 = jit_hcomp(H2, 1), w2 = jit_hcomp(H2, 2);
    static constexpr int idx = jit_herm(t1 + t2, u1 + u2, w1 + w2);
    static constexpr bool neg = ((t2 + u2 + w2) & 1) != 0;
};
template <int H1, int... H2>
__host__ __device__ inline double jit_inner(const double *R, const double *ec, std::integer_sequence<int, H2...>) {
    return (0.0 + ... + (ec[H2] * R[JitSum<H1, H2>::idx]));
}
template <int... H2>
__host__ __device__ inline void jit_ket_e(const double *Eb, int px, int py, int pz, double *ec,
                                          std::integer_sequence<int, H2...>) {
    ((ec[H2] = (JitSum<0, H2>::neg ? -1.0 : 1.0) * Eb[px + JitSum<0, H2>::t2] * Eb[py + JitSum<0, H2>::u2] *
               Eb[pz + JitSum<0, H2>::w2]),
     ...);
}
template <int LCD, int... H1>
__host__ __device__ inline double jit_outer(const double *Ea, int ox, int oy, int oz, const double *R, const double *ec,
                                            std::integer_sequence<int, H1...>) {
    return (0.0 + ... +
            (Ea[ox + JitSum<H1, 0>::t1] * Ea[oy + JitSum<H1, 0>::u1] * Ea[oz + JitSum<H1, 0>::w1] *
             jit_inner<H1>(R, ec, std::make_integer_sequence<int, jit_nherm(LCD)>{})));
}
// One component on one primitive quartet: sum over the bra Hermite triples
// of total order <= LAB and the ket triples of order <= LCD, every R index a
// compile-time constant.  The E rows hold zeros beyond each component's own
// order, so the triples outside a component's box add nothing.
template <int LAB, int LCD>
__host__ __device__ inline double jit_comp(const double *Ea, int ox, int oy, int oz, const double *Eb, int px, int py,
                                           int pz, const double *R) {
    double ec[jit_nherm(LCD)];
    jit_ket_e(Eb, px, py, pz, ec, std::make_integer_sequence<int, jit_nherm(LCD)>{});
    return jit_outer<LCD>(Ea, ox, oy, oz, R, ec, std::make_integer_sequence<int, jit_nherm(LAB)>{});
}
template <int LAB, int TMAX>
__host__ __device__ inline double jit_comp_l(int Lcd, const double *Ea, int ox, int oy, int oz, const double *Eb, int px,
                                             int py, int pz, const double *R) {
    if constexpr (TMAX >= 3 && LAB <= 2)
        if (Lcd == 3) return jit_comp<LAB, 3>(Ea, ox, oy, oz, Eb, px, py, pz, R);
    switch (Lcd) {
    case 0: return jit_comp<LAB, 0>(Ea, ox, oy, oz, Eb, px, py, pz, R);
    case 1: return jit_comp<LAB, 1>(Ea, ox, oy, oz, Eb, px, py, pz, R);
    default: return jit_comp<LAB, 2>(Ea, ox, oy, oz, Eb, px, py, pz, R);
    }
}
// The same sum for the larger classes (either side of total order 3 or 4):
// each component's own box of Hermite terms, lane-dependent loop bounds and
// indices through the LDS copy hm of jit_herm.
__host__ __device__ inline double jit_comp_box(const double *Ea, int ox, int oy, int oz, int ex, int ey, int ez,
                                               const double *Eb, int px, int py, int pz, int fx, int fy, int fz,
                                               const double *R, const short *hm) {
    double val = 0.0;
    for (int t = 0; t <= ex; t++)
        for (int u = 0; u <= ey; u++)
            for (int w = 0; w <= ez; w++) {
                const double eab = Ea[ox + t] * Ea[oy + u] * Ea[oz + w];
                double inner = 0.0;
                for (int tt = 0; tt <= fx; tt++)
                    for (int uu = 0; uu <= fy; uu++)
                        for (int ww = 0; ww <= fz; ww++) {
                            const double ecd = Eb[px + tt] * Eb[py + uu] * Eb[pz + ww];
                            const double r = R[hm[((t + tt) * 9 + (u + uu)) * 9 + (w + ww)]];
                            inner += ((tt + uu + ww) & 1) ? -ecd * r : ecd * r;
                        }
                val += eab * inner;
            }
    return val;
}

// Hermite expansion coefficients of one primitive pair along one direction,
// for a fixed (LA, LB), unrolled: E[(i * (LB + 1) + j) * (LA + LB + 1) + t],
// i <= LA, j <= LB, every t written (zero for t > i + j, so the rows can be
// summed to the class's full order).  E^{00}_0 = 1: the exponential factor
// is in the pair prefactor K.  Two-term recurrences
// E^{i+1,j}_t = E^{ij}_{t-1} / (2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}, and
// the same in j with X_PB.
template <int LA, int LB>
__host__ __device__ inline void jit_hermite_e_t(double xpa, double xpb, double i2p, double *E) {
    constexpr int NT = LA + LB + 1;
    double e[LA + 1][LB + 1][NT];
#pragma unroll
    for (int i = 0; i <= LA; i++)
#pragma unroll
        for (int j = 0; j <= LB; j++)
#pragma unroll
            for (int t = 0; t < NT; t++) e[i][j][t] = 0.0;
    e[0][0][0] = 1.0;
#pragma unroll
    for (int i = 0; i < LA; i++)
#pragma unroll
        for (int t = 0; t <= i + 1; t++)
            e[i + 1][0][t] = (t > 0 ? i2p * e[i][0][t - 1] : 0.0) + (t <= i ? xpa * e[i][0][t] : 0.0) +
                             (t + 1 <= i ? (t + 1) * e[i][0][t + 1] : 0.0);
#pragma unroll
    for (int i = 0; i <= LA; i++)
#pragma unroll
        for (int j = 0; j < LB; j++)
#pragma unroll
            for (int t = 0; t <= i + j + 1; t++)
                e[i][j + 1][t] = (t > 0 ? i2p * e[i][j][t - 1] : 0.0) + (t <= i + j ? xpb * e[i][j][t] : 0.0) +
                                 (t + 1 <= i + j ? (t + 1) * e[i][j][t + 1] : 0.0);
#pragma unroll
    for (int i = 0; i <= LA; i++)
#pragma unroll
        for (int j = 0; j <= LB; j++)
#pragma unroll
            for (int t = 0; t < NT; t++) E[(i * (LB + 1) + j) * NT + t] = e[i][j][t];
}
__host__ __device__ inline void jit_hermite_e_any(int la, int lb, double xpa, double xpb, double i2p, double *E) {
    switch (la * 3 + lb) {
    case 0: jit_hermite_e_t<0, 0>(xpa, xpb, i2p, E); break;
    case 1: jit_hermite_e_t<0, 1>(xpa, xpb, i2p, E); break;
    case 2: jit_hermite_e_t<0, 2>(xpa, xpb, i2p, E); break;
    case 3: jit_hermite_e_t<1, 0>(xpa, xpb, i2p, E); break;
    case 4: jit_hermite_e_t<1, 1>(xpa, xpb, i2p, E); break;
    case 5: jit_hermite_e_t<1, 2>(xpa, xpb, i2p, E); break;
    case 6: jit_hermite_e_t<2, 0>(xpa, xpb, i2p, E); break;
    case 7: jit_hermite_e_t<2, 1>(xpa, xpb, i2p, E); break;
    default: jit_hermite_e_t<2, 2>(xpa, xpb, i2p, E); break;
    }
}

// The in-place Hermite recursion for a fixed total order LT, in registers:
// for n = LT-1 .. 0 the entries of order <= LT-n in decreasing index (so
// the orders an entry reads still hold R^{(n+1)}), then R^{(n)}_{000};
// R^{(0)} written out at the end.
template <int H>
struct JitStep {   // R_h = X_dir R_{h1} + c R_{h2}
    static constexpr int t = jit_hcomp(H, 0), u = jit_hcomp(H, 1), w = jit_hcomp(H, 2);
    static constexpr int dir = t > 0 ? 0 : (u > 0 ? 1 : 2);
    static constexpr int c = dir == 0 ? t - 1 : (dir == 1 ? u - 1 : w - 1);
    static constexpr int h1 = dir == 0 ? jit_herm(t - 1, u, w) : (dir == 1 ? jit_herm(t, u - 1, w) : jit_herm(t, u, w - 1));
    static constexpr int h2 = c <= 0 ? -1 : (dir == 0 ? jit_herm(t - 2, u, w) : (dir == 1 ? jit_herm(t, u - 2, w) : jit_herm(t, u, w - 2)));
};
template <int H>
__host__ __device__ inline void jit_rdesc(double *r, const double *xyz) {
    if constexpr (H >= 1) {
        using S = JitStep<H>;
        if constexpr (S::h2 >= 0) r[H] = xyz[S::dir] * r[S::h1] + (double)S::c * r[S::h2];
        else r[H] = xyz[S::dir] * r[S::h1];
        jit_rdesc<H - 1>(r, xyz);
    }
}
template <int LT, int N>
__host__ __device__ inline void jit_rlevel(double *r, const double *xyz, const double *Fs) {
    if constexpr (N >= 0) {
        jit_rdesc<jit_nherm(LT - N) - 1>(r, xyz);
        r[0] = Fs[N];
        jit_rlevel<LT, N - 1>(r, xyz, Fs);
    }
}
template <int LT>
__host__ __device__ inline void jit_rrec(double X, double Y, double Z, const double *Fs, double *R) {
    double r[jit_nherm(LT)];
    const double xyz[3] = {X, Y, Z};
    r[0] = Fs[LT];
    jit_rlevel<LT, LT - 1>(r, xyz, Fs);
#pragma unroll
    for (int h = 0; h < jit_nherm(LT); h++) R[h] = r[h];
}

// x / d for x d < 2^32, with m = magic[d].
__host__ __device__ inline int jit_div(int x, unsigned long long m) {
    return (int)(((unsigned long long)(unsigned)x * m) >> 32);
}

// Input value of index k (k < 0: none).
__host__ __device__ inline double jit_in(const JitView &v, int k) {
    return k < 0 ? 0.0 : (k < v.N ? v.X[k] : v.Y0[k - v.N]);
}

// One row task on NL lanes (64 on the device, 1 in the emulator), lds of
// LDSN doubles (enough for every ket class against the task's bra class):
// bra pair a against the partner entries gent[t_e0, t_e1), its partial rows
// into part[task].
// Per quartet: the partial sums of each component over the primitive
// quartets live in nsl slices (NL / ncomp of them when a quartet has fewer
// components than lanes, so that the lanes left over share the primitive
// loop), added in slice order at the end of the quartet.
// E tables of both pairs and the Hermite integrals R of one primitive
// quartet (bra primitive pair pi, ket primitive pair pj) into Ea (esa
// doubles), Eb = Ea + esa (esb) and R = Eb + esb (nH(Lt) + Lt + 1).
__host__ __device__ inline void jit_prim(const JitView &v, int la, int lb, int lc, int ld, int esa_d, int esb_d,
                                         const double *pi, const double *pj, const double *Ag, const double *Cg,
                                         double *Ea) {
    const int Lt = la + lb + lc + ld, nH = jit_nherm(Lt);
    double *Eb = Ea + 3 * esa_d, *R = Eb + 3 * esb_d;
    const double p = pi[0], qq = pj[0];
    for (int d = 0; d < 3; d++) {
        jit_hermite_e_any(la, lb, pi[1 + d] - Ag[d], pi[1 + d] - Ag[3 + d], pi[5], Ea + d * esa_d);
        jit_hermite_e_any(lc, ld, pj[1 + d] - Cg[d], pj[1 + d] - Cg[3 + d], pj[5], Eb + d * esb_d);
    }
    const double pq = p + qq, inv_pq = 1.0 / pq, alpha = p * qq * inv_pq;
    const double X = pi[1] - pj[1], Y = pi[2] - pj[2], Z = pi[3] - pj[3];
    const double T = alpha * (X * X + Y * Y + Z * Z);
    // 2 pi^{5/2} / (p q sqrt(p + q)) K_i K_j, with 1/p = 2 pi[5]
    const double pre = 2.0 * M_PI * M_PI * sqrt(M_PI) * (4.0 * pi[5] * pj[5]) * sqrt(inv_pq) * pi[4] * pj[4];
    double *Fs = R + nH;   // pre (-2 alpha)^n F_n
    jit_boys(v, Lt, T, Fs, 1);
    double s = pre;
    for (int n = 0; n <= Lt; n++) { Fs[n] *= s; s *= -2.0 * alpha; }
    // R^{(n)} from R^{(n+1)} in place, n = Lt-1 .. 0: orders from the
    // highest down, so that the orders k-1 and k-2 an order-k entry reads
    // still hold R^{(n+1)}
    switch (Lt) {
    case 0: R[0] = Fs[0]; break;
    case 1: jit_rrec<1>(X, Y, Z, Fs, R); break;
    case 2: jit_rrec<2>(X, Y, Z, Fs, R); break;
    case 3: jit_rrec<3>(X, Y, Z, Fs, R); break;
    case 4: jit_rrec<4>(X, Y, Z, Fs, R); break;
    default: break;
    }
    const JitTabs &tb = jit_tabs();
    if (Lt > 4) R[0] = Fs[Lt];
    for (int n = Lt > 4 ? Lt - 1 : -1; n >= 0; n--) {
        for (int k = Lt - n; k >= 1; k--) {
            const int hb = k * (k + 1) * (k + 2) / 6, he = (k + 1) * (k + 2) * (k + 3) / 6;
            for (int h = hb; h < he; h++) {
                const int h1 = tb.rstep[h][0], h2 = tb.rstep[h][1], d = tb.rstep[h][3];
                const double xx = d == 0 ? X : (d == 1 ? Y : Z);
                R[h] = h2 >= 0 ? xx * R[h1] + (double)tb.rstep[h][2] * R[h2] : xx * R[h1];
            }
        }
        R[0] = Fs[n];
    }
}

// One component (E-row offsets and orders ea, ec) on one primitive quartet.
template <int TMAX>
__host__ __device__ inline double jit_cval(int Lab, int Lcd, const double *Ea, const double *Eb, const double *R,
                                           const short *ea, const short *ec, const short *hm) {
    const int ox = ea[0], oy = ea[1], oz = ea[2], px = ec[0], py = ec[1], pz = ec[2];
    if (Lab > TMAX || Lcd > TMAX || Lab + Lcd > 5)
        return jit_comp_box(Ea, ox, oy, oz, ea[3] & 15, (ea[3] >> 4) & 15, ea[3] >> 8, Eb, px, py, pz, ec[3] & 15,
                            (ec[3] >> 4) & 15, ec[3] >> 8, R, hm);
    if (Lab == 0) return jit_comp_l<0, TMAX>(Lcd, Ea, ox, oy, oz, Eb, px, py, pz, R);
    if (Lab == 1) return jit_comp_l<1, TMAX>(Lcd, Ea, ox, oy, oz, Eb, px, py, pz, R);
    if (Lab == 2) return jit_comp_l<2, TMAX>(Lcd, Ea, ox, oy, oz, Eb, px, py, pz, R);
    if constexpr (TMAX >= 3) return jit_comp_l<3, TMAX>(Lcd, Ea, ox, oy, oz, Eb, px, py, pz, R);
    return 0.0;
}

// One row task on NL lanes (64 on the device, 1 in the emulator), lds of
// LDSN doubles (enough for every ket class against the task's bra class):
// bra pair a against the partner entries gent[t_e0, t_e1), its partial rows
// into part[task].
//
// Entries are taken in passes.  A pass is either one ket (its primitive
// quartets in batches of up to NL, each component's sum over them split
// into NL / ncomp slices when a quartet has fewer components than lanes,
// the slices added in order), or up to JIT_KMAX consecutive kets of one
// class whose primitive quartets number at most JIT_PASSQ together: their
// E/R work shares the lanes, each (ket, component) sums its own primitive
// quartets in order (in the small-LDS launch).  The pass boundaries depend
// only on the task, so the summation order is fixed.  pk holds 6 * (JIT_KMAX + 1) integers of
// per-ket data (LDS on the device).
template <int NL, int LDSN, int TMAX>
__host__ __device__ void jit_task(const JitView &v, int task, int lane, double *lds, const short *hm, const short *eo,
                                  long long *pk) {
    const int a = v.t_pair[task];
    const int la = v.pinfo[4 * a], lb = v.pinfo[4 * a + 1], npa = v.pinfo[4 * a + 2], pa0 = v.pinfo[4 * a + 3];
    const double *Ag = v.pgeo + 6 * a;
    const int na = jit_ncart(la), nb = jit_ncart(lb), nab = na * nb, Lab = la + lb;
    const int esa_d = (la + 1) * (lb + 1) * (Lab + 1), esa = 3 * esa_d;
    constexpr int RPL = NL >= JIT_NAB ? 1 : JIT_NAB;   // rows per lane
    constexpr int S = JIT_KMAX + 1;                   // pk block stride
    double racc[RPL];
    for (int k = 0; k < RPL; k++) racc[k] = 0.0;
    const int te1 = v.t_e1[task];
    for (int ge = v.t_e0[task]; ge < te1;) {
        const int e = v.gent[ge];
        const int b = v.kl_pair[e];
        const int *km = v.kmap + v.kl_moff[e];
        const int lc = v.pinfo[4 * b], ld = v.pinfo[4 * b + 1], npb = v.pinfo[4 * b + 2], pb0 = v.pinfo[4 * b + 3];
        const double *Cg = v.pgeo + 6 * b;
        const int nc = jit_ncart(lc), nd = jit_ncart(ld), ncd = nc * nd, Lcd = lc + ld, Lt = Lab + Lcd;
        const int esb_d = (lc + 1) * (ld + 1) * (Lcd + 1), esb = 3 * esb_d;
        const int rsz = jit_nherm(Lt) + Lt + 1;   // R^{(0)} and the scaled Boys values
        const int ncomp = nab * ncd;
        const int perq = esa + esb + rsz;
        const JitTabs &tb = jit_tabs();
        const unsigned long long m_comp = tb.magic[ncomp], m_cd = tb.magic[ncd], m_b = tb.magic[npb];
        const short *eo_a = eo + 4 * tb.pcoff[la * 3 + lb], *eo_c = eo + 4 * tb.pcoff[lc * 3 + ld];
        // ---- the pass: kets ge .. ge + K - 1 of this class
        // (the small-LDS launch only: in the other the extra path costs spills)
        int K = 1, sq = npa * npb;
        if (LDSN == JIT_LDS_S && sq <= JIT_PASSQ)
            while (ge + K < te1 && K < JIT_KMAX) {
                const int bn = v.kl_pair[v.gent[ge + K]];
                if (v.pinfo[4 * bn] != lc || v.pinfo[4 * bn + 1] != ld) break;
                const int nq2 = npa * v.pinfo[4 * bn + 2];
                if (sq + nq2 > JIT_PASSQ || (K + 1) * (ncomp + ncd) + (sq + nq2) * perq > LDSN) break;
                sq += nq2;
                K++;
            }
        if (K > 1) {
            jit_sync();   // the previous pass's readers of pk are done
            if (lane == 0) {
                int q = 0;
                for (int k = 0; k < K; k++) {
                    const int ek = v.gent[ge + k], bk = v.kl_pair[ek], npk = v.pinfo[4 * bk + 2];
                    pk[k] = bk;
                    pk[S + k] = npk;
                    pk[2 * S + k] = v.pinfo[4 * bk + 3];
                    pk[3 * S + k] = v.kl_moff[ek];
                    pk[4 * S + k] = q;
                    pk[5 * S + k] = (long long)tb.magic[npk];
                    q += npa * npk;
                }
                pk[4 * S + K] = q;
            }
            double *acc = lds, *vin = lds + K * ncomp, *qb = vin + K * ncd;
            for (int c = lane; c < K * ncomp; c += NL) acc[c] = 0.0;
            jit_sync();
            for (int it = lane; it < K * ncd; it += NL) {   // every ket's input
                const int k = jit_div(it, m_cd), c = it - k * ncd;
                vin[it] = jit_in(v, v.kmap[pk[3 * S + k] + c]);
            }
            for (int qf = lane; qf < sq; qf += NL) {   // E and R of every primitive quartet of the pass
                int k = 0;
                while (k + 1 < K && qf >= (int)pk[4 * S + k + 1]) k++;
                const int q = qf - (int)pk[4 * S + k];
                const int i = jit_div(q, (unsigned long long)pk[5 * S + k]), j = q - i * (int)pk[S + k];
                const int bk = (int)pk[k];
                jit_prim(v, la, lb, lc, ld, esa_d, esb_d, v.prim + 6 * (size_t)(pa0 + i),
                         v.prim + 6 * (size_t)(pk[2 * S + k] + j), Ag, v.pgeo + 6 * bk, qb + (size_t)qf * perq);
            }
            jit_sync();
            for (int it = lane; it < K * ncomp; it += NL) {   // (ket, component) over its primitive quartets
                const int k = jit_div(it, m_comp), cmp = it - k * ncomp;
                const int ab = jit_div(cmp, m_cd), cd = cmp - ab * ncd;
                const short *ea = eo_a + 4 * ab, *ec = eo_c + 4 * cd;
                double sum = 0.0;
                for (int qf = (int)pk[4 * S + k]; qf < (int)pk[4 * S + k + 1]; qf++) {
                    const double *Ea = qb + (size_t)qf * perq;
                    sum += jit_cval<TMAX>(Lab, Lcd, Ea, Ea + esa, Ea + esa + esb, ea, ec, hm);
                }
                acc[it] = sum;
            }
            jit_sync();
            for (int k = 0; k < RPL; k++) {   // contract with every ket's input, kets in order
                const int r = lane + k * NL;
                if (r < nab)
                    for (int kk = 0; kk < K; kk++) {
                        double s = 0.0;
                        for (int c = 0; c < ncd; c++) s += acc[kk * ncomp + r * ncd + c] * vin[kk * ncd + c];
                        racc[k] += s;
                    }
            }
            jit_sync();
            ge += K;
            continue;
        }
        // ---- one ket
        const int nsl = ncomp >= NL ? 1 : jit_div(NL, m_comp), npart = nsl * ncomp;
        const int preg = ncomp > 64 ? ncomp : 64;   // partial-sum region (npart <= preg)
        double *acc = lds, *vin = lds + preg, *qb = vin + ncd;
        int bq = jit_div(LDSN - preg - ncd, tb.magic[perq]);
        if (bq > NL) bq = NL;
        for (int c = lane; c < npart; c += NL) acc[c] = 0.0;
        // the partner's input: on the device one value per lane, held in a
        // register and stored after the first batch's E/R work, so the
        // global load's latency hides behind it
        double vreg = 0.0;
        if constexpr (NL >= JIT_NAB) {
            if (lane < ncd) vreg = jit_in(v, km[lane]);
        } else {
            for (int c = lane; c < ncd; c += NL) vin[c] = jit_in(v, km[c]);
        }
        const int nq = npa * npb;
        for (int q0 = 0; q0 < nq; q0 += bq) {
            const int nqb = nq - q0 < bq ? nq - q0 : bq;
            jit_sync();
            for (int q = lane; q < nqb; q += NL) {   // E tables and R of one primitive quartet
                const int i = jit_div(q0 + q, m_b), j = q0 + q - i * npb;
                jit_prim(v, la, lb, lc, ld, esa_d, esb_d, v.prim + 6 * (size_t)(pa0 + i), v.prim + 6 * (size_t)(pb0 + j),
                         Ag, Cg, qb + (size_t)q * perq);
            }
            if constexpr (NL >= JIT_NAB)
                if (q0 == 0 && lane < ncd) vin[lane] = vreg;
            jit_sync();
            for (int it = lane; it < npart; it += NL) {   // components over the primitive quartets
                const int sl = jit_div(it, m_comp), cmp = it - sl * ncomp;
                const int ab = jit_div(cmp, m_cd), cd = cmp - ab * ncd;
                const short *ea = eo_a + 4 * ab, *ec = eo_c + 4 * cd;
                double sum = acc[it];
                for (int q = sl; q < nqb; q += nsl) {
                    const double *Ea = qb + (size_t)q * perq;
                    sum += jit_cval<TMAX>(Lab, Lcd, Ea, Ea + esa, Ea + esa + esb, ea, ec, hm);
                }
                acc[it] = sum;
            }
        }
        jit_sync();
        if (nsl > 1) {   // slices in order; lane cmp alone reads and writes column cmp
            for (int cmp = lane; cmp < ncomp; cmp += NL) {
                double t = acc[cmp];
                for (int sl = 1; sl < nsl; sl++) t += acc[sl * ncomp + cmp];
                acc[cmp] = t;
            }
            jit_sync();
        }
        for (int k = 0; k < RPL; k++) {   // contract the quartet with the partner's input
            const int r = lane + k * NL;
            if (r < nab) {
                double s = 0.0;
                for (int c = 0; c < ncd; c++) s += acc[r * ncd + c] * vin[c];
                racc[k] += s;
            }
        }
        jit_sync();
        ge++;
    }
    for (int k = 0; k < RPL; k++) {
        const int r = lane + k * NL;
        if (r < nab) v.part[(size_t)task * JIT_NAB + r] = racc[k];
    }
}

// Output rows of group g: the partials of its tasks, added in entry order.
__host__ __device__ inline void jit_reduce(const JitView &v, int g) {
    const int a = v.r_pair[g];
    const int nab = jit_ncart(v.pinfo[4 * a]) * jit_ncart(v.pinfo[4 * a + 1]);
    const int *om = v.omap + v.r_omap[g];
    for (int r = 0; r < nab; r++) {
        const int o = om[r];
        if (o < 0) continue;
        double s = 0.0;
        for (int k = v.r_t0[g]; k < v.r_t1[g]; k++) s += v.part[(size_t)v.r_tasks[k] * JIT_NAB + r];
        if (o < v.N) v.YJ[o] = s;
        else v.Y1[o - v.N] = s;
    }
}

__global__ void jit_reduce_kernel(JitView v, int ngroup) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ngroup) jit_reduce(v, g);
}

// The small-LDS launch asks for 3 waves per SIMD (168 VGPRs); the other is
// held to 2 by its LDS anyway.
template <int LDSN>
__global__ void __launch_bounds__(64, LDSN == JIT_LDS_S ? 3 : 2) jit_task_kernel(JitView v, int t0, int t1) {
    __shared__ double lds[LDSN];
    __shared__ short hm[729];   // jit_herm by (t, u, w), for lane-dependent lookups
    __shared__ short eo[400];   // E-row offsets by pair class and Cartesian pair
    __shared__ long long pk[6 * (JIT_KMAX + 1)];   // per-ket data of a multi-ket pass
    const int task = t0 + blockIdx.x;
    if (task >= t1) return;
    for (int k = threadIdx.x; k < 729; k += 64) hm[k] = (&jit_tabs_dev.herm[0][0][0])[k];
    for (int k = threadIdx.x; k < 400; k += 64) eo[k] = (&jit_tabs_dev.eoff[0][0])[k];
    __syncthreads();
    jit_task<64, LDSN, LDSN == JIT_LDS_S ? 2 : 3>(v, task, threadIdx.x, lds, hm, eo, pk);
}

// Up sweep of one level: y0[n] = U_n^T in_n for the listed nodes, where the
// input rows come from x (leaves) or from the children's y0 (index map uin).
// Four waves split the rows, lanes the columns; partial sums added in a fixed
// order.
__global__ void __launch_bounds__(256) jit_up_kernel(const int *__restrict__ nodes, const long long *__restrict__ uoff,
                                                     const int *__restrict__ unr, const int *__restrict__ unc,
                                                     const int *__restrict__ uin0, const int *__restrict__ y0off,
                                                     const int *__restrict__ uin, const double *__restrict__ U,
                                                     const double *__restrict__ X, double *__restrict__ Y0, int N) {
    __shared__ double red[4][64];
    const int n = nodes[blockIdx.x];
    const int nr = unr[n], ncl = unc[n], w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double *Un = U + uoff[n];
    const int *in = uin + uin0[n];
    for (int c0 = 0; c0 < ncl; c0 += 64) {
        const int c = c0 + lane;
        double s = 0.0;
        if (c < ncl)
            for (int r = w; r < nr; r += 4) {
                const int k = in[r];
                const double xv = k < N ? X[k] : Y0[k - N];
                s += Un[(size_t)r * ncl + c] * xv;
            }
        red[w][lane] = s;
        __syncthreads();
        if (w == 0 && c < ncl) Y0[y0off[n] + c] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
        __syncthreads();
    }
}

// Down sweep of one level: t = U_n y1[n]; row r goes to its leaf row of yd
// (index < N) or is added to the child's y1 (index - N).
__global__ void __launch_bounds__(256) jit_down_kernel(const int *__restrict__ nodes, const long long *__restrict__ uoff,
                                                       const int *__restrict__ unr, const int *__restrict__ unc,
                                                       const int *__restrict__ uin0, const int *__restrict__ y0off,
                                                       const int *__restrict__ uin, const double *__restrict__ U,
                                                       double *__restrict__ YD, double *__restrict__ Y1, int N) {
    const int n = nodes[blockIdx.x];
    const int nr = unr[n], ncl = unc[n];
    const double *Un = U + uoff[n];
    const double *y1 = Y1 + y0off[n];
    const int *out = uin + uin0[n];
    for (int r = threadIdx.x; r < nr; r += 256) {
        double s = 0.0;
        for (int c = 0; c < ncl; c++) s += Un[(size_t)r * ncl + c] * y1[c];
        const int k = out[r];
        if (k < N) YD[k] = s;
        else Y1[k - N] += s;
    }
}

__global__ void jit_zero_kernel(double *__restrict__ a, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        a[i] = 0.0;
}

__global__ void jit_add_kernel(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a[i] + b[i];
}

// ---------------------------------------------------------------------------
// Host plan
// ---------------------------------------------------------------------------
struct JitPlan {
    int N = 0, nY = 0, num_sp = 0, max_level = 0;
    std::vector<int> pinfo;
    std::vector<double> pgeo, prim;
    std::vector<int> kl_start, kl_count, kl_pair, kl_moff, kmap;
    std::vector<int> t_pair, t_e0, t_e1, gent, omap;
    std::vector<int> r_pair, r_omap, r_t0, r_t1, r_tasks;
    // sweeps
    std::vector<long long> uoff;
    std::vector<int> unr, unc, uin0, y0off, uin;
    std::vector<double> U;
    std::vector<std::vector<int>> level_nodes;   // nodes with a non-empty U, per level
    // work per matvec
    double quartets = 0, prim_quartets = 0, eri_values = 0;
    double cost_b = 0, cost_d = 0;   // planned task cost of coupling and near-field entries
    int n_big = 0;   // tasks [0, n_big) have a bra or ket pair with la + lb >= 3
    const double *boys = nullptr;
    int bgrid = 0, bcol = 0;
    double bstep = 0, btmax = 0;
};

static bool jit_plan(H2ERI_p h, JitPlan &J) {
    HostState *hs = host_state(h);
    const int np = h->num_sp, n_node = h->n_node, N = h->krnl_mat_size;
    if (!h->BD_JIT || !h->U || (int)hs->pairs.size() != np || np <= 0) {
        set_error(ERR_ARG, "on-the-fly B/D: no BD_JIT representation with shell-pair data");
        return false;
    }
    if (!h->J_pair || !h->J_row || !h->pt_cluster || !h->mat_cluster || !h->sp_bfp_sidx || !h->node_level) {
        set_error(ERR_ARG, "on-the-fly B/D: tree or skeleton lists missing");
        return false;
    }
    J.N = N;
    J.num_sp = np;
    J.max_level = h->max_level;
    // ---- shell pairs
    J.pinfo.resize(4 * (size_t)np);
    J.pgeo.resize(6 * (size_t)np);
    for (int i = 0; i < np; i++) {
        const PairView pv = pair_view(hs->pairs[i]);
        if (pv.la > JIT_LMAX || pv.lb > JIT_LMAX || pv.la < 0 || pv.lb < 0) {
            set_error(ERR_UNSUPPORTED, "on-the-fly B/D: shell pair %d has angular momentum above d", i);
            return false;
        }
        if (pv.np > 1296) {   // the multiply-high divisions by the primitive-pair count
            set_error(ERR_UNSUPPORTED, "on-the-fly B/D: shell pair %d has %d primitive pairs (at most 1296)", i, pv.np);
            return false;
        }
        if (jit_ncart(pv.la) * jit_ncart(pv.lb) != h->sp_nbfp[i] ||
            h->sp_bfp_sidx[i + 1] - h->sp_bfp_sidx[i] != h->sp_nbfp[i]) {
            set_error(ERR_ARG, "on-the-fly B/D: shell pair %d rows disagree with sp_nbfp", i);
            return false;
        }
        J.pinfo[4 * i] = pv.la;
        J.pinfo[4 * i + 1] = pv.lb;
        J.pinfo[4 * i + 2] = pv.np;
        J.pinfo[4 * i + 3] = (int)(J.prim.size() / 6);
        for (int d = 0; d < 3; d++) { J.pgeo[6 * i + d] = pv.A[d]; J.pgeo[6 * i + 3 + d] = pv.B[d]; }
        for (int k = 0; k < pv.np; k++) {
            J.prim.push_back(pv.p[k]);
            for (int d = 0; d < 3; d++) J.prim.push_back(pv.P[3 * k + d]);
            J.prim.push_back(pv.K[k]);
            J.prim.push_back(0.5 / pv.p[k]);
        }
    }
    if (J.prim.size() / 6 >= (size_t)INT32_MAX) { set_error(ERR_UNSUPPORTED, "on-the-fly B/D: too many primitive pairs"); return false; }
    // ---- nodes: ranks, y0 offsets, U and the sweep index maps
    auto rank = [&](int n) { return h->U[n] ? h->U[n]->ncol : 0; };
    auto urows = [&](int n) { return h->U[n] ? h->U[n]->nrow : 0; };
    J.y0off.assign(n_node, 0);
    J.unr.assign(n_node, 0);
    J.unc.assign(n_node, 0);
    J.uin0.assign(n_node, 0);
    J.uoff.assign(n_node, 0);
    long long ny = 0;
    for (int n = 0; n < n_node; n++) { J.y0off[n] = (int)ny; ny += rank(n); }
    if (ny + N >= INT32_MAX) { set_error(ERR_UNSUPPORTED, "on-the-fly B/D: index space too large"); return false; }
    J.nY = (int)ny;
    J.level_nodes.assign(h->max_level + 1, {});
    for (int n = 0; n < n_node; n++) {
        const int nr = urows(n), nc = rank(n);
        J.uin0[n] = (int)J.uin.size();
        if (nr == 0 || nc == 0) continue;
        const H2E_dense_mat_p Un = h->U[n];
        J.unr[n] = nr;
        J.unc[n] = nc;
        J.uoff[n] = (long long)J.U.size();
        for (int r = 0; r < nr; r++)
            for (int c = 0; c < nc; c++) J.U.push_back(Un->data[(size_t)r * Un->ld + c]);
        if (h->n_child[n] == 0) {
            const int s0 = h->mat_cluster[2 * n], s1 = h->mat_cluster[2 * n + 1];
            if (s1 - s0 + 1 != nr || s0 < 0 || s1 >= N) {
                set_error(ERR_ARG, "on-the-fly B/D: leaf %d has %d U rows for %d x rows", n, nr, s1 - s0 + 1);
                return false;
            }
            for (int r = 0; r < nr; r++) J.uin.push_back(s0 + r);
        } else {
            int tot = 0;
            for (int k = 0; k < h->n_child[n]; k++) {
                const int ch = h->children[(size_t)n * 8 + k];   // 8 child slots per node
                for (int r = 0; r < rank(ch); r++) J.uin.push_back(N + J.y0off[ch] + r);
                tot += rank(ch);
            }
            if (tot != nr) {
                set_error(ERR_ARG, "on-the-fly B/D: node %d has %d U rows for %d child ranks", n, nr, tot);
                return false;
            }
        }
        const int lv = h->node_level[n];
        if (lv < 0 || lv > h->max_level) { set_error(ERR_ARG, "on-the-fly B/D: node %d level %d", n, lv); return false; }
        J.level_nodes[lv].push_back(n);
    }
    // ---- partner lists: 2n = full rows of leaf n, 2n + 1 = skeleton rows of n
    auto is_leaf = [&](int n) { return h->n_child[n] == 0; };
    J.kl_start.assign(2 * (size_t)n_node, 0);
    J.kl_count.assign(2 * (size_t)n_node, 0);
    std::vector<std::vector<int>> skel_pos(n_node);   // row of the concatenated J_pair rows -> skeleton index
    for (int n = 0; n < n_node; n++) {
        if (is_leaf(n)) {
            J.kl_start[2 * n] = (int)J.kl_pair.size();
            for (int b = h->pt_cluster[2 * n]; b <= h->pt_cluster[2 * n + 1]; b++) {
                J.kl_pair.push_back(b);
                J.kl_moff.push_back((int)J.kmap.size());
                for (int c = 0; c < h->sp_nbfp[b]; c++) J.kmap.push_back(h->sp_bfp_sidx[b] + c);
            }
            J.kl_count[2 * n] = (int)J.kl_pair.size() - J.kl_start[2 * n];
        }
        const H2E_int_vec_p jp = h->J_pair[n], jr = h->J_row[n];
        if (!jp || !jr || rank(n) == 0) continue;
        if (jr->length != rank(n)) {
            set_error(ERR_ARG, "on-the-fly B/D: node %d has %d skeleton rows for rank %d", n, jr->length, rank(n));
            return false;
        }
        int tot = 0;
        for (int k = 0; k < jp->length; k++) tot += h->sp_nbfp[jp->data[k]];
        std::vector<int> &pos = skel_pos[n];
        pos.assign(tot, -1);
        for (int j = 0; j < jr->length; j++) {
            if (jr->data[j] < 0 || jr->data[j] >= tot) { set_error(ERR_ARG, "on-the-fly B/D: node %d skeleton row", n); return false; }
            pos[jr->data[j]] = j;
        }
        J.kl_start[2 * n + 1] = (int)J.kl_pair.size();
        int rr = 0;
        for (int k = 0; k < jp->length; k++) {
            const int b = jp->data[k];
            bool any = false;
            for (int c = 0; c < h->sp_nbfp[b]; c++) any |= pos[rr + c] >= 0;
            if (any) {
                J.kl_pair.push_back(b);
                J.kl_moff.push_back((int)J.kmap.size());
                for (int c = 0; c < h->sp_nbfp[b]; c++)
                    J.kmap.push_back(pos[rr + c] >= 0 ? N + J.y0off[n] + pos[rr + c] : -1);
            }
            rr += h->sp_nbfp[b];
        }
        J.kl_count[2 * n + 1] = (int)J.kl_pair.size() - J.kl_start[2 * n + 1];
    }
    // ---- partners
    std::vector<std::vector<int>> adm(n_node), nearf(n_node);
    for (int i = 0; i < h->n_r_adm_pair; i++) {
        const int n0 = h->r_adm_pairs[2 * i], n1 = h->r_adm_pairs[2 * i + 1];
        adm[n0].push_back(n1);
        adm[n1].push_back(n0);
    }
    for (int i = 0; i < h->n_leaf_node; i++) nearf[h->height_nodes[i]].push_back(h->height_nodes[i]);
    for (int i = 0; i < h->n_r_inadm_pair; i++) {
        const int n0 = h->r_inadm_pairs[2 * i], n1 = h->r_inadm_pairs[2 * i + 1];
        nearf[n0].push_back(n1);
        nearf[n1].push_back(n0);
    }
    auto lv = [&](int n) { return h->node_level[n]; };
    // output groups: one bra pair on one side against the side's partner
    // entries gent[e0, e1); split into tasks of bounded cost below
    struct GroupRec { int pair, omap, e0, e1; };
    std::vector<GroupRec> groups;
    // cost of one quartet in rough lane-steps: per primitive quartet, the
    // Hermite terms of all its components and the R table, over the 64
    // lanes, plus a fixed cost (loads, barriers)
    double cls[3][3][3][3];
    for (int la = 0; la <= JIT_LMAX; la++)
        for (int lb = 0; lb <= JIT_LMAX; lb++)
            for (int lc = 0; lc <= JIT_LMAX; lc++)
                for (int ld = 0; ld <= JIT_LMAX; ld++) {
                    double terms = 0;
                    int e[4][3];
                    for (int ia = 0; ia < jit_ncart(la); ia++)
                        for (int ib = 0; ib < jit_ncart(lb); ib++)
                            for (int ic = 0; ic < jit_ncart(lc); ic++)
                                for (int id = 0; id < jit_ncart(ld); id++) {
                                    jit_cart(la, ia, e[0][0], e[0][1], e[0][2]);
                                    jit_cart(lb, ib, e[1][0], e[1][1], e[1][2]);
                                    jit_cart(lc, ic, e[2][0], e[2][1], e[2][2]);
                                    jit_cart(ld, id, e[3][0], e[3][1], e[3][2]);
                                    double t = 1;
                                    for (int d = 0; d < 3; d++) t *= (e[0][d] + e[1][d] + 1) * (e[2][d] + e[3][d] + 1);
                                    terms += t;
                                }
                    const int Lt = la + lb + lc + ld;
                    cls[la][lb][lc][ld] = (3.0 * terms + 40.0 + (Lt + 1) * jit_nherm(Lt)) / 64.0;
                }
    auto qcost = [&](int a, int b) {
        return (double)J.pinfo[4 * a + 2] * J.pinfo[4 * b + 2] *
                   cls[J.pinfo[4 * a]][J.pinfo[4 * a + 1]][J.pinfo[4 * b]][J.pinfo[4 * b + 1]] + 40.0;
    };
    int side_e0 = 0, side_e1 = 0;
    std::vector<char> gent_d;   // per gent entry: from a near-field (D) list
    // ket pairs of total order 3 or 4 last in each side's entries: tasks are
    // cut where they start and run in the large-LDS launch
    auto high_ket = [&](int e) { return J.pinfo[4 * J.kl_pair[e]] + J.pinfo[4 * J.kl_pair[e] + 1] >= 3; };
    // and within each part grouped by pair class (la, lb), so that runs of
    // one class form multi-ket passes
    auto open_side = [&](const std::vector<int> &lists, int n_near) {   // the first n_near lists are near field
        side_e0 = (int)J.gent.size();
        for (int pass = 0; pass < 2; pass++)
            for (int cls = 0; cls < 9; cls++)
                for (size_t li = 0; li < lists.size(); li++) {
                    const int L = lists[li];
                    for (int e = J.kl_start[L]; e < J.kl_start[L] + J.kl_count[L]; e++) {
                        const int bb = J.kl_pair[e];
                        if (high_ket(e) == (pass == 1) && J.pinfo[4 * bb] * 3 + J.pinfo[4 * bb + 1] == cls) {
                            J.gent.push_back(e);
                            gent_d.push_back((int)li < n_near ? 1 : 0);
                        }
                    }
                }
        side_e1 = (int)J.gent.size();
    };
    auto add_tasks = [&](const std::vector<int> &, int a, const std::vector<int> &om) {
        GroupRec g;
        g.pair = a;
        g.omap = (int)J.omap.size();
        J.omap.insert(J.omap.end(), om.begin(), om.end());
        g.e0 = side_e0;
        g.e1 = side_e1;
        groups.push_back(g);
        for (int ge = g.e0; ge < g.e1; ge++) {
            const int b = J.kl_pair[J.gent[ge]];
            J.quartets += 1;
            J.prim_quartets += (double)J.pinfo[4 * a + 2] * J.pinfo[4 * b + 2];
            J.eri_values += (double)h->sp_nbfp[a] * h->sp_nbfp[b];
        }
    };
    for (int n = 0; n < n_node; n++) {
        if (is_leaf(n)) {   // full rows: near field and case (2)/(3) coupling with n uncompressed
            std::vector<int> lists;
            for (int m : nearf[n]) {
                if (!is_leaf(m)) { set_error(ERR_ARG, "on-the-fly B/D: near-field partner %d is not a leaf", m); return false; }
                lists.push_back(2 * m);
            }
            const int n_near = (int)lists.size();
            for (int m : adm[n])
                if (lv(n) < lv(m)) lists.push_back(2 * m + 1);
            open_side(lists, n_near);
            if (!lists.empty())
                for (int a = h->pt_cluster[2 * n]; a <= h->pt_cluster[2 * n + 1]; a++) {
                    std::vector<int> om(h->sp_nbfp[a]);
                    for (int r = 0; r < h->sp_nbfp[a]; r++) om[r] = h->sp_bfp_sidx[a] + r;
                    add_tasks(lists, a, om);
                }
        }
        std::vector<int> lists;   // skeleton rows: partners at the same or a coarser level
        for (int m : adm[n]) {
            if (lv(n) < lv(m)) continue;
            if (lv(m) < lv(n) && !is_leaf(m)) {
                set_error(ERR_ARG, "on-the-fly B/D: coarser admissible partner %d of %d is not a leaf", m, n);
                return false;
            }
            lists.push_back(lv(m) >= lv(n) ? 2 * m + 1 : 2 * m);
        }
        if (lists.empty()) continue;
        open_side(lists, 0);
        const H2E_int_vec_p jp = h->J_pair[n];
        if (!jp || rank(n) == 0) { set_error(ERR_ARG, "on-the-fly B/D: node %d has coupling blocks and no skeleton", n); return false; }
        int rr = 0;
        for (int k = 0; k < jp->length; k++) {
            const int a = jp->data[k];
            std::vector<int> om(h->sp_nbfp[a], -1);
            bool any = false;
            for (int r = 0; r < h->sp_nbfp[a]; r++) {
                const int j = skel_pos[n][rr + r];
                if (j >= 0) { om[r] = N + J.y0off[n] + j; any = true; }
            }
            if (any) add_tasks(lists, a, om);
            rr += h->sp_nbfp[a];
        }
    }
    // tasks: each group's entries cut into runs of at most `cap` cost, so that
    // no task holds the whole pass (a bra pair of contracted s shells against
    // a full near field is the longest); cap = total / 32768
    double total = 0;
    for (const GroupRec &g : groups)
        for (int ge = g.e0; ge < g.e1; ge++) {
            const double c = qcost(g.pair, J.kl_pair[J.gent[ge]]);
            total += c;
            (gent_d[ge] ? J.cost_d : J.cost_b) += c;
        }
    const double cap = std::max(total / 32768.0, 256.0);
    struct TaskRec { int pair, e0, e1, group; double cost; bool high; };
    std::vector<TaskRec> tasks;
    for (int gi = 0; gi < (int)groups.size(); gi++) {
        const GroupRec &g = groups[gi];
        int e0 = g.e0;
        double c = 0;
        for (int ge = g.e0; ge < g.e1; ge++) {
            c += qcost(g.pair, J.kl_pair[J.gent[ge]]);
            const bool cut = ge + 1 == g.e1 || high_ket(J.gent[ge + 1]) != high_ket(J.gent[ge]);
            if (c >= cap || cut) {
                tasks.push_back(TaskRec{g.pair, e0, ge + 1, gi, c, high_ket(J.gent[ge])});
                e0 = ge + 1;
                c = 0;
            }
        }
        if (g.e0 == g.e1) tasks.push_back(TaskRec{g.pair, g.e0, g.e1, gi, 0.0, false});   // writes zero rows
    }
    // tasks with a bra or ket pair of total order >= 3 first (the launch with
    // the larger LDS), each launch longest tasks first
    auto big = [&](int t) {
        return tasks[t].high || J.pinfo[4 * tasks[t].pair] + J.pinfo[4 * tasks[t].pair + 1] >= 3;
    };
    std::vector<int> order(tasks.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (big(x) != big(y)) return big(x);
        return tasks[x].cost > tasks[y].cost;
    });
    J.n_big = 0;
    std::vector<int> newid(tasks.size());
    for (size_t k = 0; k < order.size(); k++) {
        const TaskRec &t = tasks[order[k]];
        newid[order[k]] = (int)k;
        J.n_big += big(order[k]) ? 1 : 0;
        J.t_pair.push_back(t.pair);
        J.t_e0.push_back(t.e0);
        J.t_e1.push_back(t.e1);
    }
    // each group's tasks in entry order (they were created in that order)
    std::vector<std::vector<int>> gtasks(groups.size());
    for (size_t t = 0; t < tasks.size(); t++) gtasks[tasks[t].group].push_back(newid[t]);
    for (size_t gi = 0; gi < groups.size(); gi++) {
        J.r_pair.push_back(groups[gi].pair);
        J.r_omap.push_back(groups[gi].omap);
        J.r_t0.push_back((int)J.r_tasks.size());
        J.r_tasks.insert(J.r_tasks.end(), gtasks[gi].begin(), gtasks[gi].end());
        J.r_t1.push_back((int)J.r_tasks.size());
    }
    // ---- every index within its array
    const long long space = (long long)N + J.nY;
    for (int k : J.kmap)
        if (k < -1 || k >= space) { set_error(ERR_ARG, "on-the-fly B/D: input index %d out of range", k); return false; }
    for (int k : J.omap)
        if (k < -1 || k >= space) { set_error(ERR_ARG, "on-the-fly B/D: output index %d out of range", k); return false; }
    for (int k : J.uin)
        if (k < 0 || k >= space) { set_error(ERR_ARG, "on-the-fly B/D: sweep index %d out of range", k); return false; }
    for (size_t e = 0; e < J.kl_pair.size(); e++)
        if (J.kl_pair[e] < 0 || J.kl_pair[e] >= np || J.kl_moff[e] + h->sp_nbfp[J.kl_pair[e]] > (long long)J.kmap.size()) {
            set_error(ERR_ARG, "on-the-fly B/D: partner entry %zu out of range", e);
            return false;
        }
    for (size_t t = 0; t < J.t_pair.size(); t++)
        if (J.t_e0[t] < 0 || J.t_e0[t] > J.t_e1[t] || J.t_e1[t] > (int)J.gent.size()) {
            set_error(ERR_ARG, "on-the-fly B/D: task %zu out of range", t);
            return false;
        }
    for (int e : J.gent)
        if (e < 0 || e >= (int)J.kl_pair.size()) { set_error(ERR_ARG, "on-the-fly B/D: entry %d out of range", e); return false; }
    for (size_t g = 0; g < J.r_pair.size(); g++)
        if (J.r_omap[g] + h->sp_nbfp[J.r_pair[g]] > (long long)J.omap.size() || J.r_t1[g] > (int)J.r_tasks.size()) {
            set_error(ERR_ARG, "on-the-fly B/D: output group %zu out of range", g);
            return false;
        }
    for (int t : J.r_tasks)
        if (t < 0 || t >= (int)J.t_pair.size()) { set_error(ERR_ARG, "on-the-fly B/D: task id %d out of range", t); return false; }
    for (int i = 0; i < np; i++)
        if ((size_t)(J.pinfo[4 * i + 3] + J.pinfo[4 * i + 2]) * 6 > J.prim.size()) {
            set_error(ERR_ARG, "on-the-fly B/D: primitives of pair %d out of range", i);
            return false;
        }
    // every quartet class fits its launch's LDS with at least one primitive
    // quartet per batch (shells up to d; both orders <= 2 in the small launch)
    for (int la = 0; la <= JIT_LMAX; la++)
        for (int lb = 0; lb <= JIT_LMAX; lb++)
            for (int lc = 0; lc <= JIT_LMAX; lc++)
                for (int ld = 0; ld <= JIT_LMAX; ld++) {
                    const int ncomp = jit_ncart(la) * jit_ncart(lb) * jit_ncart(lc) * jit_ncart(ld);
                    const int ncd = jit_ncart(lc) * jit_ncart(ld), Lt = la + lb + lc + ld;
                    const int perq = 3 * (la + 1) * (lb + 1) * (la + lb + 1) + 3 * (lc + 1) * (ld + 1) * (lc + ld + 1) +
                                     jit_nherm(Lt) + Lt + 1;
                    const int need = (ncomp > 64 ? ncomp : 64) + ncd + perq;
                    if (need > (la + lb >= 3 || lc + ld >= 3 ? JIT_LDS : JIT_LDS_S)) {
                        set_error(ERR_ARG, "on-the-fly B/D: LDS budget of class (%d%d|%d%d)", la, lb, lc, ld);
                        return false;
                    }
                }
    J.boys = boys_table(&J.bgrid, &J.bcol, &J.bstep, &J.btmax);
    if (J.bcol < 2 * JIT_LMAX * 2 + 8) { set_error(ERR_ARG, "on-the-fly B/D: Boys table has too few orders"); return false; }
    return true;
}

// ---------------------------------------------------------------------------
// Device state
// ---------------------------------------------------------------------------
struct JitState {
    JitPlan P;
    bool ready = false;
    int *d_ints = nullptr;      // every int array, one allocation
    double *d_dbls = nullptr;   // pgeo, prim, boys, U
    long long *d_uoff = nullptr;
    double *Y0 = nullptr, *Y1 = nullptr, *YJ = nullptr, *YD = nullptr, *part = nullptr;
    JitView view{};
    const int *d_unr = nullptr, *d_unc = nullptr, *d_uin0 = nullptr, *d_y0off = nullptr, *d_uin = nullptr;
    const double *d_U = nullptr;
    std::vector<const int *> d_level;   // node lists per level
    std::vector<int> level_off;
    size_t bytes = 0;   // device memory held
    std::vector<hipEvent_t> ev;   // phase timing: 6 per vector
};

#define JIT_HIP(expr)                                                                                       \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            set_error(ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return false;                                                                                   \
        }                                                                                                   \
    } while (0)

void jit_release(JitState *&js) {
    if (!js) return;
    for (hipEvent_t e : js->ev) (void)hipEventDestroy(e);
    void *ps[] = {js->d_ints, js->d_dbls, js->d_uoff, js->Y0, js->Y1, js->YJ, js->YD, js->part};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete js;
    js = nullptr;
}

size_t jit_resident_bytes(const JitState *js) { return js ? js->bytes : 0; }

bool jit_upload(H2ERI_p h, JitState *&js) {
    if (js && js->ready) return true;
    jit_release(js);
    js = new JitState();
    JitPlan &P = js->P;
    if (!jit_plan(h, P)) { jit_release(js); return false; }
    // int arrays, concatenated
    std::vector<const std::vector<int> *> iv = {&P.pinfo, &P.r_pair, &P.r_omap, &P.kl_pair, &P.kl_moff, &P.kmap,
                                                &P.t_pair, &P.t_e0, &P.t_e1, &P.gent, &P.r_t0, &P.omap,
                                                &P.unr, &P.unc, &P.uin0, &P.y0off, &P.uin, &P.r_t1, &P.r_tasks};
    std::vector<int> ints;
    std::vector<size_t> ioff;
    for (auto *v : iv) { ioff.push_back(ints.size()); ints.insert(ints.end(), v->begin(), v->end()); }
    js->level_off.clear();
    for (auto &l : P.level_nodes) { js->level_off.push_back((int)ints.size()); ints.insert(ints.end(), l.begin(), l.end()); }
    ints.push_back(0);
    std::vector<double> dbls;
    const size_t o_geo = 0;
    dbls.insert(dbls.end(), P.pgeo.begin(), P.pgeo.end());
    const size_t o_prim = dbls.size();
    dbls.insert(dbls.end(), P.prim.begin(), P.prim.end());
    const size_t o_boys = dbls.size();
    dbls.insert(dbls.end(), P.boys, P.boys + (size_t)P.bgrid * P.bcol);
    const size_t o_U = dbls.size();
    dbls.insert(dbls.end(), P.U.begin(), P.U.end());
    dbls.push_back(0.0);
    JIT_HIP(hipMalloc(&js->d_ints, ints.size() * sizeof(int)));
    JIT_HIP(hipMemcpy(js->d_ints, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice));
    JIT_HIP(hipMalloc(&js->d_dbls, dbls.size() * sizeof(double)));
    JIT_HIP(hipMemcpy(js->d_dbls, dbls.data(), dbls.size() * sizeof(double), hipMemcpyHostToDevice));
    JIT_HIP(hipMalloc(&js->d_uoff, (P.uoff.size() + 1) * sizeof(long long)));
    JIT_HIP(hipMemcpy(js->d_uoff, P.uoff.data(), P.uoff.size() * sizeof(long long), hipMemcpyHostToDevice));
    const size_t ny = (size_t)std::max(P.nY, 1), nn = (size_t)std::max(P.N, 1);
    JIT_HIP(hipMalloc(&js->Y0, ny * sizeof(double)));
    JIT_HIP(hipMalloc(&js->Y1, ny * sizeof(double)));
    JIT_HIP(hipMalloc(&js->YJ, nn * sizeof(double)));
    JIT_HIP(hipMalloc(&js->YD, nn * sizeof(double)));
    const size_t npart = (size_t)std::max<size_t>(P.t_pair.size(), 1) * JIT_NAB;
    JIT_HIP(hipMalloc(&js->part, npart * sizeof(double)));
    js->bytes = ints.size() * sizeof(int) + dbls.size() * sizeof(double) + (P.uoff.size() + 1) * sizeof(long long) +
                (2 * ny + 2 * nn + npart) * sizeof(double);
    const int *I = js->d_ints;
    JitView &v = js->view;
    v.N = P.N;
    v.pinfo = I + ioff[0];
    v.r_pair = I + ioff[1];
    v.r_omap = I + ioff[2];
    v.kl_pair = I + ioff[3];
    v.kl_moff = I + ioff[4];
    v.kmap = I + ioff[5];
    v.t_pair = I + ioff[6];
    v.t_e0 = I + ioff[7];
    v.t_e1 = I + ioff[8];
    v.gent = I + ioff[9];
    v.r_t0 = I + ioff[10];
    v.omap = I + ioff[11];
    js->d_unr = I + ioff[12];
    js->d_unc = I + ioff[13];
    js->d_uin0 = I + ioff[14];
    js->d_y0off = I + ioff[15];
    js->d_uin = I + ioff[16];
    v.r_t1 = I + ioff[17];
    v.r_tasks = I + ioff[18];
    v.part = js->part;
    js->d_level.clear();
    for (int o : js->level_off) js->d_level.push_back(I + o);
    v.pgeo = js->d_dbls + o_geo;
    v.prim = js->d_dbls + o_prim;
    v.boys = js->d_dbls + o_boys;
    v.bgrid = P.bgrid;
    v.bcol = P.bcol;
    v.bstep = P.bstep;
    v.btmax = P.btmax;
    v.inv_bstep = 1.0 / P.bstep;
    js->d_U = js->d_dbls + o_U;
    v.Y0 = js->Y0;
    v.Y1 = js->Y1;
    v.YJ = js->YJ;
    js->ready = true;
    return true;
}

bool jit_matvec(H2ERI_p h, JitState *js, int nvec, const double *X, double *Y, void *stream, double *ms5) {
    if (!js || !js->ready) { set_error(ERR_ARG, "on-the-fly B/D: not uploaded"); return false; }
    hipStream_t s = (hipStream_t)stream;
    const JitPlan &P = js->P;
    const int N = P.N;
    if (ms5)
        while (js->ev.size() < 6 * (size_t)nvec) {
            hipEvent_t e;
            JIT_HIP(hipEventCreate(&e));
            js->ev.push_back(e);
        }
    auto mark = [&](int k, int i) -> bool {
        if (ms5) JIT_HIP(hipEventRecord(js->ev[6 * k + i], s));
        return true;
    };
    for (int k = 0; k < nvec; k++) {
        const double *x = X + (size_t)k * N;
        double *y = Y + (size_t)k * N;
        if (!mark(k, 0)) return false;
        hipLaunchKernelGGL(jit_zero_kernel, dim3(1024), dim3(256), 0, s, js->Y1, (long long)std::max(P.nY, 1));
        hipLaunchKernelGGL(jit_zero_kernel, dim3(1024), dim3(256), 0, s, js->YJ, (long long)std::max(N, 1));
        hipLaunchKernelGGL(jit_zero_kernel, dim3(1024), dim3(256), 0, s, js->YD, (long long)std::max(N, 1));
        JIT_HIP(hipGetLastError());
        if (!mark(k, 1)) return false;
        for (int L = P.max_level; L >= 0; L--) {
            const int nn = (int)P.level_nodes[L].size();
            if (nn == 0) continue;
            hipLaunchKernelGGL(jit_up_kernel, dim3(nn), dim3(256), 0, s, js->d_level[L], js->d_uoff, js->d_unr,
                               js->d_unc, js->d_uin0, js->d_y0off, js->d_uin, js->d_U, x, js->Y0, N);
            JIT_HIP(hipGetLastError());
        }
        if (!mark(k, 2)) return false;
        JitView v = js->view;
        v.X = x;
        const int ntask = (int)P.t_pair.size();
        if (P.n_big > 0) {
            hipLaunchKernelGGL(jit_task_kernel<JIT_LDS>, dim3(P.n_big), dim3(64), 0, s, v, 0, P.n_big);
            JIT_HIP(hipGetLastError());
        }
        if (ntask > P.n_big) {
            hipLaunchKernelGGL(jit_task_kernel<JIT_LDS_S>, dim3(ntask - P.n_big), dim3(64), 0, s, v, P.n_big, ntask);
            JIT_HIP(hipGetLastError());
        }
        const int ngroup = (int)P.r_pair.size();
        if (ngroup > 0) {
            hipLaunchKernelGGL(jit_reduce_kernel, dim3((ngroup + 255) / 256), dim3(256), 0, s, v, ngroup);
            JIT_HIP(hipGetLastError());
        }
        if (!mark(k, 3)) return false;
        for (int L = 0; L <= P.max_level; L++) {
            const int nn = (int)P.level_nodes[L].size();
            if (nn == 0) continue;
            hipLaunchKernelGGL(jit_down_kernel, dim3(nn), dim3(256), 0, s, js->d_level[L], js->d_uoff, js->d_unr,
                               js->d_unc, js->d_uin0, js->d_y0off, js->d_uin, js->d_U, js->YD, js->Y1, N);
            JIT_HIP(hipGetLastError());
        }
        if (!mark(k, 4)) return false;
        hipLaunchKernelGGL(jit_add_kernel, dim3((N + 255) / 256), dim3(256), 0, s, js->YD, js->YJ, y, N);
        JIT_HIP(hipGetLastError());
        if (!mark(k, 5)) return false;
    }
    if (ms5) {   // per phase: up sweep, B and D tasks (split by planned cost), down sweep, vector work
        JIT_HIP(hipEventSynchronize(js->ev[6 * (size_t)nvec - 1]));
        const double fb = P.cost_b + P.cost_d > 0 ? P.cost_b / (P.cost_b + P.cost_d) : 0.0;
        for (int k = 0; k < nvec; k++) {
            float t[5];
            for (int i = 0; i < 5; i++) JIT_HIP(hipEventElapsedTime(&t[i], js->ev[6 * k + i], js->ev[6 * k + i + 1]));
            ms5[0] += t[1];
            ms5[1] += t[2] * fb;
            ms5[2] += t[3];
            ms5[3] += t[2] * (1.0 - fb);
            ms5[4] += t[0] + t[4];
        }
    }
    (void)h;
    return true;
}

}  // namespace h2e

using namespace h2e;

extern "C" {

// Host execution of the on-the-fly matvec (tests only): the same row-task
// code on one lane, the sweeps as plain loops.  Returns 0, or -1 with the
// error set.
int H2ERI_jit_emulate(H2ERI_p h, const double *x, double *y) {
    if (!h || !x || !y) { set_error(ERR_ARG, "jit_emulate: null argument"); return -1; }
    JitPlan P;
    if (!jit_plan(h, P)) return -1;
    const int N = P.N;
    std::vector<double> Y0(std::max(P.nY, 1), 0.0), Y1(std::max(P.nY, 1), 0.0), YJ(N, 0.0), YD(N, 0.0);
    for (int L = P.max_level; L >= 0; L--)
        for (int n : P.level_nodes[L]) {
            const double *Un = P.U.data() + P.uoff[n];
            const int *in = P.uin.data() + P.uin0[n];
            for (int c = 0; c < P.unc[n]; c++) {
                double part[4] = {0, 0, 0, 0};
                for (int r = 0; r < P.unr[n]; r++) {
                    const int k = in[r];
                    part[r & 3] += Un[(size_t)r * P.unc[n] + c] * (k < N ? x[k] : Y0[k - N]);
                }
                Y0[P.y0off[n] + c] = (part[0] + part[1]) + (part[2] + part[3]);
            }
        }
    JitView v{};
    v.N = N;
    v.pinfo = P.pinfo.data(); v.pgeo = P.pgeo.data(); v.prim = P.prim.data();
    v.boys = P.boys; v.bgrid = P.bgrid; v.bcol = P.bcol; v.bstep = P.bstep; v.btmax = P.btmax; v.inv_bstep = 1.0 / P.bstep;
    v.kl_pair = P.kl_pair.data(); v.kl_moff = P.kl_moff.data(); v.kmap = P.kmap.data();
    v.t_pair = P.t_pair.data(); v.t_e0 = P.t_e0.data(); v.t_e1 = P.t_e1.data(); v.gent = P.gent.data();
    v.r_pair = P.r_pair.data(); v.r_omap = P.r_omap.data(); v.r_t0 = P.r_t0.data(); v.r_t1 = P.r_t1.data();
    v.r_tasks = P.r_tasks.data(); v.omap = P.omap.data();
    std::vector<double> part(std::max<size_t>(P.t_pair.size(), 1) * JIT_NAB, 0.0);
    v.part = part.data();
    v.X = x; v.Y0 = Y0.data(); v.YJ = YJ.data(); v.Y1 = Y1.data();
    const int ntask = (int)P.t_pair.size();
    std::atomic<int> next(0);
    auto worker = [&]() {
        std::vector<double> lds(JIT_LDS);
        std::vector<long long> pk(6 * (JIT_KMAX + 1));
        for (int t = next++; t < ntask; t = next++) {
            const short *hm = &jit_tabs_host.herm[0][0][0], *eo = &jit_tabs_host.eoff[0][0];
            if (t < P.n_big) jit_task<1, JIT_LDS, 3>(v, t, 0, lds.data(), hm, eo, pk.data());
            else jit_task<1, JIT_LDS_S, 2>(v, t, 0, lds.data(), hm, eo, pk.data());
        }
    };
    const int nth = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int i = 1; i < nth; i++) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
    for (int g = 0; g < (int)P.r_pair.size(); g++) jit_reduce(v, g);
    for (int L = 0; L <= P.max_level; L++)
        for (int n : P.level_nodes[L]) {
            const double *Un = P.U.data() + P.uoff[n];
            const int *out = P.uin.data() + P.uin0[n];
            for (int r = 0; r < P.unr[n]; r++) {
                double s = 0.0;
                for (int c = 0; c < P.unc[n]; c++) s += Un[(size_t)r * P.unc[n] + c] * Y1[P.y0off[n] + c];
                const int k = out[r];
                if (k < N) YD[k] = s;
                else Y1[k - N] += s;
            }
        }
    for (int i = 0; i < N; i++) y[i] = YD[i] + YJ[i];
    return 0;
}

// Work of one on-the-fly matvec: out4 = {row tasks, shell quartets,
// primitive quartets, integral values}.  Host only.  Returns 0 or -1.
int H2ERI_jit_stats(H2ERI_p h, double *out4) {
    if (!h || !out4) { set_error(ERR_ARG, "jit_stats: null argument"); return -1; }
    JitPlan P;
    if (!jit_plan(h, P)) return -1;
    out4[0] = (double)P.t_pair.size();
    out4[1] = P.quartets;
    out4[2] = P.prim_quartets;
    out4[3] = P.eri_values;
    return 0;
}

// B and D blocks stored dense (on = 0) or compressed to [U | V^T] where it
// pays (on = 1, the default), from the next H2ERI_build_H2 on.
void H2ERI_set_block_compression(H2ERI_p h, const int on) {
    if (!h) { set_error(ERR_ARG, "set_block_compression: null handle"); return; }
    host_state(h)->bd_dense = on == 0;
}

}  // extern "C"
