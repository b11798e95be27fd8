/* Oracle restatement of H2P-ERI's AOT H2 matvec -- TEST INFRASTRUCTURE ONLY.
 *
 * Step by step as src/H2ERI_matvec.c:
 *   fused bi-GEMV  o0 += M i0, o1 += M^T i1 with one pass over M   :24-63
 *   forward sweep  max_level -> min_adm_level; leaf y0 = U^T x_n,
 *                  inner y0 = sum_k U_k^T y0[child_k]               :93-175
 *   block product  dense: bi-GEMV; low-rank [U|V^T]:
 *                  o0 += U (V i0), o1 += V^T (U^T i1)               :225-264
 *   coupling       three level cases, thread-private y1 and y       :299-407
 *   backward sweep min_adm_level -> max_level; t = U y1; leaf y += t,
 *                  inner: push t's slices to the children's y1       :663-757
 *   near field     diagonal y_n += D_nn x_n only; inadmissible pairs
 *                  y_0 += D x_1, y_1 += D^T x_0                      :572-660
 *   reduction      y += sum of thread-private y                      :882-957
 * Parity pins (tests/test_oracle.py): equality with the dense brute force
 * orc_dense_matrix on trees with every coupling case and low-rank blocks;
 * symmetry u^T(A v) = v^T(A u) on symmetric representations; linearity;
 * end to end through the reference's J_alk32 fixture (J within the H2
 * tolerance of the fixture computed by the reference's own test).
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
static int nthreads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

/* o0 += M i0 ; o1 += M^T i1  (M row-major nrow x ncol, one pass).  The row
 * dots are SIMD reductions (several partial sums), as a BLAS dgemv's are. */
static void bi_gemv(int nrow, int ncol, const double *M, const double *i0, const double *i1, double *o0, double *o1) {
    for (int r = 0; r < nrow; r++) {
        const double *row = M + (size_t)r * ncol;
        double s = 0.0;
        if (o1) {
            const double a = i1[r];
            #pragma omp simd reduction(+:s)
            for (int c = 0; c < ncol; c++) {
                s += row[c] * i0[c];
                o1[c] += a * row[c];
            }
        } else {
            #pragma omp simd reduction(+:s)
            for (int c = 0; c < ncol; c++) s += row[c] * i0[c];
        }
        o0[r] += s;
    }
}

/* y (=|+=) M^T x, M nrow x ncol row-major */
static void gemv_t(int nrow, int ncol, const double *M, const double *x, double *y, int beta1) {
    if (!beta1) memset(y, 0, sizeof(double) * ncol);
    for (int r = 0; r < nrow; r++) {
        const double *row = M + (size_t)r * ncol;
        const double a = x[r];
        for (int c = 0; c < ncol; c++) y[c] += a * row[c];
    }
}

/* y (=|+=) M x */
static void gemv_n(int nrow, int ncol, const double *M, const double *x, double *y, int beta1) {
    for (int r = 0; r < nrow; r++) {
        const double *row = M + (size_t)r * ncol;
        double s = 0.0;
        #pragma omp simd reduction(+:s)
        for (int c = 0; c < ncol; c++) s += row[c] * x[c];
        y[r] = beta1 ? y[r] + s : s;
    }
}

/* block product with optional transposed half (o1 may be NULL) */
static void blk_bi(const orc_blk *b, const double *i0, const double *i1, double *o0, double *o1, double *tmp) {
    if (b->nrow == 0 || b->ncol == 0) return;
    if (b->ld >= 0) {
        bi_gemv(b->nrow, b->ncol, b->data, i0, i1, o0, o1);
        return;
    }
    const int k = -b->ld;
    const double *U = b->data, *VT = U + (size_t)b->nrow * k;
    gemv_t(b->ncol, k, VT, i0, tmp, 0);          /* tmp = V i0 */
    gemv_n(b->nrow, k, U, tmp, o0, 1);           /* o0 += U tmp */
    if (o1) {
        gemv_t(b->nrow, k, U, i1, tmp, 0);       /* tmp = U^T i1 */
        gemv_n(b->ncol, k, VT, tmp, o1, 1);      /* o1 += V^T tmp */
    }
}

static int rank_of(const orc_rep *r, int n) { return r->U[n].ncol; }

void orc_matvec(orc_rep *r, const double *x, double *y) {
    const int N = r->N, nn = r->n_node, nl = r->n_leaf, nt = nthreads();
    /* y0 and per-thread y1 / y */
    size_t *yo = (size_t *)malloc(sizeof(size_t) * (nn + 1));
    yo[0] = 0;
    for (int i = 0; i < nn; i++) yo[i + 1] = yo[i] + (size_t)rank_of(r, i);
    const size_t R = yo[nn];
    double *y0 = (double *)calloc(R + 1, sizeof(double));
    double *y1t = (double *)calloc((R + 1) * (size_t)nt, sizeof(double));
    double *yt = (double *)calloc((size_t)N * nt + 1, sizeof(double));
    int maxk = 1;
    for (int i = 0; i < r->n_adm; i++) if (r->B[i].ld < 0 && -r->B[i].ld > maxk) maxk = -r->B[i].ld;
    for (int i = 0; i < nl + r->n_inadm; i++) if (r->D[i].ld < 0 && -r->D[i].ld > maxk) maxk = -r->D[i].ld;
    int maxrow = 1;
    for (int i = 0; i < nn; i++) if (r->U[i].nrow > maxrow) maxrow = r->U[i].nrow;
    memset(y, 0, sizeof(double) * N);

    /* forward sweep */
    for (int lv = r->max_level; lv >= r->min_adm_level; lv--) {
        const int cnt = r->level_n_node[lv];
        const int *nodes = r->level_nodes + (size_t)lv * nl;
        #pragma omp parallel for schedule(dynamic) num_threads(nt)
        for (int j = 0; j < cnt; j++) {
            const int node = nodes[j];
            const orc_blk *U = &r->U[node];
            if (U->ncol == 0) continue;
            double *out = y0 + yo[node];
            if (r->n_child[node] == 0) {
                gemv_t(U->nrow, U->ncol, U->data, x + r->mat_cluster[2 * node], out, 0);
            } else {
                int srow = 0;
                for (int k = 0; k < r->n_child[node]; k++) {
                    const int ch = r->children[8 * node + k];
                    const int rc = rank_of(r, ch);
                    gemv_t(rc, U->ncol, U->data + (size_t)srow * U->ncol, y0 + yo[ch], out, k > 0);
                    srow += rc;
                }
            }
        }
    }
    /* coupling */
    #pragma omp parallel num_threads(nt)
    {
        const int tid = omp_get_thread_num();
        double *y1 = y1t + (R + 1) * (size_t)tid;
        double *ty = yt + (size_t)N * tid;
        double *tmp = (double *)malloc(sizeof(double) * (maxk + 1));
        #pragma omp for schedule(dynamic)
        for (int i = 0; i < r->n_adm; i++) {
            const int n0 = r->adm[2 * i], n1 = r->adm[2 * i + 1];
            const int l0 = r->node_level[n0], l1 = r->node_level[n1];
            if (l0 == l1)
                blk_bi(&r->B[i], y0 + yo[n1], y0 + yo[n0], y1 + yo[n0], y1 + yo[n1], tmp);
            else if (l0 > l1)
                blk_bi(&r->B[i], x + r->mat_cluster[2 * n1], y0 + yo[n0], y1 + yo[n0], ty + r->mat_cluster[2 * n1], tmp);
            else
                blk_bi(&r->B[i], y0 + yo[n1], x + r->mat_cluster[2 * n0], ty + r->mat_cluster[2 * n0], y1 + yo[n1], tmp);
        }
        free(tmp);
    }
    double *y1 = y1t;   /* thread 0 slab holds the sum */
    #pragma omp parallel for schedule(static) num_threads(nt)
    for (size_t e = 0; e < R; e++)
        for (int t = 1; t < nt; t++) y1[e] += y1t[(R + 1) * (size_t)t + e];
    /* backward sweep */
    for (int lv = r->min_adm_level; lv <= r->max_level; lv++) {
        const int cnt = r->level_n_node[lv];
        const int *nodes = r->level_nodes + (size_t)lv * nl;
        #pragma omp parallel num_threads(nt)
        {
            double *t = (double *)malloc(sizeof(double) * (maxrow + 1));
            #pragma omp for schedule(dynamic)
            for (int j = 0; j < cnt; j++) {
                const int node = nodes[j];
                const orc_blk *U = &r->U[node];
                if (U->ncol == 0 || U->nrow == 0) continue;
                gemv_n(U->nrow, U->ncol, U->data, y1 + yo[node], t, 0);
                if (r->n_child[node] == 0) {
                    double *ys = y + r->mat_cluster[2 * node];
                    for (int k = 0; k < U->nrow; k++) ys[k] += t[k];
                } else {
                    int off = 0;
                    for (int k = 0; k < r->n_child[node]; k++) {
                        const int ch = r->children[8 * node + k];
                        const int rc = rank_of(r, ch);
                        double *dst = y1 + yo[ch];
                        for (int l = 0; l < rc; l++) dst[l] += t[off + l];
                        off += rc;
                    }
                }
            }
            free(t);
        }
    }
    /* near field */
    #pragma omp parallel num_threads(nt)
    {
        const int tid = omp_get_thread_num();
        double *ty = yt + (size_t)N * tid;
        double *tmp = (double *)malloc(sizeof(double) * (maxk + 1));
        #pragma omp for schedule(dynamic) nowait
        for (int i = 0; i < nl; i++) {
            const int node = r->height_nodes[i];
            const int s = r->mat_cluster[2 * node];
            blk_bi(&r->D[i], x + s, NULL, ty + s, NULL, tmp);
        }
        #pragma omp for schedule(dynamic)
        for (int i = 0; i < r->n_inadm; i++) {
            const int n0 = r->inadm[2 * i], n1 = r->inadm[2 * i + 1];
            const int s0 = r->mat_cluster[2 * n0], s1 = r->mat_cluster[2 * n1];
            blk_bi(&r->D[nl + i], x + s1, x + s0, ty + s0, ty + s1, tmp);
        }
        free(tmp);
    }
    /* reduction of thread-private y */
    #pragma omp parallel for schedule(static) num_threads(nt)
    for (int e = 0; e < N; e++)
        for (int t = 0; t < nt; t++) y[e] += yt[(size_t)N * t + e];
    free(yo); free(y0); free(y1t); free(yt);
}

/* J := contract(A uncontract(D))  (src/H2ERI_build_Coulomb.c:26-152) */
void orc_build_coulomb(orc_rep *r, const double *D, double *J) {
    const int nbf = r->num_bf, np = r->num_sp, N = r->N;
    double *x = (double *)malloc(sizeof(double) * (N + 1));
    double *y = (double *)malloc(sizeof(double) * (N + 1));
    #pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads())
    for (int i = 0; i < np; i++) {
        const int s0 = r->sp_shell_idx[i], s1 = r->sp_shell_idx[i + np];
        const int srow = r->shell_bf_sidx[s0], nrow = r->shell_bf_sidx[s0 + 1] - srow;
        const int scol = r->shell_bf_sidx[s1], ncol = r->shell_bf_sidx[s1 + 1] - scol;
        const double c = (s0 == s1) ? 1.0 : 2.0;
        double *xp = x + r->sp_bfp_sidx[i];
        for (int j = 0; j < ncol; j++)
            for (int k = 0; k < nrow; k++) xp[j * nrow + k] = c * D[(size_t)(scol + j) * nbf + srow + k];
    }
    orc_matvec(r, x, y);
    memset(J, 0, sizeof(double) * (size_t)nbf * nbf);
    #pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads())
    for (int i = 0; i < np; i++) {
        const int s0 = r->sp_shell_idx[i], s1 = r->sp_shell_idx[i + np];
        const int srow = r->shell_bf_sidx[s0], nrow = r->shell_bf_sidx[s0 + 1] - srow;
        const int scol = r->shell_bf_sidx[s1], ncol = r->shell_bf_sidx[s1 + 1] - scol;
        const double c = (s0 == s1) ? 0.5 : 1.0;
        const double *yp = y + r->sp_bfp_sidx[i];
        for (int j = 0; j < ncol; j++)
            for (int k = 0; k < nrow; k++) J[(size_t)(scol + j) * nbf + srow + k] += c * yp[j * nrow + k];
    }
    for (int i = 0; i < nbf; i++) {
        for (int j = 0; j < i; j++) {
            const double v = J[(size_t)i * nbf + j] + J[(size_t)j * nbf + i];
            J[(size_t)i * nbf + j] = v;
            J[(size_t)j * nbf + i] = v;
        }
        J[(size_t)i * nbf + i] += J[(size_t)i * nbf + i];
    }
    free(x);
    free(y);
}
