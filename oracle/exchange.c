/* Oracle exchange matrix through the H2 representation -- TEST INFRASTRUCTURE ONLY.
 *
 *   K_mq = sum_{n,p} (mn|pq) D_np            (src/H2ERI_build_exchange.c:1356-1450)
 *
 * The reference evaluates this with partial multi-column H2 products, one
 * per (N,Q) shell pair (:1021-1353).  The input columns are (n,q) for n in N
 * and q in Q, gathered from D over the pairs (P,Q) (:609-688). The output rows
 * are the pairs (M,N), scattered into K (:691-771).  A screened shell pair is
 * stored once, smaller shell index first (:275-281).  Its row (a,b) stands for
 * both (ab| and (ba| when the two shells differ; a same-shell pair (M,M) has
 * one row per ordered (a,b).  Those partial products restrict
 *   A_H2 = sum_i D_ii + sum_{(i,j)} [F_ij + F_ij^T],  F_ij = P_i C_ij Q_j^T
 * (diagonal near-field blocks once; every reduced admissible / inadmissible
 * pair with its transpose, src/H2ERI_matvec.c:299-407,572-660) to the rows
 * and columns of (N,Q).  This oracle evaluates the same sum block by block:
 * each F_ij is formed in row chunks from the expanded nested bases
 * (P_i = U of a leaf, or blockdiag(P of children) U of an inner node; the
 * identity on a side that uses the raw vector) and contracted with D over the
 * orientations of its rows and columns.  For symmetric D the F^T terms
 * contribute K1^T, where K1 is the sum over the F terms, so
 *   K = K1 + K1^T + K_diag.
 * The density is screened first as the reference screens it (orc_dlist):
 * blocks (N,P) of D with every |D_np| <= 1e-10 take no part.
 * Parity pins (tests/test_oracle_exchange.py): equality with the contraction
 * of the dense brute-force matrix orc_dense_matrix; the reference fixture
 * tests/K_alk26_ccpvdz.bin (within the H2 tolerance); symmetry and linearity
 * in D for densities without sub-threshold blocks; the screening list
 * against a numpy restatement and K of a density with sub-threshold blocks
 * against the contraction with those blocks zeroed.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

static int x_threads(void) { return omp_get_max_threads(); }

typedef struct {
    int s0, m0, k0;       /* rows [s0, s0+m0); basis P (m0 x k0) or NULL = identity (k0 = m0) */
    const double *P;
    int s1, m1, k1;       /* columns [s1, s1+m1); basis Q (m1 x k1) or NULL */
    const double *Q;
    const orc_blk *C;     /* core k0 x k1 */
    int diag;             /* 1: diagonal near-field block (applied once) */
} x_term;

/* out (h x k1) = X (h x k0) * C, C dense or low-rank [U|V^T] */
static void mul_core(int h, int k0, const double *X, int ldx, const orc_blk *C, double *out, double *tmp) {
    const int k1 = C->ncol;
    memset(out, 0, sizeof(double) * (size_t)h * k1);
    if (C->ld >= 0) {
        for (int r = 0; r < h; r++)
            for (int s = 0; s < k0; s++) {
                const double a = X[(size_t)r * ldx + s];
                if (a == 0.0) continue;
                const double *crow = C->data + (size_t)s * k1;
                double *o = out + (size_t)r * k1;
                for (int t = 0; t < k1; t++) o[t] += a * crow[t];
            }
        return;
    }
    const int k = -C->ld;
    const double *U = C->data, *VT = U + (size_t)C->nrow * k;
    memset(tmp, 0, sizeof(double) * (size_t)h * k);
    for (int r = 0; r < h; r++)
        for (int s = 0; s < k0; s++) {
            const double a = X[(size_t)r * ldx + s];
            if (a == 0.0) continue;
            const double *urow = U + (size_t)s * k;
            double *o = tmp + (size_t)r * k;
            for (int l = 0; l < k; l++) o[l] += a * urow[l];
        }
    for (int r = 0; r < h; r++)
        for (int t = 0; t < k1; t++) {
            const double *vrow = VT + (size_t)t * k;
            const double *w = tmp + (size_t)r * k;
            double s = 0.0;
            for (int l = 0; l < k; l++) s += w[l] * vrow[l];
            out[(size_t)r * k1 + t] = s;
        }
}

/* out (h x ncol) = rows [r0, r0+h) of C */
static void core_rows(const orc_blk *C, int r0, int h, double *out) {
    const int k1 = C->ncol;
    if (C->ld >= 0) {
        memcpy(out, C->data + (size_t)r0 * k1, sizeof(double) * (size_t)h * k1);
        return;
    }
    const int k = -C->ld;
    const double *U = C->data, *VT = U + (size_t)C->nrow * k;
    for (int q = 0; q < h; q++)
        for (int t = 0; t < k1; t++) {
            const double *u = U + (size_t)(r0 + q) * k, *v = VT + (size_t)t * k;
            double s = 0.0;
            for (int l = 0; l < k; l++) s += u[l] * v[l];
            out[(size_t)q * k1 + t] = s;
        }
}

/* Density screening list (src/H2ERI_build_exchange.c:774-838): shell P is in
 * dlist[N] when some |D_np|, n in N and p in P, exceeds DTOL = 1e-10 (:790).
 * Fills sidx[nshell + 1] and, when list is not NULL, the lists (sorted P per
 * N); returns the total length. */
int orc_dlist(const orc_rep *r, const double *D, int *sidx, int *list) {
    const int nbf = r->num_bf, ns = r->nshell;
    const double DTOL = 1e-10;
    int pos = 0;
    sidx[0] = 0;
    for (int n = 0; n < ns; n++) {
        for (int p = 0; p < ns; p++) {
            int flag = 0;
            for (int i = r->shell_bf_sidx[n]; i < r->shell_bf_sidx[n + 1] && !flag; i++)
                for (int j = r->shell_bf_sidx[p]; j < r->shell_bf_sidx[p + 1]; j++)
                    if (fabs(D[(size_t)i * nbf + j]) > DTOL) { flag = 1; break; }
            if (flag) {
                if (list) list[pos] = p;
                pos++;
            }
        }
        sidx[n + 1] = pos;
    }
    return pos;
}

/* The reference takes, for each (N,Q), only the shells P of plist[Q] that
 * are also in dlist[N] (src/H2ERI_build_exchange.c:410-417,1397-1398): the
 * sum over (n,p) then runs over the blocks (N,P) of dlist only, which is the
 * sum with every other block of D set to zero. */
static double *screened_density(const orc_rep *r, const double *D) {
    const int nbf = r->num_bf, ns = r->nshell;
    int *sidx = (int *)malloc(sizeof(int) * (ns + 1));
    const int len = orc_dlist(r, D, sidx, NULL);
    int *list = (int *)malloc(sizeof(int) * (len > 0 ? len : 1));
    orc_dlist(r, D, sidx, list);
    double *Ds = (double *)calloc((size_t)nbf * nbf, sizeof(double));
    for (int n = 0; n < ns; n++)
        for (int k = sidx[n]; k < sidx[n + 1]; k++) {
            const int p = list[k];
            for (int i = r->shell_bf_sidx[n]; i < r->shell_bf_sidx[n + 1]; i++)
                for (int j = r->shell_bf_sidx[p]; j < r->shell_bf_sidx[p + 1]; j++)
                    Ds[(size_t)i * nbf + j] = D[(size_t)i * nbf + j];
        }
    free(sidx);
    free(list);
    return Ds;
}

void orc_build_exchange(orc_rep *r, const double *D_in, double *K) {
    const int nbf = r->num_bf, N = r->N, nn = r->n_node, nl = r->n_leaf, np = r->num_sp;
    double *D = screened_density(r, D_in);
    /* basis-function pair of every row: (a, b), and whether the shells differ */
    int *ra = (int *)malloc(sizeof(int) * N), *rb = (int *)malloc(sizeof(int) * N);
    char *sw = (char *)malloc(N);
    for (int i = 0; i < np; i++) {
        const int s0 = r->sp_shell_idx[i], s1 = r->sp_shell_idx[i + np];
        const int f0 = r->shell_bf_sidx[s0], n0 = r->shell_bf_sidx[s0 + 1] - f0;
        const int f1 = r->shell_bf_sidx[s1], n1 = r->shell_bf_sidx[s1 + 1] - f1;
        for (int j = 0; j < n1; j++)
            for (int k = 0; k < n0; k++) {
                const int x = r->sp_bfp_sidx[i] + j * n0 + k;
                ra[x] = f0 + k;
                rb[x] = f1 + j;
                sw[x] = s0 != s1;
            }
    }
    /* expanded nested bases, leaves first */
    double **P = (double **)calloc(nn, sizeof(double *));
    for (int lv = r->max_level; lv >= 0; lv--) {
        const int *nodes = r->level_nodes + (size_t)lv * nl;
        for (int j = 0; j < r->level_n_node[lv]; j++) {
            const int node = nodes[j];
            const orc_blk *U = &r->U[node];
            const int k = U->ncol;
            if (k == 0 || U->nrow == 0) continue;
            const int s = r->mat_cluster[2 * node], m = r->mat_cluster[2 * node + 1] - s + 1;
            double *Pn = (double *)calloc((size_t)m * k, sizeof(double));
            if (r->n_child[node] == 0) {
                memcpy(Pn, U->data, sizeof(double) * (size_t)m * k);
            } else {
                int off = 0;
                for (int c = 0; c < r->n_child[node]; c++) {
                    const int ch = r->children[8 * node + c];
                    const int kc = r->U[ch].ncol;
                    if (kc == 0 || !P[ch]) { off += kc; continue; }
                    const int cs = r->mat_cluster[2 * ch] - s, cm = r->mat_cluster[2 * ch + 1] - r->mat_cluster[2 * ch] + 1;
                    for (int row = 0; row < cm; row++)
                        for (int l = 0; l < kc; l++) {
                            const double a = P[ch][(size_t)row * kc + l];
                            const double *urow = U->data + (size_t)(off + l) * k;
                            double *o = Pn + (size_t)(cs + row) * k;
                            for (int t = 0; t < k; t++) o[t] += a * urow[t];
                        }
                    off += kc;
                }
            }
            P[node] = Pn;
        }
    }
    /* terms */
    const int nterm = nl + r->n_inadm + r->n_adm;
    x_term *T = (x_term *)calloc(nterm, sizeof(x_term));
    int nt = 0;
    for (int i = 0; i < nl; i++) {
        const int node = r->height_nodes[i];
        const orc_blk *b = &r->D[i];
        if (b->nrow == 0 || b->ncol == 0) continue;
        const int s = r->mat_cluster[2 * node];
        T[nt++] = (x_term){s, b->nrow, b->nrow, NULL, s, b->ncol, b->ncol, NULL, b, 1};
    }
    for (int i = 0; i < r->n_inadm; i++) {
        const orc_blk *b = &r->D[nl + i];
        if (b->nrow == 0 || b->ncol == 0) continue;
        const int n0 = r->inadm[2 * i], n1 = r->inadm[2 * i + 1];
        T[nt++] = (x_term){r->mat_cluster[2 * n0], b->nrow, b->nrow, NULL, r->mat_cluster[2 * n1], b->ncol, b->ncol,
                           NULL, b, 0};
    }
    for (int i = 0; i < r->n_adm; i++) {
        const orc_blk *b = &r->B[i];
        if (b->nrow == 0 || b->ncol == 0) continue;
        const int n0 = r->adm[2 * i], n1 = r->adm[2 * i + 1];
        const int l0 = r->node_level[n0], l1 = r->node_level[n1];
        const int s0 = r->mat_cluster[2 * n0], m0 = r->mat_cluster[2 * n0 + 1] - s0 + 1;
        const int s1 = r->mat_cluster[2 * n1], m1 = r->mat_cluster[2 * n1 + 1] - s1 + 1;
        x_term t = {s0, m0, m0, NULL, s1, m1, m1, NULL, b, 0};
        if (l0 >= l1) { if (!P[n0]) continue; t.P = P[n0]; t.k0 = r->U[n0].ncol; }
        if (l1 >= l0) { if (!P[n1]) continue; t.Q = P[n1]; t.k1 = r->U[n1].ncol; }
        if (t.k0 != b->nrow || t.k1 != b->ncol) continue;   /* inconsistent block: not part of A_H2 */
        T[nt++] = t;
    }
    const int nth = x_threads();
    const size_t nn2 = (size_t)nbf * nbf;
    double *K1t = (double *)calloc(nn2 * nth, sizeof(double));
    double *Kdt = (double *)calloc(nn2 * nth, sizeof(double));
    int maxm = 1, maxk = 1;
    for (int i = 0; i < nt; i++) {
        if (T[i].m1 > maxm) maxm = T[i].m1;
        if (T[i].k1 > maxk) maxk = T[i].k1;
        if (T[i].C->ld < 0 && -T[i].C->ld > maxk) maxk = -T[i].C->ld;
    }
    const int CH = 32;
    #pragma omp parallel num_threads(nth)
    {
        const int tid = omp_get_thread_num();
        double *T1 = (double *)malloc(sizeof(double) * CH * (size_t)maxk);
        double *tmp = (double *)malloc(sizeof(double) * CH * (size_t)maxk);
        double *F = (double *)malloc(sizeof(double) * CH * (size_t)maxm);
        #pragma omp for schedule(dynamic)
        for (int it = 0; it < nt; it++) {
            const x_term *t = &T[it];
            double *Kt = (t->diag ? Kdt : K1t) + nn2 * tid;
            for (int r0 = 0; r0 < t->m0; r0 += CH) {
                const int h = (t->m0 - r0 < CH) ? t->m0 - r0 : CH;
                if (t->P) mul_core(h, t->k0, t->P + (size_t)r0 * t->k0, t->k0, t->C, T1, tmp);   /* T1 = P C */
                else core_rows(t->C, r0, h, T1);                                              /* T1 = C rows */
                if (t->Q) {                                          /* F = T1 Q^T (h x m1) */
                    for (int q = 0; q < h; q++)
                        for (int c = 0; c < t->m1; c++) {
                            const double *qrow = t->Q + (size_t)c * t->k1;
                            const double *w = T1 + (size_t)q * t->k1;
                            double s = 0.0;
                            for (int l = 0; l < t->k1; l++) s += w[l] * qrow[l];
                            F[(size_t)q * t->m1 + c] = s;
                        }
                } else {
                    memcpy(F, T1, sizeof(double) * (size_t)h * t->m1);
                }
                /* contraction over the orientations of rows and columns */
                for (int q = 0; q < h; q++) {
                    const int gr = t->s0 + r0 + q;
                    const int a = ra[gr], b = rb[gr];
                    const double *Db = D + (size_t)b * nbf, *Da = D + (size_t)a * nbf;
                    double *Ka = Kt + (size_t)a * nbf, *Kb = Kt + (size_t)b * nbf;
                    const double *Fr = F + (size_t)q * t->m1;
                    for (int c = 0; c < t->m1; c++) {
                        const int gc = t->s1 + c;
                        const int a2 = ra[gc], b2 = rb[gc];
                        const double v = Fr[c];
                        Ka[b2] += v * Db[a2];
                        if (sw[gc]) Ka[a2] += v * Db[b2];
                        if (sw[gr]) {
                            Kb[b2] += v * Da[a2];
                            if (sw[gc]) Kb[a2] += v * Da[b2];
                        }
                    }
                }
            }
        }
        free(T1); free(tmp); free(F);
    }
    memset(K, 0, sizeof(double) * nn2);
    for (int t = 0; t < nth; t++) {
        const double *k1 = K1t + nn2 * t, *kd = Kdt + nn2 * t;
        for (int i = 0; i < nbf; i++)
            for (int j = 0; j < nbf; j++)
                K[(size_t)i * nbf + j] += k1[(size_t)i * nbf + j] + k1[(size_t)j * nbf + i] + kd[(size_t)i * nbf + j];
    }
    free(K1t); free(Kdt); free(T);
    for (int i = 0; i < nn; i++) free(P[i]);
    free(P); free(ra); free(rb); free(sw);
    free(D);
}
