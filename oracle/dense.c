/* Dense brute-force assembly of A_H2 -- TEST INFRASTRUCTURE ONLY.
 *
 * Independent restatement of what the H2 representation means, used to pin
 * the sweep-based orc_matvec:
 *   Ubig_leaf = U_leaf;  Ubig_n = blkdiag(Ubig_children) U_n
 *   admissible (n0, n1), same level : A[n0,n1] += Ubig_n0 B Ubig_n1^T
 *                        l0 > l1    : A[n0,n1] += Ubig_n0 B       (n1 a leaf)
 *                        l0 < l1    : A[n0,n1] += B Ubig_n1^T     (n0 a leaf)
 *                        and A[n1,n0] += (the same block)^T
 *   inadmissible (n0,n1)            : A[n0,n1] += D, A[n1,n0] += D^T
 *   leaf diagonal                   : A[n,n] += D_nn (no transposed half,
 *                                     src/H2ERI_matvec.c:597-617)
 * Row/column ranges are mat_cluster (src/H2ERI_partition.c:233-265).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

static double *expand_blk(const orc_blk *b) {   /* dense nrow x ncol copy */
    double *M = (double *)calloc((size_t)b->nrow * b->ncol + 1, sizeof(double));
    if (b->ld >= 0) {
        memcpy(M, b->data, sizeof(double) * (size_t)b->nrow * b->ncol);
    } else {
        const int k = -b->ld;
        const double *U = b->data, *VT = U + (size_t)b->nrow * k;
        for (int i = 0; i < b->nrow; i++)
            for (int j = 0; j < b->ncol; j++) {
                double s = 0.0;
                for (int l = 0; l < k; l++) s += U[(size_t)i * k + l] * VT[(size_t)j * k + l];
                M[(size_t)i * b->ncol + j] = s;
            }
    }
    return M;
}

/* C (m x n) = A (m x k) B (k x n), row-major */
static void mm(int m, int k, int n, const double *A, const double *B, double *C) {
    memset(C, 0, sizeof(double) * (size_t)m * n);
    for (int i = 0; i < m; i++)
        for (int l = 0; l < k; l++) {
            const double a = A[(size_t)i * k + l];
            for (int j = 0; j < n; j++) C[(size_t)i * n + j] += a * B[(size_t)l * n + j];
        }
}

static void add_block(double *A, int N, int r0, int c0, int m, int n, const double *blk) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            A[(size_t)(r0 + i) * N + c0 + j] += blk[(size_t)i * n + j];
        }
}

static void add_block_t(double *A, int N, int r0, int c0, int m, int n, const double *blk) {
    /* A[r0.., c0..] += blk^T, blk is m x n */
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) A[(size_t)(r0 + j) * N + c0 + i] += blk[(size_t)i * n + j];
}

void orc_dense_matrix(orc_rep *r, double *A) {
    const int N = r->N, nn = r->n_node;
    memset(A, 0, sizeof(double) * (size_t)N * N);
    double **Ub = (double **)calloc(nn, sizeof(double *));
    /* post-order numbering: children precede parents */
    for (int n = 0; n < nn; n++) {
        const orc_blk *U = &r->U[n];
        if (U->ncol == 0) continue;
        const int len = r->mat_cluster[2 * n + 1] - r->mat_cluster[2 * n] + 1;
        Ub[n] = (double *)calloc((size_t)len * U->ncol + 1, sizeof(double));
        if (r->n_child[n] == 0) {
            memcpy(Ub[n], U->data, sizeof(double) * (size_t)len * U->ncol);
        } else {
            int srow = 0;
            for (int k = 0; k < r->n_child[n]; k++) {
                const int ch = r->children[8 * n + k];
                const int rc = r->U[ch].ncol;
                const int clen = r->mat_cluster[2 * ch + 1] - r->mat_cluster[2 * ch] + 1;
                const int off = r->mat_cluster[2 * ch] - r->mat_cluster[2 * n];
                if (rc > 0)
                    mm(clen, rc, U->ncol, Ub[ch], U->data + (size_t)srow * U->ncol, Ub[n] + (size_t)off * U->ncol);
                srow += rc;
            }
        }
    }
    for (int i = 0; i < r->n_adm; i++) {
        const int n0 = r->adm[2 * i], n1 = r->adm[2 * i + 1];
        const int l0 = r->node_level[n0], l1 = r->node_level[n1];
        const orc_blk *b = &r->B[i];
        if (b->nrow == 0 || b->ncol == 0) continue;
        const int s0 = r->mat_cluster[2 * n0], s1 = r->mat_cluster[2 * n1];
        const int len0 = r->mat_cluster[2 * n0 + 1] - s0 + 1, len1 = r->mat_cluster[2 * n1 + 1] - s1 + 1;
        double *Bd = expand_blk(b);
        double *L = NULL, *blk = (double *)malloc(sizeof(double) * (size_t)len0 * len1 + 1);
        if (l0 >= l1) {   /* rows through Ubig_n0 */
            L = (double *)malloc(sizeof(double) * (size_t)len0 * b->ncol + 1);
            mm(len0, b->nrow, b->ncol, Ub[n0], Bd, L);
        } else {
            L = Bd;
            Bd = NULL;
        }
        if (l1 >= l0) {   /* columns through Ubig_n1^T */
            const int r1 = b->ncol;
            double *UT = (double *)malloc(sizeof(double) * (size_t)r1 * len1 + 1);
            for (int a = 0; a < len1; a++)
                for (int c = 0; c < r1; c++) UT[(size_t)c * len1 + a] = Ub[n1][(size_t)a * r1 + c];
            mm(len0, r1, len1, L, UT, blk);
            free(UT);
        } else {
            memcpy(blk, L, sizeof(double) * (size_t)len0 * len1);
        }
        add_block(A, N, s0, s1, len0, len1, blk);
        add_block_t(A, N, s1, s0, len0, len1, blk);
        free(blk); free(L); free(Bd);
    }
    for (int i = 0; i < r->n_leaf; i++) {
        const int n = r->height_nodes[i];
        const int s = r->mat_cluster[2 * n];
        double *Dd = expand_blk(&r->D[i]);
        add_block(A, N, s, s, r->D[i].nrow, r->D[i].ncol, Dd);
        free(Dd);
    }
    for (int i = 0; i < r->n_inadm; i++) {
        const int n0 = r->inadm[2 * i], n1 = r->inadm[2 * i + 1];
        const orc_blk *b = &r->D[r->n_leaf + i];
        double *Dd = expand_blk(b);
        add_block(A, N, r->mat_cluster[2 * n0], r->mat_cluster[2 * n1], b->nrow, b->ncol, Dd);
        add_block_t(A, N, r->mat_cluster[2 * n1], r->mat_cluster[2 * n0], b->nrow, b->ncol, Dd);
        free(Dd);
    }
    for (int n = 0; n < nn; n++) free(Ub[n]);
    free(Ub);
}
