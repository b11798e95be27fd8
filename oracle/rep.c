/* Reader of the "H2ERIREP" version-1 representation file and the derived
 * index structures the reference builds during partition/construction.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Derived structures follow H2P-ERI:
 *   per-node admissible / inadmissible partner lists  src/H2ERI_partition.c:273-360
 *   plist (paired shells per shell, sorted)           src/H2ERI_partition.c:371-427
 *   B/D pair -> block CSR maps with transposes as -(i+1)
 *                                                     src/H2ERI_build_H2.c:988-1011, 1189-1260
 *   CSR lookup                                        src/H2ERI_utils.c:149-164
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

static int rd(FILE *f, void *p, size_t sz, size_t n) { return n == 0 || fread(p, sz, n, f) == n; }

static int *rdi(FILE *f, size_t n, int *ok) {
    int *p = (int *)malloc(sizeof(int) * (n ? n : 1));
    if (!rd(f, p, sizeof(int), n)) *ok = 0;
    return p;
}

static int rdblk(FILE *f, orc_blk *b) {
    int h[3];
    if (!rd(f, h, sizeof(int), 3)) return 0;
    b->nrow = h[0]; b->ncol = h[1]; b->ld = h[2];
    size_t n = b->ld >= 0 ? (size_t)b->nrow * b->ncol : (size_t)(b->nrow + b->ncol) * (size_t)(-b->ld);
    b->data = (double *)malloc(sizeof(double) * (n ? n : 1));
    return rd(f, b->data, sizeof(double), n);
}

orc_rep *orc_load(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    char magic[8];
    unsigned ver, flags;
    int ok = rd(f, magic, 1, 8) && rd(f, &ver, 4, 1) && rd(f, &flags, 4, 1);
    if (!ok || memcmp(magic, "H2ERIREP", 8) != 0 || ver != 1) { fclose(f); return NULL; }
    orc_rep *r = (orc_rep *)calloc(1, sizeof(orc_rep));
    int hdr[11];
    ok = rd(f, hdr, sizeof(int), 11);
    r->natom = hdr[0]; r->nshell = hdr[1]; r->num_bf = hdr[2]; r->num_sp = hdr[3]; r->N = hdr[4];
    r->n_node = hdr[5]; r->n_leaf = hdr[6]; r->max_level = hdr[7]; r->min_adm_level = hdr[8];
    r->n_adm = hdr[9]; r->n_inadm = hdr[10];
    if (flags & 1u) {
        r->shell_bf_sidx = rdi(f, (size_t)r->nshell + 1, &ok);
        r->sp_shell_idx = rdi(f, 2 * (size_t)r->num_sp, &ok);
    }
    r->sp_bfp_sidx = rdi(f, (size_t)r->num_sp + 1, &ok);
    size_t nn = r->n_node, L = r->max_level + 1, nl = r->n_leaf;
    r->parent = rdi(f, nn, &ok); r->children = rdi(f, 8 * nn, &ok); r->n_child = rdi(f, nn, &ok);
    r->node_level = rdi(f, nn, &ok); r->node_height = rdi(f, nn, &ok);
    r->pt_cluster = rdi(f, 2 * nn, &ok); r->mat_cluster = rdi(f, 2 * nn, &ok);
    r->level_n_node = rdi(f, L, &ok); r->level_nodes = rdi(f, L * nl, &ok);
    r->height_n_node = rdi(f, L, &ok); r->height_nodes = rdi(f, L * nl, &ok);
    r->adm = rdi(f, 2 * (size_t)r->n_adm, &ok);
    r->inadm = rdi(f, 2 * (size_t)r->n_inadm, &ok);
    r->U = (orc_blk *)calloc(nn ? nn : 1, sizeof(orc_blk));
    r->B = (orc_blk *)calloc(r->n_adm ? r->n_adm : 1, sizeof(orc_blk));
    r->D = (orc_blk *)calloc(nl + r->n_inadm, sizeof(orc_blk));
    for (size_t i = 0; ok && i < nn; i++) ok = rdblk(f, &r->U[i]);
    for (int i = 0; ok && i < r->n_adm; i++) ok = rdblk(f, &r->B[i]);
    for (size_t i = 0; ok && i < nl + (size_t)r->n_inadm; i++) ok = rdblk(f, &r->D[i]);
    char tail[8];
    if (ok) ok = rd(f, tail, 1, 8) && memcmp(tail, "H2ERIEND", 8) == 0;
    fclose(f);
    if (!ok) { orc_free(r); return NULL; }
    orc_derive(r);
    return r;
}

void orc_free(orc_rep *r) {
    if (!r) return;
    int *ip[] = {r->shell_bf_sidx, r->sp_shell_idx, r->sp_bfp_sidx, r->parent, r->children, r->n_child,
                 r->node_level, r->node_height, r->pt_cluster, r->mat_cluster, r->level_n_node, r->level_nodes,
                 r->height_n_node, r->height_nodes, r->adm, r->inadm, r->node_adm_sidx, r->node_adm,
                 r->node_inadm_sidx, r->node_inadm, r->B_rowptr, r->B_colidx, r->B_val, r->D_rowptr,
                 r->D_colidx, r->D_val, r->plist_sidx, r->plist, r->plist_idx};
    for (size_t i = 0; i < sizeof(ip) / sizeof(ip[0]); i++) free(ip[i]);
    if (r->U) for (int i = 0; i < r->n_node; i++) free(r->U[i].data);
    if (r->B) for (int i = 0; i < r->n_adm; i++) free(r->B[i].data);
    if (r->D) for (int i = 0; i < r->n_leaf + r->n_inadm; i++) free(r->D[i].data);
    free(r->U); free(r->B); free(r->D);
    free(r);
}

int orc_size(const orc_rep *r) { return r->N; }
int orc_nbf(const orc_rep *r) { return r->num_bf; }
int orc_nshell(const orc_rep *r) { return r->nshell; }

/* COO -> CSR with column-sorted rows (src/H2ERI_utils.c:121-146). */
static void coo_csr(int nrow, int nnz, const int *row, const int *col, const int *val, int **rp, int **ci, int **vv) {
    int *rowptr = (int *)calloc(nrow + 1, sizeof(int));
    int *colidx = (int *)malloc(sizeof(int) * (nnz ? nnz : 1));
    int *v = (int *)malloc(sizeof(int) * (nnz ? nnz : 1));
    for (int i = 0; i < nnz; i++) rowptr[row[i] + 1]++;
    for (int i = 0; i < nrow; i++) rowptr[i + 1] += rowptr[i];
    int *pos = (int *)malloc(sizeof(int) * (nrow ? nrow : 1));
    memcpy(pos, rowptr, sizeof(int) * nrow);
    for (int i = 0; i < nnz; i++) { int p = pos[row[i]]++; colidx[p] = col[i]; v[p] = val[i]; }
    for (int i = 0; i < nrow; i++)   /* insertion sort by column inside each row */
        for (int a = rowptr[i] + 1; a < rowptr[i + 1]; a++)
            for (int b = a; b > rowptr[i] && colidx[b - 1] > colidx[b]; b--) {
                int t = colidx[b]; colidx[b] = colidx[b - 1]; colidx[b - 1] = t;
                t = v[b]; v[b] = v[b - 1]; v[b - 1] = t;
            }
    free(pos);
    *rp = rowptr; *ci = colidx; *vv = v;
}

int orc_csr_get(const int *rowptr, const int *colidx, const int *val, int row, int col) {
    for (int i = rowptr[row]; i < rowptr[row + 1]; i++)
        if (colidx[i] == col) return val[i];
    return 0;
}

void orc_derive(orc_rep *r) {
    const int nn = r->n_node, nl = r->n_leaf;
    /* admissible partner lists */
    r->node_adm_sidx = (int *)calloc(nn + 1, sizeof(int));
    r->node_adm = (int *)malloc(sizeof(int) * (2 * r->n_adm + 1));
    for (int i = 0; i < r->n_adm; i++) { r->node_adm_sidx[r->adm[2 * i] + 1]++; r->node_adm_sidx[r->adm[2 * i + 1] + 1]++; }
    for (int i = 0; i < nn; i++) r->node_adm_sidx[i + 1] += r->node_adm_sidx[i];
    int *pos = (int *)malloc(sizeof(int) * (nn + 1));
    memcpy(pos, r->node_adm_sidx, sizeof(int) * nn);
    for (int i = 0; i < r->n_adm; i++) {
        int a = r->adm[2 * i], b = r->adm[2 * i + 1];
        r->node_adm[pos[a]++] = b;
        r->node_adm[pos[b]++] = a;
    }
    /* inadmissible partner lists: leaf self first */
    r->node_inadm_sidx = (int *)calloc(nn + 1, sizeof(int));
    r->node_inadm = (int *)malloc(sizeof(int) * (2 * r->n_inadm + nl + 1));
    for (int i = 0; i < nl; i++) r->node_inadm_sidx[r->height_nodes[i] + 1]++;
    for (int i = 0; i < r->n_inadm; i++) { r->node_inadm_sidx[r->inadm[2 * i] + 1]++; r->node_inadm_sidx[r->inadm[2 * i + 1] + 1]++; }
    for (int i = 0; i < nn; i++) r->node_inadm_sidx[i + 1] += r->node_inadm_sidx[i];
    memcpy(pos, r->node_inadm_sidx, sizeof(int) * nn);
    for (int i = 0; i < nl; i++) { int n = r->height_nodes[i]; r->node_inadm[pos[n]++] = n; }
    for (int i = 0; i < r->n_inadm; i++) {
        int a = r->inadm[2 * i], b = r->inadm[2 * i + 1];
        r->node_inadm[pos[a]++] = b;
        r->node_inadm[pos[b]++] = a;
    }
    free(pos);
    /* B / D CSR maps */
    int nb = 2 * r->n_adm, nd = nl + 2 * r->n_inadm;
    int *ri = (int *)malloc(sizeof(int) * (nb + nd + 1)), *cj = (int *)malloc(sizeof(int) * (nb + nd + 1));
    int *vv = (int *)malloc(sizeof(int) * (nb + nd + 1));
    for (int i = 0; i < r->n_adm; i++) {
        ri[2 * i] = r->adm[2 * i]; cj[2 * i] = r->adm[2 * i + 1]; vv[2 * i] = i + 1;
        ri[2 * i + 1] = r->adm[2 * i + 1]; cj[2 * i + 1] = r->adm[2 * i]; vv[2 * i + 1] = -(i + 1);
    }
    coo_csr(nn, nb, ri, cj, vv, &r->B_rowptr, &r->B_colidx, &r->B_val);
    int c = 0;
    for (int i = 0; i < nl; i++) { int n = r->height_nodes[i]; ri[c] = n; cj[c] = n; vv[c] = i + 1; c++; }
    for (int i = 0; i < r->n_inadm; i++) {
        ri[c] = r->inadm[2 * i]; cj[c] = r->inadm[2 * i + 1]; vv[c] = nl + i + 1; c++;
        ri[c] = r->inadm[2 * i + 1]; cj[c] = r->inadm[2 * i]; vv[c] = -(nl + i + 1); c++;
    }
    coo_csr(nn, nd, ri, cj, vv, &r->D_rowptr, &r->D_colidx, &r->D_val);
    free(ri); free(cj); free(vv);
    /* plist */
    if (r->sp_shell_idx && r->nshell > 0) {
        const int ns = r->nshell, np = r->num_sp;
        r->plist_sidx = (int *)calloc(ns + 1, sizeof(int));
        r->plist = (int *)malloc(sizeof(int) * (2 * np + 1));
        r->plist_idx = (int *)malloc(sizeof(int) * (2 * np + 1));
        for (int i = 0; i < np; i++) {
            int s0 = r->sp_shell_idx[i], s1 = r->sp_shell_idx[i + np];
            r->plist_sidx[s0 + 1]++;
            if (s0 != s1) r->plist_sidx[s1 + 1]++;
        }
        for (int i = 0; i < ns; i++) r->plist_sidx[i + 1] += r->plist_sidx[i];
        int *p2 = (int *)malloc(sizeof(int) * (ns + 1));
        memcpy(p2, r->plist_sidx, sizeof(int) * ns);
        for (int i = 0; i < np; i++) {
            int s0 = r->sp_shell_idx[i], s1 = r->sp_shell_idx[i + np];
            int k = p2[s0]++;
            r->plist[k] = s1; r->plist_idx[k] = i;
            if (s0 != s1) { k = p2[s1]++; r->plist[k] = s0; r->plist_idx[k] = i; }
        }
        free(p2);
        for (int s = 0; s < ns; s++)   /* sort each shell's list by partner shell */
            for (int a = r->plist_sidx[s] + 1; a < r->plist_sidx[s + 1]; a++)
                for (int b = a; b > r->plist_sidx[s] && r->plist[b - 1] > r->plist[b]; b--) {
                    int t = r->plist[b]; r->plist[b] = r->plist[b - 1]; r->plist[b - 1] = t;
                    t = r->plist_idx[b]; r->plist_idx[b] = r->plist_idx[b - 1]; r->plist_idx[b - 1] = t;
                }
    }
}
