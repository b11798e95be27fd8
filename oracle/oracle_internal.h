/* Internal types of the CPU oracle (test infrastructure only). */
#ifndef H2ERI_ORACLE_INTERNAL_H
#define H2ERI_ORACLE_INTERNAL_H

#include <stddef.h>

#include "h2eri_oracle.h"

/* A stored block: ld > 0 dense nrow x ncol (row-major, ld == ncol);
 * ld < 0 low-rank [U (nrow x k) | V^T (ncol x k)], k = -ld. */
typedef struct {
    int nrow, ncol, ld;
    double *data;
} orc_blk;

struct orc_rep {
    int natom, nshell, num_bf, num_sp, N;
    int n_node, n_leaf, max_level, min_adm_level, n_adm, n_inadm;
    int *shell_bf_sidx, *sp_shell_idx, *sp_bfp_sidx;
    int *parent, *children, *n_child, *node_level, *node_height, *pt_cluster, *mat_cluster;
    int *level_n_node, *level_nodes, *height_n_node, *height_nodes;
    int *adm, *inadm;
    orc_blk *U, *B, *D;
    /* derived (reference: src/H2ERI_partition.c:273-427, src/H2ERI_build_H2.c:988-1011,1214-1260) */
    int *node_adm_sidx, *node_adm;       /* per node: admissible partners */
    int *node_inadm_sidx, *node_inadm;   /* per node: itself (leaf) + inadmissible partners */
    int *B_rowptr, *B_colidx, *B_val;    /* (n0,n1) -> +(i+1), (n1,n0) -> -(i+1) */
    int *D_rowptr, *D_colidx, *D_val;
    int *plist_sidx, *plist, *plist_idx; /* per shell: paired shells (sorted) and pair index */
};

void orc_derive(orc_rep *r);
int orc_csr_get(const int *rowptr, const int *colidx, const int *val, int row, int col);

#endif
