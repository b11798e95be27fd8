/*
 * CPU oracle for the H2 J/K apply path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C99 + OpenMP restatement of H2P-ERI's algorithm, used by tests/
 * as the parity checker and by bench.py as the timed CPU baseline.  Nothing
 * in the product library links, calls or includes this code.
 *
 * It reads a representation written by H2ERI_save_rep (format "H2ERIREP"
 * version 1) with its own reader, and follows:
 *   orc_matvec          src/H2ERI_matvec.c:24-63, 93-407, 572-757, 882-957
 *   orc_build_coulomb   src/H2ERI_build_Coulomb.c:26-152
 *   orc_build_exchange  src/H2ERI_build_exchange.c:1356-1450 (same quantity, block by block)
 *   orc_dlist           src/H2ERI_build_exchange.c:774-838 (density screening list)
 *   orc_dense_matrix    (brute force) A_H2 assembled block by block from the
 *                       expanded nested bases, independent of the sweeps
 * Parity pins: see the header comment of each source file.
 */
#ifndef H2ERI_ORACLE_H
#define H2ERI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_rep orc_rep;

orc_rep *orc_load(const char *path);
void orc_free(orc_rep *r);
int orc_size(const orc_rep *r);      /* krnl_mat_size */
int orc_nbf(const orc_rep *r);       /* num_bf (0 for synthetic representations) */
int orc_nshell(const orc_rep *r);
void orc_set_threads(int n);

/* y := A_H2 x (restatement of H2ERI_matvec, thread-private partial sums). */
void orc_matvec(orc_rep *r, const double *x, double *y);

/* J := contract(A_H2 uncontract(D)). */
void orc_build_coulomb(orc_rep *r, const double *D, double *J);

/* K := sum_{n,p} (mn|pq) D_np through the H2 representation. */
void orc_build_exchange(orc_rep *r, const double *D, double *K);
/* density screening list of src/H2ERI_build_exchange.c:774-838 (DTOL 1e-10):
 * sidx[nshell + 1], list (may be NULL: count only); returns its length */
int orc_dlist(const orc_rep *r, const double *D, int *sidx, int *list);

/* Dense brute force: writes A_H2 (krnl_mat_size^2, row-major) by assembling
 * every admissible block Ubig_i B Ubig_j^T from the expanded nested bases
 * (Ubig_i = blkdiag(Ubig_children) U_i) and every near-field block, each with
 * its transpose.  Independent of the matvec's sweep structure; small trees. */
void orc_dense_matrix(orc_rep *r, double *out);

#ifdef __cplusplus
}
#endif

#endif
