/*
 * h2eri.h -- C-ABI of the MI355X-native H2 ERI library (libh2eri_amd.so).
 *
 * Drop-in boundary for H2P-ERI's J/K apply path.  Every entry point below
 * keeps the reference's name, argument meaning and `void` return; the
 * reference declaration it replaces is cited beside it (paths relative to the
 * H2P-ERI source tree).  Entry points without a counterpart in the reference
 * are marked "additive".
 *
 * Construction (screening, partition, proxy-point ID, B/D blocks) runs on the
 * host in C++; the U, B and D blocks are then uploaded once into flat device
 * arrays and every matvec runs as hand-written gfx950 HIP kernels.
 *
 * Errors: the reference aborts through assert().  This library keeps the void
 * signatures; a failing call records an error code and message, returns
 * without touching its outputs beyond what it already wrote, and the caller
 * may query h2eri_last_error()/h2eri_last_error_msg().  There is no CPU
 * fallback: a device error is reported, never silently replaced.
 */
#ifndef H2ERI_AMD_H
#define H2ERI_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of Cartesian components of a shell with angular momentum am
 * (reference: src/CMS.h:25). */
#define NCART(am) ((am >= 0) ? ((((am) + 2) * ((am) + 1)) >> 1) : 0)

/* Contracted Gaussian shell.  Field names and meaning follow the fields of
 * the SIMINT shell struct that H2P-ERI reads (src/CMS.c:91-108,
 * src/H2ERI_shell_operations.c:159-171): after normalisation `coef[k]`
 * already contains the primitive and contraction normalisation, so the x^am
 * Cartesian component is  sum_k coef[k] * x^am * exp(-alpha[k] r^2). */
typedef struct h2eri_shell {
    int     am;
    int     nprim;
    double  x, y, z;
    double *alpha;
    double *coef;
} shell_t;

/* Dense matrix as used by the reference (src/H2ERI_aux_structs.h:150-157).
 * ld > 0: dense row-major nrow x ncol.  ld < 0: low-rank block M = U*V of
 * rank k = -ld stored as [U (nrow x k) | V^T (ncol x k)], both row-major
 * (src/H2ERI_build_H2.c:876-901). */
typedef struct H2E_dense_mat {
    int     nrow;
    int     ncol;
    int     ld;
    int     size;
    double *data;
} *H2E_dense_mat_p;

typedef struct H2E_int_vec {
    int  capacity;
    int  length;
    int *data;
} *H2E_int_vec_p;

/* Library state.  The leading fields keep the names, order and meaning of
 * src/H2ERI_typedef.h:14-130 for everything a caller of the reference reads
 * or writes (natom, nshell, shells, num_bf, n_matvec, mat_size, timers; see
 * tests/test_J_input_mat.c:22,34,47-48).  SIMINT work buffers of the
 * reference are replaced by opaque pointers; device state is appended at the
 * end (the library allocates the struct, so appending is ABI-safe). */
struct H2ERI {
    int    natom;
    int    nshell;
    int    max_am;
    int    max_shell_nbf;
    int    num_bf;
    int    num_sp;
    int    num_sp_bfp;
    int    pp_npts_layer;
    int    pp_nlayer_ext;
    int   *shell_bf_sidx;
    int   *sp_nbfp;
    int   *sp_bfp_sidx;
    int   *sp_shell_idx;
    int   *index_seq;
    int   *node_adm_pairs;
    int   *node_adm_pairs_sidx;
    int   *node_inadm_pairs;
    int   *node_inadm_pairs_sidx;
    int   *plist;
    int   *plist_idx;
    int   *plist_sidx;
    int   *dlist;
    int   *dlist_sidx;
    void **thread_Kmat_workbuf;
    double scr_tol;
    double ext_tol;
    double *sp_center;
    double *sp_extent;
    double *box_extent;
    double *unc_denmat_x;
    double *H2_matvec_y;
    shell_t *shells;
    shell_t *sp_shells;
    void    *sp;                    /* reference: SIMINT shell pairs; unused */
    H2E_int_vec_p *J_pair;
    H2E_int_vec_p *J_row;
    H2E_int_vec_p *ovlp_ff_idx;
    void    *simint_buffs;          /* reference: SIMINT buffers; unused */
    void    *eri_batch_buffs;       /* reference: ERI batch buffers; unused */
    H2E_dense_mat_p *c_B_blks;
    H2E_dense_mat_p *c_D_blks;
    void    *thread_buffs;

    int    n_thread;
    int    pt_dim;
    int    xpt_dim;
    int    krnl_dim;
    int    QR_stop_type;
    int    QR_stop_rank;
    int    n_point;
    int    krnl_mat_size;
    int    max_leaf_points;
    int    n_node;
    int    root_idx;
    int    n_leaf_node;
    int    max_child;
    int    max_neighbor;
    int    max_level;
    int    min_adm_level;
    int    n_r_inadm_pair;
    int    n_r_adm_pair;
    int    n_UJ;
    int    n_B;
    int    n_D;
    int    BD_JIT;
    int    print_timers;
    int    print_dbginfo;
    int   *parent;
    int   *children;
    int   *pt_cluster;
    int   *mat_cluster;
    int   *n_child;
    int   *node_level;
    int   *node_height;
    int   *level_n_node;
    int   *level_nodes;
    int   *height_n_node;
    int   *height_nodes;
    int   *r_inadm_pairs;
    int   *r_adm_pairs;
    int   *node_inadm_lists;
    int   *node_n_r_inadm;
    int   *node_n_r_adm;
    int   *coord_idx;
    int   *B_p2i_rowptr;
    int   *B_p2i_colidx;
    int   *B_p2i_val;
    int   *D_p2i_rowptr;
    int   *D_p2i_colidx;
    int   *D_p2i_val;
    int   *B_nrow;
    int   *B_ncol;
    int   *D_nrow;
    int   *D_ncol;
    size_t *B_ptr;
    size_t *D_ptr;
    double max_leaf_size;
    double QR_stop_tol;
    double *coord;
    double *coord0;
    double *enbox;
    double *root_enbox;
    double *B_data;
    double *D_data;
    void   *B_blk;
    void   *D_blk0;
    void   *D_blk1;
    H2E_dense_mat_p *U;
    H2E_dense_mat_p *y0;
    H2E_dense_mat_p *y1;
    int    n_matvec;
    size_t mat_size[11];
    double timers[11];

    /* ---- appended: MI355X device state (opaque) ---- */
    void  *dev;
};
typedef struct H2ERI *H2ERI_p;

/* mat_size[] / timers[] indices (src/H2ERI_typedef.h:134-158). */
enum { U_SIZE_IDX = 0, B_SIZE_IDX, D_SIZE_IDX, MV_FWD_SIZE_IDX, MV_MID_SIZE_IDX,
       MV_BWD_SIZE_IDX, MV_DEN_SIZE_IDX, MV_VOP_SIZE_IDX };
enum { PT_TIMER_IDX = 0, U_BUILD_TIMER_IDX, B_BUILD_TIMER_IDX, D_BUILD_TIMER_IDX,
       MV_FWD_TIMER_IDX, MV_MID_TIMER_IDX, MV_BWD_TIMER_IDX, MV_DEN_TIMER_IDX,
       MV_VOP_TIMER_IDX };

/* ------------------------------------------------------------------ */
/* Lifecycle (replaces src/H2ERI_typedef.h:169-179, src/H2ERI_typedef.c:303) */
void H2ERI_init(H2ERI_p *h2eri_, const double scr_tol, const double ext_tol, const double QR_tol);
void H2ERI_destroy(H2ERI_p h2eri);
void H2ERI_print_statistic(H2ERI_p h2eri);
void H2E_reset_timers(H2ERI_p h2eri);

/* Molecule input (replaces src/CMS.h:49): reads the .mol format, normalises
 * the shells. */
void CMS_read_mol_file(const char *mol_fname, int *natom_, int *nshell_, shell_t **shells_);
void CMS_destroy_shells(const int nshell, shell_t *shells);

/* Host construction (replace src/H2ERI_shell_operations.h:25,
 * src/H2ERI_partition.h:28, src/H2ERI_build_H2.h:16).  BD_JIT = 0 stores the
 * B and D blocks; BD_JIT = 1 (src/H2ERI_build_H2.c:1406-1411) stores U, the
 * skeleton lists and the block shapes only, and every matvec evaluates the B
 * and D blocks from the shell pairs on the device (shells up to d; the
 * reference's src/H2ERI_matvec.c:411-568,761-879).  A BD_JIT handle serves
 * H2ERI_matvec[_multi][_dev] and H2ERI_build_Coulomb[_multi][_dev] in FP64;
 * the entry points that plan stored blocks (graph capture, autotune, groups,
 * the allocation-free matvec) and H2ERI_build_exchange refuse it with an
 * error, as the reference's exchange does (src/H2ERI_build_exchange.c:1358). */
void H2ERI_process_shells(H2ERI_p h2eri);
void H2ERI_partition(H2ERI_p h2eri);
void H2ERI_build_H2(H2ERI_p h2eri, const int BD_JIT);

/* B and D blocks stored dense (on = 0) or compressed to [U | V^T] where it
 * pays (on = 1, the default; src/H2ERI_build_H2.c:843-903), from the next
 * H2ERI_build_H2 on.  Dense blocks are exactly what a BD_JIT matvec
 * evaluates.  No counterpart in the reference. */
void H2ERI_set_block_compression(H2ERI_p h2eri, const int on);

/* ------------------------------------------------------------------ */
/* Hot path.  Each call runs on the device; host pointers in, host pointers
 * out, exactly as the reference.                                       */

/* y := A_H2 x, x and y of length krnl_mat_size
 * (replaces src/H2ERI_matvec.h:16). */
void H2ERI_matvec(H2ERI_p h2eri, const double *x, double *y);

/* J := contract(A_H2 uncontract(D)), D and J num_bf x num_bf row-major
 * (replaces src/H2ERI_build_Coulomb.h:16). */
void H2ERI_build_Coulomb(H2ERI_p h2eri, const double *den_mat, double *J_mat);

/* K_mq := sum_{n,p} (mn|pq) D_np through the H2 representation, D symmetric,
 * D and K num_bf x num_bf row-major (replaces src/H2ERI_build_exchange.h:16).
 * Accumulates with FP64 atomics (as the reference's scatter does), so the
 * last bits of K may differ between calls. */
void H2ERI_build_exchange(H2ERI_p h2eri, const double *den_mat, double *K_mat);
/* src/H2ERI_build_exchange.h:19-22: init builds and uploads the exchange plan
 * now (otherwise the first K build does); free releases its device memory. */
void H2ERI_exchange_workbuf_init(H2ERI_p h2eri);
void H2ERI_exchange_workbuf_free(H2ERI_p h2eri);
/* The density screening list of src/H2ERI_build_exchange.c:774-838 on the
 * host: fills h2eri->dlist_sidx (nshell + 1) and h2eri->dlist with, for each
 * shell N, the shells P whose block (N,P) of D has some |D_np| > 1e-10.
 * H2ERI_build_exchange calls it (the reference's side effect on the
 * handle); every exchange build, host or device pointers, sums over those
 * blocks of D only (the device applies the same rule to its copy of D). */
void H2ERI_exchange_dlist(H2ERI_p h2eri, const double *den_mat);

/* Additive: device-resident exchange build (den_dev, K_dev: num_bf^2 doubles
 * on the library's device; enqueued on `stream`, no synchronisation), and the
 * work figures of the exchange plan: out4 = {GEMM flops done once at upload,
 * number of block terms, device bytes resident, scratch doubles (the basis
 * products plus the intermediate G of one chunk of terms)}. */
void H2ERI_build_exchange_dev(H2ERI_p h2eri, const double *den_dev, double *K_dev, void *stream);
void H2ERI_exchange_stats(H2ERI_p h2eri, double *out4);
/* Additive (tuning): bound in bytes of the exchange build's intermediate G
 * per chunk (default 2 GiB).  The block terms run in chunks whose G fits it,
 * each chunk's ket contraction then its bra contraction; with two or more
 * chunks, odd chunks run on a second stream in a second scratch region, so a
 * chunk's ket overlaps the previous chunk's bra (scratch: twice the bound).
 * A term larger than the bound is a chunk of its own.  Releases the exchange
 * plan and an exchange graph. */
void H2ERI_set_exchange_scratch(H2ERI_p h2eri, const double bytes);
/* Additive: the exchange plan built and bounds-checked on the host only (no
 * HIP call); out4 as above with out4[2] the bytes it would make resident.
 * Returns the number of block terms, -1 on failure. */
int H2ERI_exchange_plan_check(H2ERI_p h2eri, double *out4);

/* ------------------------------------------------------------------ */
/* Additive entry points (no counterpart in the reference).             */

/* Last error code (0 = none) and message of the calling thread. */
int         h2eri_last_error(void);
const char *h2eri_last_error_msg(void);
void        h2eri_clear_error(void);

/* Storage precision of the far-field U and B blocks on the device:
 * 64 (default, the drop-in) or 32 (FP32-stored U/B, FP64 D and FP64
 * accumulation).  Re-uploads U/B if the representation is resident. */
void H2ERI_set_precision(H2ERI_p h2eri, const int ub_bits);

/* Y := A_H2 X for nvec vectors stored column after column (X[k*N + i],
 * N = krnl_mat_size).  Matrix blocks are streamed once per call. */
void H2ERI_matvec_multi(H2ERI_p h2eri, const int nvec, const double *X, double *Y);

/* J_k := contract(A_H2 uncontract(D_k)) for ndens densities stored one
 * after another (each num_bf^2). */
void H2ERI_build_Coulomb_multi(H2ERI_p h2eri, const int ndens, const double *den_mats, double *J_mats);

/* Device-resident forms: x_dev / y_dev (and D/J) are device pointers on the
 * library's current device; the call enqueues on `stream` (a hipStream_t,
 * NULL = default stream) and does not synchronise. */
void H2ERI_matvec_dev(H2ERI_p h2eri, const double *x_dev, double *y_dev, void *stream);
void H2ERI_matvec_multi_dev(H2ERI_p h2eri, const int nvec, const double *X_dev, double *Y_dev, void *stream);
void H2ERI_build_Coulomb_dev(H2ERI_p h2eri, const double *den_dev, double *J_dev, void *stream);
void H2ERI_build_Coulomb_multi_dev(H2ERI_p h2eri, const int ndens, const double *den_dev, double *J_dev,
                                   void *stream);

/* No counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): H2ERI_matvec_dev that allocates nothing.  The
 * workspace of the pass (and the persistent executor's counters) lives in a
 * file-scope device array of 128 MiB that each GPU the library runs on holds
 * until the process ends; a representation whose single-vector workspace
 * needs more is rejected with an error.  The handle must be prepared by
 * H2ERI_upload after its last change of precision, executor, tile size or
 * schedule (the call fails with an error otherwise), and its device must be
 * current.  Every allocation-free call on one GPU shares that array: calls
 * that overlap on one GPU race.  Results equal H2ERI_matvec_dev's bitwise;
 * H2ERI_device_status reports a failed wait of the last such pass. */
void H2ERI_matvec_dev_noalloc(H2ERI_p h2eri, const double *x_dev, double *y_dev, void *stream);

/* HIP-graph form of the device hot path, for loops that repeat it on the
 * same buffers (SCF iterations).  kind 0 captures H2ERI_matvec_multi_dev
 * with n vectors, kind 1 H2ERI_build_Coulomb_multi_dev with n densities,
 * kind 2 H2ERI_build_exchange_dev (n = 1, the density screening included),
 * on the device buffers in_dev/out_dev; nothing runs during the capture
 * (plans and uploads happen before it).  H2ERI_exchange_workbuf_free
 * releases an exchange graph.
 * H2ERI_graph_replay enqueues the captured sequence on `stream`.  One graph
 * per handle; changing precision, tile size, schedule or executor, running
 * the autotuner, or capturing again releases it (a replay after that fails
 * with the error set).  A Coulomb graph owns its x, y and half-matrix
 * buffers, so eager calls of any size between replays leave it valid.
 * Return 0, or -1 with the error set. */
int H2ERI_graph_capture(H2ERI_p h2eri, const int kind, const int n, const double *in_dev, double *out_dev);
int H2ERI_graph_replay(H2ERI_p h2eri, void *stream);

/* Upload the stored representation to the device now (otherwise done by the
 * first hot-path call) and prepare everything a single-vector pass needs
 * besides its vectors (plan, streams, events, the allocation-free entry
 * point's scratch address).  Returns bytes resident on the device. */
size_t H2ERI_upload(H2ERI_p h2eri);

/* Versioned binary serialisation of a built representation (the
 * reference has none).  Load replaces any construction. */
int H2ERI_save_rep(H2ERI_p h2eri, const char *fname);
int H2ERI_load_rep(H2ERI_p h2eri, const char *fname);

/* Fill a partitioned handle with seeded synthetic (chemistry-free) U, B and
 * D blocks instead of ERIs: for kernel tests and benchmarks of tree shapes
 * independent of integrals.  lr_frac in [0,1] is the fraction of B and
 * off-diagonal D blocks stored low-rank, diag_lr_frac that of the diagonal
 * D blocks (dense diagonal blocks are symmetric). */
void H2ERI_build_H2_synthetic(H2ERI_p h2eri, const unsigned long long seed,
                              const int max_rank, const double lr_frac,
                              const double diag_lr_frac);
/* Partition a point cloud directly (synthetic trees): npts shell pairs at
 * coord (3 x npts, column-major as the reference's sp_center), with nbfp[i]
 * basis-function pairs each. */
void H2ERI_partition_points_synthetic(H2ERI_p h2eri, const int npts, const double *coord,
                                      const int *nbfp, const int max_leaf_points);

/* Additive measurement aid (host, OpenMP): exact rows of the ERI matrix
 * times x for npair shell pairs (indices in tree order, x in tree order):
 * the rows of pair_idx[k] (sp_nbfp of it) are written one pair after another
 * into y, each the sum over every kept shell pair j of (i|j) x_j with every
 * integral evaluated directly.  Compared with the same rows of an H2 matvec
 * it gives the compression error.  Needs a molecule processed in this
 * process (H2ERI_process_shells + H2ERI_partition). */
void H2ERI_direct_matvec_rows(H2ERI_p h2eri, const int npair, const int *pair_idx, const double *x, double *y);

/* Per-phase device timing of the last hot-path call, milliseconds, by the
 * phase each launch belongs to: [0] upward U^T sweep, [1] B coupling (first
 * two launches), [2] B coupling (last launch), [3] downward U sweep, [4]
 * near-field-only launches, [5] output gather, [6] total.  Near-field tiles
 * share the sweep launches and count in their phase.  Collected only when
 * H2ERI_set_phase_timing(h2eri, 1). */
void H2ERI_set_phase_timing(H2ERI_p h2eri, const int on);
void H2ERI_get_phase_ms(H2ERI_p h2eri, double *ms7);
/* Per-launch records of the last timed pass: for launch i, out5[5i..5i+4] =
 * phase (numbered as above), start in ms after an event recorded before the
 * pass's first launch (launches on the two streams overlap; the union of
 * the intervals [start, start + ms] is the pass's kernel time), workgroups,
 * matrix bytes streamed, ms.  Returns the number of launches (fills at most
 * max_launch). */
int H2ERI_get_launch_stats(H2ERI_p h2eri, const int max_launch, double *out5);

/* Algorithmic bytes moved by one single-vector matvec on the device with
 * the current precision (U twice, B once, D once with packed diagonal
 * blocks, vectors), and the reference-accounting bytes
 * 8*(MV_FWD+MV_MID+MV_BWD+MV_DEN) of src/H2ERI_build_H2.c:801-805,1129-1130,
 * 1369-1381.  out[0]=alg, out[1]=ref-accounting, out[2..5]=U,B,D,vector
 * bytes of the algorithmic count. */
void H2ERI_matvec_bytes(H2ERI_p h2eri, double *out6);

/* Tile size of the device bi-GEMV tasks in bytes (tuning; re-plans the
 * device layout on the next call). */
void H2ERI_set_tile_bytes(H2ERI_p h2eri, const double bytes);

/* Near-field schedule (tuning; re-plans): 0 balances the near-field tiles
 * over the launches of the sweeps, 1 gives them three launches of their own
 * ahead of the sweeps, 2 + g (g = 0..255) runs them on a second stream
 * concurrently with the sweeps as g + 1 byte-balanced groups of blocks
 * pipelined over g + 3 launches (launch k: first pass of group k, second of
 * group k - 1, third of group k - 2). */
void H2ERI_set_schedule(H2ERI_p h2eri, const int mode);

/* Matvec executor (re-plans; no reference counterpart): 0 launches every
 * stage of the plan separately, 1 runs the whole pass as one persistent
 * launch (plus a small counter-reset launch) whose workgroups take the plan's
 * tasks by ticket in a topological order and wait only for the producers of
 * their own inputs.  The two agree to rounding (the staged plan runs small
 * matrix parts as wave tiles, which sum columns in another order).  With 1, phase
 * timing sees one launch per pass: its time goes to the near-field slot of
 * H2ERI_get_phase_ms and of the reference's timers. */
void H2ERI_set_executor(H2ERI_p h2eri, const int mode);
int  H2ERI_get_executor(H2ERI_p h2eri);   /* of single-vector passes */

/* Synchronises the device and reports whether a dependency wait of the
 * persistent executor exceeded its bound (about a second) since the last
 * call; such a pass's results are invalid.  0 clean, 1 with the error set,
 * -1 on a device error.  Clears the flag.  The synchronous entry points
 * (H2ERI_matvec[_multi], H2ERI_build_Coulomb[_multi]) read the flag
 * themselves and set ERR_DEVICE for such a pass; device-pointer forms and
 * graph replays leave it to this call. */
int H2ERI_device_status(H2ERI_p h2eri);
/* Bound of every dependency wait of the persistent executor, in polls
 * (default 2^22, about a second; additive, for tests of the error path: a
 * tiny bound makes waits give up, every wave still exits and the pass is
 * reported as failed). */
void H2ERI_set_spin_limit(H2ERI_p h2eri, const int polls);

/* Queue of the nv-vector plan for the persistent executor (built when
 * needed): out4 = {tasks, tasks started without a ticket, workgroups,
 * list-scheduling estimate of the pass in microseconds}.  0, -1 on error. */
int H2ERI_queue_stats(H2ERI_p h2eri, const int nv, double *out4);

/* Molecule groups (uneven molecules in one persistent launch; no reference
 * counterpart): H2ERI_group_create plans nh uploaded representations together
 * for fixed device vectors (y_k := A_k x_k, one vector each) and orders all
 * their tasks in one queue; H2ERI_group_matvec enqueues one matvec of every
 * molecule on `stream` (a counter reset and one persistent launch;
 * graph-capturable); H2ERI_group_status synchronises and reports a dependency
 * wait past its bound (0 clean, 1 error set, -1 device error);
 * H2ERI_group_destroy frees the group.  The handles must outlive the group
 * and keep their precision. */
void *H2ERI_group_create(const int nh, H2ERI_p *handles, const double *const *X_dev, double *const *Y_dev);
int   H2ERI_group_matvec(void *group, void *stream);
int   H2ERI_group_status(void *group);
void  H2ERI_group_destroy(void *group);
/* Mode of H2ERI_group_matvec (additive): 0 (default) one grouped persistent
 * launch, 1 each molecule's own pass in turn with its handle's configuration
 * (faster when every molecule fills the GPU alone).  H2ERI_group_autotune
 * times both (reps calls each) and keeps the faster: out3 = {ms of mode 0,
 * ms of mode 1, mode kept}.  Both return 0, -1 on error. */
int   H2ERI_group_set_mode(void *group, const int mode);
int   H2ERI_group_autotune(void *group, const int reps, double *out3);

/* Autotuner: times nv-vector device matvecs (reps calls after a warm-up)
 * for every executor (H2ERI_set_executor) x tile size {64, 128, 256, 512} KB x
 * near-field schedule {0, 2, 3, 4, 6, 10} x (staged executor) wave-tile limit {0,
 * 32} KB, then, when the staged executor with a second stream wins, its
 * near-field launches capped at 3/4 and 1/2 of the workgroups the GPU holds
 * (H2ERI_set_near_field_grid), and keeps the fastest combination for later
 * passes of nv vectors (each batch width 1, 2, 4, 8, 16 has a configuration
 * of its own; the tuning setters above change all five).  out6 = {ms per
 * call, tile bytes, schedule, executor, wave-tile bytes, near-field grid} of
 * the combination kept.  Returns 0, -1 on error. */
int H2ERI_autotune(H2ERI_p h2eri, const int nv, const int reps, double *out6);

/* Staged executor: matrix parts up to `bytes` (default 32768) run as wave
 * tiles, one wave each and four per workgroup, which pays off for the many
 * small coupling blocks of globular molecules; 0 disables them (tuning;
 * re-plans; no reference counterpart). */
void H2ERI_set_wave_tile_bytes(H2ERI_p h2eri, const double bytes);

/* Workgroups of each near-field launch of the staged executor on its second
 * stream (tuning; no re-plan; releases a captured graph): 0 (default) gives
 * every task a workgroup of its own, n > 0 at most n workgroups looping over
 * the launch's tasks, so that the sweeps' launches beside them keep CUs. */
void H2ERI_set_near_field_grid(H2ERI_p h2eri, const int n);

/* Builds the device execution plan on the host only (no HIP call) and
 * checks every access of every launch against its buffer.  Returns the
 * number of tile tasks, -1 on failure. */
int H2ERI_plan_check(H2ERI_p h2eri, const int ub_bits);

/* Work figures of the nv-vector plan with the current configuration, built
 * on the host only (no HIP call): out8 = {tile tasks, reduce tasks,
 * stages, workspace elements per vector, matrix bytes streamed per pass,
 * workspace elements written per vector, workspace elements read per
 * vector, output elements}.  Returns 0, -1 on failure. */
int H2ERI_plan_stats(H2ERI_p h2eri, const int ub_bits, const int nv, double *out8);

/* Host verification of the device plan (no HIP call, tests only): executes
 * every launch of the plan for nv (1, 2, 4, 8 or 16) vectors X (nv x N,
 * vector-major) with the kernels' semantics into Y and returns the number of
 * intra-launch read/write conflicts plus output elements not written exactly
 * once (0 expected), -1 on failure.  The hot-path entry points never call it. */
int H2ERI_plan_emulate(H2ERI_p h2eri, const int ub_bits, const int nv, const double *X, double *Y);

/* Host execution of a BD_JIT matvec (no HIP call, tests only): the device's
 * row-task code on one lane, the sweeps as plain loops; y = A x.  Returns 0,
 * -1 on failure. */
int H2ERI_jit_emulate(H2ERI_p h2eri, const double *x, double *y);

/* Work of one BD_JIT matvec, counted on the host: out4 = {row tasks, shell
 * quartets, primitive quartets, integral values}.  Returns 0, -1 on failure. */
int H2ERI_jit_stats(H2ERI_p h2eri, double *out4);

/* Host verification of the persistent executor's queue (no HIP call, tests
 * only): builds the queue for `workgroups` resident workgroups, executes the
 * tasks in ticket order with the kernels' semantics into Y, and returns the
 * number of workspace reads not covered by the counters the reading task
 * waits for, reads of unwritten elements and output elements not written
 * exactly once (0 expected), -1 on failure. */
int H2ERI_queue_emulate(H2ERI_p h2eri, const int ub_bits, const int nv, const int workgroups, const double *X,
                        double *Y);

/* sizeof(struct H2ERI) as compiled into the library (ABI check for
 * bindings that mirror the struct). */
size_t h2eri_struct_size(void);

#ifdef __cplusplus
}
#endif

#endif /* H2ERI_AMD_H */
