// Lifecycle, error state, small containers and statistics output.
//
// Behaviour follows src/H2ERI_typedef.c:13-312 (init defaults, destroy,
// print_statistic, H2E_reset_timers) of H2P-ERI; the containers follow the
// semantics of src/H2ERI_aux_structs.h:150-199 (H2E_dense_mat, ld < 0 for
// low-rank blocks).
#include "h2eri_internal.h"

#include <cstdarg>
#include <omp.h>
#include <sys/time.h>

namespace h2e {

static thread_local int g_err = 0;
static thread_local char g_msg[512] = {0};

void set_error(int code, const char *fmt, ...) {
    g_err = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    fprintf(stderr, "[h2eri] error %d: %s\n", code, g_msg);
}
bool has_error() { return g_err != 0; }

double wtime() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return (double)tv.tv_sec + (double)tv.tv_usec * 1e-6;
}

H2E_dense_mat_p dmat_new(int nrow, int ncol) {
    auto *m = (H2E_dense_mat_p)calloc(1, sizeof(struct H2E_dense_mat));
    if (!m) { set_error(ERR_ALLOC, "dense mat struct"); return nullptr; }
    m->nrow = nrow > 0 ? nrow : 0;
    m->ncol = ncol > 0 ? ncol : 0;
    m->ld = m->ncol;
    m->size = m->nrow * m->ncol;
    if (m->size > 0) {
        m->data = (double *)malloc(sizeof(double) * (size_t)m->size);
        if (!m->data) set_error(ERR_ALLOC, "dense mat %d x %d", nrow, ncol);
    }
    return m;
}

void dmat_free(H2E_dense_mat_p &m) {
    if (!m) return;
    free(m->data);
    free(m);
    m = nullptr;
}

void dmat_resize(H2E_dense_mat_p m, int nrow, int ncol) {
    int ns = nrow * ncol;
    m->nrow = nrow;
    m->ncol = ncol;
    m->ld = ncol;
    if (ns > m->size) {
        free(m->data);
        m->data = (double *)malloc(sizeof(double) * (size_t)ns);
        if (!m->data) set_error(ERR_ALLOC, "dense mat %d x %d", nrow, ncol);
        m->size = ns;
    }
}

H2E_int_vec_p ivec_new(int cap) {
    auto *v = (H2E_int_vec_p)calloc(1, sizeof(struct H2E_int_vec));
    if (cap < 1) cap = 1;
    v->capacity = cap;
    v->data = (int *)malloc(sizeof(int) * (size_t)cap);
    return v;
}
void ivec_free(H2E_int_vec_p &v) {
    if (!v) return;
    free(v->data);
    free(v);
    v = nullptr;
}
void ivec_set_capacity(H2E_int_vec_p v, int cap) {
    if (cap <= v->capacity) return;
    v->data = (int *)realloc(v->data, sizeof(int) * (size_t)cap);
    v->capacity = cap;
}
void ivec_push(H2E_int_vec_p v, int x) {
    if (v->length >= v->capacity) ivec_set_capacity(v, 2 * v->capacity + 4);
    v->data[v->length++] = x;
}

static HostState *hs_get(H2ERI_p h) {
    // HostState is stored through the reference's unused `sp` slot.
    return (HostState *)h->sp;
}
HostState *host_state(H2ERI_p h) {
    if (!h->sp) h->sp = new HostState();
    return hs_get(h);
}
void host_state_free(H2ERI_p h) {
    HostState *hs = hs_get(h);
    if (!hs) return;
    for (auto *p : hs->pairs) pair_data_free(p);
    delete hs;
    h->sp = nullptr;
}

}  // namespace h2e

using namespace h2e;

extern "C" {

int h2eri_last_error(void) { return g_err; }
const char *h2eri_last_error_msg(void) { return g_msg; }
void h2eri_clear_error(void) { g_err = 0; g_msg[0] = 0; }

// src/H2ERI_typedef.c:13-78
void H2ERI_init(H2ERI_p *h2eri_, const double scr_tol, const double ext_tol, const double QR_tol) {
    if (!h2eri_) { set_error(ERR_ARG, "H2ERI_init: null handle pointer"); return; }
    auto h = (H2ERI_p)calloc(1, sizeof(struct H2ERI));
    if (!h) { set_error(ERR_ALLOC, "H2ERI_init"); *h2eri_ = nullptr; return; }
    h->scr_tol = scr_tol;
    h->ext_tol = ext_tol;
    h->pp_npts_layer = 384;
    h->pp_nlayer_ext = 3;
    h->n_thread = omp_get_max_threads();
    h->pt_dim = 3;
    h->xpt_dim = 3;
    h->krnl_dim = 1;
    h->max_child = 8;
    h->max_neighbor = 27;
    h->QR_stop_type = 1;  // relative column norm
    h->QR_stop_tol = QR_tol;
    const char *pt = getenv("H2E_PRINT_TIMERS");
    h->print_timers = (pt && atoi(pt) == 1) ? 1 : 0;
    const char *pd = getenv("H2E_PRINT_DBGINFO");
    h->print_dbginfo = (pd && atoi(pd) == 1) ? 1 : 0;
    *h2eri_ = h;
}

void CMS_destroy_shells(const int nshell, shell_t *shells) {
    if (!shells) return;
    for (int i = 0; i < nshell; i++) free_shell(shells[i]);
}

static void free_blk_array(H2E_dense_mat_p *&a, int n) {
    if (!a) return;
    for (int i = 0; i < n; i++) dmat_free(a[i]);
    free(a);
    a = nullptr;
}
static void free_ivec_array(H2E_int_vec_p *&a, int n) {
    if (!a) return;
    for (int i = 0; i < n; i++) ivec_free(a[i]);
    free(a);
    a = nullptr;
}

void H2ERI_destroy(H2ERI_p h) {
    if (!h) return;
    device_release(h);
    host_state_free(h);
    if (h->shells) { CMS_destroy_shells(h->nshell, h->shells); free(h->shells); }
    if (h->sp_shells) { CMS_destroy_shells(2 * h->num_sp, h->sp_shells); free(h->sp_shells); }
    int *iarrs[] = {h->shell_bf_sidx, h->sp_nbfp, h->sp_bfp_sidx, h->sp_shell_idx, h->index_seq,
                    h->node_adm_pairs, h->node_adm_pairs_sidx, h->node_inadm_pairs,
                    h->node_inadm_pairs_sidx, h->plist, h->plist_idx, h->plist_sidx, h->dlist,
                    h->dlist_sidx, h->parent, h->children, h->pt_cluster, h->mat_cluster,
                    h->n_child, h->node_level, h->node_height, h->level_n_node, h->level_nodes,
                    h->height_n_node, h->height_nodes, h->r_inadm_pairs, h->r_adm_pairs,
                    h->node_inadm_lists, h->node_n_r_inadm, h->node_n_r_adm, h->coord_idx,
                    h->B_p2i_rowptr, h->B_p2i_colidx, h->B_p2i_val, h->D_p2i_rowptr,
                    h->D_p2i_colidx, h->D_p2i_val, h->B_nrow, h->B_ncol, h->D_nrow, h->D_ncol};
    for (int *p : iarrs) free(p);
    free(h->B_ptr);
    free(h->D_ptr);
    double *darrs[] = {h->sp_center, h->sp_extent, h->box_extent, h->unc_denmat_x, h->H2_matvec_y,
                       h->coord, h->coord0, h->enbox, h->root_enbox};
    for (double *p : darrs) free(p);
    free_ivec_array(h->J_pair, h->n_node);
    free_ivec_array(h->J_row, h->n_node);
    free_ivec_array(h->ovlp_ff_idx, h->n_node);
    free_blk_array(h->c_B_blks, h->n_B);
    free_blk_array(h->c_D_blks, h->n_D);
    free_blk_array(h->U, h->n_node);
    free_blk_array(h->y0, h->n_node);
    free_blk_array(h->y1, h->n_node);
    free(h);
}

// src/H2ERI_typedef.c:303-312
void H2E_reset_timers(H2ERI_p h) {
    if (!h) return;
    h->n_matvec = 0;
    h->mat_size[MV_VOP_SIZE_IDX] = 0;
    for (int i = MV_FWD_TIMER_IDX; i <= MV_VOP_TIMER_IDX; i++) h->timers[i] = 0.0;
}

// src/H2ERI_typedef.c:148-300 (same sections and units; matvec phases are
// device phases measured with HIP events).
void H2ERI_print_statistic(H2ERI_p h) {
    if (!h) return;
    if (h->n_node == 0) { printf("H2ERI has nothing to report yet.\n"); return; }
    printf("================ H2ERI molecular system info ================\n");
    printf("  * Number of atoms / shells / basis functions : %d, %d, %d\n", h->natom, h->nshell, h->num_bf);
    printf("  * Number of symm-unique screened shell pairs : %d\n", h->num_sp);
    printf("==================== H2ERI H2 tree info ====================\n");
    printf("  * Number of points               : %d\n", h->n_point);
    printf("  * Kernel matrix size             : %d\n", h->krnl_mat_size);
    printf("  * Maximum points in a leaf node  : %d\n", h->max_leaf_points);
    printf("  * Maximum leaf node box size     : %e\n", h->max_leaf_size);
    printf("  * Number of levels (root at 0)   : %d\n", h->max_level + 1);
    printf("  * Number of nodes                : %d\n", h->n_node);
    printf("  * Number of nodes on each level  : ");
    for (int i = 0; i < h->max_level; i++) printf("%d, ", h->level_n_node[i]);
    printf("%d\n", h->level_n_node[h->max_level]);
    printf("  * Number of nodes on each height : ");
    for (int i = 0; i < h->max_level; i++) printf("%d, ", h->height_n_node[i]);
    printf("%d\n", h->height_n_node[h->max_level]);
    printf("  * Minimum admissible pair level  : %d\n", h->min_adm_level);
    printf("  * Number of reduced adm. pairs   : %d\n", h->n_r_adm_pair);
    printf("  * Number of reduced inadm. pairs : %d\n", h->n_r_inadm_pair);
    if (h->U == nullptr) { printf("H2ERI H2 matrix has not been constructed yet.\n"); return; }
    printf("==================== H2ERI storage info ====================\n");
    const double MB = 8.0 / 1048576.0;
    size_t *ms = h->mat_size;
    double fwd_MB = ms[MV_FWD_SIZE_IDX] * MB, mid_MB = ms[MV_MID_SIZE_IDX] * MB;
    double bwd_MB = ms[MV_BWD_SIZE_IDX] * MB, den_MB = ms[MV_DEN_SIZE_IDX] * MB;
    double mv_MB = fwd_MB + mid_MB + bwd_MB + den_MB;
    printf("  * H2 representation U, B, D     : %.2lf, %.2lf, %.2lf (MB) \n",
           ms[U_SIZE_IDX] * MB, ms[B_SIZE_IDX] * MB, ms[D_SIZE_IDX] * MB);
    int max_rank = 0;
    double sum_rank = 0, n_ne = 0;
    for (int i = 0; i < h->n_node; i++) {
        int r = h->U[i] ? h->U[i]->ncol : 0;
        if (r > 0) { sum_rank += r; n_ne += 1; if (r > max_rank) max_rank = r; }
    }
    printf("  * Max / Avg compressed rank     : %d, %.0lf \n", max_rank, n_ne > 0 ? sum_rank / n_ne : 0.0);
    printf("  * Just-In-Time B & D build      : %s\n", h->BD_JIT ? "Yes (B & D not allocated)" : "No");
    printf("==================== H2ERI timing info =====================\n");
    double *t = h->timers;
    printf("  * H2 construction time (sec)   = %.3lf \n", t[0] + t[1] + t[2] + t[3]);
    printf("      |----> Point partition     = %.3lf \n", t[PT_TIMER_IDX]);
    printf("      |----> U construction      = %.3lf \n", t[U_BUILD_TIMER_IDX]);
    printf("      |----> B construction      = %.3lf \n", t[B_BUILD_TIMER_IDX]);
    printf("      |----> D construction      = %.3lf \n", t[D_BUILD_TIMER_IDX]);
    if (h->n_matvec == 0) {
        printf("H2ERI does not have matvec timings results yet.\n");
    } else {
        double n = (double)h->n_matvec;
        double f = t[MV_FWD_TIMER_IDX] / n, m = t[MV_MID_TIMER_IDX] / n, b = t[MV_BWD_TIMER_IDX] / n;
        double d = t[MV_DEN_TIMER_IDX] / n, v = t[MV_VOP_TIMER_IDX] / n;
        double tot = f + m + b + d + v;
        auto gbs = [](double mb, double s) { return s > 0 ? mb / s / 1024.0 : 0.0; };
        printf("  * H2 matvec average time (sec) = %.6lf, %.2lf GB/s\n", tot, gbs(mv_MB, tot));
        printf("      |----> Forward transformation      = %.6lf, %.2lf GB/s\n", f, gbs(fwd_MB, f));
        if (h->BD_JIT) printf("      |----> Intermediate multiplication = %.6lf\n", m);
        else printf("      |----> Intermediate multiplication = %.6lf, %.2lf GB/s\n", m, gbs(mid_MB, m));
        printf("      |----> Backward transformation     = %.6lf, %.2lf GB/s\n", b, gbs(bwd_MB, b));
        if (h->BD_JIT) printf("      |----> Dense multiplication        = %.6lf\n", d);
        else printf("      |----> Dense multiplication        = %.6lf, %.2lf GB/s\n", d, gbs(den_MB, d));
        printf("      |----> Vector operations           = %.6lf\n", v);
    }
    printf("=============================================================\n");
}

}  // extern "C"
