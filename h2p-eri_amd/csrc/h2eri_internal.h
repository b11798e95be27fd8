// Internal helpers shared by the host construction code and the device layer.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "h2eri.h"

#define H2E_ALPHA_H2  0.999999   // admissibility coefficient (src/H2ERI_config.h:95)
#define H2E_ALPHA_SUP 1.0        // super-cell coefficient (src/H2ERI_typedef.h:160)

namespace h2e {

// ---- error state (thread local) -------------------------------------------
void set_error(int code, const char *fmt, ...);
bool has_error();

enum ErrCode {
    ERR_NONE = 0,
    ERR_ARG = 1,      // invalid argument / state
    ERR_IO = 2,       // file I/O
    ERR_ALLOC = 3,    // host allocation
    ERR_DEVICE = 4,   // HIP runtime / device
    ERR_UNSUPPORTED = 5
};

double wtime();

// ---- small containers mirroring the reference structs ----------------------
H2E_dense_mat_p dmat_new(int nrow, int ncol);
void dmat_free(H2E_dense_mat_p &m);
void dmat_resize(H2E_dense_mat_p m, int nrow, int ncol);
H2E_int_vec_p ivec_new(int cap);
void ivec_free(H2E_int_vec_p &v);
void ivec_set_capacity(H2E_int_vec_p v, int cap);
void ivec_push(H2E_int_vec_p v, int x);

// number of doubles stored in a block (dense nrow*ncol, low-rank (nrow+ncol)*k)
inline size_t blk_elems(const H2E_dense_mat_p b) {
    if (b == nullptr) return 0;
    if (b->ld >= 0) return (size_t)b->nrow * (size_t)b->ncol;
    return (size_t)(b->nrow + b->ncol) * (size_t)(-b->ld);
}

// ---- construction pieces (host) -----------------------------------------------
void normalize_shells(int nshell, shell_t *shells);
void copy_shell(const shell_t &src, shell_t &dst);
void free_shell(shell_t &s);

// ERI engine: computes the Cartesian integrals (NM|QP) of shell pair i=(M,N)
// (bra) against shell pair j=(P,Q) (ket), in the layout of
// src/H2ERI_partition.c:92-96: row = n*nbf(M)+m, col = q*nbf(P)+p.
struct PairData;   // precomputed primitive-pair data of a shell pair
struct EriWork;    // per-thread scratch of the ERI engine
EriWork *eri_work_new(int max_am);
void eri_work_free(EriWork *w);
void pair_data_build(const shell_t &M, const shell_t &N, PairData *&pd);
void pair_data_free(PairData *&pd);
int pair_nbfp(const PairData *pd);
// Writes nbfp(bra) x nbfp(ket) values into out with leading dimension ldo.
void eri_pair_block(const PairData *bra, const PairData *ket, EriWork *w, double *out, int ldo);
// Potential integrals (n m | 1/|r-C|) of pair i against unit charges at
// npt points; out is nbfp x npt with leading dimension ldo.
void nai_pair_block(const PairData *pd, int npt, const double *px, const double *py,
                    const double *pz, EriWork *w, double *out, int ldo);
double boys_f(int m, double T);   // exposed for tests
void boys_fm(int mmax, double T, double *F);
// Read-only view of a pair's primitive data (first shell A = N, second B = M):
// np kept primitive pairs, exponent sums p[k], centres P[3k..3k+2] and
// prefactors K[k] = c_a c_b exp(-ab/p |AB|^2).
struct PairView {
    int la, lb, np;
    const double *A, *B, *p, *P, *K;
};
PairView pair_view(const PairData *pd);
// The Boys-function table of boys_fm: ngrid points of spacing step up to
// tmax, ncol orders per point (row-major [grid][m]).
const double *boys_table(int *ngrid, int *ncol, double *step, double *tmax);

// Pivoted-QR interpolative decomposition over rows (A ~= U * A(J,:)).
// A is nrow x ncol row-major and is destroyed.  Returns U (nrow x r,
// row-major, dense) and sorted skeleton rows J.  tol is relative.
void id_compress_rows(int nrow, int ncol, double *A, double tol,
                      std::vector<double> &U, std::vector<int> &J);

// Host-side per-handle construction state that is not part of the C struct.
struct HostState {
    std::vector<PairData *> pairs;   // per screened shell pair (sorted order)
    int max_am = 0;
    bool synthetic = false;
    bool bd_dense = false;   // H2ERI_set_block_compression(h, 0): B and D stored dense
    double build_ms[4] = {0, 0, 0, 0};
};
HostState *host_state(H2ERI_p h);
void host_state_free(H2ERI_p h);

void partition_tree(H2ERI_p h, int npts, double *coord_in, int max_leaf_points, double max_leaf_size);
void finish_partition_metadata(H2ERI_p h);   // box extent, mat_cluster, adm lists, plist
void build_B_D_index(H2ERI_p h);             // B/D shapes, CSR maps, node_n_r_adm
void update_mat_sizes(H2ERI_p h);

// device layer (device.hip)
void device_release(H2ERI_p h);
void device_invalidate(H2ERI_p h);
// releases the handle's captured graph if it is of the given kind (2: exchange)
void device_release_graph_kind(H2ERI_p h, int kind);
// on-the-fly B/D layer (jit.hip): device state of a BD_JIT handle
struct JitState;
bool jit_upload(H2ERI_p h, JitState *&js);   // builds the task lists and uploads them
void jit_release(JitState *&js);
size_t jit_resident_bytes(const JitState *js);
// y = A x for nvec vectors of length krnl_mat_size (device pointers) on stream s;
// with ms5, waits for the work and adds its milliseconds per phase: up sweep,
// coupling (B) tasks, down sweep, near-field (D) tasks, vector work
bool jit_matvec(H2ERI_p h, JitState *js, int nvec, const double *X, double *Y, void *stream, double *ms5);
// exchange layer (exchange.hip): device state held in h->thread_Kmat_workbuf
void exchange_release(H2ERI_p h);
bool exchange_prepare(H2ERI_p h);   // plan, upload and basis products (allocates)
// enqueues one K build on `stream` (a hipStream_t) without allocating
bool exchange_enqueue(H2ERI_p h, const double *D_dev, double *K_dev, void *stream);

}  // namespace h2e
