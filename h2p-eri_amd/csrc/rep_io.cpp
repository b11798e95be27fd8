// Versioned binary serialisation of a built H2 representation.  H2P-ERI has
// no on-disk format (its struct is src/H2ERI_typedef.h:14-130); this format
// stores what the apply paths read:
//
//   char[8]  "H2ERIREP"      uint32 version (=1)     uint32 flags (bit0: chemistry metadata)
//   int32    natom nshell num_bf num_sp krnl_mat_size n_node n_leaf_node
//            max_level min_adm_level n_r_adm_pair n_r_inadm_pair
//   if chemistry: int32 shell_bf_sidx[nshell+1], sp_shell_idx[2*num_sp]
//   int32    sp_bfp_sidx[num_sp+1]
//   int32    parent, children(8/node), n_child, node_level, node_height,
//            pt_cluster(2/node), mat_cluster(2/node)            (per node)
//   int32    level_n_node[L], level_nodes[L*n_leaf], height_n_node[L],
//            height_nodes[L*n_leaf]                             (L = max_level+1)
//   int32    r_adm_pairs[2*n_adm], r_inadm_pairs[2*n_inadm]
//   blocks   U[n_node], B[n_adm], D[n_leaf+n_inadm]: int32 nrow, ncol, ld;
//            float64 data[elems] (dense: nrow*ncol; ld<0: (nrow+ncol)*(-ld))
//   char[8]  "H2ERIEND"
// All integers little-endian int32, all reals IEEE float64.
#include "h2eri_internal.h"

#include <algorithm>

namespace h2e {

namespace {

struct W {
    FILE *f;
    bool ok = true;
    template <class T> void put(const T *p, size_t n) { if (ok && n) ok = fwrite(p, sizeof(T), n, f) == n; }
    void i32(int v) { put(&v, 1); }
};
struct R {
    FILE *f;
    bool ok = true;
    template <class T> void get(T *p, size_t n) { if (ok && n) ok = fread(p, sizeof(T), n, f) == n; }
    int i32() { int v = 0; get(&v, 1); return v; }
    template <class T> T *arr(size_t n) {
        T *p = (T *)malloc(sizeof(T) * (n > 0 ? n : 1));
        get(p, n);
        return p;
    }
};

void put_blk(W &w, H2E_dense_mat_p b) {
    int nr = b ? b->nrow : 0, nc = b ? b->ncol : 0, ld = b ? b->ld : 0;
    w.i32(nr); w.i32(nc); w.i32(ld);
    if (b) w.put(b->data, blk_elems(b));
}

H2E_dense_mat_p get_blk(R &r) {
    int nr = r.i32(), nc = r.i32(), ld = r.i32();
    if (!r.ok || nr < 0 || nc < 0) { r.ok = false; return nullptr; }
    H2E_dense_mat_p b;
    if (ld >= 0) {
        b = dmat_new(nr, nc);
    } else {
        b = dmat_new(1, (nr + nc) * (-ld));
        b->nrow = nr; b->ncol = nc; b->ld = ld;
    }
    r.get(b->data, blk_elems(b));
    return b;
}

}  // namespace
}  // namespace h2e

using namespace h2e;

extern "C" {

int H2ERI_save_rep(H2ERI_p h, const char *fname) {
    if (!h || !h->U || !h->c_B_blks || !h->c_D_blks) { set_error(ERR_ARG, "save_rep: nothing built"); return -1; }
    FILE *f = fopen(fname, "wb");
    if (!f) { set_error(ERR_IO, "save_rep: cannot open %s", fname); return -1; }
    W w{f};
    const bool chem = h->shell_bf_sidx && h->sp_shell_idx && h->nshell > 0;
    w.put("H2ERIREP", 8);
    unsigned ver = 1, flags = chem ? 1u : 0u;
    w.put(&ver, 1); w.put(&flags, 1);
    int hdr[11] = {h->natom, h->nshell, h->num_bf, h->num_sp, h->krnl_mat_size, h->n_node, h->n_leaf_node,
                   h->max_level, h->min_adm_level, h->n_r_adm_pair, h->n_r_inadm_pair};
    w.put(hdr, 11);
    if (chem) {
        w.put(h->shell_bf_sidx, h->nshell + 1);
        w.put(h->sp_shell_idx, 2 * (size_t)h->num_sp);
    }
    w.put(h->sp_bfp_sidx, (size_t)h->num_sp + 1);
    const size_t nn = h->n_node, L = h->max_level + 1, nl = h->n_leaf_node;
    w.put(h->parent, nn); w.put(h->children, 8 * nn); w.put(h->n_child, nn);
    w.put(h->node_level, nn); w.put(h->node_height, nn);
    w.put(h->pt_cluster, 2 * nn); w.put(h->mat_cluster, 2 * nn);
    w.put(h->level_n_node, L); w.put(h->level_nodes, L * nl);
    w.put(h->height_n_node, L); w.put(h->height_nodes, L * nl);
    w.put(h->r_adm_pairs, 2 * (size_t)h->n_r_adm_pair);
    w.put(h->r_inadm_pairs, 2 * (size_t)h->n_r_inadm_pair);
    for (int i = 0; i < h->n_node; i++) put_blk(w, h->U[i]);
    for (int i = 0; i < h->n_B; i++) put_blk(w, h->c_B_blks[i]);
    for (int i = 0; i < h->n_D; i++) put_blk(w, h->c_D_blks[i]);
    w.put("H2ERIEND", 8);
    bool ok = w.ok;
    if (fclose(f) != 0) ok = false;
    if (!ok) { set_error(ERR_IO, "save_rep: write failed for %s", fname); return -1; }
    return 0;
}

int H2ERI_load_rep(H2ERI_p h, const char *fname) {
    if (!h) { set_error(ERR_ARG, "load_rep: null handle"); return -1; }
    if (h->n_node > 0 || h->U) { set_error(ERR_ARG, "load_rep: handle already holds a tree; use a fresh handle"); return -1; }
    FILE *f = fopen(fname, "rb");
    if (!f) { set_error(ERR_IO, "load_rep: cannot open %s", fname); return -1; }
    R r{f};
    char magic[8];
    r.get(magic, 8);
    unsigned ver = 0, flags = 0;
    r.get(&ver, 1); r.get(&flags, 1);
    if (!r.ok || memcmp(magic, "H2ERIREP", 8) != 0 || ver != 1) {
        fclose(f);
        set_error(ERR_IO, "load_rep: %s is not a version-1 representation", fname);
        return -1;
    }
    int hdr[11];
    r.get(hdr, 11);
    h->natom = hdr[0]; h->nshell = hdr[1]; h->num_bf = hdr[2]; h->num_sp = hdr[3];
    h->krnl_mat_size = h->num_sp_bfp = hdr[4]; h->n_node = hdr[5]; h->n_leaf_node = hdr[6];
    h->max_level = hdr[7]; h->min_adm_level = hdr[8]; h->n_r_adm_pair = hdr[9]; h->n_r_inadm_pair = hdr[10];
    h->root_idx = h->n_node - 1;
    h->n_point = h->num_sp;
    h->max_child = 8;
    if (flags & 1u) {
        h->shell_bf_sidx = r.arr<int>(h->nshell + 1);
        h->sp_shell_idx = r.arr<int>(2 * (size_t)h->num_sp);
    }
    h->sp_bfp_sidx = r.arr<int>((size_t)h->num_sp + 1);
    const size_t nn = h->n_node, L = h->max_level + 1, nl = h->n_leaf_node;
    h->parent = r.arr<int>(nn); h->children = r.arr<int>(8 * nn); h->n_child = r.arr<int>(nn);
    h->node_level = r.arr<int>(nn); h->node_height = r.arr<int>(nn);
    h->pt_cluster = r.arr<int>(2 * nn); h->mat_cluster = r.arr<int>(2 * nn);
    h->level_n_node = r.arr<int>(L); h->level_nodes = r.arr<int>(L * nl);
    h->height_n_node = r.arr<int>(L); h->height_nodes = r.arr<int>(L * nl);
    h->r_adm_pairs = r.arr<int>(2 * (size_t)h->n_r_adm_pair);
    h->r_inadm_pairs = r.arr<int>(2 * (size_t)h->n_r_inadm_pair);
    if (r.ok) {
        h->sp_nbfp = (int *)malloc(sizeof(int) * (h->num_sp > 0 ? h->num_sp : 1));
        for (int i = 0; i < h->num_sp; i++) h->sp_nbfp[i] = h->sp_bfp_sidx[i + 1] - h->sp_bfp_sidx[i];
        h->U = (H2E_dense_mat_p *)calloc(nn, sizeof(H2E_dense_mat_p));
        for (size_t i = 0; i < nn && r.ok; i++) h->U[i] = get_blk(r);
    }
    if (r.ok) {
        // block shapes, CSR maps and per-node lists follow from the tree
        build_B_D_index(h);
        h->c_B_blks = (H2E_dense_mat_p *)calloc(h->n_B > 0 ? h->n_B : 1, sizeof(H2E_dense_mat_p));
        for (int i = 0; i < h->n_B && r.ok; i++) h->c_B_blks[i] = get_blk(r);
        h->c_D_blks = (H2E_dense_mat_p *)calloc(h->n_D, sizeof(H2E_dense_mat_p));
        for (int i = 0; i < h->n_D && r.ok; i++) h->c_D_blks[i] = get_blk(r);
        char tail[8];
        r.get(tail, 8);
        if (r.ok && memcmp(tail, "H2ERIEND", 8) != 0) r.ok = false;
    }
    fclose(f);
    if (!r.ok) { set_error(ERR_IO, "load_rep: %s is truncated or corrupt", fname); return -1; }
    if (h->shell_bf_sidx) {
        h->num_bf = h->shell_bf_sidx[h->nshell];
        int mx = 0;
        for (int i = 0; i < h->nshell; i++) mx = std::max(mx, h->shell_bf_sidx[i + 1] - h->shell_bf_sidx[i]);
        h->max_shell_nbf = mx;
    }
    finish_partition_metadata(h);
    update_mat_sizes(h);
    return 0;
}

}  // extern "C"
