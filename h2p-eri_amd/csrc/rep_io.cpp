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
        if (!p) { ok = false; return nullptr; }
        get(p, n);
        return p;
    }
};

// Bounds of a loaded representation: counts large enough for any molecule
// this library builds and small enough that no size product overflows.
const int MAX_COUNT = 1 << 28;
const long long MAX_BLOCK_ELEMS = (1LL << 31) - 1;   // H2E_dense_mat.size is an int

bool in_range(long long v, long long lo, long long hi) { return v >= lo && v < hi; }

void put_blk(W &w, H2E_dense_mat_p b) {
    int nr = b ? b->nrow : 0, nc = b ? b->ncol : 0, ld = b ? b->ld : 0;
    w.i32(nr); w.i32(nc); w.i32(ld);
    if (b) w.put(b->data, blk_elems(b));
}

H2E_dense_mat_p get_blk(R &r) {
    int nr = r.i32(), nc = r.i32(), ld = r.i32();
    if (!r.ok || nr < 0 || nc < 0 || nr >= MAX_COUNT || nc >= MAX_COUNT || ld < -MAX_COUNT) { r.ok = false; return nullptr; }
    const long long elems = ld >= 0 ? (long long)nr * nc : (long long)(nr + nc) * (-(long long)ld);
    if (elems > MAX_BLOCK_ELEMS || (ld < 0 && (-ld > nr || -ld > nc))) { r.ok = false; return nullptr; }
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

// Header counts of a file: every count non-negative and bounded, a tree
// with at least one node and one leaf, at most 63 levels.
bool header_ok(const int *hdr) {
    for (int i = 0; i < 11; i++)
        if (!in_range(hdr[i], 0, MAX_COUNT)) return false;
    const int n_node = hdr[5], n_leaf = hdr[6], max_level = hdr[7], min_adm = hdr[8];
    return n_node >= 1 && in_range(n_leaf, 1, n_node + 1) && max_level < 63 && min_adm <= max_level + 1;
}

// Index arrays of a loaded tree: every stored index inside the array it
// addresses, shell-pair offsets monotone from 0 to krnl_mat_size, each leaf a
// non-empty contiguous range of shell pairs, and node lists naming nodes.
bool tree_ok(H2ERI_p h, bool chem) {
    const int nn = h->n_node, np = h->num_sp, nl = h->n_leaf_node, L = h->max_level + 1;
    if (h->sp_bfp_sidx[0] != 0 || h->sp_bfp_sidx[np] != h->krnl_mat_size) return false;
    for (int i = 0; i < np; i++)
        if (h->sp_bfp_sidx[i + 1] < h->sp_bfp_sidx[i]) return false;
    if (chem) {
        if (h->shell_bf_sidx[0] != 0) return false;
        for (int i = 0; i < h->nshell; i++)
            if (h->shell_bf_sidx[i + 1] < h->shell_bf_sidx[i]) return false;
        for (size_t i = 0; i < 2 * (size_t)np; i++)
            if (!in_range(h->sp_shell_idx[i], 0, h->nshell)) return false;
        // each shell pair's functions fill exactly its x range
        for (int i = 0; i < np; i++) {
            const int s0 = h->sp_shell_idx[i], s1 = h->sp_shell_idx[i + np];
            const long long nb = (long long)(h->shell_bf_sidx[s0 + 1] - h->shell_bf_sidx[s0]) *
                                 (h->shell_bf_sidx[s1 + 1] - h->shell_bf_sidx[s1]);
            if (nb != h->sp_bfp_sidx[i + 1] - h->sp_bfp_sidx[i]) return false;
        }
    }
    int leaves = 0, next_pair = 0;
    for (int i = 0; i < nn; i++) {
        if (!in_range(h->parent[i], -1, nn) || !in_range(h->n_child[i], 0, 9)) return false;
        if (!in_range(h->node_level[i], 0, L)) return false;
        for (int c = 0; c < h->n_child[i]; c++)
            if (!in_range(h->children[8 * (size_t)i + c], 0, i)) return false;   // post order: children first
        const int ps = h->pt_cluster[2 * i], pe = h->pt_cluster[2 * i + 1];
        if (!in_range(ps, 0, np) || !in_range(pe, ps, np)) return false;
        if (h->n_child[i] == 0) {
            // leaves in node order take consecutive shell-pair ranges that
            // cover every pair once (their x ranges then partition x)
            if (ps != next_pair) return false;
            next_pair = pe + 1;
            leaves++;
        }
    }
    if (leaves != nl || next_pair != np) return false;
    for (int l = 0; l < L; l++) {
        if (!in_range(h->level_n_node[l], 0, nl + 1) || !in_range(h->height_n_node[l], 0, nl + 1)) return false;
        for (int j = 0; j < h->level_n_node[l]; j++)
            if (!in_range(h->level_nodes[(size_t)l * nl + j], 0, nn)) return false;
        for (int j = 0; j < h->height_n_node[l]; j++)
            if (!in_range(h->height_nodes[(size_t)l * nl + j], 0, nn)) return false;
    }
    if (h->height_n_node[0] != nl) return false;
    for (int j = 0; j < nl; j++)
        if (h->n_child[h->height_nodes[j]] != 0) return false;
    for (size_t i = 0; i < 2 * (size_t)h->n_r_adm_pair; i++)
        if (!in_range(h->r_adm_pairs[i], 0, nn)) return false;
    for (size_t i = 0; i < 2 * (size_t)h->n_r_inadm_pair; i++)
        if (!in_range(h->r_inadm_pairs[i], 0, nn)) return false;
    for (int i = 0; i < h->n_r_inadm_pair; i++)   // near-field pairs join two leaves
        if (h->n_child[h->r_inadm_pairs[2 * i]] != 0 || h->n_child[h->r_inadm_pairs[2 * i + 1]] != 0) return false;
    return true;
}

int leaf_elems(H2ERI_p h, int node) {
    return h->sp_bfp_sidx[h->pt_cluster[2 * node + 1] + 1] - h->sp_bfp_sidx[h->pt_cluster[2 * node]];
}

// Block shapes against the tree: a leaf's U has one row per basis-function
// pair of the leaf, an inner node's U one row per column of its children's
// U; B and D blocks have the shapes the tree gives them.
bool blocks_ok(H2ERI_p h) {
    auto rows = [](H2E_dense_mat_p b) { return b ? b->nrow : 0; };
    auto cols = [](H2E_dense_mat_p b) { return b ? b->ncol : 0; };
    for (int i = 0; i < h->n_node; i++) {
        H2E_dense_mat_p u = h->U[i];
        if (!u) continue;
        if (u->ld < 0) return false;   // U blocks are dense
        if (cols(u) == 0) continue;
        int nr = 0;
        if (h->n_child[i] == 0) nr = leaf_elems(h, i);
        else
            for (int c = 0; c < h->n_child[i]; c++) nr += cols(h->U[h->children[8 * (size_t)i + c]]);
        if (rows(u) != nr) return false;
    }
    for (int i = 0; i < h->n_B; i++) {
        H2E_dense_mat_p b = h->c_B_blks[i];
        if (b && (b->nrow || b->ncol) && (b->nrow != h->B_nrow[i] || b->ncol != h->B_ncol[i])) return false;
    }
    for (int i = 0; i < h->n_D; i++) {
        H2E_dense_mat_p b = h->c_D_blks[i];
        if (b && (b->nrow || b->ncol) && (b->nrow != h->D_nrow[i] || b->ncol != h->D_ncol[i])) return false;
    }
    return true;
}

}  // namespace
}  // namespace h2e

using namespace h2e;

extern "C" {

int H2ERI_save_rep(H2ERI_p h, const char *fname) {
    if (!h || !h->U || !h->c_B_blks || !h->c_D_blks) { set_error(ERR_ARG, "save_rep: nothing built"); return -1; }
    if (h->BD_JIT) { set_error(ERR_UNSUPPORTED, "save_rep: a BD_JIT representation stores no B or D blocks"); return -1; }
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
    if (!r.ok || !header_ok(hdr)) {
        fclose(f);
        set_error(ERR_IO, "load_rep: %s has invalid header counts", fname);
        return -1;
    }
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
    if (r.ok && !tree_ok(h, (flags & 1u) != 0)) {
        fclose(f);
        set_error(ERR_IO, "load_rep: %s holds an index outside its range", fname);
        return -1;
    }
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
    if (!blocks_ok(h)) { set_error(ERR_IO, "load_rep: %s holds a block whose shape does not match the tree", fname); return -1; }
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
