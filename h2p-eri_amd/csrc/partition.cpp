// Hierarchical partition of shell-pair centres and the tree metadata the
// matvec consumes.
//
// Behaviour follows H2P-ERI:
//   - bisection octree, post-order node numbering, leaf rule
//     (npts <= max_leaf_points or box width <= max_leaf_size), root-box
//     centre shift with srand48(19241112)       src/H2ERI_partition_points.c:36-265
//   - tree -> arrays (parent, children, level/height lists in post order)
//                                               src/H2ERI_partition_points.c:272-311
//   - reduced (in)admissible pairs, min_adm_level
//                                               src/H2ERI_partition_points.c:321-426, 521-527
//   - box admissibility test                    src/H2ERI_utils.c:11-26
//   - box extent, mat_cluster, per-node pair lists, plist
//                                               src/H2ERI_partition.c:152-427
#include "h2eri_internal.h"

#include <algorithm>
#include <cmath>

namespace h2e {

namespace {

struct TNode {
    int pt_s, pt_e;
    int n_child = 0;
    TNode *child[8] = {nullptr};
    int n_node = 1;
    int po_idx = 0;
    int level = 0, height = 0;
    double enbox[6];
};

struct PartCtx {
    int n_point;
    double *coord;       // 3 x n_point column-major, permuted in place
    double *coord_tmp;
    int *coord_idx;
    int *coord_idx_tmp;
    int max_leaf_points;
    double max_leaf_size;
    int curr_po = 0;
    int n_leaf = 0;
    int max_level = 0;
};

TNode *bisect(PartCtx &c, int level, int s, int e, const double *enbox_in) {
    int npts = e - s + 1;
    if (level > c.max_level) c.max_level = level;
    double enbox[6];
    if (enbox_in == nullptr) {
        double center[3] = {0, 0, 0};
        for (int j = 0; j < 3; j++) {
            const double *cj = c.coord + (size_t)j * c.n_point;
            for (int i = s; i <= e; i++) center[j] += cj[i];
            center[j] /= (double)npts;
        }
        double semi = 0.0;
        for (int j = 0; j < 3; j++) {
            const double *cj = c.coord + (size_t)j * c.n_point;
            for (int i = s; i <= e; i++) semi = std::max(semi, std::fabs(cj[i] - center[j]));
        }
        semi += 1e-8;
        // Small fixed-seed shift of the root centre against symmetric inputs.
        double shift_scale = semi * 1e-3;
        srand48(19241112);
        for (int j = 0; j < 3; j++) center[j] += (drand48() * 0.5 + 0.25) * shift_scale;
        semi = 0.0;
        for (int j = 0; j < 3; j++) {
            const double *cj = c.coord + (size_t)j * c.n_point;
            for (int i = s; i <= e; i++) semi = std::max(semi, std::fabs(cj[i] - center[j]));
        }
        semi += 1e-8;
        for (int j = 0; j < 3; j++) {
            enbox[j] = center[j] - semi - 2e-12;
            enbox[3 + j] = 2 * semi + 4e-12;
        }
    } else {
        memcpy(enbox, enbox_in, sizeof(enbox));
    }
    TNode *node = new TNode();
    node->pt_s = s;
    node->pt_e = e;
    node->level = level;
    memcpy(node->enbox, enbox, sizeof(enbox));
    if (npts <= c.max_leaf_points || enbox[3] <= c.max_leaf_size) {
        node->po_idx = c.curr_po++;
        c.n_leaf++;
        return node;
    }
    // Octant of each point: bit j set when the point is in the upper half of dim j.
    std::vector<int> rel(3 * (size_t)npts), cidx(npts, 0);
    for (int j = 0; j < 3; j++) {
        const double *cj = c.coord + (size_t)j * c.n_point + s;
        for (int i = 0; i < npts; i++) {
            int r = (int)std::floor(2.0 * (cj[i] - enbox[j]) / enbox[3 + j]);
            if (r == 2) r = 1;
            rel[(size_t)j * npts + i] = r;
            cidx[i] += r << j;
        }
    }
    int cnt[8] = {0}, sub_rel[24] = {0}, displ[9];
    for (int i = 0; i < npts; i++) {
        cnt[cidx[i]]++;
        for (int j = 0; j < 3; j++) sub_rel[j * 8 + cidx[i]] = rel[(size_t)j * npts + i];
    }
    displ[0] = 0;
    for (int k = 0; k < 8; k++) displ[k + 1] = displ[k] + cnt[k];
    for (int j = 0; j < 3; j++)
        memcpy(c.coord_tmp + (size_t)j * c.n_point + s, c.coord + (size_t)j * c.n_point + s, sizeof(double) * npts);
    memcpy(c.coord_idx_tmp + s, c.coord_idx + s, sizeof(int) * npts);
    int pos[8];
    for (int k = 0; k < 8; k++) pos[k] = displ[k];
    for (int i = 0; i < npts; i++) {
        int dst = s + pos[cidx[i]]++;
        for (int j = 0; j < 3; j++) c.coord[(size_t)j * c.n_point + dst] = c.coord_tmp[(size_t)j * c.n_point + s + i];
        c.coord_idx[dst] = c.coord_idx_tmp[s + i];
    }
    int max_child_h = 0;
    for (int k = 0; k < 8; k++) {
        if (cnt[k] == 0) continue;
        double sub[6];
        for (int j = 0; j < 3; j++) {
            sub[j] = enbox[j] + 0.5 * enbox[3 + j] * sub_rel[j * 8 + k] - 1e-12;
            sub[3 + j] = 0.5 * enbox[3 + j] + 2e-12;
        }
        TNode *ch = bisect(c, level + 1, s + displ[k], s + displ[k + 1] - 1, sub);
        node->child[node->n_child++] = ch;
        node->n_node += ch->n_node;
        max_child_h = std::max(max_child_h, ch->height);
    }
    node->height = max_child_h + 1;
    node->po_idx = c.curr_po++;
    return node;
}

void tree_to_arrays(TNode *nd, H2ERI_p h) {
    for (int i = 0; i < nd->n_child; i++) tree_to_arrays(nd->child[i], h);
    int id = nd->po_idx;
    int *ch = h->children + (size_t)id * 8;
    for (int i = 0; i < nd->n_child; i++) {
        ch[i] = nd->child[i]->po_idx;
        h->parent[ch[i]] = id;
    }
    for (int i = nd->n_child; i < 8; i++) ch[i] = -1;
    h->pt_cluster[2 * id] = nd->pt_s;
    h->pt_cluster[2 * id + 1] = nd->pt_e;
    memcpy(h->enbox + 6 * (size_t)id, nd->enbox, sizeof(double) * 6);
    h->node_level[id] = nd->level;
    h->node_height[id] = nd->height;
    h->n_child[id] = nd->n_child;
    h->level_nodes[nd->level * h->n_leaf_node + h->level_n_node[nd->level]++] = id;
    h->height_nodes[nd->height * h->n_leaf_node + h->height_n_node[nd->height]++] = id;
}

void free_tree(TNode *nd) {
    for (int i = 0; i < nd->n_child; i++) free_tree(nd->child[i]);
    delete nd;
}

bool box_admissible(const double *b0, const double *b1, double alpha) {
    for (int i = 0; i < 3; i++) {
        double r0 = b0[3 + i], r1 = b1[3 + i];
        double c0 = b0[i] + 0.5 * r0, c1 = b1[i] + 0.5 * r1;
        double min_r = std::min(r0, r1);
        if (std::fabs(c0 - c1) >= alpha * min_r + 0.5 * (r0 + r1)) return true;
    }
    return false;
}

struct PairLists {
    std::vector<int> adm, inadm;
    int min_adm_level;
};

void reduced_pairs(H2ERI_p h, int n0, int n1, PairLists &pl) {
    const int *nch = h->n_child;
    const int *ch = h->children;
    if (n0 == n1) {
        int c = nch[n0];
        if (c == 0) return;
        const int *cn = ch + (size_t)n0 * 8;
        for (int i = 0; i < c; i++) reduced_pairs(h, cn[i], cn[i], pl);
        for (int i = 0; i < c; i++)
            for (int j = i + 1; j < c; j++) reduced_pairs(h, cn[i], cn[j], pl);
        return;
    }
    int c0 = nch[n0], c1 = nch[n1];
    int l0 = h->node_level[n0], l1 = h->node_level[n1];
    if (box_admissible(h->enbox + 6 * (size_t)n0, h->enbox + 6 * (size_t)n1, H2E_ALPHA_H2)) {
        pl.adm.push_back(n0);
        pl.adm.push_back(n1);
        pl.min_adm_level = std::min(pl.min_adm_level, std::max(l0, l1));
        return;
    }
    if (c0 == 0 && c1 == 0) {
        pl.inadm.push_back(n0);
        pl.inadm.push_back(n1);
        return;
    }
    if (c0 == 0) {
        const int *cn = ch + (size_t)n1 * 8;
        for (int j = 0; j < c1; j++) reduced_pairs(h, n0, cn[j], pl);
        return;
    }
    if (c1 == 0) {
        const int *cn = ch + (size_t)n0 * 8;
        for (int i = 0; i < c0; i++) reduced_pairs(h, cn[i], n1, pl);
        return;
    }
    const int *cn0 = ch + (size_t)n0 * 8;
    const int *cn1 = ch + (size_t)n1 * 8;
    for (int i = 0; i < c0; i++)
        for (int j = 0; j < c1; j++) reduced_pairs(h, cn0[i], cn1[j], pl);
}

template <class T>
T *dup_vec(const std::vector<T> &v) {
    T *p = (T *)malloc(sizeof(T) * (v.size() > 0 ? v.size() : 1));
    if (!v.empty()) memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

}  // namespace

// Builds the tree over npts points (coord 3 x npts, column-major).  On return
// h->coord holds the permuted coordinates and h->coord_idx the permutation.
void partition_tree(H2ERI_p h, int npts, double *coord_in, int max_leaf_points, double max_leaf_size) {
    if (max_leaf_points <= 0) max_leaf_points = 300;
    if (max_leaf_size <= 0.0) max_leaf_size = 1.0;
    h->n_point = npts;
    h->max_leaf_points = max_leaf_points;
    h->max_leaf_size = max_leaf_size;
    free(h->coord);
    free(h->coord0);
    free(h->coord_idx);
    h->coord = (double *)malloc(sizeof(double) * 3 * (size_t)npts);
    h->coord0 = (double *)malloc(sizeof(double) * 3 * (size_t)npts);
    h->coord_idx = (int *)malloc(sizeof(int) * (size_t)npts);
    memcpy(h->coord, coord_in, sizeof(double) * 3 * (size_t)npts);
    memcpy(h->coord0, coord_in, sizeof(double) * 3 * (size_t)npts);
    for (int i = 0; i < npts; i++) h->coord_idx[i] = i;
    std::vector<double> ctmp(3 * (size_t)npts);
    std::vector<int> itmp(npts);
    PartCtx c;
    c.n_point = npts;
    c.coord = h->coord;
    c.coord_tmp = ctmp.data();
    c.coord_idx = h->coord_idx;
    c.coord_idx_tmp = itmp.data();
    c.max_leaf_points = max_leaf_points;
    c.max_leaf_size = max_leaf_size;
    TNode *root = bisect(c, 0, 0, npts - 1, nullptr);

    int n_node = root->n_node;
    h->n_node = n_node;
    h->root_idx = n_node - 1;
    h->n_leaf_node = c.n_leaf;
    h->max_child = 8;
    h->max_level = c.max_level;
    int nl = c.max_level + 1;
    h->parent = (int *)malloc(sizeof(int) * n_node);
    h->children = (int *)malloc(sizeof(int) * n_node * 8);
    h->pt_cluster = (int *)malloc(sizeof(int) * n_node * 2);
    h->mat_cluster = (int *)malloc(sizeof(int) * n_node * 2);
    h->n_child = (int *)malloc(sizeof(int) * n_node);
    h->node_level = (int *)malloc(sizeof(int) * n_node);
    h->node_height = (int *)malloc(sizeof(int) * n_node);
    h->level_n_node = (int *)calloc(nl, sizeof(int));
    h->level_nodes = (int *)malloc(sizeof(int) * nl * h->n_leaf_node);
    h->height_n_node = (int *)calloc(nl, sizeof(int));
    h->height_nodes = (int *)malloc(sizeof(int) * nl * h->n_leaf_node);
    h->enbox = (double *)malloc(sizeof(double) * 6 * n_node);
    tree_to_arrays(root, h);
    h->parent[h->root_idx] = -1;
    free_tree(root);

    PairLists pl;
    pl.min_adm_level = h->max_level;
    h->min_adm_level = 0;
    reduced_pairs(h, h->root_idx, h->root_idx, pl);
    h->min_adm_level = pl.min_adm_level;
    h->n_r_adm_pair = (int)(pl.adm.size() / 2);
    h->n_r_inadm_pair = (int)(pl.inadm.size() / 2);
    h->r_adm_pairs = dup_vec(pl.adm);
    h->r_inadm_pairs = dup_vec(pl.inadm);
}

// Box extents (src/H2ERI_partition.c:152-226), mat_cluster (:233-265),
// per-node pair lists (:273-360) and plist (:371-427).
void finish_partition_metadata(H2ERI_p h) {
    const int n_node = h->n_node, nleaf = h->n_leaf_node;
    // mat_cluster: leaves in post order own consecutive x ranges
    int off = 0;
    for (int i = 0; i < n_node; i++) {
        if (h->n_child[i] == 0) {
            int ps = h->pt_cluster[2 * i], pe = h->pt_cluster[2 * i + 1];
            int nb = h->sp_bfp_sidx[pe + 1] - h->sp_bfp_sidx[ps];
            h->mat_cluster[2 * i] = off;
            h->mat_cluster[2 * i + 1] = off + nb - 1;
            off += nb;
        } else {
            const int *ch = h->children + (size_t)i * 8;
            h->mat_cluster[2 * i] = h->mat_cluster[2 * ch[0]];
            h->mat_cluster[2 * i + 1] = h->mat_cluster[2 * ch[h->n_child[i] - 1] + 1];
        }
    }
    // box extent, if shell-pair extents are known
    if (h->sp_extent != nullptr) {
        free(h->box_extent);
        h->box_extent = (double *)calloc(n_node, sizeof(double));
        const int np = h->num_sp;
        for (int lv = h->max_level; lv >= 1; lv--) {
            for (int j = 0; j < h->level_n_node[lv]; j++) {
                int node = h->level_nodes[lv * nleaf + j];
                const double *eb = h->enbox + 6 * (size_t)node;
                double ctr[3] = {eb[0] + 0.5 * eb[3], eb[1] + 0.5 * eb[4], eb[2] + 0.5 * eb[5]};
                if (h->n_child[node] == 0) {
                    double be = 0.0;
                    for (int d = 0; d < 3; d++) {
                        const double *cd = h->sp_center + (size_t)d * np;
                        double w = eb[3 + d];
                        for (int k = h->pt_cluster[2 * node]; k <= h->pt_cluster[2 * node + 1]; k++) {
                            double t = 0.5 * w - std::fabs(cd[k] - ctr[d]);
                            t = (h->sp_extent[k] - t) / w;
                            be = std::max(be, std::max(t, 0.1));
                        }
                    }
                    h->box_extent[node] = std::ceil(be);
                } else {
                    double be = 0.0;
                    const int *ch = h->children + (size_t)node * 8;
                    for (int k = 0; k < h->n_child[node]; k++) be = std::max(be, h->box_extent[ch[k]]);
                    h->box_extent[node] = std::ceil(0.5 * be);
                }
            }
        }
    }
    // node_adm_pairs / node_inadm_pairs (leaf self pair first for inadm)
    auto build_lists = [&](const int *pairs, int npair, bool with_self, int *&sidx, int *&lst) {
        free(sidx);
        free(lst);
        sidx = (int *)calloc(n_node + 1, sizeof(int));
        int total = 2 * npair + (with_self ? nleaf : 0);
        lst = (int *)malloc(sizeof(int) * (total > 0 ? total : 1));
        std::vector<int> cnt(n_node, 0);
        if (with_self)
            for (int i = 0; i < nleaf; i++) cnt[h->height_nodes[i]]++;
        for (int i = 0; i < npair; i++) { cnt[pairs[2 * i]]++; cnt[pairs[2 * i + 1]]++; }
        for (int i = 0; i < n_node; i++) sidx[i + 1] = sidx[i] + cnt[i];
        std::vector<int> pos(sidx, sidx + n_node);
        if (with_self)
            for (int i = 0; i < nleaf; i++) { int n = h->height_nodes[i]; lst[pos[n]++] = n; }
        for (int i = 0; i < npair; i++) {
            int a = pairs[2 * i], b = pairs[2 * i + 1];
            lst[pos[a]++] = b;
            lst[pos[b]++] = a;
        }
    };
    build_lists(h->r_adm_pairs, h->n_r_adm_pair, false, h->node_adm_pairs_sidx, h->node_adm_pairs);
    build_lists(h->r_inadm_pairs, h->n_r_inadm_pair, true, h->node_inadm_pairs_sidx, h->node_inadm_pairs);

    // plist: for each shell, the shells it is paired with (sorted), and the pair index
    if (h->sp_shell_idx != nullptr && h->nshell > 0) {
        const int ns = h->nshell, np = h->num_sp;
        free(h->plist); free(h->plist_idx); free(h->plist_sidx);
        h->plist = (int *)malloc(sizeof(int) * 2 * (size_t)(np > 0 ? np : 1));
        h->plist_idx = (int *)malloc(sizeof(int) * 2 * (size_t)(np > 0 ? np : 1));
        h->plist_sidx = (int *)calloc(ns + 1, sizeof(int));
        std::vector<int> cnt(ns, 0);
        for (int i = 0; i < np; i++) {
            int s0 = h->sp_shell_idx[i], s1 = h->sp_shell_idx[i + np];
            cnt[s0]++;
            if (s0 != s1) cnt[s1]++;
        }
        for (int i = 0; i < ns; i++) h->plist_sidx[i + 1] = h->plist_sidx[i] + cnt[i];
        std::vector<std::vector<std::pair<int, int>>> tmp(ns);
        for (int i = 0; i < np; i++) {
            int s0 = h->sp_shell_idx[i], s1 = h->sp_shell_idx[i + np];
            tmp[s0].push_back({s1, i});
            if (s0 != s1) tmp[s1].push_back({s0, i});
        }
        for (int s = 0; s < ns; s++) {
            std::sort(tmp[s].begin(), tmp[s].end());
            int b = h->plist_sidx[s];
            for (size_t k = 0; k < tmp[s].size(); k++) {
                h->plist[b + k] = tmp[s][k].first;
                h->plist_idx[b + k] = tmp[s][k].second;
            }
        }
    }
}

}  // namespace h2e

using namespace h2e;

extern "C" {

// Permutes the screened shell pairs into tree order and builds all tree
// metadata (src/H2ERI_partition.c:28-143, 430-454).
void H2ERI_partition(H2ERI_p h) {
    if (!h || h->num_sp <= 0 || !h->sp_center) {
        set_error(ERR_ARG, "H2ERI_partition: call H2ERI_process_shells first");
        return;
    }
    double t0 = wtime();
    const int np = h->num_sp;
    int nbfp_total = 0;
    for (int i = 0; i < np; i++) nbfp_total += NCART(h->sp_shells[i].am) * NCART(h->sp_shells[i + np].am);
    h->krnl_mat_size = nbfp_total;
    partition_tree(h, np, h->sp_center, 0, 0.0);
    memcpy(h->sp_center, h->coord, sizeof(double) * 3 * (size_t)np);
    // permute shell pairs and extents into tree order
    std::vector<shell_t> sh_new(2 * (size_t)np);
    std::vector<int> idx_new(2 * (size_t)np);
    std::vector<double> ext_new(np);
    for (int i = 0; i < np; i++) {
        int c = h->coord_idx[i];
        ext_new[i] = h->sp_extent[c];
        sh_new[i] = h->sp_shells[c];
        sh_new[i + np] = h->sp_shells[c + np];
        idx_new[i] = h->sp_shell_idx[c];
        idx_new[i + np] = h->sp_shell_idx[c + np];
    }
    memcpy(h->sp_shells, sh_new.data(), sizeof(shell_t) * 2 * (size_t)np);
    memcpy(h->sp_shell_idx, idx_new.data(), sizeof(int) * 2 * (size_t)np);
    memcpy(h->sp_extent, ext_new.data(), sizeof(double) * (size_t)np);
    free(h->index_seq);
    h->index_seq = (int *)malloc(sizeof(int) * (size_t)np);
    for (int i = 0; i < np; i++) h->index_seq[i] = i;
    // basis-function bookkeeping (src/H2ERI_partition.c:108-143)
    free(h->shell_bf_sidx);
    free(h->sp_nbfp);
    free(h->sp_bfp_sidx);
    h->shell_bf_sidx = (int *)malloc(sizeof(int) * (h->nshell + 1));
    h->sp_nbfp = (int *)malloc(sizeof(int) * (size_t)np);
    h->sp_bfp_sidx = (int *)malloc(sizeof(int) * (size_t)(np + 1));
    h->shell_bf_sidx[0] = 0;
    h->max_am = 0;
    for (int i = 0; i < h->nshell; i++) {
        h->shell_bf_sidx[i + 1] = h->shell_bf_sidx[i] + NCART(h->shells[i].am);
        h->max_am = std::max(h->max_am, h->shells[i].am);
    }
    h->num_bf = h->shell_bf_sidx[h->nshell];
    h->max_shell_nbf = NCART(h->max_am);
    h->sp_bfp_sidx[0] = 0;
    for (int i = 0; i < np; i++) {
        h->sp_nbfp[i] = NCART(h->sp_shells[i].am) * NCART(h->sp_shells[i + np].am);
        h->sp_bfp_sidx[i + 1] = h->sp_bfp_sidx[i] + h->sp_nbfp[i];
    }
    h->num_sp_bfp = h->sp_bfp_sidx[np];
    finish_partition_metadata(h);
    // primitive pair data in tree order, (N, M) order as the reference
    HostState *hs = host_state(h);
    for (auto *p : hs->pairs) pair_data_free(p);
    hs->pairs.assign(np, nullptr);
    hs->max_am = h->max_am;
    #pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < np; i++) pair_data_build(h->sp_shells[i], h->sp_shells[i + np], hs->pairs[i]);
    h->timers[PT_TIMER_IDX] = wtime() - t0;
}

void H2ERI_partition_points_synthetic(H2ERI_p h, const int npts, const double *coord, const int *nbfp,
                                      const int max_leaf_points) {
    if (!h || npts <= 0 || !coord || !nbfp) { set_error(ERR_ARG, "partition_points_synthetic: bad args"); return; }
    double t0 = wtime();
    std::vector<double> c(coord, coord + 3 * (size_t)npts);
    partition_tree(h, npts, c.data(), max_leaf_points, 0.0);
    h->num_sp = npts;
    free(h->sp_nbfp);
    free(h->sp_bfp_sidx);
    h->sp_nbfp = (int *)malloc(sizeof(int) * npts);
    h->sp_bfp_sidx = (int *)malloc(sizeof(int) * (npts + 1));
    h->sp_bfp_sidx[0] = 0;
    for (int i = 0; i < npts; i++) {
        h->sp_nbfp[i] = nbfp[h->coord_idx[i]];
        h->sp_bfp_sidx[i + 1] = h->sp_bfp_sidx[i] + h->sp_nbfp[i];
    }
    h->num_sp_bfp = h->krnl_mat_size = h->sp_bfp_sidx[npts];
    free(h->index_seq);
    h->index_seq = (int *)malloc(sizeof(int) * (size_t)npts);
    for (int i = 0; i < npts; i++) h->index_seq[i] = i;
    finish_partition_metadata(h);
    host_state(h)->synthetic = true;
    h->timers[PT_TIMER_IDX] = wtime() - t0;
}

}  // extern "C"
