// H2 representation construction on the host: nested bases U (proxy-point
// interpolative decomposition), coupling blocks B and near-field blocks D,
// with low-rank re-compression of B and D.
//
// Algorithm of H2P-ERI's H2ERI_build_H2 (src/H2ERI_build_H2.c):
//   proxy-point layers on box surfaces                        :25-103
//   far-field pairs whose extents overlap the near field      :115-259
//   skeleton row/pair extraction                              :271-335
//   U per level, max_level -> min_adm_level: [NAI(proxy) S_pp | ERI(ff) S_ff]
//   with 16-nonzero +-1 sparse sketches, column normalisation, row ID
//                                                             :351-827
//   B/D compression into [U | V^T] when (nrow+ncol) k <= 0.8 nrow ncol
//                                                             :843-903
//   B blocks for the three level cases, CSR pair->block maps  :910-1152
//   D blocks (leaf diagonal, then inadmissible pairs)         :1159-1403
// The sketch uses a seeded per-node generator instead of rand(), so a build
// is reproducible.  Diagonal D blocks are evaluated on the upper shell-pair
// triangle and mirrored, so they are exactly symmetric.
#include "h2eri_internal.h"

#include <algorithm>
#include <cmath>
#include <omp.h>

namespace h2e {

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
    int below(int n) { return (int)(next() % (uint64_t)n); }
};

int leaf_nbfp(H2ERI_p h, int node) {
    int ps = h->pt_cluster[2 * node], pe = h->pt_cluster[2 * node + 1];
    return h->sp_bfp_sidx[pe + 1] - h->sp_bfp_sidx[ps];
}

int sum_nbfp(H2ERI_p h, const int *idx, int n) {
    int s = 0;
    for (int i = 0; i < n; i++) s += h->sp_nbfp[idx[i]];
    return s;
}

// Rows of the concatenated bfp of `pairs` -> pairs holding them (compacted)
// and the row indices within the compacted concatenation.
void extract_pair_idx(H2ERI_p h, const std::vector<int> &rows, const std::vector<int> &pairs,
                      std::vector<int> &pair_sub, std::vector<int> &rows_new) {
    const int np = (int)pairs.size();
    std::vector<int> off(np + 1, 0);
    for (int i = 0; i < np; i++) off[i + 1] = off[i] + h->sp_nbfp[pairs[i]];
    std::vector<int> owner(rows.size()), flag(np, -1);
    for (size_t t = 0; t < rows.size(); t++) {
        int x = rows[t];
        int j = (int)(std::upper_bound(off.begin(), off.end(), x) - off.begin()) - 1;
        owner[t] = j;
        flag[j] = 1;
    }
    pair_sub.clear();
    std::vector<int> newoff(np, 0);
    int acc = 0;
    for (int i = 0; i < np; i++) {
        if (flag[i] < 0) continue;
        flag[i] = (int)pair_sub.size();
        pair_sub.push_back(i);
        newoff[i] = acc;
        acc += h->sp_nbfp[pairs[i]];
    }
    rows_new.resize(rows.size());
    for (size_t t = 0; t < rows.size(); t++) {
        int j = owner[t];
        rows_new[t] = rows[t] - off[j] + newoff[j];
    }
}

// ERI block of bra pairs x ket pairs (global pair indices), row-major, ld ldo.
void eri_pairs_to_mat(HostState *hs, const int *bra, int nbra, const int *ket, int nket, EriWork *w,
                      double *out, int ldo) {
    int r = 0;
    for (int i = 0; i < nbra; i++) {
        const PairData *pb = hs->pairs[bra[i]];
        int c = 0;
        for (int j = 0; j < nket; j++) {
            const PairData *pk = hs->pairs[ket[j]];
            eri_pair_block(pb, pk, w, out + (size_t)r * ldo + c, ldo);
            c += pair_nbfp(pk);
        }
        r += pair_nbfp(pb);
    }
}

// Compress a block into [U | V^T] when it pays (src/H2ERI_build_H2.c:843-903).
H2E_dense_mat_p compress_blk(int nrow, int ncol, std::vector<double> &blk, double tol) {
    if (tol < 0.0) {   // stored dense, as the on-the-fly form evaluates it
        H2E_dense_mat_p res = dmat_new(nrow, ncol);
        if ((long long)nrow * ncol > 0) memcpy(res->data, blk.data(), sizeof(double) * (size_t)nrow * ncol);
        return res;
    }
    std::vector<double> A(blk);
    std::vector<double> U;
    std::vector<int> J;
    if (nrow > 0 && ncol > 0) id_compress_rows(nrow, ncol, A.data(), tol, U, J);
    int r = (int)J.size();
    if (nrow > 0 && ncol > 0 && r == 0) {
        // numerically zero block: store an exact rank-1 zero factorisation
        // (a rank-0 block would read as ld = -0 = 0, i.e. dense)
        H2E_dense_mat_p z = dmat_new(1, nrow + ncol);
        z->nrow = nrow; z->ncol = ncol; z->ld = -1;
        memset(z->data, 0, sizeof(double) * (size_t)(nrow + ncol));
        return z;
    }
    const long long old_size = (long long)nrow * ncol;
    const long long new_size = (long long)(nrow + ncol) * r;
    H2E_dense_mat_p res;
    if (nrow == 0 || ncol == 0 || new_size > (old_size * 4) / 5) {
        res = dmat_new(nrow, ncol);
        if (old_size > 0) memcpy(res->data, blk.data(), sizeof(double) * (size_t)old_size);
    } else {
        res = dmat_new(1, (int)new_size);
        res->nrow = nrow;
        res->ncol = ncol;
        res->ld = -r;
        memcpy(res->data, U.data(), sizeof(double) * (size_t)nrow * r);
        double *VT = res->data + (size_t)nrow * r;
        for (int l = 0; l < ncol; l++)
            for (int k = 0; k < r; k++) VT[(size_t)l * r + k] = blk[(size_t)J[k] * ncol + l];
    }
    return res;
}

// Proxy points on nlayer scaled copies of a [-1,1]^3 surface grid
// (src/H2ERI_build_H2.c:25-103).
void proxy_points(double r1, double r2, int nlayer, int npts_layer, std::vector<double> &pp) {
    int nf = npts_layer / 6;
    int na = (int)std::ceil(std::sqrt((double)nf));
    npts_layer = 6 * na * na;
    int nt = nlayer * npts_layer;
    pp.assign(3 * (size_t)nt, 0.0);
    double *x = pp.data(), *y = x + nt, *z = y + nt;
    double hh = 2.0 / (na + 1);
    int idx = 0;
    for (int i = 0; i < na; i++) {
        double hi = hh * (i + 1) - 1.0;
        for (int j = 0; j < na; j++) {
            double hj = hh * (j + 1) - 1.0;
            double pts[6][3] = {{hi, hj, -1.0}, {hi, hj, 1.0}, {hi, -1.0, hj}, {hi, 1.0, hj}, {-1.0, hi, hj}, {1.0, hi, hj}};
            for (int k = 0; k < 6; k++) { x[idx] = pts[k][0]; y[idx] = pts[k][1]; z[idx] = pts[k][2]; idx++; }
        }
    }
    for (int l = 1; l < nlayer; l++) {
        memcpy(x + (size_t)l * npts_layer, x, sizeof(double) * npts_layer);
        memcpy(y + (size_t)l * npts_layer, y, sizeof(double) * npts_layer);
        memcpy(z + (size_t)l * npts_layer, z, sizeof(double) * npts_layer);
    }
    int nl1 = std::max(nlayer - 1, 1);
    double dr = ((r2 - r1) / r1) / (double)nl1;
    for (int l = 0; l < nlayer; l++) {
        double r = r1 * (1.0 + l * dr);
        for (int j = 0; j < npts_layer; j++) {
            x[(size_t)l * npts_layer + j] *= r;
            y[(size_t)l * npts_layer + j] *= r;
            z[(size_t)l * npts_layer + j] *= r;
        }
    }
}

// Far-field shell pairs whose extents overlap each node's near field
// (src/H2ERI_build_H2.c:115-259).
void calc_ovlp_ff(H2ERI_p h, std::vector<std::vector<int>> &ff) {
    const int n_node = h->n_node, np = h->n_point, nleaf = h->n_leaf_node;
    const double *cx = h->sp_center, *cy = cx + np, *cz = cy + np, *ext = h->sp_extent;
    ff.assign(n_node, {});
    ff[h->root_idx].resize(np);
    for (int i = 0; i < np; i++) ff[h->root_idx][i] = i;
    for (int lv = 0; lv < h->max_level; lv++) {
        for (int j = 0; j < h->level_n_node[lv]; j++) {
            int node = h->level_nodes[lv * nleaf + j];
            int nc = h->n_child[node];
            if (nc == 0) continue;
            const std::vector<int> &src = ff[node];
            for (int k = 0; k < nc; k++) {
                int ch = h->children[(size_t)node * 8 + k];
                const double *eb = h->enbox + 6 * (size_t)ch;
                double ctr[3] = {eb[0] + 0.5 * eb[3], eb[1] + 0.5 * eb[4], eb[2] + 0.5 * eb[5]};
                const double sc = 0.5 + H2E_ALPHA_SUP;
                double semi[3] = {sc * eb[3], sc * eb[4], sc * eb[5]};
                std::vector<int> &dst = ff[ch];
                for (int l : src) {
                    double e = ext[l];
                    double rel[3] = {cx[l] - ctr[0] + semi[0] - e, cy[l] - ctr[1] + semi[1] - e, cz[l] - ctr[2] + semi[2] - e};
                    bool adm = false;
                    for (int d = 0; d < 3 && !adm; d++) {
                        bool left = rel[d] < 0;
                        bool adm_left = std::fabs(rel[d]) >= 2.0 * e - 1e-8;
                        bool adm_right = std::fabs(rel[d]) >= 2.0 * semi[d] - 1e-8;
                        adm = (left && adm_left) || (!left && adm_right);
                    }
                    if (!adm) dst.push_back(l);
                }
            }
        }
    }
    for (int lv = h->min_adm_level; lv <= h->max_level; lv++) {
        for (int j = 0; j < h->level_n_node[lv]; j++) {
            int node = h->level_nodes[lv * nleaf + j];
            const double *eb = h->enbox + 6 * (size_t)node;
            double ctr[3] = {eb[0] + 0.5 * eb[3], eb[1] + 0.5 * eb[4], eb[2] + 0.5 * eb[5]};
            const double ac = 0.5 + H2E_ALPHA_H2;
            double semi[3] = {ac * eb[3], ac * eb[4], ac * eb[5]};
            std::vector<int> kept;
            for (int l : ff[node]) {
                if (std::fabs(cx[l] - ctr[0]) > semi[0] || std::fabs(cy[l] - ctr[1]) > semi[1] ||
                    std::fabs(cz[l] - ctr[2]) > semi[2])
                    kept.push_back(l);
            }
            ff[node].swap(kept);
        }
    }
}

// Sparse +-1 sketch: for each of ncol output columns, nnz distinct rows in [0, k).
void rand_sketch(Rng &rng, int k, int ncol, int max_nnz, std::vector<int> &idx, std::vector<double> &val, int &nnz) {
    nnz = std::min(max_nnz, k);
    idx.resize((size_t)ncol * nnz);
    val.resize((size_t)ncol * nnz);
    std::vector<char> flag(k > 0 ? k : 1, 0);
    for (int c = 0; c < ncol; c++) {
        int *ci = idx.data() + (size_t)c * nnz;
        for (int t = 0; t < nnz; t++) val[(size_t)c * nnz + t] = (rng.next() & 1) ? 1.0 : -1.0;
        int cnt = 0;
        while (cnt < nnz) {
            int r = rng.below(k);
            if (!flag[r]) { flag[r] = 1; ci[cnt++] = r; }
        }
        for (int t = 0; t < nnz; t++) flag[ci[t]] = 0;
    }
}

void build_U(H2ERI_p h, uint64_t seed) {
    HostState *hs = host_state(h);
    const int n_node = h->n_node, nleaf = h->n_leaf_node, np = h->n_point;
    h->U = (H2E_dense_mat_p *)calloc(n_node, sizeof(H2E_dense_mat_p));
    h->J_pair = (H2E_int_vec_p *)calloc(n_node, sizeof(H2E_int_vec_p));
    h->J_row = (H2E_int_vec_p *)calloc(n_node, sizeof(H2E_int_vec_p));
    if (h->n_r_adm_pair == 0) {   // no far field (e.g. a single-leaf tree): no node needs a basis
        for (int i = 0; i < n_node; i++) {
            h->U[i] = dmat_new(0, 0);
            h->J_pair[i] = ivec_new(1);
            h->J_row[i] = ivec_new(1);
        }
        return;
    }
    std::vector<std::vector<int>> ff;
    calc_ovlp_ff(h, ff);
    std::vector<char> skel(np, 1);
    // leaves above each level keep all their pairs as skeleton pairs
    std::vector<std::vector<int>> lvl_leaf(h->max_level + 1);
    for (int i = 0; i < nleaf; i++) {
        int lf = h->height_nodes[i];
        for (int j = h->node_level[lf] + 1; j <= h->max_level; j++) lvl_leaf[j].push_back(lf);
    }
    const double tol = h->QR_stop_tol;
    for (int lv = h->max_level; lv >= h->min_adm_level; lv--) {
        const int nn = h->level_n_node[lv];
        const int *lnodes = h->level_nodes + (size_t)lv * nleaf;
        #pragma omp parallel
        {
            EriWork *w = eri_work_new(hs->max_am);
            std::vector<double> pp, Apb, Abk, strip, Ub;
            std::vector<int> sk_idx, J;
            std::vector<double> sk_val;
            #pragma omp for schedule(dynamic)
            for (int jn = 0; jn < nn; jn++) {
                const int node = lnodes[jn];
                Rng rng(seed ^ ((uint64_t)(node + 1) * 0xD1B54A32D192ED03ULL));
                // (1) candidate rows
                std::vector<int> pair_idx, row_idx;
                const int nc = h->n_child[node];
                if (nc == 0) {
                    for (int k = h->pt_cluster[2 * node]; k <= h->pt_cluster[2 * node + 1]; k++) pair_idx.push_back(k);
                    int nb = sum_nbfp(h, pair_idx.data(), (int)pair_idx.size());
                    row_idx.resize(nb);
                    for (int k = 0; k < nb; k++) row_idx[k] = k;
                } else {
                    int off = 0;
                    for (int k = 0; k < nc; k++) {
                        int ch = h->children[(size_t)node * 8 + k];
                        H2E_int_vec_p jp = h->J_pair[ch], jr = h->J_row[ch];
                        for (int t = 0; t < jr->length; t++) row_idx.push_back(jr->data[t] + off);
                        for (int t = 0; t < jp->length; t++) pair_idx.push_back(jp->data[t]);
                        off += sum_nbfp(h, jp->data, jp->length);
                    }
                }
                const int nrow_all = sum_nbfp(h, pair_idx.data(), (int)pair_idx.size());
                // (2) proxy points around the node box
                const double *eb = h->enbox + 6 * (size_t)node;
                const double width = eb[3], bext = h->box_extent[node];
                const double r1 = width * (0.5 + H2E_ALPHA_SUP), r2 = width * (0.5 + bext);
                const int nlayer = 1 + (int)std::ceil((bext - H2E_ALPHA_SUP) * (h->pp_nlayer_ext - 1));
                proxy_points(r1, r2, nlayer, h->pp_npts_layer, pp);
                const int npp = (int)(pp.size() / 3);
                for (int k = 0; k < npp; k++) {
                    pp[k] += eb[0] + 0.5 * eb[3];
                    pp[npp + k] += eb[1] + 0.5 * eb[4];
                    pp[2 * (size_t)npp + k] += eb[2] + 0.5 * eb[5];
                }
                // (3) far-field pairs restricted to current skeleton pairs
                std::vector<int> ffp;
                for (int l : ff[node]) if (skel[l]) ffp.push_back(l);
                const int nff = sum_nbfp(h, ffp.data(), (int)ffp.size());
                const int ncolA = 2 * nrow_all;
                Abk.assign((size_t)nrow_all * ncolA, 0.0);
                // (4) NAI part: proxy points used by the sketch only
                int nnz;
                rand_sketch(rng, npp, nrow_all, 16, sk_idx, sk_val, nnz);
                {
                    std::vector<int> map(npp, -1), used;
                    for (int v : sk_idx) if (map[v] < 0) { map[v] = 0; }
                    for (int k = 0; k < npp; k++) if (map[k] >= 0) { map[k] = (int)used.size(); used.push_back(k); }
                    const int nu = (int)used.size();
                    std::vector<double> ux(nu), uy(nu), uz(nu);
                    for (int k = 0; k < nu; k++) { ux[k] = pp[used[k]]; uy[k] = pp[npp + used[k]]; uz[k] = pp[2 * (size_t)npp + used[k]]; }
                    Apb.assign((size_t)nrow_all * (nu > 0 ? nu : 1), 0.0);
                    int r = 0;
                    for (int pi : pair_idx) {
                        nai_pair_block(hs->pairs[pi], nu, ux.data(), uy.data(), uz.data(), w, Apb.data() + (size_t)r * nu, nu);
                        r += h->sp_nbfp[pi];
                    }
                    for (int row = 0; row < nrow_all; row++) {
                        const double *src = Apb.data() + (size_t)row * nu;
                        double *dst = Abk.data() + (size_t)row * ncolA;
                        for (int c = 0; c < nrow_all; c++) {
                            double s = 0.0;
                            for (int t = 0; t < nnz; t++) s += sk_val[(size_t)c * nnz + t] * src[map[sk_idx[(size_t)c * nnz + t]]];
                            dst[c] = s;
                        }
                    }
                }
                // (5) ERI part against the far-field pairs touched by the sketch
                if (nff > 0) {
                    rand_sketch(rng, nff, nrow_all, 16, sk_idx, sk_val, nnz);
                    std::vector<int> usedcol(nff, 0);
                    for (int v : sk_idx) usedcol[v] = 1;
                    std::vector<int> ffoff(ffp.size() + 1, 0);
                    for (size_t k = 0; k < ffp.size(); k++) ffoff[k + 1] = ffoff[k] + h->sp_nbfp[ffp[k]];
                    std::vector<int> ff1, coloff_new(ffp.size(), -1);
                    int ncol1 = 0;
                    for (size_t k = 0; k < ffp.size(); k++) {
                        bool any = false;
                        for (int c = ffoff[k]; c < ffoff[k + 1] && !any; c++) any = usedcol[c] != 0;
                        if (!any) continue;
                        coloff_new[k] = ncol1;
                        ff1.push_back(ffp[k]);
                        ncol1 += h->sp_nbfp[ffp[k]];
                    }
                    std::vector<int> colmap(nff, -1);
                    for (size_t k = 0; k < ffp.size(); k++)
                        if (coloff_new[k] >= 0)
                            for (int c = ffoff[k]; c < ffoff[k + 1]; c++) colmap[c] = coloff_new[k] + (c - ffoff[k]);
                    strip.resize((size_t)36 * (ncol1 > 0 ? ncol1 : 1));
                    int r = 0;
                    for (int pi : pair_idx) {
                        const int nb = h->sp_nbfp[pi];
                        eri_pairs_to_mat(hs, &pi, 1, ff1.data(), (int)ff1.size(), w, strip.data(), ncol1);
                        for (int row = 0; row < nb; row++) {
                            const double *src = strip.data() + (size_t)row * ncol1;
                            double *dst = Abk.data() + (size_t)(r + row) * ncolA + nrow_all;
                            for (int c = 0; c < nrow_all; c++) {
                                double s = 0.0;
                                for (int t = 0; t < nnz; t++)
                                    s += sk_val[(size_t)c * nnz + t] * src[colmap[sk_idx[(size_t)c * nnz + t]]];
                                dst[c] = s;
                            }
                        }
                        r += nb;
                    }
                }
                // (6) normalise columns, keep candidate rows, row ID
                for (int c = 0; c < ncolA; c++) {
                    double s = 0.0;
                    for (int row = 0; row < nrow_all; row++) s += Abk[(size_t)row * ncolA + c] * Abk[(size_t)row * ncolA + c];
                    if (s <= 0.0) continue;
                    double inv = 1.0 / std::sqrt(s);
                    for (int row = 0; row < nrow_all; row++) Abk[(size_t)row * ncolA + c] *= inv;
                }
                const int nsel = (int)row_idx.size();
                for (int t = 0; t < nsel; t++)
                    if (row_idx[t] != t)
                        memcpy(Abk.data() + (size_t)t * ncolA, Abk.data() + (size_t)row_idx[t] * ncolA, sizeof(double) * ncolA);
                id_compress_rows(nsel, ncolA, Abk.data(), tol, Ub, J);
                const int rank = (int)J.size();
                H2E_dense_mat_p U = dmat_new(nsel, rank);
                if (rank > 0) memcpy(U->data, Ub.data(), sizeof(double) * (size_t)nsel * rank);
                h->U[node] = U;
                std::vector<int> sub_rows(rank), pair_sub, rows_new;
                for (int k = 0; k < rank; k++) sub_rows[k] = row_idx[J[k]];
                extract_pair_idx(h, sub_rows, pair_idx, pair_sub, rows_new);
                H2E_int_vec_p jp = ivec_new((int)pair_sub.size()), jr = ivec_new(rank);
                for (size_t k = 0; k < pair_sub.size(); k++) jp->data[k] = pair_idx[pair_sub[k]];
                jp->length = (int)pair_sub.size();
                for (int k = 0; k < rank; k++) jr->data[k] = rows_new[k];
                jr->length = rank;
                h->J_pair[node] = jp;
                h->J_row[node] = jr;
            }
            eri_work_free(w);
        }
        // skeleton pairs for the next (coarser) level
        std::fill(skel.begin(), skel.end(), 0);
        for (int jn = 0; jn < nn; jn++) {
            H2E_int_vec_p jp = h->J_pair[lnodes[jn]];
            for (int k = 0; k < jp->length; k++) skel[jp->data[k]] = 1;
        }
        for (int lf : lvl_leaf[lv])
            for (int k = h->pt_cluster[2 * lf]; k <= h->pt_cluster[2 * lf + 1]; k++) skel[k] = 1;
        if (h->print_dbginfo) {
            fprintf(stderr, "[H2ERI_build_H2] U level %d: %d nodes done\n", lv, nn);
            fflush(stderr);
        }
    }
    for (int i = 0; i < n_node; i++) {
        if (!h->U[i]) { h->U[i] = dmat_new(0, 0); }
        if (!h->J_row[i]) h->J_row[i] = ivec_new(1);
        if (!h->J_pair[i]) h->J_pair[i] = ivec_new(1);
    }
}

}  // namespace

// Shapes of the B and D blocks, pair->block CSR maps (src/H2ERI_build_H2.c:932-1011,
// 1180-1260) and node_n_r_adm.
void build_B_D_index(H2ERI_p h) {
    const int nB = h->n_r_adm_pair, nleaf = h->n_leaf_node, nI = h->n_r_inadm_pair, n_node = h->n_node;
    h->n_B = nB;
    h->n_D = nleaf + nI;
    free(h->B_nrow); free(h->B_ncol); free(h->D_nrow); free(h->D_ncol); free(h->node_n_r_adm);
    h->B_nrow = (int *)malloc(sizeof(int) * (nB > 0 ? nB : 1));
    h->B_ncol = (int *)malloc(sizeof(int) * (nB > 0 ? nB : 1));
    h->D_nrow = (int *)malloc(sizeof(int) * h->n_D);
    h->D_ncol = (int *)malloc(sizeof(int) * h->n_D);
    h->node_n_r_adm = (int *)calloc(n_node, sizeof(int));
    auto rank = [&](int n) { return h->U[n] ? h->U[n]->ncol : 0; };
    std::vector<int> bi, bj, bv;
    for (int i = 0; i < nB; i++) {
        int n0 = h->r_adm_pairs[2 * i], n1 = h->r_adm_pairs[2 * i + 1];
        int l0 = h->node_level[n0], l1 = h->node_level[n1];
        h->node_n_r_adm[n0]++;
        h->node_n_r_adm[n1]++;
        if (l0 == l1) { h->B_nrow[i] = rank(n0); h->B_ncol[i] = rank(n1); }
        else if (l0 > l1) { h->B_nrow[i] = rank(n0); h->B_ncol[i] = leaf_nbfp(h, n1); }
        else { h->B_nrow[i] = leaf_nbfp(h, n0); h->B_ncol[i] = rank(n1); }
        bi.push_back(n0); bj.push_back(n1); bv.push_back(i + 1);
        bi.push_back(n1); bj.push_back(n0); bv.push_back(-(i + 1));
    }
    std::vector<int> di, dj, dv;
    for (int i = 0; i < nleaf; i++) {
        int n = h->height_nodes[i];
        h->D_nrow[i] = h->D_ncol[i] = leaf_nbfp(h, n);
        di.push_back(n); dj.push_back(n); dv.push_back(i + 1);
    }
    for (int i = 0; i < nI; i++) {
        int n0 = h->r_inadm_pairs[2 * i], n1 = h->r_inadm_pairs[2 * i + 1];
        h->D_nrow[nleaf + i] = leaf_nbfp(h, n0);
        h->D_ncol[nleaf + i] = leaf_nbfp(h, n1);
        di.push_back(n0); dj.push_back(n1); dv.push_back(nleaf + i + 1);
        di.push_back(n1); dj.push_back(n0); dv.push_back(-(nleaf + i + 1));
    }
    auto coo_to_csr = [&](const std::vector<int> &r, const std::vector<int> &c, const std::vector<int> &v,
                          int *&rowptr, int *&colidx, int *&val) {
        free(rowptr); free(colidx); free(val);
        size_t nnz = r.size();
        rowptr = (int *)calloc(n_node + 1, sizeof(int));
        colidx = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
        val = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
        std::vector<std::vector<std::pair<int, int>>> rows(n_node);
        for (size_t k = 0; k < nnz; k++) rows[r[k]].push_back({c[k], v[k]});
        size_t pos = 0;
        for (int i = 0; i < n_node; i++) {
            std::sort(rows[i].begin(), rows[i].end());
            for (auto &e : rows[i]) { colidx[pos] = e.first; val[pos] = e.second; pos++; }
            rowptr[i + 1] = (int)pos;
        }
    };
    coo_to_csr(bi, bj, bv, h->B_p2i_rowptr, h->B_p2i_colidx, h->B_p2i_val);
    coo_to_csr(di, dj, dv, h->D_p2i_rowptr, h->D_p2i_colidx, h->D_p2i_val);
}

// Storage statistics in the reference's units (src/H2ERI_build_H2.c:801-805,
// 1123-1136, 1363-1387): element counts.
void update_mat_sizes(H2ERI_p h) {
    size_t *ms = h->mat_size;
    for (int i = 0; i < 7; i++) ms[i] = 0;
    for (int i = 0; i < h->n_node; i++) {
        H2E_dense_mat_p U = h->U[i];
        if (!U || (U->nrow == 0 && U->ncol == 0)) continue;
        size_t e = (size_t)U->nrow * U->ncol;
        ms[U_SIZE_IDX] += e;
        ms[MV_FWD_SIZE_IDX] += e + U->nrow + U->ncol;
        ms[MV_BWD_SIZE_IDX] += e + U->nrow + U->ncol;
    }
    for (int i = 0; i < h->n_B; i++) {
        size_t e = blk_elems(h->c_B_blks[i]);
        ms[B_SIZE_IDX] += e;
        ms[MV_MID_SIZE_IDX] += e + 2 * (size_t)(h->B_nrow[i] + h->B_ncol[i]);
    }
    for (int i = 0; i < h->n_D; i++) {
        H2E_dense_mat_p D = h->c_D_blks[i];
        if (!D) continue;   // BD_JIT: evaluated during the matvec, not stored
        size_t e = blk_elems(D);
        ms[D_SIZE_IDX] += e;
        ms[MV_DEN_SIZE_IDX] += e + (size_t)(i < h->n_leaf_node ? 1 : 2) * (D->nrow + D->ncol);
    }
}

}  // namespace h2e

using namespace h2e;

extern "C" {

// BD_JIT = 1 (src/H2ERI_build_H2.c:1406-1411, :1016, :1265): U, the skeleton
// lists and the B/D shapes only; the matvec evaluates every B and D block from
// the shell pairs when it needs it, and c_B_blks / c_D_blks hold no blocks.
void H2ERI_build_H2(H2ERI_p h, const int BD_JIT) {
    if (!h || h->n_node <= 0) { set_error(ERR_ARG, "H2ERI_build_H2: call H2ERI_partition first"); return; }
    HostState *hs = host_state(h);
    if (hs->synthetic || hs->pairs.empty()) { set_error(ERR_ARG, "H2ERI_build_H2: no shell-pair data"); return; }
    if (BD_JIT == 1 && hs->max_am > 2) {
        set_error(ERR_UNSUPPORTED, "H2ERI_build_H2: BD_JIT=1 evaluates shells up to d (max am %d)", hs->max_am);
        return;
    }
    device_invalidate(h);
    h->BD_JIT = BD_JIT == 1 ? 1 : 0;
    double t0 = wtime();
    build_U(h, 0x5EEDULL);
    double t1 = wtime();
    build_B_D_index(h);
    if (h->BD_JIT) {
        h->c_B_blks = (H2E_dense_mat_p *)calloc(h->n_B > 0 ? h->n_B : 1, sizeof(H2E_dense_mat_p));
        h->c_D_blks = (H2E_dense_mat_p *)calloc(h->n_D > 0 ? h->n_D : 1, sizeof(H2E_dense_mat_p));
        update_mat_sizes(h);
        h->timers[U_BUILD_TIMER_IDX] = t1 - t0;
        h->timers[B_BUILD_TIMER_IDX] = 0.0;
        h->timers[D_BUILD_TIMER_IDX] = 0.0;
        return;
    }
    // a negative tolerance keeps every block dense (H2ERI_set_block_compression)
    const double tol = hs->bd_dense ? -1.0 : h->QR_stop_tol;
    const int nB = h->n_B;
    h->c_B_blks = (H2E_dense_mat_p *)calloc(nB > 0 ? nB : 1, sizeof(H2E_dense_mat_p));
    #pragma omp parallel
    {
        EriWork *w = eri_work_new(hs->max_am);
        std::vector<double> tmp, blk;
        #pragma omp for schedule(dynamic)
        for (int i = 0; i < nB; i++) {
            int n0 = h->r_adm_pairs[2 * i], n1 = h->r_adm_pairs[2 * i + 1];
            int l0 = h->node_level[n0], l1 = h->node_level[n1];
            std::vector<int> bra, ket;
            const int *rsel = nullptr, *csel = nullptr;
            int nrs = 0, ncs = 0;
            if (l0 >= l1) {
                bra.assign(h->J_pair[n0]->data, h->J_pair[n0]->data + h->J_pair[n0]->length);
                rsel = h->J_row[n0]->data; nrs = h->J_row[n0]->length;
            } else {
                for (int k = h->pt_cluster[2 * n0]; k <= h->pt_cluster[2 * n0 + 1]; k++) bra.push_back(k);
            }
            if (l1 >= l0) {
                ket.assign(h->J_pair[n1]->data, h->J_pair[n1]->data + h->J_pair[n1]->length);
                csel = h->J_row[n1]->data; ncs = h->J_row[n1]->length;
            } else {
                for (int k = h->pt_cluster[2 * n1]; k <= h->pt_cluster[2 * n1 + 1]; k++) ket.push_back(k);
            }
            int tr = sum_nbfp(h, bra.data(), (int)bra.size()), tc = sum_nbfp(h, ket.data(), (int)ket.size());
            tmp.assign((size_t)tr * tc, 0.0);
            eri_pairs_to_mat(hs, bra.data(), (int)bra.size(), ket.data(), (int)ket.size(), w, tmp.data(), tc);
            const int nr = rsel ? nrs : tr, nc = csel ? ncs : tc;
            blk.assign((size_t)nr * nc, 0.0);
            for (int r = 0; r < nr; r++) {
                int sr = rsel ? rsel[r] : r;
                for (int c = 0; c < nc; c++) blk[(size_t)r * nc + c] = tmp[(size_t)sr * tc + (csel ? csel[c] : c)];
            }
            h->c_B_blks[i] = compress_blk(nr, nc, blk, tol);
        }
        eri_work_free(w);
    }
    double t2 = wtime();
    if (h->print_dbginfo) {
        fprintf(stderr, "[H2ERI_build_H2] %d B blocks done (%.1f s)\n", nB, t2 - t0);
        fflush(stderr);
    }
    const int nleaf = h->n_leaf_node, nD = h->n_D;
    h->c_D_blks = (H2E_dense_mat_p *)calloc(nD, sizeof(H2E_dense_mat_p));
    #pragma omp parallel
    {
        EriWork *w = eri_work_new(hs->max_am);
        std::vector<double> blk;
        #pragma omp for schedule(dynamic)
        for (int i = 0; i < nD; i++) {
            int n0, n1;
            if (i < nleaf) { n0 = n1 = h->height_nodes[i]; }
            else { n0 = h->r_inadm_pairs[2 * (i - nleaf)]; n1 = h->r_inadm_pairs[2 * (i - nleaf) + 1]; }
            const int ps0 = h->pt_cluster[2 * n0], pe0 = h->pt_cluster[2 * n0 + 1];
            const int ps1 = h->pt_cluster[2 * n1], pe1 = h->pt_cluster[2 * n1 + 1];
            const int nr = h->D_nrow[i], nc = h->D_ncol[i];
            blk.assign((size_t)nr * nc, 0.0);
            if (n0 == n1) {
                // upper shell-pair triangle, mirrored
                for (int a = ps0; a <= pe0; a++) {
                    int ra = h->sp_bfp_sidx[a] - h->sp_bfp_sidx[ps0];
                    for (int b = a; b <= pe0; b++) {
                        int cb = h->sp_bfp_sidx[b] - h->sp_bfp_sidx[ps0];
                        eri_pair_block(hs->pairs[a], hs->pairs[b], w, blk.data() + (size_t)ra * nc + cb, nc);
                    }
                }
                for (int a = ps0; a <= pe0; a++) {
                    int ra = h->sp_bfp_sidx[a] - h->sp_bfp_sidx[ps0], na = h->sp_nbfp[a];
                    for (int b = a + 1; b <= pe0; b++) {
                        int cb = h->sp_bfp_sidx[b] - h->sp_bfp_sidx[ps0], nb = h->sp_nbfp[b];
                        for (int r = 0; r < na; r++)
                            for (int c = 0; c < nb; c++)
                                blk[(size_t)(cb + c) * nc + ra + r] = blk[(size_t)(ra + r) * nc + cb + c];
                    }
                    // diagonal shell-pair block: symmetrise
                    for (int r = 0; r < na; r++)
                        for (int c = r + 1; c < na; c++) {
                            double v = 0.5 * (blk[(size_t)(ra + r) * nc + ra + c] + blk[(size_t)(ra + c) * nc + ra + r]);
                            blk[(size_t)(ra + r) * nc + ra + c] = blk[(size_t)(ra + c) * nc + ra + r] = v;
                        }
                }
            } else {
                std::vector<int> bra, ket;
                for (int k = ps0; k <= pe0; k++) bra.push_back(k);
                for (int k = ps1; k <= pe1; k++) ket.push_back(k);
                eri_pairs_to_mat(hs, bra.data(), (int)bra.size(), ket.data(), (int)ket.size(), w, blk.data(), nc);
            }
            h->c_D_blks[i] = compress_blk(nr, nc, blk, tol);
        }
        eri_work_free(w);
    }
    double t3 = wtime();
    if (h->print_dbginfo) {
        fprintf(stderr, "[H2ERI_build_H2] %d D blocks done (%.1f s)\n", nD, t3 - t0);
        fflush(stderr);
    }
    update_mat_sizes(h);
    h->timers[U_BUILD_TIMER_IDX] = t1 - t0;
    h->timers[B_BUILD_TIMER_IDX] = t2 - t1;
    h->timers[D_BUILD_TIMER_IDX] = t3 - t2;
}

// Seeded chemistry-free U/B/D on an existing partition.  Ranks are random in
// [1, max_rank] (capped by the row count); entries are N(0,1)/sqrt(n); a
// fraction lr_frac of the B and off-diagonal D blocks and a fraction
// diag_lr_frac of the diagonal D blocks are stored low-rank; dense diagonal
// D blocks are symmetric.
void H2ERI_build_H2_synthetic(H2ERI_p h, const unsigned long long seed, const int max_rank, const double lr_frac,
                              const double diag_lr_frac) {
    if (!h || h->n_node <= 0 || max_rank <= 0) { set_error(ERR_ARG, "build_H2_synthetic: partition first"); return; }
    device_invalidate(h);
    h->BD_JIT = 0;
    double t0 = wtime();
    const int n_node = h->n_node, nleaf = h->n_leaf_node;
    h->U = (H2E_dense_mat_p *)calloc(n_node, sizeof(H2E_dense_mat_p));
    h->J_pair = (H2E_int_vec_p *)calloc(n_node, sizeof(H2E_int_vec_p));
    h->J_row = (H2E_int_vec_p *)calloc(n_node, sizeof(H2E_int_vec_p));
    for (int lv = h->max_level; lv >= 0; lv--) {
        for (int j = 0; j < h->level_n_node[lv]; j++) {
            int node = h->level_nodes[(size_t)lv * nleaf + j];
            if (lv < h->min_adm_level) { h->U[node] = dmat_new(0, 0); continue; }
            Rng rng(seed ^ ((uint64_t)(node + 7) * 0xA24BAED4963EE407ULL));
            int nrow;
            if (h->n_child[node] == 0) nrow = leaf_nbfp(h, node);
            else {
                nrow = 0;
                for (int k = 0; k < h->n_child[node]; k++) nrow += h->U[h->children[(size_t)node * 8 + k]]->ncol;
            }
            int r = 1 + rng.below(max_rank);
            if (r > nrow) r = nrow;
            H2E_dense_mat_p U = dmat_new(nrow, r);
            double sc = 1.0 / std::sqrt((double)(nrow > 0 ? nrow : 1));
            for (size_t k = 0; k < (size_t)nrow * r; k++) U->data[k] = rng.normal() * sc;
            h->U[node] = U;
        }
    }
    for (int i = 0; i < n_node; i++) {
        h->J_pair[i] = ivec_new(1);
        h->J_row[i] = ivec_new(1);
    }
    build_B_D_index(h);
    auto make_blk = [&](Rng &rng, int nr, int nc, bool sym) -> H2E_dense_mat_p {
        bool lr = nr > 1 && nc > 1 && rng.uniform() < (sym ? diag_lr_frac : lr_frac);
        if (lr) sym = false;
        int kmax = (int)((0.8 * nr * nc) / (double)(nr + nc));
        if (lr && kmax >= 1) {
            int k = 1 + rng.below(std::min(kmax, max_rank));
            H2E_dense_mat_p b = dmat_new(1, (nr + nc) * k);
            b->nrow = nr; b->ncol = nc; b->ld = -k;
            double sc = 1.0 / std::sqrt((double)k * std::sqrt((double)nr * nc));
            for (size_t t = 0; t < (size_t)(nr + nc) * k; t++) b->data[t] = rng.normal() * std::sqrt(sc);
            return b;
        }
        H2E_dense_mat_p b = dmat_new(nr, nc);
        double sc = 1.0 / std::sqrt((double)(nc > 0 ? nc : 1));
        for (int r = 0; r < nr; r++)
            for (int c = (sym ? r : 0); c < nc; c++) {
                double v = rng.normal() * sc;
                b->data[(size_t)r * nc + c] = v;
                if (sym) b->data[(size_t)c * nc + r] = v;
            }
        return b;
    };
    h->c_B_blks = (H2E_dense_mat_p *)calloc(h->n_B > 0 ? h->n_B : 1, sizeof(H2E_dense_mat_p));
    for (int i = 0; i < h->n_B; i++) {
        Rng rng(seed ^ ((uint64_t)(i + 11) * 0x9FB21C651E98DF25ULL));
        h->c_B_blks[i] = make_blk(rng, h->B_nrow[i], h->B_ncol[i], false);
    }
    h->c_D_blks = (H2E_dense_mat_p *)calloc(h->n_D, sizeof(H2E_dense_mat_p));
    for (int i = 0; i < h->n_D; i++) {
        Rng rng(seed ^ ((uint64_t)(i + 13) * 0xC2B2AE3D27D4EB4FULL));
        h->c_D_blks[i] = make_blk(rng, h->D_nrow[i], h->D_ncol[i], i < nleaf);
    }
    update_mat_sizes(h);
    h->timers[U_BUILD_TIMER_IDX] = wtime() - t0;
}

// Exact rows of the ERI matrix times x for the listed shell pairs (tree
// order): for pair i, y[rows of i] = sum over every kept shell pair j of
// (i|j) x_j, every integral evaluated directly, nothing compressed.  The
// rows of the listed pairs are written one pair after another into y.  A
// measurement aid for the compression error of the H2 representation
// (src/H2ERI_build_H2.c builds the same integrals block by block).
void H2ERI_direct_matvec_rows(H2ERI_p h, const int npair, const int *pair_idx, const double *x, double *y) {
    if (!h || npair < 0 || (npair > 0 && (!pair_idx || !x || !y))) { set_error(ERR_ARG, "H2ERI_direct_matvec_rows: bad argument"); return; }
    HostState *hs = host_state(h);
    const int np = h->num_sp;
    if ((int)hs->pairs.size() != np || np == 0 || hs->synthetic) {
        set_error(ERR_ARG, "H2ERI_direct_matvec_rows: needs a molecule processed in this process (shell-pair data)");
        return;
    }
    std::vector<size_t> yoff(npair + 1, 0);
    for (int k = 0; k < npair; k++) {
        if (pair_idx[k] < 0 || pair_idx[k] >= np) { set_error(ERR_ARG, "H2ERI_direct_matvec_rows: pair %d out of range", pair_idx[k]); return; }
        yoff[k + 1] = yoff[k] + h->sp_nbfp[pair_idx[k]];
    }
    const int mx = h->max_shell_nbf * h->max_shell_nbf;
    // one (bra pair, ket range) item per task; partial row sums per task
    const int KCH = 512;
    const int nch = (np + KCH - 1) / KCH;
    std::vector<double> part((size_t)npair * nch * mx, 0.0);
    #pragma omp parallel
    {
        EriWork *w = eri_work_new(hs->max_am);
        std::vector<double> blk((size_t)mx * mx);
        #pragma omp for schedule(dynamic, 1)
        for (long long t = 0; t < (long long)npair * nch; t++) {
            const int k = (int)(t / nch), c = (int)(t % nch);
            const int i = pair_idx[k], ni = h->sp_nbfp[i];
            double *acc = part.data() + (size_t)t * mx;
            for (int j = c * KCH; j < std::min(np, (c + 1) * KCH); j++) {
                const int nj = h->sp_nbfp[j];
                eri_pair_block(hs->pairs[i], hs->pairs[j], w, blk.data(), nj);
                const double *xj = x + h->sp_bfp_sidx[j];
                for (int r = 0; r < ni; r++) {
                    double s = 0.0;
                    for (int q = 0; q < nj; q++) s += blk[(size_t)r * nj + q] * xj[q];
                    acc[r] += s;
                }
            }
        }
        eri_work_free(w);
    }
    for (int k = 0; k < npair; k++) {
        const int ni = h->sp_nbfp[pair_idx[k]];
        for (int r = 0; r < ni; r++) {
            double s = 0.0;
            for (int c = 0; c < nch; c++) s += part[((size_t)k * nch + c) * mx + r];
            y[yoff[k] + r] = s;
        }
    }
}

}  // extern "C"
