// Molecule input, Schwarz screening and shell-pair extents.
//
//   .mol reader and shell normalisation        src/CMS.c:14-123
//   Schwarz screening, pairs (M,N) with M <= N src/H2ERI_shell_operations.c:22-89
//   Gaussian extent by bisection               src/H2ERI_shell_operations.c:98-128
//   shell-pair extent box and centre           src/H2ERI_shell_operations.c:137-310
//
// Schwarz values here are max_{m,n} sqrt-free |(mn|mn)| over the diagonal
// elements of the (MN|MN) block, computed with the library's own integral
// engine.
#include "h2eri_internal.h"

#include <algorithm>
#include <cmath>
#include <omp.h>

namespace h2e {

static double gauss_extent(double alpha, double coef, int am_i, double tol) {
    auto G = [&](double x) { return coef * std::exp(-alpha * x * x) * std::pow(x, (double)am_i) - tol; };
    double lower = 0.0, upper = 20.0;
    if (am_i > 0) {
        lower = std::sqrt(am_i * 0.5 / alpha);
        upper = 10.0 * lower;
    }
    if (G(lower) <= 1e-8 * tol) return 0.0;
    while (G(upper) > 0) upper *= 2.0;
    double ext = 0.5 * (upper + lower), g = G(ext);
    for (int it = 0; it < 200 && std::fabs(g) > 2e-16; it++) {
        if (g > 0.0) lower = ext; else upper = ext;
        ext = 0.5 * (upper + lower);
        g = G(ext);
    }
    return ext;
}

}  // namespace h2e

using namespace h2e;

extern "C" {

void CMS_read_mol_file(const char *fname, int *natom_, int *nshell_, shell_t **shells_) {
    FILE *f = fopen(fname, "r");
    if (!f) { set_error(ERR_IO, "cannot open mol file %s", fname); return; }
    int am_map[128];
    for (int i = 0; i < 128; i++) am_map[i] = -1;
    const char *letters = "SPDFGHIJKL";
    for (int i = 0; letters[i]; i++) am_map[(int)letters[i]] = i;
    int natom = 0;
    std::vector<shell_t> shells;
    bool ok = fscanf(f, "%d", &natom) == 1 && natom > 0;
    for (int a = 0; ok && a < natom; a++) {
        char sym[16];
        int nsh, np_all, npg_all;
        double x, y, z;
        if (fscanf(f, "%15s %d %d %d", sym, &nsh, &np_all, &npg_all) != 4) { ok = false; break; }
        if (fscanf(f, "%lf %lf %lf", &x, &y, &z) != 3) { ok = false; break; }
        for (int s = 0; ok && s < nsh; s++) {
            char type[16];
            int nprim, ngen;
            if (fscanf(f, "%15s %d %d", type, &nprim, &ngen) != 3 || nprim <= 0 || ngen <= 0 ||
                (int)strlen(type) < ngen) { ok = false; break; }
            size_t base = shells.size();
            for (int g = 0; g < ngen; g++) {
                shell_t sh;
                int am = am_map[(int)(unsigned char)type[g] & 127];
                if (am < 0) { ok = false; break; }
                sh.am = am;
                sh.nprim = nprim;
                sh.x = x; sh.y = y; sh.z = z;
                sh.alpha = (double *)malloc(sizeof(double) * nprim);
                sh.coef = (double *)malloc(sizeof(double) * nprim);
                shells.push_back(sh);
            }
            for (int k = 0; ok && k < nprim; k++) {
                double alpha;
                if (fscanf(f, "%lf", &alpha) != 1) { ok = false; break; }
                for (int g = 0; g < ngen; g++) {
                    double c;
                    if (fscanf(f, "%lf", &c) != 1) { ok = false; break; }
                    shells[base + g].alpha[k] = alpha;
                    shells[base + g].coef[k] = c;
                }
            }
        }
    }
    fclose(f);
    if (!ok) {
        for (auto &s : shells) free_shell(s);
        set_error(ERR_IO, "malformed mol file %s", fname);
        return;
    }
    normalize_shells((int)shells.size(), shells.data());
    shell_t *out = (shell_t *)malloc(sizeof(shell_t) * shells.size());
    memcpy(out, shells.data(), sizeof(shell_t) * shells.size());
    *natom_ = natom;
    *nshell_ = (int)shells.size();
    *shells_ = out;
}

void H2ERI_process_shells(H2ERI_p h) {
    if (!h || h->nshell <= 0 || !h->shells) { set_error(ERR_ARG, "H2ERI_process_shells: no shells"); return; }
    const int ns = h->nshell;
    int max_am = 0;
    for (int i = 0; i < ns; i++) max_am = std::max(max_am, h->shells[i].am);
    if (max_am > 3) { set_error(ERR_UNSUPPORTED, "angular momentum > f is not supported"); return; }
    // 1. Schwarz values (MN|MN)
    std::vector<double> scr((size_t)ns * ns, 0.0);
    double gmax = 0.0;
    #pragma omp parallel reduction(max : gmax)
    {
        EriWork *w = eri_work_new(max_am);
        std::vector<double> blk(36 * 36 * 4);
        #pragma omp for schedule(dynamic)
        for (int M = 0; M < ns; M++) {
            for (int N = M; N < ns; N++) {
                PairData *pd = nullptr;
                pair_data_build(h->shells[M], h->shells[N], pd);
                int nb = pair_nbfp(pd);
                if ((int)blk.size() < nb * nb) blk.resize((size_t)nb * nb);
                eri_pair_block(pd, pd, w, blk.data(), nb);
                double mx = 0.0;
                for (int k = 0; k < nb; k++) mx = std::max(mx, std::fabs(blk[(size_t)k * nb + k]));
                scr[(size_t)M * ns + N] = scr[(size_t)N * ns + M] = mx;
                gmax = std::max(gmax, mx);
                pair_data_free(pd);
            }
        }
        eri_work_free(w);
    }
    const double thres = h->scr_tol * h->scr_tol / gmax;
    int np = 0;
    for (int i = 0; i < ns; i++)
        for (int j = i; j < ns; j++)
            if (scr[(size_t)i * ns + j] >= thres) np++;
    h->num_sp = np;
    h->sp_shells = (shell_t *)malloc(sizeof(shell_t) * 2 * (size_t)np);
    h->sp_shell_idx = (int *)malloc(sizeof(int) * 2 * (size_t)np);
    int c = 0;
    for (int i = 0; i < ns; i++)
        for (int j = i; j < ns; j++) {
            if (scr[(size_t)i * ns + j] < thres) continue;
            copy_shell(h->shells[i], h->sp_shells[c]);
            copy_shell(h->shells[j], h->sp_shells[c + np]);
            h->sp_shell_idx[c] = i;
            h->sp_shell_idx[c + np] = j;
            c++;
        }
    // 2. extents and centres
    h->sp_extent = (double *)malloc(sizeof(double) * (size_t)np);
    h->sp_center = (double *)malloc(sizeof(double) * 3 * (size_t)np);
    const double ext_tol = h->ext_tol;
    #pragma omp parallel for schedule(dynamic, 32)
    for (int i = 0; i < np; i++) {
        const shell_t &s0 = h->sp_shells[i], &s1 = h->sp_shells[i + np];
        const int n01 = s0.nprim * s1.nprim, am01 = s0.am + s1.am;
        double dx = s0.x - s1.x, dy = s0.y - s1.y, dz = s0.z - s1.z;
        const double r01 = dx * dx + dy * dy + dz * dz, d01 = std::sqrt(r01);
        const double tol_i = ext_tol / (double)n01;
        std::vector<double> ext(n01), cx(n01), cy(n01), cz(n01);
        double max_ext = 0.0;
        for (int j0 = 0; j0 < s0.nprim; j0++)
            for (int j1 = 0; j1 < s1.nprim; j1++) {
                double a0 = s0.alpha[j0], a1 = s1.alpha[j1], ap = a0 + a1;
                double coef = s0.coef[j0] * s1.coef[j1] * std::exp(-a0 * a1 / ap * r01);
                int j = j0 * s1.nprim + j1;
                ext[j] = gauss_extent(ap, coef, am01, tol_i);
                max_ext = std::max(max_ext, ext[j]);
                cx[j] = (a0 * s0.x + a1 * s1.x) / ap;
                cy[j] = (a0 * s0.y + a1 * s1.y) / ap;
                cz[j] = (a0 * s0.z + a1 * s1.z) / ap;
            }
        double *ctr = h->sp_center;
        if (d01 < 1e-16) {
            h->sp_extent[i] = max_ext;
            ctr[i] = s0.x; ctr[np + i] = s0.y; ctr[2 * (size_t)np + i] = s0.z;
            continue;
        }
        if (max_ext < 1e-16) {
            double sx = 0, sy = 0, sz = 0;
            for (int j = 0; j < n01; j++) { sx += cx[j]; sy += cy[j]; sz += cz[j]; }
            h->sp_extent[i] = 0.0;
            ctr[i] = sx / n01; ctr[np + i] = sy / n01; ctr[2 * (size_t)np + i] = sz / n01;
            continue;
        }
        dx /= d01; dy /= d01; dz /= d01;
        // ends of each non-zero extent segment along the pair axis
        double best0 = 0.0, best1 = 0.0;
        double up[3] = {0, 0, 0}, lo[3] = {0, 0, 0};
        bool have0 = false, have1 = false;
        for (int j = 0; j < n01; j++) {
            if (ext[j] <= 1e-16) continue;
            double u[3] = {cx[j] + ext[j] * dx, cy[j] + ext[j] * dy, cz[j] + ext[j] * dz};
            double l[3] = {cx[j] - ext[j] * dx, cy[j] - ext[j] * dy, cz[j] - ext[j] * dz};
            double d0 = (u[0] - s1.x) * (u[0] - s1.x) + (u[1] - s1.y) * (u[1] - s1.y) + (u[2] - s1.z) * (u[2] - s1.z);
            double d1 = (l[0] - s0.x) * (l[0] - s0.x) + (l[1] - s0.y) * (l[1] - s0.y) + (l[2] - s0.z) * (l[2] - s0.z);
            if (d0 > best0) { best0 = d0; memcpy(up, u, sizeof(u)); have0 = true; }
            if (d1 > best1) { best1 = d1; memcpy(lo, l, sizeof(l)); have1 = true; }
        }
        if (!have0 || !have1) {   // degenerate: fall back to the largest single extent
            h->sp_extent[i] = max_ext;
            ctr[i] = 0.5 * (s0.x + s1.x); ctr[np + i] = 0.5 * (s0.y + s1.y); ctr[2 * (size_t)np + i] = 0.5 * (s0.z + s1.z);
            continue;
        }
        double ex = up[0] - lo[0], ey = up[1] - lo[1], ez = up[2] - lo[2];
        h->sp_extent[i] = 0.5 * std::sqrt(ex * ex + ey * ey + ez * ez);
        ctr[i] = 0.5 * (up[0] + lo[0]);
        ctr[np + i] = 0.5 * (up[1] + lo[1]);
        ctr[2 * (size_t)np + i] = 0.5 * (up[2] + lo[2]);
    }
}

}  // extern "C"
