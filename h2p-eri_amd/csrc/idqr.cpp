// Interpolative decomposition over rows, A ~= U * A(J,:), through a
// column-pivoted Householder QR of A^T.
//
// Follows the algorithm of H2P-ERI's H2E_ID_compress / H2E_partial_pivot_QR
// (src/H2E_ID_compress.c:51-184, 433-487, 491-576): greedy pivoting on the
// largest remaining row norm, stop when that norm falls below
// max(sqrt(ncol) * 1e-15, tol) times the largest initial row norm, norm
// downdating with exact recomputation once the downdate becomes small,
// U(J,:) = I and U(rest,:) = (R11^{-1} R12)^T, skeleton rows sorted
// ascending with U's columns reordered to match.
#include "h2eri_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace h2e {

static inline double nrm2(int n, const double *x) {
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; i++) {
        double a = std::fabs(x[i]);
        if (a == 0.0) continue;
        if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
        else ssq += (a / scale) * (a / scale);
    }
    return scale * std::sqrt(ssq);
}

void id_compress_rows(int nrow, int ncol, double *A, double tol, std::vector<double> &U, std::vector<int> &J) {
    // View: M = A^T is ncol x nrow column-major; column j of M is row j of A.
    const int m = ncol, n = nrow;
    const int ldm = ncol;
    std::vector<int> p(n);
    std::vector<double> cn(n);
    std::iota(p.begin(), p.end(), 0);
    double norm_p = 0.0;
    int piv = 0;
    for (int j = 0; j < n; j++) {
        cn[j] = nrm2(m, A + (size_t)j * ldm);
        if (cn[j] > norm_p) { norm_p = cn[j]; piv = j; }
    }
    const int max_iter = std::min(m, n);
    const double stop_norm = std::max(std::sqrt((double)m) * 1e-15, tol) * norm_p;
    std::vector<double> tmpcol(m);
    int rank = max_iter;
    for (int i = 0; i < max_iter; i++) {
        if (norm_p < stop_norm) { rank = i; break; }
        if (i != piv) {
            std::swap(p[i], p[piv]);
            std::swap(cn[i], cn[piv]);
            std::swap_ranges(A + (size_t)i * ldm, A + (size_t)i * ldm + m, A + (size_t)piv * ldm);
        }
        double *h = A + (size_t)i * ldm + i;
        const int hl = m - i;
        const double sign = h[0] > 0.0 ? 1.0 : -1.0;
        const double hn = nrm2(hl, h);
        h[0] += sign * hn;
        const double hn2 = nrm2(hl, h);
        const double inv = hn2 > 0.0 ? 1.0 / hn2 : 0.0;
        for (int k = 0; k < hl; k++) h[k] *= inv;
        for (int j = i + 1; j < n; j++) {
            double *c = A + (size_t)j * ldm + i;
            double d = 0.0;
            for (int k = 0; k < hl; k++) d += h[k] * c[k];
            d *= 2.0;
            for (int k = 0; k < hl; k++) c[k] -= d * h[k];
            if (cn[j] < stop_norm) { cn[j] = 0.0; continue; }
            double t = cn[j] * cn[j] - c[0] * c[0];
            cn[j] = (t <= 1e-10) ? nrm2(hl - 1, c + 1) : std::sqrt(t);
        }
        h[0] = -sign * hn;
        for (int k = 1; k < hl; k++) h[k] = 0.0;
        piv = i + 1;
        norm_p = 0.0;
        for (int j = i + 1; j < n; j++)
            if (cn[j] > norm_p) { norm_p = cn[j]; piv = j; }
    }
    const int r = rank;
    J.assign(p.begin(), p.begin() + r);
    U.assign((size_t)n * r, 0.0);
    if (r == 0) return;
    // E = R11^{-1} R12 (r x (n-r)); R(k, j) = A[j*ldm + k]
    const int nr = n - r;
    std::vector<double> E((size_t)r * (nr > 0 ? nr : 1));
    for (int c = 0; c < nr; c++) {
        const double *rc = A + (size_t)(r + c) * ldm;
        for (int k = r - 1; k >= 0; k--) {
            double s = rc[k];
            for (int l = k + 1; l < r; l++) s -= A[(size_t)l * ldm + k] * E[(size_t)l * nr + c];
            E[(size_t)k * nr + c] = s / A[(size_t)k * ldm + k];
        }
    }
    // sort skeleton indices, remember where each pivot column goes
    std::vector<int> ord(r);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return J[a] < J[b]; });
    std::vector<int> colpos(r);
    for (int k = 0; k < r; k++) colpos[ord[k]] = k;
    std::vector<int> Js(r);
    for (int k = 0; k < r; k++) Js[k] = J[ord[k]];
    for (int k = 0; k < r; k++) U[(size_t)p[k] * r + colpos[k]] = 1.0;
    for (int c = 0; c < nr; c++) {
        double *urow = U.data() + (size_t)p[r + c] * r;
        for (int k = 0; k < r; k++) urow[colpos[k]] = E[(size_t)k * nr + c];
    }
    J.swap(Js);
}

}  // namespace h2e
