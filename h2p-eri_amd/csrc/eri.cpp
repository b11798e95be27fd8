// Host integral engine: contracted Cartesian electron-repulsion integrals
// (ERI) and point-charge potential integrals (NAI) for Gaussian shells.
//
// It replaces the SIMINT calls of H2P-ERI (simint_compute_eri /
// simint_compute_potential, used at src/CMS.c:349-352,472-475) with an
// implementation written from the published recurrences:
//   - Boys function F_m(T): 8-term Taylor expansion on a tabulated grid for
//     T <= 40 (table from the convergent series
//     F_m(T) = e^-T sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)) and downward
//     recursion), upward recursion from F_0 = sqrt(pi/T) erf(sqrt T)/2 above.
//   - Obara-Saika vertical recurrence in the Head-Gordon-Pople form
//     (HGP 1988, J. Chem. Phys. 89, 5777) to build [e0|f0], contraction over
//     primitives, then the horizontal recurrence on bra and ket.
//   - Obara-Saika recurrence for the potential integrals
//     (Obara & Saika 1986, J. Chem. Phys. 84, 3963).
// Cartesian order within a shell is lexicographic (xx, xy, xz, yy, yz, zz),
// the SIMINT order.  Shell normalisation follows SIMINT's convention: every
// Cartesian component of a shell carries the normalisation of the x^l
// component (src/CMS.c:118 normalises through SIMINT).
#include "h2eri_internal.h"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace h2e {

// ---------------------------------------------------------------------------
// Cartesian component tables (global numbering: all components of total
// angular momentum 0, then 1, ...; lexicographic within a level).
// ---------------------------------------------------------------------------
static const int LMAX_TAB = 12;                       // up to (ff|ff) in VRR
static const int NCOMP_TAB = (LMAX_TAB + 1) * (LMAX_TAB + 2) * (LMAX_TAB + 3) / 6;

struct CartTab {
    int l[NCOMP_TAB][3];
    int L[NCOMP_TAB];
    int dir[NCOMP_TAB];          // direction decremented to reach the parent
    int dm1[NCOMP_TAB][3];       // index of g - 1_i (-1 if invalid)
    int dp1[NCOMP_TAB][3];       // index of g + 1_i (-1 if beyond table)
    int off[LMAX_TAB + 2];       // first global index of each level
};
static CartTab g_tab;
static std::once_flag g_tab_once;

static inline int ncart(int L) { return (L + 1) * (L + 2) / 2; }
static inline int comp_local(int lx, int ly, int L) {
    int a = L - lx;
    return a * (a + 1) / 2 + (L - lx - ly);
}
static inline int comp_global(int lx, int ly, int lz) {
    int L = lx + ly + lz;
    return g_tab.off[L] + comp_local(lx, ly, L);
}

static void build_tab() {
    int g = 0;
    for (int L = 0; L <= LMAX_TAB; L++) {
        g_tab.off[L] = g;
        for (int lx = L; lx >= 0; lx--)
            for (int ly = L - lx; ly >= 0; ly--) {
                int lz = L - lx - ly;
                g_tab.l[g][0] = lx; g_tab.l[g][1] = ly; g_tab.l[g][2] = lz;
                g_tab.L[g] = L;
                g_tab.dir[g] = lx > 0 ? 0 : (ly > 0 ? 1 : 2);
                g++;
            }
    }
    g_tab.off[LMAX_TAB + 1] = g;
    for (int k = 0; k < g; k++) {
        for (int i = 0; i < 3; i++) {
            int l[3] = {g_tab.l[k][0], g_tab.l[k][1], g_tab.l[k][2]};
            if (l[i] > 0) { l[i]--; g_tab.dm1[k][i] = comp_global(l[0], l[1], l[2]); l[i]++; }
            else g_tab.dm1[k][i] = -1;
            l[i]++;
            g_tab.dp1[k][i] = (g_tab.L[k] + 1 <= LMAX_TAB) ? comp_global(l[0], l[1], l[2]) : -1;
        }
    }
}

// ---------------------------------------------------------------------------
// Boys function
// ---------------------------------------------------------------------------
static const int BOYS_MMAX = 12 + 8 + 1;   // table orders (m_max + Taylor terms)
static const double BOYS_TMAX = 40.0;
static const double BOYS_H = 0.05;
static const int BOYS_NGRID = (int)(BOYS_TMAX / BOYS_H) + 2;
static double *g_boys = nullptr;   // [grid][m]
static std::once_flag g_boys_once;

static void build_boys() {
    g_boys = (double *)malloc(sizeof(double) * (size_t)BOYS_NGRID * (BOYS_MMAX + 1));
    for (int gi = 0; gi < BOYS_NGRID; gi++) {
        long double T = (long double)gi * BOYS_H;
        long double eT = expl(-T);
        // series for the top order
        int M = BOYS_MMAX;
        long double term = 1.0L / (2 * M + 1), sum = term;
        for (int k = 1; k < 2000; k++) {
            term *= 2.0L * T / (long double)(2 * M + 2 * k + 1);
            sum += term;
            if (term < sum * 1e-21L) break;
        }
        long double F[BOYS_MMAX + 1];
        F[M] = eT * sum;
        for (int m = M; m > 0; m--) F[m - 1] = (2.0L * T * F[m] + eT) / (long double)(2 * m - 1);
        for (int m = 0; m <= M; m++) g_boys[(size_t)gi * (BOYS_MMAX + 1) + m] = (double)F[m];
    }
}

// F[0..mmax] at T (mmax <= 12)
void boys_fm(int mmax, double T, double *F) {
    if (T <= BOYS_TMAX) {
        int gi = (int)(T / BOYS_H + 0.5);
        double dT = gi * BOYS_H - T;   // F_m(T) = sum_k F_{m+k}(T0) (T0-T)^k / k!
        const double *row = g_boys + (size_t)gi * (BOYS_MMAX + 1);
        // Taylor for the top order, downward recursion for the rest
        double f = 0.0, pw = 1.0;
        static const double inv_fact[8] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040};
        for (int k = 0; k < 8; k++) { f += row[mmax + k] * pw * inv_fact[k]; pw *= dT; }
        F[mmax] = f;
        if (mmax > 0) {
            double eT = std::exp(-T);
            for (int m = mmax; m > 0; m--) F[m - 1] = (2.0 * T * F[m] + eT) / (double)(2 * m - 1);
        }
    } else {
        double eT = std::exp(-T);
        double inv2T = 0.5 / T;
        F[0] = 0.5 * std::sqrt(M_PI / T) * std::erf(std::sqrt(T));
        for (int m = 0; m < mmax; m++) F[m + 1] = ((2 * m + 1) * F[m] - eT) * inv2T;
    }
}

double boys_f(int m, double T) {
    std::call_once(g_boys_once, build_boys);
    double F[BOYS_MMAX + 1];
    boys_fm(m, T, F);
    return F[m];
}

// ---------------------------------------------------------------------------
// Shells
// ---------------------------------------------------------------------------
static double dfact(int n) {   // (n)!! with (-1)!! = 1
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

// Normalises raw contraction coefficients in place so that the x^am component
// of the contracted function has unit norm.
void normalize_shells(int nshell, shell_t *shells) {
    for (int s = 0; s < nshell; s++) {
        shell_t &sh = shells[s];
        int l = sh.am;
        double df = dfact(2 * l - 1);
        for (int k = 0; k < sh.nprim; k++) {
            double a = sh.alpha[k];
            double np = std::pow(2.0 * a / M_PI, 0.75) * std::pow(4.0 * a, 0.5 * l) / std::sqrt(df);
            sh.coef[k] *= np;
        }
        double s2 = 0.0;
        for (int j = 0; j < sh.nprim; j++)
            for (int k = 0; k < sh.nprim; k++) {
                double ajk = sh.alpha[j] + sh.alpha[k];
                s2 += sh.coef[j] * sh.coef[k] * df / std::pow(2.0 * ajk, l) * std::pow(M_PI / ajk, 1.5);
            }
        double inv = 1.0 / std::sqrt(s2);
        for (int k = 0; k < sh.nprim; k++) sh.coef[k] *= inv;
    }
}

void copy_shell(const shell_t &src, shell_t &dst) {
    dst.am = src.am;
    dst.nprim = src.nprim;
    dst.x = src.x; dst.y = src.y; dst.z = src.z;
    dst.alpha = (double *)malloc(sizeof(double) * (src.nprim > 0 ? src.nprim : 1));
    dst.coef = (double *)malloc(sizeof(double) * (src.nprim > 0 ? src.nprim : 1));
    memcpy(dst.alpha, src.alpha, sizeof(double) * src.nprim);
    memcpy(dst.coef, src.coef, sizeof(double) * src.nprim);
}

void free_shell(shell_t &s) {
    free(s.alpha);
    free(s.coef);
    s.alpha = s.coef = nullptr;
    s.nprim = 0;
}

// ---------------------------------------------------------------------------
// Pair data
// ---------------------------------------------------------------------------
struct PairData {
    int la, lb;            // A = first shell, B = second shell
    double A[3], B[3], AB[3];
    int np;                // kept primitive pairs
    std::vector<double> p, P, K;   // exponent sum, centre (3*np), c_a c_b exp(-ab/p AB^2)
};

void pair_data_build(const shell_t &Ms, const shell_t &Ns, PairData *&pd) {
    // The reference builds the SIMINT pair as (N, M) (src/H2ERI_partition.c:92-96):
    // first shell A = N, second shell B = M.
    std::call_once(g_tab_once, build_tab);
    std::call_once(g_boys_once, build_boys);
    pd = new PairData();
    const shell_t &As = Ns, &Bs = Ms;
    pd->la = As.am; pd->lb = Bs.am;
    pd->A[0] = As.x; pd->A[1] = As.y; pd->A[2] = As.z;
    pd->B[0] = Bs.x; pd->B[1] = Bs.y; pd->B[2] = Bs.z;
    double r2 = 0;
    for (int d = 0; d < 3; d++) { pd->AB[d] = pd->A[d] - pd->B[d]; r2 += pd->AB[d] * pd->AB[d]; }
    pd->np = 0;
    for (int i = 0; i < As.nprim; i++)
        for (int j = 0; j < Bs.nprim; j++) {
            double a = As.alpha[i], b = Bs.alpha[j], p = a + b;
            double K = As.coef[i] * Bs.coef[j] * std::exp(-a * b / p * r2);
            // drop primitive pairs whose overlap-scale prefactor is negligible
            if (std::fabs(K) * std::pow(M_PI / p, 1.5) < 1e-20) continue;
            pd->p.push_back(p);
            for (int d = 0; d < 3; d++) pd->P.push_back((a * pd->A[d] + b * pd->B[d]) / p);
            pd->K.push_back(K);
            pd->np++;
        }
}

void pair_data_free(PairData *&pd) { delete pd; pd = nullptr; }

PairView pair_view(const PairData *pd) {
    PairView v;
    v.la = pd->la; v.lb = pd->lb; v.np = pd->np;
    v.A = pd->A; v.B = pd->B;
    v.p = pd->p.data(); v.P = pd->P.data(); v.K = pd->K.data();
    return v;
}

const double *boys_table(int *ngrid, int *ncol, double *step, double *tmax) {
    std::call_once(g_boys_once, build_boys);
    *ngrid = BOYS_NGRID;
    *ncol = BOYS_MMAX + 1;
    *step = BOYS_H;
    *tmax = BOYS_TMAX;
    return g_boys;
}
int pair_nbfp(const PairData *pd) { return ncart(pd->la) * ncart(pd->lb); }

// ---------------------------------------------------------------------------
// Work space
// ---------------------------------------------------------------------------
struct EriWork {
    std::vector<double> v1, v2, acc, t1, t2;
};
EriWork *eri_work_new(int) { return new EriWork(); }
void eri_work_free(EriWork *w) { delete w; }

// Horizontal recurrence: X holds [e][K] for all components e with
// la <= |e| <= la+lb (indexed e - off[la]).  Produces out[(a*nb+b)][K] for
// |a| = la, |b| = lb using (a, b+1_i| = (a+1_i, b| + AB_i (a, b|.
static void hrr(int la, int lb, const double *AB, const double *X, int K, double *out,
                std::vector<double> &buf) {
    const int na = ncart(la);
    if (lb == 0) {
        memcpy(out, X, sizeof(double) * (size_t)na * K);
        return;
    }
    const int e0 = g_tab.off[la];
    // level j array: [e (|e| in la..la+lb-j)][b (|b| = j)][K]
    int ne_prev = g_tab.off[la + lb + 1] - e0;
    int nb_prev = 1;
    std::vector<double> prev(X, X + (size_t)ne_prev * K);
    for (int j = 1; j <= lb; j++) {
        int ne = g_tab.off[la + lb - j + 1] - e0;
        int nb = ncart(j);
        buf.assign((size_t)ne * nb * K, 0.0);
        for (int bl = 0; bl < nb; bl++) {
            int bg = g_tab.off[j] + bl;
            int i = g_tab.dir[bg];
            int bpg = g_tab.dm1[bg][i];
            int bpl = bpg - g_tab.off[j - 1];
            double ab = AB[i];
            for (int el = 0; el < ne; el++) {
                int eg = e0 + el;
                int epg = g_tab.dp1[eg][i];
                int epl = epg - e0;
                const double *s1 = prev.data() + ((size_t)epl * nb_prev + bpl) * K;
                const double *s0 = prev.data() + ((size_t)el * nb_prev + bpl) * K;
                double *d = buf.data() + ((size_t)el * nb + bl) * K;
                for (int k = 0; k < K; k++) d[k] = s1[k] + ab * s0[k];
            }
        }
        prev.swap(buf);
        nb_prev = nb;
    }
    // prev: [a (|a| = la)][b][K]
    memcpy(out, prev.data(), sizeof(double) * (size_t)na * ncart(lb) * K);
}

// ---------------------------------------------------------------------------
// ERI of one shell quartet, written to out (nbfp(bra) x nbfp(ket), ld ldo)
// ---------------------------------------------------------------------------
void eri_pair_block(const PairData *bra, const PairData *ket, EriWork *w, double *out, int ldo) {
    const int la = bra->la, lb = bra->lb, lc = ket->la, ld = ket->lb;
    const int Lab = la + lb, Lcd = lc + ld, Lt = Lab + Lcd;
    const int ne = g_tab.off[Lab + 1];          // all e comps with |e| <= Lab
    const int nf = g_tab.off[Lcd + 1];
    const int eo = g_tab.off[la], fo = g_tab.off[lc];
    const int ner = ne - eo, nfr = nf - fo;     // needed ranges
    w->acc.assign((size_t)nfr * ner, 0.0);
    w->v1.resize((size_t)(Lt + 1) * ne);
    w->v2.resize((size_t)(Lcd + 1) * nf * ne);
    double *V1 = w->v1.data();
    double *V2 = w->v2.data();
    double F[32];
    const double pref0 = 2.0 * std::pow(M_PI, 2.5);
    for (int ib = 0; ib < bra->np; ib++) {
        const double p = bra->p[ib];
        const double *P = &bra->P[3 * ib];
        const double Kab = bra->K[ib];
        double PA[3] = {P[0] - bra->A[0], P[1] - bra->A[1], P[2] - bra->A[2]};
        for (int ik = 0; ik < ket->np; ik++) {
            const double q = ket->p[ik];
            const double *Q = &ket->P[3 * ik];
            const double pq = p + q;
            const double rho = p * q / pq;
            double PQ2 = 0;
            double W[3], WP[3], WQ[3], QC[3];
            for (int d = 0; d < 3; d++) {
                double t = P[d] - Q[d];
                PQ2 += t * t;
                W[d] = (p * P[d] + q * Q[d]) / pq;
                WP[d] = W[d] - P[d];
                WQ[d] = W[d] - Q[d];
                QC[d] = Q[d] - ket->A[d];
            }
            const double T = rho * PQ2;
            const double pre = pref0 / (p * q * std::sqrt(pq)) * Kab * ket->K[ik];
            boys_fm(Lt, T, F);
            // stage 1: [e|0]^(m)
            for (int m = 0; m <= Lt; m++) V1[(size_t)m * ne] = pre * F[m];
            const double i2p = 0.5 / p, rp = rho / p;
            for (int L = 1; L <= Lab; L++) {
                for (int g = g_tab.off[L]; g < g_tab.off[L + 1]; g++) {
                    int i = g_tab.dir[g];
                    int e = g_tab.dm1[g][i];
                    int ei = g_tab.l[e][i];
                    int e2 = ei > 0 ? g_tab.dm1[e][i] : -1;
                    double c2 = ei * i2p;
                    for (int m = 0; m <= Lt - L; m++) {
                        double v = PA[i] * V1[(size_t)m * ne + e] + WP[i] * V1[(size_t)(m + 1) * ne + e];
                        if (e2 >= 0) v += c2 * (V1[(size_t)m * ne + e2] - rp * V1[(size_t)(m + 1) * ne + e2]);
                        V1[(size_t)m * ne + g] = v;
                    }
                }
            }
            // stage 2: [e|f]^(m), layout V2[(m*nf + f)*ne + e]
            for (int m = 0; m <= Lcd; m++)
                memcpy(V2 + ((size_t)m * nf + 0) * ne, V1 + (size_t)m * ne, sizeof(double) * ne);
            const double i2q = 0.5 / q, rq = rho / q, i2pq = 0.5 / pq;
            for (int k = 1; k <= Lcd; k++) {
                for (int f = g_tab.off[k]; f < g_tab.off[k + 1]; f++) {
                    int i = g_tab.dir[f];
                    int f1 = g_tab.dm1[f][i];
                    int fi = g_tab.l[f1][i];
                    int f2 = fi > 0 ? g_tab.dm1[f1][i] : -1;
                    double cf = fi * i2q;
                    for (int m = 0; m <= Lcd - k; m++) {
                        double *dst = V2 + ((size_t)m * nf + f) * ne;
                        const double *a0 = V2 + ((size_t)m * nf + f1) * ne;
                        const double *a1 = V2 + ((size_t)(m + 1) * nf + f1) * ne;
                        for (int e = 0; e < ne; e++) dst[e] = QC[i] * a0[e] + WQ[i] * a1[e];
                        if (f2 >= 0) {
                            const double *b0 = V2 + ((size_t)m * nf + f2) * ne;
                            const double *b1 = V2 + ((size_t)(m + 1) * nf + f2) * ne;
                            for (int e = 0; e < ne; e++) dst[e] += cf * (b0[e] - rq * b1[e]);
                        }
                        for (int e = 1; e < ne; e++) {
                            int ei = g_tab.l[e][i];
                            if (ei > 0) dst[e] += ei * i2pq * a1[g_tab.dm1[e][i]];
                        }
                    }
                }
            }
            // contract needed [e|f]^(0)
            for (int f = fo; f < nf; f++) {
                const double *src = V2 + (size_t)f * ne + eo;
                double *dst = w->acc.data() + (size_t)(f - fo) * ner;
                for (int e = 0; e < ner; e++) dst[e] += src[e];
            }
        }
    }
    // ket HRR on acc[f][e] with K = ner  ->  t1[(c,d)][e]
    const int nc = ncart(lc), nd = ncart(ld), na = ncart(la), nbb = ncart(lb);
    const int ncd = nc * nd, nab = na * nbb;
    w->t1.resize((size_t)ncd * ner);
    std::vector<double> hbuf;
    hrr(lc, ld, ket->AB, w->acc.data(), ner, w->t1.data(), hbuf);
    // transpose to [e][(c,d)]
    w->t2.resize((size_t)ner * ncd);
    for (int cd = 0; cd < ncd; cd++)
        for (int e = 0; e < ner; e++) w->t2[(size_t)e * ncd + cd] = w->t1[(size_t)cd * ner + e];
    // bra HRR -> [(a,b)][(c,d)]
    w->t1.resize((size_t)nab * ncd);
    hrr(la, lb, bra->AB, w->t2.data(), ncd, w->t1.data(), hbuf);
    for (int ab = 0; ab < nab; ab++) memcpy(out + (size_t)ab * ldo, w->t1.data() + (size_t)ab * ncd, sizeof(double) * ncd);
}

// ---------------------------------------------------------------------------
// Potential integrals (a b | 1/|r - C|) of one pair against npt unit charges
// ---------------------------------------------------------------------------
void nai_pair_block(const PairData *pd, int npt, const double *px, const double *py, const double *pz,
                    EriWork *w, double *out, int ldo) {
    const int la = pd->la, lb = pd->lb, Lab = la + lb;
    const int ne = g_tab.off[Lab + 1], eo = g_tab.off[la], ner = ne - eo;
    const int na = ncart(la), nb = ncart(lb);
    w->acc.resize((size_t)ner);
    w->v1.resize((size_t)(Lab + 1) * ne);
    w->t1.resize((size_t)na * nb);
    double *V = w->v1.data();
    double F[32];
    std::vector<double> hbuf;
    for (int ip = 0; ip < npt; ip++) {
        const double C[3] = {px[ip], py[ip], pz[ip]};
        std::fill(w->acc.begin(), w->acc.end(), 0.0);
        for (int k = 0; k < pd->np; k++) {
            const double p = pd->p[k];
            const double *P = &pd->P[3 * k];
            double PA[3], PC[3], PC2 = 0;
            for (int d = 0; d < 3; d++) {
                PA[d] = P[d] - pd->A[d];
                PC[d] = P[d] - C[d];
                PC2 += PC[d] * PC[d];
            }
            boys_fm(Lab, p * PC2, F);
            const double pre = 2.0 * M_PI / p * pd->K[k];
            for (int m = 0; m <= Lab; m++) V[(size_t)m * ne] = pre * F[m];
            const double i2p = 0.5 / p;
            for (int L = 1; L <= Lab; L++)
                for (int g = g_tab.off[L]; g < g_tab.off[L + 1]; g++) {
                    int i = g_tab.dir[g];
                    int e = g_tab.dm1[g][i];
                    int ei = g_tab.l[e][i];
                    int e2 = ei > 0 ? g_tab.dm1[e][i] : -1;
                    for (int m = 0; m <= Lab - L; m++) {
                        double v = PA[i] * V[(size_t)m * ne + e] - PC[i] * V[(size_t)(m + 1) * ne + e];
                        if (e2 >= 0) v += ei * i2p * (V[(size_t)m * ne + e2] - V[(size_t)(m + 1) * ne + e2]);
                        V[(size_t)m * ne + g] = v;
                    }
                }
            for (int e = 0; e < ner; e++) w->acc[e] += V[eo + e];
        }
        hrr(la, lb, pd->AB, w->acc.data(), 1, w->t1.data(), hbuf);
        for (int ab = 0; ab < na * nb; ab++) out[(size_t)ab * ldo + ip] = w->t1[ab];
    }
}

}  // namespace h2e
