/*
 * A C caller of libh2eri_amd.so through include/H2ERI.h only, compiled with
 * gcc as a user of the reference's C API would compile theirs (test
 * infrastructure: the GPU tests run it).  It makes the calls such a caller
 * makes, in their order:
 *
 *   H2ERI_init -> CMS_read_mol_file (into the handle's natom / nshell /
 *   shells) -> H2ERI_process_shells -> H2ERI_partition -> H2ERI_build_H2
 *   -> H2ERI_build_Coulomb (warm-up + timed calls, reading n_matvec and the
 *   phase timers) and/or H2ERI_build_exchange -> H2ERI_print_statistic ->
 *   H2ERI_destroy
 *
 * (declarations: src/H2ERI_typedef.h:169-179, src/CMS.h:49,
 * src/H2ERI_shell_operations.h:25, src/H2ERI_partition.h:28,
 * src/H2ERI_build_H2.h:16, src/H2ERI_build_Coulomb.h:16,
 * src/H2ERI_build_exchange.h:16 of H2P-ERI).
 *
 * usage: h2eri_c_driver MOL D_BIN reltol J|K REF_BIN [ncall [BD_JIT]]
 * D_BIN and REF_BIN hold num_bf x num_bf float64 (row-major, raw); BD_JIT
 * (default 0) is passed to H2ERI_build_H2.  Prints
 * one line "result=J|K nbf=... relerr=... n_matvec=... ms_per_call=..." and
 * exits 0, or prints the library's error and exits 1.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "H2ERI.h"

static double *read_matrix(const char *path, size_t n) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    double *m = (double *)malloc(sizeof(double) * n);
    const size_t got = m ? fread(m, sizeof(double), n, f) : 0;
    fclose(f);
    if (got != n) {
        free(m);
        return NULL;
    }
    return m;
}

static int check(const char *where) {
    if (h2eri_last_error() == 0) return 0;
    fprintf(stderr, "%s: error %d: %s\n", where, h2eri_last_error(), h2eri_last_error_msg());
    return 1;
}

static double now_ms(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

int main(int argc, char **argv) {
    if (argc < 6 || (strcmp(argv[4], "J") != 0 && strcmp(argv[4], "K") != 0)) {
        fprintf(stderr, "usage: %s MOL D_BIN reltol J|K REF_BIN [ncall [BD_JIT]]\n", argv[0]);
        return 2;
    }
    const double reltol = atof(argv[3]);
    const int want_j = strcmp(argv[4], "J") == 0;
    const int ncall = argc > 6 ? atoi(argv[6]) : 3;
    const int bd_jit = argc > 7 ? atoi(argv[7]) : 0;

    simint_init();
    H2ERI_p h2eri;
    H2ERI_init(&h2eri, 1e-10, 1e-10, reltol);
    CMS_read_mol_file(argv[1], &h2eri->natom, &h2eri->nshell, &h2eri->shells);
    if (check("CMS_read_mol_file")) return 1;
    H2ERI_process_shells(h2eri);
    H2ERI_partition(h2eri);
    H2ERI_build_H2(h2eri, bd_jit);
    if (check("construction")) return 1;

    const int nbf = h2eri->num_bf;
    const size_t nn = (size_t)nbf * nbf;
    double *D = read_matrix(argv[2], nn);
    double *ref = read_matrix(argv[5], nn);
    double *out = (double *)malloc(sizeof(double) * nn);
    if (!D || !ref || !out) {
        fprintf(stderr, "cannot read %s / %s as %d x %d float64\n", argv[2], argv[5], nbf, nbf);
        return 1;
    }

    double ms = 0.0;
    if (want_j) {
        H2ERI_build_Coulomb(h2eri, D, out);   /* warm-up: uploads the representation */
        if (check("H2ERI_build_Coulomb")) return 1;
        h2eri->n_matvec = 0;
        for (int i = 4; i < 9; i++) h2eri->timers[i] = 0.0;
        const double t0 = now_ms();
        for (int c = 0; c < ncall; c++) H2ERI_build_Coulomb(h2eri, D, out);
        ms = (now_ms() - t0) / ncall;
        if (check("H2ERI_build_Coulomb")) return 1;
        if (h2eri->n_matvec != ncall) {
            fprintf(stderr, "n_matvec %d after %d calls\n", h2eri->n_matvec, ncall);
            return 1;
        }
    } else {
        H2ERI_exchange_workbuf_init(h2eri);
        const double t0 = now_ms();
        for (int c = 0; c < ncall; c++) H2ERI_build_exchange(h2eri, D, out);
        ms = (now_ms() - t0) / ncall;
        if (check("H2ERI_build_exchange")) return 1;
        if (!h2eri->dlist_sidx || h2eri->dlist_sidx[h2eri->nshell] <= 0) {
            fprintf(stderr, "the exchange build left no density screening list\n");
            return 1;
        }
        H2ERI_exchange_workbuf_free(h2eri);
    }
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < nn; i++) {
        const double d = out[i] - ref[i];
        num += d * d;
        den += ref[i] * ref[i];
    }
    H2ERI_print_statistic(h2eri);
    printf("result=%s nbf=%d relerr=%.6e n_matvec=%d ms_per_call=%.4f\n", want_j ? "J" : "K", nbf,
           den > 0 ? sqrt(num / den) : 0.0, h2eri->n_matvec, ms);
    fflush(stdout);
    free(D);
    free(ref);
    free(out);
    H2ERI_destroy(h2eri);
    simint_finalize();
    return 0;
}
