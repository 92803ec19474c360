/* Timing of the LASPack-compatible boundary (include/laspack.h) with the
 * call pattern of the reference's pressure step (src/navier_pressure.cpp:
 * 9-76, 150-192): every outer iteration constructs a symmetric QMatrix,
 * fills it row by row with Q_SetLen / Q_SetEntry (upper triangle, 1-based,
 * diagonal first), zeroes x, solves with CGIter + ILUPrecond and destroys the
 * matrix. The matrix is the 7-point Laplacian of an N^3 grid (diagonal
 * 6 + 1e-3, neighbours -1), with a fixed pseudo-random right-hand side.
 *
 * Prints one JSON line: host assembly time (the caller's own loop), the
 * CGIter wall time of the first call on a new matrix (host packing, colouring
 * of a new structure, PCIe transfers, device solve, download) and of a
 * second call on the unchanged matrix (transfers of x and b plus the device
 * solve), iterations and the true relative residual.
 *
 * usage: laspack_bench [N] [outer iterations]
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "laspack.h"

static double now_ms(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

int main(int argc, char **argv) {
    const size_t N = argc > 1 ? (size_t)atol(argv[1]) : 146;
    const int outer = argc > 2 ? atoi(argv[2]) : 2;
    const size_t n = N * N * N;
    Vector x, b;
    V_Constr(&x, "x", n, Normal, True);
    V_Constr(&b, "b", n, Normal, True);
    unsigned long long s = 88172645463325252ull;
    for (size_t i = 1; i <= n; ++i) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        b.Cmp[i] = (double)(s >> 11) / 9007199254740992.0 - 0.5;
    }
    double t_asm = 0., t_first = 0., t_second = 0.;
    int it1 = 0, it2 = 0;
    double acc1 = 0., acc2 = 0.;
    for (int o = 0; o < outer; ++o) {
        QMatrix A;
        double t0 = now_ms();
        Q_Constr(&A, "A", n, True, Rowws, Normal, True);
        for (size_t k = 0; k < N; ++k)
            for (size_t j = 0; j < N; ++j)
                for (size_t i = 0; i < N; ++i) {
                    const size_t r = (k * N + j) * N + i + 1;
                    size_t nb[3], m = 0;
                    if (i + 1 < N) nb[m++] = r + 1;
                    if (j + 1 < N) nb[m++] = r + N;
                    if (k + 1 < N) nb[m++] = r + N * N;
                    Q_SetLen(&A, r, m + 1);
                    Q_SetEntry(&A, r, 0, r, 6.001);
                    for (size_t e = 0; e < m; ++e) Q_SetEntry(&A, r, e + 1, nb[e], -1.0);
                }
        double t1 = now_ms();
        V_SetAllCmp(&x, 0.);
        SetRTCAccuracy(1e-8);
        CGIter(&A, &x, &b, 5000, ILUPrecond, 1.0);
        double t2 = now_ms();
        it1 = GetLastNoIter();
        acc1 = GetLastAccuracy();
        V_SetAllCmp(&x, 0.);
        CGIter(&A, &x, &b, 5000, ILUPrecond, 1.0);
        double t3 = now_ms();
        it2 = GetLastNoIter();
        acc2 = GetLastAccuracy();
        Q_Destr(&A);
        if (o > 0 || outer == 1) {  /* the first iteration also loads the GPU runtime */
            t_asm += t1 - t0;
            t_first += t2 - t1;
            t_second += t3 - t2;
        }
    }
    const int timed = outer > 1 ? outer - 1 : 1;
    if (LASResult() != LASOK) {
        fprintf(stderr, "laspack_bench: LASPack error %d\n", (int)LASResult());
        return 1;
    }
    printf("{\"n\": %zu, \"outer_timed\": %d, \"assembly_ms\": %.3f, \"cgiter_new_matrix_ms\": %.3f, "
           "\"cgiter_same_matrix_ms\": %.3f, \"iterations\": [%d, %d], \"true_relres\": [%.3e, %.3e]}\n",
           n, timed, t_asm / timed, t_first / timed, t_second / timed, it1, it2, acc1, acc2);
    V_Destr(&x);
    V_Destr(&b);
    return 0;
}
