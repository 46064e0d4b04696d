/*
 * TEST INFRASTRUCTURE ONLY: a driver for the oracle built with AddressSanitizer and
 * UndefinedBehaviorSanitizer (make -C oracle sanitize). It runs small 2D and 3D problems --
 * isotropic and per-axis relaxation maps, the exponential model, sources, air points and sensors
 * on and next to the ghost border -- forward and through the adjoint, in both precisions, so
 * every index the oracle computes is exercised under the sanitizers. It exits non-zero on a
 * non-finite result; the sanitizers abort on any out-of-bounds access or undefined operation.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fw_oracle.c"

static double frand(unsigned* s) {
    *s = *s * 1664525u + 1013904223u;
    return (double)(*s >> 8) / 16777216.0;
}

static int run_case(int ndim, int aniso, int expo, int64_t nx, int64_t ny, int64_t nz, int64_t nt) {
    const int64_t n = nx * ny * (ndim == 3 ? nz : 1);
    unsigned seed = 7u + (unsigned)(ndim * 100 + aniso * 10 + expo);
    fwo_problem P;
    memset(&P, 0, sizeof P);
    P.ndim = ndim;
    P.n_relax = 2;
    P.nx = nx;
    P.ny = ny;
    P.nz = ndim == 3 ? nz : 1;
    P.dx = P.dy = P.dz = 4.3e-5;
    P.dt = 1.1e-8;
    double* K = malloc(sizeof(double) * n);
    double* rho = malloc(sizeof(double) * n);
    double* maps[3][10];
    int32_t* dcmap = malloc(sizeof(int32_t) * n);
    double* aexp = expo ? malloc(sizeof(double) * n) : 0;
    const int nmaps = aniso ? ndim : 1;
    for (int64_t c = 0; c < n; ++c) {
        rho[c] = 950.0 + 100.0 * frand(&seed);
        const double cc = 1450.0 + 170.0 * frand(&seed);
        K[c] = rho[c] * cc * cc;
        dcmap[c] = (int32_t)(frand(&seed) * 3.0);
        if (expo) aexp[c] = 0.999 + 0.001 * frand(&seed);
    }
    for (int a = 0; a < nmaps; ++a)
        for (int m = 0; m < 10; ++m) {
            maps[a][m] = malloc(sizeof(double) * n);
            for (int64_t c = 0; c < n; ++c) {
                /* kappa >= 1 for m = 0, 1; a <= 0 and 0 < b <= 1 for the others */
                const double u = frand(&seed);
                maps[a][m][c] = m < 2 ? 1.0 + 0.1 * u : (m % 2 == 0 ? -0.05 * u : 0.9 + 0.1 * u);
            }
        }
    for (int a = 0; a < ndim && !expo; ++a) {
        double** mm = maps[aniso ? a : 0];
        P.kappa_v[a] = mm[0];
        P.kappa_p[a] = mm[1];
        for (int nu = 0; nu < 2; ++nu) {
            P.a_v[a][nu] = mm[2 + 4 * nu];
            P.b_v[a][nu] = mm[3 + 4 * nu];
            P.a_p[a][nu] = mm[4 + 4 * nu];
            P.b_p[a][nu] = mm[5 + 4 * nu];
        }
    }
    P.K = K;
    P.rho = rho;
    P.aexp = aexp;
    P.dcmap = dcmap;
    double coef[3 * 9];
    /* second-order weights with a small transverse term: stable at this time step */
    for (int s = 0; s < 3; ++s)
        for (int k = 0; k < 9; ++k) coef[s * 9 + k] = (k == 0 ? 0.96 : k == 8 ? 0.01 : 0.0) * (1 + 0.01 * s);
    P.coef = coef;
    P.ndmap = 3;
    /* sources at the first and last updated cells and in the ghost border; air next to a
     * source; sensors everywhere including the grid's last cell */
    int64_t icc[4] = {0, n - 1, n / 2, (int64_t)(8 * (ny * (ndim == 3 ? nz : 1)) + 8 * (ndim == 3 ? nz : 1) + (ndim == 3 ? 8 : 0))};
    const int64_t ntic = nt / 2;
    double* icmat = malloc(sizeof(double) * 4 * ntic);
    for (int64_t q = 0; q < 4 * ntic; ++q) icmat[q] = sin(0.3 * (double)q);
    int64_t zero[2] = {n / 2 + 1, n - 2};
    int64_t outc[5] = {0, n / 2, n / 2 + 1, icc[3], n - 1};
    P.ncoords = 4;
    P.icc = icc;
    P.icmat = icmat;
    P.ntic = ntic;
    P.nzero = 2;
    P.zero = zero;
    P.nout = 5;
    P.outc = outc;
    P.modt = 2;
    const int64_t nrec = (nt + 1) / 2;
    double* rec = calloc((size_t)(nrec * 5), sizeof(double));
    double* grad = malloc(sizeof(double) * nrec * 5);
    double* gK = malloc(sizeof(double) * n);
    double* gr = malloc(sizeof(double) * n);
    for (int64_t q = 0; q < nrec * 5; ++q) grad[q] = frand(&seed) - 0.5;
    int bad = 0;
    for (int prec = 0; prec < 2; ++prec) {
        void* h = prec ? fwo_create_f32(&P) : fwo_create_f64(&P);
        if (!h) return 1;
        if (prec) {
            fwo_run_f32(h, nt, rec);
            fwo_adjoint_f32(h, nt, grad, gK, gr);
            fwo_destroy_f32(h);
        } else {
            fwo_run_f64(h, nt, rec);
            fwo_adjoint_f64(h, nt, grad, gK, gr);
            fwo_destroy_f64(h);
        }
        for (int64_t q = 0; q < nrec * 5; ++q) bad |= !isfinite(rec[q]);
        for (int64_t c = 0; c < n; ++c) bad |= !isfinite(gK[c]) || !isfinite(gr[c]);
    }
    printf("%dD %s%s %lldx%lldx%lld, %lld steps: %s\n", ndim, expo ? "exponential" : aniso ? "per-axis" : "isotropic",
           "", (long long)nx, (long long)ny, (long long)(ndim == 3 ? nz : 1), (long long)nt, bad ? "NON-FINITE" : "ok");
    free(K);
    free(rho);
    free(dcmap);
    free(aexp);
    for (int a = 0; a < nmaps; ++a)
        for (int m = 0; m < 10; ++m) free(maps[a][m]);
    free(icmat);
    free(rec);
    free(grad);
    free(gK);
    free(gr);
    return bad;
}

int main(void) {
    int bad = 0;
    bad |= run_case(2, 0, 0, 37, 41, 1, 30);
    bad |= run_case(2, 1, 0, 35, 40, 1, 25);
    bad |= run_case(2, 0, 1, 36, 38, 1, 25);
    bad |= run_case(3, 0, 0, 21, 19, 23, 12);
    bad |= run_case(3, 1, 0, 19, 22, 20, 11);
    bad |= run_case(3, 0, 1, 20, 20, 19, 10);
    puts(bad ? "sanitize check: FAILED" : "sanitize check: ok");
    return bad;
}
