/*
 * TEST INFRASTRUCTURE ONLY. This is the oracle: a plain-C restatement of the Fullwave 2.5
 * time-stepping engine, used by tests/, the smoke check and bench.py's CPU-baseline leg to
 * check and time the HIP engine. Nothing in the product links or calls it.
 *
 * The reference ships no engine source (its engine is an absent CUDA executable launched at
 * fullwave/solver/launcher.py:197-215), so this file restates the update the reference's
 * input contract describes (fullwave/solver/input_file_writer.py:563-881) and its formulation
 * (README.md:36-44, Komatitsch & Martin 2007 CPML as cited at fullwave/solver/pml_builder.py:
 * 850-853):
 *
 *   leapfrog, v at n+1/2 and p at n+1; for every axis a (h_a the spacing):
 *     D+_a p = (1/h_a) [ sum_{k=1..8} w_k (p(+k) - p(-k+1))
 *                        + e * sum_{b != a} sum_{s=+-1} (p(+1, s e_b) - p(0, s e_b)) ]
 *     psi_a,nu <- b_v,nu psi_a,nu + a_v,nu D+_a p              (nu = 1, 2)
 *     v_a      <- v_a - dt / rho_bar_a (D+_a p / kappa_v + sum_nu psi_a,nu)
 *   then
 *     D-_a v_a = (1/h_a) [ sum_k w_k (v_a(+k-1) - v_a(-k))
 *                          + e * sum_{b != a} sum_s (v_a(0, s e_b) - v_a(-1, s e_b)) ]
 *     phi_a,nu <- b_p,nu phi_a,nu + a_p,nu D-_a v_a
 *     p        <- p - dt K sum_a (D-_a v_a / kappa_p + sum_nu phi_a,nu)
 *   then p[source s] <- icmat[s][n] (n < nTic), p[air] <- 0, and if n % modT == 0 the
 *   sensors record p.
 *
 * Exponential-attenuation model (a_exp.dat, fullwave/solver/input_file_writer.py:629-644): no
 * memory variables, kappa = 1, and both updates are scaled by the voxel's decay factor A:
 *     v_a <- A (v_a - dt / rho_bar_a D+_a p),   p <- A (p - dt K sum_a D-_a v_a).
 * Scaling both fields makes the leapfrog's per-step amplification matrix have determinant A^2,
 * so every mode decays by exactly A per step; A = exp(-dt alpha(f0) c0) with alpha in Np/m
 * (fullwave/medium.py:1068-1090) is then a frequency-independent attenuation alpha(f0).
 *
 * w_1..w_8, e come from the per-voxel coefficient set dcmap[c] (dmap table); a staggered point
 * c + 1/2 e_a uses voxel c's set and maps. rho_bar_a = (rho[c] + rho[c + e_a]) / 2. Only cells
 * with every coordinate in [8, n - 8) are updated; the 8-cell ghost border stays zero.
 *
 * The same source is compiled twice, with REAL = double and REAL = float.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int32_t ndim;
    int32_t n_relax; /* must be 2 */
    int64_t nx, ny, nz;
    double dx, dy, dz, dt;
    const double* K;
    const double* rho;
    const double* kappa_v[3]; /* velocity side, per axis (isotropic: same map) */
    const double* kappa_p[3]; /* pressure side */
    const double* a_v[3][2];
    const double* b_v[3][2];
    const double* a_p[3][2];
    const double* b_p[3][2];
    const int32_t* dcmap;
    const double* coef; /* [ndmap][9]: w_1..w_8, e */
    int64_t ndmap;
    int64_t ncoords;
    const int64_t* icc; /* linear indices */
    const double* icmat; /* [ncoords][ntic] */
    int64_t ntic;
    int64_t nzero;
    const int64_t* zero;
    int64_t nout;
    const int64_t* outc;
    int64_t modt;
    const double* aexp; /* exponential-attenuation model: per-voxel decay factor; NULL for the
                           relaxation model (then kappa/a/b are used and aexp is not) */
} fwo_problem;

#define REAL double
#define SFX(name) name##_f64
#include "fw_oracle_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX(name) name##_f32
#include "fw_oracle_impl.h"
#undef REAL
#undef SFX

int fwo_abi(void) { return 2; }
