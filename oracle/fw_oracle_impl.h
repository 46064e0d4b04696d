/*
 * TEST INFRASTRUCTURE ONLY: body of the oracle, included by fw_oracle.c once per precision
 * with REAL and SFX defined. The update is described at the top of fw_oracle.c.
 */

typedef struct {
    fwo_problem P;
    int64_t n; /* cells */
    int64_t step;
    int64_t s[3];
    REAL h[3];
    REAL dt;
    REAL* p;
    REAL* v[3];
    REAL* psi[3][2];
    REAL* phi[3][2];
    REAL* K;
    REAL* rho;
    REAL* kv[3];
    REAL* kp[3];
    REAL* av[3][2];
    REAL* bv[3][2];
    REAL* ap[3][2];
    REAL* bp[3][2];
    REAL* coef;
    REAL* aexp; /* NULL unless the exponential-attenuation model */
} SFX(fwo_state);

static REAL* SFX(fwo_conv)(const double* src, int64_t n) {
    if (!src) return 0;
    REAL* dst = (REAL*)malloc(sizeof(REAL) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) dst[i] = (REAL)src[i];
    return dst;
}

void* SFX(fwo_create)(const fwo_problem* P) {
    SFX(fwo_state)* st = (SFX(fwo_state)*)calloc(1, sizeof(SFX(fwo_state)));
    if (!st) return 0;
    st->P = *P;
    int64_t nz = P->ndim == 3 ? P->nz : 1;
    st->n = P->nx * P->ny * nz;
    st->s[0] = P->ny * nz;
    st->s[1] = nz;
    st->s[2] = 1;
    st->h[0] = (REAL)P->dx;
    st->h[1] = (REAL)P->dy;
    st->h[2] = (REAL)P->dz;
    st->dt = (REAL)P->dt;
    st->p = (REAL*)calloc((size_t)st->n, sizeof(REAL));
    for (int a = 0; a < P->ndim; ++a) {
        st->v[a] = (REAL*)calloc((size_t)st->n, sizeof(REAL));
        st->kv[a] = SFX(fwo_conv)(P->kappa_v[a], st->n);
        st->kp[a] = SFX(fwo_conv)(P->kappa_p[a], st->n);
        for (int nu = 0; nu < 2; ++nu) {
            st->psi[a][nu] = (REAL*)calloc((size_t)st->n, sizeof(REAL));
            st->phi[a][nu] = (REAL*)calloc((size_t)st->n, sizeof(REAL));
            st->av[a][nu] = SFX(fwo_conv)(P->a_v[a][nu], st->n);
            st->bv[a][nu] = SFX(fwo_conv)(P->b_v[a][nu], st->n);
            st->ap[a][nu] = SFX(fwo_conv)(P->a_p[a][nu], st->n);
            st->bp[a][nu] = SFX(fwo_conv)(P->b_p[a][nu], st->n);
        }
    }
    st->K = SFX(fwo_conv)(P->K, st->n);
    st->rho = SFX(fwo_conv)(P->rho, st->n);
    st->coef = SFX(fwo_conv)(P->coef, P->ndmap * 9);
    st->aexp = SFX(fwo_conv)(P->aexp, st->n);
    return st;
}

void SFX(fwo_destroy)(void* handle) {
    SFX(fwo_state)* st = (SFX(fwo_state)*)handle;
    if (!st) return;
    free(st->p);
    for (int a = 0; a < st->P.ndim; ++a) {
        free(st->v[a]);
        free(st->kv[a]);
        free(st->kp[a]);
        for (int nu = 0; nu < 2; ++nu) {
            free(st->psi[a][nu]);
            free(st->phi[a][nu]);
            free(st->av[a][nu]);
            free(st->bv[a][nu]);
            free(st->ap[a][nu]);
            free(st->bp[a][nu]);
        }
    }
    free(st->K);
    free(st->rho);
    free(st->coef);
    free(st->aexp);
    free(st);
}

static inline int SFX(fwo_active)(const SFX(fwo_state)* st, int64_t i, int64_t j, int64_t k) {
    const fwo_problem* P = &st->P;
    if (i < 8 || i >= P->nx - 8 || j < 8 || j >= P->ny - 8) return 0;
    if (P->ndim == 3 && (k < 8 || k >= P->nz - 8)) return 0;
    return 1;
}

/* velocity half step: v^{n+1/2} from p^n */
static void SFX(fwo_velocity)(SFX(fwo_state)* st) {
    const int nd = st->P.ndim;
    const int64_t nx = st->P.nx, ny = st->P.ny, nz = nd == 3 ? st->P.nz : 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nx; ++i) {
        for (int64_t j = 0; j < ny; ++j) {
            for (int64_t k = 0; k < nz; ++k) {
                if (!SFX(fwo_active)(st, i, j, k)) continue;
                const int64_t c = (i * ny + j) * nz + k;
                const REAL* w = st->coef + 9 * (int64_t)st->P.dcmap[c];
                for (int a = 0; a < nd; ++a) {
                    const int64_t sa = st->s[a];
                    REAL acc = 0;
                    for (int m = 1; m <= 8; ++m)
                        acc += w[m - 1] * (st->p[c + m * sa] - st->p[c - (m - 1) * sa]);
                    REAL tr = 0;
                    for (int b = 0; b < nd; ++b) {
                        if (b == a) continue;
                        const int64_t sb = st->s[b];
                        tr += (st->p[c + sa + sb] - st->p[c + sb]) + (st->p[c + sa - sb] - st->p[c - sb]);
                    }
                    const REAL D = (acc + w[8] * tr) / st->h[a];
                    const REAL rho_bar = (REAL)0.5 * (st->rho[c] + st->rho[c + sa]);
                    if (st->aexp) {
                        st->v[a][c] = st->aexp[c] * (st->v[a][c] - st->dt / rho_bar * D);
                        continue;
                    }
                    REAL mem = 0;
                    for (int nu = 0; nu < 2; ++nu) {
                        REAL q = st->bv[a][nu][c] * st->psi[a][nu][c] + st->av[a][nu][c] * D;
                        st->psi[a][nu][c] = q;
                        mem += q;
                    }
                    st->v[a][c] -= st->dt / rho_bar * (D / st->kv[a][c] + mem);
                }
            }
        }
    }
}

/* pressure step: p^{n+1} from v^{n+1/2} */
static void SFX(fwo_pressure)(SFX(fwo_state)* st) {
    const int nd = st->P.ndim;
    const int64_t nx = st->P.nx, ny = st->P.ny, nz = nd == 3 ? st->P.nz : 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nx; ++i) {
        for (int64_t j = 0; j < ny; ++j) {
            for (int64_t k = 0; k < nz; ++k) {
                if (!SFX(fwo_active)(st, i, j, k)) continue;
                const int64_t c = (i * ny + j) * nz + k;
                const REAL* w = st->coef + 9 * (int64_t)st->P.dcmap[c];
                REAL div = 0;
                for (int a = 0; a < nd; ++a) {
                    const REAL* va = st->v[a];
                    const int64_t sa = st->s[a];
                    REAL acc = 0;
                    for (int m = 1; m <= 8; ++m)
                        acc += w[m - 1] * (va[c + (m - 1) * sa] - va[c - m * sa]);
                    REAL tr = 0;
                    for (int b = 0; b < nd; ++b) {
                        if (b == a) continue;
                        const int64_t sb = st->s[b];
                        tr += (va[c + sb] - va[c - sa + sb]) + (va[c - sb] - va[c - sa - sb]);
                    }
                    const REAL D = (acc + w[8] * tr) / st->h[a];
                    if (st->aexp) {
                        div += D;
                        continue;
                    }
                    REAL mem = 0;
                    for (int nu = 0; nu < 2; ++nu) {
                        REAL q = st->bp[a][nu][c] * st->phi[a][nu][c] + st->ap[a][nu][c] * D;
                        st->phi[a][nu][c] = q;
                        mem += q;
                    }
                    div += D / st->kp[a][c] + mem;
                }
                if (st->aexp)
                    st->p[c] = st->aexp[c] * (st->p[c] - st->dt * st->K[c] * div);
                else
                    st->p[c] -= st->dt * st->K[c] * div;
            }
        }
    }
}

/* Advance nsteps; rec (may be NULL) is [ceil(total / modt)][nout] in double, indexed by the
 * global step count. Returns 0. */
int SFX(fwo_run)(void* handle, int64_t nsteps, double* rec) {
    SFX(fwo_state)* st = (SFX(fwo_state)*)handle;
    const fwo_problem* P = &st->P;
    for (int64_t it = 0; it < nsteps; ++it) {
        const int64_t n = st->step;
        SFX(fwo_velocity)(st);
        SFX(fwo_pressure)(st);
        if (n < P->ntic)
            for (int64_t s = 0; s < P->ncoords; ++s) st->p[P->icc[s]] = (REAL)P->icmat[s * P->ntic + n];
        for (int64_t q = 0; q < P->nzero; ++q) st->p[P->zero[q]] = 0;
        if (rec && P->modt > 0 && n % P->modt == 0) {
            double* row = rec + (n / P->modt) * P->nout;
            for (int64_t j = 0; j < P->nout; ++j) row[j] = (double)st->p[P->outc[j]];
        }
        st->step = n + 1;
    }
    return 0;
}

/* Copy a wave field out as double: 0 = p, 1..3 = v_x, v_y, v_z. */
int SFX(fwo_get_field)(void* handle, int which, double* out) {
    SFX(fwo_state)* st = (SFX(fwo_state)*)handle;
    const REAL* f = which == 0 ? st->p : (which - 1 < st->P.ndim ? st->v[which - 1] : 0);
    if (!f) return 1;
    for (int64_t i = 0; i < st->n; ++i) out[i] = (double)f[i];
    return 0;
}

/* Overwrite a wave field from double (for tests that start from an initial condition). */
int SFX(fwo_set_field)(void* handle, int which, const double* in) {
    SFX(fwo_state)* st = (SFX(fwo_state)*)handle;
    REAL* f = which == 0 ? st->p : (which - 1 < st->P.ndim ? st->v[which - 1] : 0);
    if (!f) return 1;
    for (int64_t i = 0; i < st->n; ++i) f[i] = (REAL)in[i];
    return 0;
}

/* ---- adjoint (reverse mode) ------------------------------------------------------------ *
 * Vector-Jacobian product of the recorded traces with respect to the K and rho maps: runs
 * nsteps from a zero state keeping a checkpoint every ceil(sqrt(nsteps)) steps, then sweeps back
 * one segment at a time: the segment's states are recomputed from its checkpoint, then the
 * transposed steps run from its last step to its first. grad_rec is [ceil(nsteps / modt)][nout];
 * gK and grho (length n) receive sum_{n,j} grad_rec[n][j] * d rec[n][j] / d K[c] (and / d rho[c]),
 * with the coefficient sets (dcmap) held fixed. The transposed stencils are applied in scatter
 * form (loops over the cells that read a value), the GPU kernels use the gather form. The
 * per-cell loops run in parallel over x rows; a scatter from row i writes rows i - 8 .. i + 8,
 * so the scatter loops take the rows in blocks of 17 and run the even blocks in parallel, then
 * the odd ones (no two threads write one value). The state is left at the end of the run. */

#define FWO_ROWBLK 17

static inline int64_t SFX(fwo_nstate)(const SFX(fwo_state)* st) {
    return st->n * (1 + st->P.ndim + (st->aexp ? 0 : 4 * st->P.ndim));
}

static void SFX(fwo_pack)(const SFX(fwo_state)* st, REAL* dst) {
    const int nd = st->P.ndim;
    const int64_t n = st->n;
    memcpy(dst, st->p, sizeof(REAL) * n);
    dst += n;
    for (int a = 0; a < nd; ++a, dst += n) memcpy(dst, st->v[a], sizeof(REAL) * n);
    if (st->aexp) return;
    for (int a = 0; a < nd; ++a)
        for (int nu = 0; nu < 2; ++nu, dst += n) memcpy(dst, st->psi[a][nu], sizeof(REAL) * n);
    for (int a = 0; a < nd; ++a)
        for (int nu = 0; nu < 2; ++nu, dst += n) memcpy(dst, st->phi[a][nu], sizeof(REAL) * n);
}

static void SFX(fwo_unpack)(SFX(fwo_state)* st, const REAL* src) {
    const int nd = st->P.ndim;
    const int64_t n = st->n;
    memcpy(st->p, src, sizeof(REAL) * n);
    src += n;
    for (int a = 0; a < nd; ++a, src += n) memcpy(st->v[a], src, sizeof(REAL) * n);
    if (st->aexp) return;
    for (int a = 0; a < nd; ++a)
        for (int nu = 0; nu < 2; ++nu, src += n) memcpy(st->psi[a][nu], src, sizeof(REAL) * n);
    for (int a = 0; a < nd; ++a)
        for (int nu = 0; nu < 2; ++nu, src += n) memcpy(st->phi[a][nu], src, sizeof(REAL) * n);
}

/* views into a packed state */
static inline const REAL* SFX(fwo_sp)(const SFX(fwo_state)* st, const REAL* s) { (void)st; return s; }
static inline const REAL* SFX(fwo_sv)(const SFX(fwo_state)* st, const REAL* s, int a) { return s + st->n * (1 + a); }
static inline const REAL* SFX(fwo_spsi)(const SFX(fwo_state)* st, const REAL* s, int a, int nu) {
    return s + st->n * (1 + st->P.ndim + 2 * a + nu);
}
static inline const REAL* SFX(fwo_sphi)(const SFX(fwo_state)* st, const REAL* s, int a, int nu) {
    return s + st->n * (1 + 3 * st->P.ndim + 2 * a + nu);
}

/* D+_a p at active cell c (the velocity update's derivative) */
static REAL SFX(fwo_dplus)(const SFX(fwo_state)* st, const REAL* p, int64_t c, int a) {
    const REAL* w = st->coef + 9 * (int64_t)st->P.dcmap[c];
    const int64_t sa = st->s[a];
    REAL acc = 0;
    for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (p[c + m * sa] - p[c - (m - 1) * sa]);
    REAL tr = 0;
    for (int b = 0; b < st->P.ndim; ++b) {
        if (b == a) continue;
        const int64_t sb = st->s[b];
        tr += (p[c + sa + sb] - p[c + sb]) + (p[c + sa - sb] - p[c - sb]);
    }
    return (acc + w[8] * tr) / st->h[a];
}

/* D-_a f at active cell c (the pressure update's derivative of v_a) */
static REAL SFX(fwo_dminus)(const SFX(fwo_state)* st, const REAL* f, int64_t c, int a) {
    const REAL* w = st->coef + 9 * (int64_t)st->P.dcmap[c];
    const int64_t sa = st->s[a];
    REAL acc = 0;
    for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (f[c + (m - 1) * sa] - f[c - m * sa]);
    REAL tr = 0;
    for (int b = 0; b < st->P.ndim; ++b) {
        if (b == a) continue;
        const int64_t sb = st->s[b];
        tr += (f[c + sb] - f[c - sa + sb]) + (f[c - sb] - f[c - sa - sb]);
    }
    return (acc + w[8] * tr) / st->h[a];
}

/* out += (D+_a)^T g, scattered from active cell c with value g */
static void SFX(fwo_dplus_t)(const SFX(fwo_state)* st, REAL* out, int64_t c, int a, REAL g) {
    const REAL* w = st->coef + 9 * (int64_t)st->P.dcmap[c];
    const int64_t sa = st->s[a];
    const REAL gh = g / st->h[a];
    for (int m = 1; m <= 8; ++m) {
        out[c + m * sa] += w[m - 1] * gh;
        out[c - (m - 1) * sa] -= w[m - 1] * gh;
    }
    for (int b = 0; b < st->P.ndim; ++b) {
        if (b == a) continue;
        const int64_t sb = st->s[b];
        out[c + sa + sb] += w[8] * gh;
        out[c + sb] -= w[8] * gh;
        out[c + sa - sb] += w[8] * gh;
        out[c - sb] -= w[8] * gh;
    }
}

/* out += (D-_a)^T h, scattered from active cell c */
static void SFX(fwo_dminus_t)(const SFX(fwo_state)* st, REAL* out, int64_t c, int a, REAL hv) {
    const REAL* w = st->coef + 9 * (int64_t)st->P.dcmap[c];
    const int64_t sa = st->s[a];
    const REAL hh = hv / st->h[a];
    for (int m = 1; m <= 8; ++m) {
        out[c + (m - 1) * sa] += w[m - 1] * hh;
        out[c - m * sa] -= w[m - 1] * hh;
    }
    for (int b = 0; b < st->P.ndim; ++b) {
        if (b == a) continue;
        const int64_t sb = st->s[b];
        out[c + sb] += w[8] * hh;
        out[c - sa + sb] -= w[8] * hh;
        out[c - sb] += w[8] * hh;
        out[c - sa - sb] -= w[8] * hh;
    }
}

/* pressure update, transposed, at every active cell of x row i: H, the phi adjoints, gK */
static void SFX(fwo_adj_pressure_row)(SFX(fwo_state)* st, int64_t i, const REAL* cur, REAL* lp, REAL* lphi[3][2],
                                      REAL* H, double* gK) {
    const fwo_problem* P = &st->P;
    const int nd = P->ndim;
    const int64_t n = st->n, ny = P->ny, nz = nd == 3 ? P->nz : 1;
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t k = 0; k < nz; ++k) {
            if (!SFX(fwo_active)(st, i, j, k)) continue;
            const int64_t c = (i * ny + j) * nz + k;
            const REAL mu = lp[c];
            REAL div = 0;
            if (!st->aexp) {
                const REAL t = -st->dt * st->K[c] * mu;
                for (int a = 0; a < nd; ++a) {
                    const REAL hfw = SFX(fwo_dminus)(st, SFX(fwo_sv)(st, cur, a), c, a);
                    REAL lh = t / st->kp[a][c];
                    for (int nu = 0; nu < 2; ++nu) {
                        const REAL tot = lphi[a][nu][c] + t;
                        lh += st->ap[a][nu][c] * tot;
                        lphi[a][nu][c] = st->bp[a][nu][c] * tot;
                        div += SFX(fwo_sphi)(st, cur, a, nu)[c];
                    }
                    div += hfw / st->kp[a][c];
                    H[a * n + c] = lh;
                }
                gK[c] += (double)(mu * (-st->dt) * div);
            } else {
                const REAL A = st->aexp[c];
                for (int a = 0; a < nd; ++a) {
                    div += SFX(fwo_dminus)(st, SFX(fwo_sv)(st, cur, a), c, a);
                    H[a * n + c] = -A * st->dt * st->K[c] * mu;
                }
                lp[c] = A * mu;
                gK[c] += (double)(mu * (-A * st->dt) * div);
            }
        }
}

/* velocity update, transposed, at every active cell of x row i: G, the psi adjoints, grho
 * (which also receives at row i + 1) */
static void SFX(fwo_adj_velocity_row)(SFX(fwo_state)* st, int64_t i, const REAL* prev, const REAL* cur, REAL* lv[3],
                                      REAL* lpsi[3][2], REAL* G, double* grho) {
    const fwo_problem* P = &st->P;
    const int nd = P->ndim;
    const int64_t n = st->n, ny = P->ny, nz = nd == 3 ? P->nz : 1;
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t k = 0; k < nz; ++k) {
            if (!SFX(fwo_active)(st, i, j, k)) continue;
            const int64_t c = (i * ny + j) * nz + k;
            for (int a = 0; a < nd; ++a) {
                const int64_t sa = st->s[a];
                const REAL nu_v = lv[a][c];
                const REAL rb = (REAL)0.5 * (st->rho[c] + st->rho[c + sa]);
                const REAL beta = st->dt / rb;
                const REAL g = SFX(fwo_dplus)(st, SFX(fwo_sp)(st, prev), c, a);
                REAL u, lg;
                if (!st->aexp) {
                    u = g / st->kv[a][c];
                    lg = -beta * nu_v / st->kv[a][c];
                    for (int q = 0; q < 2; ++q) {
                        const REAL tot = lpsi[a][q][c] - beta * nu_v;
                        lg += st->av[a][q][c] * tot;
                        lpsi[a][q][c] = st->bv[a][q][c] * tot;
                        u += SFX(fwo_spsi)(st, cur, a, q)[c];
                    }
                } else {
                    const REAL A = st->aexp[c];
                    u = A * g;
                    lg = -A * beta * nu_v;
                    lv[a][c] = A * nu_v;
                }
                G[a * n + c] = lg;
                /* v -= beta u with beta = dt / rho_bar, rho_bar = (rho_c + rho_c+e_a) / 2 */
                const double drb = (double)(nu_v * (-u)) * (double)(-st->dt / (rb * rb));
                grho[c] += 0.5 * drb;
                grho[c + sa] += 0.5 * drb;
            }
        }
}

/* the transposed derivatives of x row i scattered: kind 0 = (D-_a)^T H into lv, 1 = (D+_a)^T G
 * into lp */
static void SFX(fwo_adj_scatter_row)(SFX(fwo_state)* st, int64_t i, int kind, REAL* lp, REAL* lv[3], const REAL* H,
                                     const REAL* G) {
    const fwo_problem* P = &st->P;
    const int nd = P->ndim;
    const int64_t n = st->n, ny = P->ny, nz = nd == 3 ? P->nz : 1;
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t k = 0; k < nz; ++k) {
            if (!SFX(fwo_active)(st, i, j, k)) continue;
            const int64_t c = (i * ny + j) * nz + k;
            for (int a = 0; a < nd; ++a) {
                if (kind == 0) SFX(fwo_dminus_t)(st, lv[a], c, a, H[a * n + c]);
                else SFX(fwo_dplus_t)(st, lp, c, a, G[a * n + c]);
            }
        }
}

int SFX(fwo_adjoint)(void* handle, int64_t nsteps, const double* grad_rec, double* gK, double* grho) {
    SFX(fwo_state)* st = (SFX(fwo_state)*)handle;
    const fwo_problem* P = &st->P;
    const int nd = P->ndim;
    const int64_t n = st->n, ns = SFX(fwo_nstate)(st);
    const int relax = st->aexp == 0;
    const int64_t nx = P->nx;
    const int64_t nblk = (nx + FWO_ROWBLK - 1) / FWO_ROWBLK;
    int64_t every = 1;
    while (every * every < nsteps) ++every;
    const int64_t nck = nsteps > 0 ? (nsteps + every - 1) / every : 0;

    /* forward from zero, a checkpoint before every `every`-th step, and the final state */
    memset(st->p, 0, sizeof(REAL) * n);
    for (int a = 0; a < nd; ++a) {
        memset(st->v[a], 0, sizeof(REAL) * n);
        for (int nu = 0; nu < 2; ++nu) {
            memset(st->psi[a][nu], 0, sizeof(REAL) * n);
            memset(st->phi[a][nu], 0, sizeof(REAL) * n);
        }
    }
    st->step = 0;
    REAL* ck = (REAL*)malloc(sizeof(REAL) * (size_t)(ns * (nck + 1)));
    REAL* seg = (REAL*)malloc(sizeof(REAL) * (size_t)(ns * (every + 1)));
    REAL* lam = (REAL*)calloc((size_t)(ns + 2 * nd * n), sizeof(REAL)); /* adjoint state + H, G */
    if (!ck || !seg || !lam) {
        free(ck);
        free(seg);
        free(lam);
        return 1;
    }
    REAL* fin = ck + ns * nck;
    for (int64_t it = 0; it < nsteps; ++it) {
        if (it % every == 0) SFX(fwo_pack)(st, ck + ns * (it / every));
        SFX(fwo_run)(handle, 1, 0);
    }
    SFX(fwo_pack)(st, fin);

    REAL* lp = lam;
    REAL* lv[3] = {0, 0, 0};
    REAL* lpsi[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    REAL* lphi[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    REAL* H = lam + ns;      /* [nd][n] */
    REAL* G = H + nd * n;    /* [nd][n] */
    for (int a = 0; a < nd; ++a) {
        lv[a] = lam + n * (1 + a);
        for (int nu = 0; nu < 2; ++nu) {
            lpsi[a][nu] = relax ? lam + n * (1 + nd + 2 * a + nu) : 0;
            lphi[a][nu] = relax ? lam + n * (1 + 3 * nd + 2 * a + nu) : 0;
        }
    }
    for (int64_t c = 0; c < n; ++c) gK[c] = grho[c] = 0;

    for (int64_t sgi = nck - 1; sgi >= 0; --sgi) {
        const int64_t s0 = sgi * every, s1 = s0 + every < nsteps ? s0 + every : nsteps;
        SFX(fwo_unpack)(st, ck + ns * sgi);
        st->step = s0;
        SFX(fwo_pack)(st, seg);
        for (int64_t t = s0; t < s1; ++t) {
            SFX(fwo_run)(handle, 1, 0);
            SFX(fwo_pack)(st, seg + ns * (t - s0 + 1));
        }
        for (int64_t it = s1 - 1; it >= s0; --it) {
            const REAL* prev = seg + ns * (it - s0);     /* state before step it */
            const REAL* cur = seg + ns * (it - s0 + 1);  /* state after step it */
            /* records of step it, then the sources and pressure-release points (overwritten) */
            if (P->modt > 0 && it % P->modt == 0) {
                const double* row = grad_rec + (it / P->modt) * P->nout;
                for (int64_t j = 0; j < P->nout; ++j) lp[P->outc[j]] += (REAL)row[j];
            }
            if (it < P->ntic)
                for (int64_t q = 0; q < P->ncoords; ++q) lp[P->icc[q]] = 0;
            for (int64_t q = 0; q < P->nzero; ++q) lp[P->zero[q]] = 0;

            /* pressure update, transposed */
            memset(H, 0, sizeof(REAL) * (size_t)(nd * n));
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < nx; ++i) SFX(fwo_adj_pressure_row)(st, i, cur, lp, lphi, H, gK);
            for (int color = 0; color < 2; ++color) {
#pragma omp parallel for schedule(static)
                for (int64_t b = color; b < nblk; b += 2)
                    for (int64_t i = b * FWO_ROWBLK; i < nx && i < (b + 1) * FWO_ROWBLK; ++i)
                        SFX(fwo_adj_scatter_row)(st, i, 0, lp, lv, H, G);
            }

            /* velocity update, transposed */
            memset(G, 0, sizeof(REAL) * (size_t)(nd * n));
            for (int color = 0; color < 2; ++color) {
#pragma omp parallel for schedule(static)
                for (int64_t b = color; b < nblk; b += 2)
                    for (int64_t i = b * FWO_ROWBLK; i < nx && i < (b + 1) * FWO_ROWBLK; ++i)
                        SFX(fwo_adj_velocity_row)(st, i, prev, cur, lv, lpsi, G, grho);
            }
            for (int color = 0; color < 2; ++color) {
#pragma omp parallel for schedule(static)
                for (int64_t b = color; b < nblk; b += 2)
                    for (int64_t i = b * FWO_ROWBLK; i < nx && i < (b + 1) * FWO_ROWBLK; ++i)
                        SFX(fwo_adj_scatter_row)(st, i, 1, lp, lv, H, G);
            }
        }
    }
    SFX(fwo_unpack)(st, fin);
    st->step = nsteps;
    free(lam);
    free(seg);
    free(ck);
    return 0;
}
