// Adjoint (reverse-mode) kernels: one time step of the update transposed.
//
// The forward step is, per active cell c and axis a (h_a folded into the coefficients):
//   velocity:  g_a = D+_a p,  psi' = b psi + a g_a,  v_a' = v_a - beta_a (g_a / kappa + sum psi')
//   pressure:  h_a = D-_a v_a',  phi' = b phi + a h_a,  p' = p - dt K sum_a (h_a / kappa + sum phi')
//   then p'[source] = signal, p'[air] = 0, and the sensors record p'
// (the exponential model: v_a' = A (v_a - beta_a g_a), p' = A (p - dt K sum_a h_a)), with
// beta_a = dt / rho_bar_a. Going back through step n with the adjoint state lambda of the state
// after the step:
//   1. lambda_p[sensor] += dL/d record   (k_adj_records)
//   2. lambda_p[source] = lambda_p[air] = 0   (k_adj_mask; those values were overwritten)
//   3. pressure transposed (k_adj_pressure): t = -dt K mu, mu = lambda_p;
//        lambda_phi_tot = lambda_phi + t, H_a = t / kappa + sum a lambda_phi_tot,
//        lambda_phi = b lambda_phi_tot, dL/dK += mu * (-dt) * div (div from the stored state)
//   4. lambda_v_a += (D-_a)^T H_a   (k_adj_dminus_t, gather form)
//   5. velocity transposed (k_adj_velocity): nu = lambda_v_a;
//        lambda_psi_tot = lambda_psi - beta nu, G_a = -beta nu / kappa + sum a lambda_psi_tot,
//        lambda_psi = b lambda_psi_tot, dL/d rho_bar_a += nu * u_a * dt / rho_bar^2 with
//        u_a = g_a / kappa + sum psi'
//   6. lambda_p += sum_a (D+_a)^T G_a   (k_adj_dplus_t, gather form)
// The coefficient sets (dcmap) are constants. The transposed stencils gather: the value at x
// collects every active cell whose stencil reads x, with that cell's coefficients.
#include <cstdlib>
#include <cstring>

#include "fw_internal.h"

namespace fw {

namespace {

template <int ND>
__device__ inline bool adj_cell(const DevSim& s, int64_t& i, int64_t& j, int64_t& k) {
    if (ND == 2) {
        k = 0;
        j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        i = (int64_t)blockIdx.y * blockDim.y + threadIdx.y;
        return i >= kHalo && i < s.nx - kHalo && j >= kHalo && j < s.ny - kHalo;
    }
    k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    j = (int64_t)blockIdx.y * blockDim.y + threadIdx.y;
    i = blockIdx.z;
    return i >= kHalo && i < s.nx - kHalo && j >= kHalo && j < s.ny - kHalo && k >= kHalo && k < s.nz - kHalo;
}

// The coefficient set of cell c on axis a. UNI: one sound speed (ndmap == 1), the same set
// everywhere, so the gathers skip the per-neighbour dcmap load.
template <bool UNI>
__device__ inline const float* wts(const DevSim& s, int64_t c, int nd, int a) {
    if (UNI) return s.coef + a * kNCoef;
    return s.coef + ((int64_t)s.dcmap[c] * nd + a) * kNCoef;
}

template <int ND, bool UNI>
__device__ inline float dplus(const DevSim& s, const float* p, int64_t c, int a, const int64_t* st) {
    const float* w = wts<UNI>(s, c, ND, a);
    const int64_t sa = st[a];
    float acc = 0.f;
#pragma unroll
    for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (p[c + m * sa] - p[c - (m - 1) * sa]);
    float tr = 0.f;
#pragma unroll
    for (int b = 0; b < ND; ++b) {
        if (b == a) continue;
        const int64_t sb = st[b];
        tr += (p[c + sa + sb] - p[c + sb]) + (p[c + sa - sb] - p[c - sb]);
    }
    return acc + w[8] * tr;
}

template <int ND, bool UNI>
__device__ inline float dminus(const DevSim& s, const float* f, int64_t c, int a, const int64_t* st) {
    const float* w = wts<UNI>(s, c, ND, a);
    const int64_t sa = st[a];
    float acc = 0.f;
#pragma unroll
    for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (f[c + (m - 1) * sa] - f[c - m * sa]);
    float tr = 0.f;
#pragma unroll
    for (int b = 0; b < ND; ++b) {
        if (b == a) continue;
        const int64_t sb = st[b];
        tr += (f[c + sb] - f[c - sa + sb]) + (f[c - sb] - f[c - sa - sb]);
    }
    return acc + w[8] * tr;
}

// records of one step: one thread per distinct sensor voxel sums its sensors' weights
__global__ __launch_bounds__(256) void k_adj_records(float* lp, const float* grow, const int64_t* vox,
                                                     const int32_t* first, const int32_t* order, int64_t nvox) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvox) return;
    float acc = 0.f;
    for (int32_t q = first[t]; q < first[t + 1]; ++q) acc += grow[order[q]];
    lp[vox[t]] += acc;
}

__global__ __launch_bounds__(256) void k_adj_mask(float* lp, const int64_t* src_lin, int64_t nsrc,
                                                  const int64_t* zero_lin, int64_t nzero) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nsrc && src_lin[t] >= 0) lp[src_lin[t]] = 0.f;
    if (t < nzero) lp[zero_lin[t]] = 0.f;
}

// CODED (isotropic relaxation with codes): the relaxation coefficients of the cell come from its
// code's table row (2 bytes a cell) instead of the five per-cell maps
template <int ND, bool EXPO, bool UNI, bool CODED>
__global__ __launch_bounds__(256) void k_adj_pressure(DevSim s, AdjView w) {
    int64_t i, j, k;
    if (!adj_cell<ND>(s, i, j, k)) return;
    const int64_t c = i * s.sx + j * s.sy + k;
    const int64_t st[3] = {s.sx, s.sy, 1};
    const float mu = w.lp[c];
    float div = 0.f;
    if (EXPO) {
        const float A = s.aexp[c];
        const float t = -A * s.dt * s.K[c] * mu;
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            div += dminus<ND, UNI>(s, w.fv[a], c, a, st);
            w.H[a][c] = t;
        }
        w.lp[c] = A * mu;
        w.gK[c] += mu * (-A * s.dt) * div;
        return;
    }
    const float t = -s.dt * s.K[c] * mu;
    CodedCoef kc{};
    if (CODED) kc = cell_coef<false, UNI>(s, (uint32_t)c);
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const float h = dminus<ND, UNI>(s, w.fv[a], c, a, st);
        const float ik = CODED ? kc.ik : s.ikp[a][c];
        float lh = t * ik;
#pragma unroll
        for (int nu = 0; nu < kNRelax; ++nu) {
            const float tot = w.lphi[a][nu][c] + t;
            const float an = CODED ? (nu ? kc.a2 : kc.a1) : s.ap[a][nu][c];
            const float bn = CODED ? (nu ? kc.b2 : kc.b1) : s.bp[a][nu][c];
            lh += an * tot;
            w.lphi[a][nu][c] = bn * tot;
        }
        div += h * ik + (w.fphi[a][0][c] + w.fphi[a][1][c]);
        w.H[a][c] = lh;
    }
    w.gK[c] += mu * (-s.dt) * div;
}

// lambda_v_a += (D-_a)^T H_a at active cells
template <int ND, bool UNI>
__global__ __launch_bounds__(256) void k_adj_dminus_t(DevSim s, AdjView w) {
    int64_t i, j, k;
    if (!adj_cell<ND>(s, i, j, k)) return;
    const int64_t x = i * s.sx + j * s.sy + k;
    const int64_t st[3] = {s.sx, s.sy, 1};
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const float* H = w.H[a];
        const int64_t sa = st[a];
        float acc = 0.f;
#pragma unroll
        for (int m = 1; m <= 8; ++m) {
            const int64_t c1 = x - (m - 1) * sa, c2 = x + m * sa;
            acc += wts<UNI>(s, c1, ND, a)[m - 1] * H[c1] - wts<UNI>(s, c2, ND, a)[m - 1] * H[c2];
        }
#pragma unroll
        for (int b = 0; b < ND; ++b) {
            if (b == a) continue;
            const int64_t sb = st[b];
#pragma unroll
            for (int sg = -1; sg <= 1; sg += 2) {
                const int64_t c1 = x - sg * sb, c2 = x + sa - sg * sb;
                acc += wts<UNI>(s, c1, ND, a)[8] * H[c1] - wts<UNI>(s, c2, ND, a)[8] * H[c2];
            }
        }
        w.lv[a][x] += acc;
    }
}

template <int ND, bool EXPO, bool UNI, bool CODED>
__global__ __launch_bounds__(256) void k_adj_velocity(DevSim s, AdjView w) {
    int64_t i, j, k;
    if (!adj_cell<ND>(s, i, j, k)) return;
    const int64_t c = i * s.sx + j * s.sy + k;
    const int64_t st[3] = {s.sx, s.sy, 1};
    const float rho_c = s.rho[c];
    CodedCoef kc{};
    if (CODED) kc = cell_coef<true, UNI>(s, (uint32_t)c);
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const int64_t sa = st[a];
        const float nu = w.lv[a][c];
        const float rb = 0.5f * (rho_c + s.rho[c + sa]);
        const float beta = s.dt / rb;
        const float g = dplus<ND, UNI>(s, w.fp_prev, c, a, st);
        float u, lg;
        if (EXPO) {
            const float A = s.aexp[c];
            u = A * g;
            lg = -A * beta * nu;
            w.lv[a][c] = A * nu;
        } else {
            const float ik = CODED ? kc.ik : s.ikv[a][c];
            u = g * ik + (w.fpsi[a][0][c] + w.fpsi[a][1][c]);
            lg = -beta * nu * ik;
#pragma unroll
            for (int q = 0; q < kNRelax; ++q) {
                const float tot = w.lpsi[a][q][c] - beta * nu;
                const float aq = CODED ? (q ? kc.a2 : kc.a1) : s.av[a][q][c];
                const float bq = CODED ? (q ? kc.b2 : kc.b1) : s.bv[a][q][c];
                lg += aq * tot;
                w.lpsi[a][q][c] = bq * tot;
            }
        }
        w.G[a][c] = lg;
        w.gRb[a][c] += nu * u * s.dt / (rb * rb);
    }
}

// lambda_p += sum_a (D+_a)^T G_a at active cells
template <int ND, bool UNI>
__global__ __launch_bounds__(256) void k_adj_dplus_t(DevSim s, AdjView w) {
    int64_t i, j, k;
    if (!adj_cell<ND>(s, i, j, k)) return;
    const int64_t x = i * s.sx + j * s.sy + k;
    const int64_t st[3] = {s.sx, s.sy, 1};
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const float* G = w.G[a];
        const int64_t sa = st[a];
#pragma unroll
        for (int m = 1; m <= 8; ++m) {
            const int64_t c1 = x - m * sa, c2 = x + (m - 1) * sa;
            acc += wts<UNI>(s, c1, ND, a)[m - 1] * G[c1] - wts<UNI>(s, c2, ND, a)[m - 1] * G[c2];
        }
#pragma unroll
        for (int b = 0; b < ND; ++b) {
            if (b == a) continue;
            const int64_t sb = st[b];
#pragma unroll
            for (int sg = -1; sg <= 1; sg += 2) {
                const int64_t c1 = x - sa - sg * sb, c2 = x - sg * sb;
                acc += wts<UNI>(s, c1, ND, a)[8] * G[c1] - wts<UNI>(s, c2, ND, a)[8] * G[c2];
            }
        }
    }
    w.lp[x] += acc;
}

// The two transposed gathers of a 2D step on tiles, with the gathered field (H or G), the
// coefficient-set index of every cell it comes from and the coefficient table staged in LDS (the
// per-cell kernels above read each neighbour's index and weights from global memory). A
// workgroup of 256 lanes owns a tile of ATX rows (x) by ATY columns (y); lane (r, j) updates
// rows r, r + 4, ... of column j. Axis x's field is staged over rows [i0 - 8, i0 + ATX + 8) and
// columns [j0 - 1, j0 + ATY + 1), axis y's over rows [i0 - 1, i0 + ATX + 1) and columns
// [j0 - 8, j0 + ATY + 8): every offset either gather reads (along its axis -8..+8, across it
// -1..+1). Cells outside the grid stage 0 (H and G are 0 off the updated cells anyway). The
// arithmetic and its order are the per-cell kernels', so the results are the same bit for bit.
constexpr int ATX = 16, ATY = 64;
constexpr int AXR = ATX + 16, AXC = ATY + 2;  // axis-x staged region
constexpr int AYR = ATX + 2, AYC = ATY + 16;  // axis-y staged region
constexpr int kAdjLdsSets = 256;              // coefficient sets staged (ndmap beyond: per-cell kernels)

// DMINUS: lambda_v_a += (D-_a)^T H_a (k_adj_dminus_t); else lambda_p += sum_a (D+_a)^T G_a
template <bool UNI, bool DMINUS>
__global__ __launch_bounds__(256) void k_adj_gather_tile2d(DevSim s, AdjView w) {
    __shared__ float FX[AXR * AXC];
    __shared__ float FY[AYR * AYC];
    __shared__ int16_t DX[UNI ? 1 : AXR * AXC];
    __shared__ int16_t DY[UNI ? 1 : AYR * AYC];
    __shared__ float CT[UNI ? 2 * kNCoef : kAdjLdsSets * 2 * kNCoef];  // [set][axis][9]
    const float* fx = DMINUS ? w.H[0] : w.G[0];
    const float* fy = DMINUS ? w.H[1] : w.G[1];
    const int64_t i0 = kHalo + (int64_t)blockIdx.y * ATX, j0 = kHalo + (int64_t)blockIdx.x * ATY;
    const int nsets = UNI ? 1 : s.ndmap;
    for (int t = threadIdx.x; t < nsets * 2 * kNCoef; t += 256) CT[t] = s.coef[t];
    for (int t = threadIdx.x; t < AXR * AXC; t += 256) {
        const int r = t / AXC, cc = t - r * AXC;
        const int64_t i = i0 - 8 + r, j = j0 - 1 + cc;
        const bool in = i >= 0 && i < s.nx && j >= 0 && j < s.ny;
        FX[t] = in ? fx[i * s.sx + j] : 0.f;
        if (!UNI) DX[t] = in ? (int16_t)s.dcmap[i * s.sx + j] : 0;
    }
    for (int t = threadIdx.x; t < AYR * AYC; t += 256) {
        const int r = t / AYC, cc = t - r * AYC;
        const int64_t i = i0 - 1 + r, j = j0 - 8 + cc;
        const bool in = i >= 0 && i < s.nx && j >= 0 && j < s.ny;
        FY[t] = in ? fy[i * s.sx + j] : 0.f;
        if (!UNI) DY[t] = in ? (int16_t)s.dcmap[i * s.sx + j] : 0;
    }
    __syncthreads();
    const int jl = threadIdx.x % ATY, rl = threadIdx.x / ATY;
    const int64_t j = j0 + jl;
    if (j >= s.ny - kHalo) return;
    // staged element at (di, dj) from cell (il, jl) of the tile, for axis x's and axis y's field
    auto ex = [&](int il, int di, int dj) { return (il + 8 + di) * AXC + (jl + 1 + dj); };
    auto ey = [&](int il, int di, int dj) { return (il + 1 + di) * AYC + (jl + 8 + dj); };
    auto wx = [&](int e, int k) { return CT[(UNI ? 0 : (int)DX[e] * 2) * kNCoef + k]; };             // axis 0 set
    auto wy = [&](int e, int k) { return CT[(UNI ? kNCoef : ((int)DY[e] * 2 + 1) * kNCoef) + k]; };  // axis 1 set
    for (int il = rl; il < ATX; il += 4) {
        const int64_t i = i0 + il;
        if (i >= s.nx - kHalo) break;
        const int64_t x = i * s.sx + j;
        if (DMINUS) {
            // axis x: c1 = x - (m-1) e_x, c2 = x + m e_x; transverse c1 = x -+ e_y, c2 = x + e_x -+ e_y
            float acc = 0.f;
#pragma unroll
            for (int m = 1; m <= 8; ++m) {
                const int e1 = ex(il, -(m - 1), 0), e2 = ex(il, m, 0);
                acc += wx(e1, m - 1) * FX[e1] - wx(e2, m - 1) * FX[e2];
            }
#pragma unroll
            for (int sg = -1; sg <= 1; sg += 2) {
                const int e1 = ex(il, 0, -sg), e2 = ex(il, 1, -sg);
                acc += wx(e1, 8) * FX[e1] - wx(e2, 8) * FX[e2];
            }
            w.lv[0][x] += acc;
            acc = 0.f;
#pragma unroll
            for (int m = 1; m <= 8; ++m) {
                const int e1 = ey(il, 0, -(m - 1)), e2 = ey(il, 0, m);
                acc += wy(e1, m - 1) * FY[e1] - wy(e2, m - 1) * FY[e2];
            }
#pragma unroll
            for (int sg = -1; sg <= 1; sg += 2) {
                const int e1 = ey(il, -sg, 0), e2 = ey(il, -sg, 1);
                acc += wy(e1, 8) * FY[e1] - wy(e2, 8) * FY[e2];
            }
            w.lv[1][x] += acc;
        } else {
            // axis x: c1 = x - m e_x, c2 = x + (m-1) e_x; transverse c1 = x - e_x -+ e_y, c2 = x -+ e_y
            float acc = 0.f;
#pragma unroll
            for (int m = 1; m <= 8; ++m) {
                const int e1 = ex(il, -m, 0), e2 = ex(il, m - 1, 0);
                acc += wx(e1, m - 1) * FX[e1] - wx(e2, m - 1) * FX[e2];
            }
#pragma unroll
            for (int sg = -1; sg <= 1; sg += 2) {
                const int e1 = ex(il, -1, -sg), e2 = ex(il, 0, -sg);
                acc += wx(e1, 8) * FX[e1] - wx(e2, 8) * FX[e2];
            }
#pragma unroll
            for (int m = 1; m <= 8; ++m) {
                const int e1 = ey(il, 0, -m), e2 = ey(il, 0, m - 1);
                acc += wy(e1, m - 1) * FY[e1] - wy(e2, m - 1) * FY[e2];
            }
#pragma unroll
            for (int sg = -1; sg <= 1; sg += 2) {
                const int e1 = ey(il, -sg, -1), e2 = ey(il, -sg, 0);
                acc += wy(e1, 8) * FY[e1] - wy(e2, 8) * FY[e2];
            }
            w.lp[x] += acc;
        }
    }
}

// d/d rho[x] = sum_a (g_rb_a[x] + g_rb_a[x - e_a]) / 2, over the whole (padded) grid
__global__ __launch_bounds__(256) void k_adj_rho(DevSim s, AdjView w, float* grho) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = s.nx * s.sx;
    if (x >= n) return;
    const int64_t i = x / s.sx, rem = x - (x / s.sx) * s.sx;
    const int64_t j = s.ndim == 3 ? rem / s.sy : rem, k = s.ndim == 3 ? rem - (rem / s.sy) * s.sy : 0;
    if (j >= s.ny || k >= s.nz) {  // row padding
        grho[x] = 0.f;
        return;
    }
    const int64_t coord[3] = {i, j, k};
    const int64_t st[3] = {s.sx, s.sy, 1};
    float acc = 0.f;
    for (int a = 0; a < s.ndim; ++a) {
        acc += w.gRb[a][x];
        if (coord[a] > 0) acc += w.gRb[a][x - st[a]];
    }
    grho[x] = 0.5f * acc;
}

}  // namespace

static dim3 adj_grid(const DevSim& s) {
    if (s.ndim == 2) return dim3((unsigned)((s.ny + 63) / 64), (unsigned)((s.nx + 3) / 4), 1);
    return dim3((unsigned)((s.nz + 63) / 64), (unsigned)((s.ny + 3) / 4), (unsigned)s.nx);
}

hipError_t launch_adj_records(float* lp, const float* grow, const int64_t* vox, const int32_t* first,
                              const int32_t* order, int64_t nvox, hipStream_t st) {
    if (nvox == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_records, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, st, lp, grow, vox, first,
                       order, nvox);
    return hipGetLastError();
}

hipError_t launch_adj_mask(float* lp, const int64_t* src_lin, int64_t nsrc, const int64_t* zero_lin, int64_t nzero,
                           hipStream_t st) {
    const int64_t n = nsrc > nzero ? nsrc : nzero;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_mask, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, lp, src_lin, nsrc, zero_lin,
                       nzero);
    return hipGetLastError();
}

// the transposed gathers on LDS tiles: 2D grids of at most kAdjLdsSets coefficient sets (the
// staged set indices are 16-bit)
bool adj_tiles(const DevSim& s) { return s.ndim == 2 && s.ndmap <= kAdjLdsSets; }

// FULLWAVE_HIP_ADJ_GATHERS=percell keeps the per-cell gathers (to compare the two forms)
bool adj_percell_gathers() {
    const char* e = getenv("FULLWAVE_HIP_ADJ_GATHERS");
    return e && !strncmp(e, "percell", 7);
}

template <bool DMINUS>
void launch_gather_tile2d(const DevSim& s, const AdjView& w, hipStream_t st) {
    const dim3 g((unsigned)((s.ny - 2 * kHalo + ATY - 1) / ATY), (unsigned)((s.nx - 2 * kHalo + ATX - 1) / ATX));
    if (s.ndmap == 1) hipLaunchKernelGGL((k_adj_gather_tile2d<true, DMINUS>), g, dim3(256), 0, st, s, w);
    else hipLaunchKernelGGL((k_adj_gather_tile2d<false, DMINUS>), g, dim3(256), 0, st, s, w);
}

template <int ND, bool UNI>
void launch_adj_step_nd(const DevSim& s, const AdjView& w, dim3 g, dim3 b, hipStream_t st) {
    const bool coded = !s.aexp && s.isotropic && s.rcode;
    const bool tiles = adj_tiles(s) && !adj_percell_gathers();
    if (s.aexp) hipLaunchKernelGGL((k_adj_pressure<ND, true, UNI, false>), g, b, 0, st, s, w);
    else if (coded) hipLaunchKernelGGL((k_adj_pressure<ND, false, UNI, true>), g, b, 0, st, s, w);
    else hipLaunchKernelGGL((k_adj_pressure<ND, false, UNI, false>), g, b, 0, st, s, w);
    if (tiles) launch_gather_tile2d<true>(s, w, st);
    else hipLaunchKernelGGL((k_adj_dminus_t<ND, UNI>), g, b, 0, st, s, w);
    if (s.aexp) hipLaunchKernelGGL((k_adj_velocity<ND, true, UNI, false>), g, b, 0, st, s, w);
    else if (coded) hipLaunchKernelGGL((k_adj_velocity<ND, false, UNI, true>), g, b, 0, st, s, w);
    else hipLaunchKernelGGL((k_adj_velocity<ND, false, UNI, false>), g, b, 0, st, s, w);
    if (tiles) launch_gather_tile2d<false>(s, w, st);
    else hipLaunchKernelGGL((k_adj_dplus_t<ND, UNI>), g, b, 0, st, s, w);
}

hipError_t launch_adj_step(const DevSim& s, const AdjView& w, hipStream_t st) {
    const dim3 g = adj_grid(s), b(64, 4);
    const bool u = s.ndmap == 1;
    if (s.ndim == 2) {
        if (u) launch_adj_step_nd<2, true>(s, w, g, b, st);
        else launch_adj_step_nd<2, false>(s, w, g, b, st);
    } else {
        if (u) launch_adj_step_nd<3, true>(s, w, g, b, st);
        else launch_adj_step_nd<3, false>(s, w, g, b, st);
    }
    return hipGetLastError();
}

hipError_t launch_adj_rho(const DevSim& s, const AdjView& w, float* grho, hipStream_t st) {
    const int64_t n = s.nx * s.sx;
    hipLaunchKernelGGL(k_adj_rho, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s, w, grho);
    return hipGetLastError();
}

}  // namespace fw
