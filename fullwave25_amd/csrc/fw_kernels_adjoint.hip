// Adjoint (reverse-mode) kernels: one time step of the update transposed.
//
// The forward step is, per active cell c and axis a (h_a folded into the coefficients):
//   velocity:  g_a = D+_a p,  psi' = b psi + a g_a,  u_a = g_a / kappa + sum psi',
//              v_a' = v_a - beta_a u_a
//   pressure:  h_a = D-_a v_a',  phi' = b phi + a h_a,  div = sum_a (h_a / kappa + sum phi'),
//              p' = p - dt K div
//   then p'[source] = signal, p'[air] = 0, and the sensors record p'
// (the exponential model: u_a = A g_a, v_a' = A v_a - beta_a u_a; div = sum_a h_a,
// p' = A (p - dt K div)), with beta_a = dt / rho_bar_a. The forward kernels write u_a and div
// of every step to a tape (DevSim::tape_u, tape_div): that is all the sweep needs of the forward
// state. Going back through step n with the adjoint state lambda of the state after the step:
//   1. lambda_p[sensor] += dL/d record, then lambda_p[source] = lambda_p[air] = 0 (k_adj_sparse;
//      those values were overwritten)
//   2. pressure transposed (k_adj_pressure): t = -dt K mu, mu = lambda_p;
//        lambda_phi_tot = lambda_phi + t, H_a = t / kappa + sum a lambda_phi_tot,
//        lambda_phi = b lambda_phi_tot, dL/dK += mu * (-dt) * div
//   3. lambda_v_a += (D-_a)^T H_a   (gather form)
//   4. velocity transposed (k_adj_velocity): nu = lambda_v_a;
//        lambda_psi_tot = lambda_psi - beta nu, G_a = -beta nu / kappa + sum a lambda_psi_tot,
//        lambda_psi = b lambda_psi_tot, dL/d rho_bar_a += nu * u_a * dt / rho_bar^2
//   5. lambda_p += sum_a (D+_a)^T G_a   (gather form)
// The coefficient sets (dcmap) are constants. The transposed stencils gather: the value at x
// collects every active cell whose stencil reads x, with that cell's coefficients. In 2D, steps
// 3 and 4 run as one tiled kernel (the gather yields nu of its own cell, which step 4 needs
// alone) and step 5 as another, which also runs steps 1 and 2 of the next (earlier) step for its
// cells when that step's tape is held (they need lambda_p of the cell alone); elsewhere each step
// is its own per-cell kernel.
#include <cstdlib>
#include <cstring>
#include <type_traits>

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

// Records, then the overwritten points, of one step (see AdjSparse). One thread per distinct
// sensor voxel sums its sensors' weights; a voxel that the same step zeroes keeps no record, so
// no two threads write one address.
__global__ __launch_bounds__(256) void k_adj_sparse(float* lp, const float* grow, AdjSparse sp, int injecting) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (grow && t < sp.nvox) {
        const int m = sp.vmask[t];
        if (!(m == 1 || (m == 2 && injecting))) {
            float acc = 0.f;
            for (int32_t q = sp.first[t]; q < sp.first[t + 1]; ++q) acc += grow[sp.order[q]];
            lp[sp.vox[t]] += acc;
        }
    }
    if (injecting && t < sp.nsrc && sp.src_lin[t] >= 0) lp[sp.src_lin[t]] = 0.f;
    if (t < sp.nzero) lp[sp.zero_lin[t]] = 0.f;
}

// The pointwise steps take their operands from a struct loaded ahead of the arithmetic: the tiled
// kernels load a row's operands before the gathers that precede their use, so that the loads are
// in flight during the gathers. The per-cell kernels use the same structs and the same arithmetic.

// the relaxation coefficients of the cell once its code is in (CODED; escaped cells read the maps)
template <bool VEL>
__device__ inline CodedCoef decode_cell(const DevSim& s, int64_t c, uint32_t code) {
    if (code != kCodeEscape) return coded_coef(VEL ? s.rtab_v : s.rtab_p, code);
    if (VEL) return mapped_coef<true>(s.ikv[0], s.av[0][0], s.av[0][1], s.bv[0][0], s.bv[0][1], s.dcmap, (uint32_t)c);
    return mapped_coef<true>(s.ikp[0], s.ap[0][0], s.ap[0][1], s.bp[0][0], s.bp[0][1], s.dcmap, (uint32_t)c);
}

// operands of the pressure transposed at one cell
template <int ND, bool EXPO, bool CODED>
struct PreOps {
    float lp, K, div, gk, A;
    float phi[ND][kNRelax];
    CodedCoef kc[ND];
    uint32_t code, kind;
};

template <int ND, bool EXPO, bool CODED>
__device__ inline void pre_load(const DevSim& s, const AdjView& w, int64_t c, PreOps<ND, EXPO, CODED>& o) {
    o.lp = w.lp[c];
    o.K = s.K[c];
    o.div = w.tdiv[c];
    o.gk = w.gK[c];
    if (EXPO) {
        o.A = s.aexp[c];
        return;
    }
#pragma unroll
    for (int a = 0; a < ND; ++a) {
#pragma unroll
        for (int nu = 0; nu < kNRelax; ++nu) o.phi[a][nu] = w.lphi[a][nu][c];
        if (!CODED) o.kc[a] = CodedCoef{s.ikp[a][c], s.ap[a][0][c], s.ap[a][1][c], s.bp[a][0][c], s.bp[a][1][c], 0};
    }
    if (CODED) o.code = gld16(s.rcode, (uint32_t)c);
}

template <int ND, bool EXPO, bool CODED>
__device__ inline void pre_decode(const DevSim& s, int64_t c, PreOps<ND, EXPO, CODED>& o) {
    if (!CODED || EXPO) return;
    const CodedCoef k = decode_cell<false>(s, c, o.code);
#pragma unroll
    for (int a = 0; a < ND; ++a) o.kc[a] = k;
}

// The pressure transposed at active cell c with mu = lambda_p[c] after the step's sparse part;
// leaves lambda_p[c] = mu (STORE_MU; the exponential model always stores A mu).
template <int ND, bool EXPO, bool CODED, bool STORE_MU>
__device__ inline void pre_store(const DevSim& s, const AdjView& w, int64_t c, float mu, const PreOps<ND, EXPO, CODED>& o) {
    if (EXPO) {
        const float t = -o.A * s.dt * o.K * mu;
#pragma unroll
        for (int a = 0; a < ND; ++a) w.H[a][c] = t;
        w.lp[c] = o.A * mu;
        w.gK[c] = o.gk + mu * (-o.A * s.dt) * o.div;
        return;
    }
    const float t = -s.dt * o.K * mu;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const CodedCoef& k = o.kc[a];
        float lh = t * k.ik;
#pragma unroll
        for (int nu = 0; nu < kNRelax; ++nu) {
            const float tot = o.phi[a][nu] + t;
            lh += (nu ? k.a2 : k.a1) * tot;
            w.lphi[a][nu][c] = (nu ? k.b2 : k.b1) * tot;
        }
        w.H[a][c] = lh;
    }
    if (STORE_MU) w.lp[c] = mu;
    w.gK[c] = o.gk + mu * (-s.dt) * o.div;
}

// CODED (isotropic relaxation with codes): the relaxation coefficients of the cell come from its
// code's table row (2 bytes a cell) instead of the five per-cell maps
template <int ND, bool EXPO, bool CODED>
__global__ __launch_bounds__(256) void k_adj_pressure(DevSim s, AdjView w) {
    int64_t i, j, k;
    if (!adj_cell<ND>(s, i, j, k)) return;
    const int64_t c = i * s.sx + j * s.sy + k;
    PreOps<ND, EXPO, CODED> o;
    pre_load(s, w, c, o);
    pre_decode(s, c, o);
    pre_store<ND, EXPO, CODED, false>(s, w, c, o.lp, o);
}

// lambda_p[c] of the next step after its sparse part (as k_adj_sparse), from lp = lambda_p[c]
// after this step's gathers and kd = the cell's kind
__device__ inline float next_mu(const AdjNext& nx, int64_t c, unsigned kd, float lp) {
    if (!kd) return lp;
    if ((kd & 1u) || ((kd & 2u) && nx.injecting)) return 0.f;
    if (!(kd & 4u) || !nx.grow) return lp;
    int64_t lo = 0, hi = nx.nvox - 1;  // c is one of the sorted voxels
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (nx.vox[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    float acc = 0.f;
    for (int32_t q = nx.first[lo]; q < nx.first[lo + 1]; ++q) acc += nx.grow[nx.order[q]];
    return lp + acc;
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

// operands of the velocity transposed at one cell (nu: lambda_v as loaded)
template <int ND, bool EXPO, bool CODED>
struct VelOps {
    float nu[ND], rc, rn[ND], u[ND], grb[ND], A;
    float psi[ND][kNRelax];
    CodedCoef kc[ND];
    uint32_t code;
};

template <int ND, bool EXPO, bool CODED>
__device__ inline void vel_load(const DevSim& s, const AdjView& w, int64_t c, VelOps<ND, EXPO, CODED>& o) {
    const int64_t st[3] = {s.sx, s.sy, 1};
    o.rc = s.rho[c];
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        o.nu[a] = w.lv[a][c];
        o.rn[a] = s.rho[c + st[a]];
        o.u[a] = w.tu[a][c];
        o.grb[a] = w.gRb[a][c];
    }
    if (EXPO) {
        o.A = s.aexp[c];
        return;
    }
#pragma unroll
    for (int a = 0; a < ND; ++a) {
#pragma unroll
        for (int q = 0; q < kNRelax; ++q) o.psi[a][q] = w.lpsi[a][q][c];
        if (!CODED) o.kc[a] = CodedCoef{s.ikv[a][c], s.av[a][0][c], s.av[a][1][c], s.bv[a][0][c], s.bv[a][1][c], 0};
    }
    if (CODED) o.code = gld16(s.rcode, (uint32_t)c);
}

template <int ND, bool EXPO, bool CODED>
__device__ inline void vel_decode(const DevSim& s, int64_t c, VelOps<ND, EXPO, CODED>& o) {
    if (!CODED || EXPO) return;
    const CodedCoef k = decode_cell<true>(s, c, o.code);
#pragma unroll
    for (int a = 0; a < ND; ++a) o.kc[a] = k;
}

// The velocity transposed at active cell c with o.nu[a] = nu (lambda_v_a after the gather);
// STORE_NU: also store lambda_v = nu (the exponential model always stores A nu).
template <int ND, bool EXPO, bool CODED, bool STORE_NU>
__device__ inline void vel_store(const DevSim& s, const AdjView& w, int64_t c, const VelOps<ND, EXPO, CODED>& o) {
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const float nu = o.nu[a];
        const float rb = 0.5f * (o.rc + o.rn[a]);
        const float beta = s.dt / rb;
        float lg;
        if (EXPO) {
            lg = -o.A * beta * nu;
            w.lv[a][c] = o.A * nu;
        } else {
            const CodedCoef& k = o.kc[a];
            lg = -beta * nu * k.ik;
#pragma unroll
            for (int q = 0; q < kNRelax; ++q) {
                const float tot = o.psi[a][q] - beta * nu;
                lg += (q ? k.a2 : k.a1) * tot;
                w.lpsi[a][q][c] = (q ? k.b2 : k.b1) * tot;
            }
            if (STORE_NU) w.lv[a][c] = nu;
        }
        w.G[a][c] = lg;
        w.gRb[a][c] = o.grb[a] + nu * o.u[a] * s.dt / (rb * rb);
    }
}

template <int ND, bool EXPO, bool CODED>
__global__ __launch_bounds__(256) void k_adj_velocity(DevSim s, AdjView w) {
    int64_t i, j, k;
    if (!adj_cell<ND>(s, i, j, k)) return;
    const int64_t c = i * s.sx + j * s.sy + k;
    VelOps<ND, EXPO, CODED> o;
    vel_load(s, w, c, o);
    vel_decode(s, c, o);
    vel_store<ND, EXPO, CODED, false>(s, w, c, o);
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

// The transposed gathers of a 2D step on tiles, with the gathered field (H or G), the
// coefficient-set index of every cell it comes from and the coefficient table staged in LDS (the
// per-cell kernels above read each neighbour's index and weights from global memory). A
// workgroup of 256 lanes owns a tile of ATX rows (x) by ATY columns (y); lane (r, j) updates
// rows r, r + 4, ... of column j. Axis x's field is staged over rows [i0 - 8, i0 + ATX + 8) and
// columns [j0 - 1, j0 + ATY + 1), axis y's over rows [i0 - 1, i0 + ATX + 1) and columns
// [j0 - 8, j0 + ATY + 8): every offset either gather reads (along its axis -8..+8, across it
// -1..+1). Cells outside the grid stage 0 (H and G are 0 off the updated cells anyway). Every
// lane issues all of its staging loads before it stores any of them, so one memory latency
// covers the staging. The coefficient table takes the handle's sets only (dynamic LDS). The
// arithmetic and its order are the per-cell kernels', so the results are the same bit for bit.
constexpr int ATX = 16, ATY = 64, kAdjLanes = 256;
constexpr int AXR = ATX + 16, AXC = ATY + 2;  // axis-x staged region
constexpr int AYR = ATX + 2, AYC = ATY + 16;  // axis-y staged region
constexpr int NSX = (AXR * AXC + kAdjLanes - 1) / kAdjLanes;  // staged elements per lane
constexpr int NSY = (AYR * AYC + kAdjLanes - 1) / kAdjLanes;
constexpr int kAdjLdsSets = 256;  // coefficient sets staged (ndmap beyond: per-cell kernels)
constexpr uint32_t kXcds = 8;     // XCDs of an MI355X (the tile order is a bijection for any value)
constexpr int NCT = (kAdjLdsSets * 2 * kNCoef + kAdjLanes - 1) / kAdjLanes;

template <bool UNI>
__device__ inline void stage_gather(const DevSim& s, const uint8_t* dc8, const float* fx, const float* fy, int64_t i0,
                                    int64_t j0, float* FX, float* FY, uint8_t* DX, uint8_t* DY, float* CT) {
    const int nct = (UNI ? 1 : s.ndmap) * 2 * kNCoef;
    float vx[NSX], vy[NSY], vc[NCT];
    int32_t dx[NSX], dy[NSY];
#pragma unroll
    for (int q = 0; q < NSX; ++q) {
        const int t = threadIdx.x + q * kAdjLanes;
        const int r = t / AXC, cc = t - r * AXC;
        const int64_t i = i0 - 8 + r, j = j0 - 1 + cc;
        const bool in = t < AXR * AXC && i >= 0 && i < s.nx && j >= 0 && j < s.ny;
        vx[q] = in ? fx[i * s.sx + j] : 0.f;
        if (!UNI) dx[q] = in ? dc8[i * s.sx + j] : 0;
    }
#pragma unroll
    for (int q = 0; q < NSY; ++q) {
        const int t = threadIdx.x + q * kAdjLanes;
        const int r = t / AYC, cc = t - r * AYC;
        const int64_t i = i0 - 1 + r, j = j0 - 8 + cc;
        const bool in = t < AYR * AYC && i >= 0 && i < s.nx && j >= 0 && j < s.ny;
        vy[q] = in ? fy[i * s.sx + j] : 0.f;
        if (!UNI) dy[q] = in ? dc8[i * s.sx + j] : 0;
    }
#pragma unroll
    for (int q = 0; q < NCT; ++q) {
        const int t = threadIdx.x + q * kAdjLanes;
        vc[q] = t < nct ? s.coef[t] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < NSX; ++q) {
        const int t = threadIdx.x + q * kAdjLanes;
        if (t < AXR * AXC) {
            FX[t] = vx[q];
            if (!UNI) DX[t] = (uint8_t)dx[q];
        }
    }
#pragma unroll
    for (int q = 0; q < NSY; ++q) {
        const int t = threadIdx.x + q * kAdjLanes;
        if (t < AYR * AYC) {
            FY[t] = vy[q];
            if (!UNI) DY[t] = (uint8_t)dy[q];
        }
    }
#pragma unroll
    for (int q = 0; q < NCT; ++q) {
        const int t = threadIdx.x + q * kAdjLanes;
        if (t < nct) CT[t] = vc[q];
    }
}

// DMINUS: lambda_v_a += (D-_a)^T H_a, then the velocity transposed of the cell (EXPO, CODED as
// k_adj_velocity); else lambda_p += sum_a (D+_a)^T G_a, then with NEXT the next step's sparse
// part and pressure transposed of the cell (EXPO, CODED as k_adj_pressure; w.tdiv is then the
// next step's tape). Each lane loads the operands of its first row before the staging and those
// of row m + 1 before the gathers of row m, so their memory latency overlaps other work.
template <bool UNI, bool DMINUS, bool EXPO, bool CODED, bool NEXT>
__global__ __launch_bounds__(kAdjLanes) void k_adj_gather_tile2d(DevSim s, AdjView w, AdjNext nx) {
    __shared__ float FX[AXR * AXC];
    __shared__ float FY[AYR * AYC];
    __shared__ uint8_t DX[UNI ? 1 : AXR * AXC];
    __shared__ uint8_t DY[UNI ? 1 : AYR * AYC];
    extern __shared__ float CT[];  // [set][axis][9], the handle's sets
    constexpr int NR = ATX / 4;    // rows per lane: rl, rl + 4, ...
    // Workgroup b runs on XCD b % 8 (round-robin dispatch). XCD k takes the tiles
    // [k q + min(k, r), ...) in row-major order, so that neighbouring tiles, which stage the same
    // halo cells, share an L2: the fetch from memory drops by 40 % (the halos are 3.5 staged
    // values a cell) and the kernels run 17 % faster than with the tiles in launch order.
    const uint32_t nb = gridDim.x * gridDim.y, b = blockIdx.x + blockIdx.y * gridDim.x;
    const uint32_t xcd = b % kXcds, k = b / kXcds, q = nb / kXcds, r = nb % kXcds;
    const uint32_t tile = xcd * q + (xcd < r ? xcd : r) + k;
    const uint32_t bx = tile % gridDim.x, by = tile / gridDim.x;
    const int64_t i0 = kHalo + (int64_t)by * ATX, j0 = kHalo + (int64_t)bx * ATY;
    const int jl = threadIdx.x % ATY, rl = threadIdx.x / ATY;
    const int64_t j = j0 + jl;
    // this lane's rows that hold updated cells
    const int64_t left = s.nx - kHalo - (i0 + rl);
    const int nrows = j >= s.ny - kHalo || left <= 0 ? 0 : (int)(left >= 4 * NR ? NR : (left + 3) / 4);
    using Ops = std::conditional_t<DMINUS, VelOps<2, EXPO, CODED>, PreOps<2, EXPO, CODED>>;
    Ops ops[2];
    auto load = [&](int m, Ops& o) {
        const int64_t x = (i0 + rl + 4 * m) * s.sx + j;
        if constexpr (DMINUS) {
            vel_load(s, w, x, o);
        } else if constexpr (NEXT) {
            pre_load(s, w, x, o);
            o.kind = nx.kind[x];
        } else {
            o.lp = w.lp[x];
        }
    };
    if (nrows > 0) load(0, ops[0]);
    stage_gather<UNI>(s, w.dc8, DMINUS ? w.H[0] : w.G[0], DMINUS ? w.H[1] : w.G[1], i0, j0, FX, FY, DX, DY, CT);
    __syncthreads();
    // staged element at (di, dj) from cell (il, jl) of the tile, for axis x's and axis y's field
    auto ex = [&](int il, int di, int dj) { return (il + 8 + di) * AXC + (jl + 1 + dj); };
    auto ey = [&](int il, int di, int dj) { return (il + 1 + di) * AYC + (jl + 8 + dj); };
    auto wx = [&](int e, int k) { return CT[(UNI ? 0 : (int)DX[e] * 2) * kNCoef + k]; };             // axis 0 set
    auto wy = [&](int e, int k) { return CT[(UNI ? kNCoef : ((int)DY[e] * 2 + 1) * kNCoef) + k]; };  // axis 1 set
#pragma unroll
    for (int m = 0; m < NR; ++m) {
        if (m >= nrows) break;
        Ops& o = ops[m & 1];
        const int il = rl + 4 * m;
        const int64_t x = (i0 + il) * s.sx + j;
        if constexpr (DMINUS) vel_decode(s, x, o);
        else if constexpr (NEXT) pre_decode(s, x, o);
        if (m + 1 < nrows) load(m + 1, ops[(m + 1) & 1]);
        if constexpr (DMINUS) {
            // axis x: c1 = x - (m-1) e_x, c2 = x + m e_x; transverse c1 = x -+ e_y, c2 = x + e_x -+ e_y
            float acc = 0.f;
#pragma unroll
            for (int q = 1; q <= 8; ++q) {
                const int e1 = ex(il, -(q - 1), 0), e2 = ex(il, q, 0);
                acc += wx(e1, q - 1) * FX[e1] - wx(e2, q - 1) * FX[e2];
            }
#pragma unroll
            for (int sg = -1; sg <= 1; sg += 2) {
                const int e1 = ex(il, 0, -sg), e2 = ex(il, 1, -sg);
                acc += wx(e1, 8) * FX[e1] - wx(e2, 8) * FX[e2];
            }
            o.nu[0] = o.nu[0] + acc;
            acc = 0.f;
#pragma unroll
            for (int q = 1; q <= 8; ++q) {
                const int e1 = ey(il, 0, -(q - 1)), e2 = ey(il, 0, q);
                acc += wy(e1, q - 1) * FY[e1] - wy(e2, q - 1) * FY[e2];
            }
#pragma unroll
            for (int sg = -1; sg <= 1; sg += 2) {
                const int e1 = ey(il, -sg, 0), e2 = ey(il, -sg, 1);
                acc += wy(e1, 8) * FY[e1] - wy(e2, 8) * FY[e2];
            }
            o.nu[1] = o.nu[1] + acc;
            vel_store<2, EXPO, CODED, true>(s, w, x, o);
        } else {
            // axis x: c1 = x - m e_x, c2 = x + (m-1) e_x; transverse c1 = x - e_x -+ e_y, c2 = x -+ e_y
            float acc = 0.f;
#pragma unroll
            for (int q = 1; q <= 8; ++q) {
                const int e1 = ex(il, -q, 0), e2 = ex(il, q - 1, 0);
                acc += wx(e1, q - 1) * FX[e1] - wx(e2, q - 1) * FX[e2];
            }
#pragma unroll
            for (int sg = -1; sg <= 1; sg += 2) {
                const int e1 = ex(il, -1, -sg), e2 = ex(il, 0, -sg);
                acc += wx(e1, 8) * FX[e1] - wx(e2, 8) * FX[e2];
            }
#pragma unroll
            for (int q = 1; q <= 8; ++q) {
                const int e1 = ey(il, 0, -q), e2 = ey(il, 0, q - 1);
                acc += wy(e1, q - 1) * FY[e1] - wy(e2, q - 1) * FY[e2];
            }
#pragma unroll
            for (int sg = -1; sg <= 1; sg += 2) {
                const int e1 = ey(il, -sg, -1), e2 = ey(il, -sg, 0);
                acc += wy(e1, 8) * FY[e1] - wy(e2, 8) * FY[e2];
            }
            if constexpr (NEXT) pre_store<2, EXPO, CODED, true>(s, w, x, next_mu(nx, x, o.kind, o.lp + acc), o);
            else w.lp[x] = o.lp + acc;
        }
    }
}

__global__ __launch_bounds__(256) void k_adj_set_index(const int32_t* dcmap, uint8_t* dc8, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dc8[t] = (uint8_t)dcmap[t];
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

hipError_t launch_adj_sparse(float* lp, const float* grow, const AdjSparse& sp, bool injecting, hipStream_t st) {
    int64_t n = sp.nzero;
    if (grow && sp.nvox > n) n = sp.nvox;
    if (injecting && sp.nsrc > n) n = sp.nsrc;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_sparse, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, lp, grow, sp,
                       injecting ? 1 : 0);
    return hipGetLastError();
}

// the transposed gathers on LDS tiles: 2D grids of at most kAdjLdsSets coefficient sets (the
// staged set indices are 16-bit)
bool adj_tiles(const DevSim& s) { return s.ndim == 2 && s.ndmap <= kAdjLdsSets; }

// FULLWAVE_HIP_ADJ_GATHERS=percell keeps the per-cell kernels (to compare the two forms)
bool adj_percell_gathers() {
    const char* e = getenv("FULLWAVE_HIP_ADJ_GATHERS");
    return e && !strncmp(e, "percell", 7);
}

bool adj_fuses_next(const DevSim& s) { return adj_tiles(s) && !adj_percell_gathers(); }

hipError_t launch_adj_set_index(const int32_t* dcmap, uint8_t* dc8, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_set_index, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dcmap, dc8, n);
    return hipGetLastError();
}

template <bool DMINUS, bool EXPO, bool CODED, bool NEXT>
void launch_gather_tile2d(const DevSim& s, const AdjView& w, const AdjNext& nx, hipStream_t st) {
    const dim3 g((unsigned)((s.ny - 2 * kHalo + ATY - 1) / ATY), (unsigned)((s.nx - 2 * kHalo + ATX - 1) / ATX));
    const bool uni = s.ndmap == 1;
    const size_t lds = (size_t)(uni ? 1 : s.ndmap) * 2 * kNCoef * sizeof(float);
    if (uni)
        hipLaunchKernelGGL((k_adj_gather_tile2d<true, DMINUS, EXPO, CODED, NEXT>), g, dim3(kAdjLanes), lds, st, s, w, nx);
    else
        hipLaunchKernelGGL((k_adj_gather_tile2d<false, DMINUS, EXPO, CODED, NEXT>), g, dim3(kAdjLanes), lds, st, s, w, nx);
}

static bool adj_coded(const DevSim& s) { return !s.aexp && s.isotropic && s.rcode; }

template <int ND>
void launch_pressure_nd(const DevSim& s, const AdjView& w, dim3 g, dim3 b, hipStream_t st) {
    if (s.aexp) hipLaunchKernelGGL((k_adj_pressure<ND, true, false>), g, b, 0, st, s, w);
    else if (adj_coded(s)) hipLaunchKernelGGL((k_adj_pressure<ND, false, true>), g, b, 0, st, s, w);
    else hipLaunchKernelGGL((k_adj_pressure<ND, false, false>), g, b, 0, st, s, w);
}

template <int ND, bool UNI>
void launch_rest_nd(const DevSim& s, const AdjView& w, dim3 g, dim3 b, hipStream_t st) {
    hipLaunchKernelGGL((k_adj_dminus_t<ND, UNI>), g, b, 0, st, s, w);
    if (s.aexp) hipLaunchKernelGGL((k_adj_velocity<ND, true, false>), g, b, 0, st, s, w);
    else if (adj_coded(s)) hipLaunchKernelGGL((k_adj_velocity<ND, false, true>), g, b, 0, st, s, w);
    else hipLaunchKernelGGL((k_adj_velocity<ND, false, false>), g, b, 0, st, s, w);
    hipLaunchKernelGGL((k_adj_dplus_t<ND, UNI>), g, b, 0, st, s, w);
}

hipError_t launch_adj_pressure(const DevSim& s, const AdjView& w, hipStream_t st) {
    const dim3 g = adj_grid(s), b(64, 4);
    if (s.ndim == 2) launch_pressure_nd<2>(s, w, g, b, st);
    else launch_pressure_nd<3>(s, w, g, b, st);
    return hipGetLastError();
}

hipError_t launch_adj_rest(const DevSim& s, const AdjView& w, const AdjNext* next, hipStream_t st) {
    if (adj_fuses_next(s)) {
        const AdjNext none{};
        if (s.aexp) launch_gather_tile2d<true, true, false, false>(s, w, none, st);
        else if (adj_coded(s)) launch_gather_tile2d<true, false, true, false>(s, w, none, st);
        else launch_gather_tile2d<true, false, false, false>(s, w, none, st);
        if (!next) {
            launch_gather_tile2d<false, false, false, false>(s, w, none, st);
        } else {
            AdjView wn = w;  // the next step's pressure transposed reads its own tape
            wn.tdiv = next->tdiv;
            if (s.aexp) launch_gather_tile2d<false, true, false, true>(s, wn, *next, st);
            else if (adj_coded(s)) launch_gather_tile2d<false, false, true, true>(s, wn, *next, st);
            else launch_gather_tile2d<false, false, false, true>(s, wn, *next, st);
        }
        return hipGetLastError();
    }
    if (next) return hipErrorInvalidValue;  // the per-cell kernels do not take the next step
    const dim3 g = adj_grid(s), b(64, 4);
    const bool u = s.ndmap == 1;
    if (s.ndim == 2) {
        if (u) launch_rest_nd<2, true>(s, w, g, b, st);
        else launch_rest_nd<2, false>(s, w, g, b, st);
    } else {
        if (u) launch_rest_nd<3, true>(s, w, g, b, st);
        else launch_rest_nd<3, false>(s, w, g, b, st);
    }
    return hipGetLastError();
}

hipError_t launch_adj_rho(const DevSim& s, const AdjView& w, float* grho, hipStream_t st) {
    const int64_t n = s.nx * s.sx;
    hipLaunchKernelGGL(k_adj_rho, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s, w, grho);
    return hipGetLastError();
}

}  // namespace fw
