// Streaming (x-marching) kernels for gfx950.
//
// 2D grid [nx][ny], y contiguous. A workgroup owns TY consecutive columns (one per lane, 64-wide
// wavefronts, TY / 64 waves) and a chunk of rows; it marches down the chunk one row per
// iteration. The 16-point axial (x) stencil reads a per-lane register queue of the last 16 rows
// of the input field (one new coalesced row load per iteration). The lateral (y) stencil and the
// transverse corrections read rows staged in a 4-slot LDS ring (current row with an 8-cell halo
// on both sides); one barrier per iteration. Every map and memory variable is read and written
// once per cell, coalesced.
//
// Arithmetic order is the per-cell kernels' (fw_kernels_simple.hip), so both paths compute the
// same sums.
#include "fw_internal.h"

namespace fw {

hipError_t launch_velocity_simple(const DevSim& s, hipStream_t st);
hipError_t launch_pressure_simple(const DevSim& s, hipStream_t st);

namespace {

constexpr int kRing = 4;  // LDS ring slots (rows i-1, i, i+1 in use, one being refilled)
constexpr int kChunk2D = 64;  // rows per workgroup

template <int TY>
struct Ring2D {
    float row[kRing][TY + 2 * kHalo];
};

// coefficient set for an axis: from the per-voxel table, or the single set when ndmap == 1
template <bool UNIFORM>
__device__ inline const float* coef_row(const float* lds_coef, const int32_t* dcmap, int64_t c, int axis,
                                        int ndim) {
    if (UNIFORM) return lds_coef + axis * kNCoef;
    return lds_coef + ((int64_t)dcmap[c] * ndim + axis) * kNCoef;
}

__device__ inline void load_coef_table(float* lds_coef, const float* coef, int n) {
    for (int t = threadIdx.x; t < n; t += blockDim.x) lds_coef[t] = coef[t];
}

// ---------------------------------------------------------------------------------------------
// velocity, 2D
template <int TY, bool UNIFORM, bool ISO>
__global__ __launch_bounds__(TY) void k_velocity_2d(DevSim s) {
    extern __shared__ float lds_coef[];
    __shared__ Ring2D<TY> ring;
    const int tid = threadIdx.x;
    const int64_t ny = s.ny, nx = s.nx;
    const int64_t j0 = (int64_t)blockIdx.x * TY;
    const int64_t j = j0 + tid;
    const int64_t i0 = kHalo + (int64_t)blockIdx.y * kChunk2D;
    const int64_t i1 = min(i0 + kChunk2D, nx - kHalo);
    const bool col_ok = j < ny;
    const bool active = j >= kHalo && j < ny - kHalo;
    const float* __restrict__ p = s.p;

    load_coef_table(lds_coef, s.coef, (UNIFORM ? 1 : s.ndmap) * 2 * kNCoef);

    // halo columns of a row: lanes 0..7 take j0-8..j0-1, lanes 8..15 take j0+TY..j0+TY+7
    auto stage_row = [&](int64_t r, float own) {
        float* dst = ring.row[r & (kRing - 1)];
        dst[kHalo + tid] = own;
        if (tid < 2 * kHalo) {
            const int64_t hj = tid < kHalo ? j0 - kHalo + tid : j0 + TY + (tid - kHalo);
            const int slot = tid < kHalo ? tid : TY + tid;
            dst[slot] = (hj >= 0 && hj < ny) ? p[r * ny + hj] : 0.f;
        }
    };

    float q[16];  // q[m] = p(i - 7 + m, j)
#pragma unroll
    for (int m = 0; m < 16; ++m) q[m] = col_ok ? p[(i0 - 7 + m) * ny + j] : 0.f;
    stage_row(i0 - 1, q[6]);
    stage_row(i0, q[7]);
    float rho_i = col_ok ? s.rho[i0 * ny + j] : 0.f;

    for (int64_t i = i0; i < i1; ++i) {
        if (i > i0) {
#pragma unroll
            for (int m = 0; m < 15; ++m) q[m] = q[m + 1];
            q[15] = col_ok ? p[(i + 8) * ny + j] : 0.f;
        }
        stage_row(i + 1, q[8]);
        __syncthreads();
        const int64_t c = i * ny + j;
        const float rho_next = col_ok ? s.rho[c + ny] : 0.f;
        if (active) {
            const float* rm = ring.row[(i - 1) & (kRing - 1)] + kHalo + tid;
            const float* r0 = ring.row[i & (kRing - 1)] + kHalo + tid;
            const float* rp = ring.row[(i + 1) & (kRing - 1)] + kHalo + tid;
            // x: axial from the register queue, transverse from rows i, i+1 at j-1, j+1
            {
                const float* w = coef_row<UNIFORM>(lds_coef, s.dcmap, c, 0, 2);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (q[7 + m] - q[8 - m]);
                float tr = 0.f;
                tr += (rp[1] - r0[1]) + (rp[-1] - r0[-1]);
                const float D = acc + w[8] * tr;
                const float b1 = s.bv[0][0][c], b2 = s.bv[0][1][c];
                const float a1 = s.av[0][0][c], a2 = s.av[0][1][c];
                const float qq1 = b1 * s.psi[0][0][c] + a1 * D;
                const float qq2 = b2 * s.psi[0][1][c] + a2 * D;
                s.psi[0][0][c] = qq1;
                s.psi[0][1][c] = qq2;
                const float rho_bar = 0.5f * (rho_i + rho_next);
                s.v[0][c] -= s.dt / rho_bar * (D * s.ikv[0][c] + (qq1 + qq2));
            }
            // y: axial from row i in LDS, transverse from rows i-1, i+1 at j, j+1
            {
                const float* w = coef_row<UNIFORM>(lds_coef, s.dcmap, c, 1, 2);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (r0[m] - r0[-(m - 1)]);
                float tr = 0.f;
                tr += (rp[1] - rp[0]) + (rm[1] - rm[0]);
                const float D = acc + w[8] * tr;
                const int ax = ISO ? 0 : 1;
                const float b1 = s.bv[ax][0][c], b2 = s.bv[ax][1][c];
                const float a1 = s.av[ax][0][c], a2 = s.av[ax][1][c];
                const float qq1 = b1 * s.psi[1][0][c] + a1 * D;
                const float qq2 = b2 * s.psi[1][1][c] + a2 * D;
                s.psi[1][0][c] = qq1;
                s.psi[1][1][c] = qq2;
                const float rho_bar = 0.5f * (rho_i + s.rho[c + 1]);
                const float ik = ISO ? s.ikv[0][c] : s.ikv[1][c];
                s.v[1][c] -= s.dt / rho_bar * (D * ik + (qq1 + qq2));
            }
        }
        rho_i = rho_next;
    }
}

// ---------------------------------------------------------------------------------------------
// pressure, 2D
template <int TY, bool UNIFORM, bool ISO>
__global__ __launch_bounds__(TY) void k_pressure_2d(DevSim s) {
    extern __shared__ float lds_coef[];
    __shared__ Ring2D<TY> ring_y;  // vy rows i-1, i, i+1 (row i with the 8-cell halo)
    __shared__ float ring_x[kRing][TY + 2];  // vx rows i-1, i with one halo column per side
    const int tid = threadIdx.x;
    const int64_t ny = s.ny, nx = s.nx;
    const int64_t j0 = (int64_t)blockIdx.x * TY;
    const int64_t j = j0 + tid;
    const int64_t i0 = kHalo + (int64_t)blockIdx.y * kChunk2D;
    const int64_t i1 = min(i0 + kChunk2D, nx - kHalo);
    const bool col_ok = j < ny;
    const bool active = j >= kHalo && j < ny - kHalo;
    const float* __restrict__ vx = s.v[0];
    const float* __restrict__ vy = s.v[1];

    load_coef_table(lds_coef, s.coef, (UNIFORM ? 1 : s.ndmap) * 2 * kNCoef);

    auto stage_vy = [&](int64_t r, float own) {
        float* dst = ring_y.row[r & (kRing - 1)];
        dst[kHalo + tid] = own;
        if (tid < 2 * kHalo) {
            const int64_t hj = tid < kHalo ? j0 - kHalo + tid : j0 + TY + (tid - kHalo);
            const int slot = tid < kHalo ? tid : TY + tid;
            dst[slot] = (hj >= 0 && hj < ny) ? vy[r * ny + hj] : 0.f;
        }
    };
    auto stage_vx = [&](int64_t r, float own) {
        float* dst = ring_x[r & (kRing - 1)];
        dst[1 + tid] = own;
        if (tid < 2) {
            const int64_t hj = tid == 0 ? j0 - 1 : j0 + TY;
            dst[tid == 0 ? 0 : TY + 1] = (hj >= 0 && hj < ny) ? vx[r * ny + hj] : 0.f;
        }
    };

    float q[16];  // q[m] = vx(i - 8 + m, j)
#pragma unroll
    for (int m = 0; m < 16; ++m) q[m] = col_ok ? vx[(i0 - 8 + m) * ny + j] : 0.f;
    float vy_next = col_ok ? vy[i0 * ny + j] : 0.f;
    stage_vy(i0 - 1, col_ok ? vy[(i0 - 1) * ny + j] : 0.f);
    stage_vy(i0, vy_next);
    stage_vx(i0 - 1, q[7]);

    for (int64_t i = i0; i < i1; ++i) {
        if (i > i0) {
#pragma unroll
            for (int m = 0; m < 15; ++m) q[m] = q[m + 1];
            q[15] = col_ok ? vx[(i + 7) * ny + j] : 0.f;
        }
        stage_vx(i, q[8]);
        stage_vy(i + 1, col_ok ? vy[(i + 1) * ny + j] : 0.f);
        __syncthreads();
        const int64_t c = i * ny + j;
        if (active) {
            const float* xm = ring_x[(i - 1) & (kRing - 1)] + 1 + tid;
            const float* x0 = ring_x[i & (kRing - 1)] + 1 + tid;
            const float* ym = ring_y.row[(i - 1) & (kRing - 1)] + kHalo + tid;
            const float* y0 = ring_y.row[i & (kRing - 1)] + kHalo + tid;
            const float* yp = ring_y.row[(i + 1) & (kRing - 1)] + kHalo + tid;
            float div = 0.f;
            {
                const float* w = coef_row<UNIFORM>(lds_coef, s.dcmap, c, 0, 2);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (q[7 + m] - q[8 - m]);
                float tr = 0.f;
                tr += (x0[1] - xm[1]) + (x0[-1] - xm[-1]);
                const float D = acc + w[8] * tr;
                const float qq1 = s.bp[0][0][c] * s.phi[0][0][c] + s.ap[0][0][c] * D;
                const float qq2 = s.bp[0][1][c] * s.phi[0][1][c] + s.ap[0][1][c] * D;
                s.phi[0][0][c] = qq1;
                s.phi[0][1][c] = qq2;
                div += D * s.ikp[0][c] + (qq1 + qq2);
            }
            {
                const float* w = coef_row<UNIFORM>(lds_coef, s.dcmap, c, 1, 2);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (y0[m - 1] - y0[-m]);
                float tr = 0.f;
                tr += (yp[0] - yp[-1]) + (ym[0] - ym[-1]);
                const float D = acc + w[8] * tr;
                const int ax = ISO ? 0 : 1;
                const float qq1 = s.bp[ax][0][c] * s.phi[1][0][c] + s.ap[ax][0][c] * D;
                const float qq2 = s.bp[ax][1][c] * s.phi[1][1][c] + s.ap[ax][1][c] * D;
                s.phi[1][0][c] = qq1;
                s.phi[1][1][c] = qq2;
                const float ik = ISO ? s.ikp[0][c] : s.ikp[1][c];
                div += D * ik + (qq1 + qq2);
            }
            s.p[c] -= s.dt * s.K[c] * div;
        }
    }
}

int coef_lds_bytes(const DevSim& s, bool uniform) { return (uniform ? 1 : s.ndmap) * s.ndim * kNCoef * 4; }

}  // namespace

// the per-voxel coefficient table goes to LDS; beyond this many sets the per-cell kernels run
constexpr int kMaxLdsCoefSets = 1024;

bool streaming_supported(const DevSim& s) {
    if (s.ndim != 2) return false;
    if (s.ndmap > kMaxLdsCoefSets) return false;
    return s.nx >= 2 * kHalo + 1 && s.ny >= 2 * kHalo + 1;
}

template <bool U, bool I>
static void launch_v2(const DevSim& s, hipStream_t st, dim3 grid, int lds) {
    hipLaunchKernelGGL((k_velocity_2d<256, U, I>), grid, dim3(256), lds, st, s);
}
template <bool U, bool I>
static void launch_p2(const DevSim& s, hipStream_t st, dim3 grid, int lds) {
    hipLaunchKernelGGL((k_pressure_2d<256, U, I>), grid, dim3(256), lds, st, s);
}

static dim3 grid_2d(const DevSim& s) {
    const int64_t rows = s.nx - 2 * kHalo;
    return dim3((unsigned)((s.ny + 255) / 256), (unsigned)((rows + kChunk2D - 1) / kChunk2D), 1);
}

hipError_t launch_velocity(const DevSim& s, hipStream_t st) {
    if (!streaming_supported(s)) return launch_velocity_simple(s, st);
    const bool u = s.ndmap == 1;
    const int lds = coef_lds_bytes(s, u);
    const dim3 g = grid_2d(s);
    if (u && s.isotropic) launch_v2<true, true>(s, st, g, lds);
    else if (u) launch_v2<true, false>(s, st, g, lds);
    else if (s.isotropic) launch_v2<false, true>(s, st, g, lds);
    else launch_v2<false, false>(s, st, g, lds);
    return hipGetLastError();
}

hipError_t launch_pressure(const DevSim& s, hipStream_t st) {
    if (!streaming_supported(s)) return launch_pressure_simple(s, st);
    const bool u = s.ndmap == 1;
    const int lds = coef_lds_bytes(s, u);
    const dim3 g = grid_2d(s);
    if (u && s.isotropic) launch_p2<true, true>(s, st, g, lds);
    else if (u) launch_p2<true, false>(s, st, g, lds);
    else if (s.isotropic) launch_p2<false, true>(s, st, g, lds);
    else launch_p2<false, false>(s, st, g, lds);
    return hipGetLastError();
}

}  // namespace fw
