// Streaming (x-marching) kernels for gfx950: the 2D kernels and the dispatch for both kernel
// families (3D marching kernels are in fw_kernels_stream3d.hip).
//
// 2D grid [nx][ny], y contiguous. A workgroup owns TY consecutive columns (one per lane, 64-wide
// wavefronts, TY / 64 waves) and a chunk of rows; it marches down the chunk one row per
// iteration. The 16-point axial (x) stencil reads a per-lane register queue of the last 16 rows
// of the input field. The lateral (y) stencil and the transverse corrections read rows staged
// in a 4-slot LDS ring (current row with an 8-cell halo on both sides); one barrier per row.
// Every map and memory variable is read and written once per cell, coalesced, and the operands
// of row r + 1 are loaded while row r is computed (two register sets, loop unrolled by two), so
// each workgroup keeps a row of loads in flight across the barrier.
//
// Arithmetic order is the per-cell kernels' (fw_kernels_simple.hip), so both paths compute the
// same sums.
#include "fw_internal.h"

namespace fw {

hipError_t launch_velocity_simple(const DevSim& s, hipStream_t st);
hipError_t launch_pressure_simple(const DevSim& s, hipStream_t st);
hipError_t launch_velocity_3d(const DevSim& s, const LaunchCfg& cfg, hipStream_t st);
hipError_t launch_pressure_3d(const DevSim& s, const LaunchCfg& cfg, hipStream_t st);

namespace {

constexpr int kRing = 4;  // LDS ring slots (rows r-1, r, r+1 in use, one being refilled)

template <bool UNIFORM>
__device__ inline const float* coef_row(const float* lds_coef, int dc, int axis) {
    if (UNIFORM) return lds_coef + axis * kNCoef;
    return lds_coef + (dc * 2 + axis) * kNCoef;
}

// ---------------------------------------------------------------------------------------------
// velocity, 2D

// operands of one row (r) of the velocity update for one lane
template <bool ISO>
struct VelRow {
    static constexpr int NA = ISO ? 1 : 2;
    float pq;         // p(r + 8, j): next element of the x queue
    float halo;       // lanes 0..15: p(r + 1, halo column) for LDS row r + 1
    float rho_next;   // rho(r + 1, j)
    float rho_right;  // rho(r, j + 1)
    int dc;
    float ik[NA], b1[NA], b2[NA], a1[NA], a2[NA];
    float ps1[2], ps2[2], v[2];
};

// The velocity update of one workgroup's tile: columns [bx * TY, bx * TY + TY), rows chunk `by`.
// CODED (isotropic only): the relaxation coefficients come from the code table (rcode, rtab_v)
template <int TY, bool UNIFORM, bool ISO, bool CODED>
__device__ __forceinline__ void velocity_2d_tile(const DevSim& s, unsigned bx, unsigned by, int chunk,
                                                 float* lds_coef) {
    __shared__ float ring[kRing][TY + 2 * kHalo];
    const int tid = threadIdx.x;
    const int64_t ny = s.ny, nx = s.nx;
    const int64_t j0 = (int64_t)bx * TY;
    const int64_t j = j0 + tid;
    const int64_t i0 = kHalo + (int64_t)by * chunk;
    const int64_t i1 = min(i0 + chunk, nx - kHalo);
    const bool col_ok = j < ny;
    const bool active = j >= kHalo && j < ny - kHalo;
    const bool halo_lane = tid < 2 * kHalo;
    const int64_t hj = tid < kHalo ? j0 - kHalo + tid : j0 + TY + (tid - kHalo);
    const bool halo_ok = halo_lane && hj >= 0 && hj < ny;
    const int halo_slot = tid < kHalo ? tid : TY + tid;
    const float* __restrict__ p = s.p;

    const int ncoef = (UNIFORM ? 1 : s.ndmap) * 2 * kNCoef;
    for (int t = tid; t < ncoef; t += TY) lds_coef[t] = s.coef[t];

    auto load = [&](VelRow<ISO>& R, int64_t r) {
        const uint32_t c = (uint32_t)(r * ny + j);
        R.pq = col_ok ? gld32((const float*)p, c + (uint32_t)(8 * ny)) : 0.f;
        R.rho_next = col_ok ? gld32((const float*)s.rho, c + (uint32_t)(ny)) : 0.f;
        R.halo = halo_ok ? gld32((const float*)p, (uint32_t)((r + 1) * ny + hj)) : 0.f;
        if (active) {
            R.rho_right = gld32((const float*)s.rho, c + (uint32_t)(1));
            if (CODED) {
                const CodedCoef k = cell_coef<true, UNIFORM>(s, c);
                R.ik[0] = k.ik;
                R.a1[0] = k.a1;
                R.a2[0] = k.a2;
                R.b1[0] = k.b1;
                R.b2[0] = k.b2;
                R.dc = UNIFORM ? 0 : k.dc;
            } else {
                R.dc = UNIFORM ? 0 : gld32((const int32_t*)s.dcmap, c);
#pragma unroll
                for (int a = 0; a < VelRow<ISO>::NA; ++a) {
                    R.ik[a] = gld32((const float*)s.ikv[a], c);
                    R.b1[a] = gld32((const float*)s.bv[a][0], c);
                    R.b2[a] = gld32((const float*)s.bv[a][1], c);
                    R.a1[a] = gld32((const float*)s.av[a][0], c);
                    R.a2[a] = gld32((const float*)s.av[a][1], c);
                }
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                R.ps1[a] = gldnt(s.psi[a][0], c);
                R.ps2[a] = gldnt(s.psi[a][1], c);
                R.v[a] = gld32((const float*)s.v[a], c);
            }
        }
    };

    float q[16];  // q[m] = p(r - 7 + m, j)
#pragma unroll
    for (int m = 0; m < 16; ++m) q[m] = col_ok ? gld32((const float*)p, (uint32_t)((i0 - 7 + m) * ny + j)) : 0.f;
    for (int64_t r = i0 - 1; r <= i0; ++r) {
        float* dst = ring[r & (kRing - 1)];
        dst[kHalo + tid] = r < i0 ? q[6] : q[7];
        if (halo_lane) dst[halo_slot] = halo_ok ? gld32((const float*)p, (uint32_t)(r * ny + hj)) : 0.f;
    }
    float rho_i = col_ok ? gld32((const float*)s.rho, (uint32_t)(i0 * ny + j)) : 0.f;

    auto step = [&](int64_t r, const VelRow<ISO>& R) {
        if (r > i0) {
#pragma unroll
            for (int m = 0; m < 15; ++m) q[m] = q[m + 1];
            q[15] = R.pq;
        }
        {
            float* dst = ring[(r + 1) & (kRing - 1)];
            dst[kHalo + tid] = q[8];
            if (halo_lane) dst[halo_slot] = R.halo;
        }
        __syncthreads();
        if (active) {
            const uint32_t c = (uint32_t)(r * ny + j);
            const float* rm = ring[(r - 1) & (kRing - 1)] + kHalo + tid;
            const float* r0 = ring[r & (kRing - 1)] + kHalo + tid;
            const float* rp = ring[(r + 1) & (kRing - 1)] + kHalo + tid;
            {  // x: axial from the register queue, transverse from rows r, r+1 at j-1, j+1
                const float* w = coef_row<UNIFORM>(lds_coef, R.dc, 0);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (q[7 + m] - q[8 - m]);
                float tr = 0.f;
                tr += (rp[1] - r0[1]) + (rp[-1] - r0[-1]);
                const float D = acc + w[8] * tr;
                const float qq1 = R.b1[0] * R.ps1[0] + R.a1[0] * D;
                const float qq2 = R.b2[0] * R.ps2[0] + R.a2[0] * D;
                gstnt(s.psi[0][0], c, qq1);
                gstnt(s.psi[0][1], c, qq2);
                const float rho_bar = 0.5f * (rho_i + R.rho_next);
                gst32(s.v[0], c, R.v[0] - s.dt / rho_bar * (D * R.ik[0] + (qq1 + qq2)));
            }
            {  // y: axial from row r in LDS, transverse from rows r-1, r+1 at j, j+1
                constexpr int ax = ISO ? 0 : 1;
                const float* w = coef_row<UNIFORM>(lds_coef, R.dc, 1);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (r0[m] - r0[-(m - 1)]);
                float tr = 0.f;
                tr += (rp[1] - rp[0]) + (rm[1] - rm[0]);
                const float D = acc + w[8] * tr;
                const float qq1 = R.b1[ax] * R.ps1[1] + R.a1[ax] * D;
                const float qq2 = R.b2[ax] * R.ps2[1] + R.a2[ax] * D;
                gstnt(s.psi[1][0], c, qq1);
                gstnt(s.psi[1][1], c, qq2);
                const float rho_bar = 0.5f * (rho_i + R.rho_right);
                gst32(s.v[1], c, R.v[1] - s.dt / rho_bar * (D * R.ik[ax] + (qq1 + qq2)));
            }
        }
        rho_i = R.rho_next;
    };

    VelRow<ISO> A, B;
    load(A, i0);
    for (int64_t r = i0; r < i1; r += 2) {
        if (r + 1 < i1) load(B, r + 1);
        step(r, A);
        if (r + 1 >= i1) break;
        if (r + 2 < i1) load(A, r + 2);
        step(r + 1, B);
    }
}

template <int TY, bool UNIFORM, bool ISO, bool CODED>
__global__ __launch_bounds__(TY) void k_velocity_2d(DevSim batch, int chunk) {
    extern __shared__ float lds_coef[];
    velocity_2d_tile<TY, UNIFORM, ISO, CODED>(shot_view(batch, blockIdx.z), blockIdx.x, blockIdx.y, chunk, lds_coef);
}

// ---------------------------------------------------------------------------------------------
// pressure, 2D

template <bool ISO>
struct PreRow {
    static constexpr int NA = ISO ? 1 : 2;
    float vxq;      // vx(r + 7, j): next element of the x queue
    float vy_own;   // vy(r + 1, j) for LDS row r + 1
    float vy_halo;  // lanes 0..15: vy(r + 1, halo column)
    float vx_halo;  // lanes 0..1: vx(r, j0 - 1) / vx(r, j0 + TY)
    float K, p;
    int dc;
    float ik[NA], b1[NA], b2[NA], a1[NA], a2[NA];
    float ph1[2], ph2[2];
};

// The pressure update of one workgroup's tile (as velocity_2d_tile).
template <int TY, bool UNIFORM, bool ISO, bool CODED>
__device__ __forceinline__ void pressure_2d_tile(const DevSim& s, unsigned bx, unsigned by, int chunk,
                                                 float* lds_coef) {
    __shared__ float ring_y[kRing][TY + 2 * kHalo];  // vy rows r-1, r, r+1 (8-cell halo)
    __shared__ float ring_x[kRing][TY + 2];          // vx rows r-1, r (1-cell halo)
    const int tid = threadIdx.x;
    const int64_t ny = s.ny, nx = s.nx;
    const int64_t j0 = (int64_t)bx * TY;
    const int64_t j = j0 + tid;
    const int64_t i0 = kHalo + (int64_t)by * chunk;
    const int64_t i1 = min(i0 + chunk, nx - kHalo);
    const bool col_ok = j < ny;
    const bool active = j >= kHalo && j < ny - kHalo;
    const bool halo_lane = tid < 2 * kHalo;
    const int64_t hj = tid < kHalo ? j0 - kHalo + tid : j0 + TY + (tid - kHalo);
    const bool halo_ok = halo_lane && hj >= 0 && hj < ny;
    const int halo_slot = tid < kHalo ? tid : TY + tid;
    const bool xh_lane = tid < 2;
    const int64_t xhj = tid == 0 ? j0 - 1 : j0 + TY;
    const bool xh_ok = xh_lane && xhj >= 0 && xhj < ny;
    const int xh_slot = tid == 0 ? 0 : TY + 1;
    const float* __restrict__ vx = s.v[0];
    const float* __restrict__ vy = s.v[1];

    const int ncoef = (UNIFORM ? 1 : s.ndmap) * 2 * kNCoef;
    for (int t = tid; t < ncoef; t += TY) lds_coef[t] = s.coef[t];

    auto load = [&](PreRow<ISO>& R, int64_t r) {
        const uint32_t c = (uint32_t)(r * ny + j);
        R.vxq = col_ok ? gld32((const float*)vx, c + (uint32_t)(7 * ny)) : 0.f;
        R.vy_own = col_ok ? gld32((const float*)vy, c + (uint32_t)(ny)) : 0.f;
        R.vy_halo = halo_ok ? gld32((const float*)vy, (uint32_t)((r + 1) * ny + hj)) : 0.f;
        R.vx_halo = xh_ok ? gld32((const float*)vx, (uint32_t)(r * ny + xhj)) : 0.f;
        if (active) {
            R.K = gld32((const float*)s.K, c);
            R.p = gld32((const float*)s.p, c);
            if (CODED) {
                const CodedCoef k = cell_coef<false, UNIFORM>(s, c);
                R.ik[0] = k.ik;
                R.a1[0] = k.a1;
                R.a2[0] = k.a2;
                R.b1[0] = k.b1;
                R.b2[0] = k.b2;
                R.dc = UNIFORM ? 0 : k.dc;
            } else {
                R.dc = UNIFORM ? 0 : gld32((const int32_t*)s.dcmap, c);
#pragma unroll
                for (int a = 0; a < PreRow<ISO>::NA; ++a) {
                    R.ik[a] = gld32((const float*)s.ikp[a], c);
                    R.b1[a] = gld32((const float*)s.bp[a][0], c);
                    R.b2[a] = gld32((const float*)s.bp[a][1], c);
                    R.a1[a] = gld32((const float*)s.ap[a][0], c);
                    R.a2[a] = gld32((const float*)s.ap[a][1], c);
                }
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                R.ph1[a] = gldnt(s.phi[a][0], c);
                R.ph2[a] = gldnt(s.phi[a][1], c);
            }
        }
    };

    float q[16];  // q[m] = vx(r - 8 + m, j)
#pragma unroll
    for (int m = 0; m < 16; ++m) q[m] = col_ok ? gld32((const float*)vx, (uint32_t)((i0 - 8 + m) * ny + j)) : 0.f;
    for (int64_t r = i0 - 1; r <= i0; ++r) {
        float* dst = ring_y[r & (kRing - 1)];
        dst[kHalo + tid] = col_ok ? gld32((const float*)vy, (uint32_t)(r * ny + j)) : 0.f;
        if (halo_lane) dst[halo_slot] = halo_ok ? gld32((const float*)vy, (uint32_t)(r * ny + hj)) : 0.f;
    }
    {
        float* dst = ring_x[(i0 - 1) & (kRing - 1)];
        dst[1 + tid] = q[7];
        if (xh_lane) dst[xh_slot] = xh_ok ? gld32((const float*)vx, (uint32_t)((i0 - 1) * ny + xhj)) : 0.f;
    }

    auto step = [&](int64_t r, const PreRow<ISO>& R) {
        if (r > i0) {
#pragma unroll
            for (int m = 0; m < 15; ++m) q[m] = q[m + 1];
            q[15] = R.vxq;
        }
        {
            float* dx_ = ring_x[r & (kRing - 1)];
            dx_[1 + tid] = q[8];
            if (xh_lane) dx_[xh_slot] = R.vx_halo;
            float* dy_ = ring_y[(r + 1) & (kRing - 1)];
            dy_[kHalo + tid] = R.vy_own;
            if (halo_lane) dy_[halo_slot] = R.vy_halo;
        }
        __syncthreads();
        if (active) {
            const uint32_t c = (uint32_t)(r * ny + j);
            const float* xm = ring_x[(r - 1) & (kRing - 1)] + 1 + tid;
            const float* x0 = ring_x[r & (kRing - 1)] + 1 + tid;
            const float* ym = ring_y[(r - 1) & (kRing - 1)] + kHalo + tid;
            const float* y0 = ring_y[r & (kRing - 1)] + kHalo + tid;
            const float* yp = ring_y[(r + 1) & (kRing - 1)] + kHalo + tid;
            float div = 0.f;
            {
                const float* w = coef_row<UNIFORM>(lds_coef, R.dc, 0);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (q[7 + m] - q[8 - m]);
                float tr = 0.f;
                tr += (x0[1] - xm[1]) + (x0[-1] - xm[-1]);
                const float D = acc + w[8] * tr;
                const float qq1 = R.b1[0] * R.ph1[0] + R.a1[0] * D;
                const float qq2 = R.b2[0] * R.ph2[0] + R.a2[0] * D;
                gstnt(s.phi[0][0], c, qq1);
                gstnt(s.phi[0][1], c, qq2);
                div += D * R.ik[0] + (qq1 + qq2);
            }
            {
                constexpr int ax = ISO ? 0 : 1;
                const float* w = coef_row<UNIFORM>(lds_coef, R.dc, 1);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (y0[m - 1] - y0[-m]);
                float tr = 0.f;
                tr += (yp[0] - yp[-1]) + (ym[0] - ym[-1]);
                const float D = acc + w[8] * tr;
                const float qq1 = R.b1[ax] * R.ph1[1] + R.a1[ax] * D;
                const float qq2 = R.b2[ax] * R.ph2[1] + R.a2[ax] * D;
                gstnt(s.phi[1][0], c, qq1);
                gstnt(s.phi[1][1], c, qq2);
                div += D * R.ik[ax] + (qq1 + qq2);
            }
            gst32(s.p, c, R.p - s.dt * R.K * div);
        }
    };

    PreRow<ISO> A, B;
    load(A, i0);
    for (int64_t r = i0; r < i1; r += 2) {
        if (r + 1 < i1) load(B, r + 1);
        step(r, A);
        if (r + 1 >= i1) break;
        if (r + 2 < i1) load(A, r + 2);
        step(r + 1, B);
    }
}

// at most 96 VGPRs (five waves per SIMD instead of four): measured 3-15 % faster across the
// 128/256-lane shapes, for a few bytes of spill per lane
// gs.g non-null: each workgroup then also does the step's source, pressure-release and sensor
// work of its tile (tile blockIdx.y * gridDim.x + blockIdx.x of shot blockIdx.z), as the group
// kernels do
template <int TY, bool UNIFORM, bool ISO, bool CODED>
__global__ __launch_bounds__(TY) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_pressure_2d(DevSim batch, int chunk,
                                                                                               GroupStep gs) {
    extern __shared__ float lds_coef[];
    TileWork w{};
    if (gs.g) {
        const GroupBoundary& e = gs.g[blockIdx.z];
        w = tile_work(e, blockIdx.y * gridDim.x + blockIdx.x, e.step0 + gs.it + (gs.base ? *gs.base : 0));
    }
    pressure_2d_tile<TY, UNIFORM, ISO, CODED>(shot_view(batch, blockIdx.z), blockIdx.x, blockIdx.y, chunk, lds_coef);
    if (!gs.g) return;
    __syncthreads();
    tile_boundary(w);
}

// ---------------------------------------------------------------------------------------------
// grouped launches: the tiles of several grids (different sizes and media, isotropic
// relaxation) in one launch. Workgroup b belongs to the last entry with start <= b.

template <int TY>
__global__ __launch_bounds__(TY) void k_velocity_2d_group(const GroupTile* __restrict__ t, int n) {
    extern __shared__ float lds_coef[];
    const int m = group_entry(t, n);
    const DevSim s = t[m].s;
    const int chunk = t[m].chunk;
    const unsigned local = blockIdx.x - t[m].start;
    const unsigned bx = local % t[m].gx, by = local / t[m].gx;
    if (s.rcode) velocity_2d_tile<TY, false, true, true>(s, bx, by, chunk, lds_coef);
    else velocity_2d_tile<TY, false, true, false>(s, bx, by, chunk, lds_coef);
}

// the pressure update of the tile, then its share of the step's boundary work (gs.g non-null)
template <int TY>
__global__ __launch_bounds__(TY) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_pressure_2d_group(
    const GroupTile* __restrict__ t, int n, GroupStep gs) {
    extern __shared__ float lds_coef[];
    const int m = group_entry(t, n);
    const DevSim s = t[m].s;
    const int chunk = t[m].chunk;
    const unsigned local = blockIdx.x - t[m].start;
    const unsigned bx = local % t[m].gx, by = local / t[m].gx;
    TileWork w{};
    if (gs.g) w = tile_work(gs.g[m], local, gs.g[m].step0 + gs.it + (gs.base ? *gs.base : 0));
    if (s.rcode) pressure_2d_tile<TY, false, true, true>(s, bx, by, chunk, lds_coef);
    else pressure_2d_tile<TY, false, true, false>(s, bx, by, chunk, lds_coef);
    if (!gs.g) return;
    __syncthreads();
    tile_boundary(w);
}

template <int TY>
hipError_t launch_2d(const DevSim& s, int chunk, bool velocity, const GroupStep& gs, hipStream_t st) {
    const bool u = s.ndmap == 1;
    const int lds = (u ? 1 : s.ndmap) * 2 * kNCoef * 4;
    const int64_t rows = s.nx - 2 * kHalo;
    const dim3 g((unsigned)((s.ny + TY - 1) / TY), (unsigned)((rows + chunk - 1) / chunk), (unsigned)s.nshots);
    const dim3 b(TY);
    const bool coded = s.isotropic && s.rcode;
    if (velocity) {
        if (u && coded) hipLaunchKernelGGL((k_velocity_2d<TY, true, true, true>), g, b, lds, st, s, chunk);
        else if (u && s.isotropic) hipLaunchKernelGGL((k_velocity_2d<TY, true, true, false>), g, b, lds, st, s, chunk);
        else if (u) hipLaunchKernelGGL((k_velocity_2d<TY, true, false, false>), g, b, lds, st, s, chunk);
        else if (coded) hipLaunchKernelGGL((k_velocity_2d<TY, false, true, true>), g, b, lds, st, s, chunk);
        else if (s.isotropic) hipLaunchKernelGGL((k_velocity_2d<TY, false, true, false>), g, b, lds, st, s, chunk);
        else hipLaunchKernelGGL((k_velocity_2d<TY, false, false, false>), g, b, lds, st, s, chunk);
    } else {
        if (u && coded) hipLaunchKernelGGL((k_pressure_2d<TY, true, true, true>), g, b, lds, st, s, chunk, gs);
        else if (u && s.isotropic) hipLaunchKernelGGL((k_pressure_2d<TY, true, true, false>), g, b, lds, st, s, chunk, gs);
        else if (u) hipLaunchKernelGGL((k_pressure_2d<TY, true, false, false>), g, b, lds, st, s, chunk, gs);
        else if (coded) hipLaunchKernelGGL((k_pressure_2d<TY, false, true, true>), g, b, lds, st, s, chunk, gs);
        else if (s.isotropic) hipLaunchKernelGGL((k_pressure_2d<TY, false, true, false>), g, b, lds, st, s, chunk, gs);
        else hipLaunchKernelGGL((k_pressure_2d<TY, false, false, false>), g, b, lds, st, s, chunk, gs);
    }
    return hipGetLastError();
}

hipError_t launch_2d_any(const DevSim& s, const LaunchCfg& cfg, bool velocity, const GroupStep& gs, hipStream_t st) {
    const int chunk = cfg.chunk < 2 ? 2 : cfg.chunk;
    switch (cfg.width) {
        case 64: return launch_2d<64>(s, chunk, velocity, gs, st);
        case 256: return launch_2d<256>(s, chunk, velocity, gs, st);
        default: return launch_2d<128>(s, chunk, velocity, gs, st);
    }
}

template <int TY>
hipError_t launch_group(const GroupTile* t, int n, uint32_t blocks, int lds, bool velocity, const GroupStep& gs,
                        hipStream_t st) {
    if (velocity) hipLaunchKernelGGL((k_velocity_2d_group<TY>), dim3(blocks), dim3(TY), lds, st, t, n);
    else hipLaunchKernelGGL((k_pressure_2d_group<TY>), dim3(blocks), dim3(TY), lds, st, t, n, gs);
    return hipGetLastError();
}

}  // namespace

unsigned group_tiles_x(int64_t ny, int width) { return (unsigned)((ny + width - 1) / width); }
unsigned group_tiles_y(int64_t nx, int chunk) { return (unsigned)((nx - 2 * kHalo + chunk - 1) / chunk); }

hipError_t launch_group_2d(const GroupTile* t, int n, uint32_t blocks, int width, int max_ndmap, bool velocity,
                           const GroupStep& gs, hipStream_t st) {
    if (n <= 0 || blocks == 0) return hipSuccess;
    const int lds = max_ndmap * 2 * kNCoef * 4;
    switch (width) {
        case 64: return launch_group<64>(t, n, blocks, lds, velocity, gs, st);
        case 128: return launch_group<128>(t, n, blocks, lds, velocity, gs, st);
        default: return launch_group<256>(t, n, blocks, lds, velocity, gs, st);
    }
}

// the per-voxel coefficient table goes to LDS; beyond this many sets the 2D grids run the
// per-cell kernels
constexpr int kMaxLdsCoefSets2D = 1024;

bool streaming_supported(const DevSim& s) {
    if (s.aexp) return false;  // the exponential model runs the per-cell kernels
    if (s.ndim == 3) return true;  // large 3D tables are read from global memory
    // 2D: the coefficient table fits LDS, and 32-bit byte offsets cover every array
    return s.ndmap <= kMaxLdsCoefSets2D && s.nx * s.ny < ((int64_t)1 << 30);
}

bool tile2d_supported(const DevSim& s);
hipError_t launch_tile2d(const DevSim& s, const LaunchCfg& cfg, bool velocity, const GroupStep& gs, hipStream_t st);

bool tape_capable(const DevSim& s, const LaunchCfg& cfg) {
    return cfg.variant == 0 || (cfg.variant == 2 && tile2d_supported(s));
}

// A step that writes the adjoint tape runs the tiled 2D kernels when `cfg` asks for them and the
// per-cell kernels otherwise (the only families that write it).
hipError_t launch_velocity(const DevSim& s, const LaunchCfg& cfg, hipStream_t st) {
    if (cfg.variant == 2 && tile2d_supported(s)) return launch_tile2d(s, cfg, true, GroupStep{}, st);
    if (s.tape_div) return launch_velocity_simple(s, st);
    if ((cfg.variant == 3 || cfg.variant == 4) && s.ndim == 3 && streaming_supported(s))
        return launch_pipe3d(s, cfg, true, st);
    if (cfg.variant == 0 || !streaming_supported(s)) return launch_velocity_simple(s, st);
    if (s.ndim == 3) return launch_velocity_3d(s, cfg, st);
    return launch_2d_any(s, cfg, true, GroupStep{}, st);
}

bool pressure_fuses_boundary(const DevSim& s, const LaunchCfg& cfg) {
    if (s.tape_div || s.ndim != 2) return false;
    if (cfg.variant == 2) return tile2d_supported(s);
    return cfg.variant == 1 && streaming_supported(s);
}

hipError_t launch_pressure(const DevSim& s, const LaunchCfg& cfg, hipStream_t st, const GroupStep& gs) {
    if (gs.g && !pressure_fuses_boundary(s, cfg)) return hipErrorInvalidValue;
    if (cfg.variant == 2 && tile2d_supported(s)) return launch_tile2d(s, cfg, false, gs, st);
    if (s.tape_div) return launch_pressure_simple(s, st);
    if ((cfg.variant == 3 || cfg.variant == 4) && s.ndim == 3 && streaming_supported(s))
        return launch_pipe3d(s, cfg, false, st);
    if (cfg.variant == 0 || !streaming_supported(s)) return launch_pressure_simple(s, st);
    if (s.ndim == 3) return launch_pressure_3d(s, cfg, st);
    return launch_2d_any(s, cfg, false, gs, st);
}

}  // namespace fw
