// Streaming (x-marching) kernels for 3D grids on gfx950.
//
// Grid [nx][ny][nz], z contiguous. A workgroup owns a TY x TZ tile of (y, z) columns — TZ = 64
// lanes of one wavefront along z, TY = 8 or 16 wavefronts along y — and a chunk of x planes that it
// marches through one plane per iteration. The axial x stencil reads a per-lane register queue
// of the last 16 planes of the input field. The y and z stencils and the transverse
// corrections read the tile of the current planes staged in 4-slot LDS rings (8-cell halo
// along the axis a field is differentiated on, 1 cell across). One barrier per plane.
//
// Arithmetic order is the per-cell kernels' (fw_kernels_simple.hip).
#include "fw_internal.h"

namespace fw {

namespace {

constexpr int TZ = 64;
// TY (rows of the tile along y, one wavefront each) is a template parameter: 8 or 16
constexpr int kRing3 = 4;

// a staged tile: rows y in [j0 - HY, j0 + TY + HY), columns z in [k0 - HZ, k0 + TZ + HZ)
template <int TY, int HY, int HZ>
struct Tile {
    static constexpr int PY = TY + 2 * HY;
    static constexpr int PZ = TZ + 2 * HZ;
    static constexpr int N = PY * PZ;
    float v[kRing3][N];
    __device__ float* slot(int64_t r) { return v[r & (kRing3 - 1)]; }
    // element (j, k) of the tile, relative to the thread's own (ty, tz)
    __device__ static int at(int ty, int tz, int dy, int dz) { return (ty + HY + dy) * PZ + (tz + HZ + dz); }
    __device__ void stage(const float* __restrict__ f, int64_t r, int64_t j0, int64_t k0, const DevSim& s) {
        float* dst = slot(r);
        const int64_t base = r * s.sx;
        for (int t = threadIdx.x; t < N; t += TZ * TY) {
            const int yy = t / PZ, zz = t - (t / PZ) * PZ;
            const int64_t jj = j0 - HY + yy, kk = k0 - HZ + zz;
            dst[t] = (jj >= 0 && jj < s.ny && kk >= 0 && kk < s.nz) ? f[base + jj * s.sy + kk] : 0.f;
        }
    }
};

// coefficient tables of up to kLdsSets3D sets are copied to LDS; larger ones are read from
// global memory (a few KB, cache resident)
constexpr int kLdsSets3D = 128;

template <bool UNIFORM>
__device__ inline const float* coef3(const float* lds_coef, const int32_t* dcmap, int64_t c, int axis) {
    if (UNIFORM) return lds_coef + axis * kNCoef;
    return lds_coef + ((int64_t)dcmap[c] * 3 + axis) * kNCoef;
}

__device__ inline const float* coef_base_3d(const DevSim& s, float* lds_coef, bool uniform) {
    const int nsets = uniform ? 1 : s.ndmap;
    if (nsets > kLdsSets3D) return s.coef;
    for (int t = threadIdx.x; t < nsets * 3 * kNCoef; t += blockDim.x) lds_coef[t] = s.coef[t];
    return lds_coef;
}

template <int TY, bool UNIFORM, bool ISO>
__global__ __launch_bounds__(TZ * TY) void k_velocity_3d(DevSim batch, int chunk) {
    unsigned zl;
    const DevSim s = shot_view(batch, shot_of_block(batch, zl));
    extern __shared__ float lds_coef[];
    __shared__ Tile<TY, kHalo, kHalo> P;  // p planes, 8-cell halo in y and z
    const int tz = threadIdx.x % TZ, ty = threadIdx.x / TZ;
    const int64_t k0 = (int64_t)blockIdx.x * TZ, j0 = (int64_t)blockIdx.y * TY;
    const int64_t k = k0 + tz, j = j0 + ty;
    const int64_t i0 = kHalo + (int64_t)zl * chunk;
    const int64_t i1 = min(i0 + chunk, s.nx - kHalo);
    const bool in_grid = j < s.ny && k < s.nz;
    const bool active = j >= kHalo && j < s.ny - kHalo && k >= kHalo && k < s.nz - kHalo;
    const int64_t sx = s.sx, sy = s.sy;
    const float* __restrict__ p = s.p;
    const float* coefs = coef_base_3d(s, lds_coef, UNIFORM);

    const int64_t col = j * sy + k;
    float q[16];  // q[m] = p(i - 7 + m, j, k)
#pragma unroll
    for (int m = 0; m < 16; ++m) q[m] = in_grid ? p[(i0 - 7 + m) * sx + col] : 0.f;
    P.stage(p, i0 - 1, j0, k0, s);
    P.stage(p, i0, j0, k0, s);
    float rho_i = in_grid ? s.rho[i0 * sx + col] : 0.f;

    for (int64_t i = i0; i < i1; ++i) {
        if (i > i0) {
#pragma unroll
            for (int m = 0; m < 15; ++m) q[m] = q[m + 1];
            q[15] = in_grid ? p[(i + 8) * sx + col] : 0.f;
        }
        P.stage(p, i + 1, j0, k0, s);
        __syncthreads();
        const int64_t c = i * sx + col;
        const float rho_next = in_grid ? s.rho[c + sx] : 0.f;
        if (active) {
            const float* pm = P.slot(i - 1);
            const float* p0 = P.slot(i);
            const float* pp = P.slot(i + 1);
            auto A = [&](const float* pl, int dy, int dz) { return pl[Tile<TY, kHalo, kHalo>::at(ty, tz, dy, dz)]; };
            // x
            {
                const float* w = coef3<UNIFORM>(coefs, s.dcmap, c, 0);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (q[7 + m] - q[8 - m]);
                float tr = 0.f;
                tr += (A(pp, 1, 0) - A(p0, 1, 0)) + (A(pp, -1, 0) - A(p0, -1, 0));
                tr += (A(pp, 0, 1) - A(p0, 0, 1)) + (A(pp, 0, -1) - A(p0, 0, -1));
                const float D = acc + w[8] * tr;
                const float qq1 = s.bv[0][0][c] * s.psi[0][0][c] + s.av[0][0][c] * D;
                const float qq2 = s.bv[0][1][c] * s.psi[0][1][c] + s.av[0][1][c] * D;
                s.psi[0][0][c] = qq1;
                s.psi[0][1][c] = qq2;
                const float rho_bar = 0.5f * (rho_i + rho_next);
                s.v[0][c] -= s.dt / rho_bar * (D * s.ikv[0][c] + (qq1 + qq2));
            }
            // y
            {
                const float* w = coef3<UNIFORM>(coefs, s.dcmap, c, 1);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (A(p0, m, 0) - A(p0, -(m - 1), 0));
                float tr = 0.f;
                tr += (A(pp, 1, 0) - A(pp, 0, 0)) + (A(pm, 1, 0) - A(pm, 0, 0));
                tr += (A(p0, 1, 1) - A(p0, 0, 1)) + (A(p0, 1, -1) - A(p0, 0, -1));
                const float D = acc + w[8] * tr;
                const int ax = ISO ? 0 : 1;
                const float qq1 = s.bv[ax][0][c] * s.psi[1][0][c] + s.av[ax][0][c] * D;
                const float qq2 = s.bv[ax][1][c] * s.psi[1][1][c] + s.av[ax][1][c] * D;
                s.psi[1][0][c] = qq1;
                s.psi[1][1][c] = qq2;
                const float rho_bar = 0.5f * (rho_i + s.rho[c + sy]);
                s.v[1][c] -= s.dt / rho_bar * (D * s.ikv[ax][c] + (qq1 + qq2));
            }
            // z
            {
                const float* w = coef3<UNIFORM>(coefs, s.dcmap, c, 2);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (A(p0, 0, m) - A(p0, 0, -(m - 1)));
                float tr = 0.f;
                tr += (A(pp, 0, 1) - A(pp, 0, 0)) + (A(pm, 0, 1) - A(pm, 0, 0));
                tr += (A(p0, 1, 1) - A(p0, 1, 0)) + (A(p0, -1, 1) - A(p0, -1, 0));
                const float D = acc + w[8] * tr;
                const int ax = ISO ? 0 : 2;
                const float qq1 = s.bv[ax][0][c] * s.psi[2][0][c] + s.av[ax][0][c] * D;
                const float qq2 = s.bv[ax][1][c] * s.psi[2][1][c] + s.av[ax][1][c] * D;
                s.psi[2][0][c] = qq1;
                s.psi[2][1][c] = qq2;
                const float rho_bar = 0.5f * (rho_i + s.rho[c + 1]);
                s.v[2][c] -= s.dt / rho_bar * (D * s.ikv[ax][c] + (qq1 + qq2));
            }
        }
        rho_i = rho_next;
    }
}

template <int TY, bool UNIFORM, bool ISO>
__global__ __launch_bounds__(TZ * TY) void k_pressure_3d(DevSim batch, int chunk) {
    unsigned zl;
    const DevSim s = shot_view(batch, shot_of_block(batch, zl));
    extern __shared__ float lds_coef[];
    __shared__ Tile<TY, 1, 1> VX;       // vx planes i-1, i
    __shared__ Tile<TY, kHalo, 1> VY;   // vy planes i-1, i, i+1; 8-cell halo in y
    __shared__ Tile<TY, 1, kHalo> VZ;   // vz planes i-1, i, i+1; 8-cell halo in z
    const int tz = threadIdx.x % TZ, ty = threadIdx.x / TZ;
    const int64_t k0 = (int64_t)blockIdx.x * TZ, j0 = (int64_t)blockIdx.y * TY;
    const int64_t k = k0 + tz, j = j0 + ty;
    const int64_t i0 = kHalo + (int64_t)zl * chunk;
    const int64_t i1 = min(i0 + chunk, s.nx - kHalo);
    const bool in_grid = j < s.ny && k < s.nz;
    const bool active = j >= kHalo && j < s.ny - kHalo && k >= kHalo && k < s.nz - kHalo;
    const int64_t sx = s.sx, sy = s.sy;
    const float* __restrict__ vx = s.v[0];
    const float* __restrict__ vy = s.v[1];
    const float* __restrict__ vz = s.v[2];
    const float* coefs = coef_base_3d(s, lds_coef, UNIFORM);

    const int64_t col = j * sy + k;
    float q[16];  // q[m] = vx(i - 8 + m, j, k)
#pragma unroll
    for (int m = 0; m < 16; ++m) q[m] = in_grid ? vx[(i0 - 8 + m) * sx + col] : 0.f;
    VX.stage(vx, i0 - 1, j0, k0, s);
    VY.stage(vy, i0 - 1, j0, k0, s);
    VY.stage(vy, i0, j0, k0, s);
    VZ.stage(vz, i0 - 1, j0, k0, s);
    VZ.stage(vz, i0, j0, k0, s);

    for (int64_t i = i0; i < i1; ++i) {
        if (i > i0) {
#pragma unroll
            for (int m = 0; m < 15; ++m) q[m] = q[m + 1];
            q[15] = in_grid ? vx[(i + 7) * sx + col] : 0.f;
        }
        VX.stage(vx, i, j0, k0, s);
        VY.stage(vy, i + 1, j0, k0, s);
        VZ.stage(vz, i + 1, j0, k0, s);
        __syncthreads();
        const int64_t c = i * sx + col;
        if (active) {
            const float* xm = VX.slot(i - 1);
            const float* x0 = VX.slot(i);
            const float* ym = VY.slot(i - 1);
            const float* y0 = VY.slot(i);
            const float* yp = VY.slot(i + 1);
            const float* zm = VZ.slot(i - 1);
            const float* z0 = VZ.slot(i);
            const float* zp = VZ.slot(i + 1);
            auto X = [&](const float* pl, int dy, int dz) { return pl[Tile<TY, 1, 1>::at(ty, tz, dy, dz)]; };
            auto Y = [&](const float* pl, int dy, int dz) { return pl[Tile<TY, kHalo, 1>::at(ty, tz, dy, dz)]; };
            auto Z = [&](const float* pl, int dy, int dz) { return pl[Tile<TY, 1, kHalo>::at(ty, tz, dy, dz)]; };
            float div = 0.f;
            {
                const float* w = coef3<UNIFORM>(coefs, s.dcmap, c, 0);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (q[7 + m] - q[8 - m]);
                float tr = 0.f;
                tr += (X(x0, 1, 0) - X(xm, 1, 0)) + (X(x0, -1, 0) - X(xm, -1, 0));
                tr += (X(x0, 0, 1) - X(xm, 0, 1)) + (X(x0, 0, -1) - X(xm, 0, -1));
                const float D = acc + w[8] * tr;
                const float qq1 = s.bp[0][0][c] * s.phi[0][0][c] + s.ap[0][0][c] * D;
                const float qq2 = s.bp[0][1][c] * s.phi[0][1][c] + s.ap[0][1][c] * D;
                s.phi[0][0][c] = qq1;
                s.phi[0][1][c] = qq2;
                div += D * s.ikp[0][c] + (qq1 + qq2);
            }
            {
                const float* w = coef3<UNIFORM>(coefs, s.dcmap, c, 1);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (Y(y0, m - 1, 0) - Y(y0, -m, 0));
                float tr = 0.f;
                tr += (Y(yp, 0, 0) - Y(yp, -1, 0)) + (Y(ym, 0, 0) - Y(ym, -1, 0));
                tr += (Y(y0, 0, 1) - Y(y0, -1, 1)) + (Y(y0, 0, -1) - Y(y0, -1, -1));
                const float D = acc + w[8] * tr;
                const int ax = ISO ? 0 : 1;
                const float qq1 = s.bp[ax][0][c] * s.phi[1][0][c] + s.ap[ax][0][c] * D;
                const float qq2 = s.bp[ax][1][c] * s.phi[1][1][c] + s.ap[ax][1][c] * D;
                s.phi[1][0][c] = qq1;
                s.phi[1][1][c] = qq2;
                div += D * s.ikp[ax][c] + (qq1 + qq2);
            }
            {
                const float* w = coef3<UNIFORM>(coefs, s.dcmap, c, 2);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (Z(z0, 0, m - 1) - Z(z0, 0, -m));
                float tr = 0.f;
                tr += (Z(zp, 0, 0) - Z(zp, 0, -1)) + (Z(zm, 0, 0) - Z(zm, 0, -1));
                tr += (Z(z0, 1, 0) - Z(z0, 1, -1)) + (Z(z0, -1, 0) - Z(z0, -1, -1));
                const float D = acc + w[8] * tr;
                const int ax = ISO ? 0 : 2;
                const float qq1 = s.bp[ax][0][c] * s.phi[2][0][c] + s.ap[ax][0][c] * D;
                const float qq2 = s.bp[ax][1][c] * s.phi[2][1][c] + s.ap[ax][1][c] * D;
                s.phi[2][0][c] = qq1;
                s.phi[2][1][c] = qq2;
                div += D * s.ikp[ax][c] + (qq1 + qq2);
            }
            s.p[c] -= s.dt * s.K[c] * div;
        }
    }
}

dim3 grid_3d(const DevSim& s, int ty, int chunk) {
    return dim3((unsigned)((s.nz + TZ - 1) / TZ), (unsigned)((s.ny + ty - 1) / ty),
                (unsigned)((s.nx - 2 * kHalo + chunk - 1) / chunk * s.nshots));
}

}  // namespace

static int coef_lds_bytes_3d(const DevSim& s) {
    const int nsets = s.ndmap == 1 ? 1 : s.ndmap;
    return nsets > kLdsSets3D ? 0 : nsets * 3 * kNCoef * 4;
}

template <int TY>
void launch_3d(bool velocity, const DevSim& s, int lds, int chunk, hipStream_t st) {
    const bool u = s.ndmap == 1;
    const dim3 g = grid_3d(s, TY, chunk), b(TZ * TY);
    if (velocity) {
        if (u && s.isotropic) hipLaunchKernelGGL((k_velocity_3d<TY, true, true>), g, b, lds, st, s, chunk);
        else if (u) hipLaunchKernelGGL((k_velocity_3d<TY, true, false>), g, b, lds, st, s, chunk);
        else if (s.isotropic) hipLaunchKernelGGL((k_velocity_3d<TY, false, true>), g, b, lds, st, s, chunk);
        else hipLaunchKernelGGL((k_velocity_3d<TY, false, false>), g, b, lds, st, s, chunk);
    } else {
        if (u && s.isotropic) hipLaunchKernelGGL((k_pressure_3d<TY, true, true>), g, b, lds, st, s, chunk);
        else if (u) hipLaunchKernelGGL((k_pressure_3d<TY, true, false>), g, b, lds, st, s, chunk);
        else if (s.isotropic) hipLaunchKernelGGL((k_pressure_3d<TY, false, true>), g, b, lds, st, s, chunk);
        else hipLaunchKernelGGL((k_pressure_3d<TY, false, false>), g, b, lds, st, s, chunk);
    }
}

// cfg.width is the workgroup size: 512 (tiles of 8 rows in y) or 1024 (16 rows)
hipError_t launch_velocity_3d(const DevSim& s, const LaunchCfg& cfg, hipStream_t st) {
    const int chunk = cfg.chunk < 1 ? 1 : cfg.chunk;
    if (cfg.width == 1024) launch_3d<16>(true, s, coef_lds_bytes_3d(s), chunk, st);
    else launch_3d<8>(true, s, coef_lds_bytes_3d(s), chunk, st);
    return hipGetLastError();
}

hipError_t launch_pressure_3d(const DevSim& s, const LaunchCfg& cfg, hipStream_t st) {
    const int chunk = cfg.chunk < 1 ? 1 : cfg.chunk;
    if (cfg.width == 1024) launch_3d<16>(false, s, coef_lds_bytes_3d(s), chunk, st);
    else launch_3d<8>(false, s, coef_lds_bytes_3d(s), chunk, st);
    return hipGetLastError();
}

}  // namespace fw
