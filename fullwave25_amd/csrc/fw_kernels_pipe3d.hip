// Software-pipelined x-marching kernels for 3D grids (kernel variant 3).
//
// Same tiling and arithmetic as fw_kernels_stream3d.hip (a TY x 64 tile of (y, z) columns per
// workgroup, 64 lanes along z, a chunk of x planes marched one per iteration, the axial x
// stencil in a 16-deep register queue, the y/z stencils from planes staged in 4-slot LDS
// rings), with the memory latency taken off the critical path:
//   * the plane that enters the LDS ring at iteration i is loaded into registers at
//     iteration i - 1 and only written to LDS at iteration i (each lane owns fixed elements of
//     the staged tile; their in-plane offsets are computed once);
//   * every per-cell operand of plane i + 1 (the next queue element, rho, the memory
//     variables, a/b, 1/kappa, the velocities or pressure, the coefficient-set index) is loaded
//     at iteration i, before the barrier, and consumed one iteration later.
// One barrier per plane. A 4-slot ring is enough: at iteration i the lanes write slot i + 1
// (mod 4) while lanes still in iteration i - 1 read slots i - 2 .. i.
// With coded relaxation coefficients, a cell's code is loaded one iteration before the iteration
// that gathers its table row: memory operations complete in issue order for the wait counters,
// so a row gather issued right after its code load waits for every load issued before the code,
// and each plane took two memory round trips (measured: pressure kernel 9.99 -> 9.51 ms, velocity
// 10.95 -> 10.80 ms on the 904 x 904 x 776 grid, with the row read as two vector loads).
#include "fw_internal.h"

namespace fw {

namespace {

constexpr int PZ3 = 64;  // lanes along z
constexpr int kRingP = 4;
constexpr int kLdsSetsP = 128;
// grids whose axes share their weights (SAME kernels): one set per sound speed, staged in LDS
// up to this many sound speeds (9 KB)
constexpr int kLdsSetsSame = 256;

// global memory at 32-bit element indices (gld32/gst32 in fw_internal.h)
template <class T>
__device__ __forceinline__ T gld(const T* __restrict__ base, uint32_t i) { return gld32(base, i); }
template <class T>
__device__ __forceinline__ void gst(T* base, uint32_t i, T v) { gst32(base, i, v); }

// The staged tile: rows y in [j0 - HY, j0 + TY + HY), columns z in [k0 - HZ, k0 + 64 + HZ).
template <int TY, int HY, int HZ>
struct PTile {
    static constexpr int PY = TY + 2 * HY;
    static constexpr int PZ = PZ3 + 2 * HZ;
    static constexpr int N = PY * PZ;
    static constexpr int NT = PZ3 * TY;
    static constexpr int R = (N + NT - 1) / NT;  // elements per lane
    float v[kRingP][N];
    __device__ float* slot(int64_t r) { return v[r & (kRingP - 1)]; }
    __device__ static int at(int ty, int tz, int dy, int dz) { return (ty + HY + dy) * PZ + (tz + HZ + dz); }
};

// A lane's share of one staged plane: fixed offsets within a plane, values in flight.
template <int TY, int HY, int HZ>
struct Stager {
    using T = PTile<TY, HY, HZ>;
    int32_t off[T::R];  // in-plane offset, or -1 outside the grid / past the tile
    float val[T::R];
    __device__ void init(const DevSim& s, int64_t j0, int64_t k0) {
#pragma unroll
        for (int r = 0; r < T::R; ++r) {
            const int e = threadIdx.x + r * T::NT;
            off[r] = -1;
            if (e < T::N) {
                const int yy = e / T::PZ, zz = e - (e / T::PZ) * T::PZ;
                const int64_t jj = j0 - HY + yy, kk = k0 - HZ + zz;
                if (jj >= 0 && jj < s.ny && kk >= 0 && kk < s.nz) off[r] = (int32_t)(jj * s.sy + kk);
            }
        }
    }
    __device__ void load(const float* __restrict__ f, int64_t plane, const DevSim& s) {
        const float* base = f + plane * s.sx;
#pragma unroll
        for (int r = 0; r < T::R; ++r) val[r] = off[r] >= 0 ? gld(base, (uint32_t)off[r]) : 0.f;
    }
    __device__ void store(T& tile, int64_t plane) const {
        float* dst = tile.slot(plane);
#pragma unroll
        for (int r = 0; r < T::R; ++r) {
            const int e = threadIdx.x + r * T::NT;
            if (r < T::R - 1 || e < T::N) dst[e] = val[r];
        }
    }
};

// As Stager, for a field whose interior values (one per lane, the lane's own column) are
// already in the lanes' registers (the axial register queue): the lanes load only the halo ring
// of the staged plane, and each lane writes its own interior element from a register. Saves
// the interior's second read of every plane (a quarter to a third of the staging loads).
template <int TY, int HY, int HZ>
struct HaloStager {
    using T = PTile<TY, HY, HZ>;
    static constexpr int NH = T::N - TY * PZ3;          // halo elements of a plane
    static constexpr int R = (NH + T::NT - 1) / T::NT;  // halo elements per lane
    int32_t off[R];  // in-plane offset, or -1 outside the grid / past the halo
    int32_t idx[R];  // element of the staged plane, or -1 past the halo
    float val[R];
    // the h-th halo element: HY whole rows above the interior rows, the 2 HZ side cells of
    // each interior row, HY whole rows below
    __device__ static int halo_element(int h) {
        constexpr int top = HY * T::PZ, side = TY * 2 * HZ;
        if (h < top) return h;
        h -= top;
        if (h < side) {
            const int zz = h % (2 * HZ);
            return (HY + h / (2 * HZ)) * T::PZ + (zz < HZ ? zz : zz + PZ3);
        }
        return (HY + TY) * T::PZ + (h - side);
    }
    __device__ void init(const DevSim& s, int64_t j0, int64_t k0) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int h = threadIdx.x + r * T::NT;
            off[r] = -1;
            idx[r] = -1;
            if (h < NH) {
                const int e = halo_element(h);
                idx[r] = e;
                const int yy = e / T::PZ, zz = e - (e / T::PZ) * T::PZ;
                const int64_t jj = j0 - HY + yy, kk = k0 - HZ + zz;
                if (jj >= 0 && jj < s.ny && kk >= 0 && kk < s.nz) off[r] = (int32_t)(jj * s.sy + kk);
            }
        }
    }
    __device__ void load(const float* __restrict__ f, int64_t plane, const DevSim& s) {
        const float* base = f + plane * s.sx;
#pragma unroll
        for (int r = 0; r < R; ++r) val[r] = off[r] >= 0 ? gld(base, (uint32_t)off[r]) : 0.f;
    }
    // own: this lane's interior value of the plane (0 outside the grid)
    __device__ void store(T& tile, int64_t plane, float own, int ty, int tz) const {
        float* dst = tile.slot(plane);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (idx[r] >= 0) dst[idx[r]] = val[r];
        dst[T::at(ty, tz, 0, 0)] = own;
    }
};

// SAME kernels (every axis has the same weights): each workgroup stages one set per sound speed
// in LDS, [set][9], and a cell reads its weights there (LDS reads instead of three vector-memory
// loads a cell from the global table; neighbouring cells mostly share a set, so the reads are
// mostly broadcasts).
template <bool UNIFORM>
__device__ inline void stage_same_sets(const DevSim& s, float* lds) {
    const int nsets = UNIFORM ? 1 : s.ndmap;
    for (int t = threadIdx.x; t < nsets * kNCoef; t += blockDim.x)
        lds[t] = s.coef[(t / kNCoef) * 3 * kNCoef + t % kNCoef];
}
__device__ inline void same_weights(float (&w)[kNCoef], const float* lds, int32_t dc) {
#pragma unroll
    for (int m = 0; m < kNCoef; ++m) w[m] = lds[dc * kNCoef + m];
}

// the code of cell c (coded grids, active cells), else 0
template <bool CODED>
__device__ inline uint32_t load_code(const DevSim& s, uint32_t c, bool active) {
    return CODED && active ? gld16(s.rcode, c) : 0u;
}

// The weights of axis `axis` for coefficient set dc. SAME (equal spacings: every axis has the
// same weights, DevSim::coef_axes_same): axis 0's, so a cell reads its 9 weights once instead of
// three times (27 loads a cell on grids whose sets are read from the global table).
template <bool UNIFORM, bool SAME>
__device__ inline const float* coefp(const float* coefs, int32_t dc, int axis) {
    const int a = SAME ? 0 : axis;
    if (UNIFORM) return coefs + a * kNCoef;
    return coefs + ((int64_t)dc * 3 + a) * kNCoef;
}

__device__ inline const float* coef_table(const DevSim& s, float* lds_coef, bool uniform) {
    const int nsets = uniform ? 1 : s.ndmap;
    if (nsets > kLdsSetsP) return s.coef;
    for (int t = threadIdx.x; t < nsets * 3 * kNCoef; t += blockDim.x) lds_coef[t] = s.coef[t];
    return lds_coef;
}

// per-cell operands of one plane for the velocity update
template <bool ISO>
struct VelOps {
    static constexpr int NA = ISO ? 1 : 3;
    float qnew;            // p(i + 8): the queue's next element
    float rho_x;           // rho at c + e_x: the next plane's rho_i
    float rho_y, rho_z;    // rho at c + e_y, c + e_z (from the LDS row, after the barrier)
    float ryn, rzn;        // the next plane's rho beyond the tile (RhoEdge), edge lanes only
    float psi[3][2];
    float v[3];
    float a[NA][2], b[NA][2];
    float ik[NA];
    int32_t dc;
    bool esc;  // coded grid, cell without a code: the coefficients come from the maps (fix_escape)
};

// The coefficients of a cell without a code (o.esc) from the per-cell maps (coded grids).
template <bool VEL, bool UNIFORM, class Ops>
__device__ inline void fix_escape(Ops& o, const DevSim& s, uint32_t c) {
    if (!o.esc) return;
    const CodedCoef k = escape_coef<VEL, UNIFORM>(s, c);
    o.a[0][0] = k.a1;
    o.a[0][1] = k.a2;
    o.b[0][0] = k.b1;
    o.b[0][1] = k.b2;
    o.ik[0] = k.ik;
    o.dc = UNIFORM ? 0 : k.dc;
}

// CODED (isotropic only): the relaxation coefficients come from row `code` of the code table
// (rtab_v), `code` being the cell's code (rcode), loaded beforehand. The row is read as one 16-byte
// and one 8-byte load by every lane (a cell without a code reads row 0 and is fixed up after the
// barrier by fix_escape): with the rare map reads in a branch of their own, the compiler does not
// turn the gather into six lane-selected loads.
template <bool UNIFORM, bool ISO, bool CODED>
__device__ inline void load_vel(VelOps<ISO>& o, const DevSim& s, uint32_t c, bool in_grid, bool active,
                                uint32_t code) {
    const uint32_t sx = (uint32_t)s.sx;
    o.qnew = in_grid ? gld(s.p, c + 8 * sx) : 0.f;
    o.rho_x = in_grid ? gld(s.rho, c + sx) : 0.f;
    if (!active) return;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        o.v[a] = gld((const float*)s.v[a], c);
        o.psi[a][0] = gldnt(s.psi[a][0], c);
        o.psi[a][1] = gldnt(s.psi[a][1], c);
    }
    if (CODED) {
        o.esc = code == kCodeEscape;
        const CodedCoef k = coded_coef(s.rtab_v, o.esc ? 0u : code);
        o.a[0][0] = k.a1;
        o.a[0][1] = k.a2;
        o.b[0][0] = k.b1;
        o.b[0][1] = k.b2;
        o.ik[0] = k.ik;
        o.dc = UNIFORM ? 0 : k.dc;
    } else {
        o.dc = UNIFORM ? 0 : gld(s.dcmap, c);
#pragma unroll
        for (int a = 0; a < VelOps<ISO>::NA; ++a) {
            o.a[a][0] = gld(s.av[a][0], c);
            o.a[a][1] = gld(s.av[a][1], c);
            o.b[a][0] = gld(s.bv[a][0], c);
            o.b[a][1] = gld(s.bv[a][1], c);
            o.ik[a] = gld(s.ikv[a], c);
        }
    }
}

// The rho neighbours of the velocity update (c + e_y, c + e_z) come from a plane of rho in LDS
// that every lane fills with its own rho_i; the row below the tile and the column right of it
// are loaded by a few lanes only: the lanes of the tile's last row (the row below, one load in
// one wavefront) and lanes 0 .. TY - 1 of the first row (one cell of the right column each).
// Two loads a workgroup and plane instead of two a cell. In-plane offsets:
struct RhoEdge {
    uint32_t row, colz;       // (j + 1, k) for row lanes; (j0 + tz, k0 + 64) for column lanes
    bool row_lane, col_lane;
};
// the next plane's edge values of a lane (plane base `base`)
template <class Ops>
__device__ inline void load_rho_edge(Ops& o, const DevSim& s, uint32_t base, const RhoEdge& e) {
    o.ryn = e.row_lane ? gld(s.rho, base + e.row) : 0.f;
    o.rzn = e.col_lane ? gld(s.rho, base + e.colz) : 0.f;
}

// AHEAD: per-cell operands loaded one iteration ahead (more registers); otherwise at the top
// of their own iteration, before the barrier
template <int TY, bool UNIFORM, bool ISO, bool AHEAD, bool CODED, bool SAME = false>
__global__ __launch_bounds__(PZ3 * TY) void k_velocity_pipe3d(DevSim batch, int chunk) {
    using TP = PTile<TY, kHalo, kHalo>;
    unsigned zl;
    const DevSim s = shot_view(batch, shot_of_block(batch, zl));
    extern __shared__ float lds_coef[];
    __shared__ TP P;
    __shared__ float RH[2][TY + 1][PZ3 + 1];  // rho of plane i (slot i % 2), with the row below and column right
    const int tz = threadIdx.x % PZ3, ty = threadIdx.x / PZ3;
    const int64_t k0 = (int64_t)blockIdx.x * PZ3, j0 = (int64_t)blockIdx.y * TY;
    const int64_t k = k0 + tz, j = j0 + ty;
    const int64_t i0 = kHalo + (int64_t)zl * chunk;
    const int64_t i1 = min(i0 + chunk, s.nx - kHalo);
    const bool in_grid = j < s.ny && k < s.nz;
    const bool active = j >= kHalo && j < s.ny - kHalo && k >= kHalo && k < s.nz - kHalo;
    const uint32_t sx = (uint32_t)s.sx, col = (uint32_t)(j * s.sy + k);
    const float* coefs = SAME ? nullptr : coef_table(s, lds_coef, UNIFORM);
    if (SAME) stage_same_sets<UNIFORM>(s, lds_coef);

    HaloStager<TY, kHalo, kHalo> st;
    st.init(s, j0, k0);
    float q[16];  // q[m] = p(i - 7 + m) in the stencil; p(i - 8 + m) at the top of iteration i
#pragma unroll
    for (int m = 0; m < 15; ++m) q[m + 1] = in_grid ? gld((const float*)s.p, (uint32_t)(i0 - 7 + m) * sx + col) : 0.f;
    st.load(s.p, i0 - 1, s);
    st.store(P, i0 - 1, q[7], ty, tz);
    st.load(s.p, i0, s);
    st.store(P, i0, q[8], ty, tz);
    st.load(s.p, i0 + 1, s);
    float rho_i = in_grid ? gld(s.rho, (uint32_t)i0 * sx + col) : 0.f;
    RhoEdge edge;
    edge.row_lane = ty == TY - 1 && active;
    edge.row = col + (uint32_t)s.sy;
    edge.col_lane = ty == 0 && tz < TY && j0 + tz < s.ny && k0 + PZ3 < s.nz;
    edge.colz = (uint32_t)((j0 + tz) * s.sy + k0 + PZ3);
    VelOps<ISO> cur, nxt;
    load_rho_edge(cur, s, (uint32_t)i0 * sx, edge);  // plane i0's edge values
    float ryn_i = cur.ryn, rzn_i = cur.rzn;
    // ca: the code of the plane whose operands this iteration loads (i, or i + 1 with AHEAD)
    constexpr int kAhead = AHEAD ? 1 : 0;
    if (AHEAD) {
        load_vel<UNIFORM, ISO, CODED>(cur, s, (uint32_t)i0 * sx + col, in_grid, active,
                                      load_code<CODED>(s, (uint32_t)i0 * sx + col, active));
        load_rho_edge(cur, s, (uint32_t)(i0 + 1) * sx, edge);
    }
    uint32_t ca = i0 + kAhead < i1 ? load_code<CODED>(s, (uint32_t)(i0 + kAhead) * sx + col, active) : 0u;

    for (int64_t i = i0; i < i1; ++i) {
        st.store(P, i + 1, q[9], ty, tz);
        {  // rho of plane i: own cell, the row below the tile, the column right of it
            float(*rh)[PZ3 + 1] = RH[i & 1];
            rh[ty][tz] = rho_i;
            if (edge.row_lane) rh[TY][tz] = ryn_i;
            if (edge.col_lane) rh[tz][PZ3] = rzn_i;
        }
        const bool more = i + 1 < i1;
        const uint32_t cb = i + kAhead + 1 < i1 ? load_code<CODED>(s, (uint32_t)(i + kAhead + 1) * sx + col, active) : 0u;
        if (!AHEAD) {
            load_vel<UNIFORM, ISO, CODED>(cur, s, (uint32_t)i * sx + col, in_grid, active, ca);
            load_rho_edge(cur, s, (uint32_t)(i + 1) * sx, edge);  // i + 1 <= nx - 8
        }
        if (more) {
            st.load(s.p, i + 2, s);
            if (AHEAD) {
                load_vel<UNIFORM, ISO, CODED>(nxt, s, (uint32_t)(i + 1) * sx + col, in_grid, active, ca);
                load_rho_edge(nxt, s, (uint32_t)(i + 2) * sx, edge);  // i + 2 <= nx - 8
            }
        }
        ca = cb;
#pragma unroll
        for (int m = 0; m < 15; ++m) q[m] = q[m + 1];
        q[15] = cur.qnew;
        __syncthreads();
        if (active) {
            const uint32_t c = (uint32_t)i * sx + col;
            if (CODED) fix_escape<true, UNIFORM>(cur, s, c);
            float ws[kNCoef];
            if (SAME) same_weights(ws, lds_coef, UNIFORM ? 0 : cur.dc);
            cur.rho_y = RH[i & 1][ty + 1][tz];
            cur.rho_z = RH[i & 1][ty][tz + 1];
            const float* pm = P.slot(i - 1);
            const float* p0 = P.slot(i);
            const float* pp = P.slot(i + 1);
            auto A = [&](const float* pl, int dy, int dz) { return pl[TP::at(ty, tz, dy, dz)]; };
            float D[3];
            {
                const float* w = SAME ? ws : coefp<UNIFORM, SAME>(coefs, cur.dc, 0);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (q[7 + m] - q[8 - m]);
                float tr = 0.f;
                tr += (A(pp, 1, 0) - A(p0, 1, 0)) + (A(pp, -1, 0) - A(p0, -1, 0));
                tr += (A(pp, 0, 1) - A(p0, 0, 1)) + (A(pp, 0, -1) - A(p0, 0, -1));
                D[0] = acc + w[8] * tr;
            }
            {
                const float* w = SAME ? ws : coefp<UNIFORM, SAME>(coefs, cur.dc, 1);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (A(p0, m, 0) - A(p0, -(m - 1), 0));
                float tr = 0.f;
                tr += (A(pp, 1, 0) - A(pp, 0, 0)) + (A(pm, 1, 0) - A(pm, 0, 0));
                tr += (A(p0, 1, 1) - A(p0, 0, 1)) + (A(p0, 1, -1) - A(p0, 0, -1));
                D[1] = acc + w[8] * tr;
            }
            {
                const float* w = SAME ? ws : coefp<UNIFORM, SAME>(coefs, cur.dc, 2);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (A(p0, 0, m) - A(p0, 0, -(m - 1)));
                float tr = 0.f;
                tr += (A(pp, 0, 1) - A(pp, 0, 0)) + (A(pm, 0, 1) - A(pm, 0, 0));
                tr += (A(p0, 1, 1) - A(p0, 1, 0)) + (A(p0, -1, 1) - A(p0, -1, 0));
                D[2] = acc + w[8] * tr;
            }
            const float rnb[3] = {cur.rho_x, cur.rho_y, cur.rho_z};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ax = ISO ? 0 : a;
                const float qq1 = cur.b[ax][0] * cur.psi[a][0] + cur.a[ax][0] * D[a];
                const float qq2 = cur.b[ax][1] * cur.psi[a][1] + cur.a[ax][1] * D[a];
                gstnt(s.psi[a][0], c, qq1);
                gstnt(s.psi[a][1], c, qq2);
                const float rho_bar = 0.5f * (rho_i + rnb[a]);
                gst(s.v[a], c, cur.v[a] - s.dt / rho_bar * (D[a] * cur.ik[ax] + (qq1 + qq2)));
            }
        }
        rho_i = cur.rho_x;
        ryn_i = cur.ryn;
        rzn_i = cur.rzn;
        if (AHEAD) cur = nxt;
    }
}

// per-cell operands of one plane for the pressure update
template <bool ISO>
struct PreOps {
    static constexpr int NA = ISO ? 1 : 3;
    float qnew;  // vx(i + 7)
    float phi[3][2];
    float a[NA][2], b[NA][2];
    float ik[NA];
    float K, p;
    int32_t dc;
    bool esc;  // as VelOps::esc
};


template <bool UNIFORM, bool ISO, bool CODED>
__device__ inline void load_pre(PreOps<ISO>& o, const DevSim& s, uint32_t c, bool in_grid, bool active,
                                uint32_t code) {
    o.qnew = in_grid ? gld((const float*)s.v[0], c + 7 * (uint32_t)s.sx) : 0.f;
    if (!active) return;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        o.phi[a][0] = gldnt(s.phi[a][0], c);
        o.phi[a][1] = gldnt(s.phi[a][1], c);
    }
    if (CODED) {
        o.esc = code == kCodeEscape;
        const CodedCoef k = coded_coef(s.rtab_p, o.esc ? 0u : code);
        o.a[0][0] = k.a1;
        o.a[0][1] = k.a2;
        o.b[0][0] = k.b1;
        o.b[0][1] = k.b2;
        o.ik[0] = k.ik;
        o.dc = UNIFORM ? 0 : k.dc;
    } else {
        o.dc = UNIFORM ? 0 : gld(s.dcmap, c);
#pragma unroll
        for (int a = 0; a < PreOps<ISO>::NA; ++a) {
            o.a[a][0] = gld(s.ap[a][0], c);
            o.a[a][1] = gld(s.ap[a][1], c);
            o.b[a][0] = gld(s.bp[a][0], c);
            o.b[a][1] = gld(s.bp[a][1], c);
            o.ik[a] = gld(s.ikp[a], c);
        }
    }
    o.K = gld(s.K, c);
    o.p = gld((const float*)s.p, c);
}

template <int TY, bool UNIFORM, bool ISO, bool AHEAD, bool CODED, bool SAME = false>
__global__ __launch_bounds__(PZ3 * TY) void k_pressure_pipe3d(DevSim batch, int chunk) {
    using TX_ = PTile<TY, 1, 1>;
    using TY_ = PTile<TY, kHalo, 1>;
    using TZ_ = PTile<TY, 1, kHalo>;
    unsigned zl;
    const DevSim s = shot_view(batch, shot_of_block(batch, zl));
    extern __shared__ float lds_coef[];
    __shared__ TX_ VX;  // vx planes i-1, i
    __shared__ TY_ VY;  // vy planes i-1, i, i+1
    __shared__ TZ_ VZ;  // vz planes i-1, i, i+1
    const int tz = threadIdx.x % PZ3, ty = threadIdx.x / PZ3;
    const int64_t k0 = (int64_t)blockIdx.x * PZ3, j0 = (int64_t)blockIdx.y * TY;
    const int64_t k = k0 + tz, j = j0 + ty;
    const int64_t i0 = kHalo + (int64_t)zl * chunk;
    const int64_t i1 = min(i0 + chunk, s.nx - kHalo);
    const bool in_grid = j < s.ny && k < s.nz;
    const bool active = j >= kHalo && j < s.ny - kHalo && k >= kHalo && k < s.nz - kHalo;
    const uint32_t sx = (uint32_t)s.sx, col = (uint32_t)(j * s.sy + k);
    const float* coefs = SAME ? nullptr : coef_table(s, lds_coef, UNIFORM);
    if (SAME) stage_same_sets<UNIFORM>(s, lds_coef);

    HaloStager<TY, 1, 1> sx_;
    Stager<TY, kHalo, 1> sy_;
    Stager<TY, 1, kHalo> sz_;
    sx_.init(s, j0, k0);
    sy_.init(s, j0, k0);
    sz_.init(s, j0, k0);
    float q[16];  // q[m] = vx(i - 8 + m) in the stencil; vx(i - 9 + m) at the top of iteration i
#pragma unroll
    for (int m = 0; m < 15; ++m) q[m + 1] = in_grid ? gld((const float*)s.v[0], (uint32_t)(i0 - 8 + m) * sx + col) : 0.f;
    sx_.load(s.v[0], i0 - 1, s);
    sx_.store(VX, i0 - 1, q[8], ty, tz);
    sy_.load(s.v[1], i0 - 1, s);
    sy_.store(VY, i0 - 1);
    sy_.load(s.v[1], i0, s);
    sy_.store(VY, i0);
    sz_.load(s.v[2], i0 - 1, s);
    sz_.store(VZ, i0 - 1);
    sz_.load(s.v[2], i0, s);
    sz_.store(VZ, i0);
    sx_.load(s.v[0], i0, s);
    sy_.load(s.v[1], i0 + 1, s);
    sz_.load(s.v[2], i0 + 1, s);
    PreOps<ISO> cur, nxt;
    // ca: the code of the plane whose operands this iteration loads (i, or i + 1 with AHEAD)
    constexpr int kAhead = AHEAD ? 1 : 0;
    if (AHEAD)
        load_pre<UNIFORM, ISO, CODED>(cur, s, (uint32_t)i0 * sx + col, in_grid, active,
                                      load_code<CODED>(s, (uint32_t)i0 * sx + col, active));
    uint32_t ca = i0 + kAhead < i1 ? load_code<CODED>(s, (uint32_t)(i0 + kAhead) * sx + col, active) : 0u;

    for (int64_t i = i0; i < i1; ++i) {
        sx_.store(VX, i, q[9], ty, tz);
        sy_.store(VY, i + 1);
        sz_.store(VZ, i + 1);
        const bool more = i + 1 < i1;
        const uint32_t cb = i + kAhead + 1 < i1 ? load_code<CODED>(s, (uint32_t)(i + kAhead + 1) * sx + col, active) : 0u;
        if (!AHEAD) load_pre<UNIFORM, ISO, CODED>(cur, s, (uint32_t)i * sx + col, in_grid, active, ca);
        if (more) {
            sx_.load(s.v[0], i + 1, s);
            sy_.load(s.v[1], i + 2, s);
            sz_.load(s.v[2], i + 2, s);
            if (AHEAD) load_pre<UNIFORM, ISO, CODED>(nxt, s, (uint32_t)(i + 1) * sx + col, in_grid, active, ca);
        }
        ca = cb;
#pragma unroll
        for (int m = 0; m < 15; ++m) q[m] = q[m + 1];
        q[15] = cur.qnew;
        __syncthreads();
        if (active) {
            const uint32_t c = (uint32_t)i * sx + col;
            if (CODED) fix_escape<false, UNIFORM>(cur, s, c);
            float ws[kNCoef];
            if (SAME) same_weights(ws, lds_coef, UNIFORM ? 0 : cur.dc);
            const float* xm = VX.slot(i - 1);
            const float* x0 = VX.slot(i);
            const float* ym = VY.slot(i - 1);
            const float* y0 = VY.slot(i);
            const float* yp = VY.slot(i + 1);
            const float* zm = VZ.slot(i - 1);
            const float* z0 = VZ.slot(i);
            const float* zp = VZ.slot(i + 1);
            auto X = [&](const float* pl, int dy, int dz) { return pl[TX_::at(ty, tz, dy, dz)]; };
            auto Y = [&](const float* pl, int dy, int dz) { return pl[TY_::at(ty, tz, dy, dz)]; };
            auto Z = [&](const float* pl, int dy, int dz) { return pl[TZ_::at(ty, tz, dy, dz)]; };
            float D[3];
            {
                const float* w = SAME ? ws : coefp<UNIFORM, SAME>(coefs, cur.dc, 0);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (q[7 + m] - q[8 - m]);
                float tr = 0.f;
                tr += (X(x0, 1, 0) - X(xm, 1, 0)) + (X(x0, -1, 0) - X(xm, -1, 0));
                tr += (X(x0, 0, 1) - X(xm, 0, 1)) + (X(x0, 0, -1) - X(xm, 0, -1));
                D[0] = acc + w[8] * tr;
            }
            {
                const float* w = SAME ? ws : coefp<UNIFORM, SAME>(coefs, cur.dc, 1);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (Y(y0, m - 1, 0) - Y(y0, -m, 0));
                float tr = 0.f;
                tr += (Y(yp, 0, 0) - Y(yp, -1, 0)) + (Y(ym, 0, 0) - Y(ym, -1, 0));
                tr += (Y(y0, 0, 1) - Y(y0, -1, 1)) + (Y(y0, 0, -1) - Y(y0, -1, -1));
                D[1] = acc + w[8] * tr;
            }
            {
                const float* w = SAME ? ws : coefp<UNIFORM, SAME>(coefs, cur.dc, 2);
                float acc = 0.f;
#pragma unroll
                for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (Z(z0, 0, m - 1) - Z(z0, 0, -m));
                float tr = 0.f;
                tr += (Z(zp, 0, 0) - Z(zp, 0, -1)) + (Z(zm, 0, 0) - Z(zm, 0, -1));
                tr += (Z(z0, 1, 0) - Z(z0, 1, -1)) + (Z(z0, -1, 0) - Z(z0, -1, -1));
                D[2] = acc + w[8] * tr;
            }
            float div = 0.f;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ax = ISO ? 0 : a;
                const float qq1 = cur.b[ax][0] * cur.phi[a][0] + cur.a[ax][0] * D[a];
                const float qq2 = cur.b[ax][1] * cur.phi[a][1] + cur.a[ax][1] * D[a];
                gstnt(s.phi[a][0], c, qq1);
                gstnt(s.phi[a][1], c, qq2);
                div += D[a] * cur.ik[ax] + (qq1 + qq2);
            }
            gst(s.p, c, cur.p - s.dt * cur.K * div);
        }
        if (AHEAD) cur = nxt;
    }
}

template <int TY, bool AHEAD>
void launch_pipe(bool velocity, const DevSim& s, int chunk, hipStream_t st) {
    const bool u = s.ndmap == 1;
    const int nsets = u ? 1 : s.ndmap;
    const bool same = s.coef_axes_same != 0 && nsets <= kLdsSetsSame;
    const int lds = same ? nsets * kNCoef * 4 : (nsets > kLdsSetsP ? 0 : nsets * 3 * kNCoef * 4);
    const dim3 g((unsigned)((s.nz + PZ3 - 1) / PZ3), (unsigned)((s.ny + TY - 1) / TY),
                 (unsigned)((s.nx - 2 * kHalo + chunk - 1) / chunk * s.nshots));
    const dim3 b(PZ3 * TY);
    const bool coded = s.isotropic && s.rcode;
    if (velocity) {
        if (u && coded && same)
            hipLaunchKernelGGL((k_velocity_pipe3d<TY, true, true, AHEAD, true, true>), g, b, lds, st, s, chunk);
        else if (coded && same)
            hipLaunchKernelGGL((k_velocity_pipe3d<TY, false, true, AHEAD, true, true>), g, b, lds, st, s, chunk);
        else if (u && coded) hipLaunchKernelGGL((k_velocity_pipe3d<TY, true, true, AHEAD, true>), g, b, lds, st, s, chunk);
        else if (u && s.isotropic) hipLaunchKernelGGL((k_velocity_pipe3d<TY, true, true, AHEAD, false>), g, b, lds, st, s, chunk);
        else if (u) hipLaunchKernelGGL((k_velocity_pipe3d<TY, true, false, AHEAD, false>), g, b, lds, st, s, chunk);
        else if (coded) hipLaunchKernelGGL((k_velocity_pipe3d<TY, false, true, AHEAD, true>), g, b, lds, st, s, chunk);
        else if (s.isotropic) hipLaunchKernelGGL((k_velocity_pipe3d<TY, false, true, AHEAD, false>), g, b, lds, st, s, chunk);
        else hipLaunchKernelGGL((k_velocity_pipe3d<TY, false, false, AHEAD, false>), g, b, lds, st, s, chunk);
    } else {
        if (u && coded && same)
            hipLaunchKernelGGL((k_pressure_pipe3d<TY, true, true, AHEAD, true, true>), g, b, lds, st, s, chunk);
        else if (coded && same)
            hipLaunchKernelGGL((k_pressure_pipe3d<TY, false, true, AHEAD, true, true>), g, b, lds, st, s, chunk);
        else if (u && coded) hipLaunchKernelGGL((k_pressure_pipe3d<TY, true, true, AHEAD, true>), g, b, lds, st, s, chunk);
        else if (u && s.isotropic) hipLaunchKernelGGL((k_pressure_pipe3d<TY, true, true, AHEAD, false>), g, b, lds, st, s, chunk);
        else if (u) hipLaunchKernelGGL((k_pressure_pipe3d<TY, true, false, AHEAD, false>), g, b, lds, st, s, chunk);
        else if (coded) hipLaunchKernelGGL((k_pressure_pipe3d<TY, false, true, AHEAD, true>), g, b, lds, st, s, chunk);
        else if (s.isotropic) hipLaunchKernelGGL((k_pressure_pipe3d<TY, false, true, AHEAD, false>), g, b, lds, st, s, chunk);
        else hipLaunchKernelGGL((k_pressure_pipe3d<TY, false, false, AHEAD, false>), g, b, lds, st, s, chunk);
    }
}

}  // namespace

// cfg.width is the workgroup size: 512 (tiles of 8 rows in y) or 1024 (16 rows); variant 3
// loads the per-cell operands one iteration ahead, variant 4 at the top of their iteration
hipError_t launch_pipe3d(const DevSim& s, const LaunchCfg& cfg, bool velocity, hipStream_t st) {
    const int chunk = cfg.chunk < 1 ? 1 : cfg.chunk;
    const bool ahead = cfg.variant == 3;
    if (cfg.width == 1024 && ahead) launch_pipe<16, true>(velocity, s, chunk, st);
    else if (cfg.width == 1024) launch_pipe<16, false>(velocity, s, chunk, st);
    else if (ahead) launch_pipe<8, true>(velocity, s, chunk, st);
    else launch_pipe<8, false>(velocity, s, chunk, st);
    return hipGetLastError();
}

}  // namespace fw
