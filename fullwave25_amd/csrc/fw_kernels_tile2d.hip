// Tiled 2D kernels with 4-wide vector access (gfx950).
//
// A workgroup of 256 lanes (4 wavefronts) owns a tile of TX rows x 256 columns. It first stages
// the input field(s) of the tile plus the stencil halo into LDS with 16-byte loads, then every
// lane computes 4 consecutive cells (one float4 column quad) on TX / 4 rows. All per-cell maps
// and memory variables move as float4 (one 16-byte load or store per quad), which is what lets
// the HBM stream approach the copy-kernel rate; the stencil operands come from LDS. Rows of a
// tile are independent once the tile is staged, so each lane issues the loads of all its rows
// without waiting at barriers.
//
// Requires ny % 4 == 0 (16-byte aligned rows). Arithmetic order is the per-cell kernels'.
#include "fw_internal.h"

namespace fw {

namespace {

constexpr int kTW = 256;        // columns per tile
constexpr int kLanes = 256;     // threads per workgroup
constexpr int kQuads = kTW / 4; // column quads per tile row (64: one wavefront)

struct alignas(16) F4 {
    float x, y, z, w;
    __device__ float& operator[](int i) { return (&x)[i]; }
    __device__ float operator[](int i) const { return (&x)[i]; }
};
struct alignas(16) I4 {
    int x, y, z, w;
    __device__ int operator[](int i) const { return (&x)[i]; }
};

// Vector access at a 32-bit element index: the array base stays in scalar registers and each
// lane supplies a 32-bit byte offset (grids here have fewer than 2^30 cells per array, checked
// on the host), which saves a 64-bit address per access.
__device__ inline const char* at_(const void* base, uint32_t i) { return reinterpret_cast<const char*>(base) + (i << 2); }
__device__ inline F4 ld4(const float* base, uint32_t i) { return *reinterpret_cast<const F4*>(at_(base, i)); }
__device__ inline I4 ldi4(const int32_t* base, uint32_t i) { return *reinterpret_cast<const I4*>(at_(base, i)); }
__device__ inline float ld1(const float* base, uint32_t i) { return *reinterpret_cast<const float*>(at_(base, i)); }
// four 16-bit codes at element index i (i % 4 == 0: 8-byte aligned)
__device__ inline uint2 ldc4(const uint16_t* base, uint32_t i) {
    return *reinterpret_cast<const uint2*>(reinterpret_cast<const char*>(base) + (i << 1));
}

__device__ inline void st4(float* base, uint32_t i, const F4& v) {
    *reinterpret_cast<F4*>(reinterpret_cast<char*>(base) + (i << 2)) = v;
}
// non-temporal 16-byte access for the memory variables (read and written once per step)
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ inline F4 ld4nt(const float* base, uint32_t i) {
    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(at_(base, i)));
    return F4{v.x, v.y, v.z, v.w};
}
__device__ inline void st4nt(float* base, uint32_t i, const F4& v) {
    const f4v w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<f4v*>(reinterpret_cast<char*>(base) + (i << 2)));
}

// the relaxation coefficients and coefficient-set indices of a quad: from the code table
// (CODED: `cw`, the quad's four codes, loaded beforehand) or from the per-cell maps. Each coded
// cell reads its table row as one 16-byte and one 8-byte load (a cell without a code reads row
// 0 first); the rare cells without a code then take their maps in a branch of their own, so
// that the compiler does not turn the row reads into lane-selected 4-byte loads.
template <bool CODED, bool UNIFORM, bool VEL, int NA>
__device__ inline void quad_coef(const DevSim& s, uint32_t c, uint2 cw, F4* ik, F4* b1, F4* b2, F4* a1, F4* a2,
                                 I4& dc) {
    const float* const* ik_m = VEL ? s.ikv : s.ikp;
    const float* const (*a_m)[kNRelax] = VEL ? s.av : s.ap;
    const float* const (*b_m)[kNRelax] = VEL ? s.bv : s.bp;
    if (CODED) {
        const uint32_t cs[4] = {cw.x & 0xffffu, cw.x >> 16, cw.y & 0xffffu, cw.y >> 16};
        int d[4];
        bool esc = false;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const CodedCoef k = coded_coef(VEL ? s.rtab_v : s.rtab_p, cs[e] != kCodeEscape ? cs[e] : 0u);
            ik[0][e] = k.ik;
            a1[0][e] = k.a1;
            a2[0][e] = k.a2;
            b1[0][e] = k.b1;
            b2[0][e] = k.b2;
            d[e] = UNIFORM ? 0 : k.dc;
            esc = esc || cs[e] == kCodeEscape;
        }
        if (esc) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (cs[e] != kCodeEscape) continue;
                const CodedCoef k = mapped_coef<UNIFORM>(ik_m[0], a_m[0][0], a_m[0][1], b_m[0][0], b_m[0][1],
                                                         s.dcmap, c + e);
                ik[0][e] = k.ik;
                a1[0][e] = k.a1;
                a2[0][e] = k.a2;
                b1[0][e] = k.b1;
                b2[0][e] = k.b2;
                d[e] = UNIFORM ? 0 : k.dc;
            }
        }
        dc = I4{d[0], d[1], d[2], d[3]};
        return;
    }
    dc = UNIFORM ? I4{0, 0, 0, 0} : ldi4(s.dcmap, c);
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        ik[a] = ld4(ik_m[a], c);
        b1[a] = ld4(b_m[a][0], c);
        b2[a] = ld4(b_m[a][1], c);
        a1[a] = ld4(a_m[a][0], c);
        a2[a] = ld4(a_m[a][1], c);
    }
}

// The codes of the quad at row i, column jq of a coded grid, for rows the tile updates (i <
// nx - 8; ny % 4 == 0, so the quad lies in its row); else 0. A tile loads the codes of a lane's
// next row one row ahead: its row gathers then do not wait behind the row's other loads (the
// memory wait counter completes loads in issue order).
template <bool CODED>
__device__ inline uint2 quad_codes(const DevSim& s, int64_t i, int64_t jq) {
    if (!CODED || i >= s.nx - kHalo) return uint2{0u, 0u};
    return ldc4(s.rcode, (uint32_t)(i * s.ny + jq));
}

// LDS tile layout: each staged row of `pitch` floats is split into four planes of pitch / 4
// floats by tile column % 4, each plane holding its columns in order. The lanes of a wavefront
// own consecutive column quads, so when they all read the same element of their quads (at the
// same stencil offset) they touch consecutive words: the LDS serves these single-float reads
// without bank conflicts, where a row-major tile (lanes 4 floats apart) conflicts four ways.
//
// Stage rows [r0, r0 + nrows) x columns [j0 - hl, j0 + kTW + hr) of f into lds (tile column 0
// is j0 - hl). hl and hr are multiples of 4; ny % 4 == 0, so every quad is either entirely
// inside [0, ny) or entirely outside (then zero). Rows outside [0, nx) are zero.
__device__ inline void stage_tile(float* lds, const float* __restrict__ f, int64_t r0, int nrows, int64_t j0,
                                  int hl, int hr, int pitch, int64_t nx, int64_t ny) {
    const int qpr = (kTW + hl + hr) / 4;  // quads per row
    const int total = nrows * qpr;
    for (int t = threadIdx.x; t < total; t += kLanes) {
        const int row = t / qpr, qd = t - row * qpr;
        const int64_t r = r0 + row;
        const int64_t jj = j0 - hl + 4 * qd;
        F4 v{0.f, 0.f, 0.f, 0.f};
        if (r >= 0 && r < nx && jj >= 0 && jj < ny) v = ld4(f, (uint32_t)(r * ny + jj));
        float* dst = lds + row * pitch + qd;
        const int plane = pitch / 4;
        dst[0] = v.x;
        dst[plane] = v.y;
        dst[2 * plane] = v.z;
        dst[3 * plane] = v.w;
    }
}

// Element at row offset dr and column offset dc from tile row `row` of the lane's quad base
// `lb` (= lds + row * pitch + qd): K = hl + e + dc is the column within the quad grid (hl + e
// for the lane's own cell e); every index part is a compile-time constant after unrolling.
template <int PITCH, int HL>
__device__ inline float tl(const float* lb, int dr, int col) {
    const int K = HL + col;  // >= 0 for every offset used
    return lb[dr * PITCH + (K & 3) * (PITCH / 4) + (K >> 2)];
}

// LDS coefficient sets: [set][axis][kCS], or [set][kCS] when both axes share their sets (equal
// spacings; `axes` = 1), which halves the table and lets one more workgroup fit on a CU. Each
// set of 9 weights is padded to 12 floats, so a cell reads its weights as two 16-byte loads
// and one 4-byte load.
constexpr int kCS = 12;
template <bool UNIFORM>
__device__ inline const float* cset(const float* lds_coef, int dc, int axis, int axes) {
    return lds_coef + ((UNIFORM ? 0 : dc) * axes + (axes > 1 ? axis : 0)) * kCS;
}

struct W9 {
    float w[9];
};
__device__ inline W9 weights(const float* c) {
    const F4 a = *reinterpret_cast<const F4*>(c), b = *reinterpret_cast<const F4*>(c + 4);
    return W9{{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c[8]}};
}

// Stage the coefficient sets into LDS (all axes, or axis 0 only when they are the same).
template <bool UNIFORM>
__device__ inline int stage_coef(float* lds_coef, const DevSim& s) {
    const int axes = s.coef_axes_same ? 1 : 2;
    const int nsets = UNIFORM ? 1 : s.ndmap;
    for (int t = threadIdx.x; t < nsets * axes * kNCoef; t += kLanes) {
        const int set = t / kNCoef, k = t - set * kNCoef;  // set: (sound speed) * axes + axis
        lds_coef[set * kCS + k] = s.coef[(axes == 2 ? set : set * 2) * kNCoef + k];
    }
    return axes;
}

// velocity: p tile rows i0-7 .. i0+TX+7 (TX + 15 rows), columns j0-8 .. j0+TW+8
template <int TX>
struct VelTile {
    static constexpr int PITCH = kTW + 16;
    static constexpr int NROWS = TX + 15;
    float P[NROWS * PITCH];
};

// The velocity update of tile (bx, by) of grid s: columns [bx * 256, bx * 256 + 256), rows
// [8 + by * TX, 8 + by * TX + TX). TAPE: also the adjoint tape (DevSim::tape_u; 0 in the ghost
// columns of a quad).
template <int TX, bool UNIFORM, bool ISO, bool CODED, bool TAPE = false>
__device__ inline void velocity_tile(const DevSim& s, unsigned bx, unsigned by, VelTile<TX>& L, float* lds_coef) {
    constexpr int PITCH = VelTile<TX>::PITCH;
    constexpr int NROWS = VelTile<TX>::NROWS;
    float* P = L.P;
    const int64_t nx = s.nx, ny = s.ny;
    const int64_t j0 = (int64_t)bx * kTW;
    const int64_t i0 = kHalo + (int64_t)by * TX;
    const int qd = threadIdx.x % kQuads;
    const int ty = threadIdx.x / kQuads;
    const int64_t jq = j0 + 4 * qd;  // first column of this lane's quad
    uint2 cw_next = jq < ny ? quad_codes<CODED>(s, i0 + ty, jq) : uint2{0u, 0u};
    const int axes = stage_coef<UNIFORM>(lds_coef, s);
    stage_tile(P, s.p, i0 - 7, NROWS, j0, 8, 8, PITCH, nx, ny);
    __syncthreads();

    if (jq >= ny) return;
    const float dt = s.dt;
#pragma unroll 1
    for (int m = 0; m < TX / 4; ++m) {
        const int lr = ty + 4 * m;  // row within the tile
        const int64_t i = i0 + lr;
        if (i >= nx - kHalo) break;
        const uint32_t c = (uint32_t)(i * ny + jq);
        const uint2 cw = cw_next;
        if (m + 1 < TX / 4) cw_next = quad_codes<CODED>(s, i + 4, jq);
        const F4 rho0 = ld4(s.rho, c);
        const F4 rho1 = ld4(s.rho, c + (uint32_t)ny);
        const float rho_r = (jq + 4 < ny) ? ld1(s.rho, c + 4) : 0.f;
        I4 dc;
        F4 ik[2], b1[2], b2[2], a1[2], a2[2];
        constexpr int NA = ISO ? 1 : 2;
        quad_coef<CODED, UNIFORM, true, NA>(s, c, cw, ik, b1, b2, a1, a2, dc);
        F4 ps1[2], ps2[2], v[2];
        F4 tu[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (TAPE) tu[a] = F4{0.f, 0.f, 0.f, 0.f};
            ps1[a] = ld4nt(s.psi[a][0], c);
            ps2[a] = ld4nt(s.psi[a][1], c);
            v[a] = ld4(s.v[a], c);
        }
        // LDS row of tile row lr is lr + 7; column j is tile column j - j0 + 8
        const float* lb = P + (lr + 7) * PITCH + qd;
        auto pa = [&](int e, int dr, int dc_) { return tl<PITCH, 8>(lb, dr, e + dc_); };
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t j = jq + e;
            if (j < kHalo || j >= ny - kHalo) continue;
            {
                const W9 wv = weights(cset<UNIFORM>(lds_coef, dc[e], 0, axes));
                float acc = 0.f;
#pragma unroll
                for (int k = 1; k <= 8; ++k) acc += wv.w[k - 1] * (pa(e, k, 0) - pa(e, -(k - 1), 0));
                float tr = 0.f;
                tr += (pa(e, 1, 1) - pa(e, 0, 1)) + (pa(e, 1, -1) - pa(e, 0, -1));
                const float D = acc + wv.w[8] * tr;
                const float q1 = b1[0][e] * ps1[0][e] + a1[0][e] * D;
                const float q2 = b2[0][e] * ps2[0][e] + a2[0][e] * D;
                ps1[0][e] = q1;
                ps2[0][e] = q2;
                const float rho_bar = 0.5f * (rho0[e] + rho1[e]);
                const float u = D * ik[0][e] + (q1 + q2);
                v[0][e] -= dt / rho_bar * u;
                if (TAPE) tu[0][e] = u;
            }
            {
                constexpr int ax = ISO ? 0 : 1;
                const W9 wv = weights(cset<UNIFORM>(lds_coef, dc[e], 1, axes));
                float acc = 0.f;
#pragma unroll
                for (int k = 1; k <= 8; ++k) acc += wv.w[k - 1] * (pa(e, 0, k) - pa(e, 0, -(k - 1)));
                float tr = 0.f;
                tr += (pa(e, 1, 1) - pa(e, 1, 0)) + (pa(e, -1, 1) - pa(e, -1, 0));
                const float D = acc + wv.w[8] * tr;
                const float q1 = b1[ax][e] * ps1[1][e] + a1[ax][e] * D;
                const float q2 = b2[ax][e] * ps2[1][e] + a2[ax][e] * D;
                ps1[1][e] = q1;
                ps2[1][e] = q2;
                const float rho_right = e < 3 ? rho0[e + 1] : rho_r;
                const float rho_bar = 0.5f * (rho0[e] + rho_right);
                const float u = D * ik[ax][e] + (q1 + q2);
                v[1][e] -= dt / rho_bar * u;
                if (TAPE) tu[1][e] = u;
            }
        }
        // cells outside the active columns keep their loaded values, so whole quads store back
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            st4nt(s.psi[a][0], c, ps1[a]);
            st4nt(s.psi[a][1], c, ps2[a]);
            st4(s.v[a], c, v[a]);
            if (TAPE) st4(s.tape_u[a], c, tu[a]);
        }
    }
}

// pressure: vx tile rows i0-8 .. i0+TX+6 (TX + 15 rows), columns j0-4 .. j0+TW+4;
//           vy tile rows i0-1 .. i0+TX (TX + 2 rows), columns j0-8 .. j0+TW+8
template <int TX>
struct PreTile {
    static constexpr int XP = kTW + 8;
    static constexpr int XR = TX + 15;
    static constexpr int YP = kTW + 16;
    static constexpr int YR = TX + 2;
    float VX[XR * XP];
    float VY[YR * YP];
};

// The pressure update of tile (bx, by) of grid s (as velocity_tile; TAPE: DevSim::tape_div).
template <int TX, bool UNIFORM, bool ISO, bool CODED, bool TAPE = false>
__device__ inline void pressure_tile(const DevSim& s, unsigned bx, unsigned by, PreTile<TX>& L, float* lds_coef) {
    constexpr int XP = PreTile<TX>::XP;
    constexpr int XR = PreTile<TX>::XR;
    constexpr int YP = PreTile<TX>::YP;
    constexpr int YR = PreTile<TX>::YR;
    float* VX = L.VX;
    float* VY = L.VY;
    const int64_t nx = s.nx, ny = s.ny;
    const int64_t j0 = (int64_t)bx * kTW;
    const int64_t i0 = kHalo + (int64_t)by * TX;
    const int qd = threadIdx.x % kQuads;
    const int ty = threadIdx.x / kQuads;
    const int64_t jq = j0 + 4 * qd;
    uint2 cw_next = jq < ny ? quad_codes<CODED>(s, i0 + ty, jq) : uint2{0u, 0u};
    const int axes = stage_coef<UNIFORM>(lds_coef, s);
    stage_tile(VX, s.v[0], i0 - 8, XR, j0, 4, 4, XP, nx, ny);
    stage_tile(VY, s.v[1], i0 - 1, YR, j0, 8, 8, YP, nx, ny);
    __syncthreads();

    if (jq >= ny) return;
    const float dt = s.dt;
#pragma unroll 1
    for (int m = 0; m < TX / 4; ++m) {
        const int lr = ty + 4 * m;
        const int64_t i = i0 + lr;
        if (i >= nx - kHalo) break;
        const uint32_t c = (uint32_t)(i * ny + jq);
        const uint2 cw = cw_next;
        if (m + 1 < TX / 4) cw_next = quad_codes<CODED>(s, i + 4, jq);
        F4 pv = ld4(s.p, c);
        const F4 K = ld4(s.K, c);
        I4 dc;
        constexpr int NA = ISO ? 1 : 2;
        F4 ik[2], b1[2], b2[2], a1[2], a2[2];
        quad_coef<CODED, UNIFORM, false, NA>(s, c, cw, ik, b1, b2, a1, a2, dc);
        F4 ph1[2], ph2[2];
        F4 td = F4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            ph1[a] = ld4nt(s.phi[a][0], c);
            ph2[a] = ld4nt(s.phi[a][1], c);
        }
        const float* xb = VX + (lr + 8) * XP + qd;  // vx(i, .): tile column j - j0 + 4
        const float* yb = VY + (lr + 1) * YP + qd;  // vy(i, .): tile column j - j0 + 8
        auto xa = [&](int e, int dr, int dc_) { return tl<XP, 4>(xb, dr, e + dc_); };
        auto ya = [&](int e, int dr, int dc_) { return tl<YP, 8>(yb, dr, e + dc_); };
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t j = jq + e;
            if (j < kHalo || j >= ny - kHalo) continue;
            float div = 0.f;
            {
                const W9 wv = weights(cset<UNIFORM>(lds_coef, dc[e], 0, axes));
                float acc = 0.f;
#pragma unroll
                for (int k = 1; k <= 8; ++k) acc += wv.w[k - 1] * (xa(e, k - 1, 0) - xa(e, -k, 0));
                float tr = 0.f;
                tr += (xa(e, 0, 1) - xa(e, -1, 1)) + (xa(e, 0, -1) - xa(e, -1, -1));
                const float D = acc + wv.w[8] * tr;
                const float q1 = b1[0][e] * ph1[0][e] + a1[0][e] * D;
                const float q2 = b2[0][e] * ph2[0][e] + a2[0][e] * D;
                ph1[0][e] = q1;
                ph2[0][e] = q2;
                div += D * ik[0][e] + (q1 + q2);
            }
            {
                constexpr int ax = ISO ? 0 : 1;
                const W9 wv = weights(cset<UNIFORM>(lds_coef, dc[e], 1, axes));
                float acc = 0.f;
#pragma unroll
                for (int k = 1; k <= 8; ++k) acc += wv.w[k - 1] * (ya(e, 0, k - 1) - ya(e, 0, -k));
                float tr = 0.f;
                tr += (ya(e, 1, 0) - ya(e, 1, -1)) + (ya(e, -1, 0) - ya(e, -1, -1));
                const float D = acc + wv.w[8] * tr;
                const float q1 = b1[ax][e] * ph1[1][e] + a1[ax][e] * D;
                const float q2 = b2[ax][e] * ph2[1][e] + a2[ax][e] * D;
                ph1[1][e] = q1;
                ph2[1][e] = q2;
                div += D * ik[ax][e] + (q1 + q2);
            }
            pv[e] -= dt * K[e] * div;
            if (TAPE) td[e] = div;
        }
        st4(s.p, c, pv);
        if (TAPE) st4(s.tape_div, c, td);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            st4nt(s.phi[a][0], c, ph1[a]);
            st4nt(s.phi[a][1], c, ph2[a]);
        }
    }
}

template <int TX, bool UNIFORM, bool ISO, bool CODED, bool TAPE>
__global__ __launch_bounds__(kLanes) void k_velocity_tile(DevSim batch) {
    __shared__ __attribute__((aligned(16))) VelTile<TX> L;
    extern __shared__ __attribute__((aligned(16))) float lds_coef[];
    velocity_tile<TX, UNIFORM, ISO, CODED, TAPE>(shot_view(batch, blockIdx.z), blockIdx.x, blockIdx.y, L, lds_coef);
}

// gs.g non-null: each workgroup then also does the step's source, pressure-release and sensor
// work of its tile (tile blockIdx.y * gridDim.x + blockIdx.x of shot blockIdx.z), as the group
// kernels do
template <int TX, bool UNIFORM, bool ISO, bool CODED, bool TAPE>
__global__ __launch_bounds__(kLanes) void k_pressure_tile(DevSim batch, GroupStep gs) {
    __shared__ __attribute__((aligned(16))) PreTile<TX> L;
    extern __shared__ __attribute__((aligned(16))) float lds_coef[];
    TileWork w{};
    if (gs.g) {
        const GroupBoundary& e = gs.g[blockIdx.z];
        w = tile_work(e, blockIdx.y * gridDim.x + blockIdx.x, e.step0 + gs.it + (gs.base ? *gs.base : 0));
    }
    pressure_tile<TX, UNIFORM, ISO, CODED, TAPE>(shot_view(batch, blockIdx.z), blockIdx.x, blockIdx.y, L, lds_coef);
    if (!gs.g) return;
    __syncthreads();
    tile_boundary(w);
}

// grouped launches: the tiles of several grids (different sizes and media, isotropic
// relaxation) in one 1D launch. ALL_CODED: every grid reads coded coefficients; otherwise each
// grid takes the coded or the per-cell-map path by its own handle.
template <int TX, bool ALL_CODED>
__global__ __launch_bounds__(kLanes) void k_velocity_tile_group(const GroupTile* __restrict__ t, int n) {
    __shared__ __attribute__((aligned(16))) VelTile<TX> L;
    extern __shared__ __attribute__((aligned(16))) float lds_coef[];
    const int m = group_entry(t, n);
    const DevSim s = t[m].s;
    const unsigned local = blockIdx.x - t[m].start;
    const unsigned bx = local % t[m].gx, by = local / t[m].gx;
    if (ALL_CODED || s.rcode) velocity_tile<TX, false, true, true>(s, bx, by, L, lds_coef);
    else velocity_tile<TX, false, true, false>(s, bx, by, L, lds_coef);
}

// the pressure update of the tile, then its share of the step's boundary work (gs.g non-null)
template <int TX, bool ALL_CODED>
__global__ __launch_bounds__(kLanes) void k_pressure_tile_group(const GroupTile* __restrict__ t, int n, GroupStep gs) {
    __shared__ __attribute__((aligned(16))) PreTile<TX> L;
    extern __shared__ __attribute__((aligned(16))) float lds_coef[];
    const int m = group_entry(t, n);
    const DevSim s = t[m].s;
    const unsigned local = blockIdx.x - t[m].start;
    const unsigned bx = local % t[m].gx, by = local / t[m].gx;
    TileWork w{};
    if (gs.g) w = tile_work(gs.g[m], local, gs.g[m].step0 + gs.it + (gs.base ? *gs.base : 0));
    if (ALL_CODED || s.rcode) pressure_tile<TX, false, true, true>(s, bx, by, L, lds_coef);
    else pressure_tile<TX, false, true, false>(s, bx, by, L, lds_coef);
    if (!gs.g) return;
    __syncthreads();
    tile_boundary(w);
}

template <int TX, bool ALL_CODED>
hipError_t launch_group_tile(const GroupTile* t, int n, uint32_t blocks, int lds, bool velocity, const GroupStep& gs,
                             hipStream_t st) {
    if (velocity) hipLaunchKernelGGL((k_velocity_tile_group<TX, ALL_CODED>), dim3(blocks), dim3(kLanes), lds, st, t, n);
    else hipLaunchKernelGGL((k_pressure_tile_group<TX, ALL_CODED>), dim3(blocks), dim3(kLanes), lds, st, t, n, gs);
    return hipGetLastError();
}

template <int TX>
hipError_t launch_group_tile_any(const GroupTile* t, int n, uint32_t blocks, int lds, bool all_coded, bool velocity,
                                 const GroupStep& gs, hipStream_t st) {
    return all_coded ? launch_group_tile<TX, true>(t, n, blocks, lds, velocity, gs, st)
                     : launch_group_tile<TX, false>(t, n, blocks, lds, velocity, gs, st);
}

template <int TX, bool TAPE = false>
hipError_t launch_tile(const DevSim& s, bool velocity, const GroupStep& gs, hipStream_t st) {
    const bool u = s.ndmap == 1;
    const int lds = (u ? 1 : s.ndmap) * (s.coef_axes_same ? 1 : 2) * kCS * 4;
    const int64_t rows = s.nx - 2 * kHalo;
    const dim3 g((unsigned)((s.ny + kTW - 1) / kTW), (unsigned)((rows + TX - 1) / TX), (unsigned)s.nshots);
    const dim3 b(kLanes);
    const bool coded = s.isotropic && s.rcode;
    if (velocity) {
        if (u && coded) hipLaunchKernelGGL((k_velocity_tile<TX, true, true, true, TAPE>), g, b, lds, st, s);
        else if (u && s.isotropic) hipLaunchKernelGGL((k_velocity_tile<TX, true, true, false, TAPE>), g, b, lds, st, s);
        else if (u) hipLaunchKernelGGL((k_velocity_tile<TX, true, false, false, TAPE>), g, b, lds, st, s);
        else if (coded) hipLaunchKernelGGL((k_velocity_tile<TX, false, true, true, TAPE>), g, b, lds, st, s);
        else if (s.isotropic) hipLaunchKernelGGL((k_velocity_tile<TX, false, true, false, TAPE>), g, b, lds, st, s);
        else hipLaunchKernelGGL((k_velocity_tile<TX, false, false, false, TAPE>), g, b, lds, st, s);
    } else {
        if (u && coded) hipLaunchKernelGGL((k_pressure_tile<TX, true, true, true, TAPE>), g, b, lds, st, s, gs);
        else if (u && s.isotropic) hipLaunchKernelGGL((k_pressure_tile<TX, true, true, false, TAPE>), g, b, lds, st, s, gs);
        else if (u) hipLaunchKernelGGL((k_pressure_tile<TX, true, false, false, TAPE>), g, b, lds, st, s, gs);
        else if (coded) hipLaunchKernelGGL((k_pressure_tile<TX, false, true, true, TAPE>), g, b, lds, st, s, gs);
        else if (s.isotropic) hipLaunchKernelGGL((k_pressure_tile<TX, false, true, false, TAPE>), g, b, lds, st, s, gs);
        else hipLaunchKernelGGL((k_pressure_tile<TX, false, false, false, TAPE>), g, b, lds, st, s, gs);
    }
    return hipGetLastError();
}

}  // namespace

// The most sound speeds (coefficient sets) the tiled kernels stage: two padded sets (one per
// axis, unequal spacings) of every sound speed next to the largest tile must fit in the 160 KiB of
// LDS a workgroup can have on gfx950, for both kernels and for the group kernels (which stage the
// largest member's sets next to the same tiles).
constexpr int kMaxTileSets = 768;
constexpr size_t kLdsPerWorkgroup = size_t(160) << 10;
static_assert(sizeof(PreTile<32>) + size_t(2) * kMaxTileSets * kCS * sizeof(float) <= kLdsPerWorkgroup,
              "tiled pressure kernel: 32-row tile plus the staged sets exceeds the LDS of a workgroup");
static_assert(sizeof(VelTile<32>) + size_t(2) * kMaxTileSets * kCS * sizeof(float) <= kLdsPerWorkgroup,
              "tiled velocity kernel: 32-row tile plus the staged sets exceeds the LDS of a workgroup");

bool tile2d_supported(const DevSim& s) {
    return !s.aexp && s.ndim == 2 && s.ny % 4 == 0 && s.ndmap <= kMaxTileSets && s.nx * s.ny < ((int64_t)1 << 30);
}

unsigned group_tile_rows(int64_t nx, int rows) { return (unsigned)((nx - 2 * kHalo + rows - 1) / rows); }

hipError_t launch_group_tile2d(const GroupTile* t, int n, uint32_t blocks, int rows, int coef_sets, bool all_coded,
                               bool velocity, const GroupStep& gs, hipStream_t st) {
    if (n <= 0 || blocks == 0) return hipSuccess;
    const int lds = coef_sets * kCS * 4;  // the largest member's staged sets (stage_coef)
    switch (rows) {
        case 8: return launch_group_tile_any<8>(t, n, blocks, lds, all_coded, velocity, gs, st);
        case 32: return launch_group_tile_any<32>(t, n, blocks, lds, all_coded, velocity, gs, st);
        default: return launch_group_tile_any<16>(t, n, blocks, lds, all_coded, velocity, gs, st);
    }
}

hipError_t launch_tile2d(const DevSim& s, const LaunchCfg& cfg, bool velocity, const GroupStep& gs, hipStream_t st) {
    if (s.tape_div) return launch_tile<8, true>(s, velocity, GroupStep{}, st);  // taped steps: 8-row tiles
    switch (cfg.chunk) {
        case 8: return launch_tile<8>(s, velocity, gs, st);
        case 32: return launch_tile<32>(s, velocity, gs, st);
        default: return launch_tile<16>(s, velocity, gs, st);
    }
}

}  // namespace fw
