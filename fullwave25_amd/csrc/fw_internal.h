// Internal interface between the C-ABI host code (fw_api.cpp) and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fw {

constexpr int kHalo = 8;      // axial stencil half-width == ghost border width
constexpr int kNCoef = 9;     // w_1..w_8, e
constexpr int kMaxDim = 3;
constexpr int kNRelax = 2;

// Device view of one simulation. Fields and maps are float32 over the full grid
// [nx][ny][nz] (nz == 1 in 2D), C order. Cells with every coordinate in [8, n - 8) are
// updated; the ghost border is never written and stays zero.
struct DevSim {
    int ndim;
    int64_t nx, ny, nz;
    int64_t sx, sy;  // strides of x and y (z stride is 1)
    float dt;

    float* p;
    float* v[kMaxDim];
    float* psi[kMaxDim][kNRelax];  // velocity-side memory variables
    float* phi[kMaxDim][kNRelax];  // pressure-side memory variables

    const float* K;
    const float* rho;
    const float* ikv[kMaxDim];  // 1 / kappa, velocity side, per axis
    const float* ikp[kMaxDim];  // 1 / kappa, pressure side, per axis
    const float* av[kMaxDim][kNRelax];
    const float* bv[kMaxDim][kNRelax];
    const float* ap[kMaxDim][kNRelax];
    const float* bp[kMaxDim][kNRelax];
    const int32_t* dcmap;
    const float* coef;  // [ndmap][ndim][9], w and e already divided by the axis spacing
    int32_t ndmap;
    int32_t coef_axes_same;  // 1: every axis has the same set for each dcmap value (equal spacings)
    int32_t isotropic;
    // exponential-attenuation model: no memory variables; v and p are scaled by aexp after
    // their update (aexp == nullptr for the relaxation model)
    const float* aexp;

    // a batch of `nshots` transmits on this medium: every wave field (p, v, psi, phi) holds
    // nshots consecutive grids, shot_stride floats apart; the maps are shared
    int32_t nshots;
    int64_t shot_stride;

    // isotropic relaxation model, coded coefficients (fw_kernels_codes.hip): rcode[c] selects
    // row rcode[c] of rtab_v / rtab_p, kCodeStride floats {1/kappa, a1, a2, b1, b2, dcmap (int
    // bits), 0, 0}; rcode[c] == kCodeEscape: the cell reads the per-cell maps and dcmap. The
    // marching kernels read these instead of ikv/av/bv, ikp/ap/bp and dcmap (rcode == nullptr:
    // per-cell maps only). The per-cell maps stay valid for every other kernel.
    const uint16_t* rcode;
    const float* rtab_v;
    const float* rtab_p;

    // adjoint tape of the step (null: none). When tape_div is set, the velocity kernel also
    // writes, at every updated cell, tape_u[a][c] = D+_a p / kappa + psi1' + psi2' (exponential
    // model: A D+_a p), and the pressure kernel tape_div[c] = sum_a (D-_a v_a' / kappa + phi1' +
    // phi2') (exponential model: sum_a D-_a v_a'): what the adjoint needs of the forward state.
    // Single-transmit handles only; the per-cell and tiled 2D kernels write it (launch_velocity
    // / launch_pressure take those while a tape is set).
    float* tape_u[kMaxDim];
    float* tape_div;
};

constexpr int kCodeVals = 11;          // values per tuple: velocity side, pressure side, dcmap
constexpr int kCodeStride = 8;         // floats per table row (one 16-byte and one 8-byte load)
constexpr uint32_t kCodeEscape = 0xffff;  // the cell's tuple has no code
constexpr int kMaxCodes = 65535;       // codes 0 .. 65534
constexpr uint32_t kCodeSlotsSmall = 1u << 16;  // hash table slots, first try
constexpr uint32_t kCodeSlots = 1u << 22;       // hash table slots, when the first try overflows
constexpr int kCodeMaxProbe = 128;
static_assert(kCodeStride == 8, "coded_coef addresses rows as code << 5 bytes");

// the maps of a tuple: ikv, av1, av2, bv1, bv2, ikp, ap1, ap2, bp1, bp2 and dcmap (int32)
struct CodeMaps {
    const float* f[kCodeVals - 1];
    const int32_t* dcmap;
};

// Global memory at a 32-bit element index: the array base stays in scalar registers and each
// lane supplies a 32-bit byte offset, which the hardware adds (one VGPR per access instead of a
// 64-bit address). Only for arrays of fewer than 2^30 elements; the launchers check the grid.
template <class T>
__device__ __forceinline__ T gld32(const T* __restrict__ base, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (i << 2));
}
template <class T>
__device__ __forceinline__ void gst32(T* base, uint32_t i, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (i << 2)) = v;
}

// Non-temporal access (the streaming hint) for the memory variables, which each step reads and
// writes exactly once and whose next use is a whole step away: they then do not displace the
// fields the next kernel reads back from the caches. Measured 9 % faster on the 2D marching
// kernels and 2-4 % on the 3D ones.
__device__ __forceinline__ float gldnt(const float* __restrict__ base, uint32_t i) {
    return __builtin_nontemporal_load(reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (i << 2)));
}
__device__ __forceinline__ void gstnt(float* base, uint32_t i, float v) {
    __builtin_nontemporal_store(v, reinterpret_cast<float*>(reinterpret_cast<char*>(base) + (i << 2)));
}
template <class T>
__device__ __forceinline__ T gld16(const T* __restrict__ base, uint32_t i) {  // 2-byte elements
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (i << 1));
}

// Relaxation coefficients and coefficient-set index of one cell.
struct CodedCoef {
    float ik, a1, a2, b1, b2;
    int32_t dc;
};
// from row `code` of a code table
__device__ __forceinline__ CodedCoef coded_coef(const float* __restrict__ tab, uint32_t code) {
    const char* row = reinterpret_cast<const char*>(tab) + (code << 5);
    const float4 q = *reinterpret_cast<const float4*>(row);
    const float2 r = *reinterpret_cast<const float2*>(row + 16);
    return CodedCoef{q.x, q.y, q.z, q.w, r.x, __float_as_int(r.y)};
}
// from the per-cell maps (a cell without a code); dc only where the set index varies
template <bool UNIFORM>
__device__ __forceinline__ CodedCoef mapped_coef(const float* ik, const float* a1, const float* a2,
                                                 const float* b1, const float* b2, const int32_t* dcmap,
                                                 uint32_t c) {
    return CodedCoef{gld32(ik, c), gld32(a1, c), gld32(a2, c), gld32(b1, c), gld32(b2, c),
                     UNIFORM ? 0 : gld32(dcmap, c)};
}
// the per-cell map coefficients of cell c (a cell without a code), velocity (VEL) or pressure side
template <bool VEL, bool UNIFORM>
__device__ __forceinline__ CodedCoef escape_coef(const DevSim& s, uint32_t c) {
    if (VEL) return mapped_coef<UNIFORM>(s.ikv[0], s.av[0][0], s.av[0][1], s.bv[0][0], s.bv[0][1], s.dcmap, c);
    return mapped_coef<UNIFORM>(s.ikp[0], s.ap[0][0], s.ap[0][1], s.bp[0][0], s.bp[0][1], s.dcmap, c);
}
// the coefficients of cell c, whose code is `code`, for the velocity (VEL) or pressure update of
// a coded grid
template <bool VEL, bool UNIFORM>
__device__ __forceinline__ CodedCoef coef_of_code(const DevSim& s, uint32_t c, uint32_t code) {
    if (code != kCodeEscape) return coded_coef(VEL ? s.rtab_v : s.rtab_p, code);
    if (VEL) return mapped_coef<UNIFORM>(s.ikv[0], s.av[0][0], s.av[0][1], s.bv[0][0], s.bv[0][1], s.dcmap, c);
    return mapped_coef<UNIFORM>(s.ikp[0], s.ap[0][0], s.ap[0][1], s.bp[0][0], s.bp[0][1], s.dcmap, c);
}
// the same, reading the code first (the row gather then waits for the code)
template <bool VEL, bool UNIFORM>
__device__ __forceinline__ CodedCoef cell_coef(const DevSim& s, uint32_t c) {
    return coef_of_code<VEL, UNIFORM>(s, c, gld16(s.rcode, c));
}

// The fields of one shot of the batch.
__host__ __device__ inline DevSim shot_view(const DevSim& s, int shot) {
    DevSim t = s;
    const int64_t o = (int64_t)shot * s.shot_stride;
    t.p += o;
    for (int a = 0; a < kMaxDim; ++a) {
        t.v[a] = s.v[a] ? s.v[a] + o : nullptr;
        for (int nu = 0; nu < kNRelax; ++nu) {
            t.psi[a][nu] = s.psi[a][nu] ? s.psi[a][nu] + o : nullptr;
            t.phi[a][nu] = s.phi[a][nu] ? s.phi[a][nu] + o : nullptr;
        }
    }
    return t;
}

// Launch grids carry the shots in z: gridDim.z = (blocks per shot in z) * nshots. Returns the
// shot of this workgroup and, in `zl`, its z block within the shot.
__device__ inline int shot_of_block(const DevSim& s, unsigned& zl) {
    const unsigned zper = gridDim.z / (unsigned)s.nshots;
    const unsigned shot = blockIdx.z / zper;
    zl = blockIdx.z - shot * zper;
    return (int)shot;
}

// Largest z extent of a launch grid.
constexpr unsigned kMaxGridZ = 65535;

// Source injection, pressure release and sensor gather of one step.
struct DevBoundary {
    int64_t ncoords, nzero, nout, ntic;
    const int64_t* src_lin;    // -1 = dropped (duplicate or air point)
    const float* icmat_t;      // [ntic][ncoords]
    const int64_t* zero_lin;
    const int64_t* out_lin;
    const int32_t* out_src;    // source index at the sensor's voxel, or -1
    const int8_t* out_air;     // 1 if the sensor's voxel is a pressure-release point
    float* rec;                // [nshots][n_records][nout]
    int64_t icmat_stride;      // floats between the source matrices of consecutive shots
    int64_t rec_stride;        // floats between the records of consecutive shots
};

// Adjoint sweep of one time step (fw_kernels_adjoint.hip): the adjoint state (lp, lv, lpsi,
// lphi, laid out like the forward state), the step's tape (DevSim::tape_u / tape_div as the
// forward kernels wrote them), per-axis scratch H and G, and the gradient accumulators.
struct AdjView {
    float* lp;
    float* lv[kMaxDim];
    float* lpsi[kMaxDim][kNRelax];
    float* lphi[kMaxDim][kNRelax];
    const float* tu[kMaxDim];
    const float* tdiv;
    const uint8_t* dc8;  // coefficient-set index per device cell (the tiled 2D kernels; ndmap <= 256)
    float* H[kMaxDim];
    float* G[kMaxDim];
    float* gK;               // d/dK, summed over steps
    float* gRb[kMaxDim];     // d/d rho_bar_a, summed over steps
};

// One grid (or one shot of a batch) of a grouped launch: its workgroups are
// [start, start + gx * gy) of the launch, gx tiles across the contiguous axis.
struct GroupTile {
    DevSim s;
    uint32_t start;
    uint32_t gx;
    int32_t chunk;  // rows per workgroup for this grid
};

// The entry of a grouped launch that owns workgroup blockIdx.x: the last with start <= it
// (binary search over the table; every lane of the workgroup reads the same entries).
__device__ inline int group_entry(const GroupTile* __restrict__ t, int n) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t[mid].start <= blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The source, pressure-release and sensor points of one grid (or shot), for a pressure launch
// that does the step's boundary work (grouped launches, 2D single grids): each workgroup, once
// its tile is updated, injects the sources, zeroes the pressure-release points and records the
// sensors whose voxels its tile owns. Points are grouped by owning tile, tile t's being
// src[off[t] .. off[t + 1]), zero[off[T + 1 + t] .. off[T + 2 + t]) and
// out[off[2 T + 2 + t] .. off[2 T + 3 + t]) for T tiles; each entry carries what the pass needs
// (no index to look up), so the work after the tile is one load and one store a point. A voxel in
// the ghost border belongs to the tile that holds its row and column clamped to the updated region.
struct GroupBoundary {
    float* p;             // this shot's pressure field
    DevBoundary b;        // icmat_t and rec already offset to this shot
    int64_t modt;
    int64_t step0;        // the grid's step when the current run call began
    int32_t ntiles;       // T
    const int32_t* off;   // [3][T + 1]
    const int4* src;      // {source row, 0, voxel (int64, low and high word)}: sources not dropped
    const int64_t* zero;  // pressure-release voxels
    const int4* out;      // {sensor column, source row at the voxel or -1, voxel or -1 on a release point}
};

// What a workgroup needs for the boundary work of its tile at step n, read when it starts (so
// that the work after the tile waits for no table lookups): the tile's point ranges and the
// grid's pointers and sizes.
struct TileWork {
    float* p;
    const float* icmat;  // row n of the source signals
    float* rec;          // record row of step n (null: step n records nothing)
    const int4* src;
    const int64_t* zero;
    const int4* out;
    int32_t s0, s1, z0, z1, o0, o1;
    bool inject;
};
__device__ inline TileWork tile_work(const GroupBoundary& e, unsigned tile, int64_t n) {
    const int32_t t1 = e.ntiles + 1;
    TileWork w;
    w.p = e.p;
    w.inject = n < e.b.ntic;
    w.icmat = e.b.icmat_t + n * e.b.ncoords;
    w.rec = (e.modt > 0 && n % e.modt == 0) ? e.b.rec + (n / e.modt) * e.b.nout : nullptr;
    w.src = e.src;
    w.zero = e.zero;
    w.out = e.out;
    w.s0 = e.off[tile];
    w.s1 = e.off[tile + 1];
    w.z0 = e.off[t1 + tile];
    w.z1 = e.off[t1 + tile + 1];
    w.o0 = e.off[2 * t1 + tile];
    w.o1 = e.off[2 * t1 + tile + 1];
    return w;
}

// Step n's sources, pressure release and sensor records at the points of one tile (`w`), by the
// lanes of one workgroup after it has updated the tile's pressure (the caller synchronizes the
// workgroup first). The semantics are the plain boundary pass's (k_boundary): a source dropped
// for a pressure-release point or a later repeat is not listed, and a sensor on a source or
// release point records the value written there, so no lane reads what another lane of this pass
// writes.
__device__ inline void tile_boundary(const TileWork& w) {
    const int32_t lane = (int32_t)threadIdx.x, step = (int32_t)blockDim.x;
    if (w.inject) {
        for (int32_t k = w.s0 + lane; k < w.s1; k += step) {
            const int4 q = w.src[k];
            w.p[(int64_t)(((uint64_t)(uint32_t)q.w << 32) | (uint32_t)q.z)] = w.icmat[q.x];
        }
    }
    for (int32_t k = w.z0 + lane; k < w.z1; k += step) w.p[w.zero[k]] = 0.f;
    if (!w.rec) return;
    for (int32_t k = w.o0 + lane; k < w.o1; k += step) {
        const int4 q = w.out[k];
        const int64_t lin = (int64_t)(((uint64_t)(uint32_t)q.w << 32) | (uint32_t)q.z);
        float val = 0.f;  // a pressure-release point
        if (lin >= 0) val = (w.inject && q.y >= 0) ? w.icmat[q.y] : w.p[lin];
        w.rec[q.x] = val;
    }
}

// The boundary work of a group's pressure launch at iteration `it` of a run (null `g`: none,
// as when the kernel shapes are timed); the step of entry m is g[m].step0 + it (+ *base).
struct GroupStep {
    const GroupBoundary* g;
    const int64_t* base;
    int64_t it;
};

enum KernelKind { kVelocity = 0, kPressure = 1, kBoundary = 2, kNumKinds = 3 };

// Kernel family and shape for the two stencil kernels of a step.
//   variant 0: per-cell kernels (one thread per cell)
//   variant 1: x-marching kernels; `width` lanes per workgroup along the contiguous axis
//              (2D: 64, 128 or 256; 3D: fixed 64 x 8 tile) and `chunk` planes per workgroup
//   variant 2: 2D tiled kernels with 4-wide vector access; tiles of `chunk` rows (8, 16, 32)
//              by 256 columns (needs ny % 4 == 0)
//   variant 3: 3D x-marching kernels with software pipelining (planes and per-cell operands
//              loaded one iteration ahead); `width` = workgroup size 512 or 1024
//   variant 4: as 3, with the per-cell operands loaded at the top of their own iteration
struct LaunchCfg {
    int variant = 0;
    int width = 128;
    int chunk = 32;
};

bool streaming_supported(const DevSim& s);
bool tile2d_supported(const DevSim& s);
hipError_t launch_velocity(const DevSim& s, const LaunchCfg& cfg, hipStream_t st);
hipError_t launch_pipe3d(const DevSim& s, const LaunchCfg& cfg, bool velocity, hipStream_t st);
// gs.g non-null (2D marching and tiled kernels, untaped steps: pressure_fuses_boundary): the
// pressure launch also does the step's sources, pressure release and sensors (tile_boundary;
// gs.g holds one entry per shot, whose points are grouped by the tiles of `cfg`)
hipError_t launch_pressure(const DevSim& s, const LaunchCfg& cfg, hipStream_t st, const GroupStep& gs = GroupStep{});
bool pressure_fuses_boundary(const DevSim& s, const LaunchCfg& cfg);
hipError_t launch_boundary(const DevSim& s, const DevBoundary& b, int64_t step, int64_t modt,
                           hipStream_t st);
hipError_t launch_boundary_dev(const DevSim& s, const DevBoundary& b, const int64_t* step, int64_t modt,
                               hipStream_t st);
hipError_t launch_tick(int64_t* step, int64_t set, hipStream_t st);
hipError_t launch_advance(int64_t* counter, int64_t k, hipStream_t st);
// grouped 2D x-marching kernels (isotropic relaxation); `width` 64/128/256 lanes, `chunk` rows
unsigned group_tiles_x(int64_t ny, int width);
unsigned group_tiles_y(int64_t nx, int chunk);
// the pressure launches (velocity == false) also do each member's boundary work of the step (gs)
hipError_t launch_group_2d(const GroupTile* t, int n, uint32_t blocks, int width, int max_ndmap, bool velocity,
                           const GroupStep& gs, hipStream_t st);
// grouped tiled 2D kernels (isotropic relaxation, ny % 4 == 0): 256 columns x `rows` rows
// (8, 16 or 32) per workgroup; coef_sets: the most coefficient sets any member stages in LDS
// (ndmap, times 2 unless its axes share their sets); all_coded: every member has codes (the
// kernel then holds the coded path only, fewer registers)
unsigned group_tile_rows(int64_t nx, int rows);
hipError_t launch_group_tile2d(const GroupTile* t, int n, uint32_t blocks, int rows, int coef_sets, bool all_coded,
                               bool velocity, const GroupStep& gs, hipStream_t st);
hipError_t launch_reciprocal(float* data, int64_t n, hipStream_t st);
// [nshots][ncoords][ntic] -> [nshots][ntic][ncoords]
hipError_t launch_transpose_icmat(const float* in, float* out, int64_t nshots, int64_t ncoords, int64_t ntic,
                                  hipStream_t st);
hipError_t launch_copy(const void* src, void* dst, int64_t n16, hipStream_t st);
// streaming probe (fw_device_stream_bandwidth): up to kMaxStreams read and write arrays
constexpr int kMaxStreams = 16;
struct StreamSet {
    const float* r[kMaxStreams];
    float* w[kMaxStreams];
};
hipError_t launch_streams(const StreamSet& a, int nr, int nw, int64_t n, hipStream_t st);
hipError_t launch_code_insert(const CodeMaps& m, int64_t n, unsigned long long* keys, uint32_t nslots, uint32_t* count,
                              int32_t* rep, int32_t* status, hipStream_t st);
hipError_t launch_code_assign(const CodeMaps& m, int64_t n, const unsigned long long* keys, uint32_t nslots,
                              const uint16_t* slot_code, uint16_t* code, int32_t* status, hipStream_t st);
hipError_t launch_code_table(const CodeMaps& m, int32_t ncodes, const int32_t* rep_of_code, float* tab_v, float* tab_p,
                             hipStream_t st);
hipError_t launch_code_verify(const CodeMaps& m, int64_t n, const uint16_t* code, int32_t ncodes, const float* tab_v,
                              const float* tab_p, int32_t* status, hipStream_t st);
// The sparse part of a reverse step, in one launch: the records' weights added at the distinct
// sensor voxels (grow: the step's row of weights, or null when the step records nothing), then
// the source points (nsrc of them while the step injects, else 0) and the pressure-release
// points zeroed. vmask[t] says whether voxel t is one of those (1: pressure release, 2: a
// source point); its record is then dropped, as the zeroing would overwrite it.
struct AdjSparse {
    const int64_t* vox;
    const int32_t* first;
    const int32_t* order;
    const int8_t* vmask;
    int64_t nvox;
    const int64_t* src_lin;
    int64_t nsrc;
    const int64_t* zero_lin;
    int64_t nzero;
};
hipError_t launch_adj_sparse(float* lp, const float* grow, const AdjSparse& sp, bool injecting, hipStream_t st);
// The start of the next (earlier) step of the sweep, done by launch_adj_rest when
// adj_fuses_next: that step's sparse part (as k_adj_sparse; kind[c]: bit 0 pressure release,
// bit 1 source point, bit 2 sensor voxel, per device cell) and its pressure transposed (tdiv:
// that step's tape).
struct AdjNext {
    const float* tdiv;
    const float* grow;  // the step's record weights, or null
    int injecting;
    const uint8_t* kind;
    const int64_t* vox;  // sorted distinct sensor voxels
    const int32_t* first;
    const int32_t* order;
    int64_t nvox;
};
// One step of the sweep, after launch_adj_sparse: launch_adj_pressure (the pressure transposed),
// then launch_adj_rest (the rest of the step; with `next` when adj_fuses_next, also the start of
// the next step, whose launch_adj_sparse and launch_adj_pressure are then skipped).
bool adj_tiles(const DevSim& s);  // the grid takes the tiled 2D reverse kernels
bool adj_fuses_next(const DevSim& s);
// the set-index map the tiled kernels read (AdjView::dc8), from dcmap; only where adj_fuses_next
hipError_t launch_adj_set_index(const int32_t* dcmap, uint8_t* dc8, int64_t n, hipStream_t st);
hipError_t launch_adj_pressure(const DevSim& s, const AdjView& w, hipStream_t st);
hipError_t launch_adj_rest(const DevSim& s, const AdjView& w, const AdjNext* next, hipStream_t st);
// the forward kernels that write a tape (DevSim::tape_div set): the per-cell kernels, and the
// tiled 2D kernels (8 rows per workgroup) where tile2d_supported
bool tape_capable(const DevSim& s, const LaunchCfg& cfg);
hipError_t launch_adj_rho(const DevSim& s, const AdjView& w, float* grho, hipStream_t st);

}  // namespace fw
