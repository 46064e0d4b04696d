// Relaxation coefficient codes (isotropic relaxation model).
//
// The ten per-cell relaxation maps (1/kappa, a1, a2, b1, b2 on the velocity side and on the
// pressure side) are built from a lookup table of attenuation pairs and a PML ramp, and the
// coefficient-set map (dcmap) from a few sound speeds, so a grid holds few distinct tuples of
// these eleven values where it matters (the interior and the PML faces), while the maps cost
// 44 bytes per cell in every step. The encoder gives the 65535 most frequent tuples a 16-bit
// code and builds two tables of codes x 8 words {1/kappa, a1, a2, b1, b2, dcmap, 0, 0}
// (velocity side, pressure side). The marching kernels then read 2 bytes per cell and kernel
// and gather the row from a table that stays in the caches; cells of the remaining tuples
// (PML edges and corners of 3D grids) carry kCodeEscape and read the per-cell maps as before.
//
// Encoding runs once per medium, on the device:
//   1. insert: every cell hashes its tuple (the raw bits of the eleven words) into an
//      open-addressing table of 64-bit keys (linear probing, bounded probe count), counts it
//      and keeps the lowest cell holding it as the slot's representative (an atomic minimum;
//      for grids up to 2^24 cells a table of 2^16 slots first, 2^22 if that one overflows);
//   2. the host picks the most frequent occupied slots (n <= 65535) and numbers them in the
//      order of their representative cells: tuples that first occur close together get
//      neighbouring codes, so the cells of a row that hold different tuples (the ramp of the
//      absorbing layer along the fastest axis) read rows that share cache lines of the tables;
//   3. assign: every cell finds its key again and stores the slot's code (or the escape);
//   4. table: each code copies its representative cell's eleven words into the tables;
//   5. verify: every coded cell compares its eleven words, bit for bit, with its code's row.
// A failure (probe limit, a hash collision caught by the check) leaves the handle on the
// per-cell maps, so a coded run reads exactly the values of the maps.
#include "fw_internal.h"

namespace fw {

namespace {

__device__ inline uint64_t fmix64(uint64_t h) {  // MurmurHash3 finaliser
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

__device__ inline void tuple_of(const CodeMaps& m, int64_t c, uint32_t t[kCodeVals]) {
#pragma unroll
    for (int k = 0; k < kCodeVals - 1; ++k) t[k] = __float_as_uint(m.f[k][c]);
    t[kCodeVals - 1] = (uint32_t)m.dcmap[c];
}

__device__ inline uint64_t tuple_hash(const uint32_t t[kCodeVals]) {
    uint64_t h = 0x9e3779b97f4a7c15ull;
#pragma unroll
    for (int k = 0; k < kCodeVals; ++k) h = fmix64(h ^ ((uint64_t)t[k] + 0x632be59bd9b4e019ull * (uint64_t)(k + 1)));
    return h == 0 ? 1 : h;  // 0 marks an empty slot
}

// Counts and lowest cells gather first in a small direct-mapped table per workgroup in LDS and
// reach the global table once per workgroup and slot: the few tuples of a grid's interior would
// otherwise take an atomic per wavefront on a handful of addresses, which serialise (3D
// inclusion grid, 654 M cells: 224 -> 72 ms). A slot whose LDS entry another slot holds goes to
// the global table directly.
constexpr int kInsertThreads = 256;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

__global__ __launch_bounds__(kInsertThreads) void k_code_insert(CodeMaps m, int64_t n, unsigned long long* keys,
                                                                uint32_t mask, uint32_t* count, int32_t* rep,
                                                                int32_t* status) {
    __shared__ uint32_t lslot[kInsertThreads];
    __shared__ uint32_t lcount[kInsertThreads];
    __shared__ int32_t lmin[kInsertThreads];
    lslot[threadIdx.x] = kNoSlot;
    lcount[threadIdx.x] = 0;
    lmin[threadIdx.x] = 0x7fffffff;
    __syncthreads();
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        uint32_t t[kCodeVals];
        tuple_of(m, c, t);
        const unsigned long long h = tuple_hash(t);
        // lanes holding the first active lane's tuple are counted by one of them
        const uint32_t lo = (uint32_t)h, hi = (uint32_t)(h >> 32);
        const bool same = lo == (uint32_t)__builtin_amdgcn_readfirstlane((int)lo) &&
                          hi == (uint32_t)__builtin_amdgcn_readfirstlane((int)hi);
        const unsigned long long group = __ballot(same);
        const bool leader = same && (int)__lane_id() == __ffsll((long long)group) - 1;
        if (same && !leader) continue;
        uint32_t slot = (uint32_t)h & mask;
        bool placed = false;
        for (int probe = 0; probe < kCodeMaxProbe; ++probe) {
            unsigned long long k = __hip_atomic_load(&keys[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (k == 0) {
                k = atomicCAS(&keys[slot], 0ull, h);
                if (k == 0) k = h;
            }
            if (k == h) {
                placed = true;
                break;
            }
            slot = (slot + 1) & mask;
        }
        if (placed) {
            const uint32_t add = same ? (uint32_t)__popcll(group) : 1u;  // the leader is the group's lowest cell
            const uint32_t e = slot & (kInsertThreads - 1);
            const uint32_t held = atomicCAS(&lslot[e], kNoSlot, slot);
            if (held == kNoSlot || held == slot) {
                atomicAdd(&lcount[e], add);
                atomicMin(&lmin[e], (int32_t)c);
            } else {
                atomicAdd(&count[slot], add);
                atomicMin(&rep[slot], (int32_t)c);
            }
        } else {
            status[0] = 1;  // every writer stores the same value
        }
    }
    __syncthreads();
    const uint32_t held = lslot[threadIdx.x];
    if (held != kNoSlot) {
        atomicAdd(&count[held], lcount[threadIdx.x]);
        atomicMin(&rep[held], lmin[threadIdx.x]);
    }
}

__global__ __launch_bounds__(256) void k_code_assign(CodeMaps m, int64_t n, const unsigned long long* keys,
                                                     uint32_t mask, const uint16_t* slot_code, uint16_t* code,
                                                     int32_t* status) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        uint32_t t[kCodeVals];
        tuple_of(m, c, t);
        const unsigned long long h = tuple_hash(t);
        uint32_t slot = (uint32_t)h & mask;
        bool found = false;
        for (int probe = 0; probe < kCodeMaxProbe; ++probe) {
            const unsigned long long k = keys[slot];
            if (k == h) {
                found = true;
                break;
            }
            if (k == 0) break;
            slot = (slot + 1) & mask;
        }
        code[c] = found ? slot_code[slot] : (uint16_t)kCodeEscape;
        if (!found) status[1] = 1;
    }
}

__global__ __launch_bounds__(256) void k_code_table(CodeMaps m, int32_t ncodes, const int32_t* rep_of_code,
                                                    float* tab_v, float* tab_p) {
    const int32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ncodes) return;
    const int64_t c = rep_of_code[id];
    float* rv = tab_v + (int64_t)id * kCodeStride;
    float* rp = tab_p + (int64_t)id * kCodeStride;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        rv[k] = m.f[k][c];
        rp[k] = m.f[5 + k][c];
    }
    rv[5] = rp[5] = __int_as_float(m.dcmap[c]);
    rv[6] = rv[7] = rp[6] = rp[7] = 0.f;
}

__global__ __launch_bounds__(256) void k_code_verify(CodeMaps m, int64_t n, const uint16_t* code, int32_t ncodes,
                                                     const float* tab_v, const float* tab_p, int32_t* status) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t id = code[c];
        if (id == kCodeEscape) continue;
        bool same = id < (uint32_t)ncodes;
        if (same) {
            const float* rv = tab_v + (int64_t)id * kCodeStride;
            const float* rp = tab_p + (int64_t)id * kCodeStride;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                same &= __float_as_uint(m.f[k][c]) == __float_as_uint(rv[k]);
                same &= __float_as_uint(m.f[5 + k][c]) == __float_as_uint(rp[k]);
            }
            same &= m.dcmap[c] == __float_as_int(rv[5]) && m.dcmap[c] == __float_as_int(rp[5]);
        }
        if (!same) status[2] = 1;
    }
}

dim3 code_grid(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 256 * 64) b = 256 * 64;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

}  // namespace

// nslots: a power of two
hipError_t launch_code_insert(const CodeMaps& m, int64_t n, unsigned long long* keys, uint32_t nslots, uint32_t* count,
                              int32_t* rep, int32_t* status, hipStream_t st) {
    hipLaunchKernelGGL(k_code_insert, code_grid(n), dim3(kInsertThreads), 0, st, m, n, keys, nslots - 1, count, rep, status);
    return hipGetLastError();
}

hipError_t launch_code_assign(const CodeMaps& m, int64_t n, const unsigned long long* keys, uint32_t nslots,
                              const uint16_t* slot_code, uint16_t* code, int32_t* status, hipStream_t st) {
    hipLaunchKernelGGL(k_code_assign, code_grid(n), dim3(256), 0, st, m, n, keys, nslots - 1, slot_code, code, status);
    return hipGetLastError();
}

hipError_t launch_code_table(const CodeMaps& m, int32_t ncodes, const int32_t* rep_of_code, float* tab_v, float* tab_p,
                             hipStream_t st) {
    if (ncodes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_code_table, dim3((unsigned)((ncodes + 255) / 256)), dim3(256), 0, st, m, ncodes, rep_of_code,
                       tab_v, tab_p);
    return hipGetLastError();
}

hipError_t launch_code_verify(const CodeMaps& m, int64_t n, const uint16_t* code, int32_t ncodes, const float* tab_v,
                              const float* tab_p, int32_t* status, hipStream_t st) {
    hipLaunchKernelGGL(k_code_verify, code_grid(n), dim3(256), 0, st, m, n, code, ncodes, tab_v, tab_p, status);
    return hipGetLastError();
}

}  // namespace fw
