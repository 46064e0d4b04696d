// Per-cell kernels: one thread per grid cell, stencil operands straight from global memory.
// They are the plain statement of the update (same arithmetic order as the streaming
// kernels in fw_kernels_stream.hip) and are used for grids the streaming
// kernels do not take, and as the second GPU implementation the tests compare against.
#include "fw_internal.h"

namespace fw {

namespace {

template <int ND>
__device__ inline bool cell_of_thread(const DevSim& s, unsigned zl, int64_t& i, int64_t& j, int64_t& k) {
    if (ND == 2) {
        k = 0;
        j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        i = (int64_t)blockIdx.y * blockDim.y + threadIdx.y;
        return i >= kHalo && i < s.nx - kHalo && j >= kHalo && j < s.ny - kHalo;
    }
    k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    j = (int64_t)blockIdx.y * blockDim.y + threadIdx.y;
    i = zl;
    return i >= kHalo && i < s.nx - kHalo && j >= kHalo && j < s.ny - kHalo && k >= kHalo &&
           k < s.nz - kHalo;
}

// TAPE: also write the adjoint tape of the step (DevSim::tape_u, tape_div)
template <int ND, bool EXPO, bool TAPE>
__global__ __launch_bounds__(256) void k_velocity_simple(DevSim batch) {
    unsigned zl;
    const DevSim s = shot_view(batch, shot_of_block(batch, zl));
    int64_t i, j, k;
    if (!cell_of_thread<ND>(s, zl, i, j, k)) return;
    const int64_t c = i * s.sx + j * s.sy + k;
    const float* cf = s.coef + (int64_t)s.dcmap[c] * ND * kNCoef;
    const int64_t stride[3] = {s.sx, s.sy, 1};
    const float* p = s.p;
    const float rho_c = s.rho[c];
    const float decay = EXPO ? s.aexp[c] : 1.f;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const float* w = cf + a * kNCoef;
        const int64_t sa = stride[a];
        float acc = 0.f;
#pragma unroll
        for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (p[c + m * sa] - p[c - (m - 1) * sa]);
        float tr = 0.f;
#pragma unroll
        for (int b = 0; b < ND; ++b) {
            if (b == a) continue;
            const int64_t sb = stride[b];
            tr += (p[c + sa + sb] - p[c + sb]) + (p[c + sa - sb] - p[c - sb]);
        }
        const float D = acc + w[8] * tr;
        const float rho_bar = 0.5f * (rho_c + s.rho[c + sa]);
        if (EXPO) {
            s.v[a][c] = decay * (s.v[a][c] - s.dt / rho_bar * D);
            if (TAPE) s.tape_u[a][c] = decay * D;
            continue;
        }
        const float q1 = s.bv[a][0][c] * s.psi[a][0][c] + s.av[a][0][c] * D;
        const float q2 = s.bv[a][1][c] * s.psi[a][1][c] + s.av[a][1][c] * D;
        s.psi[a][0][c] = q1;
        s.psi[a][1][c] = q2;
        const float u = D * s.ikv[a][c] + (q1 + q2);
        s.v[a][c] -= s.dt / rho_bar * u;
        if (TAPE) s.tape_u[a][c] = u;
    }
}

template <int ND, bool EXPO, bool TAPE>
__global__ __launch_bounds__(256) void k_pressure_simple(DevSim batch) {
    unsigned zl;
    const DevSim s = shot_view(batch, shot_of_block(batch, zl));
    int64_t i, j, k;
    if (!cell_of_thread<ND>(s, zl, i, j, k)) return;
    const int64_t c = i * s.sx + j * s.sy + k;
    const float* cf = s.coef + (int64_t)s.dcmap[c] * ND * kNCoef;
    const int64_t stride[3] = {s.sx, s.sy, 1};
    float div = 0.f;
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        const float* w = cf + a * kNCoef;
        const float* va = s.v[a];
        const int64_t sa = stride[a];
        float acc = 0.f;
#pragma unroll
        for (int m = 1; m <= 8; ++m) acc += w[m - 1] * (va[c + (m - 1) * sa] - va[c - m * sa]);
        float tr = 0.f;
#pragma unroll
        for (int b = 0; b < ND; ++b) {
            if (b == a) continue;
            const int64_t sb = stride[b];
            tr += (va[c + sb] - va[c - sa + sb]) + (va[c - sb] - va[c - sa - sb]);
        }
        const float D = acc + w[8] * tr;
        if (EXPO) {
            div += D;
            continue;
        }
        const float q1 = s.bp[a][0][c] * s.phi[a][0][c] + s.ap[a][0][c] * D;
        const float q2 = s.bp[a][1][c] * s.phi[a][1][c] + s.ap[a][1][c] * D;
        s.phi[a][0][c] = q1;
        s.phi[a][1][c] = q2;
        div += D * s.ikp[a][c] + (q1 + q2);
    }
    if (EXPO)
        s.p[c] = s.aexp[c] * (s.p[c] - s.dt * s.K[c] * div);
    else
        s.p[c] -= s.dt * s.K[c] * div;
    if (TAPE) s.tape_div[c] = div;
}

// Source injection, then pressure release, then sensor gather, for step `n`. Sources that
// sit on a pressure-release point or repeat a later source are dropped on the host, and
// sensors on such points read the value written there, so no two threads touch one address.
// The shot is blockIdx.y.
__global__ __launch_bounds__(256) void k_boundary(DevSim s, DevBoundary b, int64_t n, int64_t rec_row) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t shot = blockIdx.y;
    float* p = s.p + shot * s.shot_stride;
    const float* icmat_t = b.icmat_t + shot * b.icmat_stride;
    const bool inject = n < b.ntic;
    if (inject && t < b.ncoords) {
        const int64_t lin = b.src_lin[t];
        if (lin >= 0) p[lin] = icmat_t[n * b.ncoords + t];
    }
    if (t < b.nzero) p[b.zero_lin[t]] = 0.f;
    if (rec_row >= 0 && t < b.nout) {
        float val;
        if (b.out_air[t]) {
            val = 0.f;
        } else {
            const int32_t src = b.out_src[t];
            val = (inject && src >= 0) ? icmat_t[n * b.ncoords + src] : p[b.out_lin[t]];
        }
        b.rec[shot * b.rec_stride + rec_row * b.nout + t] = val;
    }
}

// The same pass with the step read from device memory (steps replayed from a captured graph).
__global__ __launch_bounds__(256) void k_boundary_dev(DevSim s, DevBoundary b, const int64_t* step, int64_t modt) {
    const int64_t n = *step;
    const int64_t rec_row = (modt > 0 && n % modt == 0) ? n / modt : -1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t shot = blockIdx.y;
    float* p = s.p + shot * s.shot_stride;
    const float* icmat_t = b.icmat_t + shot * b.icmat_stride;
    const bool inject = n < b.ntic;
    if (inject && t < b.ncoords) {
        const int64_t lin = b.src_lin[t];
        if (lin >= 0) p[lin] = icmat_t[n * b.ncoords + t];
    }
    if (t < b.nzero) p[b.zero_lin[t]] = 0.f;
    if (rec_row >= 0 && t < b.nout) {
        float val;
        if (b.out_air[t]) {
            val = 0.f;
        } else {
            const int32_t src = b.out_src[t];
            val = (inject && src >= 0) ? icmat_t[n * b.ncoords + src] : p[b.out_lin[t]];
        }
        b.rec[shot * b.rec_stride + rec_row * b.nout + t] = val;
    }
}

// Advance (or set, with set >= 0) the device step counter; one lane writes it.
__global__ void k_tick(int64_t* step, int64_t set) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step = set >= 0 ? set : *step + 1;
}

// Add k to a device counter; one lane writes it.
__global__ void k_advance(int64_t* counter, int64_t k) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter += k;
}

// One 16-byte element per lane: the fastest plain copy measured on the chip (a grid-stride loop
// over the same buffers reached 78 % of its rate).
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

// `nr` read streams and `nw` write streams of one float per lane at the same element index: the
// access pattern of the stencil kernels' per-cell operands without their stencils (the velocity
// kernel of a 3D grid reads 12 such streams and writes 9).
__global__ __launch_bounds__(256) void k_streams(StreamSet a, int nr, int nw, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float v[kMaxStreams];
#pragma unroll
    for (int r = 0; r < kMaxStreams; ++r) v[r] = r < nr ? a.r[r][t] : 0.f;
    float sum = 0.f;
#pragma unroll
    for (int r = 0; r < kMaxStreams; ++r) sum += v[r];
#pragma unroll
    for (int w = 0; w < kMaxStreams; ++w)
        if (w < nw) a.w[w][t] = sum * (float)(w + 1);
}

__global__ __launch_bounds__(256) void k_reciprocal(float* data, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) data[t] = data[t] != 0.f ? 1.0f / data[t] : 0.f;  // zero: row padding
}

// Source signals from the file layout [shot][ncoords][ntic] (source-major) to the device
// layout [shot][ntic][ncoords] (one coalesced row per step), in 32 x 32 tiles through LDS:
// both the reads and the writes of a tile are row-contiguous. The grid strides over the tiles
// (every lane of a workgroup runs the same number of iterations, so the barriers are uniform).
__global__ __launch_bounds__(256) void k_transpose_icmat(const float* __restrict__ in, float* __restrict__ out,
                                                         int64_t nq, int64_t nt, int64_t ntiles) {
    __shared__ float tile[32][33];
    const int64_t tq = (nq + 31) / 32, tn = (nt + 31) / 32, per = tq * tn;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // ty: 0 .. 7
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t b = t / per, r = t - b * per;
        const int64_t q0 = (r / tn) * 32, n0 = (r % tn) * 32;
        const float* src = in + b * nq * nt;
        float* dst = out + b * nq * nt;
        for (int k = ty; k < 32; k += 8) {
            const int64_t q = q0 + k, n = n0 + tx;
            if (q < nq && n < nt) tile[k][tx] = src[q * nt + n];
        }
        __syncthreads();
        for (int k = ty; k < 32; k += 8) {
            const int64_t n = n0 + k, q = q0 + tx;
            if (q < nq && n < nt) dst[n * nq + q] = tile[tx][k];
        }
        __syncthreads();
    }
}

}  // namespace

static dim3 simple_grid(const DevSim& s) {
    if (s.ndim == 2) return dim3((unsigned)((s.ny + 63) / 64), (unsigned)((s.nx + 3) / 4), (unsigned)s.nshots);
    return dim3((unsigned)((s.nz + 63) / 64), (unsigned)((s.ny + 3) / 4), (unsigned)(s.nx * s.nshots));
}

template <bool TAPE>
static void launch_simple(const DevSim& s, bool velocity, hipStream_t st) {
    const dim3 g = simple_grid(s), b(64, 4);
    if (velocity) {
        if (s.ndim == 2 && s.aexp) hipLaunchKernelGGL((k_velocity_simple<2, true, TAPE>), g, b, 0, st, s);
        else if (s.ndim == 2) hipLaunchKernelGGL((k_velocity_simple<2, false, TAPE>), g, b, 0, st, s);
        else if (s.aexp) hipLaunchKernelGGL((k_velocity_simple<3, true, TAPE>), g, b, 0, st, s);
        else hipLaunchKernelGGL((k_velocity_simple<3, false, TAPE>), g, b, 0, st, s);
    } else {
        if (s.ndim == 2 && s.aexp) hipLaunchKernelGGL((k_pressure_simple<2, true, TAPE>), g, b, 0, st, s);
        else if (s.ndim == 2) hipLaunchKernelGGL((k_pressure_simple<2, false, TAPE>), g, b, 0, st, s);
        else if (s.aexp) hipLaunchKernelGGL((k_pressure_simple<3, true, TAPE>), g, b, 0, st, s);
        else hipLaunchKernelGGL((k_pressure_simple<3, false, TAPE>), g, b, 0, st, s);
    }
}

hipError_t launch_velocity_simple(const DevSim& s, hipStream_t st) {
    if (s.tape_div) launch_simple<true>(s, true, st);
    else launch_simple<false>(s, true, st);
    return hipGetLastError();
}

hipError_t launch_pressure_simple(const DevSim& s, hipStream_t st) {
    if (s.tape_div) launch_simple<true>(s, false, st);
    else launch_simple<false>(s, false, st);
    return hipGetLastError();
}

hipError_t launch_boundary(const DevSim& s, const DevBoundary& b, int64_t step, int64_t modt,
                           hipStream_t st) {
    int64_t n = b.ncoords;
    if (b.nzero > n) n = b.nzero;
    if (b.nout > n) n = b.nout;
    if (n == 0) return hipSuccess;
    const int64_t rec_row = (modt > 0 && step % modt == 0) ? step / modt : -1;
    if (rec_row < 0 && step >= b.ntic && b.nzero == 0) return hipSuccess;
    hipLaunchKernelGGL(k_boundary, dim3((unsigned)((n + 255) / 256), (unsigned)s.nshots), dim3(256), 0, st, s, b,
                       step, rec_row);
    return hipGetLastError();
}

hipError_t launch_boundary_dev(const DevSim& s, const DevBoundary& b, const int64_t* step, int64_t modt,
                               hipStream_t st) {
    int64_t n = b.ncoords;
    if (b.nzero > n) n = b.nzero;
    if (b.nout > n) n = b.nout;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_boundary_dev, dim3((unsigned)((n + 255) / 256), (unsigned)s.nshots), dim3(256), 0, st, s, b,
                       step, modt);
    return hipGetLastError();
}

hipError_t launch_streams(const StreamSet& a, int nr, int nw, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_streams, dim3((unsigned)blocks), dim3(256), 0, st, a, nr, nw, n);
    return hipGetLastError();
}

hipError_t launch_tick(int64_t* step, int64_t set, hipStream_t st) {
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, step, set);
    return hipGetLastError();
}

hipError_t launch_advance(int64_t* counter, int64_t k, hipStream_t st) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, counter, k);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, int64_t n16, hipStream_t st) {
    if (n16 == 0) return hipSuccess;
    const int64_t blocks = (n16 + 255) / 256;
    if (blocks > 0x7fffffff) return hipErrorInvalidValue;  // 2^39 bytes and more: not a probe size
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, st, (const float4*)src, (float4*)dst, n16);
    return hipGetLastError();
}

hipError_t launch_transpose_icmat(const float* in, float* out, int64_t nshots, int64_t ncoords, int64_t ntic,
                                  hipStream_t st) {
    const int64_t ntiles = nshots * ((ncoords + 31) / 32) * ((ntic + 31) / 32);
    if (ntiles == 0) return hipSuccess;
    const int64_t blocks = ntiles < 256 * 64 ? ntiles : 256 * 64;
    hipLaunchKernelGGL(k_transpose_icmat, dim3((unsigned)blocks), dim3(256), 0, st, in, out, ncoords, ntic, ntiles);
    return hipGetLastError();
}

hipError_t launch_reciprocal(float* data, int64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_reciprocal, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, data, n);
    return hipGetLastError();
}

}  // namespace fw
