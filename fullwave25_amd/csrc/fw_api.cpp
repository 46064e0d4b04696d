// C-ABI of the engine: handle management, uploads, the step loop and the directory contract.
// Declarations and the contract are in include/fullwave_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <set>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fullwave_hip.h"
#include "fw_internal.h"
#include "fw_sim.h"


namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define FW_HIP(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(FW_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

const char* kAxesV[3] = {"u", "w", "v"};  // velocity-side map suffix per axis
const char* kAxesP[3] = {"x", "y", "z"};  // pressure-side map suffix per axis

}  // namespace


namespace {

int64_t cells(const fw_desc& d) { return d.nx * d.ny * (d.ndim == 3 ? d.nz : 1); }

// Device layout: 3D grids pad each z row to a multiple of 32 floats (128 bytes), so every row
// starts on a cache-line boundary; 2D grids keep rows of ny floats.
int64_t row_pitch(const fw_desc& d) { return d.ndim == 3 ? (d.nz + 31) / 32 * 32 : d.ny; }

// Every copy and fill goes through the handle's own stream and is waited for: the handle's
// kernels run on that (non-blocking) stream, which does not order itself after work on the
// null stream, so a copy issued there could still be in flight when a kernel reads the data.

// Grids whose device rows are padded (3D) move through a device staging buffer: host <->
// staging as one contiguous copy, staging <-> the padded rows as a device-side pitched copy. A
// pitched copy straight from pageable host memory moves row by row, at a few hundred MB/s.
constexpr size_t kStageBytes = size_t(64) << 20;
constexpr size_t kFixedStageBytes = size_t(4) << 20;  // the same, for a fixed-region handle

size_t stage_cap(const fw_sim* s) { return s->fixed_base ? kFixedStageBytes : kStageBytes; }

// The next 256-byte aligned `bytes` of a fixed-region handle's region.
hipError_t fixed_carve(const fw_sim* s, void** p, size_t bytes) {
    const size_t off = (s->fixed_used + 255) / 256 * 256;
    if (off > s->fixed_size || bytes > s->fixed_size - off) return hipErrorOutOfMemory;
    *p = s->fixed_base + off;
    s->fixed_used = off + bytes;
    return hipSuccess;
}

// A temporary device buffer. On a fixed-region handle it is carved after everything else and
// handed back on release (temporaries are released in the reverse order of their creation).
struct Stage {
    void* p = nullptr;
    bool async = false;
    bool fixed = false;
    size_t mark = 0;
    hipError_t get(const fw_sim* s, size_t bytes) {
        if (s->fixed_base) {
            fixed = true;
            mark = s->fixed_used;
            return fixed_carve(s, &p, bytes);
        }
        hipError_t e = hipMallocAsync(&p, bytes, s->stream);
        if (e == hipSuccess) {
            async = true;
            return e;
        }
        (void)hipGetLastError();
        return hipMalloc(&p, bytes);
    }
    hipError_t release(const fw_sim* s) {
        void* q = p;
        p = nullptr;
        if (!q) return hipSuccess;
        if (fixed) {  // later work on the stream may still read it: wait before handing it back
            const hipError_t e = hipStreamSynchronize(s->stream);
            s->fixed_used = mark;
            return e;
        }
        if (async) return hipFreeAsync(q, s->stream);
        hipError_t e = hipStreamSynchronize(s->stream);
        hipError_t e2 = hipFree(q);
        return e != hipSuccess ? e : e2;
    }
};

hipError_t copy_sync(const fw_sim* s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(s->stream);
}

// Host-to-device copies go through pinned staging buffers: a copy from pageable memory pays a
// fixed cost per call in the runtime (about half a millisecond for a map of a few MB), and a
// handle's maps are a dozen such copies. Buffers of kPinnedBytes come from a process-wide pool
// (allocated on first use and kept, at most kPinnedKeep); a copy takes two, so the host fills
// one while the DMA drains the other. Without pinned memory: the pageable copy.
constexpr size_t kPinnedBytes = size_t(4) << 20;
constexpr size_t kPinnedKeep = 16;

std::mutex g_pinned_mu;
std::vector<void*> g_pinned_free;

void* pinned_get() {
    {
        std::lock_guard<std::mutex> lock(g_pinned_mu);
        if (!g_pinned_free.empty()) {
            void* b = g_pinned_free.back();
            g_pinned_free.pop_back();
            return b;
        }
    }
    void* b = nullptr;
    if (hipHostMalloc(&b, kPinnedBytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return b;
}

void pinned_put(void* b) {
    if (!b) return;
    {
        std::lock_guard<std::mutex> lock(g_pinned_mu);
        if (g_pinned_free.size() < kPinnedKeep) {
            g_pinned_free.push_back(b);
            return;
        }
    }
    (void)hipHostFree(b);
}

// synchronous: returns when the data is on the device
hipError_t h2d(const fw_sim* s, void* dst, const void* src, size_t bytes) {
    // a fixed-region handle takes no pinned buffers: the pageable copy
    if (s->fixed_base || bytes < (size_t(256) << 10)) return copy_sync(s, dst, src, bytes, hipMemcpyHostToDevice);
    void* buf[2] = {pinned_get(), pinned_get()};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = buf[0] && buf[1] ? hipSuccess : hipErrorOutOfMemory;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
    bool used[2] = {false, false};
    size_t off = 0;
    for (int i = 0; off < bytes && e == hipSuccess; i ^= 1) {
        const size_t n = bytes - off < kPinnedBytes ? bytes - off : kPinnedBytes;
        if (used[i]) e = hipEventSynchronize(ev[i]);  // the DMA out of buf[i] has finished
        if (e != hipSuccess) break;
        memcpy(buf[i], (const char*)src + off, n);
        e = hipMemcpyAsync((char*)dst + off, buf[i], n, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipEventRecord(ev[i], s->stream);
        used[i] = true;
        off += n;
    }
    const hipError_t e2 = hipStreamSynchronize(s->stream);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    pinned_put(buf[0]);
    pinned_put(buf[1]);
    if (!used[0] && e == hipErrorOutOfMemory) {  // no pinned memory: the pageable copy
        (void)hipGetLastError();
        return copy_sync(s, dst, src, bytes, hipMemcpyHostToDevice);
    }
    return e != hipSuccess ? e : e2;
}

// compact host grid -> padded device grid (elements of `esize` bytes)
hipError_t copy_grid_in(const fw_sim* s, void* dst, const void* src, size_t esize) {
    const size_t row = (size_t)s->row_len * esize, prow = (size_t)s->pitch * esize, rows = (size_t)s->rows;
    hipError_t e;
    if (row == prow) return h2d(s, dst, src, row * rows);
    const size_t per = stage_cap(s) / row > 0 ? stage_cap(s) / row : 1;
    Stage st;
    e = st.get(s, per * row);
    for (size_t r0 = 0; r0 < rows && e == hipSuccess; r0 += per) {
        const size_t n = per < rows - r0 ? per : rows - r0;
        e = h2d(s, st.p, (const char*)src + r0 * row, n * row);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync((char*)dst + r0 * prow, prow, st.p, row, row, n, hipMemcpyDeviceToDevice, s->stream);
    }
    const hipError_t e2 = st.release(s);
    const hipError_t e3 = hipStreamSynchronize(s->stream);
    return e != hipSuccess ? e : e2 != hipSuccess ? e2 : e3;
}

// padded device grid -> compact host grid
hipError_t copy_grid_out(const fw_sim* s, void* dst, const void* src, size_t esize) {
    const size_t row = (size_t)s->row_len * esize, prow = (size_t)s->pitch * esize, rows = (size_t)s->rows;
    hipError_t e;
    if (row == prow) {
        e = hipMemcpyAsync(dst, src, row * rows, hipMemcpyDeviceToHost, s->stream);
        return e != hipSuccess ? e : hipStreamSynchronize(s->stream);
    }
    const size_t per = stage_cap(s) / row > 0 ? stage_cap(s) / row : 1;
    Stage st;
    e = st.get(s, per * row);
    for (size_t r0 = 0; r0 < rows && e == hipSuccess; r0 += per) {
        const size_t n = per < rows - r0 ? per : rows - r0;
        e = hipMemcpy2DAsync(st.p, row, (const char*)src + r0 * prow, prow, row, n, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync((char*)dst + r0 * row, st.p, n * row, hipMemcpyDeviceToHost, s->stream);
    }
    const hipError_t e2 = st.release(s);
    const hipError_t e3 = hipStreamSynchronize(s->stream);
    return e != hipSuccess ? e : e2 != hipSuccess ? e2 : e3;
}

hipError_t fill_zero_sync(const fw_sim* s, void* dst, size_t bytes) {
    hipError_t e = hipMemsetAsync(dst, 0, bytes, s->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(s->stream);
}

std::vector<std::string> required_maps(const fw_desc& d) {
    std::vector<std::string> names = {"K", "rho", "dcmap", "dmap", "icmat", "icc", "outc"};
    if (d.ncoordszero > 0) names.push_back("icczero");
    if (d.attenuation_model == 1) {
        names.push_back("a_exp");
        return names;
    }
    const int nax = d.isotropic ? 1 : d.ndim;
    for (int a = 0; a < nax; ++a) {
        names.push_back(std::string("kappa") + kAxesV[a]);
        names.push_back(std::string("kappa") + kAxesP[a]);
        for (int nu = 1; nu <= fw::kNRelax; ++nu) {
            for (const char* ab : {"apml", "bpml"}) {
                names.push_back(std::string(ab) + kAxesV[a] + std::to_string(nu));
                names.push_back(std::string(ab) + kAxesP[a] + std::to_string(nu));
            }
        }
    }
    return names;
}

bool is_voxel_map(const fw_desc& d, const std::string& name) {
    if (name == "K" || name == "rho" || name == "c" || name == "beta" || name == "dcmap") return true;
    if (d.attenuation_model == 1) return name == "a_exp";
    const int nax = d.isotropic ? 1 : d.ndim;
    for (int a = 0; a < nax; ++a) {
        if (name == std::string("kappa") + kAxesV[a] || name == std::string("kappa") + kAxesP[a]) return true;
        for (int nu = 1; nu <= fw::kNRelax; ++nu)
            for (const char* ab : {"apml", "bpml"})
                if (name == std::string(ab) + kAxesV[a] + std::to_string(nu) ||
                    name == std::string(ab) + kAxesP[a] + std::to_string(nu))
                    return true;
    }
    return false;
}

// Device memory of a handle is stream-ordered on the handle's stream: a free does not wait
// for the whole device (other handles' kernels), and the device's memory pool keeps freed
// blocks (up to kPoolKeep bytes) for the next handle, so a server creating and destroying many
// small handles does not pay an allocation and a device-wide wait per array. Where the
// runtime has no stream-ordered allocator, hipMalloc.
constexpr uint64_t kPoolKeep = 8ull << 30;

hipError_t dev_malloc(fw_sim* s, void** p, size_t bytes) {
    if (s->fixed_base) return fixed_carve(s, p, bytes);
    static std::mutex mu;
    static std::set<int> pools_set;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (pools_set.insert(s->device).second) {
            hipMemPool_t pool = nullptr;
            if (hipDeviceGetDefaultMemPool(&pool, s->device) == hipSuccess && pool) {
                uint64_t keep = kPoolKeep;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            }
            (void)hipGetLastError();
        }
    }
    hipError_t e = hipMallocAsync(p, bytes, s->stream);
    if (e == hipSuccess) return e;
    (void)hipGetLastError();
    e = hipMalloc(p, bytes);
    if (e == hipSuccess) s->plain_alloc.insert(*p);
    return e;
}

void dev_free(fw_sim* s, void* p) {
    if (!p || s->fixed_base) return;
    auto it = s->plain_alloc.find(p);
    if (it != s->plain_alloc.end()) {
        s->plain_alloc.erase(it);
        (void)hipStreamSynchronize(s->stream);
        (void)hipFree(p);
    } else if (hipFreeAsync(p, s->stream) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s->stream);
        (void)hipFree(p);
    }
}

int alloc(fw_sim* s, const std::string& name, size_t bytes, void** out) {
    auto it = s->dev.find(name);
    if (it != s->dev.end()) {
        *out = it->second;
        return FW_OK;
    }
    auto a = s->arena_slot.find(name);
    if (a != s->arena_slot.end() && bytes == (size_t)s->ncell * 4) {
        s->dev[name] = a->second;
        *out = a->second;
        return FW_OK;
    }
    void* p = nullptr;
    hipError_t e = dev_malloc(s, &p, bytes ? bytes : 4);
    if (e != hipSuccess && s->fixed_base)
        return fail(FW_ERR_NOMEM, "'%s' (%zu bytes) does not fit in the %zu-byte fixed device region (%zu used)",
                    name.c_str(), bytes, s->fixed_size, s->fixed_used);
    if (e != hipSuccess) return fail(FW_ERR_NOMEM, "hipMalloc(%zu) for '%s' failed: %s", bytes, name.c_str(), hipGetErrorString(e));
    s->dev[name] = p;
    *out = p;
    return FW_OK;
}

int check_coords(const fw_desc& d, int64_t sx, int64_t sy, const std::vector<int32_t>& c, int64_t n,
                 const char* what, std::vector<int64_t>& lin) {
    const int nd = d.ndim;
    if ((int64_t)c.size() < n * nd) return fail(FW_ERR_INVALID, "%s list holds fewer than %lld points", what, (long long)n);
    lin.resize((size_t)n);
    const int64_t dims[3] = {d.nx, d.ny, nd == 3 ? d.nz : 1};
    const int64_t stride[3] = {sx, nd == 3 ? sy : 1, 1};
    for (int64_t q = 0; q < n; ++q) {
        int64_t idx = 0;
        for (int a = 0; a < nd; ++a) {
            const int64_t x = c[(size_t)(q * nd + a)];
            if (x < 0 || x >= dims[a])
                return fail(FW_ERR_INVALID, "%s coordinate %lld (axis %d) = %lld outside [0, %lld)", what,
                            (long long)q, a, (long long)x, (long long)dims[a]);
            idx += x * stride[a];
        }
        lin[(size_t)q] = idx;
    }
    return FW_OK;
}

// Time one configuration of one stencil kernel, `reps` launches after one warm-up.
int time_config(fw_sim* s, const fw::LaunchCfg& cfg, bool velocity, int reps, float* ms) {
    auto launch = [&] {
        return velocity ? fw::launch_velocity(s->ds, cfg, s->stream) : fw::launch_pressure(s->ds, cfg, s->stream);
    };
    FW_HIP(launch());
    FW_HIP(hipEventRecord(s->ev[0], s->stream));
    for (int r = 0; r < reps; ++r) FW_HIP(launch());
    FW_HIP(hipEventRecord(s->ev[1], s->stream));
    FW_HIP(hipEventSynchronize(s->ev[1]));
    FW_HIP(hipEventElapsedTime(ms, s->ev[0], s->ev[1]));
    return FW_OK;
}

// Whether `c` can run this handle's grid.
int check_cfg(fw_sim* s, const fw::LaunchCfg& c) {
    if (c.variant == 0) return FW_OK;
    if (c.variant == 2) {
        if (!fw::tile2d_supported(s->ds))
            return fail(FW_ERR_UNSUPPORTED,
                        "tile kernels need a 2D relaxation grid with ny %% 4 == 0 and at most 768 sound speeds");
        if (c.width != 256 || (c.chunk != 8 && c.chunk != 16 && c.chunk != 32))
            return fail(FW_ERR_INVALID, "tile kernels: width 256 and rows 8, 16 or 32");
        return FW_OK;
    }
    if (c.variant == 3 || c.variant == 4) {
        if (s->ndim != 3 || !fw::streaming_supported(s->ds))
            return fail(FW_ERR_UNSUPPORTED, "pipelined kernels need a 3D relaxation grid");
        if (s->ncell >= ((int64_t)1 << 30))  // 32-bit byte offsets per array
            return fail(FW_ERR_UNSUPPORTED, "pipelined kernels take grids of fewer than 2^30 cells");
        if (c.width != 512 && c.width != 1024)
            return fail(FW_ERR_INVALID, "pipelined kernels: workgroup size (width) must be 512 or 1024");
        if (c.chunk < 1 || c.chunk > 4096) return fail(FW_ERR_INVALID, "pipelined kernels: chunk out of range");
        return FW_OK;
    }
    if (c.variant == 1) {
        if (!fw::streaming_supported(s->ds)) return fail(FW_ERR_UNSUPPORTED, "this grid needs the per-cell kernels");
        if (s->ndim == 2 && c.width != 64 && c.width != 128 && c.width != 256)
            return fail(FW_ERR_INVALID, "x-marching kernels: width must be 64, 128 or 256");
        if (s->ndim == 3 && c.width != 512 && c.width != 1024)
            return fail(FW_ERR_INVALID, "3D x-marching kernels: workgroup size (width) must be 512 or 1024");
        if (c.chunk < 2 || c.chunk > 4096) return fail(FW_ERR_INVALID, "x-marching kernels: chunk out of range");
        return FW_OK;
    }
    return fail(FW_ERR_INVALID, "kernel variant must be 0 to 4");
}

int parse_kernel_env(fw_sim* s, const char* env, fw::LaunchCfg* out) {
    const bool stream_ok = fw::streaming_supported(s->ds);
    fw::LaunchCfg c;
    c.width = s->ndim == 3 ? 512 : 128;
    c.chunk = 16;
    if (!strncmp(env, "percell", 7)) {
        c.variant = 0;
    } else if (!strncmp(env, "tile", 4)) {
        if (!fw::tile2d_supported(s->ds))
            return fail(FW_ERR_UNSUPPORTED, "FULLWAVE_HIP_KERNELS=tile needs a 2D grid with ny %% 4 == 0");
        c.variant = 2;
        c.width = 256;
        int ch = 16;
        c.chunk = sscanf(env, "tile:%d", &ch) == 1 ? ch : 16;
        if (c.chunk != 8 && c.chunk != 16 && c.chunk != 32)
            return fail(FW_ERR_INVALID, "FULLWAVE_HIP_KERNELS: tile rows must be 8, 16 or 32");
    } else if (!strncmp(env, "pipe", 4) || !strncmp(env, "pipelite", 8)) {
        const bool lite = !strncmp(env, "pipelite", 8);
        c.variant = lite ? 4 : 3;
        int w = 0, ch = 0;
        const int n = sscanf(env + (lite ? 8 : 4), ":%d:%d", &w, &ch);
        if (n >= 1) c.width = w;
        if (n >= 2) c.chunk = ch;
        int rc = check_cfg(s, c);
        if (rc != FW_OK) return rc;
    } else if (!strncmp(env, "stream", 6)) {
        if (!stream_ok) return fail(FW_ERR_UNSUPPORTED, "FULLWAVE_HIP_KERNELS=stream: this grid needs the per-cell kernels");
        c.variant = 1;
        int w = 0, ch = 0;
        const int n = sscanf(env, "stream:%d:%d", &w, &ch);
        if (n >= 1) c.width = w;
        if (n >= 2) c.chunk = ch;
        int rc = check_cfg(s, c);
        if (rc != FW_OK) return rc;
    } else {
        return fail(FW_ERR_INVALID,
                    "FULLWAVE_HIP_KERNELS must be 'percell', 'stream[:width[:chunk]]', 'tile[:rows]' or 'pipe[:width[:chunk]]'");
    }
    *out = c;
    return FW_OK;
}

int choose_kernels_unpinned(fw_sim* s);

// Pick the kernel family and shape of each stencil kernel for this grid.
// FULLWAVE_HIP_KERNELS overrides both: "percell", "stream[:width[:chunk]]" or "tile[:rows]",
// or each on its own as "<velocity>/<pressure>" (e.g. "tile:8/stream:256:8").
// Otherwise every candidate is timed on the handle's own arrays while the wave state is still
// all zero (the update then leaves it zero, so the timing runs change nothing), and the
// fastest velocity kernel and the fastest pressure kernel are kept.
int choose_kernels(fw_sim* s) {
    int rc0 = choose_kernels_unpinned(s);
    if (rc0 != FW_OK) return rc0;
    for (int kind = 0; kind < 2; ++kind) {
        if (!s->pinned[kind]) continue;
        if ((rc0 = check_cfg(s, s->pin[kind])) != FW_OK) return rc0;
        (kind == 0 ? s->cfg_v : s->cfg_p) = s->pin[kind];
    }
    return FW_OK;
}

int choose_kernels_unpinned(fw_sim* s) {
    const char* env = getenv("FULLWAVE_HIP_KERNELS");
    if (env && *env) {
        const char* slash = strchr(env, '/');
        if (!slash) {
            int rc = parse_kernel_env(s, env, &s->cfg_v);
            if (rc != FW_OK) return rc;
            s->cfg_p = s->cfg_v;
            return FW_OK;
        }
        char vel[64];
        const size_t n = (size_t)(slash - env);
        if (n == 0 || n >= sizeof(vel)) return fail(FW_ERR_INVALID, "FULLWAVE_HIP_KERNELS: bad '<velocity>/<pressure>' pair");
        memcpy(vel, env, n);
        vel[n] = 0;
        int rc = parse_kernel_env(s, vel, &s->cfg_v);
        if (rc != FW_OK) return rc;
        return parse_kernel_env(s, slash + 1, &s->cfg_p);
    }
    const bool stream_ok = fw::streaming_supported(s->ds);
    fw::LaunchCfg def;
    def.variant = stream_ok ? 1 : 0;
    def.width = s->ndim == 3 ? 512 : 256;
    def.chunk = 8;
    s->cfg_v = s->cfg_p = def;
    // a fixed-region handle has no timing events: the default shapes
    if (s->fields_set || s->fixed_base || !stream_ok || (s->pinned[0] && s->pinned[1])) return FW_OK;
    std::vector<fw::LaunchCfg> cands;
    fw::LaunchCfg pc;
    pc.variant = 0;
    cands.push_back(pc);
    for (int ch : {8, 16, 32, 48, 64, 96, 128}) {  // beyond 32 planes per workgroup: 3D only
        if (s->ndim == 2 && ch > 32) continue;
        if (s->ndim == 2) {
            for (int w : {128, 256}) {
                fw::LaunchCfg c;
                c.variant = 1;
                c.width = w;
                c.chunk = ch;
                cands.push_back(c);
            }
        } else {
            for (int v : {1, 3, 4})
                for (int w : {512, 1024}) {
                    fw::LaunchCfg c;
                    c.variant = v;
                    c.width = w;
                    c.chunk = ch;
                    if (v == 1 || s->ncell < ((int64_t)1 << 30)) cands.push_back(c);
                }
        }
    }
    if (fw::tile2d_supported(s->ds)) {
        for (int ch : {8, 16, 32}) {
            fw::LaunchCfg c;
            c.variant = 2;
            c.width = 256;
            c.chunk = ch;
            cands.push_back(c);
        }
    }
    for (int kind = 0; kind < 2; ++kind) {
        if (s->pinned[kind]) continue;
        float best = 1e30f;
        for (const auto& c : cands) {
            float ms = 0.f;
            int rc = time_config(s, c, kind == 0, 3, &ms);
            if (rc != FW_OK) return rc;
            if (ms < best) {
                best = ms;
                (kind == 0 ? s->cfg_v : s->cfg_p) = c;
            }
        }
    }
    return FW_OK;
}

bool codes_wanted(const fw_sim* s) {
    if (s->fixed_base) return false;  // coding needs temporary tables of up to 100 MB
    if (s->codes_mode >= 0) return s->codes_mode == 1;
    const char* env = getenv("FULLWAVE_HIP_CODES");
    return !(env && env[0] == '0');
}

void drop_codes(fw_sim* s) {
    for (const char* name : {"_rcode", "_rtab_v", "_rtab_p"}) {
        auto it = s->dev.find(name);
        if (it == s->dev.end()) continue;
        dev_free(s, it->second);
        s->dev.erase(it);
    }
    s->ds.rcode = nullptr;
    s->ds.rtab_v = nullptr;
    s->ds.rtab_p = nullptr;
    s->ncodes = 0;
    s->coded_cells = 0;
}

// Code the isotropic relaxation maps and dcmap (kernels and the steps in fw_kernels_codes.hip).
// Leaves the handle on the per-cell maps when the model is not isotropic relaxation, when codes
// are switched off, when the codes would cover less than half of the cells, or when any check
// fails; device errors are errors.
int encode_relaxation(fw_sim* s) {
    drop_codes(s);
    const fw_desc& d = s->d;
    if (d.attenuation_model != 0 || !d.isotropic || !codes_wanted(s)) return FW_OK;
    if (s->ncell >= ((int64_t)1 << 30)) return FW_OK;  // the marching kernels' 32-bit offsets
    fw::CodeMaps m{};
    const char* names[fw::kCodeVals - 1] = {"kappau", "apmlu1", "apmlu2", "bpmlu1", "bpmlu2",
                                            "kappax", "apmlx1", "apmlx2", "bpmlx1", "bpmlx2"};
    for (int k = 0; k < fw::kCodeVals - 1; ++k) m.f[k] = (const float*)s->dev[names[k]];
    m.dcmap = (const int32_t*)s->dev["dcmap"];

    // large grids (3D) hold more tuples than the small table takes: start with the large one
    size_t nslots = s->ncell > ((int64_t)1 << 24) ? fw::kCodeSlots : fw::kCodeSlotsSmall;
    // temporaries are stream-ordered (no device-wide wait when they are freed, which would
    // stall a server's other jobs)
    Stage keys_b, count_b, rep_b, status_b, slot_b, repc_b;
    unsigned long long* keys = nullptr;
    uint32_t* count = nullptr;
    int32_t* rep = nullptr;
    int32_t* status = nullptr;
    uint16_t* slot_code_d = nullptr;
    int32_t* rep_of_code_d = nullptr;
    auto free_table = [&] {
        (void)keys_b.release(s);
        (void)count_b.release(s);
        (void)rep_b.release(s);
    };
    auto run = [&]() -> int {
        FW_HIP(status_b.get(s, 16));
        status = (int32_t*)status_b.p;
        int32_t st[4] = {0, 0, 0, 0};
        for (;;) {
            FW_HIP(keys_b.get(s, nslots * 8));
            FW_HIP(count_b.get(s, nslots * 4));
            FW_HIP(rep_b.get(s, nslots * 4));
            keys = (unsigned long long*)keys_b.p;
            count = (uint32_t*)count_b.p;
            rep = (int32_t*)rep_b.p;
            FW_HIP(hipMemsetAsync(keys, 0, nslots * 8, s->stream));
            FW_HIP(hipMemsetAsync(count, 0, nslots * 4, s->stream));
            FW_HIP(hipMemsetAsync(rep, 0x7f, nslots * 4, s->stream));  // above every cell index
            FW_HIP(hipMemsetAsync(status, 0, 16, s->stream));
            FW_HIP(fw::launch_code_insert(m, s->ncell, keys, (uint32_t)nslots, count, rep, status, s->stream));
            FW_HIP(copy_sync(s, st, status, 16, hipMemcpyDeviceToHost));
            if (!st[0]) break;
            if (nslots == fw::kCodeSlots) return FW_OK;  // probe limit: far too many distinct tuples
            free_table();
            nslots = fw::kCodeSlots;
        }
        std::vector<unsigned long long> hk(nslots);
        std::vector<uint32_t> hc(nslots);
        std::vector<int32_t> hr(nslots);
        FW_HIP(copy_sync(s, hk.data(), keys, nslots * 8, hipMemcpyDeviceToHost));
        FW_HIP(copy_sync(s, hc.data(), count, nslots * 4, hipMemcpyDeviceToHost));
        FW_HIP(copy_sync(s, hr.data(), rep, nslots * 4, hipMemcpyDeviceToHost));
        std::vector<uint32_t> occupied;
        for (size_t q = 0; q < nslots; ++q)
            if (hk[q]) occupied.push_back((uint32_t)q);
        if (occupied.empty()) return FW_OK;
        // the most frequent tuples get codes (ties: lower slot first, so the choice is fixed)
        std::sort(occupied.begin(), occupied.end(), [&](uint32_t x, uint32_t y) {
            return hc[x] != hc[y] ? hc[x] > hc[y] : x < y;
        });
        if (occupied.size() > (size_t)fw::kMaxCodes) occupied.resize((size_t)fw::kMaxCodes);
        int64_t covered = 0;
        for (uint32_t q : occupied) covered += hc[q];
        if (2 * covered < s->ncell) return FW_OK;
        // codes in the order of each tuple's first cell: cells close together that hold
        // different tuples gather rows close together (3D pressure kernel on the 904 x 904 x 776
        // inclusion grid: 8.68 -> 8.61 ms a step; 8.34 ms if every cell read the same row)
        std::sort(occupied.begin(), occupied.end(), [&](uint32_t x, uint32_t y) { return hr[x] < hr[y]; });
        std::vector<uint16_t> slot_code(nslots, (uint16_t)fw::kCodeEscape);
        std::vector<int32_t> rep_of_code(occupied.size());
        for (size_t i = 0; i < occupied.size(); ++i) {
            slot_code[occupied[i]] = (uint16_t)i;
            rep_of_code[i] = hr[occupied[i]];
        }
        const int32_t n = (int32_t)occupied.size();
        FW_HIP(slot_b.get(s, nslots * 2));
        FW_HIP(repc_b.get(s, (size_t)n * 4));
        slot_code_d = (uint16_t*)slot_b.p;
        rep_of_code_d = (int32_t*)repc_b.p;
        FW_HIP(copy_sync(s, slot_code_d, slot_code.data(), nslots * 2, hipMemcpyHostToDevice));
        FW_HIP(copy_sync(s, rep_of_code_d, rep_of_code.data(), (size_t)n * 4, hipMemcpyHostToDevice));
        void *code, *tv, *tp;
        int r;
        if ((r = alloc(s, "_rcode", (size_t)s->ncell * 2, &code)) != FW_OK) return r;
        if ((r = alloc(s, "_rtab_v", (size_t)n * fw::kCodeStride * 4, &tv)) != FW_OK) return r;
        if ((r = alloc(s, "_rtab_p", (size_t)n * fw::kCodeStride * 4, &tp)) != FW_OK) return r;
        FW_HIP(fw::launch_code_assign(m, s->ncell, keys, (uint32_t)nslots, slot_code_d, (uint16_t*)code, status,
                                      s->stream));
        FW_HIP(fw::launch_code_table(m, n, rep_of_code_d, (float*)tv, (float*)tp, s->stream));
        FW_HIP(fw::launch_code_verify(m, s->ncell, (const uint16_t*)code, n, (const float*)tv, (const float*)tp,
                                      status, s->stream));
        FW_HIP(copy_sync(s, st, status, 16, hipMemcpyDeviceToHost));
        if (st[1] || st[2]) return FW_OK;  // a key not found again, or a collision
        s->ds.rcode = (const uint16_t*)code;
        s->ds.rtab_v = (const float*)tv;
        s->ds.rtab_p = (const float*)tp;
        s->ncodes = n;
        s->coded_cells = covered;
        return FW_OK;
    };
    const int rc = run();
    free_table();
    (void)status_b.release(s);
    (void)slot_b.release(s);
    (void)repc_b.release(s);
    const hipError_t e = hipStreamSynchronize(s->stream);
    if (rc != FW_OK || s->ncodes == 0) drop_codes(s);
    if (rc == FW_OK && e != hipSuccess) return fail(FW_ERR_DEVICE, "coding: %s", hipGetErrorString(e));
    return rc;
}

// New codes for a handle that has run: a captured graph holds the old table addresses.
int recode(fw_sim* s) {
    FW_HIP(hipStreamSynchronize(s->stream));
    if (s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    return encode_relaxation(s);
}

// The boundary entries of a handle whose pressure kernel does the step's boundary work: one per
// shot, with the handle's current pointers (step0 0: the launch passes the absolute step).
std::vector<fw::GroupBoundary> fused_entries(const fw_sim* s, const fw::GroupBoundary& lists) {
    std::vector<fw::GroupBoundary> out;
    for (int b = 0; b < s->nshots; ++b) {
        fw::GroupBoundary e = lists;
        e.p = s->ds.p + (int64_t)b * s->ds.shot_stride;
        e.b = s->db;
        e.b.icmat_t = s->db.icmat_t + (int64_t)b * s->db.icmat_stride;
        e.b.rec = s->db.rec + (int64_t)b * s->db.rec_stride;
        e.modt = s->d.modt;
        e.step0 = 0;
        out.push_back(e);
    }
    return out;
}

// Group the handle's points by the tiles of its pressure kernel when that kernel can do the
// step's boundary work (2D marching or tiled kernels; not for a fixed-region handle, whose
// footprint does not count these lists).
int setup_fused_boundary(fw_sim* s, const fw::BoundaryLists& bl) {
    s->fused_boundary = false;
    if (s->fixed_base || !fw::pressure_fuses_boundary(s->ds, s->cfg_p)) return FW_OK;
    const char* env = getenv("FULLWAVE_HIP_FUSED_BOUNDARY");  // "0": a launch of its own (tests)
    if (env && env[0] == '0') return FW_OK;
    const fw::LaunchCfg& c = s->cfg_p;
    const int tile_w = c.variant == 2 ? 256 : (c.width == 64 || c.width == 256 ? c.width : 128);
    const int rows = c.variant == 2 ? c.chunk : (c.chunk < 2 ? 2 : c.chunk);
    const unsigned gx = (unsigned)((s->ds.ny + tile_w - 1) / tile_w);
    std::vector<int32_t> pts;
    const fw::PointTable pt = fw::point_table(s->ds.nx, s->ds.ny, gx, tile_w, rows, bl, &pts);
    void* p;
    int rc;
    if ((rc = alloc(s, "_fpoints", pts.size() * 4, &p)) != FW_OK) return rc;
    FW_HIP(copy_sync(s, p, pts.data(), pts.size() * 4, hipMemcpyHostToDevice));
    const int32_t* base = (const int32_t*)p;
    fw::GroupBoundary lists_e{};
    lists_e.ntiles = pt.ntiles;
    lists_e.off = base + pt.off;
    lists_e.src = reinterpret_cast<const int4*>(base + pt.src);
    lists_e.zero = reinterpret_cast<const int64_t*>(base + pt.zero);
    lists_e.out = reinterpret_cast<const int4*>(base + pt.out);
    s->fbound_h = fused_entries(s, lists_e);
    if ((rc = alloc(s, "_fbound", s->fbound_h.size() * sizeof(fw::GroupBoundary), &p)) != FW_OK) return rc;
    s->fbound = (fw::GroupBoundary*)p;
    FW_HIP(copy_sync(s, s->fbound, s->fbound_h.data(), s->fbound_h.size() * sizeof(fw::GroupBoundary),
                     hipMemcpyHostToDevice));
    s->fused_boundary = true;
    return FW_OK;
}

// Resolve coordinate lists into device index arrays and build the coefficient table.
int finalize(fw_sim* s) {
    const fw_desc& d = s->d;
    for (const auto& name : required_maps(d))
        if (!s->have.count(name)) return fail(FW_ERR_STATE, "input '%s' has not been uploaded", name.c_str());

    int rc;
    fw::BoundaryLists bl;
    if ((rc = fw::resolve_boundary(d, s->ds.sx, s->ds.sy, s->icc, s->outc, s->zero, &bl)) != FW_OK) return rc;
    const std::vector<int64_t>& src_lin = bl.src_lin;
    const std::vector<int64_t>& zero = bl.zero_lin;
    const std::vector<int64_t>& out = bl.out_lin;
    const std::vector<int32_t>& out_src = bl.out_src;
    const std::vector<int8_t>& out_air = bl.out_air;

    void* p;
    auto up = [&](const char* name, const void* h, size_t bytes) -> int {
        int r = alloc(s, name, bytes, &p);
        if (r != FW_OK) return r;
        if (bytes) FW_HIP(copy_sync(s, p, h, bytes, hipMemcpyHostToDevice));
        return FW_OK;
    };
    if ((rc = up("_src_lin", src_lin.data(), src_lin.size() * 8)) != FW_OK) return rc;
    s->db.src_lin = (const int64_t*)s->dev["_src_lin"];
    if ((rc = up("_zero_lin", zero.data(), zero.size() * 8)) != FW_OK) return rc;
    s->db.zero_lin = (const int64_t*)s->dev["_zero_lin"];
    s->out_lin_h = out;
    if ((rc = up("_out_lin", out.data(), out.size() * 8)) != FW_OK) return rc;
    s->db.out_lin = (const int64_t*)s->dev["_out_lin"];
    if ((rc = up("_out_src", out_src.data(), out_src.size() * 4)) != FW_OK) return rc;
    s->db.out_src = (const int32_t*)s->dev["_out_src"];
    if ((rc = up("_out_air", out_air.data(), out_air.size())) != FW_OK) return rc;
    s->db.out_air = (const int8_t*)s->dev["_out_air"];
    s->db.ncoords = d.ncoords;
    s->db.nzero = (int64_t)zero.size();
    s->db.nout = d.ncoordsout;
    s->db.ntic = d.ntic;

    int axes_same = 1;
    const std::vector<float> coef = fw::coefficient_table(d, s->dmap, &axes_same);
    if ((rc = up("_coef", coef.data(), coef.size() * 4)) != FW_OK) return rc;
    s->ds.coef = (const float*)s->dev["_coef"];
    s->ds.coef_axes_same = axes_same;

    // dcmap range check on the host would need a download; the kernels trust it, so check it
    // here once against ndmap
    {
        std::vector<int32_t> dc((size_t)s->ncell_c);
        FW_HIP(copy_grid_out(s, dc.data(), s->dev["dcmap"], 4));
        for (int32_t v : dc)
            if (v < 0 || v >= d.ndmap) return fail(FW_ERR_INVALID, "dcmap value %d outside [0, %lld)", v, (long long)d.ndmap);
    }

    s->rec_cap = d.modt > 0 ? (d.nt + d.modt - 1) / d.modt : 0;
    const size_t rec_bytes = (size_t)(s->nshots * s->rec_cap * d.ncoordsout) * 4;
    if ((rc = alloc(s, "_rec", rec_bytes, &p)) != FW_OK) return rc;
    s->db.rec = (float*)p;
    s->db.rec_stride = s->rec_cap * d.ncoordsout;
    FW_HIP(fill_zero_sync(s, p, rec_bytes));

    // map pointers
    const int nd = s->ndim;
    const int nax = d.isotropic ? 1 : nd;
    s->ds.aexp = d.attenuation_model == 1 ? (const float*)s->dev["a_exp"] : nullptr;
    for (int a = 0; a < nd && d.attenuation_model == 0; ++a) {
        const int src_axis = nax == 1 ? 0 : a;
        s->ds.ikv[a] = (const float*)s->dev[std::string("kappa") + kAxesV[src_axis]];
        s->ds.ikp[a] = (const float*)s->dev[std::string("kappa") + kAxesP[src_axis]];
        for (int nu = 0; nu < fw::kNRelax; ++nu) {
            const std::string n = std::to_string(nu + 1);
            s->ds.av[a][nu] = (const float*)s->dev[std::string("apml") + kAxesV[src_axis] + n];
            s->ds.bv[a][nu] = (const float*)s->dev[std::string("bpml") + kAxesV[src_axis] + n];
            s->ds.ap[a][nu] = (const float*)s->dev[std::string("apml") + kAxesP[src_axis] + n];
            s->ds.bp[a][nu] = (const float*)s->dev[std::string("bpml") + kAxesP[src_axis] + n];
        }
    }
    s->ds.K = (const float*)s->dev["K"];
    s->ds.rho = (const float*)s->dev["rho"];
    s->ds.dcmap = (const int32_t*)s->dev["dcmap"];
    s->ds.ndmap = (int32_t)d.ndmap;
    s->ds.isotropic = d.isotropic;
    if ((rc = encode_relaxation(s)) != FW_OK) return rc;
    if ((rc = choose_kernels(s)) != FW_OK) return rc;
    if ((rc = setup_fused_boundary(s, bl)) != FW_OK) return rc;
    s->ready = true;
    return FW_OK;
}

int launch_step(fw_sim* s, bool timed, double* ms) {
    hipError_t e;
    auto timed_launch = [&](int kind, auto&& fn) -> int {
        if (timed) FW_HIP(hipEventRecord(s->ev[0], s->stream));
        e = fn();
        if (e != hipSuccess) return fail(FW_ERR_DEVICE, "kernel launch failed: %s", hipGetErrorString(e));
        if (timed) {
            FW_HIP(hipEventRecord(s->ev[1], s->stream));
            FW_HIP(hipEventSynchronize(s->ev[1]));
            float t = 0.f;
            FW_HIP(hipEventElapsedTime(&t, s->ev[0], s->ev[1]));
            ms[kind] += t;
        }
        return FW_OK;
    };
    int rc;
    rc = timed_launch(fw::kVelocity, [&] { return fw::launch_velocity(s->ds, s->cfg_v, s->stream); });
    if (rc != FW_OK) return rc;
    if (s->fused_boundary && !s->ds.tape_div) {  // the pressure launch does the boundary work
        const fw::GroupStep gs{s->fbound, nullptr, s->step};
        rc = timed_launch(fw::kPressure, [&] { return fw::launch_pressure(s->ds, s->cfg_p, s->stream, gs); });
        if (rc != FW_OK) return rc;
    } else {
        rc = timed_launch(fw::kPressure, [&] { return fw::launch_pressure(s->ds, s->cfg_p, s->stream); });
        if (rc != FW_OK) return rc;
        rc = timed_launch(fw::kBoundary,
                          [&] { return fw::launch_boundary(s->ds, s->db, s->step, s->d.modt, s->stream); });
        if (rc != FW_OK) return rc;
    }
    s->step += 1;
    s->state_zero = false;
    return FW_OK;
}

// Capture graph_steps steps (velocity, pressure, boundary pass with the device step counter,
// counter advance) into one graph, once per handle.
int capture_graph(fw_sim* s) {
    if (!s->step_dev) FW_HIP(dev_malloc(s, (void**)&s->step_dev, sizeof(int64_t)));
    hipGraph_t graph = nullptr;
    FW_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int64_t it = 0; it < s->graph_steps && e == hipSuccess; ++it) {
        e = fw::launch_velocity(s->ds, s->cfg_v, s->stream);
        if (e == hipSuccess && s->fused_boundary) {
            e = fw::launch_pressure(s->ds, s->cfg_p, s->stream, fw::GroupStep{s->fbound, s->step_dev, 0});
        } else {
            if (e == hipSuccess) e = fw::launch_pressure(s->ds, s->cfg_p, s->stream);
            if (e == hipSuccess) e = fw::launch_boundary_dev(s->ds, s->db, s->step_dev, s->d.modt, s->stream);
        }
        if (e == hipSuccess) e = fw::launch_tick(s->step_dev, -1, s->stream);
    }
    hipError_t e2 = hipStreamEndCapture(s->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(FW_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    e = hipGraphInstantiate(&s->graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        s->graph = nullptr;
        return fail(FW_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    return FW_OK;
}

// Replay the captured steps as often as they fit in *left; leaves the rest in *left.
int run_graph(fw_sim* s, int64_t* left) {
    if (!s->graph) {
        int rc = capture_graph(s);
        if (rc != FW_OK) return rc;
    }
    FW_HIP(fw::launch_tick(s->step_dev, s->step, s->stream));  // device counter = host step
    while (*left >= s->graph_steps) {
        FW_HIP(hipGraphLaunch(s->graph, s->stream));
        s->step += s->graph_steps;
        s->state_zero = false;
        *left -= s->graph_steps;
    }
    return FW_OK;
}

int run_impl(fw_sim* s, int64_t n_steps, bool timed, double* ms) {
    if (!s) return fail(FW_ERR_INVALID, "null handle");
    if (n_steps < 0) return fail(FW_ERR_INVALID, "n_steps < 0");
    DeviceGuard g(s->device);
    if (!s->ready) {
        int rc = finalize(s);
        if (rc != FW_OK) return rc;
    }
    if (s->step + n_steps > s->d.nt)
        return fail(FW_ERR_INVALID, "running %lld steps from step %lld exceeds nt = %lld", (long long)n_steps,
                    (long long)s->step, (long long)s->d.nt);
    if (s->fused_boundary) {  // the entries follow the handle's buffers (a source matrix re-uploaded)
        std::vector<fw::GroupBoundary> now = fused_entries(s, s->fbound_h[0]);
        if (memcmp(now.data(), s->fbound_h.data(), now.size() * sizeof(fw::GroupBoundary)) != 0) {
            s->fbound_h = now;
            FW_HIP(hipMemcpyAsync(s->fbound, s->fbound_h.data(), now.size() * sizeof(fw::GroupBoundary),
                                  hipMemcpyHostToDevice, s->stream));
            FW_HIP(hipStreamSynchronize(s->stream));
        }
    }
    int64_t left = n_steps;
    if (!timed && s->graph_steps > 0 && left >= s->graph_steps) {
        int rc = run_graph(s, &left);
        if (rc != FW_OK) return rc;
    }
    for (int64_t it = 0; it < left; ++it) {
        int rc = launch_step(s, timed, ms);
        if (rc != FW_OK) return rc;
    }
    FW_HIP(hipStreamSynchronize(s->stream));
    return FW_OK;
}

}  // namespace

namespace fw {

int resolve_boundary(const fw_desc& d, int64_t sx, int64_t sy, const std::vector<int32_t>& icc,
                     const std::vector<int32_t>& outc, const std::vector<int32_t>& zero_c, BoundaryLists* out) {
    std::vector<int64_t> src, zero, sen;
    int rc;
    if ((rc = check_coords(d, sx, sy, icc, d.ncoords, "source", src)) != FW_OK) return rc;
    if ((rc = check_coords(d, sx, sy, outc, d.ncoordsout, "sensor", sen)) != FW_OK) return rc;
    if ((rc = check_coords(d, sx, sy, zero_c, d.ncoordszero, "pressure-release", zero)) != FW_OK) return rc;
    // pressure release wins over injection; a repeated source point keeps its last signal
    std::unordered_map<int64_t, int32_t> air, last_src;
    for (int64_t q = 0; q < d.ncoordszero; ++q) air[zero[(size_t)q]] = 1;
    for (int64_t q = 0; q < d.ncoords; ++q) last_src[src[(size_t)q]] = (int32_t)q;
    out->src_lin.assign((size_t)d.ncoords, -1);
    for (int64_t q = 0; q < d.ncoords; ++q) {
        const int64_t l = src[(size_t)q];
        out->src_lin[(size_t)q] = (air.count(l) || last_src[l] != q) ? -1 : l;
    }
    out->out_src.assign((size_t)d.ncoordsout, -1);
    out->out_air.assign((size_t)d.ncoordsout, 0);
    for (int64_t q = 0; q < d.ncoordsout; ++q) {
        const int64_t l = sen[(size_t)q];
        out->out_air[(size_t)q] = air.count(l) ? 1 : 0;
        auto it = last_src.find(l);
        out->out_src[(size_t)q] = it == last_src.end() ? -1 : it->second;
    }
    // one writer per pressure-release address
    std::sort(zero.begin(), zero.end());
    zero.erase(std::unique(zero.begin(), zero.end()), zero.end());
    out->zero_lin = std::move(zero);
    out->out_lin = std::move(sen);
    return FW_OK;
}

std::vector<float> coefficient_table(const fw_desc& d, const std::vector<float>& dmap, int* axes_same) {
    const int nd = d.ndim == 3 ? 3 : 2;
    const double h[3] = {d.dx, d.dy, d.dz};
    std::vector<float> coef((size_t)d.ndmap * nd * kNCoef);
    if (dmap.size() < (size_t)(9 * 2 * d.ndmap)) return {};
    for (int64_t m = 0; m < d.ndmap; ++m) {
        for (int a = 0; a < nd; ++a) {
            float* row = &coef[(size_t)((m * nd + a) * kNCoef)];
            for (int k = 1; k <= 8; ++k) row[k - 1] = (float)((double)dmap[(size_t)((k * 2 + 0) * d.ndmap + m)] / h[a]);
            row[8] = (float)((double)dmap[(size_t)((1 * 2 + 1) * d.ndmap + m)] / h[a]);
        }
    }
    int same = 1;
    for (int64_t m = 0; m < d.ndmap && same; ++m)
        for (int a = 1; a < nd; ++a)
            if (memcmp(&coef[(size_t)(m * nd * kNCoef)], &coef[(size_t)((m * nd + a) * kNCoef)], kNCoef * sizeof(float)) != 0)
                same = 0;
    if (axes_same) *axes_same = same;
    return coef;
}

hipError_t grid_to_host(const fw_sim* s, void* dst, const void* src, size_t esize) {
    return copy_grid_out(s, dst, src, esize);
}

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int prepare(fw_sim* s) { return s->ready ? FW_OK : finalize(s); }

int forward_step(fw_sim* s) { return launch_step(s, false, nullptr); }

}  // namespace fw

extern "C" {

int fw_abi_version(void) { return FW_ABI_VERSION; }

const char* fw_version_string(void) { return "fullwave25_amd engine 0.1 (gfx950, HIP)"; }

const char* fw_last_error(void) { return g_last_error.c_str(); }

int fw_device_count(int* count) {
    if (!count) return fail(FW_ERR_INVALID, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return fail(FW_ERR_DEVICE, "hipGetDeviceCount failed: %s", hipGetErrorString(e));
    }
    *count = n;
    return FW_OK;
}

int fw_device_copy_bandwidth(int device, size_t bytes, int repeats, double* gbs) {
    if (!gbs || repeats < 1 || bytes < 16) return fail(FW_ERR_INVALID, "need gbs, repeats >= 1, bytes >= 16");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(FW_ERR_DEVICE, "device %d not available", device);
    DeviceGuard g(device);
    const int64_t n16 = (int64_t)(bytes / 16);
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FW_OK;
    if (hipMalloc(&a, n16 * 16) != hipSuccess || hipMalloc(&b, n16 * 16) != hipSuccess) {
        rc = fail(FW_ERR_NOMEM, "hipMalloc of 2 x %lld bytes failed", (long long)(n16 * 16));
    } else if (hipMemset(a, 0, n16 * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
               hipEventCreate(&e1) != hipSuccess) {
        rc = fail(FW_ERR_DEVICE, "setup failed");
    } else {
        hipError_t e = fw::launch_copy(a, b, n16, nullptr);  // warm-up
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < repeats && e == hipSuccess; ++r) e = fw::launch_copy(a, b, n16, nullptr);
        (void)hipEventRecord(e1, nullptr);
        float ms = 0.f;
        if (e != hipSuccess || hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(FW_ERR_DEVICE, "copy kernel failed");
        else
            *gbs = 2.0 * (double)(n16 * 16) * repeats / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int fw_device_stream_bandwidth(int device, int64_t cells, int32_t nread, int32_t nwrite, int32_t repeats,
                               double* gbs) {
    if (!gbs || repeats < 1 || cells < 1 || nread < 0 || nwrite < 1 || nread > fw::kMaxStreams ||
        nwrite > fw::kMaxStreams)
        return fail(FW_ERR_INVALID, "need gbs, repeats >= 1, cells >= 1, 0..%d read and 1..%d write streams",
                    fw::kMaxStreams, fw::kMaxStreams);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(FW_ERR_DEVICE, "device %d not available", device);
    DeviceGuard g(device);
    const size_t slot = ((size_t)cells * 4 + 255) / 256 * 256;
    char* base = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FW_OK;
    if (hipMalloc((void**)&base, slot * (size_t)(nread + nwrite)) != hipSuccess) {
        rc = fail(FW_ERR_NOMEM, "hipMalloc of %d arrays of %lld floats failed", (int)(nread + nwrite), (long long)cells);
    } else if (hipMemset(base, 0, slot * (size_t)(nread + nwrite)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
               hipEventCreate(&e1) != hipSuccess) {
        rc = fail(FW_ERR_DEVICE, "setup failed");
    } else {
        fw::StreamSet a{};
        for (int k = 0; k < nread; ++k) a.r[k] = (const float*)(base + slot * (size_t)k);
        for (int k = 0; k < nwrite; ++k) a.w[k] = (float*)(base + slot * (size_t)(nread + k));
        hipError_t e = fw::launch_streams(a, nread, nwrite, cells, nullptr);  // warm-up
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < repeats && e == hipSuccess; ++r) e = fw::launch_streams(a, nread, nwrite, cells, nullptr);
        (void)hipEventRecord(e1, nullptr);
        float ms = 0.f;
        if (e != hipSuccess || hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(FW_ERR_DEVICE, "stream kernel failed");
        else
            *gbs = 4.0 * (double)cells * (nread + nwrite) * repeats / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (base) (void)hipFree(base);
    return rc;
}

int fw_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(FW_ERR_INVALID, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(FW_ERR_DEVICE, "device %d not available", device);
    DeviceGuard g(device);
    FW_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return FW_OK;
}

}  // extern "C"

namespace {

// fw_sim_create; with `fixed` not null, a fixed-region handle on [fixed, fixed + fixed_bytes)
int create_impl(const fw_desc* desc, int device, void* fixed, size_t fixed_bytes, fw_sim** out) {
    if (!desc || !out) return fail(FW_ERR_INVALID, "null argument");
    *out = nullptr;
    const fw_desc& d = *desc;
    if (d.abi_version != FW_ABI_VERSION) return fail(FW_ERR_INVALID, "abi_version %d, library has %d", d.abi_version, FW_ABI_VERSION);
    if (d.ndim != 2 && d.ndim != 3) return fail(FW_ERR_INVALID, "ndim must be 2 or 3, got %d", d.ndim);
    if (d.n_relax != 2) return fail(FW_ERR_UNSUPPORTED, "only n_relax = 2 is supported, got %d", d.n_relax);
    const int64_t nz = d.ndim == 3 ? d.nz : 1;
    if (d.nx < 2 * fw::kHalo + 1 || d.ny < 2 * fw::kHalo + 1 || (d.ndim == 3 && nz < 2 * fw::kHalo + 1))
        return fail(FW_ERR_INVALID, "every axis needs at least %d cells (8 ghost cells per side)", 2 * fw::kHalo + 1);
    if (d.nt < 0 || d.ncoords < 0 || d.ncoordsout < 0 || d.ncoordszero < 0 || d.ntic < 0)
        return fail(FW_ERR_INVALID, "negative count in descriptor");
    if (d.modt < 1) return fail(FW_ERR_INVALID, "modt must be >= 1");
    if (d.ndmap < 1) return fail(FW_ERR_INVALID, "ndmap must be >= 1");
    if (d.attenuation_model != 0 && d.attenuation_model != 1)
        return fail(FW_ERR_INVALID, "attenuation_model must be 0 (relaxation) or 1 (exponential)");
    const int64_t nshots = d.n_shots < 1 ? 1 : d.n_shots;
    // launch grids carry the shots in z: 3D grids spend up to nx z blocks per shot
    if (nshots * (d.ndim == 3 ? d.nx : 1) > (int64_t)fw::kMaxGridZ)
        return fail(FW_ERR_UNSUPPORTED, "%lld shots of this grid exceed one launch; split the batch (at most %lld)",
                    (long long)nshots, (long long)(fw::kMaxGridZ / (d.ndim == 3 ? d.nx : 1)));
    if (!(d.dt > 0) || !(d.dx > 0) || !(d.dy > 0) || (d.ndim == 3 && !(d.dz > 0)))
        return fail(FW_ERR_INVALID, "dt and grid spacings must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(FW_ERR_DEVICE, "device %d not available (%d devices)", device, ndev);

    fw_sim* s = new fw_sim();
    s->d = d;
    s->d.nz = nz;
    s->device = device;
    s->ndim = d.ndim;
    s->ncell_c = cells(s->d);
    s->pitch = row_pitch(s->d);
    s->rows = d.ndim == 3 ? d.nx * d.ny : d.nx;
    s->row_len = d.ndim == 3 ? nz : d.ny;
    s->ncell = s->rows * s->pitch;
    s->nshots = (int)nshots;
    s->d.n_shots = (int32_t)nshots;
    DeviceGuard g(device);
    auto cleanup = [&](int rc) {
        fw_sim_destroy(s);
        return rc;
    };
    if (fixed) {
        s->fixed_base = (char*)fixed;
        s->fixed_size = fixed_bytes;
        s->stream = hipStreamPerThread;
    } else {
        if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(FW_ERR_DEVICE, "hipStreamCreate failed"));
        if (hipEventCreate(&s->ev[0]) != hipSuccess || hipEventCreate(&s->ev[1]) != hipSuccess)
            return cleanup(fail(FW_ERR_DEVICE, "hipEventCreate failed"));
    }
    const int nd = d.ndim;
    const bool relax = d.attenuation_model == 0;  // the exponential model has no memory variables
    const int64_t field = s->ncell * nshots;  // one wave field of every shot
    s->nstate = (int64_t)(1 + nd + (relax ? 4 * nd : 0)) * field;
    if (dev_malloc(s, (void**)&s->state, (size_t)s->nstate * 4) != hipSuccess)
        return cleanup(fail(FW_ERR_NOMEM, "hipMalloc of %lld state floats failed", (long long)s->nstate));
    if (fill_zero_sync(s, s->state, (size_t)s->nstate * 4) != hipSuccess)
        return cleanup(fail(FW_ERR_DEVICE, "hipMemset failed"));
    {  // the per-cell input maps in one allocation, whose slots alloc() hands out
        const size_t slot = ((size_t)s->ncell * 4 + 255) / 256 * 256;
        std::vector<std::string> grid_maps;
        for (const std::string& n : required_maps(d))
            if (n != "dmap" && n != "icmat" && n != "icc" && n != "outc" && n != "icczero") grid_maps.push_back(n);
        if (dev_malloc(s, &s->arena, slot * grid_maps.size()) != hipSuccess)
            return cleanup(fail(FW_ERR_NOMEM, "device allocation of %zu input maps failed", grid_maps.size()));
        for (size_t i = 0; i < grid_maps.size(); ++i) s->arena_slot[grid_maps[i]] = (char*)s->arena + i * slot;
    }
    float* f = s->state;
    s->ds.p = f;
    f += field;
    for (int a = 0; a < nd; ++a, f += field) s->ds.v[a] = f;
    for (int a = 0; a < nd && relax; ++a)
        for (int nu = 0; nu < 2; ++nu, f += field) s->ds.psi[a][nu] = f;
    for (int a = 0; a < nd && relax; ++a)
        for (int nu = 0; nu < 2; ++nu, f += field) s->ds.phi[a][nu] = f;
    s->ds.nshots = (int32_t)nshots;
    s->ds.shot_stride = s->ncell;
    s->ds.ndim = nd;
    s->ds.nx = d.nx;
    s->ds.ny = d.ny;
    s->ds.nz = nz;
    s->ds.sy = d.ndim == 3 ? s->pitch : 1;
    s->ds.sx = d.ndim == 3 ? d.ny * s->pitch : d.ny;
    s->ds.dt = (float)d.dt;
    *out = s;
    return FW_OK;
}

}  // namespace

extern "C" {

int fw_sim_create(const fw_desc* desc, int device, fw_sim** out) {
    return create_impl(desc, device, nullptr, 0, out);
}

int fw_sim_upload(fw_sim* s, const char* name_c, const void* host, size_t bytes) {
    if (!s || !name_c || (!host && bytes)) return fail(FW_ERR_INVALID, "null argument");
    const std::string name(name_c);
    const fw_desc& d = s->d;
    DeviceGuard g(s->device);
    if (s->ready && (name == "icc" || name == "outc" || name == "icczero" || name == "dmap"))
        return fail(FW_ERR_STATE, "'%s' cannot change after the first run; create a new handle", name_c);
    void* p = nullptr;
    int rc;
    if (name == "c" || name == "beta") {
        // c enters only through dcmap; beta (nonlinearity) is not modelled by this engine
        if (bytes != (size_t)s->ncell_c * 4) return fail(FW_ERR_INVALID, "'%s' must be %lld bytes", name_c, (long long)s->ncell_c * 4);
        s->have[name] = true;
        return FW_OK;
    }
    if (is_voxel_map(d, name)) {
        if (bytes != (size_t)s->ncell_c * 4)
            return fail(FW_ERR_INVALID, "'%s' must be %lld bytes, got %zu", name_c, (long long)s->ncell_c * 4, bytes);
        if ((rc = alloc(s, name, (size_t)s->ncell * 4, &p)) != FW_OK) return rc;
        if (s->pitch != s->row_len) FW_HIP(fill_zero_sync(s, p, (size_t)s->ncell * 4));  // row padding
        FW_HIP(copy_grid_in(s, p, host, 4));
        if (name.rfind("kappa", 0) == 0) {
            FW_HIP(fw::launch_reciprocal((float*)p, s->ncell, s->stream));
            FW_HIP(hipStreamSynchronize(s->stream));
        }
        s->have[name] = true;
        const bool relaxation_map = name.rfind("kappa", 0) == 0 || name.rfind("apml", 0) == 0 ||
                                    name.rfind("bpml", 0) == 0 || name == "dcmap";
        if (s->ready && relaxation_map && d.isotropic) return recode(s);
        return FW_OK;
    }
    if (name == "dmap") {
        if (bytes != (size_t)(9 * 2 * d.ndmap) * 4) return fail(FW_ERR_INVALID, "'dmap' must be 9*2*ndmap floats");
        s->dmap.assign((const float*)host, (const float*)host + 9 * 2 * d.ndmap);
        s->have[name] = true;
        return FW_OK;
    }
    if (name == "icc" || name == "outc" || name == "icczero") {
        const int64_t n = name == "icc" ? d.ncoords : name == "outc" ? d.ncoordsout : d.ncoordszero;
        if (bytes != (size_t)(n * s->ndim) * 4)
            return fail(FW_ERR_INVALID, "'%s' must be [%lld, %d] int32", name_c, (long long)n, s->ndim);
        auto& dst = name == "icc" ? s->icc : name == "outc" ? s->outc : s->zero;
        dst.assign((const int32_t*)host, (const int32_t*)host + n * s->ndim);
        s->have[name] = true;
        return FW_OK;
    }
    if (name == "icmat") {
        const int64_t per_shot = d.ncoords * d.ntic;
        if (bytes != (size_t)(s->nshots * per_shot) * 4)
            return fail(FW_ERR_INVALID, "'icmat' must be [n_shots, ncoords, ntic] float32 (%lld bytes)",
                        (long long)(s->nshots * per_shot * 4));
        // device layout [shot][ntic][ncoords] (one coalesced row per step): the caller's
        // source-major matrix is copied as it is and transposed on the device
        s->db.icmat_stride = per_shot;
        if ((rc = alloc(s, name, bytes, &p)) != FW_OK) return rc;
        if (bytes) {
            Stage src;
            hipError_t e = src.get(s, bytes);
            if (e == hipSuccess) e = h2d(s, src.p, host, bytes);
            if (e == hipSuccess)
                e = fw::launch_transpose_icmat((const float*)src.p, (float*)p, s->nshots, d.ncoords, d.ntic, s->stream);
            const hipError_t e2 = src.release(s);
            const hipError_t e3 = hipStreamSynchronize(s->stream);
            FW_HIP(e != hipSuccess ? e : e2 != hipSuccess ? e2 : e3);
        }
        s->db.icmat_t = (const float*)p;
        s->have[name] = true;
        return FW_OK;
    }
    return fail(FW_ERR_INVALID, "unknown input '%s' for a %dD %s simulation", name_c, s->ndim,
                d.isotropic ? "isotropic" : "anisotropic");
}

int fw_sim_run(fw_sim* s, int64_t n_steps) { return run_impl(s, n_steps, false, nullptr); }

int fw_sim_run_timed(fw_sim* s, int64_t n_steps, double* ms, int32_t* n_kinds) {
    if (!ms) return fail(FW_ERR_INVALID, "null ms");
    if (n_kinds) *n_kinds = fw::kNumKinds;
    return run_impl(s, n_steps, true, ms);
}

int fw_sim_reset(fw_sim* s) {
    if (!s) return fail(FW_ERR_INVALID, "null handle");
    DeviceGuard g(s->device);
    FW_HIP(hipMemsetAsync(s->state, 0, (size_t)s->nstate * 4, s->stream));
    if (s->db.rec && s->rec_cap)
        FW_HIP(hipMemsetAsync(s->db.rec, 0, (size_t)(s->nshots * s->rec_cap * s->d.ncoordsout) * 4, s->stream));
    FW_HIP(hipStreamSynchronize(s->stream));
    s->step = 0;
    s->state_zero = true;
    return FW_OK;
}

int fw_sim_step_count(const fw_sim* s, int64_t* steps) {
    if (!s || !steps) return fail(FW_ERR_INVALID, "null argument");
    *steps = s->step;
    return FW_OK;
}

int fw_sim_record_count(const fw_sim* s, int64_t* n) {
    if (!s || !n) return fail(FW_ERR_INVALID, "null argument");
    *n = s->d.modt > 0 ? (s->step + s->d.modt - 1) / s->d.modt : 0;
    return FW_OK;
}

int fw_sim_download_sensors(fw_sim* s, float* out, size_t bytes) {
    if (!s || (!out && bytes)) return fail(FW_ERR_INVALID, "null argument");
    int64_t nrec = 0;
    fw_sim_record_count(s, &nrec);
    const size_t row = (size_t)(nrec * s->d.ncoordsout) * 4;  // one shot's records
    const size_t want = row * (size_t)s->nshots;
    if (bytes != want)
        return fail(FW_ERR_INVALID, "sensor buffer must be %zu bytes (%d shots x %lld records), got %zu", want,
                    s->nshots, (long long)nrec, bytes);
    if (!want) return FW_OK;
    DeviceGuard g(s->device);
    FW_HIP(hipMemcpy2DAsync(out, row, s->db.rec, (size_t)s->db.rec_stride * 4, row, (size_t)s->nshots,
                            hipMemcpyDeviceToHost, s->stream));
    FW_HIP(hipStreamSynchronize(s->stream));
    return FW_OK;
}

int fw_sim_download_field(fw_sim* s, const char* name, float* out, size_t bytes) {
    if (!s || !name || !out) return fail(FW_ERR_INVALID, "null argument");
    const float* src = nullptr;
    if (!strcmp(name, "p")) src = s->ds.p;
    else if (!strcmp(name, "vx")) src = s->ds.v[0];
    else if (!strcmp(name, "vy")) src = s->ds.v[1];
    else if (!strcmp(name, "vz") && s->ndim == 3) src = s->ds.v[2];
    if (!src) return fail(FW_ERR_INVALID, "unknown field '%s'", name);
    const size_t one = (size_t)s->ncell_c * 4;
    if (bytes != one && bytes != one * (size_t)s->nshots)
        return fail(FW_ERR_INVALID, "field buffer must be %zu bytes (first shot) or %zu (every shot)", one,
                    one * (size_t)s->nshots);
    DeviceGuard g(s->device);
    for (size_t b = 0; b * one < bytes; ++b)
        FW_HIP(copy_grid_out(s, (char*)out + b * one, src + b * s->ncell, 4));
    return FW_OK;
}

int fw_sim_set_field(fw_sim* s, const char* name, const float* values, size_t bytes) {
    if (!s || !name || !values) return fail(FW_ERR_INVALID, "null argument");
    float* dst = nullptr;
    if (!strcmp(name, "p")) dst = s->ds.p;
    else if (!strcmp(name, "vx")) dst = s->ds.v[0];
    else if (!strcmp(name, "vy")) dst = s->ds.v[1];
    else if (!strcmp(name, "vz") && s->ndim == 3) dst = s->ds.v[2];
    if (!dst) return fail(FW_ERR_INVALID, "unknown field '%s'", name);
    const size_t one = (size_t)s->ncell_c * 4;
    if (bytes != one && bytes != one * (size_t)s->nshots)
        return fail(FW_ERR_INVALID, "field buffer must be %zu bytes (first shot) or %zu (every shot)", one,
                    one * (size_t)s->nshots);
    DeviceGuard g(s->device);
    for (size_t b = 0; b * one < bytes; ++b)
        FW_HIP(copy_grid_in(s, dst + b * s->ncell, (const char*)values + b * one, 4));
    if (!s->ready) s->fields_set = true;
    s->state_zero = false;
    return FW_OK;
}

int fw_sim_kernel_config(fw_sim* s, int32_t kind, int32_t* variant, int32_t* width, int32_t* chunk) {
    if (!s || !variant || !width || !chunk) return fail(FW_ERR_INVALID, "null argument");
    if (kind != 0 && kind != 1) return fail(FW_ERR_INVALID, "kind must be 0 (velocity) or 1 (pressure)");
    if (!s->ready) {
        DeviceGuard g(s->device);
        int rc = finalize(s);
        if (rc != FW_OK) return rc;
    }
    const fw::LaunchCfg& c = kind == 0 ? s->cfg_v : s->cfg_p;
    *variant = c.variant;
    *width = c.width;
    *chunk = c.chunk;
    return FW_OK;
}

int fw_sim_set_graph_steps(fw_sim* s, int64_t steps) {
    if (!s) return fail(FW_ERR_INVALID, "null handle");
    if (steps < 0 || steps > 100000) return fail(FW_ERR_INVALID, "graph steps must be in [0, 100000]");
    DeviceGuard g(s->device);
    if (s->graph) {
        FW_HIP(hipStreamSynchronize(s->stream));
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    s->graph_steps = steps;
    return FW_OK;
}

int fw_sim_set_relaxation_codes(fw_sim* s, int32_t enable) {
    if (!s) return fail(FW_ERR_INVALID, "null handle");
    if (enable != 0 && enable != 1) return fail(FW_ERR_INVALID, "enable must be 0 or 1");
    s->codes_mode = enable;
    if (!s->ready) return FW_OK;
    DeviceGuard g(s->device);
    return recode(s);
}

int fw_sim_relaxation_codes(fw_sim* s, int32_t* n_codes, double* coded_fraction) {
    if (!s || !n_codes) return fail(FW_ERR_INVALID, "null argument");
    if (!s->ready) {
        DeviceGuard g(s->device);
        int rc = finalize(s);
        if (rc != FW_OK) return rc;
    }
    *n_codes = s->ncodes;
    if (coded_fraction) *coded_fraction = s->ncell ? (double)s->coded_cells / (double)s->ncell : 0.0;
    return FW_OK;
}

int fw_sim_set_kernel_config(fw_sim* s, int32_t kind, int32_t variant, int32_t width, int32_t chunk) {
    if (!s) return fail(FW_ERR_INVALID, "null handle");
    if (kind != 0 && kind != 1) return fail(FW_ERR_INVALID, "kind must be 0 (velocity) or 1 (pressure)");
    if (s->ready) return fail(FW_ERR_STATE, "the kernels are fixed at the first run; set them before it");
    if (variant < 0 || variant > 4) return fail(FW_ERR_INVALID, "kernel variant must be 0 to 4");
    fw::LaunchCfg c;
    c.variant = variant;
    c.width = width;
    c.chunk = chunk;
    s->pin[kind] = c;
    s->pinned[kind] = true;
    return FW_OK;
}

int fw_sim_destroy(fw_sim* s) {
    if (!s) return FW_OK;
    DeviceGuard g(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    dev_free(s, s->step_dev);
    for (auto& kv : s->dev)
        if (!s->arena_slot.count(kv.first)) dev_free(s, kv.second);
    dev_free(s, s->arena);
    dev_free(s, s->state);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& e : s->ev)
        if (e) (void)hipEventDestroy(e);
    if (s->stream && !s->fixed_base) (void)hipStreamDestroy(s->stream);  // not the per-thread stream
    delete s;
    return FW_OK;
}

// ---- directory contract -----------------------------------------------------------------

static bool read_file(const std::string& path, std::vector<char>& buf) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return false;
    const std::streamsize n = f.tellg();
    f.seekg(0);
    buf.resize((size_t)n);
    return (bool)f.read(buf.data(), n) || n == 0;
}

static bool file_exists(const std::string& path) {
    std::ifstream f(path);
    return (bool)f;
}

}  // extern "C"

namespace fw {

int read_dir_desc(const char* sim_dir, fw_desc* out) {
    if (!sim_dir) return fail(FW_ERR_INVALID, "null directory");
    const std::string dir = std::string(sim_dir) + "/";
    std::vector<char> buf;
    auto scalar_i = [&](const char* stem, int64_t* v) -> int {
        if (!read_file(dir + stem + ".dat", buf) || buf.size() < 4) return fail(FW_ERR_IO, "cannot read %s%s.dat", dir.c_str(), stem);
        int32_t x;
        memcpy(&x, buf.data(), 4);
        *v = x;
        return FW_OK;
    };
    auto scalar_f = [&](const char* stem, double* v) -> int {
        if (!read_file(dir + stem + ".dat", buf) || buf.size() < 4) return fail(FW_ERR_IO, "cannot read %s%s.dat", dir.c_str(), stem);
        float x;
        memcpy(&x, buf.data(), 4);
        *v = x;
        return FW_OK;
    };
    fw_desc d{};
    d.abi_version = FW_ABI_VERSION;
    d.n_relax = 2;
    d.ndim = file_exists(dir + "nZ.dat") ? 3 : 2;
    d.isotropic = file_exists(dir + "kappaw.dat") ? 0 : 1;
    d.attenuation_model = file_exists(dir + "a_exp.dat") ? 1 : 0;
    int rc;
    if ((rc = scalar_i("nX", &d.nx)) || (rc = scalar_i("nY", &d.ny)) || (rc = scalar_i("nT", &d.nt)) ||
        (rc = scalar_i("ncoords", &d.ncoords)) || (rc = scalar_i("ncoordsout", &d.ncoordsout)) ||
        (rc = scalar_i("nTic", &d.ntic)) || (rc = scalar_i("modT", &d.modt)) || (rc = scalar_i("ndmap", &d.ndmap)) ||
        (rc = scalar_f("dX", &d.dx)) || (rc = scalar_f("dY", &d.dy)) || (rc = scalar_f("dT", &d.dt)) ||
        (rc = scalar_f("c0", &d.c0)))
        return rc;
    if (file_exists(dir + "ncoordszero.dat")) {
        if ((rc = scalar_i("ncoordszero", &d.ncoordszero))) return rc;
    }
    d.nz = 1;
    if (d.ndim == 3 && ((rc = scalar_i("nZ", &d.nz)) || (rc = scalar_f("dZ", &d.dz)))) return rc;
    *out = d;
    return FW_OK;
}

size_t fixed_footprint(const fw_desc& d) {
    // in floating point first: the sizes come from files and may be absurd
    const int nd = d.ndim == 3 ? 3 : 2;
    const double nz = nd == 3 ? (double)d.nz : 1.0;
    const double pitch = nd == 3 ? std::ceil(nz / 32.0) * 32.0 : (double)d.ny;
    const double rows = nd == 3 ? (double)d.nx * (double)d.ny : (double)d.nx;
    const double ncell = rows * pitch;
    const double nshots = d.n_shots < 1 ? 1.0 : (double)d.n_shots;
    const bool relax = d.attenuation_model == 0;
    double nmaps = 0;
    for (const std::string& n : required_maps(d))
        if (n != "dmap" && n != "icmat" && n != "icc" && n != "outc" && n != "icczero") nmaps += 1;
    const double rec = d.modt > 0 ? std::ceil((double)d.nt / (double)d.modt) : 0.0;
    const double icmat = nshots * (double)d.ncoords * (double)d.ntic * 4;
    double b = (1 + nd + (relax ? 4 * nd : 0)) * ncell * nshots * 4  // wave state
               + nmaps * ncell * 4                                   // input maps
               + icmat                                               // source matrix
               + (double)d.ncoords * 8 + (double)d.ncoordszero * 8 + (double)d.ncoordsout * 13  // index lists
               + (double)d.ndmap * nd * kNCoef * 4                   // coefficient table
               + nshots * rec * (double)d.ncoordsout * 4;            // records
    // largest temporary: the source matrix before its transpose, or a row-staging buffer
    const double stage = pitch != (nd == 3 ? nz : (double)d.ny) ? (double)kFixedStageBytes : 0.0;
    b += std::max(icmat, stage);
    b += (nmaps + 16) * 256;  // alignment of every array
    if (!(b < 1e18)) return (size_t)-1;
    return (size_t)b;
}

int run_dir_impl(const char* sim_dir, int device, void* fixed, size_t fixed_bytes) {
    fw_desc d{};
    int rc;
    if ((rc = read_dir_desc(sim_dir, &d)) != FW_OK) return rc;
    const std::string dir = std::string(sim_dir) + "/";
    std::vector<char> buf;
    fw_sim* s = nullptr;
    if ((rc = create_impl(&d, device, fixed, fixed_bytes, &s)) != FW_OK) return rc;
    std::vector<std::string> names = required_maps(d);
    for (const auto& name : names) {
        if (name == "icczero" && d.ncoordszero == 0) continue;
        if (!read_file(dir + name + ".dat", buf)) {
            fw_sim_destroy(s);
            return fail(FW_ERR_IO, "cannot read %s%s.dat", dir.c_str(), name.c_str());
        }
        if ((rc = fw_sim_upload(s, name.c_str(), buf.data(), buf.size())) != FW_OK) {
            std::string msg = std::string("file ") + name + ".dat: " + g_last_error;
            fw_sim_destroy(s);
            return fail(rc, "%s", msg.c_str());
        }
    }
    if ((rc = fw_sim_run(s, d.nt)) != FW_OK) {
        std::string msg = g_last_error;
        fw_sim_destroy(s);
        return fail(rc, "%s", msg.c_str());
    }
    int64_t nrec = 0;
    fw_sim_record_count(s, &nrec);
    std::vector<float> rec((size_t)(nrec * d.ncoordsout));
    if ((rc = fw_sim_download_sensors(s, rec.data(), rec.size() * 4)) != FW_OK) {
        std::string msg = g_last_error;
        fw_sim_destroy(s);
        return fail(rc, "%s", msg.c_str());
    }
    fw_sim_destroy(s);
    std::ofstream out(dir + "genout.dat", std::ios::binary | std::ios::trunc);
    if (!out) return fail(FW_ERR_IO, "cannot write %sgenout.dat", dir.c_str());
    out.write((const char*)rec.data(), (std::streamsize)(rec.size() * 4));
    if (!out) return fail(FW_ERR_IO, "short write to %sgenout.dat", dir.c_str());
    return FW_OK;
}

}  // namespace fw

extern "C" {

int fw_run_dir(const char* sim_dir, int device) { return fw::run_dir_impl(sim_dir, device, nullptr, 0); }

}  // extern "C"
