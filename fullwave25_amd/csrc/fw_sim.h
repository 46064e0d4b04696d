// The engine handle, shared by the forward (fw_api.cpp) and adjoint (fw_adjoint.cpp) host code.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../include/fullwave_hip.h"
#include "fw_internal.h"

struct fw_sim {
    fw_desc d{};

    int device = 0;
    hipStream_t stream = nullptr;
    int64_t ncell = 0;    // cells of one device grid (3D: z rows padded to `pitch`)
    int64_t ncell_c = 0;  // cells of the grid as the caller sees it (nx * ny * nz)
    int64_t pitch = 0;    // floats per contiguous row on the device (3D: nz rounded up to 32)
    int64_t rows = 0;     // contiguous rows (2D: nx, 3D: nx * ny)
    int64_t row_len = 0;  // cells per row as the caller sees it (2D: ny, 3D: nz)
    int ndim = 2;
    int nshots = 1;  // transmits sharing this medium, advanced together
    fw::DevSim ds{};
    fw::DevBoundary db{};
    float* state = nullptr;  // p, v[ndim], psi[ndim][2], phi[ndim][2]
    int64_t nstate = 0;
    std::map<std::string, void*> dev;  // every other device allocation, by input name
    std::set<void*> plain_alloc;       // allocations made with hipMalloc (see dev_malloc)
    void* arena = nullptr;             // every per-cell input map, one allocation
    std::map<std::string, void*> arena_slot;
    std::map<std::string, bool> have;
    std::vector<int32_t> icc, outc, zero;
    std::vector<float> dmap;
    bool ready = false;
    int64_t step = 0;
    int64_t rec_cap = 0;
    fw::LaunchCfg cfg_v, cfg_p;  // velocity and pressure kernels, chosen independently
    // 2D marching or tiled pressure kernels: the step's sources, pressure release and sensors
    // run inside the pressure launch (one entry per shot, points grouped by the kernel's tiles)
    bool fused_boundary = false;
    fw::GroupBoundary* fbound = nullptr;
    std::vector<fw::GroupBoundary> fbound_h;  // what fbound holds
    bool pinned[2] = {false, false};  // set by fw_sim_set_kernel_config: not autotuned
    fw::LaunchCfg pin[2];
    bool fields_set = false;  // a field was set before the first run: do not tune on it
    bool state_zero = true;   // every wave field is known to be zero (created or reset, not run since)
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<int64_t> out_lin_h;  // sensor voxels (linear indices), host copy
    // steady steps replayed from a captured HIP graph (graph_steps > 0): the step number then
    // lives in device memory (step_dev) and the graph advances it
    int64_t graph_steps = 0;
    int64_t* step_dev = nullptr;
    hipGraphExec_t graph = nullptr;
    // coded relaxation coefficients: -1 = as FULLWAVE_HIP_CODES says (default on), 0 off, 1 on;
    // ncodes = codes in use (0: the marching kernels read the per-cell maps), covering
    // coded_cells of the ncell device cells (the others read the per-cell maps)
    int codes_mode = -1;
    int32_t ncodes = 0;
    int64_t coded_cells = 0;
    // fixed-region handle (fw_run_dir_noalloc): every device array is carved out of
    // [fixed_base, fixed_base + fixed_size) in order and nothing is freed; no device allocation,
    // pinned host buffer, stream or event is created (the handle runs on the per-thread stream)
    char* fixed_base = nullptr;
    size_t fixed_size = 0;
    mutable size_t fixed_used = 0;
};


namespace fw {
// host helpers of fw_api.cpp used by the adjoint
int set_error(int code, const char* fmt, ...);
int prepare(fw_sim* s);                    // check inputs, resolve coordinates, pick kernels
int forward_step(fw_sim* s);               // one time step on the handle's stream
hipError_t grid_to_host(const fw_sim* s, void* dst, const void* src, size_t esize);  // padded -> compact
// Host-side set-up of a handle, without device access.
// Source, pressure-release and sensor points as one time step applies them.
struct BoundaryLists {
    std::vector<int64_t> src_lin;   // per source point: device index, or -1 (repeated later, or air)
    std::vector<int64_t> zero_lin;  // distinct pressure-release device indices, ascending
    std::vector<int64_t> out_lin;   // per sensor: device index
    std::vector<int32_t> out_src;   // per sensor: the source point last written there, or -1
    std::vector<int8_t> out_air;    // per sensor: 1 on a pressure-release point
};
// Coordinate lists (int32 [n, ndim]) of the grid of `d` to device indices (x stride sx, y stride
// sy, z stride 1; in 2D the y stride is 1 and sy is ignored); out-of-grid coordinates fail with
// FW_ERR_INVALID. Air wins over a source at the same point; a repeated source point keeps its last
// row; a sensor on a source or air point reads the written value.
int resolve_boundary(const fw_desc& d, int64_t sx, int64_t sy, const std::vector<int32_t>& icc,
                     const std::vector<int32_t>& outc, const std::vector<int32_t>& zero, BoundaryLists* out);
// The points of a 2D grid (linear indices `lin`, entries < 0 left out) by owning tile of a
// grouped launch (GroupBoundary): tiles of `tile_w` columns by `rows` rows from row kHalo, gx
// across; tile t holds (*idx)[(*off)[t] .. (*off)[t + 1]), indices into `lin` in list order. A
// point in the ghost border belongs to the tile of its row clamped to the updated rows.
void points_by_tile(int64_t nx, int64_t ny, unsigned gx, int tile_w, int rows, const std::vector<int64_t>& lin,
                    std::vector<int32_t>* off, std::vector<int32_t>* idx);
// The point table of one grid for a pressure launch that does the step's boundary work
// (GroupBoundary), appended to `tab` (int32 words; `tab` starts 16-byte aligned on the device):
// the word offsets of its parts and its tile count.
struct PointTable {
    size_t off, src, zero, out;
    int32_t ntiles;
};
PointTable point_table(int64_t nx, int64_t ny, unsigned gx, int tile_w, int rows, const BoundaryLists& bl,
                       std::vector<int32_t>* tab);
// Coefficient table [ndmap][ndim][9] (w_1..w_8, e) from dmap [9][2][ndmap] with each axis's spacing
// folded in; *axes_same = 1 when every axis has the same set for each sound speed.
std::vector<float> coefficient_table(const fw_desc& d, const std::vector<float>& dmap, int* axes_same);
// directory contract (fw_run_dir): the scalar files of sim_dir as a descriptor
int read_dir_desc(const char* sim_dir, fw_desc* d);
// device bytes a fixed-region handle of `d` needs (input maps, wave state, records, source
// matrix, index lists, coefficient table and the largest transient staging buffer)
size_t fixed_footprint(const fw_desc& d);
// fw_run_dir with every device array of the run in [fixed, fixed + fixed_bytes) when fixed is
// not null (the caller has checked fixed_footprint); else on ordinary device allocations
int run_dir_impl(const char* sim_dir, int device, void* fixed, size_t fixed_bytes);
}  // namespace fw
