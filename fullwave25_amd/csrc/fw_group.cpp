// Grouped time steps: several engine handles (different grids and media) advanced together.
// Every step is one velocity launch and one pressure launch over the tiles of all members
// (marching kernels in fw_kernels_stream.hip or tiled kernels in fw_kernels_tile2d.hip); each
// workgroup of the pressure launch also injects the sources, zeroes the pressure-release points
// and records the sensors of its own tile once the tile is updated (tile_boundary in
// fw_internal.h), so many small or uneven jobs fill the GPU as one job does, with two launches
// a step instead of three per job.
#include <hip/hip_runtime.h>

#include <set>
#include <vector>

#include "../../include/fullwave_hip.h"
#include "fw_internal.h"
#include "fw_sim.h"

struct fw_group {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<fw_sim*> sims;
    std::vector<fw::DevSim> views;  // each member's device view when the group was made
    std::vector<int> shots_of;        // boundary entries per member (its shots)
    std::vector<fw::GroupBoundary> bounds_h;
    fw::GroupTile* tiles = nullptr;
    fw::GroupBoundary* bounds = nullptr;
    int32_t* points = nullptr;  // every member's points grouped by tile (GroupBoundary lists)
    int ntiles = 0;
    uint32_t blocks = 0;
    int variant = 1;  // 1: marching kernels (width lanes, chunk rows); 2: tiled (256 x chunk)
    int width = 256, chunk = 16, max_ndmap = 1;
    int coef_sets = 1;      // LDS coefficient sets of the tiled kernels (largest member)
    bool all_coded = true;  // every member reads coded relaxation coefficients
    // steady steps replayed from a captured graph of graph_steps steps (0: plain launches); the
    // boundary pass adds the device counter base_dev, which each replay advances
    int64_t graph_steps = 0;
    int64_t* base_dev = nullptr;
    hipGraphExec_t graph = nullptr;
};

namespace fw {

void points_by_tile(int64_t nx, int64_t ny, unsigned gx, int tile_w, int rows, const std::vector<int64_t>& lin,
                    std::vector<int32_t>* off, std::vector<int32_t>* idx) {
    const unsigned gy = (unsigned)((nx - 2 * kHalo + rows - 1) / rows);
    const size_t ntiles = (size_t)gx * gy;
    std::vector<int64_t> owner(lin.size(), -1);
    off->assign(ntiles + 1, 0);
    for (size_t q = 0; q < lin.size(); ++q) {
        if (lin[q] < 0) continue;
        int64_t i = lin[q] / ny;
        const int64_t j = lin[q] % ny;
        i = i < kHalo ? kHalo : (i > nx - kHalo - 1 ? nx - kHalo - 1 : i);
        const int64_t t = ((i - kHalo) / rows) * (int64_t)gx + j / tile_w;
        owner[q] = t;
        ++(*off)[(size_t)t + 1];
    }
    for (size_t t = 0; t < ntiles; ++t) (*off)[t + 1] += (*off)[t];
    idx->assign((size_t)(*off)[ntiles], 0);
    std::vector<int32_t> fill(off->begin(), off->end() - 1);
    for (size_t q = 0; q < lin.size(); ++q)
        if (owner[q] >= 0) (*idx)[(size_t)fill[(size_t)owner[q]]++] = (int32_t)q;
}

PointTable point_table(int64_t nx, int64_t ny, unsigned gx, int tile_w, int rows, const BoundaryLists& bl,
                       std::vector<int32_t>* tab) {
    auto align4 = [&] {
        while (tab->size() % 4) tab->push_back(0);
    };
    auto push_lin = [&](int64_t lin) {
        tab->push_back((int32_t)(uint32_t)((uint64_t)lin & 0xffffffffu));
        tab->push_back((int32_t)(uint32_t)((uint64_t)lin >> 32));
    };
    std::vector<int32_t> off[3], idx[3];
    const std::vector<int64_t>* lists[3] = {&bl.src_lin, &bl.zero_lin, &bl.out_lin};
    for (int k = 0; k < 3; ++k) points_by_tile(nx, ny, gx, tile_w, rows, *lists[k], &off[k], &idx[k]);
    PointTable pt{};
    pt.ntiles = (int32_t)(off[0].size() - 1);
    align4();
    pt.off = tab->size();
    for (int k = 0; k < 3; ++k) tab->insert(tab->end(), off[k].begin(), off[k].end());
    align4();
    pt.src = tab->size();
    for (int32_t q : idx[0]) {  // {source row, 0, voxel}
        tab->push_back(q);
        tab->push_back(0);
        push_lin(bl.src_lin[(size_t)q]);
    }
    pt.zero = tab->size();
    for (int32_t q : idx[1]) push_lin(bl.zero_lin[(size_t)q]);
    align4();
    pt.out = tab->size();
    for (int32_t q : idx[2]) {  // {sensor column, source row at the voxel, voxel or -1 on a release point}
        tab->push_back(q);
        tab->push_back(bl.out_src[(size_t)q]);
        push_lin(bl.out_air[(size_t)q] ? -1 : bl.out_lin[(size_t)q]);
    }
    return pt;
}

}  // namespace fw

namespace {

#define GR_HIP(expr)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fw::set_error(FW_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct Guard {
    int prev = -1;
    explicit Guard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        (void)hipSetDevice(dev);
    }
    ~Guard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// the group's tables are stream-ordered allocations on its own stream: freeing them does not
// wait for the whole device (other handles' work)
void free_table(fw_group* g, void* p) {
    if (!p) return;
    if (hipFreeAsync(p, g->stream) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(g->stream);
        (void)hipFree(p);
    }
}

hipError_t alloc_table(fw_group* g, void** p, size_t bytes) {
    hipError_t e = hipMallocAsync(p, bytes, g->stream);
    if (e == hipSuccess) return e;
    (void)hipGetLastError();
    return hipMalloc(p, bytes);
}

void release(fw_group* g) {
    if (!g) return;
    if (g->graph) (void)hipGraphExecDestroy(g->graph);
    if (g->stream) {
        free_table(g, g->tiles);
        free_table(g, g->bounds);
        free_table(g, g->points);
        free_table(g, g->base_dev);
        (void)hipStreamSynchronize(g->stream);
        (void)hipStreamDestroy(g->stream);
    }
    delete g;
}

int build(fw_group* g) {
    std::vector<fw::GroupTile> tiles;
    std::vector<fw::GroupBoundary> bounds;
    std::vector<int32_t> pts;
    std::vector<fw::PointTable> slots;  // per bounds entry
    uint64_t blocks = 0;
    for (fw_sim* s : g->sims) {
        unsigned gx, gy;
        int64_t chunk = g->chunk;
        if (g->variant == 2) {
            gx = fw::group_tiles_x(s->ds.ny, 256);
            gy = fw::group_tile_rows(s->ds.nx, g->chunk);
        } else {
            gx = fw::group_tiles_x(s->ds.ny, g->width);
            // a grid too small to give ~128 workgroups at the group's row count takes fewer rows
            // per workgroup (no fewer than 2), so that its tiles do not march alone at the end
            const int64_t rows = s->ds.nx - 2 * fw::kHalo;
            chunk = rows * gx / 128;
            if (chunk > g->chunk) chunk = g->chunk;
            if (chunk < 2) chunk = 2;
            gy = fw::group_tiles_y(s->ds.nx, (int)chunk);
        }
        // the member's points by owning tile, shared by its shots
        fw::BoundaryLists bl;
        int rc = fw::resolve_boundary(s->d, s->ds.sx, s->ds.sy, s->icc, s->outc, s->zero, &bl);
        if (rc != FW_OK) return rc;
        const int tile_w = g->variant == 2 ? 256 : g->width;
        const fw::PointTable ps = fw::point_table(s->ds.nx, s->ds.ny, gx, tile_w, (int)chunk, bl, &pts);
        for (int b = 0; b < s->nshots; ++b) {
            fw::GroupTile t{};
            t.s = fw::shot_view(s->ds, b);
            t.start = (uint32_t)blocks;
            t.gx = gx;
            t.chunk = (int32_t)chunk;
            tiles.push_back(t);
            blocks += (uint64_t)gx * gy;
            fw::GroupBoundary e{};
            e.p = t.s.p;
            e.b = s->db;
            e.b.icmat_t = s->db.icmat_t + (int64_t)b * s->db.icmat_stride;
            e.b.rec = s->db.rec + (int64_t)b * s->db.rec_stride;
            e.modt = s->d.modt;
            e.step0 = s->step;
            bounds.push_back(e);
            slots.push_back(ps);
        }
        if (s->ds.ndmap > g->max_ndmap) g->max_ndmap = s->ds.ndmap;
        const int sets = s->ds.ndmap * (s->ds.coef_axes_same ? 1 : 2);
        if (sets > g->coef_sets) g->coef_sets = sets;
        if (!s->ds.rcode) g->all_coded = false;
        g->shots_of.push_back(s->nshots);
    }
    if (blocks >= (1ull << 32)) return fw::set_error(FW_ERR_INVALID, "the group has too many tiles for one launch");
    if (bounds.size() > fw::kMaxGridZ) return fw::set_error(FW_ERR_INVALID, "a group holds at most 65535 grids");
    if (pts.size() >= ((size_t)1 << 31)) return fw::set_error(FW_ERR_INVALID, "the group has too many points");
    g->blocks = (uint32_t)blocks;
    g->ntiles = (int)tiles.size();
    GR_HIP(alloc_table(g, (void**)&g->tiles, tiles.size() * sizeof(fw::GroupTile)));
    GR_HIP(alloc_table(g, (void**)&g->bounds, bounds.size() * sizeof(fw::GroupBoundary)));
    GR_HIP(alloc_table(g, (void**)&g->points, pts.size() * sizeof(int32_t)));
    for (size_t k = 0; k < bounds.size(); ++k) {
        const fw::PointTable& ps = slots[k];
        fw::GroupBoundary& e = bounds[k];
        e.ntiles = ps.ntiles;
        e.off = g->points + ps.off;
        e.src = reinterpret_cast<const int4*>(g->points + ps.src);
        e.zero = reinterpret_cast<const int64_t*>(g->points + ps.zero);
        e.out = reinterpret_cast<const int4*>(g->points + ps.out);
    }
    GR_HIP(hipMemcpyAsync(g->points, pts.data(), pts.size() * sizeof(int32_t), hipMemcpyHostToDevice, g->stream));
    GR_HIP(hipMemcpyAsync(g->tiles, tiles.data(), tiles.size() * sizeof(fw::GroupTile), hipMemcpyHostToDevice,
                          g->stream));
    GR_HIP(hipMemcpyAsync(g->bounds, bounds.data(), bounds.size() * sizeof(fw::GroupBoundary), hipMemcpyHostToDevice,
                          g->stream));
    GR_HIP(hipStreamSynchronize(g->stream));
    g->bounds_h = bounds;
    return FW_OK;
}

void free_tables(fw_group* g) {
    free_table(g, g->tiles);
    free_table(g, g->bounds);
    free_table(g, g->points);
    g->tiles = nullptr;
    g->bounds = nullptr;
    g->points = nullptr;
}

// The velocity launch, or the pressure launch with the boundary work of iteration `it` of a run
// (base: the device counter a graph replay adds; boundary == false: the stencil alone, as when
// shapes are timed).
hipError_t launch_stencil(const fw_group* g, bool velocity, bool boundary = false, const int64_t* base = nullptr,
                          int64_t it = 0) {
    const fw::GroupStep gs{boundary ? g->bounds : nullptr, base, it};
    if (g->variant == 2)
        return fw::launch_group_tile2d(g->tiles, g->ntiles, g->blocks, g->chunk, g->coef_sets, g->all_coded, velocity,
                                       gs, g->stream);
    return fw::launch_group_2d(g->tiles, g->ntiles, g->blocks, g->width, g->max_ndmap, velocity, gs, g->stream);
}

// Milliseconds of `reps` grouped velocity + pressure launches with the current tables.
int time_shape(fw_group* g, int reps, float* ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    GR_HIP(hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    auto launch = [&] {
        hipError_t r = launch_stencil(g, true);
        return r != hipSuccess ? r : launch_stencil(g, false);
    };
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e0, g->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e1, g->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    GR_HIP(e);
    return FW_OK;
}

bool tiles_supported(const fw_group* g) {
    for (fw_sim* s : g->sims)
        if (!fw::tile2d_supported(s->ds)) return false;
    return true;
}

// Pick the group's kernel family, lanes per workgroup and rows per workgroup by timing the
// candidates (those left open: `tune_variant`, `tune_width`, `tune_chunk`), when every member's
// wave fields are known to be zero (the update then leaves them zero, so the timing launches
// change nothing); otherwise the marching kernels at 256 lanes and 16 rows, or the tiled ones
// at 16 rows.
int choose_shape(fw_group* g, bool tune_variant, bool tune_width, bool tune_chunk) {
    bool zero = true;
    for (fw_sim* s : g->sims) zero = zero && s->state_zero;
    if (!tune_variant && !tune_width && !tune_chunk) return build(g);
    if (!zero) {
        if (tune_variant) g->variant = 1;
        if (tune_width) g->width = 256;
        if (tune_chunk) g->chunk = 16;
        return build(g);
    }
    struct Shape {
        int variant, width, chunk;
    };
    std::vector<Shape> cands;
    const int v0 = g->variant, w0 = g->width, c0 = g->chunk;
    if (tune_variant || v0 == 1)
        for (int w : {64, 128, 256})
            for (int c : {8, 16, 32})
                if ((tune_width || w == w0) && (tune_chunk || c == c0)) cands.push_back({1, w, c});
    if ((tune_variant && tiles_supported(g)) || v0 == 2)
        for (int c : {8, 16, 32})
            if (tune_chunk || c == c0) cands.push_back({2, 256, c});
    float best = 1e30f;
    Shape pick = cands.empty() ? Shape{1, 256, 16} : cands[0];
    for (const Shape& c : cands) {
        g->variant = c.variant;
        g->width = c.width;
        g->chunk = c.chunk;
        int rc = build(g);
        float ms = 0.f;
        if (rc == FW_OK) rc = time_shape(g, 3, &ms);
        free_tables(g);
        if (rc != FW_OK) return rc;
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    g->variant = pick.variant;
    g->width = pick.width;
    g->chunk = pick.chunk;
    return build(g);
}

// Capture graph_steps group steps (velocity and pressure launches, the pressure launch's
// boundary work reading the device counter) and one counter advance into a graph, once per group.
int capture_graph(fw_group* g) {
    if (!g->base_dev) GR_HIP(alloc_table(g, (void**)&g->base_dev, sizeof(int64_t)));
    hipGraph_t graph = nullptr;
    GR_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int64_t j = 0; j < g->graph_steps && e == hipSuccess; ++j) {
        e = launch_stencil(g, true);
        if (e == hipSuccess) e = launch_stencil(g, false, true, g->base_dev, j);
    }
    if (e == hipSuccess) e = fw::launch_advance(g->base_dev, g->graph_steps, g->stream);
    const hipError_t e2 = hipStreamEndCapture(g->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fw::set_error(FW_ERR_DEVICE, "group graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    e = hipGraphInstantiate(&g->graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        g->graph = nullptr;
        return fw::set_error(FW_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    return FW_OK;
}

}  // namespace

extern "C" {

int fw_group_set_graph_steps(fw_group* g, int64_t steps) {
    if (!g) return fw::set_error(FW_ERR_INVALID, "null group");
    if (steps < 0) return fw::set_error(FW_ERR_INVALID, "graph steps must be >= 0");
    if (steps != g->graph_steps && g->graph) {
        Guard guard(g->device);
        (void)hipStreamSynchronize(g->stream);
        (void)hipGraphExecDestroy(g->graph);
        g->graph = nullptr;
    }
    g->graph_steps = steps;
    return FW_OK;
}

int fw_group_create_shaped(fw_sim* const* sims, int32_t n, int32_t variant, int32_t width, int32_t chunk,
                           fw_group** out) {
    if (!sims || !out || n < 1) return fw::set_error(FW_ERR_INVALID, "null argument or empty group");
    *out = nullptr;
    if (variant < 0 || variant > 2)
        return fw::set_error(FW_ERR_INVALID, "group variant must be 1 (marching), 2 (tiled) or 0 (timed)");
    if (variant == 2) {
        if (width != 0 && width != 256) return fw::set_error(FW_ERR_INVALID, "tiled groups are 256 columns wide");
        if (chunk != 0 && chunk != 8 && chunk != 16 && chunk != 32)
            return fw::set_error(FW_ERR_INVALID, "tiled group rows must be 8, 16 or 32 (0: timed)");
    } else {
        if (width != 0 && width != 64 && width != 128 && width != 256)
            return fw::set_error(FW_ERR_INVALID, "group width must be 64, 128 or 256 (0: timed)");
        if (chunk != 0 && (chunk < 2 || chunk > 4096))
            return fw::set_error(FW_ERR_INVALID, "group chunk must be in [2, 4096] (0: timed)");
        if (variant == 0 && (width != 0 || chunk != 0))
            return fw::set_error(FW_ERR_INVALID, "a timed kernel family takes a timed width and chunk (0, 0)");
    }
    std::set<fw_sim*> seen;
    for (int32_t i = 0; i < n; ++i) {
        fw_sim* s = sims[i];
        if (!s) return fw::set_error(FW_ERR_INVALID, "null handle in the group");
        if (!seen.insert(s).second) return fw::set_error(FW_ERR_INVALID, "a handle appears twice in the group");
        if (s->device != sims[0]->device) return fw::set_error(FW_ERR_INVALID, "group members must share one device");
        if (s->ndim != 2 || s->d.attenuation_model != 0 || !s->d.isotropic)
            return fw::set_error(FW_ERR_UNSUPPORTED, "groups take 2D isotropic relaxation simulations");
    }
    Guard guard(sims[0]->device);
    for (int32_t i = 0; i < n; ++i) {
        int rc = fw::prepare(sims[i]);
        if (rc != FW_OK) return rc;
        if (!fw::streaming_supported(sims[i]->ds))
            return fw::set_error(FW_ERR_UNSUPPORTED, "group member %d needs the per-cell kernels", (int)i);
        if (variant == 2 && !fw::tile2d_supported(sims[i]->ds))
            return fw::set_error(FW_ERR_UNSUPPORTED,
                                 "group member %d cannot run the tiled kernels (ny %% 4 != 0 or more than 768 sound "
                                 "speeds)",
                                 (int)i);
    }
    fw_group* g = new fw_group();
    g->device = sims[0]->device;
    g->variant = variant == 0 ? 1 : variant;
    g->width = variant == 2 ? 256 : width;
    g->chunk = chunk;
    g->sims.assign(sims, sims + n);
    for (fw_sim* s : g->sims) g->views.push_back(s->ds);
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        release(g);
        return fw::set_error(FW_ERR_DEVICE, "hipStreamCreateWithFlags failed: %s", hipGetErrorString(e));
    }
    int rc = choose_shape(g, variant == 0, variant != 2 && width == 0, chunk == 0);
    if (rc != FW_OK) {
        release(g);
        return rc;
    }
    *out = g;
    return FW_OK;
}

int fw_group_create(fw_sim* const* sims, int32_t n, int32_t width, int32_t chunk, fw_group** out) {
    return fw_group_create_shaped(sims, n, 1, width, chunk, out);
}

int fw_group_config(const fw_group* g, int32_t* width, int32_t* chunk) {
    if (!g || !width || !chunk) return fw::set_error(FW_ERR_INVALID, "null argument");
    *width = g->width;
    *chunk = g->chunk;
    return FW_OK;
}

int fw_group_shape(const fw_group* g, int32_t* variant, int32_t* width, int32_t* chunk) {
    if (!g || !variant || !width || !chunk) return fw::set_error(FW_ERR_INVALID, "null argument");
    *variant = g->variant;
    *width = g->width;
    *chunk = g->chunk;
    return FW_OK;
}

int fw_group_run(fw_group* g, int64_t n_steps) {
    if (!g) return fw::set_error(FW_ERR_INVALID, "null group");
    if (n_steps < 0) return fw::set_error(FW_ERR_INVALID, "n_steps < 0");
    size_t e = 0;
    for (size_t i = 0; i < g->sims.size(); ++i) {
        fw_sim* s = g->sims[i];
        const fw::DevSim& v = g->views[i];
        if (s->ds.rcode != v.rcode || s->ds.rtab_v != v.rtab_v || s->ds.rtab_p != v.rtab_p || s->ds.p != v.p)
            return fw::set_error(FW_ERR_STATE, "group member %d changed since the group was made", (int)i);
        if (s->step + n_steps > s->d.nt)
            return fw::set_error(FW_ERR_INVALID, "group member %d has %lld steps left", (int)i,
                                 (long long)(s->d.nt - s->step));
        for (int b = 0; b < g->shots_of[i]; ++b) g->bounds_h[e++].step0 = s->step;  // members may differ
    }
    Guard guard(g->device);
    for (fw_sim* s : g->sims) GR_HIP(hipStreamSynchronize(s->stream));  // members idle
    GR_HIP(hipMemcpyAsync(g->bounds, g->bounds_h.data(), g->bounds_h.size() * sizeof(fw::GroupBoundary),
                          hipMemcpyHostToDevice, g->stream));
    int64_t done = 0;
    if (g->graph_steps > 0 && n_steps >= g->graph_steps) {
        if (!g->graph) {
            int rc = capture_graph(g);
            if (rc != FW_OK) return rc;
        }
        GR_HIP(fw::launch_tick(g->base_dev, 0, g->stream));  // replays count from this call's start
        for (; done + g->graph_steps <= n_steps; done += g->graph_steps) GR_HIP(hipGraphLaunch(g->graph, g->stream));
    }
    for (int64_t it = done; it < n_steps; ++it) {
        GR_HIP(launch_stencil(g, true));
        GR_HIP(launch_stencil(g, false, true, nullptr, it));
    }
    GR_HIP(hipStreamSynchronize(g->stream));
    for (fw_sim* s : g->sims) {
        s->step += n_steps;
        if (n_steps > 0) s->state_zero = false;
    }
    return FW_OK;
}

int fw_group_destroy(fw_group* g) {
    if (!g) return FW_OK;
    Guard guard(g->device);
    release(g);
    return FW_OK;
}

}  // extern "C"
