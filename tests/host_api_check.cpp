// Host-side checks of the engine library under AddressSanitizer and UndefinedBehaviorSanitizer,
// without a GPU (built and run by tests/test_sanitize.py). Covers what the C ABI computes on the
// host before any device call: descriptor validation, the directory contract's scalar files and
// the allocation-free run's size check, coordinate-to-index resolution of sources, air points and
// sensors (repeats, overlaps, the grid's corners, out-of-grid points), the coefficient table, and
// the grouping of those points by the tile of a grouped launch that owns them, and the point table
// the pressure launches read.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "../include/fullwave_hip.h"
#include "../fullwave25_amd/csrc/fw_sim.h"

static int failures = 0;
#define CHECK(cond)                                                                   \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                               \
        }                                                                             \
    } while (0)

static fw_desc base_desc(int ndim) {
    fw_desc d{};
    d.abi_version = FW_ABI_VERSION;
    d.ndim = ndim;
    d.isotropic = 1;
    d.n_relax = 2;
    d.nx = 37;
    d.ny = 29;
    d.nz = ndim == 3 ? 23 : 1;
    d.nt = 10;
    d.dx = d.dy = d.dz = 4.3e-5;
    d.dt = 1.1e-8;
    d.c0 = 1540;
    d.modt = 1;
    d.ndmap = 3;
    return d;
}

static void descriptors() {
    fw_sim* h = nullptr;
    fw_desc d = base_desc(2);
    d.ndim = 4;
    CHECK(fw_sim_create(&d, 0, &h) == FW_ERR_INVALID && !h);
    d = base_desc(2);
    d.n_relax = 3;
    CHECK(fw_sim_create(&d, 0, &h) == FW_ERR_UNSUPPORTED);
    d = base_desc(3);
    d.nz = 16;
    CHECK(fw_sim_create(&d, 0, &h) == FW_ERR_INVALID);
    d = base_desc(2);
    d.modt = 0;
    CHECK(fw_sim_create(&d, 0, &h) == FW_ERR_INVALID);
    d = base_desc(2);
    d.dt = -1;
    CHECK(fw_sim_create(&d, 0, &h) == FW_ERR_INVALID);
    d = base_desc(3);
    d.nx = 100000;
    d.n_shots = 1000;
    CHECK(fw_sim_create(&d, 0, &h) == FW_ERR_UNSUPPORTED);
    CHECK(fw_sim_create(nullptr, 0, &h) == FW_ERR_INVALID);
    CHECK(std::strlen(fw_last_error()) > 0);
    CHECK(fw_sim_run(nullptr, 1) == FW_ERR_INVALID);
    CHECK(fw_sim_destroy(nullptr) == FW_OK);
    CHECK(fw_run_dir(nullptr, 0) == FW_ERR_INVALID);
    CHECK(fw_run_dir_noalloc(nullptr, 0) == FW_ERR_INVALID);
    CHECK(fw_group_destroy(nullptr) == FW_OK);
}

static void write_i32(const std::string& p, int32_t v) {
    std::ofstream(p, std::ios::binary).write((const char*)&v, 4);
}
static void write_f32(const std::string& p, float v) {
    std::ofstream(p, std::ios::binary).write((const char*)&v, 4);
}

static void directory(const char* tmp) {
    const std::string dir = std::string(tmp) + "/";
    const char* ints[] = {"nX", "nY", "nT", "ncoords", "ncoordsout", "nTic", "modT", "ndmap"};
    const int32_t vals[] = {40000, 40000, 100, 1, 1, 100, 1, 1};
    for (int i = 0; i < 8; ++i) write_i32(dir + ints[i] + ".dat", vals[i]);
    for (const char* f : {"dX", "dY", "dT", "c0"}) write_f32(dir + f + ".dat", 1e-4f);
    fw_desc d{};
    CHECK(fw::read_dir_desc(tmp, &d) == FW_OK && d.nx == 40000 && d.ndim == 2 && d.isotropic == 1);
    // 1.6 G cells: far beyond the allocation-free array, refused before any device call
    CHECK(fw::fixed_footprint(d) > (size_t)FW_NOALLOC_SCRATCH_BYTES);
    CHECK(fw_run_dir_noalloc(tmp, 0) == FW_ERR_UNSUPPORTED);
    // absurd sizes saturate instead of overflowing
    d.nx = d.ny = (int64_t)1 << 40;
    CHECK(fw::fixed_footprint(d) == (size_t)-1);
    // a small run fits
    d = base_desc(3);
    d.ncoords = 4;
    d.ntic = 10;
    d.ncoordsout = 7;
    CHECK(fw::fixed_footprint(d) < (size_t)FW_NOALLOC_SCRATCH_BYTES);
    // a truncated scalar file
    std::ofstream(dir + "nX.dat", std::ios::binary | std::ios::trunc);
    CHECK(fw::read_dir_desc(tmp, &d) == FW_ERR_IO);
    CHECK(fw_run_dir((dir + "missing").c_str(), 0) == FW_ERR_IO);
}

// resolve_boundary on random point lists against the semantics written out directly
static void boundary(int ndim, unsigned seed) {
    std::mt19937 rng(seed);
    fw_desc d = base_desc(ndim);
    const int64_t nz = ndim == 3 ? d.nz : 1;
    const int64_t pitch = ndim == 3 ? (d.nz + 31) / 32 * 32 : d.ny;
    const int64_t sy = ndim == 3 ? pitch : 1, sx = ndim == 3 ? d.ny * pitch : d.ny;
    auto point = [&](std::vector<int32_t>& v) {
        // corners and edges often, so the first and last cells of every axis are covered
        const int64_t dims[3] = {d.nx, d.ny, nz};
        for (int a = 0; a < ndim; ++a) {
            const unsigned r = rng() % 8;
            v.push_back((int32_t)(r == 0 ? 0 : r == 1 ? dims[a] - 1 : (int64_t)(rng() % dims[a])));
        }
    };
    std::vector<int32_t> icc, outc, zero;
    d.ncoords = 60;
    d.ncoordsout = 50;
    d.ncoordszero = 30;
    for (int q = 0; q < d.ncoords; ++q) point(icc);
    for (int q = 0; q < 10; ++q) icc[(size_t)(ndim * (rng() % d.ncoords))] = icc[0];  // repeats
    for (int q = 0; q < d.ncoordsout; ++q) {
        if (q % 3 == 0) {  // on a source point
            const size_t s = (size_t)(rng() % d.ncoords) * ndim;
            outc.insert(outc.end(), icc.begin() + s, icc.begin() + s + ndim);
        } else {
            point(outc);
        }
    }
    for (int q = 0; q < d.ncoordszero; ++q) {
        if (q % 2 == 0) {  // on a source point
            const size_t s = (size_t)(rng() % d.ncoords) * ndim;
            zero.insert(zero.end(), icc.begin() + s, icc.begin() + s + ndim);
        } else {
            point(zero);
        }
    }
    fw::BoundaryLists bl;
    CHECK(fw::resolve_boundary(d, sx, sy, icc, outc, zero, &bl) == FW_OK);
    auto lin = [&](const std::vector<int32_t>& v, int64_t q) {
        int64_t idx = (int64_t)v[(size_t)(q * ndim)] * sx;
        if (ndim == 2) return idx + v[(size_t)(q * ndim + 1)];
        return idx + (int64_t)v[(size_t)(q * ndim + 1)] * sy + v[(size_t)(q * ndim + 2)];
    };
    const int64_t ncell = (ndim == 3 ? d.nx * d.ny : d.nx) * pitch;
    std::set<int64_t> air;
    for (int64_t q = 0; q < d.ncoordszero; ++q) air.insert(lin(zero, q));
    CHECK(bl.zero_lin.size() == air.size());
    for (size_t i = 0; i < bl.zero_lin.size(); ++i) CHECK(air.count(bl.zero_lin[i]) && (i == 0 || bl.zero_lin[i - 1] < bl.zero_lin[i]));
    for (int64_t q = 0; q < d.ncoords; ++q) {
        const int64_t l = lin(icc, q);
        bool later = false;
        for (int64_t r = q + 1; r < d.ncoords; ++r) later |= lin(icc, r) == l;
        CHECK(bl.src_lin[(size_t)q] == ((later || air.count(l)) ? -1 : l));
        CHECK(l >= 0 && l < ncell);
    }
    for (int64_t q = 0; q < d.ncoordsout; ++q) {
        const int64_t l = lin(outc, q);
        CHECK(bl.out_lin[(size_t)q] == l);
        CHECK(bl.out_air[(size_t)q] == (air.count(l) ? 1 : 0));
        int32_t last = -1;
        for (int64_t r = 0; r < d.ncoords; ++r)
            if (lin(icc, r) == l) last = (int32_t)r;
        CHECK(bl.out_src[(size_t)q] == last);
    }
    // a coordinate outside the grid, on each axis
    for (int a = 0; a < ndim; ++a) {
        std::vector<int32_t> bad = outc;
        bad[(size_t)(7 * ndim + a)] = a == 0 ? (int32_t)d.nx : -1;
        CHECK(fw::resolve_boundary(d, sx, sy, icc, bad, zero, &bl) == FW_ERR_INVALID);
    }
    // a list shorter than its count
    std::vector<int32_t> short_list(icc.begin(), icc.begin() + ndim * 3);
    CHECK(fw::resolve_boundary(d, sx, sy, short_list, outc, zero, &bl) == FW_ERR_INVALID);
}

static void coefficients(int ndim) {
    fw_desc d = base_desc(ndim);
    d.ndmap = 5;
    std::vector<float> dmap((size_t)(9 * 2 * d.ndmap));
    for (size_t i = 0; i < dmap.size(); ++i) dmap[i] = 0.01f * (float)(i + 1);
    int same = -1;
    std::vector<float> coef = fw::coefficient_table(d, dmap, &same);
    CHECK(coef.size() == (size_t)(d.ndmap * ndim * 9) && same == 1);
    for (int64_t m = 0; m < d.ndmap; ++m)
        for (int k = 1; k <= 8; ++k)
            CHECK(coef[(size_t)((m * ndim) * 9 + k - 1)] == (float)((double)dmap[(size_t)((k * 2) * d.ndmap + m)] / d.dx));
    d.dy = 1.25 * d.dx;
    coef = fw::coefficient_table(d, dmap, &same);
    CHECK(same == 0);
    dmap.resize(10);  // shorter than 9 x 2 x ndmap
    CHECK(fw::coefficient_table(d, dmap, &same).empty());
}

// The points of a grouped launch by owning tile: every point in exactly one tile, the one whose
// columns hold its column and whose rows hold its row (a ghost-border row taken to the nearest
// updated row), each tile's points in list order, skipped entries (< 0) in none.
static void point_tiles(unsigned seed) {
    std::mt19937 rng(seed);
    const int64_t nx = 17 + (int64_t)(rng() % 300), ny = 17 + (int64_t)(rng() % 700);
    const int widths[] = {64, 128, 256};
    const int tile_w = widths[rng() % 3];
    const int rows = 2 + (int)(rng() % 31);
    const unsigned gx = (unsigned)((ny + tile_w - 1) / tile_w);
    const int64_t gy = (nx - 16 + rows - 1) / rows;
    std::vector<int64_t> lin(rng() % 400);
    for (auto& l : lin) l = (rng() % 7 == 0) ? -1 : (int64_t)(rng() % (uint64_t)(nx * ny));
    std::vector<int32_t> off, idx;
    fw::points_by_tile(nx, ny, gx, tile_w, rows, lin, &off, &idx);
    CHECK(off.size() == (size_t)(gx * gy + 1) && off[0] == 0);
    size_t valid = 0;
    for (int64_t l : lin) valid += l >= 0;
    CHECK((size_t)off.back() == valid && idx.size() == valid);
    std::vector<int> seen(lin.size(), 0);
    for (size_t t = 0; t + 1 < off.size(); ++t) {
        CHECK(off[t] <= off[t + 1]);
        const int64_t bx = (int64_t)(t % gx), by = (int64_t)(t / gx);
        for (int32_t k = off[t]; k < off[t + 1]; ++k) {
            const int32_t q = idx[(size_t)k];
            CHECK(q >= 0 && (size_t)q < lin.size() && lin[(size_t)q] >= 0);
            if (q < 0 || (size_t)q >= lin.size() || lin[(size_t)q] < 0) continue;
            ++seen[(size_t)q];
            if (k > off[t]) CHECK(idx[(size_t)k - 1] < q);
            int64_t i = lin[(size_t)q] / ny;
            const int64_t j = lin[(size_t)q] % ny;
            i = std::max<int64_t>(8, std::min<int64_t>(i, nx - 9));
            CHECK(j >= bx * tile_w && j < (bx + 1) * tile_w);
            CHECK(i >= 8 + by * rows && i < 8 + (by + 1) * rows);
        }
    }
    for (size_t q = 0; q < lin.size(); ++q) CHECK(seen[q] == (lin[q] >= 0 ? 1 : 0));

    // the point table built from such lists: every listed point once, in its owning tile's range,
    // with the fields the boundary pass reads (the voxel, the source row, -1 on a release point)
    fw::BoundaryLists bl;
    bl.src_lin = lin;
    for (size_t q = 0; q < lin.size(); ++q)
        if (lin[q] >= 0 && rng() % 3 == 0) bl.zero_lin.push_back(lin[q]);
    for (size_t q = 0; q < lin.size(); ++q) {
        bl.out_lin.push_back(lin[q] >= 0 ? lin[q] : (int64_t)(rng() % (uint64_t)(nx * ny)));
        bl.out_src.push_back(rng() % 2 ? (int32_t)q : -1);
        bl.out_air.push_back((int8_t)(rng() % 5 == 0));
    }
    std::vector<int32_t> tab(3, 7);  // the table is appended to
    const fw::PointTable pt = fw::point_table(nx, ny, gx, tile_w, rows, bl, &tab);
    CHECK(pt.ntiles == (int32_t)(gx * gy) && pt.off % 4 == 0 && pt.src % 4 == 0 && pt.zero % 2 == 0 && pt.out % 4 == 0);
    auto lin_at = [&](size_t w) { return (int64_t)(((uint64_t)(uint32_t)tab[w + 1] << 32) | (uint32_t)tab[w]); };
    const size_t t1 = (size_t)pt.ntiles + 1;
    std::vector<int> src_seen(lin.size(), 0), out_seen(lin.size(), 0);
    size_t zeros = 0;
    for (size_t t = 0; t < (size_t)pt.ntiles; ++t) {
        for (int32_t k = tab[pt.off + t]; k < tab[pt.off + t + 1]; ++k) {
            const size_t w = pt.src + 4 * (size_t)k;
            const int32_t q = tab[w];
            CHECK(q >= 0 && (size_t)q < lin.size() && tab[w + 1] == 0 && lin_at(w + 2) == lin[(size_t)q]);
            if (q >= 0 && (size_t)q < lin.size()) ++src_seen[(size_t)q];
        }
        for (int32_t k = tab[pt.off + t1 + t]; k < tab[pt.off + t1 + t + 1]; ++k) {
            CHECK(std::find(bl.zero_lin.begin(), bl.zero_lin.end(), lin_at(pt.zero + 2 * (size_t)k)) != bl.zero_lin.end());
            ++zeros;
        }
        for (int32_t k = tab[pt.off + 2 * t1 + t]; k < tab[pt.off + 2 * t1 + t + 1]; ++k) {
            const size_t w = pt.out + 4 * (size_t)k;
            const int32_t q = tab[w];
            CHECK(q >= 0 && (size_t)q < lin.size());
            if (q < 0 || (size_t)q >= lin.size()) continue;
            ++out_seen[(size_t)q];
            CHECK(tab[w + 1] == bl.out_src[(size_t)q]);
            CHECK(lin_at(w + 2) == (bl.out_air[(size_t)q] ? -1 : bl.out_lin[(size_t)q]));
        }
    }
    CHECK(zeros == bl.zero_lin.size());
    for (size_t q = 0; q < lin.size(); ++q) CHECK(src_seen[q] == (lin[q] >= 0 ? 1 : 0) && out_seen[q] == 1);
}

int main(int argc, char** argv) {
    const char* tmp = argc > 1 ? argv[1] : "/tmp";
    descriptors();
    directory(tmp);
    for (unsigned seed = 1; seed <= 20; ++seed) {
        boundary(2, seed);
        boundary(3, seed);
    }
    coefficients(2);
    coefficients(3);
    for (unsigned seed = 1; seed <= 50; ++seed) point_tiles(seed);
    std::printf("host api check: %s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
