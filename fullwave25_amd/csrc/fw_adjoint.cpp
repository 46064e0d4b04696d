// Reverse-mode sweep (gradients of a functional of the recorded traces with respect to the K and
// rho maps) on a tape of the forward run. Declarations in include/fullwave_hip.h.
//
// What the transposed step needs of the forward run is u_a (the velocity update's corrected
// gradient) and div (the pressure update's corrected divergence) of every step: ndim + 1 grids,
// which the forward kernels write when the handle's DevSim carries tape pointers. The run is cut
// into segments of C steps. fw_adj_forward runs the handle from rest, copies the whole wave state
// to a checkpoint before every segment and tapes the last segment as it goes. fw_adj_backward
// walks the segments from the last: a segment whose tape is not held is recomputed from its
// checkpoint with the tape on, then the transposed steps run from the segment's end back to its
// start (the kernels in fw_kernels_adjoint.hip). With C >= nT the forward run tapes every step
// and nothing is recomputed. Device memory: ceil(nT / C) + 1 copies of the state, C (ndim + 1)
// tape grids and 3 ndim + 2 scratch grids.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>
#include <vector>

#include "fw_sim.h"

struct fw_adj {
    fw_sim* f = nullptr;
    int64_t every = 0;       // segment length C
    int64_t nckpt = 0;       // segments
    float* ckpt = nullptr;   // [nckpt][nstate]: the state before each segment
    float* tape = nullptr;   // [C][ndim + 1][ncell]: u_a, div of each step of one segment
    int64_t tape_seg = -1;   // the segment whose tape is held
    float* lam = nullptr;    // adjoint state [nstate]
    float* work = nullptr;   // H[nd], G[nd], gRb[nd], gK, grho: (3 nd + 2) grids
    float* grec = nullptr;   // [n_records][nout]
    int64_t* vox = nullptr;  // distinct sensor voxels
    int32_t* first = nullptr;
    int32_t* order = nullptr;
    int8_t* vmask = nullptr;  // per distinct voxel: 1 pressure release, 2 source point, 0 neither
    uint8_t* kind = nullptr;  // per device cell (2D): bit 0 pressure release, 1 source point, 2 sensor voxel
    uint8_t* dc8 = nullptr;   // per device cell: coefficient-set index (grids of the tiled reverse kernels)
    int64_t nvox = 0;
    bool forward_done = false;
};

namespace {

#define ADJ_HIP(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
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

// pointers into a packed state [p | v[nd] | psi[nd][2] | phi[nd][2]] (psi/phi absent for the
// exponential model)
struct StateView {
    float* p;
    float* v[3] = {nullptr, nullptr, nullptr};
    float* psi[3][2] = {};
    float* phi[3][2] = {};
};

StateView view_of(float* base, const fw_sim* s) {
    StateView w;
    const int64_t n = s->ncell;
    const int nd = s->ndim;
    w.p = base;
    for (int a = 0; a < nd; ++a) w.v[a] = base + n * (1 + a);
    if (s->d.attenuation_model == 0) {
        for (int a = 0; a < nd; ++a)
            for (int nu = 0; nu < 2; ++nu) {
                w.psi[a][nu] = base + n * (1 + nd + 2 * a + nu);
                w.phi[a][nu] = base + n * (1 + 3 * nd + 2 * a + nu);
            }
    }
    return w;
}

int copy_state(fw_sim* s, float* dst, const float* src) {
    ADJ_HIP(hipMemcpyAsync(dst, src, (size_t)s->nstate * 4, hipMemcpyDeviceToDevice, s->stream));
    return FW_OK;
}

// the tape grids of step `slot` of the held segment
float* tape_slot(const fw_adj* a, int64_t slot) { return a->tape + slot * (a->f->ndim + 1) * a->f->ncell; }

// one forward step that writes its tape into `slot`
int taped_step(fw_adj* a, int64_t slot) {
    fw_sim* f = a->f;
    float* t = tape_slot(a, slot);
    for (int ax = 0; ax < f->ndim; ++ax) f->ds.tape_u[ax] = t + ax * f->ncell;
    f->ds.tape_div = t + f->ndim * f->ncell;
    const int rc = fw::forward_step(f);
    for (int ax = 0; ax < fw::kMaxDim; ++ax) f->ds.tape_u[ax] = nullptr;
    f->ds.tape_div = nullptr;
    return rc;
}

// Segment length for checkpoint_every == 0: the whole run taped when that fits in half of the
// device's free memory, else the C that minimises ceil(nT / C) states plus C tapes.
int64_t auto_segment(const fw_sim* f, size_t free_bytes) {
    const int64_t nt = std::max<int64_t>(f->d.nt, 1);
    const double tape_step = (double)(f->ndim + 1) * (double)f->ncell * 4.0;
    if (tape_step * (double)nt <= 0.5 * (double)free_bytes) return nt;
    const double c = std::sqrt((double)nt * (double)f->nstate * 4.0 / tape_step);
    return std::min<int64_t>(nt, std::max<int64_t>(1, (int64_t)std::ceil(c)));
}

// Kernels that write the tape (the tiled 2D or per-cell families): the tiled ones (8 rows) where
// the grid takes them, else the per-cell ones; a kernel the caller pinned stays if it writes
// the tape.
void tape_kernels(fw_sim* f) {
    fw::LaunchCfg t;
    t.variant = 2;
    t.width = 256;
    t.chunk = 8;
    if (!fw::tape_capable(f->ds, t)) t.variant = 0;
    for (int kind = 0; kind < 2; ++kind) {
        fw::LaunchCfg& c = kind == 0 ? f->cfg_v : f->cfg_p;
        if (f->pinned[kind] && fw::tape_capable(f->ds, c)) continue;
        c = t;
    }
}

}  // namespace

extern "C" {

int fw_adj_create(fw_sim* f, int64_t checkpoint_every, fw_adj** out) {
    if (!f || !out) return fw::set_error(FW_ERR_INVALID, "null argument");
    *out = nullptr;
    if (f->nshots != 1) return fw::set_error(FW_ERR_UNSUPPORTED, "the adjoint runs single-transmit handles");
    if (checkpoint_every < 0) return fw::set_error(FW_ERR_INVALID, "checkpoint_every must be >= 0");
    Guard g(f->device);
    int rc = fw::prepare(f);
    if (rc != FW_OK) return rc;
    tape_kernels(f);
    size_t free_b = 0, total_b = 0;
    if (checkpoint_every == 0) ADJ_HIP(hipMemGetInfo(&free_b, &total_b));
    fw_adj* a = new fw_adj();
    a->f = f;
    a->every = checkpoint_every ? std::min<int64_t>(checkpoint_every, std::max<int64_t>(f->d.nt, 1))
                                : auto_segment(f, free_b);
    a->nckpt = (f->d.nt + a->every - 1) / a->every;
    const int nd = f->ndim;
    const int64_t n = f->ncell;
    const int64_t nrec = f->rec_cap;
    auto cleanup = [&](int code) {
        fw_adj_destroy(a);
        return code;
    };
    auto dalloc = [&](void** p, size_t bytes) -> bool { return hipMalloc(p, bytes ? bytes : 4) == hipSuccess; };
    if (!dalloc((void**)&a->ckpt, (size_t)(a->nckpt * f->nstate) * 4) ||
        !dalloc((void**)&a->tape, (size_t)(a->every * (nd + 1) * n) * 4) ||
        !dalloc((void**)&a->lam, (size_t)f->nstate * 4) || !dalloc((void**)&a->work, (size_t)((3 * nd + 2) * n) * 4) ||
        !dalloc((void**)&a->grec, (size_t)(nrec * f->d.ncoordsout) * 4))
        return cleanup(fw::set_error(FW_ERR_NOMEM,
                                     "adjoint buffers (%lld checkpoints of %lld floats, a tape of %lld steps) do not fit",
                                     (long long)a->nckpt, (long long)f->nstate, (long long)a->every));
    // distinct sensor voxels and the sensors on each (several sensors may share a voxel)
    std::vector<int32_t> idx(f->out_lin_h.size());
    for (size_t q = 0; q < idx.size(); ++q) idx[q] = (int32_t)q;
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return f->out_lin_h[x] < f->out_lin_h[y]; });
    std::vector<int64_t> vox;
    std::vector<int32_t> first;
    for (size_t q = 0; q < idx.size(); ++q) {
        if (q == 0 || f->out_lin_h[idx[q]] != f->out_lin_h[idx[q - 1]]) {
            vox.push_back(f->out_lin_h[idx[q]]);
            first.push_back((int32_t)q);
        }
    }
    first.push_back((int32_t)idx.size());
    a->nvox = (int64_t)vox.size();
    // which of those voxels a step overwrites (pressure release always, source points while the
    // step injects): the resolved point lists, read back from the handle
    std::vector<int64_t> src((size_t)f->db.ncoords), zero((size_t)f->db.nzero);
    if (!src.empty())
        ADJ_HIP(hipMemcpyAsync(src.data(), f->db.src_lin, src.size() * 8, hipMemcpyDeviceToHost, f->stream));
    if (!zero.empty())
        ADJ_HIP(hipMemcpyAsync(zero.data(), f->db.zero_lin, zero.size() * 8, hipMemcpyDeviceToHost, f->stream));
    ADJ_HIP(hipStreamSynchronize(f->stream));
    const std::set<int64_t> src_set(src.begin(), src.end()), zero_set(zero.begin(), zero.end());
    std::vector<int8_t> vmask(vox.size());
    for (size_t t = 0; t < vox.size(); ++t) vmask[t] = zero_set.count(vox[t]) ? 1 : src_set.count(vox[t]) ? 2 : 0;
    // the same per cell, for the reverse kernels that run a step's sparse part cell by cell
    std::vector<uint8_t> kind;
    if (nd == 2) {
        kind.assign((size_t)n, 0);
        for (int64_t z : zero) kind[(size_t)z] |= 1;
        for (int64_t q : src)
            if (q >= 0) kind[(size_t)q] |= 2;
        for (int64_t v : vox) kind[(size_t)v] |= 4;
    }
    if (!dalloc((void**)&a->vox, vox.size() * 8) || !dalloc((void**)&a->first, first.size() * 4) ||
        !dalloc((void**)&a->order, idx.size() * 4) || !dalloc((void**)&a->vmask, vmask.size()) ||
        !dalloc((void**)&a->kind, kind.size()))
        return cleanup(fw::set_error(FW_ERR_NOMEM, "adjoint sensor tables"));
    // on the handle's stream, waited for (its kernels do not order after the null stream)
    bool ok = true;
    if (!vox.empty())
        ok = hipMemcpyAsync(a->vox, vox.data(), vox.size() * 8, hipMemcpyHostToDevice, f->stream) == hipSuccess &&
             hipMemcpyAsync(a->order, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, f->stream) == hipSuccess &&
             hipMemcpyAsync(a->vmask, vmask.data(), vmask.size(), hipMemcpyHostToDevice, f->stream) == hipSuccess;
    if (!kind.empty())
        ok = ok && hipMemcpyAsync(a->kind, kind.data(), kind.size(), hipMemcpyHostToDevice, f->stream) == hipSuccess;
    if (ok && fw::adj_tiles(f->ds)) {
        if (!dalloc((void**)&a->dc8, (size_t)n)) return cleanup(fw::set_error(FW_ERR_NOMEM, "adjoint set-index map"));
        ok = fw::launch_adj_set_index(f->ds.dcmap, a->dc8, n, f->stream) == hipSuccess;
    }
    ok = ok && hipMemcpyAsync(a->first, first.data(), first.size() * 4, hipMemcpyHostToDevice, f->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(f->stream) == hipSuccess;
    if (!ok) return cleanup(fw::set_error(FW_ERR_DEVICE, "upload of the sensor tables failed"));
    *out = a;
    return FW_OK;
}

int fw_adj_checkpoint_every(const fw_adj* a, int64_t* every) {
    if (!a || !every) return fw::set_error(FW_ERR_INVALID, "null argument");
    *every = a->every;
    return FW_OK;
}

int fw_adj_forward(fw_adj* a) {
    if (!a) return fw::set_error(FW_ERR_INVALID, "null handle");
    fw_sim* f = a->f;
    f->state_zero = false;
    Guard g(f->device);
    int rc = fw_sim_reset(f);
    if (rc != FW_OK) return rc;
    a->tape_seg = -1;
    const int64_t last = a->nckpt - 1, s_last = last * a->every;
    for (int64_t n = 0; n < f->d.nt; ++n) {
        if (n % a->every == 0 && (rc = copy_state(f, a->ckpt + (n / a->every) * f->nstate, f->state)) != FW_OK) return rc;
        rc = n >= s_last ? taped_step(a, n - s_last) : fw::forward_step(f);
        if (rc != FW_OK) return rc;
    }
    ADJ_HIP(hipStreamSynchronize(f->stream));
    a->tape_seg = last;
    a->forward_done = true;
    return FW_OK;
}

int fw_adj_backward(fw_adj* a, const float* grad_rec, size_t grad_bytes, float* grad_k, float* grad_rho,
                    size_t grid_bytes) {
    if (!a || !grad_rec || !grad_k || !grad_rho) return fw::set_error(FW_ERR_INVALID, "null argument");
    fw_sim* f = a->f;
    f->state_zero = false;
    if (!a->forward_done) return fw::set_error(FW_ERR_STATE, "run fw_adj_forward first");
    const int64_t nrec = f->rec_cap, nout = f->d.ncoordsout, n = f->ncell, nt = f->d.nt;
    const int64_t n_c = f->ncell_c;
    if (grad_bytes != (size_t)(nrec * nout) * 4)
        return fw::set_error(FW_ERR_INVALID, "grad_rec must be [%lld, %lld] float32", (long long)nrec, (long long)nout);
    if (grid_bytes != (size_t)n_c * 4)
        return fw::set_error(FW_ERR_INVALID, "gradient buffers must be %lld bytes", (long long)n_c * 4);
    Guard g(f->device);
    hipStream_t st = f->stream;
    const int nd = f->ndim;
    ADJ_HIP(hipMemcpyAsync(a->grec, grad_rec, grad_bytes, hipMemcpyHostToDevice, st));
    ADJ_HIP(hipStreamSynchronize(st));  // grad_rec is the caller's buffer, read during the call only
    ADJ_HIP(hipMemsetAsync(a->lam, 0, (size_t)f->nstate * 4, st));
    ADJ_HIP(hipMemsetAsync(a->work, 0, (size_t)((3 * nd + 2) * n) * 4, st));

    fw::AdjView w{};
    const StateView L = view_of(a->lam, f);
    w.lp = L.p;
    for (int ax = 0; ax < nd; ++ax) {
        w.lv[ax] = L.v[ax];
        for (int nu = 0; nu < 2; ++nu) {
            w.lpsi[ax][nu] = L.psi[ax][nu];
            w.lphi[ax][nu] = L.phi[ax][nu];
        }
        w.H[ax] = a->work + ax * n;
        w.G[ax] = a->work + (nd + ax) * n;
        w.gRb[ax] = a->work + (2 * nd + ax) * n;
    }
    w.gK = a->work + 3 * nd * n;
    w.dc8 = a->dc8;
    float* grho = a->work + (3 * nd + 1) * n;
    fw::AdjSparse sp{};
    sp.vox = a->vox;
    sp.first = a->first;
    sp.order = a->order;
    sp.vmask = a->vmask;
    sp.nvox = a->nvox;
    sp.src_lin = f->db.src_lin;
    sp.nsrc = f->db.ncoords;
    sp.zero_lin = f->db.zero_lin;
    sp.nzero = f->db.nzero;

    // step nn's record weights (null: the step records nothing)
    auto rec_row = [&](int64_t nn) -> const float* {
        return f->d.modt > 0 && nn % f->d.modt == 0 ? a->grec + (nn / f->d.modt) * nout : nullptr;
    };
    const bool fuse = fw::adj_fuses_next(f->ds);
    bool head_done = false;  // step nn's sparse part and pressure transposed already ran
    int rc;
    for (int64_t k = a->nckpt - 1; k >= 0; --k) {
        const int64_t s0 = k * a->every, s1 = std::min(s0 + a->every, nt);
        if (a->tape_seg != k) {
            a->tape_seg = -1;
            if ((rc = copy_state(f, f->state, a->ckpt + k * f->nstate)) != FW_OK) return rc;
            f->step = s0;
            for (int64_t nn = s0; nn < s1; ++nn)
                if ((rc = taped_step(a, nn - s0)) != FW_OK) return rc;
            a->tape_seg = k;
        }
        for (int64_t nn = s1 - 1; nn >= s0; --nn) {
            const float* t = tape_slot(a, nn - s0);
            for (int ax = 0; ax < nd; ++ax) w.tu[ax] = t + ax * n;
            w.tdiv = t + nd * n;
            if (!head_done) {
                ADJ_HIP(fw::launch_adj_sparse(w.lp, rec_row(nn), sp, nn < f->d.ntic, st));
                ADJ_HIP(fw::launch_adj_pressure(f->ds, w, st));
            }
            // the next step's start joins this one's last kernel while its tape is held
            head_done = fuse && nn > s0;
            fw::AdjNext nx{};
            if (head_done) {
                nx.tdiv = tape_slot(a, nn - 1 - s0) + nd * n;
                nx.grow = rec_row(nn - 1);
                nx.injecting = nn - 1 < f->d.ntic;
                nx.kind = a->kind;
                nx.vox = a->vox;
                nx.first = a->first;
                nx.order = a->order;
                nx.nvox = a->nvox;
            }
            ADJ_HIP(fw::launch_adj_rest(f->ds, w, head_done ? &nx : nullptr, st));
        }
    }
    ADJ_HIP(fw::launch_adj_rho(f->ds, w, grho, st));
    ADJ_HIP(hipStreamSynchronize(st));
    ADJ_HIP(fw::grid_to_host(f, grad_k, w.gK, 4));
    ADJ_HIP(fw::grid_to_host(f, grad_rho, grho, 4));
    return FW_OK;
}

int fw_adj_destroy(fw_adj* a) {
    if (!a) return FW_OK;
    Guard g(a->f->device);
    if (a->f->stream) (void)hipStreamSynchronize(a->f->stream);
    for (void* p : {(void*)a->ckpt, (void*)a->tape, (void*)a->lam, (void*)a->work, (void*)a->grec, (void*)a->vox,
                    (void*)a->first, (void*)a->order, (void*)a->vmask, (void*)a->kind, (void*)a->dc8})
        if (p) (void)hipFree(p);
    delete a;
    return FW_OK;
}

}  // extern "C"
