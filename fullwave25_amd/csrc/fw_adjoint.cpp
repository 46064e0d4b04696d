// Reverse-mode sweep (gradients of a functional of the recorded traces with respect to the K and
// rho maps) with checkpointed recomputation. Declarations in include/fullwave_hip.h.
//
// fw_adj_forward runs the forward handle from rest and copies its whole wave state (p, v, psi,
// phi) to a checkpoint before every C-th step. fw_adj_backward walks the segments from the
// last: it restores the segment's checkpoint into the forward handle, recomputes the segment's
// C steps while keeping the state after each one, then applies the transposed steps from the
// segment's end back to its start (the kernels in fw_kernels_adjoint.hip). Device memory:
// (ceil(nT / C) + C + 1) copies of the state plus 3 ndim + 1 scratch grids.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "fw_sim.h"

struct fw_adj {
    fw_sim* f = nullptr;
    int64_t every = 0;     // checkpoint spacing C
    int64_t nckpt = 0;
    float* ckpt = nullptr;  // [nckpt][nstate]
    float* seg = nullptr;   // [C][nstate]
    float* lam = nullptr;   // adjoint state [nstate]
    float* work = nullptr;  // H[nd], G[nd], gK, gRb[nd], grho: (3 nd + 2) grids
    float* grec = nullptr;  // [n_records][nout]
    int64_t* vox = nullptr;  // distinct sensor voxels
    int32_t* first = nullptr;
    int32_t* order = nullptr;
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

}  // namespace

extern "C" {

int fw_adj_create(fw_sim* f, int64_t checkpoint_every, fw_adj** out) {
    if (!f || !out) return fw::set_error(FW_ERR_INVALID, "null argument");
    *out = nullptr;
    if (f->nshots != 1) return fw::set_error(FW_ERR_UNSUPPORTED, "the adjoint runs single-transmit handles");
    if (checkpoint_every < 1) return fw::set_error(FW_ERR_INVALID, "checkpoint_every must be >= 1");
    Guard g(f->device);
    int rc = fw::prepare(f);
    if (rc != FW_OK) return rc;
    fw_adj* a = new fw_adj();
    a->f = f;
    a->every = std::min<int64_t>(checkpoint_every, std::max<int64_t>(f->d.nt, 1));
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
        !dalloc((void**)&a->seg, (size_t)(a->every * f->nstate) * 4) ||
        !dalloc((void**)&a->lam, (size_t)f->nstate * 4) || !dalloc((void**)&a->work, (size_t)((3 * nd + 2) * n) * 4) ||
        !dalloc((void**)&a->grec, (size_t)(nrec * f->d.ncoordsout) * 4))
        return cleanup(fw::set_error(FW_ERR_NOMEM, "adjoint buffers (%lld checkpoints + %lld steps of %lld floats) do not fit",
                                     (long long)a->nckpt, (long long)a->every, (long long)f->nstate));
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
    if (!dalloc((void**)&a->vox, vox.size() * 8) || !dalloc((void**)&a->first, first.size() * 4) ||
        !dalloc((void**)&a->order, idx.size() * 4))
        return cleanup(fw::set_error(FW_ERR_NOMEM, "adjoint sensor tables"));
    // on the handle's stream, waited for (its kernels do not order after the null stream)
    bool ok = true;
    if (!vox.empty())
        ok = hipMemcpyAsync(a->vox, vox.data(), vox.size() * 8, hipMemcpyHostToDevice, f->stream) == hipSuccess &&
             hipMemcpyAsync(a->order, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, f->stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(a->first, first.data(), first.size() * 4, hipMemcpyHostToDevice, f->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(f->stream) == hipSuccess;
    if (!ok) return cleanup(fw::set_error(FW_ERR_DEVICE, "upload of the sensor tables failed"));
    *out = a;
    return FW_OK;
}

int fw_adj_forward(fw_adj* a) {
    if (!a) return fw::set_error(FW_ERR_INVALID, "null handle");
    fw_sim* f = a->f;
    f->state_zero = false;
    Guard g(f->device);
    int rc = fw_sim_reset(f);
    if (rc != FW_OK) return rc;
    for (int64_t n = 0; n < f->d.nt; ++n) {
        if (n % a->every == 0 && (rc = copy_state(f, a->ckpt + (n / a->every) * f->nstate, f->state)) != FW_OK) return rc;
        if ((rc = fw::forward_step(f)) != FW_OK) return rc;
    }
    ADJ_HIP(hipStreamSynchronize(f->stream));
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
    float* grho = a->work + (3 * nd + 1) * n;

    int rc;
    for (int64_t k = a->nckpt - 1; k >= 0; --k) {
        const int64_t s0 = k * a->every, s1 = std::min(s0 + a->every, nt);
        float* ck = a->ckpt + k * f->nstate;
        if ((rc = copy_state(f, f->state, ck)) != FW_OK) return rc;
        f->step = s0;
        for (int64_t nn = s0; nn < s1; ++nn) {
            if ((rc = fw::forward_step(f)) != FW_OK) return rc;
            if ((rc = copy_state(f, a->seg + (nn - s0) * f->nstate, f->state)) != FW_OK) return rc;
        }
        for (int64_t nn = s1 - 1; nn >= s0; --nn) {
            const StateView cur = view_of(a->seg + (nn - s0) * f->nstate, f);
            const StateView prev = view_of(nn == s0 ? ck : a->seg + (nn - s0 - 1) * f->nstate, f);
            if (f->d.modt > 0 && nn % f->d.modt == 0)
                ADJ_HIP(fw::launch_adj_records(w.lp, a->grec + (nn / f->d.modt) * nout, a->vox, a->first, a->order,
                                               a->nvox, st));
            ADJ_HIP(fw::launch_adj_mask(w.lp, f->db.src_lin, nn < f->d.ntic ? f->db.ncoords : 0, f->db.zero_lin,
                                        f->db.nzero, st));
            w.fp_prev = prev.p;
            for (int ax = 0; ax < nd; ++ax) {
                w.fv[ax] = cur.v[ax];
                for (int nu = 0; nu < 2; ++nu) {
                    w.fpsi[ax][nu] = cur.psi[ax][nu];
                    w.fphi[ax][nu] = cur.phi[ax][nu];
                }
            }
            ADJ_HIP(fw::launch_adj_step(f->ds, w, st));
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
    for (void* p : {(void*)a->ckpt, (void*)a->seg, (void*)a->lam, (void*)a->work, (void*)a->grec, (void*)a->vox,
                    (void*)a->first, (void*)a->order})
        if (p) (void)hipFree(p);
    delete a;
    return FW_OK;
}

}  // extern "C"
