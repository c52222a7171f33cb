// box.h -- the state of one simulation box behind the C-ABI handle, shared
// by the single-box entry points (ehd_api.cpp) and the batch (batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "ehd_error.h"
#include "ehd_internal.h"

#define HIPCHK(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return ehd::set_error(EHD_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
    } while (0)

struct ehd_box {
    ehd_params prm;
    ehd_derived der;
    ehd::LbmParams lp;
    ehd::Prec prec;
    int nx, ny, nz, nxh;
    size_t S, nxy, esz;
    hipStream_t st = nullptr;
    void* f[2] = {nullptr, nullptr};
    void* f0[2] = {nullptr, nullptr};
    void* h[2] = {nullptr, nullptr};
    void* h0[2] = {nullptr, nullptr};
    int cur = 0;
    bool direct = true;        // state holds pre-collision populations
    bool ebuf = true;          // next collide takes E from ex/ey/ez
    bool initialised = false;
    void *rho = nullptr, *ux = nullptr, *uy = nullptr, *uz = nullptr, *q = nullptr, *phi = nullptr;
    double* phibar = nullptr;
    double *ex = nullptr, *ey = nullptr, *ez = nullptr;
    void* spec = nullptr;
    void* twx = nullptr;
    void* twy = nullptr;
    double* diag_scr = nullptr;
    double* diag_out = nullptr;
    double* shift_rs = nullptr;   // FP32S storage shifts (device)
    double* shift_qs = nullptr;
    float* shift_f = nullptr;
    ehd::PoissonJob* pjob = nullptr;   // the box's one-entry Poisson job table (device)
    ehd::PoissonJob pjob_h{};          // host copy (batches build their tables from it)
    void* lbm_items[2] = {nullptr, nullptr};   // FP32S: one-item kernel tables per buffer parity
    unsigned lbm_chunk = 0;              // FP32S LBM XCD tile order (lbm_xcd_chunk)
    int tune_lpb = 0, tune_ylines = 0;   // Poisson lines per block (0: largest that fits)
    bool debug_sync = false;   // synchronise and check after every launch (ehd_set_debug_sync)
    long long step = 0;
    double t = 0.0;
    double current_hs = 0.0;
    // profiling
    int prof_every = 0;   // events around the launches of every prof_every-th step; 0: off
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> lbm_pairs, poi_pairs, step_pairs;
    double lbm_ms = 0.0, poi_ms = 0.0, step_ms = 0.0;
    long long lbm_n = 0, poi_n = 0, step_n = 0;
};

namespace ehd {
PopBufs pop_bufs(ehd_box* b);
MacroBufs macro_bufs(ehd_box* b);
// one LBM launch on the box's stream (macro: also write rho and u), flips the buffers
// target_items: FP32S one-item tables (per buffer parity) carrying
// consistent-start targets; null for an ordinary step.
int run_lbm(ehd_box* b, bool macro, void* const* target_items = nullptr);
// one Poisson solve on the box's stream
int run_poisson(ehd_box* b);

// Consistent-start targets of one FP32S box: [ux, uy, uz, q] from a
// fields9 array on the device and the two TARGET item tables; released on
// destruction (after the box's stream has drained).
struct ConsistentTargets {
    double* tgt = nullptr;
    void* items[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    int make(ehd_box* b, const double* fields9);
    ~ConsistentTargets();
};
PoissonGroup poisson_group_of(const ehd_box* b);
}  // namespace ehd
