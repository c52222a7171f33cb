// batch.cpp -- many independent boxes stepped together (include/ehd.h,
// ehd_batch_*): one persistent LBM launch over all FP32S boxes, the Poisson
// passes batched per (nx, ny) group, the steady step pair captured in a HIP
// graph, and an autotuner for the launch shapes.  Every box runs exactly the
// arithmetic of a single-box step (same node function, same FFT line
// programs, same z-solve partition), so results are bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "box.h"

using namespace ehd;

namespace {

constexpr int kMaxGroups = 16;
// streams the Poisson groups of a batch spread over, the batch stream included
constexpr int kPoissonStreams = 4;

struct Group {
    PoissonGroup g{};
    PoissonJob* djobs = nullptr;   // device table of the group's jobs
    std::vector<int> members;
    // chains: the FP32S item tables of the members (tiles numbered from 0),
    // their tile count and XCD order, and the stream the group runs on
    void* items[2] = {nullptr, nullptr};
    unsigned tiles = 0;
    unsigned chunk = 0;
    int stream = 0;                // 0: the batch stream, k: side stream k-1
};

PopBufs pop_bufs_at(ehd_box* b, int cur)
{
    const int A = cur, B = 1 - cur;
    return PopBufs{b->f[A], b->f0[A], b->h[A], b->h0[A], b->f[B], b->f0[B], b->h[B], b->h0[B]};
}

}  // namespace

struct ehd_batch {
    std::vector<ehd_box*> boxes;
    Prec prec = Prec::F32S;
    hipStream_t st = nullptr;
    int flags = 0;
    // FP32S: one item table per buffer parity
    void* ditems[2] = {nullptr, nullptr};
    unsigned total_tiles = 0;
    unsigned lbm_grid = 0;
    unsigned lbm_chunk = 0;   // XCD tile order of the LBM launch (0: linear)
    double lbm_ms = 0.0;
    std::vector<Group> groups;
    std::vector<double> group_ms;
    // graph of the steady step pair, per starting parity
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    // side streams (default priority) on which Poisson groups run beside the
    // batch stream, and the events that fork them from it and join them back
    std::vector<hipStream_t> side;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    // one LBM launch and Poisson chain per group, each group on its stream,
    // instead of one LBM launch for the whole batch
    bool chains = false;
    double step_ms[2] = {0.0, 0.0};   // timed step: one LBM launch / chains (tuning)
};

namespace {

int batch_lbm(ehd_batch* bt, int cur, bool first, bool ebuf, bool macro)
{
    if (bt->prec == Prec::F32S) {
        HIPCHK(launch_lbm_f32s_batch(bt->ditems[cur], (int)bt->boxes.size(), bt->total_tiles, bt->lbm_grid,
                                     bt->lbm_chunk, first, ebuf, macro, bt->st));
    } else {
        for (ehd_box* b : bt->boxes)
            HIPCHK(launch_lbm_f64(b->lp, pop_bufs_at(b, cur), macro_bufs(b), first, ebuf, macro, bt->st));
    }
    return EHD_OK;
}

// The groups touch disjoint boxes, so they run concurrently: group i on
// stream i mod (1 + side streams), stream 0 being the batch stream.  The side
// streams start after the work queued so far on the batch stream (the LBM
// launch) and the batch stream continues after all of them; under graph
// capture the same fork and join become the graph's branches.
hipStream_t group_stream(const ehd_batch* bt, const Group& gr)
{
    return gr.stream == 0 ? bt->st : bt->side[gr.stream - 1];
}

// the side streams start after the work queued so far on the batch stream
int fork_side(ehd_batch* bt)
{
    HIPCHK(hipEventRecord(bt->fork, bt->st));
    for (hipStream_t s : bt->side) HIPCHK(hipStreamWaitEvent(s, bt->fork, 0));
    return EHD_OK;
}

// the batch stream continues after all the side streams
int join_side(ehd_batch* bt)
{
    for (size_t k = 0; k < bt->side.size(); ++k) {
        HIPCHK(hipEventRecord(bt->join[k], bt->side[k]));
        HIPCHK(hipStreamWaitEvent(bt->st, bt->join[k], 0));
    }
    return EHD_OK;
}

int batch_poisson(ehd_batch* bt)
{
    if (bt->side.empty()) {
        for (const Group& gr : bt->groups) HIPCHK(launch_poisson_group(bt->prec, gr.g, bt->st));
        return EHD_OK;
    }
    int rc = fork_side(bt);
    if (rc) return rc;
    for (const Group& gr : bt->groups) HIPCHK(launch_poisson_group(bt->prec, gr.g, group_stream(bt, gr)));
    return join_side(bt);
}

// One step of every group on the group's own stream (chains), between a
// fork_side and a join_side: a group's LBM launch waits only for its own
// previous Poisson solve, so one group's Poisson passes overlap another's
// LBM launch.
int chain_step(ehd_batch* bt, int cur, bool first, bool ebuf, bool macro)
{
    for (const Group& gr : bt->groups) {
        const hipStream_t s = group_stream(bt, gr);
        if (bt->prec == Prec::F32S) {
            HIPCHK(launch_lbm_f32s_batch(gr.items[cur], (int)gr.members.size(), gr.tiles, 0, gr.chunk, first, ebuf,
                                         macro, s));
        } else {
            for (int m : gr.members) {
                ehd_box* b = bt->boxes[m];
                HIPCHK(launch_lbm_f64(b->lp, pop_bufs_at(b, cur), macro_bufs(b), first, ebuf, macro, s));
            }
        }
        HIPCHK(launch_poisson_group(bt->prec, gr.g, s));
    }
    return EHD_OK;
}

// one coupled step of the whole batch, ending on the batch stream
int batch_step_once(ehd_batch* bt, int cur, bool first, bool ebuf, bool macro)
{
    int rc;
    if (!bt->chains) {
        if ((rc = batch_lbm(bt, cur, first, ebuf, macro))) return rc;
        return batch_poisson(bt);
    }
    if ((rc = fork_side(bt))) return rc;
    if ((rc = chain_step(bt, cur, first, ebuf, macro))) return rc;
    return join_side(bt);
}

void drop_graphs(ehd_batch* bt)
{
    for (auto& e : bt->exec)
        if (e) {
            (void)hipGraphExecDestroy(e);
            e = nullptr;
        }
}

// Capture the two steady steps (direct = ebuf = false, no macroscopic
// output) that start from buffer parity cur.
int capture_pair(ehd_batch* bt, int cur)
{
    HIPCHK(hipStreamBeginCapture(bt->st, hipStreamCaptureModeThreadLocal));
    int rc = EHD_OK;
    if (bt->chains) {
        // the chains run both steps before they join
        rc = fork_side(bt);
        for (int k = 0; k < 2 && rc == EHD_OK; ++k) rc = chain_step(bt, (cur + k) & 1, false, false, false);
        if (rc == EHD_OK) rc = join_side(bt);
    } else {
        for (int k = 0; k < 2 && rc == EHD_OK; ++k) {
            rc = batch_lbm(bt, (cur + k) & 1, false, false, false);
            if (rc == EHD_OK) rc = batch_poisson(bt);
        }
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(bt->st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return set_error(EHD_EHIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&bt->exec[cur], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return set_error(EHD_EHIP, "graph instantiation failed: %s", hipGetErrorString(e));
    return EHD_OK;
}

void free_batch(ehd_batch* bt)
{
    if (bt->st) (void)hipStreamSynchronize(bt->st);
    for (hipStream_t s : bt->side) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    bt->side.clear();
    for (hipEvent_t e : bt->join) (void)hipEventDestroy(e);
    bt->join.clear();
    if (bt->fork) (void)hipEventDestroy(bt->fork);
    bt->fork = nullptr;
    drop_graphs(bt);
    for (void* p : bt->ditems)
        if (p) (void)hipFree(p);
    for (Group& g : bt->groups) {
        if (g.djobs) (void)hipFree(g.djobs);
        for (void* p : g.items)
            if (p) (void)hipFree(p);
    }
    for (ehd_box* b : bt->boxes) ehd_destroy(b);
    if (bt->st) (void)hipStreamDestroy(bt->st);
}

// Divisors of ny/2 up to the largest line count that fits, largest first.
std::vector<int> x_candidates(int ny, int maxl)
{
    std::vector<int> c;
    for (int l = maxl; l >= 1 && c.size() < 4; --l)
        if ((ny / 2) % l == 0) c.push_back(l);
    return c;
}

template <class F>
int time_ms(ehd_batch* bt, int reps, F launch, double* ms)
{
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = launch();   // warm-up
    if (rc == EHD_OK) {
        (void)hipEventRecord(e0, bt->st);
        for (int r = 0; r < reps && rc == EHD_OK; ++r) rc = launch();
        (void)hipEventRecord(e1, bt->st);
        (void)hipEventSynchronize(e1);
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        *ms = t / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// Autotuner: before the boxes are initialised (their buffers hold no state
// yet), time the steady-state kernels for each candidate launch shape and
// keep the fastest.  None of the shapes changes any result.
int tune(ehd_batch* bt)
{
    const int reps = 5;
    if (bt->prec == Prec::F32S) {
        const unsigned T = bt->total_tiles;
        std::vector<unsigned> cands = {0u};
        for (unsigned g : {1024u, 2048u, 4096u, 8192u})
            if (g < T) cands.push_back(g);
        std::vector<unsigned> chunks = {bt->lbm_chunk};
        if (bt->lbm_chunk) chunks.push_back(0u);
        double best = 1e30;
        unsigned bestg = 0, bestc = bt->lbm_chunk;
        for (unsigned ch : chunks)
            for (unsigned g : cands) {
                bt->lbm_grid = g;
                bt->lbm_chunk = ch;
                double ms = 0;
                int rc = time_ms(bt, reps, [&] { return batch_lbm(bt, 0, false, false, false); }, &ms);
                if (rc) return rc;
                if (ms < best) {
                    best = ms;
                    bestg = g;
                    bestc = ch;
                }
            }
        bt->lbm_grid = bestg;
        bt->lbm_chunk = bestc;
        bt->lbm_ms = best;
    }
    bt->group_ms.assign(bt->groups.size(), 0.0);
    for (size_t gi = 0; gi < bt->groups.size(); ++gi) {
        Group& gr = bt->groups[gi];
        const PoissonShape ps = poisson_shape(bt->prec, gr.g.nx, gr.g.ny, gr.g.max_n + 2);
        double best = 1e30;
        int bl = 0, by = 0;
        for (int xl : x_candidates(gr.g.ny, ps.max_x_lines))
            for (int yl : {ps.max_y_lines, ps.max_y_lines / 2, ps.max_y_lines / 4}) {
                if (yl < 1) continue;
                gr.g.lpb = xl;
                gr.g.ylines = yl;
                double ms = 0;
                int rc = time_ms(bt, reps, [&]() -> int {
                    HIPCHK(launch_poisson_group(bt->prec, gr.g, bt->st));
                    return EHD_OK;
                }, &ms);
                if (rc) return rc;
                if (ms < best) {
                    best = ms;
                    bl = xl;
                    by = yl;
                }
            }
        gr.g.lpb = bl;
        gr.g.ylines = by;
        bt->group_ms[gi] = best;
        for (int m : gr.members) {
            bt->boxes[m]->tune_lpb = bl;
            bt->boxes[m]->tune_ylines = by;
        }
    }
    // one LBM launch for the whole batch, or one chain per group: a step of
    // each, run as the batch will run it (the captured steady pair replayed,
    // or eager)
    if (!bt->side.empty() && !(bt->flags & EHD_BATCH_CHAINS)) {
        const bool graph = (bt->flags & EHD_BATCH_GRAPH) != 0;
        for (int mode = 0; mode < 2; ++mode) {
            bt->chains = (mode == 1);
            int rc;
            if (graph) {
                if ((rc = capture_pair(bt, 0))) return rc;
                rc = time_ms(bt, reps, [&]() -> int {
                    HIPCHK(hipGraphLaunch(bt->exec[0], bt->st));
                    return EHD_OK;
                }, &bt->step_ms[mode]);
                drop_graphs(bt);
                bt->step_ms[mode] *= 0.5;   // two steps per replay
            } else {
                rc = time_ms(bt, reps, [&] { return batch_step_once(bt, 0, false, false, false); }, &bt->step_ms[mode]);
            }
            if (rc) return rc;
        }
        bt->chains = bt->step_ms[1] < bt->step_ms[0];
    }
    HIPCHK(hipStreamSynchronize(bt->st));
    return EHD_OK;
}

int check_uniform(ehd_batch* bt)
{
    const ehd_box* b0 = bt->boxes[0];
    for (const ehd_box* b : bt->boxes) {
        if (!b->initialised) return set_error(EHD_ESTATE, "batch box not initialised");
        if (b->cur != b0->cur || b->direct != b0->direct || b->ebuf != b0->ebuf)
            return set_error(EHD_ESTATE, "batch boxes are out of step (a box was stepped alone)");
    }
    return EHD_OK;
}

}  // namespace

extern "C" {

int ehd_batch_create(const ehd_params* params, int nboxes, int flags, ehd_batch** out)
{
    if (!out) return set_error(EHD_EINVAL, "null output");
    *out = nullptr;
    if (!params || nboxes < 1) return set_error(EHD_EINVAL, "need at least one box");
    for (int i = 1; i < nboxes; ++i)
        if (params[i].precision != params[0].precision) return set_error(EHD_EINVAL, "a batch has one precision");
    ehd_batch* bt = new (std::nothrow) ehd_batch();
    if (!bt) return set_error(EHD_ENOMEM, "host allocation failed");
    auto bail = [&](int code) {
        free_batch(bt);
        delete bt;
        return code;
    };
    bt->flags = flags;
    bt->prec = params[0].precision == EHD_FP64 ? Prec::F64 : Prec::F32S;
    if (hipStreamCreateWithFlags(&bt->st, hipStreamNonBlocking) != hipSuccess)
        return bail(set_error(EHD_EHIP, "hipStreamCreate failed"));
    for (int i = 0; i < nboxes; ++i) {
        ehd_box* b = nullptr;
        int rc = ehd_create(&params[i], &b);
        if (rc) return bail(rc);
        bt->boxes.push_back(b);
    }
    // Poisson groups by (nx, ny); the z-solve partition follows n, so boxes of
    // one (nx, ny) whose plane counts need different partitions get separate groups
    for (int i = 0; i < nboxes; ++i) {
        ehd_box* b = bt->boxes[i];
        const int zp = poisson_shape(b->prec, b->nx, b->ny, b->nz).zs_partition;
        Group* gr = nullptr;
        for (Group& g : bt->groups)
            if (g.g.nx == b->nx && g.g.ny == b->ny && g.g.zs_partition == zp) gr = &g;
        if (!gr) {
            if ((int)bt->groups.size() == kMaxGroups) return bail(set_error(EHD_EINVAL, "more than 16 Poisson groups"));
            bt->groups.emplace_back();
            gr = &bt->groups.back();
            gr->g.nx = b->nx;
            gr->g.ny = b->ny;
            gr->g.split_mean = b->prec == Prec::F32S ? 1 : 0;
            gr->g.zs_partition = zp;
            gr->g.twx = b->twx;
            gr->g.twy = b->twy;
        }
        gr->members.push_back(i);
        gr->g.max_n = std::max(gr->g.max_n, b->nz - 2);
    }
    for (Group& gr : bt->groups) {
        // the largest box of a group fixes the z-solve tile; its partition must serve every member
        if (poisson_shape(bt->prec, gr.g.nx, gr.g.ny, gr.g.max_n + 2).zs_partition != gr.g.zs_partition)
            return bail(set_error(EHD_EINVAL, "boxes of one (nx, ny) need different z-solve partitions"));
        std::vector<PoissonJob> jobs;
        for (int m : gr.members) jobs.push_back(bt->boxes[m]->pjob_h);
        gr.g.njobs = (int)jobs.size();
        if (hipMalloc((void**)&gr.djobs, jobs.size() * sizeof(PoissonJob)) != hipSuccess)
            return bail(set_error(EHD_ENOMEM, "job table allocation failed"));
        if (hipMemcpy(gr.djobs, jobs.data(), jobs.size() * sizeof(PoissonJob), hipMemcpyHostToDevice) != hipSuccess)
            return bail(set_error(EHD_EHIP, "job table upload failed"));
        gr.g.jobs = gr.djobs;
    }
    // up to kPoissonStreams - 1 side streams for the groups (with the batch
    // stream that is the number of hardware queues a process gets by default)
    const int nside = std::min((int)bt->groups.size(), kPoissonStreams) - 1;
    if (nside > 0) {
        if (hipEventCreateWithFlags(&bt->fork, hipEventDisableTiming) != hipSuccess)
            return bail(set_error(EHD_EHIP, "hipEventCreate failed"));
        for (int k = 0; k < nside; ++k) {
            hipStream_t s = nullptr;
            hipEvent_t e = nullptr;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
                return bail(set_error(EHD_EHIP, "hipStreamCreate failed"));
            bt->side.push_back(s);
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                return bail(set_error(EHD_EHIP, "hipEventCreate failed"));
            bt->join.push_back(e);
        }
        // groups to streams, largest first onto the least loaded stream (by nodes)
        std::vector<size_t> order(bt->groups.size());
        std::vector<double> nodes(bt->groups.size(), 0.0), load(nside + 1, 0.0);
        for (size_t i = 0; i < order.size(); ++i) {
            order[i] = i;
            for (int m : bt->groups[i].members) nodes[i] += (double)bt->boxes[m]->S;
        }
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return nodes[a] > nodes[b]; });
        for (size_t i : order) {
            const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            bt->groups[i].stream = k;
            load[k] += nodes[i];
        }
    }
    if (bt->prec == Prec::F32S) {
        const size_t isz = lbm_f32s_batch_item_size();
        for (int cur = 0; cur < 2; ++cur) {
            std::vector<unsigned char> h(isz * nboxes);
            unsigned tile0 = 0;
            for (int i = 0; i < nboxes; ++i) {
                ehd_box* b = bt->boxes[i];
                lbm_f32s_batch_item(h.data() + i * isz, b->lp, pop_bufs_at(b, cur), macro_bufs(b), tile0);
                tile0 += (unsigned)((b->S + 255) / 256);
            }
            bt->total_tiles = tile0;
            if (hipMalloc(&bt->ditems[cur], h.size()) != hipSuccess)
                return bail(set_error(EHD_ENOMEM, "item table allocation failed"));
            if (hipMemcpy(bt->ditems[cur], h.data(), h.size(), hipMemcpyHostToDevice) != hipSuccess)
                return bail(set_error(EHD_EHIP, "item table upload failed"));
        }
        // XCD tile order: superblocks are whole planes when every box has the same plane size
        bt->lbm_chunk = bt->boxes[0]->lbm_chunk;
        for (const ehd_box* b : bt->boxes)
            if (b->lbm_chunk != bt->lbm_chunk) bt->lbm_chunk = 0;
        // per-group item tables for the chains (the boxes of a group share their plane size)
        if (!bt->side.empty())
            for (Group& gr : bt->groups) {
                for (int cur = 0; cur < 2; ++cur) {
                    std::vector<unsigned char> h(isz * gr.members.size());
                    unsigned tile0 = 0;
                    for (size_t j = 0; j < gr.members.size(); ++j) {
                        ehd_box* b = bt->boxes[gr.members[j]];
                        lbm_f32s_batch_item(h.data() + j * isz, b->lp, pop_bufs_at(b, cur), macro_bufs(b), tile0);
                        tile0 += (unsigned)((b->S + 255) / 256);
                    }
                    gr.tiles = tile0;
                    if (hipMalloc(&gr.items[cur], h.size()) != hipSuccess)
                        return bail(set_error(EHD_ENOMEM, "item table allocation failed"));
                    if (hipMemcpy(gr.items[cur], h.data(), h.size(), hipMemcpyHostToDevice) != hipSuccess)
                        return bail(set_error(EHD_EHIP, "item table upload failed"));
                }
                gr.chunk = bt->boxes[gr.members[0]]->lbm_chunk;
            }
    }
    bt->chains = (flags & EHD_BATCH_CHAINS) != 0 && !bt->side.empty();
    if (flags & EHD_BATCH_TUNE) {
        int rc = tune(bt);
        if (rc) return bail(rc);
    }
    *out = bt;
    return EHD_OK;
}

void ehd_batch_destroy(ehd_batch* bt)
{
    if (!bt) return;
    free_batch(bt);
    delete bt;
}

int ehd_batch_init(ehd_batch* bt, const double* const* fields9)
{
    if (!bt) return set_error(EHD_EINVAL, "null batch");
    for (size_t i = 0; i < bt->boxes.size(); ++i) {
        int rc = ehd_init(bt->boxes[i], fields9 ? fields9[i] : nullptr);
        if (rc) return rc;
    }
    return EHD_OK;
}

int ehd_batch_init_consistent(ehd_batch* bt, const double* const* fields9, int iterations)
{
    if (!bt || !fields9) return set_error(EHD_EINVAL, "null argument");
    if (bt->prec != Prec::F32S) return set_error(EHD_EINVAL, "the consistent start is implemented for FP32S boxes");
    if (iterations < 0) return set_error(EHD_EINVAL, "iterations < 0");
    const size_t n = bt->boxes.size();
    for (size_t i = 0; i < n; ++i)
        if (fields9[i] && bt->boxes[i]->prm.init_mode != EHD_INIT_FIELDS)
            return set_error(EHD_EINVAL, "box %zu: the consistent start needs EHD_INIT_FIELDS", i);
    int rc = ehd_batch_init(bt, fields9);
    if (rc || iterations == 0) return rc;
    // targets of the boxes that have fields; TARGET item tables of the whole batch
    std::vector<double*> tg(n, nullptr);
    struct Release {
        std::vector<double*>& tg;
        hipStream_t st;
        ~Release()
        {
            (void)hipStreamSynchronize(st);
            for (double* p : tg)
                if (p) (void)hipFree(p);
        }
    } release{tg, bt->st};
    const int src[4] = {3, 4, 5, 1};   // fields9 -> ux, uy, uz, q
    for (size_t i = 0; i < n; ++i) {
        if (!fields9[i]) continue;
        const size_t S = bt->boxes[i]->S;
        if (hipMalloc(&tg[i], 4 * S * sizeof(double)) != hipSuccess) {
            tg[i] = nullptr;
            return set_error(EHD_ENOMEM, "target allocation failed");
        }
        for (int k = 0; k < 4; ++k)
            HIPCHK(hipMemcpyAsync(tg[i] + k * S, fields9[i] + src[k] * S, S * sizeof(double), hipMemcpyHostToDevice,
                                  bt->st));
    }
    // two launches per iteration: the targeted boxes through the TARGET
    // kernel, the others through the ordinary kernel (so that they step
    // exactly as they do in ehd_batch_step); each set has its own item
    // tables and tile numbering
    const size_t isz = lbm_f32s_batch_item_size();
    struct Set {
        void* items[2] = {nullptr, nullptr};
        int n = 0;
        unsigned tiles = 0;
    } sets[2];   // 0: ordinary, 1: targeted
    struct ReleaseSets {
        Set* s;
        hipStream_t st;
        ~ReleaseSets()
        {
            (void)hipStreamSynchronize(st);
            for (int k = 0; k < 2; ++k)
                for (int c = 0; c < 2; ++c)
                    if (s[k].items[c]) (void)hipFree(s[k].items[c]);
        }
    } release_sets{sets, bt->st};
    for (int k = 0; k < 2; ++k) {
        for (int cur = 0; cur < 2; ++cur) {
            std::vector<unsigned char> h;
            unsigned tile0 = 0;
            int m = 0;
            for (size_t i = 0; i < n; ++i) {
                if ((tg[i] != nullptr) != (k == 1)) continue;
                ehd_box* b = bt->boxes[i];
                h.resize(isz * (m + 1));
                lbm_f32s_batch_item(h.data() + m * isz, b->lp, pop_bufs_at(b, cur), macro_bufs(b), tile0, tg[i]);
                tile0 += (unsigned)((b->S + 255) / 256);
                ++m;
            }
            sets[k].n = m;
            sets[k].tiles = tile0;
            if (m == 0) continue;
            if (hipMalloc(&sets[k].items[cur], h.size()) != hipSuccess) {
                sets[k].items[cur] = nullptr;
                return set_error(EHD_ENOMEM, "item table allocation failed");
            }
            HIPCHK(hipMemcpyAsync(sets[k].items[cur], h.data(), h.size(), hipMemcpyHostToDevice, bt->st));
            HIPCHK(hipStreamSynchronize(bt->st));
        }
    }
    for (ehd_box* b : bt->boxes) HIPCHK(hipStreamSynchronize(b->st));
    int cur = bt->boxes[0]->cur;
    bool direct = bt->boxes[0]->direct, ebuf = bt->boxes[0]->ebuf;
    for (int i = 0; i < iterations; ++i) {
        const bool macro = (i == iterations - 1);
        for (int k = 0; k < 2; ++k)
            if (sets[k].n > 0)
                HIPCHK(launch_lbm_f32s_batch(sets[k].items[cur], sets[k].n, sets[k].tiles, 0, bt->lbm_chunk, direct,
                                             ebuf, macro, bt->st, k == 1));
        if ((rc = batch_poisson(bt))) return rc;
        cur = 1 - cur;
        direct = false;
        ebuf = false;
        for (ehd_box* b : bt->boxes) {
            b->step += 1;
            b->t = b->t + b->der.dt;
        }
    }
    for (ehd_box* b : bt->boxes) {
        b->cur = cur;
        b->direct = direct;
        b->ebuf = ebuf;
    }
    HIPCHK(hipStreamSynchronize(bt->st));
    return EHD_OK;
}

int ehd_batch_step(ehd_batch* bt, int nsteps)
{
    if (!bt) return set_error(EHD_EINVAL, "null batch");
    if (nsteps < 0) return set_error(EHD_EINVAL, "nsteps < 0");
    if (nsteps == 0) return EHD_OK;
    int rc = check_uniform(bt);
    if (rc) return rc;
    for (ehd_box* b : bt->boxes) HIPCHK(hipStreamSynchronize(b->st));
    int cur = bt->boxes[0]->cur;
    bool direct = bt->boxes[0]->direct, ebuf = bt->boxes[0]->ebuf;
    const bool graph = (bt->flags & EHD_BATCH_GRAPH) != 0;
    int i = 0;
    auto advance = [&](int n) {
        for (ehd_box* b : bt->boxes)
            for (int k = 0; k < n; ++k) {
                b->step += 1;
                b->t = b->t + b->der.dt;   // main.cu:208, one addition per step as a single box does
            }
    };
    while (i < nsteps) {
        const bool last = (i == nsteps - 1);
        if (graph && !direct && !ebuf && i + 2 <= nsteps - 1) {
            if (!bt->exec[cur] && (rc = capture_pair(bt, cur))) return rc;
            HIPCHK(hipGraphLaunch(bt->exec[cur], bt->st));
            i += 2;
            advance(2);
            continue;
        }
        if ((rc = batch_step_once(bt, cur, direct, ebuf, last))) return rc;
        cur = 1 - cur;
        direct = false;
        ebuf = false;
        i += 1;
        advance(1);
    }
    for (ehd_box* b : bt->boxes) {
        b->cur = cur;
        b->direct = direct;
        b->ebuf = ebuf;
    }
    HIPCHK(hipStreamSynchronize(bt->st));
    return EHD_OK;
}

int ehd_batch_diagnostics(ehd_batch* bt, ehd_diag* out)
{
    if (!bt || !out) return set_error(EHD_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(bt->st));
    int bad = 0;
    long long nodes = 0;
    for (size_t i = 0; i < bt->boxes.size(); ++i) {
        const int rc = ehd_diagnostics(bt->boxes[i], &out[i]);
        if (rc == EHD_ENONFINITE) {
            ++bad;
            nodes += out[i].nonfinite;
        } else if (rc) {
            return rc;
        }
    }
    if (bad) return set_error(EHD_ENONFINITE, "watchdog: %d of %zu boxes non-finite (%lld nodes)", bad,
                              bt->boxes.size(), nodes);
    return EHD_OK;
}

int ehd_batch_run_to_steady(ehd_batch* bt, int check_every, double tol, long long max_steps, ehd_steady* out)
{
    if (!bt || !out) return set_error(EHD_EINVAL, "null argument");
    if (check_every < 1 || !(tol >= 0.0) || max_steps < 0) return set_error(EHD_EINVAL, "bad steady-state criteria");
    const size_t n = bt->boxes.size();
    std::vector<ehd_diag> d(n);
    // a box the watchdog flags is marked failed and left out of the test;
    // the others go on (the boxes' buffers are disjoint)
    auto diagnose = [&]() -> int {
        const int r = ehd_batch_diagnostics(bt, d.data());
        return (r == EHD_ENONFINITE) ? EHD_OK : r;
    };
    int rc = diagnose();
    if (rc) return rc;
    std::vector<double> prev(n);
    size_t done = 0;
    for (size_t i = 0; i < n; ++i) {
        prev[i] = d[i].nusselt;
        out[i] = ehd_steady{max_steps, 0, 0, d[i].nusselt, d[i].current, 0.0};
        if (d[i].nonfinite) {
            out[i].failed = 1;
            out[i].steps = 0;
            ++done;
        }
    }
    long long taken = 0;
    while (taken < max_steps && done < n) {
        const int k = (int)std::min<long long>(check_every, max_steps - taken);
        if ((rc = ehd_batch_step(bt, k))) return rc;
        taken += k;
        if ((rc = diagnose())) return rc;
        for (size_t i = 0; i < n; ++i) {
            if (out[i].converged || out[i].failed) continue;
            out[i].nusselt = d[i].nusselt;
            out[i].current = d[i].current;
            if (d[i].nonfinite) {
                out[i].failed = 1;
                out[i].steps = taken;
                ++done;
                continue;
            }
            const double delta = std::fabs(d[i].nusselt - prev[i]);
            prev[i] = d[i].nusselt;
            out[i].delta = delta;
            if (k == check_every && delta < tol) {
                out[i].converged = 1;
                out[i].steps = taken;
                ++done;
            }
        }
    }
    for (size_t i = 0; i < n; ++i)
        if (!out[i].converged && !out[i].failed) out[i].steps = taken;
    return EHD_OK;
}

ehd_box* ehd_batch_box(ehd_batch* bt, int i)
{
    if (!bt || i < 0 || i >= (int)bt->boxes.size()) {
        set_error(EHD_EINVAL, "no box %d in the batch", i);
        return nullptr;
    }
    return bt->boxes[i];
}

int ehd_batch_tuning(ehd_batch* bt, ehd_tuning* out)
{
    if (!bt || !out) return set_error(EHD_EINVAL, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->lbm_tiles = (int)bt->total_tiles;
    out->lbm_grid = (int)(bt->lbm_grid ? bt->lbm_grid : bt->total_tiles);
    out->lbm_ms = bt->lbm_ms;
    out->lbm_chunk = (int)bt->lbm_chunk;
    out->chains = bt->chains ? 1 : 0;
    out->step_ms[0] = bt->step_ms[0];
    out->step_ms[1] = bt->step_ms[1];
    out->ngroups = (int)bt->groups.size();
    for (size_t i = 0; i < bt->groups.size(); ++i) {
        const Group& g = bt->groups[i];
        const PoissonShape ps = poisson_shape(bt->prec, g.g.nx, g.g.ny, g.g.max_n + 2);
        out->group[i].nx = g.g.nx;
        out->group[i].ny = g.g.ny;
        out->group[i].nboxes = (int)g.members.size();
        out->group[i].x_lines = g.g.lpb ? g.g.lpb : ps.max_x_lines;
        out->group[i].y_lines = g.g.ylines ? g.g.ylines : ps.max_y_lines;
        out->group[i].ms = i < bt->group_ms.size() ? bt->group_ms[i] : 0.0;
    }
    return EHD_OK;
}

}  // extern "C"
