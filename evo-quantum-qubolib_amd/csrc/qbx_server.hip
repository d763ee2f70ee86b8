// qbx_server.hip -- continuous-batching server over the exhaustive walk kernels
// (C ABI in include/qbx.h, "serving").
//
// A worker thread owns three streams for scheduling steps, one stream for admissions, and a
// pool of device slots (tables and reduction state of one request each).
//   admission  oldest first, while a slot is free and the admitted work (sum of 2^n) stays
//              within the budget; all requests admitted together are staged in one mapped
//              pinned buffer that their prep launch reads directly (no copy call: the
//              runtime's host-to-device copy has blocked the calling thread for 6-7 ms now
//              and then, and every request behind it waited).
//   step       one finite persistent walk launch over chunks of every admitted request:
//              requests with little work left go first (short requests are not held behind
//              long ones), the oldest request always gets a share (no starvation).  A
//              request may span several steps; its on-device reduction state (running
//              minimum, candidate list, done counter) carries over and the chunk that
//              finishes last writes its answer, exactly as in a single solve.
//   overlap    steps rotate over the three streams, up to three in flight: the host work
//              between one step's completion and the next launch overlaps the two steps
//              still queued; a step's result read-back waits for the previous step too.
//   completion a request is done once every step holding its chunks has completed; its
//              answer is read from the step's copy of the result array.
// Nothing on the device waits for the host: every launch is finite.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "qbx_device.h"
#include "qbx_host.h"

using namespace qbx;

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t id = 0;
    int kind = KIND_I32;
    int n = 0;
    int tie_B = -1;
    std::shared_ptr<char[]> qblk;   // matrices of the submit call (shared by its requests)
    size_t q_off = 0, q_bytes = 0;   // this request's n*n elements of the kind's type
    double t_submit = 0, t_admit = 0;
    int slot = -1;
    Plan plan;
    uint32_t next_chunk = 0;
    int64_t last_step = -1;
    double work = 0;       // 2^n
    int status = QBX_OK;
    std::string msg;
};

// per-slot device arena (sized for max_n; 8-byte elements serve both kinds)
struct SlotLayout {
    size_t off_W, off_d, off_delta, off_red, off_cand, off_slots, bytes;
    uint32_t max_chunks;
};

constexpr int kStageSlots = 2;
// Steps in flight: the worker's host work for a step (completions, admission, chunk choice,
// launch) overlaps the device work of the other queued steps.
constexpr int kDepth = 3;

}  // namespace

struct qbx_server {
    qbx_ctx* ctx = nullptr;
    int device = 0;
    qbx_server_opts_t opts;
    // walk bits per n: the built-in batch rule.  The context's tuned batch table is not used:
    // it maximises the throughput of full batches with few long lanes per problem, which
    // lengthens a lone request's walk (n = 20: 0.15 ms -> 0.68 ms p50 at 20k requests/s).
    WalkTable walk;
    SlotLayout L;
    int blocks_per_cu[kNumVariants];
    hipStream_t st[kDepth] = {};
    hipStream_t st_admit = nullptr;
    hipEvent_t ev_step[kDepth] = {};
    hipEvent_t ev_start[kDepth] = {};
    hipEvent_t ev_admit = nullptr;
    hipEvent_t ev_ref = nullptr;        // recorded at start: steps' times are measured from it
    double busy_end = 0.0;              // end of the union of step intervals so far (ms from ev_ref)
    hipEvent_t ev_stage[kStageSlots] = {nullptr, nullptr};
    bool stage_used[kStageSlots] = {false, false};
    // device
    char* d_slots = nullptr;
    InstDev* d_inst = nullptr;
    ResultDev* d_results = nullptr;
    TileRange* d_ranges[kDepth] = {};
    unsigned long long* d_tickets[kDepth] = {};
    size_t stage_bytes = 0;
    // pinned host
    char* h_stage[kStageSlots] = {nullptr, nullptr};   // mapped, coherent: read by qbx_prep
    char* m_stage[kStageSlots] = {nullptr, nullptr};   // their device addresses
    TileRange* h_ranges[kDepth] = {};
    ResultDev* h_results[kDepth] = {};
    // shared with submitters
    std::mutex m;
    std::condition_variable cv_in, cv_out;
    std::deque<Req*> incoming;
    std::deque<qbx_request_result_t> done;
    std::vector<Req*> req_pool;         // finished requests, reused by submitters
    bool stop = false;
    uint64_t next_id = 1;
    std::thread worker;
    Clock::time_point t0;
    // worker only
    std::deque<Req*> waiting;
    std::vector<Req*> active;
    std::vector<int> free_slots;
    std::vector<Req*> retired;          // finished since the last publish_done
    double inflight_work = 0;
    int stage_next = 0;
    // stats
    std::atomic<uint64_t> n_submitted{0}, n_completed{0}, n_steps{0}, n_chunks{0}, n_admissions{0};
    std::atomic<double> busy_ms{0.0};   // written by the worker, read by qbx_server_stats
    int worker_rc = QBX_OK;
    std::string worker_msg;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    }
};

namespace {

size_t stage_need(const qbx_server* S) {
    // [InstDev x max_inflight][slot index x max_inflight][Q x max_inflight]
    const size_t q = (size_t)S->opts.max_n * S->opts.max_n * 8;
    return align256(sizeof(InstDev) * S->opts.max_inflight) + align256(4 * (size_t)S->opts.max_inflight) +
           align256(q) * S->opts.max_inflight;
}

// The result record of a finished request, appended to `out` (published by publish_done).
void finish_request(qbx_server* S, Req* r, const ResultDev* res, std::vector<qbx_request_result_t>& done) {
    qbx_request_result_t out;
    memset(&out, 0, sizeof out);
    out.id = r->id;
    out.n = r->n;
    out.status = r->status;
    out.t_submit_ms = r->t_submit;
    out.t_admit_ms = r->t_admit;
    if (res && r->status == QBX_OK) {
        out.status = res->status;
        out.rank = res->rank;
        out.state = res->state;
        if (r->kind == KIND_I32) {
            out.E_int = (int64_t)res->Ebits;
            out.E = (double)out.E_int;
        } else {
            double e;
            memcpy(&e, &res->Ebits, sizeof e);
            out.E = e;
            out.E_int = (int64_t)e;
        }
    }
    out.t_done_ms = S->now_ms();
    done.push_back(out);
    S->retired.push_back(r);
}

// Hands finished requests to the pollers: one lock and one wake-up for all of them.  Their
// Req records go back to the pool (allocation and release stay off the request path).
void publish_done(qbx_server* S, std::vector<qbx_request_result_t>& done) {
    if (done.empty()) return;
    constexpr size_t kPoolMax = 16384;
    {
        std::lock_guard<std::mutex> g(S->m);
        S->done.insert(S->done.end(), done.begin(), done.end());
        while (!S->retired.empty() && S->req_pool.size() < kPoolMax) {
            S->req_pool.push_back(S->retired.back());
            S->retired.pop_back();
        }
    }
    for (Req* r : S->retired) delete r;
    S->retired.clear();
    S->n_completed += done.size();
    S->cv_out.notify_all();
    done.clear();
}

int admit(qbx_server* S, std::vector<Req*>& batch) {
    const int b = S->stage_next;
    S->stage_next = (b + 1) % kStageSlots;
    if (S->stage_used[b]) HIP_TRY(hipEventSynchronize(S->ev_stage[b]));   // staging buffer free again
    char* h = S->h_stage[b];
    char* dv = S->m_stage[b];
    const size_t off_idx = align256(sizeof(InstDev) * S->opts.max_inflight);
    const size_t off_q = off_idx + align256(4 * (size_t)S->opts.max_inflight);
    InstDev* hi = (InstDev*)h;
    uint32_t* hidx = (uint32_t*)(h + off_idx);
    size_t used = off_q;    // matrices packed back to back (256-B aligned), only what they hold
    for (size_t j = 0; j < batch.size(); ++j) {
        Req* r = batch[j];
        const Plan& p = r->plan;
        char* base = S->d_slots + (size_t)r->slot * S->L.bytes;
        InstDev& I = hi[j];
        memset(&I, 0, sizeof I);
        I.n = p.n; I.B = p.B; I.L = p.L; I.M = p.M; I.kind = r->kind; I.status = 0;
        I.tiles = p.tiles; I.tpc = p.tpc; I.chunks = p.chunks; I.expect = p.chunks; I.flags = plan_flags(p, 0);
        I.Q = dv + used;
        I.W = base + S->L.off_W;
        I.d = base + S->L.off_d;
        I.delta = base + S->L.off_delta;
        I.red = (RedDev*)(base + S->L.off_red);
        I.cand = (uint32_t*)(base + S->L.off_cand);
        I.slots = (SlotDev*)(base + S->L.off_slots);
        I.result = S->d_results + r->slot;
        hidx[j] = (uint32_t)r->slot;
        memcpy(h + used, r->qblk.get() + r->q_off, r->q_bytes);
        used = align256(used + r->q_bytes);
        r->qblk.reset();
    }
    hipLaunchKernelGGL(qbx_prep, dim3((unsigned)batch.size()), dim3(256), 0, S->st_admit, S->d_inst,
                       (const uint32_t*)(dv + off_idx), (unsigned long long*)nullptr, (const InstDev*)dv);
    HIP_TRY(hipGetLastError());
    bool seed = false;
    for (const Req* r : batch) seed = seed || (r->kind == KIND_F64 && r->plan.n >= SEED_MIN_N);
    if (seed) {   // float requests large enough: their shared bound starts from a seed
        hipLaunchKernelGGL(qbx_seed, dim3((unsigned)batch.size()), dim3(64 * SEED_WAVES), 0, S->st_admit, S->d_inst,
                           (const uint32_t*)(dv + off_idx));
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(S->ev_admit, S->st_admit));
    HIP_TRY(hipEventRecord(S->ev_stage[b], S->st_admit));
    S->stage_used[b] = true;
    S->n_admissions += batch.size();
    return QBX_OK;
}

// Marks every request whose chunks have all run in completed steps as done.
int complete_step(qbx_server* S, int64_t k) {
    const int s = (int)(k % kDepth);
    HIP_TRY(hipEventSynchronize(S->ev_step[s]));
    // device busy time: the union of the steps' intervals (steps on different streams overlap,
    // and complete in launch order)
    float t_beg = 0.f, t_end = 0.f;
    if (hipEventElapsedTime(&t_beg, S->ev_ref, S->ev_start[s]) == hipSuccess &&
        hipEventElapsedTime(&t_end, S->ev_ref, S->ev_step[s]) == hipSuccess) {
        const double b = std::max<double>(t_beg, S->busy_end);
        if (t_end > b) S->busy_ms.store(S->busy_ms.load(std::memory_order_relaxed) + (t_end - b), std::memory_order_relaxed);
        S->busy_end = std::max<double>(S->busy_end, t_end);
    }
    std::vector<Req*> keep;
    keep.reserve(S->active.size());
    std::vector<qbx_request_result_t> done;
    for (Req* r : S->active) {
        if (r->next_chunk >= r->plan.chunks && r->last_step <= k) {
            const ResultDev res = S->h_results[s][r->slot];
            S->free_slots.push_back(r->slot);
            S->inflight_work -= r->work;
            finish_request(S, r, &res, done);
        } else {
            keep.push_back(r);
        }
    }
    S->active.swap(keep);
    publish_done(S, done);
    return QBX_OK;
}

int launch_step(qbx_server* S, int64_t k, std::vector<std::pair<Req*, uint32_t>>& take) {
    const int s = (int)(k % kDepth);
    // group by kernel variant; one range per request
    std::stable_sort(take.begin(), take.end(), [](const std::pair<Req*, uint32_t>& a, const std::pair<Req*, uint32_t>& b) {
        return a.first->plan.variant < b.first->plan.variant;
    });
    TileRange* hr = S->h_ranges[s];
    struct Group { int variant, first, cnt; unsigned long long total; };
    std::vector<Group> groups;
    unsigned long long prefix = 0;
    for (size_t j = 0; j < take.size(); ++j) {
        Req* r = take[j].first;
        if (groups.empty() || groups.back().variant != r->plan.variant) {
            groups.push_back({r->plan.variant, (int)j, 0, 0});
            prefix = 0;
        }
        TileRange& R = hr[j];
        R.inst = (uint32_t)r->slot;
        R.c_begin = r->next_chunk;
        R.c_count = take[j].second;
        R.pad = 0;
        R.prefix = prefix;
        prefix += take[j].second;
        groups.back().cnt++;
        groups.back().total = prefix;
        r->next_chunk += take[j].second;
        r->last_step = k;
        S->n_chunks += take[j].second;
    }
    hipStream_t st = S->st[s];
    HIP_TRY(hipStreamWaitEvent(st, S->ev_admit, 0));
    HIP_TRY(hipMemcpyAsync(S->d_ranges[s], hr, sizeof(TileRange) * take.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(S->d_tickets[s], 0, sizeof(unsigned long long) * MAX_GROUPS, st));
    HIP_TRY(hipEventRecord(S->ev_start[s], st));
    for (size_t g = 0; g < groups.size(); ++g) {
        const Group& G = groups[g];
        const unsigned long long full = (unsigned long long)S->blocks_per_cu[G.variant] * S->ctx->prop.multiProcessorCount;
        const unsigned long long grid = std::max(1ull, std::min(full, G.total));
        walk_fn_t fn = (walk_fn_t)kVariants[G.variant].fn;
        hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(64), 0, st, (const InstDev*)S->d_inst,
                           (const TileRange*)S->d_ranges[s] + G.first, G.cnt, G.total, S->d_tickets[s] + g, 0);
        HIP_TRY(hipGetLastError());
    }
    // the result copy must also see chunks of earlier steps finished on the other streams:
    // step k - 1's copy waited for step k - 2's, and so on
    if (k > 0) HIP_TRY(hipStreamWaitEvent(st, S->ev_step[(k - 1) % kDepth], 0));
    HIP_TRY(hipMemcpyAsync(S->h_results[s], S->d_results, sizeof(ResultDev) * S->opts.max_inflight,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(S->ev_step[s], st));
    S->n_steps++;
    return QBX_OK;
}

// Chunks for the next step: requests with the least work left first, the oldest admitted
// request always included; about step_states states in total.
void choose_chunks(qbx_server* S, std::vector<std::pair<Req*, uint32_t>>& take) {
    take.clear();
    std::vector<Req*> order;
    for (Req* r : S->active)
        if (r->next_chunk < r->plan.chunks) order.push_back(r);
    if (order.empty()) return;
    Req* oldest = order.front();
    auto left = [](const Req* r) {
        return (double)(r->plan.chunks - r->next_chunk) * ldexp(1.0, chunk_work_log2(r->plan));
    };
    std::stable_sort(order.begin(), order.end(), [&](const Req* a, const Req* b) { return left(a) < left(b); });
    double budget = S->opts.step_states;
    bool oldest_in = false;
    for (Req* r : order) {
        const double cw = ldexp(1.0, chunk_work_log2(r->plan));
        const uint32_t rem = r->plan.chunks - r->next_chunk;
        uint32_t k;
        if (budget <= 0) {
            if (r != oldest || oldest_in) continue;
            k = 1;
        } else {
            const double fit = floor(budget / cw);
            k = (uint32_t)std::min<double>(rem, std::max(1.0, fit));
        }
        take.emplace_back(r, k);
        budget -= k * cw;
        if (r == oldest) oldest_in = true;
    }
}

void worker_main(qbx_server* S) {
    if (hipSetDevice(S->device) != hipSuccess) {
        S->worker_rc = QBX_EDEVICE;
        S->worker_msg = "hipSetDevice failed in the server thread";
    }
    int64_t launched = 0;
    std::deque<int64_t> inflight;
    std::vector<std::pair<Req*, uint32_t>> take;
    for (;;) {
        bool stopping;
        {
            std::unique_lock<std::mutex> lk(S->m);
            if (S->incoming.empty() && S->waiting.empty() && S->active.empty() && inflight.empty() && !S->stop)
                S->cv_in.wait_for(lk, std::chrono::milliseconds(2));
            while (!S->incoming.empty()) {
                S->waiting.push_back(S->incoming.front());
                S->incoming.pop_front();
            }
            stopping = S->stop;
        }
        if (S->worker_rc != QBX_OK) {   // device failure: fail everything still queued
            std::vector<qbx_request_result_t> failed;
            for (Req* r : S->waiting) { r->status = S->worker_rc; finish_request(S, r, nullptr, failed); }
            S->waiting.clear();
            for (Req* r : S->active) { r->status = S->worker_rc; finish_request(S, r, nullptr, failed); }
            S->active.clear();
            publish_done(S, failed);
            if (stopping) break;
            continue;
        }
        // admission: oldest first, slots and work budget permitting
        std::vector<Req*> batch;
        std::vector<qbx_request_result_t> rejected;
        while (!S->waiting.empty()) {
            Req* r = S->waiting.front();
            if (r->status != QBX_OK) {
                S->waiting.pop_front();
                finish_request(S, r, nullptr, rejected);
                continue;
            }
            if (S->free_slots.empty()) break;
            if ((!S->active.empty() || !batch.empty()) && S->inflight_work + r->work > S->opts.work_budget) break;
            S->waiting.pop_front();
            r->slot = S->free_slots.back();
            S->free_slots.pop_back();
            r->t_admit = S->now_ms();
            S->inflight_work += r->work;
            batch.push_back(r);
            S->active.push_back(r);
        }
        publish_done(S, rejected);
        if (!batch.empty()) {
            const int rc = admit(S, batch);
            if (rc != QBX_OK) { S->worker_rc = rc; S->worker_msg = "admission failed"; continue; }
        }
        // make room for the next step's buffers
        if (inflight.size() >= (size_t)kDepth) {
            const int rc = complete_step(S, inflight.front());
            inflight.pop_front();
            if (rc != QBX_OK) { S->worker_rc = rc; S->worker_msg = "step failed"; }
            continue;
        }
        choose_chunks(S, take);
        if (take.empty()) {
            if (!inflight.empty()) {
                const int rc = complete_step(S, inflight.front());
                inflight.pop_front();
                if (rc != QBX_OK) { S->worker_rc = rc; S->worker_msg = "step failed"; }
                continue;
            }
            if (stopping && S->waiting.empty() && S->active.empty()) break;
            continue;
        }
        const int rc = launch_step(S, launched, take);
        if (rc != QBX_OK) { S->worker_rc = rc; S->worker_msg = "step launch failed"; continue; }
        inflight.push_back(launched);
        launched++;
    }
    while (!inflight.empty()) {
        complete_step(S, inflight.front());
        inflight.pop_front();
    }
}

int server_alloc(qbx_server* S) {
    const int max_n = S->opts.max_n;
    // largest chunk count of any n <= max_n under the serving tiling
    qbx_opts_t o = {-1, S->opts.walk_bits, -1, 0};
    uint32_t max_chunks = 1;
    for (int kind = 0; kind < 2; ++kind)
        for (int n = 1; n <= max_n; ++n) {
            Plan p;
            if (plan_problem(kind, n, o, PLAN_BATCH, 0, p, &S->walk) == QBX_OK) max_chunks = std::max(max_chunks, p.chunks);
        }
    SlotLayout& L = S->L;
    size_t off = 0;
    L.off_W = off; off = align256(off + table_W_bytes(KIND_F64, max_n));
    L.off_d = off; off = align256(off + table_d_bytes(KIND_F64));
    L.off_delta = off; off = align256(off + table_delta_bytes(KIND_F64, 7));
    L.off_red = off; off = align256(off + sizeof(RedDev));
    L.off_cand = off; off = align256(off + 4 * (size_t)max_chunks);
    L.off_slots = off; off = align256(off + sizeof(SlotDev) * (size_t)max_chunks);
    L.bytes = off;
    L.max_chunks = max_chunks;
    const int K = S->opts.max_inflight;
    S->stage_bytes = stage_need(S);
    if (hipMalloc((void**)&S->d_slots, L.bytes * K) != hipSuccess) return set_err(QBX_ENOMEM, "server slot arena");
    HIP_TRY(hipMalloc((void**)&S->d_inst, sizeof(InstDev) * K));
    HIP_TRY(hipMalloc((void**)&S->d_results, sizeof(ResultDev) * K));
    HIP_TRY(hipMemset(S->d_results, 0, sizeof(ResultDev) * K));
    for (int b = 0; b < kStageSlots; ++b) {
        HIP_TRY(hipHostMalloc((void**)&S->h_stage[b], S->stage_bytes, hipHostMallocMapped | hipHostMallocCoherent));
        HIP_TRY(hipHostGetDevicePointer((void**)&S->m_stage[b], S->h_stage[b], 0));
        HIP_TRY(hipEventCreateWithFlags(&S->ev_stage[b], hipEventDisableTiming));
    }
    for (int s = 0; s < kDepth; ++s) {
        HIP_TRY(hipStreamCreateWithFlags(&S->st[s], hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&S->ev_step[s]));
        HIP_TRY(hipEventCreate(&S->ev_start[s]));
        HIP_TRY(hipMalloc((void**)&S->d_ranges[s], sizeof(TileRange) * K));
        HIP_TRY(hipMalloc((void**)&S->d_tickets[s], sizeof(unsigned long long) * MAX_GROUPS));
        HIP_TRY(hipHostMalloc((void**)&S->h_ranges[s], sizeof(TileRange) * K, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc((void**)&S->h_results[s], sizeof(ResultDev) * K, hipHostMallocDefault));
    }
    HIP_TRY(hipStreamCreateWithFlags(&S->st_admit, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&S->ev_admit, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(S->ev_admit, S->st_admit));
    HIP_TRY(hipEventCreate(&S->ev_ref));
    HIP_TRY(hipEventRecord(S->ev_ref, S->st_admit));
    for (int v = 0; v < kNumVariants; ++v) {
        int b = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kVariants[v].fn, 64, 0));
        S->blocks_per_cu[v] = b > 0 ? b : 1;
    }
    for (int i = K - 1; i >= 0; --i) S->free_slots.push_back(i);
    return QBX_OK;
}

void server_free(qbx_server* S) {
    for (int s = 0; s < kDepth; ++s) {
        if (S->st[s]) (void)hipStreamSynchronize(S->st[s]);
    }
    if (S->st_admit) (void)hipStreamSynchronize(S->st_admit);
    for (int b = 0; b < kStageSlots; ++b) {
        if (S->h_stage[b]) (void)hipHostFree(S->h_stage[b]);
        if (S->ev_stage[b]) (void)hipEventDestroy(S->ev_stage[b]);
    }
    for (int s = 0; s < kDepth; ++s) {
        if (S->d_ranges[s]) (void)hipFree(S->d_ranges[s]);
        if (S->d_tickets[s]) (void)hipFree(S->d_tickets[s]);
        if (S->h_ranges[s]) (void)hipHostFree(S->h_ranges[s]);
        if (S->h_results[s]) (void)hipHostFree(S->h_results[s]);
        if (S->ev_step[s]) (void)hipEventDestroy(S->ev_step[s]);
        if (S->ev_start[s]) (void)hipEventDestroy(S->ev_start[s]);
        if (S->st[s]) (void)hipStreamDestroy(S->st[s]);
    }
    if (S->ev_admit) (void)hipEventDestroy(S->ev_admit);
    if (S->ev_ref) (void)hipEventDestroy(S->ev_ref);
    if (S->st_admit) (void)hipStreamDestroy(S->st_admit);
    if (S->d_slots) (void)hipFree(S->d_slots);
    if (S->d_inst) (void)hipFree(S->d_inst);
    if (S->d_results) (void)hipFree(S->d_results);
}

// The plan of a request of size n, checked against the server's limits.
int plan_request(qbx_server* S, int kind, int n, int tie_B, int index, Plan& p) {
    if (n < 1 || n > S->opts.max_n)
        return set_err(QBX_EINVAL, "request %d: n=%d outside [1, %d] (server max_n)", index, n, S->opts.max_n);
    qbx_opts_t o = {tie_B, S->opts.walk_bits, -1, 0};
    const int rc = plan_problem(kind, n, o, PLAN_BATCH, index, p, &S->walk);
    if (rc != QBX_OK) return rc;
    if (p.chunks > S->L.max_chunks) return set_err(QBX_EINVAL, "request %d: walk_bits too small for the slot size", index);
    return QBX_OK;
}

// count fresh Req records: from the pool of finished ones, else newly allocated.
void new_requests(qbx_server* S, Req** out, int count) {
    int got = 0;
    {
        std::lock_guard<std::mutex> g(S->m);
        while (got < count && !S->req_pool.empty()) {
            out[got++] = S->req_pool.back();
            S->req_pool.pop_back();
        }
    }
    for (int i = 0; i < got; ++i) *out[i] = Req();
    for (; got < count; ++got) out[got] = new Req();
}

// Request r for the matrix at blk + off (not yet queued).
void fill_request(Req* r, int kind, const std::shared_ptr<char[]>& blk, size_t off, int tie_B, const Plan& plan,
                  size_t elem) {
    r->kind = kind;
    r->n = plan.n;
    r->tie_B = tie_B;
    r->work = ldexp(1.0, plan.n);
    r->plan = plan;
    r->qblk = blk;
    r->q_off = off;
    r->q_bytes = elem * (size_t)plan.n * plan.n;
}

// Queue requests (ids in order, under one lock, one wake-up of the worker), or none.
int enqueue_requests(qbx_server* S, Req** reqs, int count, uint64_t* ids) {
    {
        std::lock_guard<std::mutex> g(S->m);
        if (S->stop) {
            for (int i = 0; i < count; ++i) delete reqs[i];
            return set_err(QBX_EBUSY, "server is stopping");
        }
        const double t = S->now_ms();
        for (int i = 0; i < count; ++i) {
            reqs[i]->id = S->next_id++;
            reqs[i]->t_submit = t;
            ids[i] = reqs[i]->id;
            S->incoming.push_back(reqs[i]);
        }
    }
    S->n_submitted += count;
    S->cv_in.notify_one();
    return QBX_OK;
}

template <class T>
int submit(qbx_server* S, int kind, const T* Q, int n, int tie_B, uint64_t* id) {
    if (!S || !Q || !id) return set_err(QBX_EINVAL, "null argument");
    Plan p;
    const int rc = plan_request(S, kind, n, tie_B, 0, p);
    if (rc != QBX_OK) return rc;
    const size_t bytes = sizeof(T) * (size_t)n * n;
    std::shared_ptr<char[]> blk(new char[bytes]);
    memcpy(blk.get(), Q, bytes);
    Req* r;
    new_requests(S, &r, 1);
    fill_request(r, kind, blk, 0, tie_B, p, sizeof(T));
    return enqueue_requests(S, &r, 1, id);
}

// A burst: every request is checked first (all are queued, or none), then they are copied
// and queued in chunks (one matrix block per chunk), so the worker starts on the first
// chunk while later ones are still being copied.
template <class T>
int submit_many(qbx_server* S, int kind, int count, const T* Qpacked, const int64_t* offsets, const int32_t* ns,
                int tie_B, uint64_t* ids) {
    if (!S || count < 0 || (count > 0 && (!Qpacked || !offsets || !ns || !ids)))
        return set_err(QBX_EINVAL, "null argument");
    // one plan per size (the plan depends on n alone here), checked for every request
    Plan plans[MAX_N + 1];
    bool planned[MAX_N + 1] = {};
    for (int i = 0; i < count; ++i) {
        if (offsets[i] < 0) return set_err(QBX_EINVAL, "request %d: negative offset", i);
        const int n = ns[i];
        if (n >= 1 && n <= MAX_N && planned[n]) continue;
        Plan p;
        const int rc = plan_request(S, kind, n, tie_B, i, p);
        if (rc != QBX_OK) return rc;
        plans[n] = p;
        planned[n] = true;
    }
    constexpr int kChunk = 256;
    Req* reqs[kChunk];
    for (int i0 = 0; i0 < count; i0 += kChunk) {
        const int c = std::min(kChunk, count - i0);
        size_t total = 0;
        for (int i = 0; i < c; ++i) total += sizeof(T) * (size_t)ns[i0 + i] * ns[i0 + i];
        std::shared_ptr<char[]> blk(new char[total]);
        new_requests(S, reqs, c);
        size_t off = 0;
        for (int i = 0; i < c; ++i) {
            const Plan& p = plans[ns[i0 + i]];
            fill_request(reqs[i], kind, blk, off, tie_B, p, sizeof(T));
            memcpy(blk.get() + off, Qpacked + offsets[i0 + i], reqs[i]->q_bytes);
            off += reqs[i]->q_bytes;
        }
        const int rc = enqueue_requests(S, reqs, c, ids + i0);
        if (rc != QBX_OK) return rc;
    }
    return QBX_OK;
}

}  // namespace

extern "C" {

int qbx_server_start(qbx_ctx_t* ctx, const qbx_server_opts_t* opts, qbx_server_t** out) {
    if (!ctx || !out) return set_err(QBX_EINVAL, "null argument");
    *out = nullptr;
    qbx_server* S = new qbx_server();
    S->ctx = ctx;
    S->walk.clear();
    S->device = ctx->device;
    qbx_server_opts_t o = {40, 512, 1099511627776.0 /* 2^40 */, 17179869184.0 /* 2^34 */, -1, 0};
    if (opts) {
        if (opts->max_n > 0) o.max_n = opts->max_n;
        if (opts->max_inflight > 0) o.max_inflight = opts->max_inflight;
        if (opts->work_budget > 0) o.work_budget = opts->work_budget;
        if (opts->step_states > 0) o.step_states = opts->step_states;
        o.walk_bits = opts->walk_bits;
        o.flags = opts->flags;
    }
    if (o.max_n > MAX_N || o.flags != 0 || o.max_inflight > 65536) {
        delete S;
        return set_err(QBX_EINVAL, "bad server options (max_n <= %d, flags 0, max_inflight <= 65536)", MAX_N);
    }
    S->opts = o;
    S->t0 = Clock::now();
    if (hipSetDevice(S->device) != hipSuccess) { delete S; return set_err(QBX_EDEVICE, "hipSetDevice failed"); }
    const int rc = server_alloc(S);
    if (rc != QBX_OK) {
        server_free(S);
        delete S;
        return rc;
    }
    S->worker = std::thread(worker_main, S);
    *out = S;
    return QBX_OK;
}

int qbx_server_submit_i32(qbx_server_t* s, const int32_t* Q, int n, int tie_B, uint64_t* id) {
    return submit<int32_t>(s, KIND_I32, Q, n, tie_B, id);
}

int qbx_server_submit_f64(qbx_server_t* s, const double* Q, int n, int tie_B, uint64_t* id) {
    return submit<double>(s, KIND_F64, Q, n, tie_B, id);
}

int qbx_server_submit_many_i32(qbx_server_t* s, int count, const int32_t* Qpacked, const int64_t* offsets,
                               const int32_t* ns, int tie_B, uint64_t* ids) {
    return submit_many<int32_t>(s, KIND_I32, count, Qpacked, offsets, ns, tie_B, ids);
}

int qbx_server_submit_many_f64(qbx_server_t* s, int count, const double* Qpacked, const int64_t* offsets,
                               const int32_t* ns, int tie_B, uint64_t* ids) {
    return submit_many<double>(s, KIND_F64, count, Qpacked, offsets, ns, tie_B, ids);
}

int qbx_server_poll(qbx_server_t* S, qbx_request_result_t* out, int max, int wait_ms, int* count) {
    if (!S || !out || !count || max < 0) return set_err(QBX_EINVAL, "null argument");
    std::unique_lock<std::mutex> lk(S->m);
    if (S->done.empty() && wait_ms > 0)
        S->cv_out.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !S->done.empty(); });
    int k = 0;
    while (k < max && !S->done.empty()) {
        out[k++] = S->done.front();
        S->done.pop_front();
    }
    *count = k;
    return QBX_OK;
}

int qbx_server_stats(qbx_server_t* S, qbx_server_stats_t* out) {
    if (!S || !out) return set_err(QBX_EINVAL, "null argument");
    out->submitted = S->n_submitted;
    out->completed = S->n_completed;
    out->steps = S->n_steps;
    out->chunks = S->n_chunks;
    out->admissions = S->n_admissions;
    out->busy_ms = S->busy_ms.load(std::memory_order_relaxed);
    return QBX_OK;
}

int qbx_server_stop(qbx_server_t* S) {
    if (!S) return QBX_OK;
    {
        std::lock_guard<std::mutex> g(S->m);
        S->stop = true;
    }
    S->cv_in.notify_all();
    if (S->worker.joinable()) S->worker.join();
    const int rc = S->worker_rc;
    std::string msg = S->worker_msg;
    server_free(S);
    for (Req* r : S->incoming) delete r;
    for (Req* r : S->req_pool) delete r;
    for (Req* r : S->retired) delete r;
    delete S;
    if (rc != QBX_OK) return set_err(rc, "server worker failed: %s", msg.c_str());
    return QBX_OK;
}

}  // extern "C"
