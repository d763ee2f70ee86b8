// qbx_server.hip -- continuous-batching server over the exhaustive walk kernels
// (C ABI in include/qbx.h, "serving").
//
// A worker thread owns two streams for scheduling steps, one stream for admissions, and a
// pool of device slots (tables and reduction state of one request each).
//   admission  oldest first, while a slot is free and the admitted work (sum of 2^n) stays
//              within the budget; all requests admitted together are staged through one
//              pinned buffer, one host-to-device copy and one prep launch.
//   step       one finite persistent walk launch over chunks of every admitted request:
//              requests with little work left go first (short requests are not held behind
//              long ones), the oldest request always gets a share (no starvation).  A
//              request may span several steps; its on-device reduction state (running
//              minimum, candidate list, done counter) carries over and the chunk that
//              finishes last writes its answer, exactly as in a single solve.
//   overlap    steps alternate between the two streams, so the tail of one step overlaps
//              the next; a step's result read-back waits for the previous step too.
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
    std::vector<char> Q;   // n*n elements of the kind's type
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
    hipStream_t st[2] = {nullptr, nullptr};
    hipStream_t st_admit = nullptr;
    hipEvent_t ev_step[2] = {nullptr, nullptr};
    hipEvent_t ev_start[2] = {nullptr, nullptr};
    hipEvent_t ev_admit = nullptr;
    hipEvent_t ev_stage[kStageSlots] = {nullptr, nullptr};
    bool stage_used[kStageSlots] = {false, false};
    // device
    char* d_slots = nullptr;
    InstDev* d_inst = nullptr;
    ResultDev* d_results = nullptr;
    char* d_stage[kStageSlots] = {nullptr, nullptr};
    TileRange* d_ranges[2] = {nullptr, nullptr};
    unsigned long long* d_tickets[2] = {nullptr, nullptr};
    size_t stage_bytes = 0;
    // pinned host
    char* h_stage[kStageSlots] = {nullptr, nullptr};
    TileRange* h_ranges[2] = {nullptr, nullptr};
    ResultDev* h_results[2] = {nullptr, nullptr};
    // shared with submitters
    std::mutex m;
    std::condition_variable cv_in, cv_out;
    std::deque<Req*> incoming;
    std::deque<qbx_request_result_t> done;
    bool stop = false;
    uint64_t next_id = 1;
    std::thread worker;
    Clock::time_point t0;
    // worker only
    std::deque<Req*> waiting;
    std::vector<Req*> active;
    std::vector<int> free_slots;
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

void finish_request(qbx_server* S, Req* r, const ResultDev* res) {
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
    {
        std::lock_guard<std::mutex> g(S->m);
        S->done.push_back(out);
    }
    S->cv_out.notify_all();
    S->n_completed++;
    delete r;
}

int admit(qbx_server* S, std::vector<Req*>& batch) {
    const int b = S->stage_next;
    S->stage_next = (b + 1) % kStageSlots;
    if (S->stage_used[b]) HIP_TRY(hipEventSynchronize(S->ev_stage[b]));   // staging buffer free again
    char* h = S->h_stage[b];
    char* dv = S->d_stage[b];
    const size_t off_idx = align256(sizeof(InstDev) * S->opts.max_inflight);
    const size_t off_q = off_idx + align256(4 * (size_t)S->opts.max_inflight);
    const size_t qstride = align256((size_t)S->opts.max_n * S->opts.max_n * 8);
    InstDev* hi = (InstDev*)h;
    uint32_t* hidx = (uint32_t*)(h + off_idx);
    size_t used = off_q;
    for (size_t j = 0; j < batch.size(); ++j) {
        Req* r = batch[j];
        const Plan& p = r->plan;
        char* base = S->d_slots + (size_t)r->slot * S->L.bytes;
        InstDev& I = hi[j];
        memset(&I, 0, sizeof I);
        I.n = p.n; I.B = p.B; I.L = p.L; I.M = p.M; I.kind = r->kind; I.status = 0;
        I.tiles = p.tiles; I.tpc = p.tpc; I.chunks = p.chunks; I.expect = p.chunks; I.flags = plan_flags(p, 0);
        I.Q = dv + off_q + j * qstride;
        I.W = base + S->L.off_W;
        I.d = base + S->L.off_d;
        I.delta = base + S->L.off_delta;
        I.red = (RedDev*)(base + S->L.off_red);
        I.cand = (uint32_t*)(base + S->L.off_cand);
        I.slots = (SlotDev*)(base + S->L.off_slots);
        I.result = S->d_results + r->slot;
        hidx[j] = (uint32_t)r->slot;
        memcpy(h + off_q + j * qstride, r->Q.data(), r->Q.size());
        used = off_q + (j + 1) * qstride;
        r->Q.clear();
        r->Q.shrink_to_fit();
    }
    HIP_TRY(hipMemcpyAsync(dv, h, used, hipMemcpyHostToDevice, S->st_admit));
    hipLaunchKernelGGL(qbx_prep, dim3((unsigned)batch.size()), dim3(256), 0, S->st_admit, S->d_inst,
                       (const uint32_t*)(dv + off_idx), (unsigned long long*)nullptr, (const InstDev*)dv);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(S->ev_admit, S->st_admit));
    HIP_TRY(hipEventRecord(S->ev_stage[b], S->st_admit));
    S->stage_used[b] = true;
    S->n_admissions += batch.size();
    return QBX_OK;
}

// Marks every request whose chunks have all run in completed steps as done.
int complete_step(qbx_server* S, int64_t k) {
    const int s = (int)(k & 1);
    HIP_TRY(hipEventSynchronize(S->ev_step[s]));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, S->ev_start[s], S->ev_step[s]) == hipSuccess)
        S->busy_ms.store(S->busy_ms.load(std::memory_order_relaxed) + ms, std::memory_order_relaxed);
    std::vector<Req*> keep;
    keep.reserve(S->active.size());
    for (Req* r : S->active) {
        if (r->next_chunk >= r->plan.chunks && r->last_step <= k) {
            const ResultDev res = S->h_results[s][r->slot];
            S->free_slots.push_back(r->slot);
            S->inflight_work -= r->work;
            finish_request(S, r, &res);
        } else {
            keep.push_back(r);
        }
    }
    S->active.swap(keep);
    return QBX_OK;
}

int launch_step(qbx_server* S, int64_t k, std::vector<std::pair<Req*, uint32_t>>& take) {
    const int s = (int)(k & 1);
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
    // the result copy must also see chunks of earlier steps finished on the other stream
    if (k > 0) HIP_TRY(hipStreamWaitEvent(st, S->ev_step[s ^ 1], 0));
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
            for (Req* r : S->waiting) { r->status = S->worker_rc; finish_request(S, r, nullptr); }
            S->waiting.clear();
            for (Req* r : S->active) { r->status = S->worker_rc; finish_request(S, r, nullptr); }
            S->active.clear();
            if (stopping) break;
            continue;
        }
        // admission: oldest first, slots and work budget permitting
        std::vector<Req*> batch;
        while (!S->waiting.empty()) {
            Req* r = S->waiting.front();
            if (r->status != QBX_OK) {
                S->waiting.pop_front();
                finish_request(S, r, nullptr);
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
        if (!batch.empty()) {
            const int rc = admit(S, batch);
            if (rc != QBX_OK) { S->worker_rc = rc; S->worker_msg = "admission failed"; continue; }
        }
        // make room for the next step's buffers
        if (inflight.size() >= 2) {
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
        HIP_TRY(hipMalloc((void**)&S->d_stage[b], S->stage_bytes));
        HIP_TRY(hipHostMalloc((void**)&S->h_stage[b], S->stage_bytes, hipHostMallocDefault));
        HIP_TRY(hipEventCreateWithFlags(&S->ev_stage[b], hipEventDisableTiming));
    }
    for (int s = 0; s < 2; ++s) {
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
    for (int v = 0; v < kNumVariants; ++v) {
        int b = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kVariants[v].fn, 64, 0));
        S->blocks_per_cu[v] = b > 0 ? b : 1;
    }
    for (int i = K - 1; i >= 0; --i) S->free_slots.push_back(i);
    return QBX_OK;
}

void server_free(qbx_server* S) {
    for (int s = 0; s < 2; ++s) {
        if (S->st[s]) (void)hipStreamSynchronize(S->st[s]);
    }
    if (S->st_admit) (void)hipStreamSynchronize(S->st_admit);
    for (int b = 0; b < kStageSlots; ++b) {
        if (S->d_stage[b]) (void)hipFree(S->d_stage[b]);
        if (S->h_stage[b]) (void)hipHostFree(S->h_stage[b]);
        if (S->ev_stage[b]) (void)hipEventDestroy(S->ev_stage[b]);
    }
    for (int s = 0; s < 2; ++s) {
        if (S->d_ranges[s]) (void)hipFree(S->d_ranges[s]);
        if (S->d_tickets[s]) (void)hipFree(S->d_tickets[s]);
        if (S->h_ranges[s]) (void)hipHostFree(S->h_ranges[s]);
        if (S->h_results[s]) (void)hipHostFree(S->h_results[s]);
        if (S->ev_step[s]) (void)hipEventDestroy(S->ev_step[s]);
        if (S->ev_start[s]) (void)hipEventDestroy(S->ev_start[s]);
        if (S->st[s]) (void)hipStreamDestroy(S->st[s]);
    }
    if (S->ev_admit) (void)hipEventDestroy(S->ev_admit);
    if (S->st_admit) (void)hipStreamDestroy(S->st_admit);
    if (S->d_slots) (void)hipFree(S->d_slots);
    if (S->d_inst) (void)hipFree(S->d_inst);
    if (S->d_results) (void)hipFree(S->d_results);
}

// A request for problem Q of size n, planned and copied (not yet queued).
template <class T>
int make_request(qbx_server* S, int kind, const T* Q, int n, int tie_B, int index, Req** out) {
    if (n < 1 || n > S->opts.max_n)
        return set_err(QBX_EINVAL, "request %d: n=%d outside [1, %d] (server max_n)", index, n, S->opts.max_n);
    Req* r = new Req();
    r->kind = kind;
    r->n = n;
    r->tie_B = tie_B;
    r->work = ldexp(1.0, n);
    qbx_opts_t o = {tie_B, S->opts.walk_bits, -1, 0};
    const int rc = plan_problem(kind, n, o, PLAN_BATCH, index, r->plan, &S->walk);
    if (rc != QBX_OK) { delete r; return rc; }
    if (r->plan.chunks > S->L.max_chunks) {
        delete r;
        return set_err(QBX_EINVAL, "request %d: walk_bits too small for the slot size", index);
    }
    r->Q.resize(sizeof(T) * (size_t)n * n);
    memcpy(r->Q.data(), Q, r->Q.size());
    *out = r;
    return QBX_OK;
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
    Req* r = nullptr;
    const int rc = make_request<T>(S, kind, Q, n, tie_B, 0, &r);
    if (rc != QBX_OK) return rc;
    return enqueue_requests(S, &r, 1, id);
}

template <class T>
int submit_many(qbx_server* S, int kind, int count, const T* Qpacked, const int64_t* offsets, const int32_t* ns,
                int tie_B, uint64_t* ids) {
    if (!S || count < 0 || (count > 0 && (!Qpacked || !offsets || !ns || !ids)))
        return set_err(QBX_EINVAL, "null argument");
    std::vector<Req*> reqs((size_t)count, nullptr);
    for (int i = 0; i < count; ++i) {
        const int rc = make_request<T>(S, kind, Qpacked + offsets[i], ns[i], tie_B, i, &reqs[i]);
        if (rc != QBX_OK) {
            for (int j = 0; j < i; ++j) delete reqs[j];
            return rc;
        }
    }
    return count ? enqueue_requests(S, reqs.data(), count, ids) : QBX_OK;
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
    qbx_server_opts_t o = {40, 256, 1099511627776.0 /* 2^40 */, 17179869184.0 /* 2^34 */, -1, 0};
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
    delete S;
    if (rc != QBX_OK) return set_err(rc, "server worker failed: %s", msg.c_str());
    return QBX_OK;
}

}  // extern "C"
