// qbx.hip -- gfx950 kernels and the C ABI of libqbx.so (declared in include/qbx.h).
//
// Kernels
//   qbx_prep         one workgroup per problem: bit-ordered W and d tables, the inner-block
//                    step constants delta_t, the energy bound (int32 admission / FP32
//                    filter margin), and reset of the per-problem reduction state.
//                    Replaces the reference's cache building (qbf.py:246-289,
//                    qbf1.py:105-118), which there costs GEMMs over 2^B x B state tables.
//   qbx_walk_i32<M>  persistent walk over a queue of 64-lane tiles; exact int32 energies
//                    (v_add3_u32 / v_min3_i32 inner loop).  Replaces the hot loop
//                    qbf.py:188-230 and the QB_F1 kernel qbf1.py:179-249.
//   qbx_walk_f64<M>  same walk with an FP32 block filter and FP64 exact re-walks.
//
// Reduction of a problem: each tile writes its (E, rank) slot, takes an atomicMin on an
// order-preserving 64-bit key of E and, when its key is not above the previous minimum,
// appends its tile id to a candidate list.  Every tile that holds the final minimum E
// appended itself (at its commit the running minimum was >= its E), so the last tile to
// finish scans only the candidate list and writes the lexicographic (E, rank) minimum.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <stdarg.h>
#include <algorithm>
#include <string>
#include <vector>

#include "qbx_device.h"
#include "qbx_host.h"

// =============================================================================== host
using namespace qbx;

namespace qbx {
static thread_local std::string g_err;

int set_err(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
}  // namespace qbx

namespace {

int ensure_capacity(qbx_ctx* c, size_t dbytes, size_t hbytes) {
    if (dbytes > c->dcap) {
        const size_t want = std::max(dbytes, c->dcap + c->dcap / 2);
        if (c->dbuf) HIP_TRY(hipFree(c->dbuf));
        c->dbuf = nullptr;
        c->dcap = 0;
        if (hipMalloc((void**)&c->dbuf, want) != hipSuccess)
            return set_err(QBX_ENOMEM, "hipMalloc(%zu) failed", want);
        c->dcap = want;
    }
    if (hbytes > c->hcap) {
        const size_t want = std::max(hbytes, c->hcap + c->hcap / 2);
        if (c->hbuf) HIP_TRY(hipHostFree(c->hbuf));
        c->hbuf = nullptr;
        c->hcap = 0;
        if (hipHostMalloc((void**)&c->hbuf, want, hipHostMallocDefault) != hipSuccess)
            return set_err(QBX_ENOMEM, "hipHostMalloc(%zu) failed", want);
        c->hcap = want;
    }
    return QBX_OK;
}

template <class T>
int run_batch(qbx_ctx* c, int kind, int count, const T* Qpacked, const int64_t* offsets,
              const int32_t* ns, const qbx_opts_t* opts, qbx_result_t* out, bool single) {
    if (!c) return set_err(QBX_EINVAL, "null context");
    if (count <= 0) return set_err(QBX_EINVAL, "count must be positive, got %d", count);
    if (!Qpacked || !offsets || !ns || !out) return set_err(QBX_EINVAL, "null argument");
    qbx_opts_t o = {-1, -1, -1, 0};
    if (opts) o = *opts;
    if (o.flags != 0) return set_err(QBX_EINVAL, "opts.flags must be 0");
    const size_t esize = sizeof(T);

    std::vector<Plan> plans(count);
    size_t off = 0;
    const size_t off_inst = off;
    off = align256(off + sizeof(InstDev) * count);
    const size_t off_ranges = off;
    off = align256(off + sizeof(TileRange) * count);
    const PlanMode mode = single ? PLAN_SINGLE : PLAN_BATCH;
    for (int i = 0; i < count; ++i) {
        const int n = ns[i];
        if (offsets[i] < 0) return set_err(QBX_EINVAL, "problem %d: negative offset", i);
        Plan& p = plans[i];
        const int rc = plan_problem(kind, n, o, mode, i, p);
        if (rc) return rc;
        p.off_Q = off;
        off = align256(off + esize * (size_t)n * n);
    }
    const size_t stage_bytes = off;   // [instances | ranges | Q data] staged from host
    for (int i = 0; i < count; ++i) {
        Plan& p = plans[i];
        p.off_W = off;
        off = align256(off + esize * (size_t)p.n * LDW);
        p.off_d = off;
        off = align256(off + esize * 64);
        p.off_delta = off;
        off = align256(off + (kind == KIND_I32 ? 4 : 12) * (size_t)(1u << p.M));
    }
    const size_t off_red = off;
    off = align256(off + sizeof(RedDev) * count);
    for (int i = 0; i < count; ++i) {
        plans[i].off_cand = off;
        off = align256(off + 4 * (size_t)plans[i].chunks);
    }
    for (int i = 0; i < count; ++i) {
        plans[i].off_slots = off;
        off = align256(off + sizeof(SlotDev) * (size_t)plans[i].chunks);
    }
    const size_t off_res = off;
    off = align256(off + sizeof(ResultDev) * count);
    const size_t off_ticket = off;
    off = align256(off + sizeof(unsigned long long) * MAX_GROUPS);
    const size_t total_bytes = off;

    int rc = ensure_capacity(c, total_bytes, std::max(stage_bytes, sizeof(ResultDev) * (size_t)count));
    if (rc) return rc;
    char* h = c->hbuf;
    char* dv = c->dbuf;

    // order: group by kernel variant, then longest tiles first (LPT), then problem index
    std::vector<int> order(count);
    for (int i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (plans[a].variant != plans[b].variant) return plans[a].variant < plans[b].variant;
        return chunk_work_log2(plans[a]) > chunk_work_log2(plans[b]);   // longest chunks first
    });

    InstDev* hinst = (InstDev*)(h + off_inst);
    TileRange* hr = (TileRange*)(h + off_ranges);
    for (int i = 0; i < count; ++i) {
        const Plan& p = plans[i];
        InstDev& I = hinst[i];
        memset(&I, 0, sizeof I);
        I.n = p.n; I.B = p.B; I.L = p.L; I.M = p.M; I.kind = kind; I.status = 0; I.tiles = p.tiles;
        I.tpc = p.tpc; I.chunks = p.chunks;
        I.Q = dv + p.off_Q;
        I.W = dv + p.off_W;
        I.d = dv + p.off_d;
        I.delta = dv + p.off_delta;
        I.red = (RedDev*)(dv + off_red) + i;
        I.cand = (uint32_t*)(dv + p.off_cand);
        I.slots = (SlotDev*)(dv + p.off_slots);
        I.result = (ResultDev*)(dv + off_res) + i;
        memcpy(h + p.off_Q, Qpacked + offsets[i], esize * (size_t)p.n * p.n);
    }
    // one range per problem; groups of equal variant are contiguous in `order`
    struct Group { int variant; int first; int cnt; unsigned long long total; };
    std::vector<Group> groups;
    unsigned long long prefix = 0;
    for (int r = 0; r < count; ++r) {
        const int i = order[r];
        if (groups.empty() || groups.back().variant != plans[i].variant) {
            groups.push_back({plans[i].variant, r, 0, 0});
            prefix = 0;
        }
        TileRange& R = hr[r];
        R.inst = (uint32_t)i; R.c_begin = 0; R.c_count = plans[i].chunks; R.pad = 0; R.prefix = prefix;
        prefix += plans[i].chunks;
        groups.back().cnt++;
        groups.back().total = prefix;
    }
    if ((int)groups.size() > MAX_GROUPS) return set_err(QBX_EINVAL, "too many kernel variants in one batch");

    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(dv, h, stage_bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    unsigned long long* dticket = (unsigned long long*)(dv + off_ticket);
    hipLaunchKernelGGL(qbx_prep, dim3(count), dim3(256), 0, c->stream, (InstDev*)(dv + off_inst),
                       (const uint32_t*)nullptr, dticket, (const InstDev*)nullptr);
    HIP_TRY(hipGetLastError());
    for (size_t g = 0; g < groups.size(); ++g) {
        const Group& G = groups[g];
        const int v = G.variant;
        const unsigned long long grid_full = (unsigned long long)c->blocks_per_cu[v] * c->prop.multiProcessorCount;
        const unsigned long long grid = std::max(1ull, std::min(grid_full, G.total));
        walk_fn_t fn = (walk_fn_t)kVariants[v].fn;
        hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(64), 0, c->stream,
                           (const InstDev*)(dv + off_inst), (const TileRange*)(dv + off_ranges) + G.first,
                           G.cnt, G.total, dticket + g, 0);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipMemcpyAsync(h, dv + off_res, sizeof(ResultDev) * count, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    const ResultDev* hres = (const ResultDev*)h;
    int first_bad = QBX_OK;
    for (int i = 0; i < count; ++i) {
        const Plan& p = plans[i];
        qbx_result_t& R = out[i];
        memset(&R, 0, sizeof R);
        R.n = p.n; R.tie_B = p.B; R.walk_bits = p.L; R.inner_bits = p.M; R.lane_bits = p.H;
        R.states = (p.n >= 64) ? ~0ull : (1ull << p.n);
        R.tiles = p.tiles;
        R.kernel_ms = ms;
        R.status = hres[i].status;
        R.rank = hres[i].rank;
        R.state = hres[i].state;
        if (kind == KIND_I32) {
            R.E_int = (int64_t)hres[i].Ebits;
            R.E = (double)R.E_int;
        } else {
            double e;
            memcpy(&e, &hres[i].Ebits, sizeof e);
            R.E = e;
            R.E_int = (int64_t)e;
        }
        if (R.status != 0 && first_bad == QBX_OK) {
            first_bad = R.status;
            set_err(R.status, "problem %d: energy bound outside the %s path", i,
                    kind == KIND_I32 ? "int32" : "float");
        }
    }
    return first_bad;
}

}  // namespace

extern "C" {

const char* qbx_last_error(void) { return g_err.c_str(); }

int qbx_open(int device, qbx_ctx_t** out) {
    if (!out) return set_err(QBX_EINVAL, "null output pointer");
    *out = nullptr;
    int cnt = 0;
    HIP_TRY(hipGetDeviceCount(&cnt));
    if (device < 0 || device >= cnt) return set_err(QBX_EINVAL, "device %d not present (%d devices)", device, cnt);
    qbx_ctx* c = new qbx_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipGetDeviceProperties(&c->prop, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    for (int v = 0; v < kNumVariants && e == hipSuccess; ++v) {
        int b = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kVariants[v].fn, 64, 0);
        c->blocks_per_cu[v] = b > 0 ? b : 1;
    }
    if (e != hipSuccess) {
        set_err(QBX_EDEVICE, "device %d initialisation failed: %s", device, hipGetErrorString(e));
        qbx_close(c);
        return QBX_EDEVICE;
    }
    *out = c;
    return QBX_OK;
}

int qbx_close(qbx_ctx_t* c) {
    if (!c) return QBX_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->dbuf) hipFree(c->dbuf);
    if (c->hbuf) hipHostFree(c->hbuf);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return QBX_OK;
}

int qbx_device_info(qbx_ctx_t* c, qbx_devinfo_t* info) {
    if (!c || !info) return set_err(QBX_EINVAL, "null argument");
    memset(info, 0, sizeof *info);
    strncpy(info->name, c->prop.gcnArchName, sizeof(info->name) - 1);
    info->compute_units = c->prop.multiProcessorCount;
    info->clock_khz = c->prop.clockRate;
    info->simd_per_cu = 4;
    info->wave_size = c->prop.warpSize;
    info->total_mem = (int64_t)c->prop.totalGlobalMem;
    info->valu_peak_ops = (double)info->compute_units * 4.0 * 16.0 * (double)info->clock_khz * 1e3;
    return QBX_OK;
}

int qbx_solve_i32(qbx_ctx_t* c, const int32_t* Q, int n, const qbx_opts_t* opts, qbx_result_t* out) {
    const int64_t off = 0;
    const int32_t nn = n;
    return run_batch<int32_t>(c, KIND_I32, 1, Q, &off, &nn, opts, out, true);
}

int qbx_solve_f64(qbx_ctx_t* c, const double* Q, int n, const qbx_opts_t* opts, qbx_result_t* out) {
    const int64_t off = 0;
    const int32_t nn = n;
    return run_batch<double>(c, KIND_F64, 1, Q, &off, &nn, opts, out, true);
}

int qbx_batch_i32(qbx_ctx_t* c, int count, const int32_t* Qpacked, const int64_t* offsets,
                  const int32_t* ns, const qbx_opts_t* opts, qbx_result_t* out) {
    return run_batch<int32_t>(c, KIND_I32, count, Qpacked, offsets, ns, opts, out, false);
}

int qbx_batch_f64(qbx_ctx_t* c, int count, const double* Qpacked, const int64_t* offsets,
                  const int32_t* ns, const qbx_opts_t* opts, qbx_result_t* out) {
    return run_batch<double>(c, KIND_F64, count, Qpacked, offsets, ns, opts, out, false);
}

}  // extern "C"
