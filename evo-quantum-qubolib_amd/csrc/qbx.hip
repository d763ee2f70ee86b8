// qbx.hip -- gfx950 kernels and the C ABI of libqbx.so (declared in include/qbx.h).
//
// Kernels
//   qbx_prep         one workgroup per problem: bit-ordered W and d tables, the inner-block
//                    trie constants c_y, the energy bound (int32 admission), for float
//                    problems the quantized int32 copies of the tables the filter walks,
//                    and reset of the per-problem reduction state.
//                    Replaces the reference's cache building (qbf.py:246-289,
//                    qbf1.py:105-118), which there costs GEMMs over 2^B x B state tables.
//   qbx_walk_i32<M>  persistent walk over a queue of 64-lane tiles; exact int32 energies
//                    (v_add3_u32 / v_min3_i32 inner loop).  Replaces the hot loop
//                    qbf.py:188-230 and the QB_F1 kernel qbf1.py:179-249.
//   qbx_walk_f64<M>  float problems: the same int32 walk over quantized tables as a filter
//                    with a proven error bound, FP64 exact re-walks of the blocks it passes.
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
// the calling thread's last message (a fixed buffer: setting it never allocates)
static thread_local char g_err[512];

int set_err(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}
// Device arena of one batch: [instances | ranges | Q data] (staged from the host in one
// copy), then the tables qbx_prep writes, the reduction state, the results and the tickets.
struct Group { int variant; int first; int cnt; unsigned long long total; };
struct Layout {
    int kind = 0, count = 0;
    size_t esize = 0;
    std::vector<Plan> plans;
    std::vector<int> order;      // launch order: variant, then longest chunks first
    std::vector<Group> groups;   // one persistent launch per variant
    size_t off_inst = 0, off_ranges = 0, stage_bytes = 0, off_red = 0, off_res = 0, off_ticket = 0;
    size_t total_bytes = 0;
    // one part of a single problem's chunks (rank-space shard), or all of them
    uint32_t part_begin = 0, part_count = 0;   // part_count 0: every chunk
    int32_t flags = 0;                         // qbx_opts_t.flags of the call
    // matrices generated on the device: a small record per problem is staged instead of Q
    // (RandomQUBO: the PCG64 state; TSP: the city distances)
    bool gen = false;
    bool gen_tsp = false;
    int gen_dtype = 0;
    int64_t gen_low = 0, gen_high = 0;        // RandomQUBO bounds
    size_t off_rng = 0;
    // float problems large enough for qbx_seed (one launch after qbx_prep seeds them)
    bool seed = false;
    // per-reference-prefix records of a single solve (qbx_solve_minima_*)
    bool minima = false;
    uint64_t nprefix = 0;                      // 2^(n - B)
    size_t off_minima = 0, off_ticks = 0;      // device sections, nprefix x 8 B each
    size_t h_minima = 0;                       // host copy: minima then ticks, after the results
};
}  // namespace qbx

// A batch shape (kinds, sizes, tiling) with its own device arena, pinned buffers and the
// captured graph of its launches; qbx_batch_plan_run_* refills Q and replays the graph.
struct qbx_batch_plan {
    qbx_ctx* ctx = nullptr;
    qbx::Layout lay;
    char* dv = nullptr;
    char* h_stage = nullptr;
    char* h_res = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

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

int make_layout(qbx_ctx* c, int kind, int count, const int32_t* ns, const qbx_opts_t* opts, bool single,
                Layout& lay, size_t gen_rec = 0, bool minima = false) {
    if (count <= 0) return set_err(QBX_EINVAL, "count must be positive, got %d", count);
    if (!ns) return set_err(QBX_EINVAL, "null argument");
    qbx_opts_t o = {-1, -1, -1, 0};
    if (opts) o = *opts;
    if (o.flags & ~(QBX_FLAG_FLOAT_INT32_FILTER | QBX_FLAG_FLOAT_PACKED_FILTER))
        return set_err(QBX_EINVAL, "unknown opts.flags bits 0x%x", o.flags);
    if ((o.flags & QBX_FLAG_FLOAT_INT32_FILTER) && (o.flags & QBX_FLAG_FLOAT_PACKED_FILTER))
        return set_err(QBX_EINVAL, "opts.flags: at most one float filter may be forced");
    lay.flags = o.flags;
    lay.kind = kind;
    lay.count = count;
    lay.esize = kind == KIND_I32 ? sizeof(int32_t) : sizeof(double);
    lay.plans.assign(count, Plan());
    std::vector<Plan>& plans = lay.plans;
    size_t off = 0;
    lay.off_inst = off;
    off = align256(off + sizeof(InstDev) * count);
    lay.off_ranges = off;
    off = align256(off + sizeof(TileRange) * count);
    lay.gen = gen_rec > 0;
    lay.off_rng = off;
    if (lay.gen) off = align256(off + gen_rec * count);
    const size_t q_begin = off;
    const PlanMode mode = single ? PLAN_SINGLE : PLAN_BATCH;
    for (int i = 0; i < count; ++i) {
        Plan& p = plans[i];
        const int rc = plan_problem(kind, ns[i], o, mode, i, p, &c->walk, minima);
        if (rc) return rc;
        p.off_Q = off;
        off = align256(off + lay.esize * (size_t)p.n * p.n);
        if (kind == KIND_F64 && p.n >= SEED_MIN_N) lay.seed = true;
    }
    // staged from the host: [instances | ranges | Q data], or [instances | ranges | generator
    // records] when the matrices are generated on the device
    lay.stage_bytes = lay.gen ? q_begin : off;
    for (int i = 0; i < count; ++i) {
        Plan& p = plans[i];
        p.off_W = off;
        off = align256(off + table_W_bytes(kind, p.n));
        p.off_d = off;
        off = align256(off + table_d_bytes(kind));
        p.off_delta = off;
        off = align256(off + table_delta_bytes(kind, p.M));
    }
    lay.off_red = off;
    off = align256(off + sizeof(RedDev) * count);
    for (int i = 0; i < count; ++i) {
        plans[i].off_cand = off;
        off = align256(off + 4 * (size_t)plans[i].chunks);
    }
    for (int i = 0; i < count; ++i) {
        plans[i].off_slots = off;
        off = align256(off + sizeof(SlotDev) * (size_t)plans[i].chunks);
    }
    lay.off_res = off;
    off = align256(off + sizeof(ResultDev) * count);
    lay.off_ticket = off;
    off = align256(off + sizeof(unsigned long long) * MAX_GROUPS);
    if (minima) {
        const int A = plans[0].n - plans[0].B;
        if (count != 1) return set_err(QBX_EINVAL, "per-prefix records are kept for single solves");
        if (A > QBX_MAX_PREFIX_BITS)
            return set_err(QBX_EINVAL, "per-prefix records of 2^%d prefixes: at most 2^%d (raise tie_B)", A,
                           QBX_MAX_PREFIX_BITS);
        lay.minima = true;
        lay.nprefix = 1ull << A;
        lay.off_minima = off;
        off = align256(off + 8 * (size_t)lay.nprefix);
        lay.off_ticks = off;
        off = align256(off + 8 * (size_t)lay.nprefix);
        lay.h_minima = align256(sizeof(ResultDev) * (size_t)count);
    }
    lay.total_bytes = off;

    lay.order.resize(count);
    for (int i = 0; i < count; ++i) lay.order[i] = i;
    std::stable_sort(lay.order.begin(), lay.order.end(), [&](int a, int b) {
        if (plans[a].variant != plans[b].variant) return plans[a].variant < plans[b].variant;
        return chunk_work_log2(plans[a]) > chunk_work_log2(plans[b]);
    });
    lay.groups.clear();
    for (int r = 0; r < count; ++r) {
        const int i = lay.order[r];
        if (lay.groups.empty() || lay.groups.back().variant != plans[i].variant)
            lay.groups.push_back({plans[i].variant, r, 0, 0});
        Group& G = lay.groups.back();
        G.cnt++;
        G.total += lay.part_count ? lay.part_count : plans[i].chunks;
    }
    if ((int)lay.groups.size() > MAX_GROUPS) return set_err(QBX_EINVAL, "too many kernel variants in one batch");
    return QBX_OK;
}

// Descriptor of one problem whose sections sit in the device arena at dv (red and result:
// its reduction state and result record there).
void fill_inst(const Plan& p, int kind, int32_t flags, uint32_t expect, char* dv, RedDev* red, ResultDev* result,
               InstDev& I) {
    memset(&I, 0, sizeof I);
    I.n = p.n; I.B = p.B; I.L = p.L; I.M = p.M; I.kind = kind; I.status = 0; I.tiles = p.tiles;
    I.tpc = p.tpc; I.chunks = p.chunks; I.flags = plan_flags(p, flags);
    I.expect = expect;
    I.Q = dv + p.off_Q;
    I.W = dv + p.off_W;
    I.d = dv + p.off_d;
    I.delta = dv + p.off_delta;
    I.red = red;
    I.cand = (uint32_t*)(dv + p.off_cand);
    I.slots = (SlotDev*)(dv + p.off_slots);
    I.result = result;
}

// Instance descriptors and tile ranges (device pointers into the arena at dv) into the
// host staging buffer h.
void stage_descriptors(const Layout& lay, char* h, char* dv) {
    InstDev* hinst = (InstDev*)(h + lay.off_inst);
    for (int i = 0; i < lay.count; ++i) {
        const Plan& p = lay.plans[i];
        InstDev& I = hinst[i];
        fill_inst(p, lay.kind, lay.flags, lay.part_count ? lay.part_count : p.chunks, dv,
                  (RedDev*)(dv + lay.off_red) + i, (ResultDev*)(dv + lay.off_res) + i, I);
        if (lay.minima) {
            I.minima = (uint64_t*)(dv + lay.off_minima);
            I.ticks = (uint64_t*)(dv + lay.off_ticks);
            I.mshift = p.B - p.L;
        }
    }
    TileRange* hr = (TileRange*)(h + lay.off_ranges);
    // tickets count from 0 within each variant's launch
    unsigned long long prefix = 0;
    size_t next_group = 0;
    for (int r = 0; r < lay.count; ++r) {
        const int i = lay.order[r];
        if (next_group < lay.groups.size() && lay.groups[next_group].first == r) {
            prefix = 0;
            ++next_group;
        }
        TileRange& R = hr[r];
        const uint32_t cnt = lay.part_count ? lay.part_count : lay.plans[i].chunks;
        R.inst = (uint32_t)i; R.c_begin = lay.part_begin; R.c_count = cnt; R.pad = 0; R.prefix = prefix;
        prefix += cnt;
    }
}

template <class T>
int stage_matrices(const Layout& lay, char* h, const T* Qpacked, const int64_t* offsets) {
    if (!Qpacked || !offsets) return set_err(QBX_EINVAL, "null argument");
    for (int i = 0; i < lay.count; ++i) {
        if (offsets[i] < 0) return set_err(QBX_EINVAL, "problem %d: negative offset", i);
        const Plan& p = lay.plans[i];
        memcpy(h + p.off_Q, Qpacked + offsets[i], lay.esize * (size_t)p.n * p.n);
    }
    return QBX_OK;
}

// The generator kernel of a device-drawn batch (matrices into the problems' Q sections).
hipError_t launch_gen(const Layout& lay, char* dv, hipStream_t stream) {
    const dim3 grid((unsigned)((lay.count + 63) / 64));
    if (lay.gen_tsp)
        hipLaunchKernelGGL(qbx_gen_tsp, grid, dim3(64), 0, stream, (const InstDev*)(dv + lay.off_inst), lay.count,
                           (const double*)(dv + lay.off_rng), lay.gen_dtype);
    else
        hipLaunchKernelGGL(qbx_gen_random, grid, dim3(64), 0, stream, (const InstDev*)(dv + lay.off_inst), lay.count,
                           (const uint64_t*)(dv + lay.off_rng), lay.gen_low, lay.gen_high, lay.gen_dtype);
    return hipGetLastError();
}

// The work of one batch on `stream`: staged copy in, table build, one persistent launch per
// variant, results out.  Issues no memset and no allocation, so it can be captured in a
// graph; ev_begin / ev_end (nullable, not used under capture) bracket the kernels.
int enqueue(qbx_ctx* c, hipStream_t stream, const Layout& lay, char* dv, const char* h_stage, char* h_res,
            hipEvent_t ev_begin, hipEvent_t ev_end) {
    HIP_TRY(hipMemcpyAsync(dv, h_stage, lay.stage_bytes, hipMemcpyHostToDevice, stream));
    if (ev_begin) HIP_TRY(hipEventRecord(ev_begin, stream));
    if (lay.gen) {
        HIP_TRY(launch_gen(lay, dv, stream));
    }
    if (lay.minima) {   // fill kernels, not memsets: the enqueued work stays graph-safe
        const unsigned blocks = (unsigned)std::min<uint64_t>(4096, (lay.nprefix + 255) / 256);
        hipLaunchKernelGGL(qbx_fill_u64, dim3(blocks), dim3(256), 0, stream, (uint64_t*)(dv + lay.off_minima),
                           (uint64_t)lay.nprefix, ~0ull);
        hipLaunchKernelGGL(qbx_fill_u64, dim3(blocks), dim3(256), 0, stream, (uint64_t*)(dv + lay.off_ticks),
                           (uint64_t)lay.nprefix, 0ull);
        HIP_TRY(hipGetLastError());
    }
    unsigned long long* dticket = (unsigned long long*)(dv + lay.off_ticket);
    hipLaunchKernelGGL(qbx_prep, dim3(lay.count), dim3(256), 0, stream, (InstDev*)(dv + lay.off_inst),
                       (const uint32_t*)nullptr, dticket, (const InstDev*)nullptr);
    HIP_TRY(hipGetLastError());
    if (lay.seed) {
        hipLaunchKernelGGL(qbx_seed, dim3(lay.count), dim3(64 * SEED_WAVES), 0, stream, (InstDev*)(dv + lay.off_inst),
                           (const uint32_t*)nullptr);
        HIP_TRY(hipGetLastError());
    }
    for (size_t g = 0; g < lay.groups.size(); ++g) {
        const Group& G = lay.groups[g];
        const int v = G.variant;
        const unsigned long long grid_full = (unsigned long long)c->blocks_per_cu[v] * c->prop.multiProcessorCount;
        const unsigned long long grid = std::max(1ull, std::min(grid_full, G.total));
        walk_fn_t fn = (walk_fn_t)kVariants[v].fn;
        hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(64), 0, stream,
                           (const InstDev*)(dv + lay.off_inst), (const TileRange*)(dv + lay.off_ranges) + G.first,
                           G.cnt, G.total, dticket + g, 0);
        HIP_TRY(hipGetLastError());
    }
    if (ev_end) HIP_TRY(hipEventRecord(ev_end, stream));
    HIP_TRY(hipMemcpyAsync(h_res, dv + lay.off_res, sizeof(ResultDev) * lay.count, hipMemcpyDeviceToHost, stream));
    if (lay.minima) {   // minima and ticks are adjacent sections: one copy
        HIP_TRY(hipMemcpyAsync(h_res + lay.h_minima, dv + lay.off_minima, lay.off_ticks - lay.off_minima +
                               8 * (size_t)lay.nprefix, hipMemcpyDeviceToHost, stream));
    }
    return QBX_OK;
}

// The caller's result record of one problem from its device result.
void fill_result(const Plan& p, int kind, const ResultDev& h, float ms, qbx_result_t& R) {
    memset(&R, 0, sizeof R);
    R.n = p.n; R.tie_B = p.B; R.walk_bits = p.L; R.inner_bits = p.M; R.lane_bits = p.H;
    R.states = (p.n >= 64) ? ~0ull : (1ull << p.n);
    R.tiles = p.tiles;
    R.kernel_ms = ms;
    R.status = h.status;
    R.path = h.path;
    R.rank = h.rank;
    R.state = h.state;
    if (kind == KIND_I32) {
        R.E_int = (int64_t)h.Ebits;
        R.E = (double)R.E_int;
    } else {
        double e;
        memcpy(&e, &h.Ebits, sizeof e);
        R.E = e;
        R.E_int = (int64_t)e;
    }
}

int collect(const Layout& lay, const ResultDev* hres, float ms, qbx_result_t* out) {
    int first_bad = QBX_OK;
    for (int i = 0; i < lay.count; ++i) {
        qbx_result_t& R = out[i];
        fill_result(lay.plans[i], lay.kind, hres[i], ms, R);
        if (R.status != 0 && first_bad == QBX_OK) {
            first_bad = R.status;
            set_err(R.status, "problem %d: energy bound outside the %s path", i,
                    lay.kind == KIND_I32 ? "int32" : "float");
        }
    }
    return first_bad;
}

// Per-prefix outputs of qbx_solve_minima_*: minima as int64 (integer kind) or double.
struct MinimaOut {
    void* minima;
    double* wave_ms;    // nullable
};

// Keys and ticks of the host copy into the caller's arrays.
void write_minima(const qbx_ctx* c, const Layout& lay, const char* h, const MinimaOut& mo) {
    const uint64_t* keys = (const uint64_t*)(h + lay.h_minima);
    const uint64_t* ticks = (const uint64_t*)(h + lay.h_minima + (lay.off_ticks - lay.off_minima));
    for (uint64_t i = 0; i < lay.nprefix; ++i) {
        if (lay.kind == KIND_I32) {
            ((int64_t*)mo.minima)[i] = (int64_t)(keys[i] ^ 0x8000000000000000ull);
        } else {
            const uint64_t k = keys[i];
            const uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
            double e;
            memcpy(&e, &u, sizeof e);
            ((double*)mo.minima)[i] = e;
        }
        if (mo.wave_ms) mo.wave_ms[i] = c->wall_khz > 0 ? (double)ticks[i] / 64.0 / (double)c->wall_khz : 0.0;
    }
}

template <class T>
int run_batch(qbx_ctx* c, int kind, int count, const T* Qpacked, const int64_t* offsets,
              const int32_t* ns, const qbx_opts_t* opts, qbx_result_t* out, bool single,
              int part = 0, int parts = 1, const MinimaOut* mo = nullptr) {
    if (!c) return set_err(QBX_EINVAL, "null context");
    if (!Qpacked || !offsets || !ns || !out) return set_err(QBX_EINVAL, "null argument");
    if (mo && !mo->minima) return set_err(QBX_EINVAL, "null argument");
    Layout lay;
    int rc = make_layout(c, kind, count, ns, opts, single, lay, 0, mo != nullptr);
    if (rc) return rc;
    if (parts != 1) {   // one contiguous range of the (single) problem's chunks
        const uint32_t chunks = lay.plans[0].chunks;
        if (count != 1 || parts < 1 || part < 0 || part >= parts || (uint32_t)parts > chunks)
            return set_err(QBX_EINVAL, "part %d of %d: need 0 <= part < parts <= %u chunks", part, parts, chunks);
        const uint32_t b = (uint32_t)((uint64_t)chunks * (uint64_t)part / (uint64_t)parts);
        const uint32_t e = (uint32_t)((uint64_t)chunks * (uint64_t)(part + 1) / (uint64_t)parts);
        lay.part_begin = b;
        lay.part_count = e - b;
        lay.groups[0].total = e - b;
    }
    size_t hneed = std::max(lay.stage_bytes, sizeof(ResultDev) * (size_t)count);
    if (lay.minima) hneed = std::max(hneed, lay.h_minima + 16 * (size_t)lay.nprefix);
    rc = ensure_capacity(c, lay.total_bytes, hneed);
    if (rc) return rc;
    char* h = c->hbuf;
    char* dv = c->dbuf;
    stage_descriptors(lay, h, dv);
    rc = stage_matrices(lay, h, Qpacked, offsets);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    rc = enqueue(c, c->stream, lay, dv, h, h, c->ev0, c->ev1);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    rc = collect(lay, (const ResultDev*)h, ms, out);
    if (mo && rc == QBX_OK) write_minima(c, lay, h, *mo);
    if (lay.part_count) {   // states of this part only
        const Plan& p = lay.plans[0];
        const uint64_t t0 = (uint64_t)lay.part_begin * p.tpc;
        const uint64_t t1 = std::min<uint64_t>((uint64_t)(lay.part_begin + lay.part_count) * p.tpc, p.tiles);
        out[0].tiles = t1 - t0;
        out[0].states = (uint64_t)ldexp((double)(t1 - t0) / (double)p.tiles, p.n);
    }
    return rc;
}

// Device-generated batches (RandomQUBO, or TSP when tsp): the layout and the staged
// records, PCG64 states for RandomQUBO, for TSP the city distances drawn and measured here
// with the C library's pow (tsp_distances).  For TSP, ns holds the city counts and flow /
// fhigh the coordinate bounds.
int gen_prepare(qbx_ctx* c, int count, const qbx_pcg64_t* rng, const int32_t* ns, int64_t low, int64_t high,
                int dtype, const qbx_opts_t* opts, Layout& lay, bool tsp = false, double flow = 0.0,
                double fhigh = 0.0) {
    if (!c) return set_err(QBX_EINVAL, "null context");
    if (!rng || !ns) return set_err(QBX_EINVAL, "null argument");
    const int32_t* n_cities = ns;
    std::vector<int32_t> sizes;
    if (tsp) {
        if (dtype != GEN_FLOAT32 && dtype != GEN_FLOAT64)
            return set_err(QBX_EINVAL, "TSP cities must be float32 or float64 (dtype %d)", dtype);
        if (!(flow <= fhigh) || !(fhigh - flow < 1.0e300))
            return set_err(QBX_EINVAL, "unsupported TSP coordinate bounds [%g, %g)", flow, fhigh);
        sizes.resize(count > 0 ? (size_t)count : 0u);
        for (int i = 0; i < count; ++i) {
            if (ns[i] < 1 || ns[i] > TSP_MAX_CITIES)
                return set_err(QBX_EINVAL, "problem %d: %d cities outside [1, %d]", i, ns[i], TSP_MAX_CITIES);
            sizes[i] = ns[i] * ns[i];
        }
        ns = sizes.data();
    } else if (!gen_valid(low, high, dtype)) {
        return set_err(QBX_EINVAL, "unsupported generator bounds [%lld, %lld) for dtype %d", (long long)low,
                       (long long)high, dtype);
    }
    const int kind = (dtype == GEN_INT16 || dtype == GEN_INT32) ? KIND_I32 : KIND_F64;
    const size_t rec = tsp ? TSP_MAX_PAIRS * sizeof(double) : 4 * sizeof(uint64_t);
    int rc = make_layout(c, kind, count, ns, opts, false, lay, rec);
    if (rc) return rc;
    lay.gen_tsp = tsp;
    lay.gen_dtype = dtype;
    lay.gen_low = low;
    lay.gen_high = high;
    rc = ensure_capacity(c, lay.total_bytes, std::max(lay.total_bytes, sizeof(ResultDev) * (size_t)count));
    if (rc) return rc;
    stage_descriptors(lay, c->hbuf, c->dbuf);
    if (tsp) {
        double* hd = (double*)(c->hbuf + lay.off_rng);
        for (int i = 0; i < count; ++i) {
            Pcg64 g;
            g.hi = rng[i].state_hi;
            g.lo = rng[i].state_lo;
            g.inc_hi = rng[i].inc_hi;
            g.inc_lo = rng[i].inc_lo;
            g.buf32 = 0;
            g.buf16 = 0;
            g.has32 = g.has16 = 0;
            double* d = hd + (size_t)i * TSP_MAX_PAIRS;
            if (dtype == GEN_FLOAT32) tsp_distances<float>(g, n_cities[i], flow, fhigh, d);
            else tsp_distances<double>(g, n_cities[i], flow, fhigh, d);
        }
        return QBX_OK;
    }
    uint64_t* hr = (uint64_t*)(c->hbuf + lay.off_rng);
    for (int i = 0; i < count; ++i) {
        hr[4 * i + 0] = rng[i].state_hi;
        hr[4 * i + 1] = rng[i].state_lo;
        hr[4 * i + 2] = rng[i].inc_hi;
        hr[4 * i + 3] = rng[i].inc_lo;
    }
    return QBX_OK;
}

void plan_free(qbx_batch_plan* P) {
    if (!P) return;
    if (P->ctx) hipSetDevice(P->ctx->device);
    if (P->exec) hipGraphExecDestroy(P->exec);
    if (P->graph) hipGraphDestroy(P->graph);
    if (P->dv) hipFree(P->dv);
    if (P->h_stage) hipHostFree(P->h_stage);
    if (P->h_res) hipHostFree(P->h_res);
    if (P->ev0) hipEventDestroy(P->ev0);
    if (P->ev1) hipEventDestroy(P->ev1);
    delete P;
}

int plan_build(qbx_batch_plan* P) {
    qbx_ctx* c = P->ctx;
    const Layout& lay = P->lay;
    HIP_TRY(hipSetDevice(c->device));
    if (hipMalloc((void**)&P->dv, lay.total_bytes) != hipSuccess)
        return set_err(QBX_ENOMEM, "hipMalloc(%zu) failed", lay.total_bytes);
    if (hipHostMalloc((void**)&P->h_stage, lay.stage_bytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&P->h_res, sizeof(ResultDev) * lay.count, hipHostMallocDefault) != hipSuccess)
        return set_err(QBX_ENOMEM, "hipHostMalloc failed");
    memset(P->h_stage, 0, lay.stage_bytes);
    stage_descriptors(lay, P->h_stage, P->dv);
    HIP_TRY(hipEventCreate(&P->ev0));
    HIP_TRY(hipEventCreate(&P->ev1));
    // the host thread owns the context stream, so nothing else is queued on it meanwhile
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(c, c->stream, lay, P->dv, P->h_stage, P->h_res, nullptr, nullptr);
    const hipError_t e = hipStreamEndCapture(c->stream, &P->graph);
    if (rc) return rc;
    if (e != hipSuccess) return set_err(QBX_EDEVICE, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    HIP_TRY(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    return QBX_OK;
}

template <class T>
int plan_run(qbx_batch_plan* P, int kind, const T* Qpacked, const int64_t* offsets, qbx_result_t* out) {
    if (!P || !out) return set_err(QBX_EINVAL, "null argument");
    if (P->lay.kind != kind)
        return set_err(QBX_EINVAL, "plan was made for %s matrices", P->lay.kind == KIND_I32 ? "int32" : "double");
    int rc = stage_matrices(P->lay, P->h_stage, Qpacked, offsets);
    if (rc) return rc;
    qbx_ctx* c = P->ctx;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipEventRecord(P->ev0, c->stream));
    HIP_TRY(hipGraphLaunch(P->exec, c->stream));
    HIP_TRY(hipEventRecord(P->ev1, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, P->ev0, P->ev1));
    return collect(P->lay, (const ResultDev*)P->h_res, ms, out);
}

// Sections of one single int32 solve in one buffer, in make_layout's order for count = 1:
// [InstDev | TileRange | Q] (staged), tables, reduction state, candidate ids, result
// slots, the result, the tickets.
struct SingleLayout {
    size_t off_inst, off_ranges, stage_bytes, off_red, off_res, off_ticket, total;
};

void single_layout(Plan& p, int kind, SingleLayout& S) {
    size_t off = 0;
    S.off_inst = off;
    off = align256(off + sizeof(InstDev));
    S.off_ranges = off;
    off = align256(off + sizeof(TileRange));
    p.off_Q = off;
    off = align256(off + (kind == KIND_I32 ? 4 : 8) * (size_t)p.n * p.n);
    S.stage_bytes = off;
    p.off_W = off;
    off = align256(off + table_W_bytes(kind, p.n));
    p.off_d = off;
    off = align256(off + table_d_bytes(kind));
    p.off_delta = off;
    off = align256(off + table_delta_bytes(kind, p.M));
    S.off_red = off;
    off = align256(off + sizeof(RedDev));
    p.off_cand = off;
    off = align256(off + 4 * (size_t)p.chunks);
    p.off_slots = off;
    off = align256(off + sizeof(SlotDev) * (size_t)p.chunks);
    S.off_res = off;
    off = align256(off + sizeof(ResultDev));
    S.off_ticket = off;
    off = align256(off + sizeof(unsigned long long) * MAX_GROUPS);
    S.total = off;
}

// Host staging of qbx_solve_i32_noalloc, on the caller's stack: descriptor, tile range and
// the largest Q (56 x 56 int32).
constexpr size_t NOALLOC_STAGE = 256 + 256 + 12800;
static_assert(sizeof(InstDev) <= 256 && sizeof(TileRange) <= 256, "staging sections");

}  // namespace

// Device scratch of qbx_solve_i32_noalloc: a file-scope array each GPU holds (from module
// load to the end of the process), shared by every call on that GPU.  Host code never uses
// its name as an address: qbx_open takes its per-device address with hipGetSymbolAddress.
static __device__ __align__(256) char scratch[QBX_NOALLOC_SCRATCH_BYTES];

extern "C" {

const char* qbx_last_error(void) { return g_err; }

int qbx_open(int device, qbx_ctx_t** out) {
    if (!out) return set_err(QBX_EINVAL, "null output pointer");
    *out = nullptr;
    int cnt = 0;
    HIP_TRY(hipGetDeviceCount(&cnt));
    if (device < 0 || device >= cnt) return set_err(QBX_EINVAL, "device %d not present (%d devices)", device, cnt);
    qbx_ctx* c = new qbx_ctx();
    c->device = device;
    c->walk.clear();
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipGetDeviceProperties(&c->prop, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device);
    // this device's copy of the no-alloc scratch array (device `device` is current here)
    if (e == hipSuccess) e = hipGetSymbolAddress((void**)&c->scratch, HIP_SYMBOL(scratch));
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

int qbx_solve_i32_noalloc(qbx_ctx_t* c, const int32_t* Q, int n, const qbx_opts_t* opts, qbx_result_t* out) {
    if (!c || !Q || !out) return set_err(QBX_EINVAL, "null argument");
    if (!c->scratch) return set_err(QBX_EDEVICE, "no scratch address for device %d", c->device);
    qbx_opts_t o = {-1, -1, -1, 0};
    if (opts) o = *opts;
    if (o.flags & ~(QBX_FLAG_FLOAT_INT32_FILTER | QBX_FLAG_FLOAT_PACKED_FILTER))
        return set_err(QBX_EINVAL, "unknown opts.flags bits 0x%x", o.flags);
    Plan p;
    int rc = plan_problem(KIND_I32, n, o, PLAN_SINGLE, 0, p, &c->walk);
    if (rc) return rc;
    SingleLayout S;
    single_layout(p, KIND_I32, S);
    if (S.total > (size_t)QBX_NOALLOC_SCRATCH_BYTES)
        return set_err(QBX_EINVAL,
                       "n=%d with walk_bits=%d needs %zu bytes of device scratch; qbx_solve_i32_noalloc has %d "
                       "(use qbx_solve_i32, or more walk bits)", n, p.L, S.total, QBX_NOALLOC_SCRATCH_BYTES);
    alignas(256) char h[NOALLOC_STAGE];
    if (S.stage_bytes > sizeof h) return set_err(QBX_EINVAL, "staging of n=%d exceeds %zu bytes", n, sizeof h);
    char* dv = c->scratch;    // this device's scratch array
    fill_inst(p, KIND_I32, o.flags, p.chunks, dv, (RedDev*)(dv + S.off_red), (ResultDev*)(dv + S.off_res),
              *(InstDev*)(h + S.off_inst));
    TileRange& R = *(TileRange*)(h + S.off_ranges);
    R.inst = 0; R.c_begin = 0; R.c_count = p.chunks; R.pad = 0; R.prefix = 0;
    memcpy(h + p.off_Q, Q, sizeof(int32_t) * (size_t)n * n);
    HIP_TRY(hipSetDevice(c->device));
    // everything below is enqueued on the context stream, and the stream is drained before
    // returning on every path (the staging buffer lives on this stack frame)
    hipError_t e = hipMemcpyAsync(dv, h, S.stage_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    if (e == hipSuccess) {
        // qbx_prep resets the reduction state and the tickets: nothing starts from the last
        // call's contents
        hipLaunchKernelGGL(qbx_prep, dim3(1), dim3(256), 0, c->stream, (InstDev*)(dv + S.off_inst),
                           (const uint32_t*)nullptr, (unsigned long long*)(dv + S.off_ticket), (const InstDev*)nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        const int v = p.variant;
        const unsigned long long grid_full = (unsigned long long)c->blocks_per_cu[v] * c->prop.multiProcessorCount;
        const unsigned long long grid = std::max(1ull, std::min(grid_full, (unsigned long long)p.chunks));
        hipLaunchKernelGGL((walk_fn_t)kVariants[v].fn, dim3((unsigned)grid), dim3(64), 0, c->stream,
                           (const InstDev*)(dv + S.off_inst), (const TileRange*)(dv + S.off_ranges), 1,
                           (unsigned long long)p.chunks, (unsigned long long*)(dv + S.off_ticket), 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    ResultDev hres;
    if (e == hipSuccess) e = hipMemcpyAsync(&hres, dv + S.off_res, sizeof hres, hipMemcpyDeviceToHost, c->stream);
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_err(QBX_EDEVICE, "qbx_solve_i32_noalloc: %s", hipGetErrorString(e));
    if (es != hipSuccess) return set_err(QBX_EDEVICE, "hipStreamSynchronize failed: %s", hipGetErrorString(es));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    fill_result(p, KIND_I32, hres, ms, *out);
    if (out->status != 0) return set_err(out->status, "energy bound outside the int32 path");
    return QBX_OK;
}

int qbx_solve_minima_i32(qbx_ctx_t* c, const int32_t* Q, int n, const qbx_opts_t* opts, int64_t* minima,
                         double* wave_ms, qbx_result_t* out) {
    const int64_t off = 0;
    const int32_t nn = n;
    const MinimaOut mo{minima, wave_ms};
    return run_batch<int32_t>(c, KIND_I32, 1, Q, &off, &nn, opts, out, true, 0, 1, &mo);
}

int qbx_solve_minima_f64(qbx_ctx_t* c, const double* Q, int n, const qbx_opts_t* opts, double* minima,
                         double* wave_ms, qbx_result_t* out) {
    const int64_t off = 0;
    const int32_t nn = n;
    const MinimaOut mo{minima, wave_ms};
    return run_batch<double>(c, KIND_F64, 1, Q, &off, &nn, opts, out, true, 0, 1, &mo);
}

int qbx_set_walk_table(qbx_ctx_t* c, int kind, int mode, const int32_t* walk_bits, int count) {
    if (!c) return set_err(QBX_EINVAL, "null context");
    if (kind != QBX_KIND_I32 && kind != QBX_KIND_F64) return set_err(QBX_EINVAL, "kind must be 0 or 1");
    if (mode != QBX_MODE_SINGLE && mode != QBX_MODE_BATCH) return set_err(QBX_EINVAL, "mode must be 0 or 1");
    if (count < 0 || count > MAX_N + 1 || (count > 0 && !walk_bits))
        return set_err(QBX_EINVAL, "count must be in [0, %d] with a non-null table", MAX_N + 1);
    for (int n = 0; n < count; ++n)
        if (walk_bits[n] > MAX_N) return set_err(QBX_EINVAL, "walk_bits[%d]=%d exceeds %d", n, walk_bits[n], MAX_N);
    for (int n = 0; n <= MAX_N; ++n)
        c->walk.L[kind][mode][n] = (int8_t)((n < count && walk_bits[n] >= 0) ? walk_bits[n] : -1);
    return QBX_OK;
}

int qbx_plan_info(qbx_ctx_t* c, int kind, int mode, int n, const qbx_opts_t* opts, qbx_result_t* out) {
    if (!c || !out) return set_err(QBX_EINVAL, "null argument");
    if (kind != QBX_KIND_I32 && kind != QBX_KIND_F64) return set_err(QBX_EINVAL, "kind must be 0 or 1");
    if (mode != QBX_MODE_SINGLE && mode != QBX_MODE_BATCH) return set_err(QBX_EINVAL, "mode must be 0 or 1");
    qbx_opts_t o = {-1, -1, -1, 0};
    if (opts) o = *opts;
    Plan p;
    const int rc = plan_problem(kind, n, o, mode == QBX_MODE_SINGLE ? PLAN_SINGLE : PLAN_BATCH, 0, p, &c->walk);
    if (rc) return rc;
    memset(out, 0, sizeof *out);
    out->n = p.n; out->tie_B = p.B; out->walk_bits = p.L; out->inner_bits = p.M; out->lane_bits = p.H;
    out->states = (p.n >= 64) ? ~0ull : (1ull << p.n);
    out->tiles = p.tiles;
    return QBX_OK;
}

int qbx_batch_plan_create(qbx_ctx_t* c, int kind, int count, const int32_t* ns, const qbx_opts_t* opts,
                          qbx_batch_plan_t** out) {
    if (!c || !out) return set_err(QBX_EINVAL, "null argument");
    *out = nullptr;
    if (kind != QBX_KIND_I32 && kind != QBX_KIND_F64) return set_err(QBX_EINVAL, "kind must be 0 or 1");
    qbx_batch_plan* P = new qbx_batch_plan();
    P->ctx = c;
    int rc = make_layout(c, kind, count, ns, opts, false, P->lay);
    if (rc == QBX_OK) rc = plan_build(P);
    if (rc != QBX_OK) {
        plan_free(P);
        return rc;
    }
    *out = P;
    return QBX_OK;
}

int qbx_batch_plan_run_i32(qbx_batch_plan_t* P, const int32_t* Qpacked, const int64_t* offsets, qbx_result_t* out) {
    return plan_run<int32_t>(P, KIND_I32, Qpacked, offsets, out);
}

int qbx_batch_plan_run_f64(qbx_batch_plan_t* P, const double* Qpacked, const int64_t* offsets, qbx_result_t* out) {
    return plan_run<double>(P, KIND_F64, Qpacked, offsets, out);
}

int qbx_batch_plan_destroy(qbx_batch_plan_t* P) {
    if (P && P->ctx) {
        hipSetDevice(P->ctx->device);
        hipStreamSynchronize(P->ctx->stream);
    }
    plan_free(P);
    return QBX_OK;
}

static int gen_batch(qbx_ctx* c, const Layout& lay, qbx_result_t* out) {
    HIP_TRY(hipSetDevice(c->device));
    const int rc = enqueue(c, c->stream, lay, c->dbuf, c->hbuf, c->hbuf, c->ev0, c->ev1);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    return collect(lay, (const ResultDev*)c->hbuf, ms, out);
}

static int gen_read(qbx_ctx* c, const Layout& lay, void* Qout) {
    const int count = lay.count;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(c->dbuf, c->hbuf, lay.stage_bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_gen(lay, c->dbuf, c->stream));
    // the Q section is contiguous from the first problem's matrix to the last one's
    const size_t q0 = lay.plans[0].off_Q;
    const Plan& pl = lay.plans[count - 1];
    const size_t q1 = pl.off_Q + lay.esize * (size_t)pl.n * pl.n;
    HIP_TRY(hipMemcpyAsync(c->hbuf + q0, c->dbuf + q0, q1 - q0, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    char* dst = (char*)Qout;
    for (int i = 0; i < count; ++i) {
        const size_t bytes = lay.esize * (size_t)lay.plans[i].n * lay.plans[i].n;
        memcpy(dst, c->hbuf + lay.plans[i].off_Q, bytes);
        dst += bytes;
    }
    return QBX_OK;
}

int qbx_pcg64_from_seeds(int count, const uint64_t* seeds, qbx_pcg64_t* out) {
    if (count < 0) return set_err(QBX_EINVAL, "count must be non-negative, got %d", count);
    if (count > 0 && (!seeds || !out)) return set_err(QBX_EINVAL, "null argument");
    for (int i = 0; i < count; ++i) {
        Pcg64 g;
        pcg64_from_seed(seeds[i], g);
        out[i].state_hi = g.hi;
        out[i].state_lo = g.lo;
        out[i].inc_hi = g.inc_hi;
        out[i].inc_lo = g.inc_lo;
    }
    return QBX_OK;
}

int qbx_batch_random_qubo(qbx_ctx_t* c, int count, const qbx_pcg64_t* rng, const int32_t* ns, int64_t low,
                          int64_t high, int dtype, const qbx_opts_t* opts, qbx_result_t* out) {
    if (!out) return set_err(QBX_EINVAL, "null argument");
    Layout lay;
    const int rc = gen_prepare(c, count, rng, ns, low, high, dtype, opts, lay);
    return rc ? rc : gen_batch(c, lay, out);
}

int qbx_random_qubo_read(qbx_ctx_t* c, int count, const qbx_pcg64_t* rng, const int32_t* ns, int64_t low,
                         int64_t high, int dtype, void* Qout) {
    if (!Qout) return set_err(QBX_EINVAL, "null argument");
    Layout lay;
    const int rc = gen_prepare(c, count, rng, ns, low, high, dtype, nullptr, lay);
    return rc ? rc : gen_read(c, lay, Qout);
}

int qbx_batch_random_tsp(qbx_ctx_t* c, int count, const qbx_pcg64_t* rng, const int32_t* n_cities, double low,
                         double high, int dtype, const qbx_opts_t* opts, qbx_result_t* out) {
    if (!out) return set_err(QBX_EINVAL, "null argument");
    Layout lay;
    const int rc = gen_prepare(c, count, rng, n_cities, 0, 0, dtype, opts, lay, true, low, high);
    return rc ? rc : gen_batch(c, lay, out);
}

int qbx_random_tsp_read(qbx_ctx_t* c, int count, const qbx_pcg64_t* rng, const int32_t* n_cities, double low,
                        double high, int dtype, double* Qout) {
    if (!Qout) return set_err(QBX_EINVAL, "null argument");
    Layout lay;
    const int rc = gen_prepare(c, count, rng, n_cities, 0, 0, dtype, nullptr, lay, true, low, high);
    return rc ? rc : gen_read(c, lay, Qout);
}

int qbx_solve_part_i32(qbx_ctx_t* c, const int32_t* Q, int n, const qbx_opts_t* opts, int part, int parts,
                       qbx_result_t* out) {
    const int64_t off = 0;
    const int32_t nn = n;
    return run_batch<int32_t>(c, KIND_I32, 1, Q, &off, &nn, opts, out, true, part, parts);
}

int qbx_solve_part_f64(qbx_ctx_t* c, const double* Q, int n, const qbx_opts_t* opts, int part, int parts,
                       qbx_result_t* out) {
    const int64_t off = 0;
    const int32_t nn = n;
    return run_batch<double>(c, KIND_F64, 1, Q, &off, &nn, opts, out, true, part, parts);
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
