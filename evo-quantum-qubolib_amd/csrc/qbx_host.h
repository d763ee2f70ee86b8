// qbx_host.h -- host-side helpers shared by qbx.hip (single / batch) and qbx_server.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdarg.h>
#include <string>

#include "qbx_device.h"

namespace qbx {

// last error message of the calling thread (defined in qbx.hip)
int set_err(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ::qbx::set_err(QBX_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

typedef void (*walk_fn_t)(const InstDev*, const TileRange*, int, unsigned long long,
                          unsigned long long*, int);

struct Variant {
    int kind, M, OMAX;
    bool rec;          // records per-reference-prefix minima (single solves of qbx_solve_minima_*)
    const void* fn;
};

// available (kind, M, rec) instantiations, largest M first within each kind (each
// translation unit launches its own copies of the static kernels)
static const Variant kVariants[] = {
    {KIND_I32, 7, 13, false, (const void*)qbx_walk_i32<7, 13, false>},
    {KIND_I32, 6, 14, false, (const void*)qbx_walk_i32<6, 14, false>},
    {KIND_I32, 4, 16, false, (const void*)qbx_walk_i32<4, 16, false>},
    {KIND_I32, 1, 19, false, (const void*)qbx_walk_i32<1, 19, false>},
    {KIND_F64, 7, 13, false, (const void*)qbx_walk_f64<7, 13, false>},
    {KIND_F64, 6, 14, false, (const void*)qbx_walk_f64<6, 14, false>},
    {KIND_F64, 4, 16, false, (const void*)qbx_walk_f64<4, 16, false>},
    {KIND_F64, 1, 19, false, (const void*)qbx_walk_f64<1, 19, false>},
    {KIND_I32, 7, 13, true, (const void*)qbx_walk_i32<7, 13, true>},
    {KIND_I32, 6, 14, true, (const void*)qbx_walk_i32<6, 14, true>},
    {KIND_I32, 4, 16, true, (const void*)qbx_walk_i32<4, 16, true>},
    {KIND_I32, 1, 19, true, (const void*)qbx_walk_i32<1, 19, true>},
    {KIND_F64, 7, 13, true, (const void*)qbx_walk_f64<7, 13, true>},
    {KIND_F64, 6, 14, true, (const void*)qbx_walk_f64<6, 14, true>},
    {KIND_F64, 4, 16, true, (const void*)qbx_walk_f64<4, 16, true>},
    {KIND_F64, 1, 19, true, (const void*)qbx_walk_f64<1, 19, true>},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

constexpr int MAX_LANE_BITS = 37;   // tiles = 2^(H-6) < 2^31
constexpr uint32_t MAX_CHUNKS = 1u << 22;

static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// largest available M <= min(want_M, n) for this kind (and record flag), or -1
static inline int pick_variant(int kind, int want_M, int n, bool rec = false) {
    int best = -1;
    for (int v = 0; v < kNumVariants; ++v) {
        if (kVariants[v].kind != kind || kVariants[v].rec != rec) continue;
        if (kVariants[v].M <= want_M && kVariants[v].M <= n) {
            if (best < 0 || kVariants[v].M > kVariants[best].M) best = v;
        }
    }
    return best;
}

enum PlanMode { PLAN_SINGLE = 0, PLAN_BATCH = 1 };

// Default walk bits per lane (measured on MI355X): a lone problem is cut into many tiles --
// L = n - 16, capped at 18 (n = 28: 12, 32: 16, 36 and 40: 18; the float walk measured best
// with the same rule); inside a batch or a serving step many problems share the GPU, so
// fewer, larger tiles per problem keep the per-lane set-up small.
static inline int default_walk_bits(int n, int M, int OMAX, PlanMode mode) {
    const bool single = mode == PLAN_SINGLE;
    int L = single ? n - 16 : n - 10;
    const int cap = single ? 18 : 16;
    if (L > cap) L = cap;
    if (L < M) L = M;
    if (L < n - MAX_LANE_BITS) L = n - MAX_LANE_BITS;
    if (L > M + OMAX) L = M + OMAX;
    if (L > n) L = n;
    return L;
}

// Per-n walk bits chosen by an autotuner (qbx_set_walk_table), indexed
// [kind][mode][n]; entries < 0 fall back to default_walk_bits().
struct WalkTable {
    int8_t L[2][2][MAX_N + 1];
    void clear() {
        for (auto& a : L)
            for (auto& b : a)
                for (auto& v : b) v = -1;
    }
};

struct Plan {
    int n, B, L, M, kind, variant, H, mode;
    uint32_t tiles, tpc, chunks;
    size_t off_Q, off_W, off_d, off_delta, off_cand, off_slots;
};

// Descriptor flags of a problem (qbx_opts_t.flags of its call): with neither float filter
// forced, float problems below these sizes take the int32 filter.  Measured on MI355X,
// N(0,1) problems: alone, n = 26 0.16 ms packed vs 0.10 ms int32, n = 30 0.37 vs 0.56 ms;
// in batches, n = 20 0.35 vs 0.31 us per problem, n = 24 0.71 vs 0.91 us.
constexpr int QRP_MIN_N_SINGLE = 28, QRP_MIN_N_BATCH = 22;
static inline int32_t plan_flags(const Plan& p, int32_t flags) {
    const int32_t forced = QBX_FLAG_FLOAT_INT32_FILTER | QBX_FLAG_FLOAT_PACKED_FILTER;
    if (p.kind == KIND_F64 && !(flags & forced) &&
        p.n < (p.mode == PLAN_SINGLE ? QRP_MIN_N_SINGLE : QRP_MIN_N_BATCH))
        flags |= QBX_FLAG_FLOAT_INT32_FILTER;
    return flags;
}

// Tiling of one problem; returns QBX_OK or an error (message set).  cap_L_at_B: lanes no
// wider than the tie width B (every lane then lies inside one reference prefix), for
// per-prefix records; needs B >= 1.
static inline int plan_problem(int kind, int n, const qbx_opts_t& o, PlanMode mode, int index, Plan& p,
                               const WalkTable* table = nullptr, bool cap_L_at_B = false) {
    if (n < 1 || n > MAX_N) return set_err(QBX_EINVAL, "problem %d: n=%d outside [1, %d]", index, n, MAX_N);
    p.n = n;
    p.kind = kind;
    p.mode = mode;
    p.B = (o.tie_B < 0) ? (n < 20 ? n : 20) : o.tie_B;
    if (p.B > n) return set_err(QBX_EINVAL, "problem %d: tie_B=%d exceeds n=%d", index, p.B, n);
    if (cap_L_at_B && p.B < 1) return set_err(QBX_EINVAL, "per-prefix records need tie_B >= 1");
    int wantM = (o.inner_bits < 0) ? 7 : o.inner_bits;
    if (cap_L_at_B && wantM > p.B) wantM = p.B;
    p.variant = pick_variant(kind, wantM, n, cap_L_at_B);
    if (p.variant < 0) return set_err(QBX_EINVAL, "problem %d: no kernel for inner_bits=%d", index, wantM);
    p.M = kVariants[p.variant].M;
    const int OMAX = kVariants[p.variant].OMAX;
    int L;
    if (o.walk_bits >= 0) {
        L = o.walk_bits < n ? o.walk_bits : n;
    } else {
        L = default_walk_bits(n, p.M, OMAX, mode);
        const int t = table ? table->L[kind][mode][n] : -1;
        // a tuned entry is used only where it is valid for the variant picked here
        if (t >= p.M && t <= p.M + OMAX && t <= n && n - t <= MAX_LANE_BITS) L = t;
    }
    if (L < p.M) L = p.M;
    if (cap_L_at_B && L > p.B) L = p.B;
    if (L > p.M + OMAX || L > n)
        return set_err(QBX_EINVAL, "problem %d: walk_bits=%d outside [%d, %d]", index, L, p.M,
                       (n < p.M + OMAX ? n : p.M + OMAX));
    if (n - L > MAX_LANE_BITS)
        return set_err(QBX_EINVAL, "problem %d: walk_bits=%d leaves %d lane bits (max %d)", index, L, n - L,
                       MAX_LANE_BITS);
    p.L = L;
    p.H = n - L;
    p.tiles = p.H > 6 ? (1u << (p.H - 6)) : 1u;
    // tiles per chunk: amortise the per-ticket work (atomics, table loads, the final re-walk)
    // over about 2^22 states, but keep enough chunks to spread one problem over the GPU
    // (single solves) or over a few waves (batches and serving steps)
    uint32_t tpc = (22 - 6 - L) > 0 ? (1u << (22 - 6 - L)) : 1u;
    const uint32_t min_chunks = mode == PLAN_SINGLE ? 8192u : 16u;
    const uint32_t cap = p.tiles / min_chunks;
    if (tpc > cap) tpc = cap;
    if (tpc < 1u) tpc = 1u;
    // at most MAX_CHUNKS chunks (per-chunk result slots and atomics stay bounded for n > 44)
    const uint32_t floor_tpc = (uint32_t)(((uint64_t)p.tiles + MAX_CHUNKS - 1) / MAX_CHUNKS);
    if (tpc < floor_tpc) tpc = floor_tpc;
    p.tpc = tpc;
    p.chunks = (p.tiles + tpc - 1) / tpc;
    return QBX_OK;
}

// log2 of the states one chunk of this plan walks (LPT ordering key)
static inline int chunk_work_log2(const Plan& p) {
    return p.L + 6 + (31 - __builtin_clz(p.tpc));
}

}  // namespace qbx

struct qbx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    char* dbuf = nullptr;
    size_t dcap = 0;
    char* hbuf = nullptr;   // pinned staging
    size_t hcap = 0;
    int blocks_per_cu[qbx::kNumVariants];
    int wall_khz = 0;       // rate of the device's wall clock (wall_clock64)
    char* scratch = nullptr;  // this device's qbx_solve_i32_noalloc scratch (hipGetSymbolAddress)
    qbx::WalkTable walk;    // tuned per-n walk bits (all -1 until qbx_set_walk_table)
};

