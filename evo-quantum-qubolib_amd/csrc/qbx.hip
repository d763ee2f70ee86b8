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

#include "qbx_walk.h"
#include "qbx.h"

namespace qbx {

struct SlotDev {
    uint64_t Ebits;
    uint64_t rank;
};
struct RedDev {
    unsigned long long minkey;
    uint32_t cand_count;
    uint32_t done;
};
struct ResultDev {
    uint64_t Ebits;
    uint64_t rank;
    uint64_t state;
    int32_t status;
    int32_t pad;
};

enum { KIND_I32 = 0, KIND_F64 = 1 };

struct InstDev {
    int32_t n, B, L, M;
    int32_t kind;
    int32_t status;          // written by qbx_prep: 0 or QBX_ERANGE
    uint32_t tiles;          // 64-lane tiles (lane ranks tile*64 .. tile*64+63)
    uint32_t tpc;            // tiles per chunk: one ticket claims one chunk
    uint32_t chunks;         // ceil(tiles / tpc); reduction slots are per chunk
    uint32_t pad0;
    const void* Q;           // source Q: int32 or double, row-major, variable order
    void* W;                 // n x LDW, bit order
    void* d;                 // n
    void* delta;             // int32[2^M]  |  double[2^M] followed by float[2^M]
    RedDev* red;
    uint32_t* cand;          // chunks entries
    SlotDev* slots;          // chunks entries
    ResultDev* result;
    double margin;           // FP32 filter margin (float kind)
    int64_t bound;           // sum |d| + sum_{a<b} |W|  (int kind)
    double fbound;           // same in double (float kind)
};

struct TileRange {
    uint32_t inst;
    uint32_t c_begin;        // first chunk of the range
    uint32_t c_count;
    uint32_t pad;
    uint64_t prefix;         // first ticket of this range
};

constexpr int MAX_GROUPS = 8;  // one persistent launch (and ticket) per inner-bits variant

// ------------------------------------------------------------------------------- prep
__global__ __launch_bounds__(256) void qbx_prep(InstDev* insts, unsigned long long* tickets) {
    InstDev& I = insts[blockIdx.x];
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && tid < MAX_GROUPS) tickets[tid] = 0ull;
    const int n = I.n, M = I.M;
    __shared__ double red_f[256];
    __shared__ long long red_i[256];
    if (I.kind == KIND_I32) {
        const int32_t* Q = (const int32_t*)I.Q;
        int32_t* W = (int32_t*)I.W;
        int32_t* d = (int32_t*)I.d;
        int32_t* dl = (int32_t*)I.delta;
        long long acc = 0;
        for (int idx = tid; idx < n * LDW; idx += blockDim.x) {
            const int a = idx / LDW, b = idx - a * LDW;
            const long long w = w_elem<int32_t, long long>(Q, n, a, b);
            W[idx] = (int32_t)w;
            if (b > a && b < n) acc += w < 0 ? -w : w;
        }
        for (int a = tid; a < n; a += blockDim.x) {
            const long long v = d_elem<int32_t, long long>(Q, n, a);
            d[a] = (int32_t)v;
            acc += v < 0 ? -v : v;
        }
        for (uint32_t t = tid; t < (1u << M); t += blockDim.x)
            dl[t] = (int32_t)delta_elem<int32_t, long long>(Q, n, M, t);
        red_i[tid] = acc;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) red_i[tid] += red_i[tid + s];
            __syncthreads();
        }
        if (tid == 0) {
            I.bound = red_i[0];
            I.status = red_i[0] > 0x7ffffffeLL ? QBX_ERANGE : 0;
        }
    } else {
        const double* Q = (const double*)I.Q;
        double* W = (double*)I.W;
        double* d = (double*)I.d;
        double* dld = (double*)I.delta;
        float* dls = (float*)(dld + (1 << M));
        double acc = 0.0;
        for (int idx = tid; idx < n * LDW; idx += blockDim.x) {
            const int a = idx / LDW, b = idx - a * LDW;
            const double w = w_elem<double, double>(Q, n, a, b);
            W[idx] = w;
            if (b > a && b < n) acc += fabs(w);
        }
        for (int a = tid; a < n; a += blockDim.x) {
            const double v = d_elem<double, double>(Q, n, a);
            d[a] = v;
            acc += fabs(v);
        }
        for (uint32_t t = tid; t < (1u << M); t += blockDim.x) {
            const double v = delta_elem<double, double>(Q, n, M, t);
            dld[t] = v;
            dls[t] = (float)v;
        }
        red_f[tid] = acc;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) red_f[tid] += red_f[tid + s];
            __syncthreads();
        }
        if (tid == 0) {
            I.fbound = red_f[0];
            I.margin = filter_margin(red_f[0], M);
            // the FP32 filter needs finite values well inside float range
            I.status = (red_f[0] <= 1.0e30) ? 0 : QBX_ERANGE;
        }
    }
    if (tid == 0) {
        I.red->minkey = ~0ull;
        I.red->cand_count = 0u;
        I.red->done = 0u;
    }
}

// ------------------------------------------------------------- tile commit / finalize
__device__ __forceinline__ uint64_t ld_relaxed(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t key_of_bits(uint64_t Ebits, int kind) {
    if (kind == KIND_I32) return key_of_i64((int64_t)Ebits);
    return (Ebits & 0x8000000000000000ull) ? ~Ebits : (Ebits | 0x8000000000000000ull);
}

// inverse of key_of_f64 (+inf for the empty key)
__device__ __forceinline__ double energy_of_key(uint64_t k) {
    if (k == ~0ull) return __builtin_inf();
    const uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    union { uint64_t u; double d; } v;
    v.u = u;
    return v.d;
}

__device__ void finalize_problem(const InstDev& I, int lane) {
    __threadfence();
    const uint32_t cc = ld_relaxed(&I.red->cand_count);
    uint64_t bkey = ~0ull, brank = ~0ull, bbits = 0;
    for (uint32_t j = lane; j < cc; j += 64) {
        const uint32_t tile = ld_relaxed(&I.cand[j]);
        const uint64_t Eb = ld_relaxed(&I.slots[tile].Ebits);
        const uint64_t rk = ld_relaxed(&I.slots[tile].rank);
        const uint64_t k = key_of_bits(Eb, I.kind);
        if (k < bkey || (k == bkey && rk < brank)) { bkey = k; brank = rk; bbits = Eb; }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t k2 = __shfl_xor(bkey, off);
        const uint64_t r2 = __shfl_xor(brank, off);
        const uint64_t b2 = __shfl_xor(bbits, off);
        if (k2 < bkey || (k2 == bkey && r2 < brank)) { bkey = k2; brank = r2; bbits = b2; }
    }
    if (lane == 0) {
        I.result->Ebits = bbits;
        I.result->rank = brank;
        I.result->state = rank_to_state(brank, I.B);
        I.result->status = I.status;
        I.result->pad = (int32_t)cc;
    }
}

__device__ __forceinline__ void commit_tile(const InstDev& I, uint32_t tile, uint64_t key,
                                            uint64_t Ebits, uint64_t rank, int lane) {
    int last = 0;
    if (lane == 0) {
        const unsigned long long old = atomicMin(&I.red->minkey, (unsigned long long)key);
        if (key <= old) {
            // only candidate tiles are ever read back, so only they write a result slot
            I.slots[tile].Ebits = Ebits;
            I.slots[tile].rank = rank;
            __threadfence();
            const uint32_t ci = atomicAdd(&I.red->cand_count, 1u);
            I.cand[ci] = tile;
        }
        __threadfence();
        const uint32_t done = atomicAdd(&I.red->done, 1u);
        last = (done == I.chunks - 1u) ? 1 : 0;
    }
    last = __shfl(last, 0);
    if (last) finalize_problem(I, lane);
}

__device__ __forceinline__ bool next_tile(const TileRange* __restrict__ ranges, int nranges,
                                          unsigned long long total, unsigned long long* ticket,
                                          int lane, uint32_t& inst, uint32_t& chunk) {
    unsigned long long t = 0;
    if (lane == 0) t = atomicAdd(ticket, 1ull);
    // readfirstlane: the ticket (and so the problem, its tables and every loop bound derived
    // from them) is wave-uniform and lives in SGPRs; W rows are then scalar loads
    const uint32_t tlo = __builtin_amdgcn_readfirstlane((uint32_t)t);
    const uint32_t thi = __builtin_amdgcn_readfirstlane((uint32_t)(t >> 32));
    t = ((unsigned long long)thi << 32) | tlo;
    if (t >= total) return false;
    int lo = 0, hi = nranges - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ranges[mid].prefix <= t) lo = mid; else hi = mid - 1;
    }
    inst = ranges[lo].inst;
    chunk = ranges[lo].c_begin + (uint32_t)(t - ranges[lo].prefix);
    return true;
}

// ----------------------------------------------------------------------- int32 walk
__device__ __forceinline__ uint32_t uni32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    return ((uint64_t)uni32((uint32_t)(v >> 32)) << 32) | uni32((uint32_t)v);
}

template <int M, int OMAX>
__global__ __launch_bounds__(64) void qbx_walk_i32(const InstDev* __restrict__ insts,
                                                   const TileRange* __restrict__ ranges, int nranges,
                                                   unsigned long long total,
                                                   unsigned long long* ticket, int zero) {
    const int lane = threadIdx.x;
    uint32_t inst, chunk;
    while (next_tile(ranges, nranges, total, ticket, lane, inst, chunk)) {
        const InstDev& I = insts[inst];
        // each lane's best block over the chunk: (E, lane rank, block key)
        int32_t Eb = 0x7fffffff;
        uint64_t lamb = ~0ull;
        uint32_t Kb = 0xffffffffu;
        const int n = I.n, L = I.L, B = I.B, H = n - L;
        const int32_t* W = (const int32_t*)I.W;
        const int32_t* d = (const int32_t*)I.d;
        // inner-step constants: uniform values kept in VGPRs (the lane-dependent zero
        // offset keeps them out of the scalar register file)
        int32_t dl[1 << M];
        const int32_t* dsrc = (const int32_t*)I.delta;
#pragma unroll
        for (int k = 0; k < (1 << M); ++k) dl[k] = dsrc[k + (lane & zero)];
        if (I.status == 0) {
            const uint32_t t0 = chunk * I.tpc;
            const uint32_t t1 = (t0 + I.tpc < I.tiles) ? t0 + I.tpc : I.tiles;
            for (uint32_t tile = t0; tile < t1; ++tile) {
                const uint64_t lam = (uint64_t)tile * 64u + (uint64_t)lane;
                if (H >= 6 || lam < (1ull << H)) {
                    const int64_t key = lane_best_block_i32<M, OMAX>(W, d, n, L, B, lam, dl);
                    if (key_E(key) < Eb) { Eb = key_E(key); lamb = lam; Kb = key_K(key); }
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const int32_t E2 = __shfl_xor(Eb, off);
            const uint64_t l2 = __shfl_xor(lamb, off);
            const uint32_t K2 = __shfl_xor(Kb, off);
            if (E2 < Eb || (E2 == Eb && (l2 < lamb || (l2 == lamb && K2 < Kb)))) { Eb = E2; lamb = l2; Kb = K2; }
        }
        // the wave's winning block is now uniform: re-walk it once for its lowest inner rank
        Eb = (int32_t)uni32((uint32_t)Eb);
        lamb = uni64(lamb);
        Kb = uni32(Kb);
        uint64_t rank = ~0ull;
        if (Eb != 0x7fffffff)
            rank = block_rank_hi(lamb, L, M, Kb) | block_rank_lo_i32<M>(W, d, n, L, B, lamb, Eb, Kb, dl);
        commit_tile(I, chunk, key_of_i64(Eb), (uint64_t)(int64_t)Eb, rank, lane);
    }
}

// ------------------------------------------------------------------ float-filter walk
// Running upper bound shared by every wave of one problem: the problem's reduction key
// (an exact energy of some state, or empty).  publish() takes the wave minimum of the
// lanes' exact bests and lowers the key when it improved on what this wave last offered.
// All 64 lanes must call it together (it shuffles across the wave).
struct WaveBound {
    unsigned long long* minkey;
    double offered;
    __device__ double refresh() const { return energy_of_key(ld_relaxed((const uint64_t*)minkey)); }
    __device__ void publish(double e) {
        for (int off = 32; off > 0; off >>= 1) {
            const double e2 = __shfl_xor(e, off);
            e = e2 < e ? e2 : e;
        }
        if (e < offered) {
            offered = e;
            if (threadIdx.x == 0) atomicMin(minkey, (unsigned long long)key_of_f64(e));
        }
    }
};

template <int M, int OMAX>
__global__ __launch_bounds__(64) void qbx_walk_f64(const InstDev* __restrict__ insts,
                                                   const TileRange* __restrict__ ranges, int nranges,
                                                   unsigned long long total,
                                                   unsigned long long* ticket, int zero) {
    const int lane = threadIdx.x;
    uint32_t inst, chunk;
    while (next_tile(ranges, nranges, total, ticket, lane, inst, chunk)) {
        const InstDev& I = insts[inst];
        double E = __builtin_inf();
        uint64_t rank = ~0ull;
        if (I.status == 0) {
            const int n = I.n, L = I.L, H = n - L;
            const double* dld = (const double*)I.delta;
            const float* dsrc = (const float*)(dld + (1 << M));
            float dls[1 << M];
#pragma unroll
            for (int k = 0; k < (1 << M); ++k) dls[k] = dsrc[k + (lane & zero)];
            const uint32_t t0 = chunk * I.tpc;
            const uint32_t t1 = (t0 + I.tpc < I.tiles) ? t0 + I.tpc : I.tiles;
            // branch and bound across waves through the problem's running minimum key
            WaveBound wb{&I.red->minkey, __builtin_inf()};
            for (uint32_t tile = t0; tile < t1; ++tile) {
                // every lane walks (the walk shuffles across the wave); a lane past the end
                // of a small problem repeats lane 0's work and its result is dropped
                const uint64_t lam0 = (uint64_t)tile * 64u;
                const uint64_t lam = lam0 + (uint64_t)lane;
                const bool active = H >= 6 || lam < (1ull << H);
                const LaneOutF64 o = walk_lane_f64<M, OMAX>((const double*)I.W, (const double*)I.d, n, L, I.B,
                                                            active ? lam : lam0, dls, dld, I.margin, wb);
                if (active && (o.E < E || (o.E == E && o.rank < rank))) { E = o.E; rank = o.rank; }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double E2 = __shfl_xor(E, off);
            const uint64_t r2 = __shfl_xor(rank, off);
            if (E2 < E || (E2 == E && r2 < rank)) { E = E2; rank = r2; }
        }
        union { double d; uint64_t u; } cv;
        cv.d = E;
        commit_tile(I, chunk, key_of_f64(E), cv.u, rank, lane);
    }
}

}  // namespace qbx

// =============================================================================== host
using namespace qbx;

namespace {

thread_local std::string g_err;

int set_err(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_err(QBX_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

typedef void (*walk_fn_t)(const InstDev*, const TileRange*, int, unsigned long long,
                          unsigned long long*, int);

struct Variant {
    int kind, M, OMAX;
    const void* fn;
};

// available (kind, M) instantiations, largest M first within each kind
const Variant kVariants[] = {
    {KIND_I32, 7, 13, (const void*)qbx_walk_i32<7, 13>},
    {KIND_I32, 6, 14, (const void*)qbx_walk_i32<6, 14>},
    {KIND_I32, 4, 16, (const void*)qbx_walk_i32<4, 16>},
    {KIND_I32, 1, 19, (const void*)qbx_walk_i32<1, 19>},
    {KIND_F64, 6, 14, (const void*)qbx_walk_f64<6, 14>},
    {KIND_F64, 4, 16, (const void*)qbx_walk_f64<4, 16>},
    {KIND_F64, 1, 19, (const void*)qbx_walk_f64<1, 19>},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

constexpr int MAX_N = 56;
constexpr int MAX_LANE_BITS = 37;   // tiles = 2^(H-6) < 2^31

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

struct qbx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    char* dbuf = nullptr;
    size_t dcap = 0;
    char* hbuf = nullptr;   // pinned staging
    size_t hcap = 0;
    int blocks_per_cu[kNumVariants];
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

int pick_variant(int kind, int want_M, int n) {
    // largest available M <= min(want_M, n) for this kind
    int best = -1;
    for (int v = 0; v < kNumVariants; ++v) {
        if (kVariants[v].kind != kind) continue;
        if (kVariants[v].M <= want_M && kVariants[v].M <= n) {
            if (best < 0 || kVariants[v].M > kVariants[best].M) best = v;
        }
    }
    return best;
}

// Default walk bits per lane (measured on MI355X, int32 kernel): a lone problem is cut
// into about 2^10 .. 2^22 tiles -- L = n - 16, capped at 18 (n = 28: 12, 32: 16, 36 and
// 40: 18); inside a batch many problems share the GPU, so fewer, larger tiles per problem
// keep the per-lane set-up small.
int default_walk_bits(int kind, int n, int M, int OMAX, bool single) {
    (void)kind;   // both walks measured best with the same tiling (float n = 34: L = 18)
    int L = single ? n - 16 : n - 10;
    const int cap = single ? 18 : 16;
    if (L > cap) L = cap;
    if (L < M) L = M;
    if (L < n - MAX_LANE_BITS) L = n - MAX_LANE_BITS;
    if (L > M + OMAX) L = M + OMAX;
    if (L > n) L = n;
    return L;
}

struct Plan {
    int n, B, L, M, kind, variant, H;
    uint32_t tiles, tpc, chunks;
    size_t off_Q, off_W, off_d, off_delta, off_cand, off_slots;
};

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
    const size_t off_Qbase = off;
    for (int i = 0; i < count; ++i) {
        const int n = ns[i];
        if (n < 1 || n > MAX_N) return set_err(QBX_EINVAL, "problem %d: n=%d outside [1, %d]", i, n, MAX_N);
        if (offsets[i] < 0) return set_err(QBX_EINVAL, "problem %d: negative offset", i);
        Plan& p = plans[i];
        p.n = n;
        p.kind = kind;
        p.B = (o.tie_B < 0) ? std::min(n, 20) : o.tie_B;
        if (p.B > n) return set_err(QBX_EINVAL, "problem %d: tie_B=%d exceeds n=%d", i, p.B, n);
        const int wantM = (o.inner_bits < 0) ? (kind == KIND_I32 ? 7 : 6) : o.inner_bits;
        p.variant = pick_variant(kind, wantM, n);
        if (p.variant < 0) return set_err(QBX_EINVAL, "problem %d: no kernel for inner_bits=%d", i, wantM);
        p.M = kVariants[p.variant].M;
        const int OMAX = kVariants[p.variant].OMAX;
        int L = (o.walk_bits < 0) ? default_walk_bits(kind, n, p.M, OMAX, single) : std::min(o.walk_bits, n);
        if (L < p.M) L = p.M;
        if (L > p.M + OMAX || L > n)
            return set_err(QBX_EINVAL, "problem %d: walk_bits=%d outside [%d, %d]", i, L, p.M, std::min(n, p.M + OMAX));
        if (n - L > MAX_LANE_BITS)
            return set_err(QBX_EINVAL, "problem %d: walk_bits=%d leaves %d lane bits (max %d)", i, L, n - L, MAX_LANE_BITS);
        p.L = L;
        p.H = n - L;
        p.tiles = p.H > 6 ? (1u << (p.H - 6)) : 1u;
        // tiles per chunk: amortise the per-ticket work (atomics, table loads, the final
        // re-walk) over about 2^22 states, but keep enough chunks to spread one problem
        // over the GPU (single solves) or over a few waves (inside a batch)
        {
            uint32_t tpc = (22 - 6 - L) > 0 ? (1u << (22 - 6 - L)) : 1u;
            const uint32_t min_chunks = single ? 8192u : 16u;
            const uint32_t cap = p.tiles / min_chunks;
            if (tpc > cap) tpc = cap;
            if (tpc < 1u) tpc = 1u;
            p.tpc = tpc;
            p.chunks = (p.tiles + tpc - 1) / tpc;
        }
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
        // longest chunks first: chunk work = tpc * 64 * 2^L states
        const int wa = plans[a].L + 31 - __builtin_clz(plans[a].tpc);
        const int wb = plans[b].L + 31 - __builtin_clz(plans[b].tpc);
        return wa > wb;
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
    hipLaunchKernelGGL(qbx_prep, dim3(count), dim3(256), 0, c->stream, (InstDev*)(dv + off_inst), dticket);
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
