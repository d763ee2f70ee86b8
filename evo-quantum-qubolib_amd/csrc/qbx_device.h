// qbx_device.h -- device data structures and kernels shared by the single/batch entry
// points (qbx.hip) and the serving loop (qbx_server.hip).  See qbx.hip for the overview.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qbx_walk.h"
#include "qbx_gen.h"
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
    int32_t path;            // QBX_PATH_*
};

enum { KIND_I32 = 0, KIND_F64 = 1 };
static_assert(PAIR_BOUND == QBX_PAIR_BOUND, "pair walk bound of the header and the walk differ");

struct InstDev {
    int32_t n, B, L, M;
    int32_t kind;
    int32_t status;          // written by qbx_prep: 0 or QBX_ERANGE
    uint32_t tiles;          // 64-lane tiles (lane ranks tile*64 .. tile*64+63)
    uint32_t tpc;            // tiles per chunk: one ticket claims one chunk
    uint32_t chunks;         // ceil(tiles / tpc); reduction slots are per chunk
    uint32_t expect;         // chunks committed by the launch(es) of this problem: all of
                             // them, or one part's (the last commit finalizes)
    const void* Q;           // source Q: int32 or double, row-major, variable order
    void* W;                 // n x LDW, bit order: int32 W followed by packed W 65537 |
                             // double followed by int32 Wq and packed coarse Wq
    void* d;                 // 64 entries: int32 | double followed by int32 dq
    void* delta;             // 2^M entries: int32 trie constants followed by the same times
                             // 65537 | FP64 Gray step constants followed by the int32 trie
                             // constants of the quantized tables and their packed coarse copies
    RedDev* red;
    uint32_t* cand;          // chunks entries
    SlotDev* slots;          // chunks entries
    ResultDev* result;
    double qscale;           // float kind: quantization scale 2^s of Wq, dq, dlq
    double qmargin;          // float kind: filter margin in quantized units
    int64_t bound;           // sum |d| + sum_{a<b} |W|  (int kind; quantized for float)
    double fbound;           // same in double (float kind)
    int32_t path;            // set by qbx_prep: QBX_PATH_INT32 | _PAIR16 | _PAIR16_REBASED (int kind),
                             // QBX_PATH_FLOAT | _FLOAT_PAIR16 (float kind)
    int32_t slack;           // rebased pair walks: PAIR_BOUND - D, the guard's room
    int32_t qshift;          // packed float filter: coarse scale 2^-qshift of the quantized one
    int32_t flags;           // qbx_opts_t.flags (QBX_FLAG_FLOAT_INT32_FILTER)
    // per-reference-prefix records (nullable; single solves with L <= B): the prefix of rank
    // r is r >> B, so lane lambda (ranks lambda 2^L ..) lies in prefix lambda >> mshift
    uint64_t* minima;        // order-preserving key of each prefix's least energy (starts ~0)
    uint64_t* ticks;         // wall-clock ticks x 64 of the waves that walked each prefix, prorated
                             // by lanes (starts 0)
    int32_t mshift;          // B - L: lanes per prefix = 2^mshift
    int32_t pad_;
};

// Table regions of one problem (bytes), shared by the batch and serving layouts.
__host__ __device__ inline size_t table_W_bytes(int kind, int n) {
    return (size_t)n * LDW * (kind == KIND_I32 ? 8 : 16);
}
__host__ __device__ inline size_t table_d_bytes(int kind) { return 64 * (kind == KIND_I32 ? 4 : 12); }
__host__ __device__ inline size_t table_delta_bytes(int kind, int M) {
    return (size_t)(1u << M) * (kind == KIND_I32 ? 8 : 16);
}

struct TileRange {
    uint32_t inst;
    uint32_t c_begin;        // first chunk of the range
    uint32_t c_count;
    uint32_t pad;
    uint64_t prefix;         // first ticket of this range
};

constexpr int MAX_GROUPS = 8;  // one persistent launch (and ticket) per inner-bits variant

// ----------------------------------------------------- float filter: a seeded bound
// The float filters walk fastest when their threshold is tight from the first block: with
// no bound, every lane records blocks until its own minimum and the problem's shared bound
// have come down (measured at n = 34: the first three of 32 groups took 3.4x a later one).
// So qbx_seed seeds the shared bound with the energy of a good state: 16 one-flip descents
// (best improvement, FP32), one per wave, each lane holding one variable's flip gain, the
// wave's best move found by a DPP min-reduction of (gain, variable) keys; the least result
// is re-evaluated in FP64.  Any state's energy bounds the answer from above; the seed adds
// 2^-40 of the energy bound (far above the rounding of any FP64 energy sum of at most
// 56 x 57 / 2 terms), so that the walk's own evaluation of that state never lies above it
// and its tile still commits as a candidate.
// the device library's wave reductions (DPP sequences; the f32 sum is declared by the HIP
// headers only under HIP_ENABLE_EXTRA_WARP_SYNC_TYPES)
extern "C" __device__ __attribute__((const)) unsigned int __ockl_wfred_min_u32(unsigned int);
extern "C" __device__ __attribute__((const)) float __ockl_wfred_add_f32(float);

constexpr int SEED_MIN_N = 26;   // smaller problems walk too briefly to gain from a seed
constexpr int SEED_WAVES = 16;   // descents per problem (one workgroup of 16 waves)

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// order-preserving key of a float's bits, low 6 bits replaced by c (an index < 64)
__device__ __forceinline__ uint32_t seed_key(float g, uint32_t c) {
    const uint32_t u = __float_as_uint(g);
    return (((u & 0x80000000u) ? ~u : (u | 0x80000000u)) & ~63u) | c;
}

// qbx_seed: one workgroup of SEED_WAVES waves per problem (idx as for qbx_prep, which it
// follows on the same stream).  It reads the prepared bit-ordered FP64 tables W and d, and
// sets the problem's reduction key to the seed; problems it does not serve keep the empty
// key: integer problems, float problems out of range or below SEED_MIN_N, and shares of a
// solve (expect < chunks: the seed's state may lie outside the share).
static __global__ __launch_bounds__(64 * SEED_WAVES) void qbx_seed(InstDev* insts, const uint32_t* idx) {
    InstDev& I = insts[idx ? idx[blockIdx.x] : blockIdx.x];
    const int n = I.n;
    if (I.kind != KIND_F64 || I.status != 0 || n < SEED_MIN_N || I.expect != I.chunks) return;
    const double* W = (const double*)I.W;
    const double* d = (const double*)I.d;
    const int c = (int)__lane_id();            // this lane's variable (bit position c)
    const int wv = (int)(threadIdx.x >> 6);
    __shared__ float wf[MAX_N * 64];           // w_ab in FP32 (row a, column b), d_a on the diagonal
    __shared__ double wd[MAX_N * 64];          // the same in FP64, for the exact energy at the end
    {
        // every load of a thread in flight before its first LDS write
        constexpr int PER = (MAX_N * 64 + 64 * SEED_WAVES - 1) / (64 * SEED_WAVES);
        double v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = (int)threadIdx.x + u * 64 * SEED_WAVES;
            const int a = i >> 6, b = i & 63;
            v[u] = (i < MAX_N * 64 && a < n && b < n) ? (a == b ? d[a] : W[a * LDW + b]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = (int)threadIdx.x + u * 64 * SEED_WAVES;
            if (i < MAX_N * 64) {
                wf[i] = (float)v[u];
                wd[i] = v[u];
            }
        }
    }
    __syncthreads();
    // the descent of this wave: x wave-uniform, lane c holds g_c, the energy change of
    // flipping variable c (+1 for the lanes past n: never a move)
    uint64_t x = splitmix64((uint64_t)wv * 0x100000001b3ull + (uint64_t)n) & lowmask64(n);
    float h = c < n ? wf[c * 64 + c] : 0.0f;    // d_c + sum_{j != c, x_j} w_jc
    for (uint64_t r = x; r; r &= r - 1ull) {
        const int j = __builtin_ctzll(r);
        if (j != c) h += wf[j * 64 + c];
    }
    const bool xc = (x >> c) & 1ull;
    float E = __ockl_wfred_add_f32(xc ? 0.5f * (wf[c * 64 + c] + h) : 0.0f);
    float g = c < n ? (xc ? -h : h) : 1.0f;
    for (int step = 0; step < 2 * n; ++step) {
        const uint32_t m = __ockl_wfred_min_u32(seed_key(g, (uint32_t)c));
        if (m >= 0x80000000u) break;           // no move lowers the energy
        const int bi = (int)(m & 63u);
        E += __uint_as_float(~(m & ~63u));      // the gain, to 2^-17 relative (a heuristic)
        const float sg = ((x >> bi) & 1ull) ? -1.0f : 1.0f;   // the change of x_bi
        x ^= 1ull << bi;
        const float w = wf[bi * 64 + c];
        g = (c == bi) ? -g : g + (((x >> c) & 1ull) ? -sg : sg) * w;
    }
    __shared__ float best_E[SEED_WAVES];
    __shared__ uint64_t best_x[SEED_WAVES];
    if (c == 0) {
        best_E[wv] = E == E ? E : __builtin_inff();
        best_x[wv] = x;
    }
    __syncthreads();
    if (wv != 0) return;
    // the least of the descents, then its exact energy: lane c adds x_c (d_c + sum_{b < c,
    // x_b} W_bc), from the FP64 copy in LDS
    float eb = best_E[0];
    uint64_t xb = best_x[0];
    for (int k = 1; k < SEED_WAVES; ++k)
        if (best_E[k] < eb) {
            eb = best_E[k];
            xb = best_x[k];
        }
    double v = 0.0;
    for (int b = 0; b < n; ++b) {
        const double w = wd[b * 64 + c];
        v += (b < c && ((xb >> b) & 1ull)) ? w : 0.0;
    }
    const int cd = c < n ? c : 0;       // the diagonal entry d_c (index kept inside the table)
    v = (c < n && ((xb >> c) & 1ull)) ? v + wd[cd * 65] : 0.0;
    for (int o = 32; o > 0; o >>= 1) v += wave_xor(v, o);
    if (c == 0 && v - v == 0.0) I.red->minkey = key_of_f64(v + I.fbound * 0x1p-40);
}

// ------------------------------------------------------------------------------- prep
// idx (nullable): problem index per workgroup; tickets (nullable): zeroed by workgroup 0;
// staged (nullable): descriptors to install first (staged[b] -> insts[idx[b]]), so that a
// batch of admissions needs this one launch.  idx, staged and the source matrix I.Q may lie
// in mapped host memory: each is read once (the matrix into LDS, where the table loops read
// it), and nothing reads them after this kernel.
static __global__ __launch_bounds__(256) void qbx_prep(InstDev* insts, const uint32_t* idx,
                                                      unsigned long long* tickets,
                                                      const InstDev* staged) {
    InstDev& I = insts[idx ? idx[blockIdx.x] : blockIdx.x];
    const int tid = threadIdx.x;
    if (tickets && blockIdx.x == 0 && tid < MAX_GROUPS) tickets[tid] = 0ull;
    if (staged) {
        if (tid == 0) I = staged[blockIdx.x];
        __syncthreads();
    }
    const int n = I.n, M = I.M;
    __shared__ double red_f[256];
    __shared__ long long red_i[256];
    __shared__ uint32_t q_lds[MAX_N * MAX_N * 2];
    {
        // batches of 8 loads per thread, all in flight before the first LDS write (one at a
        // time, each load waited for its memory round trip: most of this kernel's time, and
        // more when the matrix lies in mapped host memory); indices clamped into the matrix
        const uint32_t* src = (const uint32_t*)I.Q;
        const int words = n * n * (I.kind == KIND_I32 ? 1 : 2);
        const int step = (int)blockDim.x;
        for (int w0 = tid; w0 < words; w0 += 8 * step) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int w = w0 + u * step;
                v[u] = src[w < words ? w : words - 1];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int w = w0 + u * step;
                if (w < words) q_lds[w] = v[u];
            }
        }
        __syncthreads();
    }
    if (I.kind == KIND_I32) {
        const int32_t* Q = (const int32_t*)q_lds;
        int32_t* W = (int32_t*)I.W;
        int32_t* d = (int32_t*)I.d;
        int32_t* dl = (int32_t*)I.delta;
        uint32_t* Wp = (uint32_t*)(W + (size_t)n * LDW);
        uint32_t* dlp = (uint32_t*)(dl + (1 << M));
        long long acc = 0;
        for (int idx = tid; idx < n * LDW; idx += blockDim.x) {
            const int a = idx / LDW, b = idx - a * LDW;
            const long long w = w_elem<int32_t, long long>(Q, n, a, b);
            W[idx] = (int32_t)w;
            Wp[idx] = pack_both((uint32_t)w);
            if (b > a && b < n) acc += w < 0 ? -w : w;
        }
        for (int a = tid; a < n; a += blockDim.x) {
            const long long v = d_elem<int32_t, long long>(Q, n, a);
            d[a] = (int32_t)v;
            acc += v < 0 ? -v : v;
        }
        for (uint32_t t = tid; t < (1u << M); t += blockDim.x) {
            dl[t] = (int32_t)trie_elem<int32_t, long long>(Q, n, M, t);
            dlp[t] = pack_both((uint32_t)dl[t]);
        }
        red_i[tid] = acc;
        // the rebased walk's guard constant and field estimate (rp_guard_D, rp_worth_it):
        // inner bit k's share on thread k, from the matrix in LDS, beside the reduction
        // (one thread reading the tables back from memory took most of this kernel's time)
        __shared__ long long sh_Dk[8];
        __shared__ double sh_Sk[8];
        if (tid < M) {
            const int k = tid;
            long long Dk = 0;
            for (int l = k + 1; l < M; ++l) Dk += llabs(w_elem<int32_t, long long>(Q, n, k, l));
            Dk += llabs(w_elem<int32_t, long long>(Q, n, M, k));
            for (int i = 0; i < GROUP_BITS && M + 1 + i < I.L; ++i)
                Dk += llabs(w_elem<int32_t, long long>(Q, n, M + 1 + i, k));
            double mean = (double)(int32_t)d_elem<int32_t, long long>(Q, n, k), var = 0.0;
            for (int j = M; j < n; ++j) {
                const double w = (double)(int32_t)w_elem<int32_t, long long>(Q, n, k, j);
                mean += 0.5 * w;
                var += 0.25 * w * w;
            }
            sh_Dk[k] = Dk;
            sh_Sk[k] = sqrt(mean * mean + var);
        }
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) red_i[tid] += red_i[tid + s];
            __syncthreads();
        }
        if (tid == 0) {
            I.bound = red_i[0];
            I.status = red_i[0] > 0x7ffffffeLL ? QBX_ERANGE : 0;
            // the walk: packed pairs of absolute 16-bit energies, of block-relative ones, or int32
            I.path = QBX_PATH_INT32;
            I.slack = 0;
            if (I.L > M && red_i[0] <= PAIR_BOUND) {
                I.path = QBX_PATH_PAIR16;
            } else if (I.status == 0 && I.L > M) {
                int64_t D = 0;
                double S = 0.0;
                for (int k = 0; k < M; ++k) {
                    D += sh_Dk[k];
                    S += sh_Sk[k];
                }
                if (D < PAIR_BOUND && (double)D + 1.5 * S <= (double)PAIR_BOUND) {
                    I.path = QBX_PATH_PAIR16_REBASED;
                    I.slack = (int32_t)(PAIR_BOUND - D);
                }
            }
        }
    } else {
        // FP64 tables for the exact re-walks, then their quantized int32 copies for the filter
        const double* Q = (const double*)q_lds;
        double* W = (double*)I.W;
        double* d = (double*)I.d;
        double* dld = (double*)I.delta;
        int32_t* Wq = (int32_t*)(W + (size_t)n * LDW);
        int32_t* dq = (int32_t*)(d + 64);
        int32_t* dlq = (int32_t*)(dld + (1 << M));
        double acc = 0.0;
        bool finite = true;
        for (int idx = tid; idx < n * LDW; idx += blockDim.x) {
            const int a = idx / LDW, b = idx - a * LDW;
            const double w = w_elem<double, double>(Q, n, a, b);
            W[idx] = w;
            finite = finite && (w - w == 0.0);
            if (b > a && b < n) acc += fabs(w);
        }
        for (int a = tid; a < 64; a += blockDim.x) {
            const double v = a < n ? d_elem<double, double>(Q, n, a) : 0.0;
            d[a] = v;
            finite = finite && (v - v == 0.0);
            acc += fabs(v);
        }
        for (uint32_t t = tid; t < (1u << M); t += blockDim.x)
            dld[t] = delta_elem<double, double>(Q, n, M, t);
        red_f[tid] = finite ? acc : __builtin_inf();
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) red_f[tid] += red_f[tid + s];
            __syncthreads();
        }
        const double fb = red_f[0];
        const bool ok = fb <= 1.0e300;
        const double scale = ok ? q_scale(fb, n) : 1.0;
        __syncthreads();
        // the quantized table also into LDS, where the steps below read it (W and d are
        // recomputed from the matrix in LDS: no read of what this kernel just wrote to memory)
        __shared__ int32_t wq_lds[MAX_N * LDW];
        long long qacc = 0;
        for (int idx = tid; idx < n * LDW; idx += blockDim.x) {
            const int a = idx / LDW, b = idx - a * LDW;
            const int32_t q = ok ? (int32_t)rint(w_elem<double, double>(Q, n, a, b) * scale) : 0;
            Wq[idx] = q;
            wq_lds[idx] = q;
            if (b > a && b < n) qacc += q < 0 ? -(long long)q : q;
        }
        for (int a = tid; a < 64; a += blockDim.x) {
            const int32_t q = ok ? (int32_t)rint((a < n ? d_elem<double, double>(Q, n, a) : 0.0) * scale) : 0;
            dq[a] = q;
            qacc += q < 0 ? -(long long)q : q;
        }
        red_i[tid] = qacc;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) red_i[tid] += red_i[tid + s];
            __syncthreads();
        }
        // trie constants of the filter from the quantized couplings (exact int32 sums)
        for (uint32_t t = tid; t < (1u << M); t += blockDim.x)
            dlq[t] = (int32_t)trie_from_table<int32_t, long long>(wq_lds, M, t);
        // the packed filter's shift choice (it reads the inner, pair and group bits' rows):
        // thread 0 takes the estimate, then one thread per candidate shift tests it
        __shared__ double sh_S, sh_rms[8], sh_din[8];
        __shared__ int sh_start;
        __shared__ int sh_ok[32];
        __shared__ long long sh_D[32];
        // finite energies whose quantized bound fits int32 (q_scale guarantees it)
        const bool in_range = ok && red_i[0] <= 0x7ffffffeLL;
        const bool try_packed = in_range && !(I.flags & QBX_FLAG_FLOAT_INT32_FILTER) && I.L > M + 1;
        __syncthreads();
        if (try_packed && tid < M) qrp_shift_term(wq_lds, dq, n, M, tid, &sh_rms[tid], &sh_din[tid]);
        __syncthreads();
        if (tid == 0) {
            double S = 0.0, Din = 0.0;
            for (int k = 0; k < M; ++k) {
                S += sh_rms[k];
                Din += sh_din[k];
            }
            sh_S = S;
            sh_start = try_packed ? qrp_shift_first(S, Din, M) : -1;
        }
        __syncthreads();
        const int s0 = sh_start;
        if (tid < 31) {
            int64_t D = 0;
            sh_ok[tid] = (s0 >= 0 && tid >= s0 && qrp_shift_ok(wq_lds, M, I.L, tid, sh_S, &D)) ? 1 : 0;
            sh_D[tid] = D;
        }
        __syncthreads();
        if (tid == 0) {
            I.fbound = fb;
            I.qscale = scale;
            I.qmargin = q_margin(n);
            I.bound = red_i[0];
            I.status = in_range ? 0 : QBX_ERANGE;
            // the packed filter wherever a coarse scale leaves its guard room: the smallest
            // shift from the estimate up that passes (what qrp_shift finds serially)
            int sh = -1;
            int64_t D = 0;
            for (int s = s0 < 0 ? 31 : s0; s <= 30; ++s)
                if (sh_ok[s]) {
                    sh = s;
                    D = sh_D[s];
                    break;
                }
            I.path = sh >= 0 ? QBX_PATH_FLOAT_PAIR16 : QBX_PATH_FLOAT;
            I.qshift = sh >= 0 ? sh : 0;
            I.slack = sh >= 0 ? (int32_t)(PAIR_BOUND - D) : 0;
        }
        __syncthreads();
        if (I.path == QBX_PATH_FLOAT_PAIR16) {
            // coarse packed copies: rnd(v / 2^sh) in both halves (|.| < 2^15 by the guard)
            const int sh = I.qshift;
            uint32_t* Wpc = (uint32_t*)(Wq + (size_t)n * LDW);
            uint32_t* dlpc = (uint32_t*)(dlq + (1 << M));
            for (int idx = tid; idx < n * LDW; idx += blockDim.x) Wpc[idx] = pack_both((uint32_t)rnd_shr(wq_lds[idx], sh));
            for (uint32_t t = tid; t < (1u << M); t += blockDim.x) dlpc[t] = pack_both((uint32_t)rnd_shr(dlq[t], sh));
        }
    }
    if (tid == 0) {
        I.red->minkey = ~0ull;
        I.red->cand_count = 0u;
        I.red->done = 0u;
    }
}

// ------------------------------------------------------------ RandomQUBO generation
// One thread per problem: the problem's matrix, drawn from its PCG64 state (4 words per
// problem: state hi, lo, increment hi, lo) exactly as numpy draws it, written where the
// host would have staged it (I.Q, in the element type of the problem's kind).
static __global__ __launch_bounds__(64) void qbx_gen_random(const InstDev* insts, int count, const uint64_t* rng,
                                                           int64_t low, int64_t high, int dtype) {
    const int i = (int)(blockIdx.x * 64u + threadIdx.x);
    if (i >= count) return;
    const InstDev& I = insts[i];
    Pcg64 g;
    g.hi = rng[4 * i + 0];
    g.lo = rng[4 * i + 1];
    g.inc_hi = rng[4 * i + 2];
    g.inc_lo = rng[4 * i + 3];
    g.buf32 = 0;
    g.buf16 = 0;
    g.has32 = g.has16 = 0;
    if (I.kind == KIND_I32) gen_random_qubo<int32_t>(g, I.n, low, high, dtype, (int32_t*)I.Q);
    else gen_random_qubo<double>(g, I.n, low, high, dtype, (double*)I.Q);
}

// One thread per problem: the TSP QUBO of n_cities = sqrt(n) cities from their distances
// (TSP_MAX_PAIRS doubles per problem, the upper triangle row by row; tsp.py:285-321), as
// double, into I.Q.  dtype: the cities' type (GEN_FLOAT32 | GEN_FLOAT64).
static __global__ __launch_bounds__(64) void qbx_gen_tsp(const InstDev* insts, int count, const double* dist,
                                                        int dtype) {
    const int i = (int)(blockIdx.x * 64u + threadIdx.x);
    if (i >= count) return;
    const InstDev& I = insts[i];
    int nc = 1;
    while ((nc + 1) * (nc + 1) <= I.n) ++nc;
    const double* d = dist + (size_t)i * TSP_MAX_PAIRS;
    if (dtype == GEN_FLOAT32) tsp_qubo<float>(nc, d, (double*)I.Q);
    else tsp_qubo<double>(nc, d, (double*)I.Q);
}

// Grid-stride fill of a device buffer (graph-safe, where a memset node is not).
static __global__ __launch_bounds__(256) void qbx_fill_u64(uint64_t* p, uint64_t count, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256u) p[i] = v;
}

// Per-reference-prefix records of one tile (all 64 lanes call it together): key is the
// order-preserving key of the lane's least energy (~0 for a lane past the problem's end),
// dt the wave's wall-clock ticks for the tile.  Each aligned run of min(2^mshift, 64) lanes
// belongs to one prefix: its minimum key is folded into the prefix's entry, and the run's
// share of the tile's time (its lanes / 64 of dt) into the prefix's ticks (kept x 64).
__device__ __forceinline__ void record_prefixes(const InstDev& I, uint64_t lam0, uint64_t key, uint64_t dt,
                                                int lane) {
    const int gw = I.mshift < 6 ? I.mshift : 6;
    for (int s = 0; s < gw; ++s) {
        const uint64_t k2 = wave_xor(key, 1 << s);
        key = k2 < key ? k2 : key;
    }
    if ((lane & ((1 << gw) - 1)) == 0 && key != ~0ull) {
        const uint64_t p = (lam0 + (uint64_t)lane) >> I.mshift;
        atomicMin((unsigned long long*)(I.minima + p), (unsigned long long)key);
        atomicAdd((unsigned long long*)(I.ticks + p), (unsigned long long)(dt << gw));
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

static __device__ void finalize_problem(const InstDev& I, int lane) {
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
        I.result->path = I.path;
    }
}

// Commit of one chunk, or of a wave's run of `count` chunks of one problem (its best state
// under the chunk slot `tile`, one of the run's): the reduction key, a result slot and a
// candidate entry if the run may hold the answer, then the chunk count; the commit that
// completes the count finalizes.  Only candidates write what the finalizer reads (slot and
// entry), so only they fence before counting: the finalizer's acquire fence synchronizes
// with each candidate's release fence through the read-modify-write chain on `done` (the
// other commits' relaxed increments belong to it).  An agent-scope fence writes back and
// invalidates the L2 of the wave's XCD; on every chunk it cost about 14 us per wave at the
// end of a one-tile-per-wave solve.
__device__ __forceinline__ void commit_tile(const InstDev& I, uint32_t tile, uint64_t key,
                                            uint64_t Ebits, uint64_t rank, int lane, uint32_t count = 1u) {
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
            __threadfence();
        }
        const uint32_t done = atomicAdd(&I.red->done, count);
        last = (done + count == I.expect) ? 1 : 0;
    }
    last = __shfl(last, 0);
    if (last) finalize_problem(I, lane);
}

// Tickets of a launch (its counter starts at 0, reset by qbx_prep or the server's step):
// a workgroup's first ticket is its own index, the later ones gridDim.x + the counter's
// value.  All workgroups start at once, and one atomic per workgroup on the counter's one
// address queued them behind each other: at 2048 workgroups the last got its first ticket
// 26 us after the launch.  Every launch's grid is at most its ticket count.
__device__ __forceinline__ bool next_tile(const TileRange* __restrict__ ranges, int nranges,
                                          unsigned long long total, unsigned long long* ticket,
                                          int lane, bool& first, uint32_t& inst, uint32_t& chunk) {
    unsigned long long t = 0;
    if (first) {
        t = blockIdx.x;
        first = false;
    } else {
        if ((unsigned long long)gridDim.x >= total) return false;   // the first tickets were all
        if (lane == 0) t = (unsigned long long)gridDim.x + atomicAdd(ticket, 1ull);
        // readfirstlane: the ticket (and so the problem, its tables and every loop bound
        // derived from them) is wave-uniform and lives in SGPRs; W rows are then scalar loads
        const uint32_t tlo = __builtin_amdgcn_readfirstlane((uint32_t)t);
        const uint32_t thi = __builtin_amdgcn_readfirstlane((uint32_t)(t >> 32));
        t = ((unsigned long long)thi << 32) | tlo;
    }
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
// K uniform table entries into VGPRs (they are too many for the scalar registers), 16 bytes
// per load where K allows: src is 16-byte aligned (every table section is), and z is a
// lane-dependent zero that keeps the compiler from treating the values as uniform.
template <int K>
__device__ __forceinline__ void load_uniform_vgprs(uint32_t* dst, const uint32_t* __restrict__ src, uint32_t z) {
    if constexpr (K % 4 == 0) {
        const uint4* s4 = (const uint4*)src;
#pragma unroll
        for (int k = 0; k < K / 4; ++k) {
            const uint4 v = s4[k + z];
            dst[4 * k] = v.x;
            dst[4 * k + 1] = v.y;
            dst[4 * k + 2] = v.z;
            dst[4 * k + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) dst[k] = src[k + z];
    }
}

__device__ __forceinline__ uint32_t uni32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    return ((uint64_t)uni32((uint32_t)(v >> 32)) << 32) | uni32((uint32_t)v);
}

// (E, lane rank, block key) minimum over the wave, made uniform
__device__ __forceinline__ void wave_best_block(int32_t& Eb, uint64_t& lamb, uint32_t& Kb) {
    for (int off = 32; off > 0; off >>= 1) {
        const int32_t E2 = __shfl_xor(Eb, off);
        const uint64_t l2 = __shfl_xor(lamb, off);
        const uint32_t K2 = __shfl_xor(Kb, off);
        if (E2 < Eb || (E2 == Eb && (l2 < lamb || (l2 == lamb && K2 < Kb)))) { Eb = E2; lamb = l2; Kb = K2; }
    }
    Eb = (int32_t)uni32((uint32_t)Eb);
    lamb = uni64(lamb);
    Kb = uni32(Kb);
}

// The end of a wave's run of chunks of one problem: its best block (wave-uniform E, lane
// rank, block key over the run) is re-walked once for its lowest inner rank, with the trie
// constants read from memory, and the run commits once, counting all its chunks.
template <int M>
__device__ __forceinline__ void commit_run_i32(const InstDev& I, uint32_t chunk, uint32_t count, int32_t Eb,
                                               uint64_t lamb, uint32_t Kb, int lane) {
    uint64_t rank = ~0ull;
    if (Eb != 0x7fffffff)
        rank = block_rank_hi(lamb, I.L, M, Kb) |
               block_rank_lo_i32<M>((const int32_t*)I.W, (const int32_t*)I.d, I.n, I.L, I.B, lamb, Eb, Kb,
                                    (const int32_t*)I.delta);
    commit_tile(I, chunk, key_of_i64(Eb), (uint64_t)(int64_t)Eb, rank, lane, count);
}

// The walk qbx_prep chose per problem: the packed pair walk (bound <= PAIR_BOUND), the
// rebased pair walk (block-relative 16-bit energies, guarded per group), or the int32 walk.
// All give the same lane keys; the winning block is re-walked by the int32 code either way.
// A wave keeps the best block of its consecutive chunks of one problem and commits them
// together when its next ticket belongs to another problem or none is left: (E, lane rank,
// block key) orders blocks as their states' ranks, so the run's best is the best of its
// chunks' bests.  (Per-chunk commits were three atomics per chunk: 3.1 MB of writes per
// n = 40 launch.)
// REC: per-reference-prefix records (I.minima).  A template flag, so that the walk without
// records compiles exactly as if they did not exist (its kernel is at the scalar-register
// ceiling, and a timestamp kept live across the walk costs it 3%).
template <int M, int OMAX, bool REC>
static __global__ __launch_bounds__(64) void qbx_walk_i32(const InstDev* __restrict__ insts,
                                                   const TileRange* __restrict__ ranges, int nranges,
                                                   unsigned long long total,
                                                   unsigned long long* ticket, int zero) {
    const int lane = threadIdx.x;
    uint32_t inst, chunk;
    bool first = true;
    // the current run (wave-uniform): its problem, last chunk, chunk count and best block
    uint32_t run_inst = 0, run_chunk = 0, run_count = 0;
    int32_t rEb = 0x7fffffff;
    uint64_t rlamb = ~0ull;
    uint32_t rKb = 0xffffffffu;
    for (;;) {
        const bool more = next_tile(ranges, nranges, total, ticket, lane, first, inst, chunk);
        if (run_count != 0u && (!more || inst != run_inst)) {
            commit_run_i32<M>(insts[run_inst], run_chunk, run_count, rEb, rlamb, rKb, lane);
            run_count = 0u;
            rEb = 0x7fffffff;
            rlamb = ~0ull;
            rKb = 0xffffffffu;
        }
        if (!more) break;
        const InstDev& I = insts[inst];
        // each lane's best block over the chunk: (E, lane rank, block key)
        int32_t Eb = 0x7fffffff;
        uint64_t lamb = ~0ull;
        uint32_t Kb = 0xffffffffu;
        const int n = I.n, L = I.L, B = I.B, H = n - L;
        const int32_t* W = (const int32_t*)I.W;
        const int32_t* d = (const int32_t*)I.d;
        const int32_t* dsrc = (const int32_t*)I.delta;
        const uint32_t t0 = chunk * I.tpc;
        const uint32_t t1 = (t0 + I.tpc < I.tiles) ? t0 + I.tpc : I.tiles;
        const int path = I.path;
        constexpr bool rec = REC;
        if (I.status == 0 && (path == QBX_PATH_PAIR16 || path == QBX_PATH_PAIR16_REBASED)) {
            const uint32_t* Wp = (const uint32_t*)(W + (size_t)n * LDW);
            // packed trie constants: uniform values kept in VGPRs (the lane-dependent zero
            // offset keeps them out of the scalar register file)
            uint32_t dlp[1 << M];
            const uint32_t* psrc = (const uint32_t*)(dsrc + (1 << M));
            load_uniform_vgprs<1 << M>(dlp, psrc, lane & zero);
            for (uint32_t tile = t0; tile < t1; ++tile) {
                const uint64_t lam = (uint64_t)tile * 64u + (uint64_t)lane;
                const uint64_t c0 = rec ? wall_clock64() : 0ull;
                uint64_t lkey = ~0ull;
                if (H >= 6 || lam < (1ull << H)) {
                    const int64_t key = path == QBX_PATH_PAIR16
                                            ? lane_best_block_pair<M, OMAX>(W, Wp, d, n, L, B, lam, dlp)
                                            : lane_best_block_rp<M, OMAX>(W, Wp, d, n, L, B, lam, dlp, I.slack,
                                                                          nullptr);
                    if (key_E(key) < Eb) { Eb = key_E(key); lamb = lam; Kb = key_K(key); }
                    lkey = key_of_i64(key_E(key));
                }
                if (rec) record_prefixes(I, (uint64_t)tile * 64u, lkey, wall_clock64() - c0, lane);
            }
        } else {
            int32_t dl[1 << M];
            load_uniform_vgprs<1 << M>((uint32_t*)dl, (const uint32_t*)dsrc, lane & zero);
            if (I.status == 0) {
                for (uint32_t tile = t0; tile < t1; ++tile) {
                    const uint64_t lam = (uint64_t)tile * 64u + (uint64_t)lane;
                    const uint64_t c0 = rec ? wall_clock64() : 0ull;
                    uint64_t lkey = ~0ull;
                    if (H >= 6 || lam < (1ull << H)) {
                        const int64_t key = lane_best_block_i32<M, OMAX>(W, d, n, L, B, lam, dl);
                        if (key_E(key) < Eb) { Eb = key_E(key); lamb = lam; Kb = key_K(key); }
                        lkey = key_of_i64(key_E(key));
                    }
                    if (rec) record_prefixes(I, (uint64_t)tile * 64u, lkey, wall_clock64() - c0, lane);
                }
            }
        }
        wave_best_block(Eb, lamb, Kb);
        if (Eb < rEb || (Eb == rEb && (lamb < rlamb || (lamb == rlamb && Kb < rKb)))) {
            rEb = Eb;
            rlamb = lamb;
            rKb = Kb;
        }
        run_inst = inst;
        run_chunk = chunk;
        ++run_count;
    }
}

// ------------------------------------------------------------------ float-filter walk
// Running upper bound shared by every wave of one problem: the problem's reduction key
// (an upper bound on the energy of some state, or empty; qbx_prep may seed it).
// publish() takes the wave minimum of the lanes' exact bests and lowers the key when it
// improves on the least value this wave has offered or read (note()), so that waves which
// hold nothing better than the shared bound send no atomic.  All 64 lanes must call
// publish() together (it shuffles across the wave).
struct WaveBound {
    unsigned long long* minkey;
    double offered;
    bool live;     // false: no shared bound, every lane finds its own exact minimum
    __device__ double refresh() const {
        return live ? energy_of_key(ld_relaxed((const uint64_t*)minkey)) : __builtin_inf();
    }
    __device__ void note(double b) { offered = b < offered ? b : offered; }
    __device__ void publish(double e) {
        if (!live) return;
        for (int off = 32; off > 0; off >>= 1) {
            const double e2 = wave_xor(e, off);
            e = e2 < e ? e2 : e;
        }
        if (e < offered) {
            offered = e;
            if (threadIdx.x == 0) atomicMin(minkey, (unsigned long long)key_of_f64(e));
        }
    }
};

// Two waves per SIMD (at most 256 VGPRs): the filters' latency hides behind the other wave.
template <int M, int OMAX, bool REC>
static __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void qbx_walk_f64(const InstDev* __restrict__ insts,
                                                   const TileRange* __restrict__ ranges, int nranges,
                                                   unsigned long long total,
                                                   unsigned long long* ticket, int zero) {
    const int lane = threadIdx.x;
    uint32_t inst, chunk;
    bool first = true;
    while (next_tile(ranges, nranges, total, ticket, lane, first, inst, chunk)) {
        const InstDev& I = insts[inst];
        double E = __builtin_inf();
        uint64_t rank = ~0ull;
        if (I.status == 0) {
            const int n = I.n, L = I.L, H = n - L;
            const double* W = (const double*)I.W;
            const double* d = (const double*)I.d;
            const double* dld = (const double*)I.delta;
            const int32_t* Wq = (const int32_t*)(W + (size_t)n * LDW);
            const int32_t* dq = (const int32_t*)(d + 64);
            const int32_t* dsrc = (const int32_t*)(dld + (1 << M));
            const uint32_t t0 = chunk * I.tpc;
            const uint32_t t1 = (t0 + I.tpc < I.tiles) ? t0 + I.tpc : I.tiles;
            // branch and bound across waves through the problem's running minimum key; with
            // per-prefix records every lane needs its own exact minimum, so no shared bound
            constexpr bool rec = REC;
            WaveBound wb{&I.red->minkey, __builtin_inf(), !rec};
            // per-lane candidate records, lane-interleaved (entry k of lane l at k * 64 + l)
            __shared__ uint32_t qc[QCAP * 64];
            __shared__ int32_t qmn[QCAP * 64];
            // every lane walks (the walks shuffle across the wave); a lane past the end of a
            // small problem repeats lane 0's work and its result is dropped
            if (I.path == QBX_PATH_FLOAT_PAIR16) {
                const uint32_t* Wpc = (const uint32_t*)(Wq + (size_t)n * LDW);
                const uint32_t* psrc = (const uint32_t*)(dsrc + (1 << M));
                uint32_t dlp[1 << M];
                // the walk without records makes its opaque zero here: kept from the kernel's
                // start it costs a scratch slot (the recording walk allocates best as it was)
                uint32_t z = (uint32_t)(lane & zero);
                if constexpr (!REC) QBX_OPAQUE(z);
                load_uniform_vgprs<1 << M>(dlp, psrc, z);
                for (uint32_t tile = t0; tile < t1; ++tile) {
                    const uint64_t lam0 = (uint64_t)tile * 64u;
                    const uint64_t lam = lam0 + (uint64_t)lane;
                    const bool active = H >= 6 || lam < (1ull << H);
                    const uint64_t c0 = rec ? wall_clock64() : 0ull;
                    const LaneOutF64 o = walk_lane_qrp<M, OMAX, !REC>(Wq, dq, Wpc, W, d, n, L, I.B, active ? lam : lam0,
                                                                dlp, dsrc, dld, I.qscale, I.qmargin, I.qshift,
                                                                I.slack, wb, qc + lane, qmn + lane, 64, nullptr);
                    if (active && (o.E < E || (o.E == E && o.rank < rank))) { E = o.E; rank = o.rank; }
                    if (rec) record_prefixes(I, lam0, active ? key_of_f64(o.E) : ~0ull, wall_clock64() - c0, lane);
                }
            } else {
                int32_t dlq[1 << M];
                uint32_t z = (uint32_t)(lane & zero);
                if constexpr (!REC) QBX_OPAQUE(z);
                load_uniform_vgprs<1 << M>((uint32_t*)dlq, (const uint32_t*)dsrc, z);
                for (uint32_t tile = t0; tile < t1; ++tile) {
                    const uint64_t lam0 = (uint64_t)tile * 64u;
                    const uint64_t lam = lam0 + (uint64_t)lane;
                    const bool active = H >= 6 || lam < (1ull << H);
                    const uint64_t c0 = rec ? wall_clock64() : 0ull;
                    const LaneOutF64 o = walk_lane_q32<M, OMAX, !REC>(Wq, dq, W, d, n, L, I.B, active ? lam : lam0, dlq,
                                                                dld, I.qscale, I.qmargin, wb, qc + lane, qmn + lane,
                                                                64);
                    if (active && (o.E < E || (o.E == E && o.rank < rank))) { E = o.E; rank = o.rank; }
                    if (rec) record_prefixes(I, lam0, active ? key_of_f64(o.E) : ~0ull, wall_clock64() - c0, lane);
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double E2 = wave_xor(E, off);
            const uint64_t r2 = wave_xor(rank, off);
            if (E2 < E || (E2 == E && r2 < rank)) { E = E2; rank = r2; }
        }
        union { double d; uint64_t u; } cv;
        cv.d = E;
        commit_tile(I, chunk, key_of_f64(E), cv.u, rank, lane);
    }
}

}  // namespace qbx

