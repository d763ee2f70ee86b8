// qbx_cpu.cpp -- libqbx_cpu.so: the exhaustive solver on the host CPU (C ABI in
// include/qbx_cpu.h).  Serves device='cpu' (the reference QBF's default device,
// src/qubo_lib/abc/solver_abc.py:62-64).
//
// The lane walk is qbx_walk.h, the very code the gfx950 kernels run, compiled for the host:
// the state space is cut into 2^H lanes of 2^L states (lane lambda owns the ranks
// [lambda 2^L, (lambda + 1) 2^L)), OpenMP threads take lanes from a dynamic schedule, and
// the lanes' keys reduce as the device's waves reduce them:
//   * integer Q: each lane's best block (E, block key) by the int32 trie walk; the
//     lexicographic (E, lane, key) minimum over all lanes is re-walked once for its lowest
//     inner rank (the device's per-chunk winner re-walk, done once here);
//   * float Q: the int32 filter over tables quantized to 2^s units with the FP64 re-walk
//     of every block it passes; the lanes share one running upper bound on the answer
//     (an atomic double, where the device shares the problem's reduction key), and the
//     answer is the FP64 (E, rank) minimum over the lanes.
// The tables are the ones qbx_prep builds on the device (bit order, rows padded to LDW),
// so the inner-block width M (the same default, 7) gives the same FP64 energies as the GPU.
#include <math.h>
#include <omp.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <string>
#include <vector>

#include "qbx.h"
#include "qbx_cpu.h"
#include "qbx_walk.h"

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

constexpr int MAX_N = 56;

// Walk widths: lanes of 2^L states, L = n - 12 within the variant's range, so that a solve
// of n >= 19 has 4096 lanes or more to spread over the threads while each lane's set-up
// (O(n) per lane bit) stays small against its 2^L states.
struct Tiling {
    int n, B, M, OMAX, L, H;
};

// cap_L_at_B: lanes no wider than the tie width B, so that every lane lies inside one
// reference prefix (per-prefix records; needs B >= 1).
int plan(int n, const qbx_opts_t* opts, Tiling& t, bool cap_L_at_B = false) {
    qbx_opts_t o = {-1, -1, -1, 0};
    if (opts) o = *opts;
    if (n < 1 || n > MAX_N) return set_err(QBX_EINVAL, "n=%d outside [1, %d]", n, MAX_N);
    if (o.flags & ~(QBX_FLAG_FLOAT_INT32_FILTER | QBX_FLAG_FLOAT_PACKED_FILTER))
        return set_err(QBX_EINVAL, "unknown opts.flags bits 0x%x", o.flags);
    t.n = n;
    t.B = o.tie_B < 0 ? (n < 20 ? n : 20) : o.tie_B;
    if (t.B > n) return set_err(QBX_EINVAL, "tie_B=%d exceeds n=%d", t.B, n);
    if (cap_L_at_B && t.B < 1) return set_err(QBX_EINVAL, "per-prefix records need tie_B >= 1");
    if (cap_L_at_B && n - t.B > QBX_MAX_PREFIX_BITS)
        return set_err(QBX_EINVAL, "per-prefix records of 2^%d prefixes: at most 2^%d (raise tie_B)", n - t.B,
                       QBX_MAX_PREFIX_BITS);
    int want = o.inner_bits < 0 ? 7 : o.inner_bits;
    if (cap_L_at_B && want > t.B) want = t.B;
    static const int kM[4] = {7, 6, 4, 1}, kO[4] = {13, 14, 16, 19};
    t.M = -1;
    for (int v = 0; v < 4; ++v)
        if (kM[v] <= want && kM[v] <= n) {
            t.M = kM[v];
            t.OMAX = kO[v];
            break;
        }
    if (t.M < 0) return set_err(QBX_EINVAL, "no walk for inner_bits=%d", want);
    int L = o.walk_bits >= 0 ? (o.walk_bits < n ? o.walk_bits : n) : n - 12;
    if (o.walk_bits < 0) {
        if (L > t.M + t.OMAX) L = t.M + t.OMAX;
        if (L > n) L = n;
    }
    if (L < t.M) L = t.M;
    if (cap_L_at_B && L > t.B) L = t.B;
    if (L > t.M + t.OMAX || L > n)
        return set_err(QBX_EINVAL, "walk_bits=%d outside [%d, %d]", L, t.M, n < t.M + t.OMAX ? n : t.M + t.OMAX);
    t.L = L;
    t.H = n - L;
    return QBX_OK;
}

void fill_result(const Tiling& t, qbx_result_t* out, double ms, int path) {
    memset(out, 0, sizeof *out);
    out->n = t.n;
    out->tie_B = t.B;
    out->walk_bits = t.L;
    out->inner_bits = t.M;
    out->lane_bits = t.H;
    out->states = 1ull << t.n;
    out->tiles = (1ull << t.H) >= 64 ? (1ull << t.H) / 64 : 1;
    out->kernel_ms = ms;
    out->path = path;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Per-reference-prefix records (qbx_cpu_solve_minima_*): lane lambda lies in prefix
// lambda >> (B - L); each prefix keeps the order-preserving key of its least energy and
// the nanoseconds its lanes took.
struct Prefixes {
    std::vector<uint64_t> key, ns;
    int shift = 0;
    bool on = false;
    void init(const Tiling& t) {
        on = true;
        shift = t.B - t.L;
        key.assign((size_t)1 << (t.n - t.B), ~0ull);
        ns.assign((size_t)1 << (t.n - t.B), 0ull);
    }
    void add(uint64_t lam, uint64_t k, uint64_t dns) {
        uint64_t* slot = &key[lam >> shift];
        uint64_t cur = __atomic_load_n(slot, __ATOMIC_RELAXED);
        while (k < cur && !__atomic_compare_exchange_n(slot, &cur, k, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
        __atomic_fetch_add(&ns[lam >> shift], dns, __ATOMIC_RELAXED);
    }
};

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------------------ integer Q
template <int M, int OMAX>
int solve_i32_t(const int32_t* Q, const Tiling& t, int threads, qbx_result_t* out, Prefixes* pre) {
    const int n = t.n, L = t.L, B = t.B;
    std::vector<int32_t> W((size_t)n * LDW, 0), d(64, 0), dl(1u << M, 0);
    long long bound = 0;
    for (int a = 0; a < n; ++a) {
        const long long v = d_elem<int32_t, long long>(Q, n, a);
        bound += v < 0 ? -v : v;
        for (int b = 0; b < LDW; ++b) {
            const long long w = w_elem<int32_t, long long>(Q, n, a, b);
            if (b > a && b < n) bound += w < 0 ? -w : w;
            if (bound > 0x7ffffffeLL) break;
            W[(size_t)a * LDW + b] = (int32_t)w;
        }
        if (bound > 0x7ffffffeLL)
            return set_err(QBX_ERANGE, "energy bound above 2^31 - 2: outside the int32 path");
        d[a] = (int32_t)v;
    }
    for (uint32_t y = 0; y < (1u << M); ++y) dl[y] = (int32_t)trie_from_table<int32_t, long long>(W.data(), M, y);

    const auto t0 = std::chrono::steady_clock::now();
    const long long lanes = 1ll << t.H;
    int32_t bE = 0x7fffffff;
    uint64_t bLam = ~0ull;
    uint32_t bK = 0xffffffffu;
#pragma omp parallel num_threads(threads > 0 ? threads : omp_get_max_threads())
    {
        int32_t E = 0x7fffffff;
        uint64_t lamE = ~0ull;
        uint32_t K = 0xffffffffu;
#pragma omp for schedule(dynamic, 16) nowait
        for (long long lam = 0; lam < lanes; ++lam) {
            const uint64_t c0 = pre ? now_ns() : 0;
            const int64_t key = lane_best_block_i32<M, OMAX>(W.data(), d.data(), n, L, B, (uint64_t)lam, dl.data());
            // lanes of one thread come in increasing order: a strict test keeps the lowest
            if (key_E(key) < E) { E = key_E(key); lamE = (uint64_t)lam; K = key_K(key); }
            if (pre) pre->add((uint64_t)lam, key_of_i64(key_E(key)), now_ns() - c0);
        }
#pragma omp critical(qbx_cpu_reduce)
        {
            if (E < bE || (E == bE && (lamE < bLam || (lamE == bLam && K < bK)))) { bE = E; bLam = lamE; bK = K; }
        }
    }
    const uint64_t rank = block_rank_hi(bLam, L, M, bK) |
                          block_rank_lo_i32<M>(W.data(), d.data(), n, L, B, bLam, bE, bK, dl.data());
    fill_result(t, out, ms_since(t0), QBX_PATH_INT32);
    out->E_int = bE;
    out->E = (double)bE;
    out->rank = rank;
    out->state = rank_to_state(rank, B);
    return QBX_OK;
}

// -------------------------------------------------------------------------- float Q
// Running upper bound on the answer shared by the threads (the device shares it through
// the problem's reduction key): the least value any lane has offered.
struct AtomicBound {
    std::atomic<double>* v;
    double refresh() const { return v->load(std::memory_order_relaxed); }
    void publish(double e) {
        double cur = v->load(std::memory_order_relaxed);
        while (e < cur && !v->compare_exchange_weak(cur, e, std::memory_order_relaxed)) {
        }
    }
    void note(double) {}
};

// No shared bound: every lane finds its own exact minimum (per-prefix records).
struct NoBound {
    double refresh() const { return __builtin_inf(); }
    void publish(double) {}
    void note(double) {}
};

template <int M, int OMAX, class Bound>
LaneOutF64 walk_f64_lane(const std::vector<int32_t>& Wq, const std::vector<int32_t>& dq, const std::vector<double>& W,
                         const std::vector<double>& d, const Tiling& t, uint64_t lam, const std::vector<int32_t>& dlq,
                         const std::vector<double>& dld, double scale, double qm, Bound& src, uint32_t* bc,
                         int32_t* bm) {
    return walk_lane_q32<M, OMAX>(Wq.data(), dq.data(), W.data(), d.data(), t.n, t.L, t.B, lam, dlq.data(),
                                  dld.data(), scale, qm, src, bc, bm, 1);
}

template <int M, int OMAX>
int solve_f64_t(const double* Q, const Tiling& t, int threads, qbx_result_t* out, Prefixes* pre) {
    const int n = t.n, B = t.B;
    std::vector<double> W((size_t)n * LDW, 0.0), d(64, 0.0), dld(1u << M, 0.0);
    std::vector<int32_t> Wq((size_t)n * LDW, 0), dq(64, 0), dlq(1u << M, 0);
    double bound = 0.0;
    for (int a = 0; a < n; ++a) {
        d[a] = d_elem<double, double>(Q, n, a);
        bound += fabs(d[a]);
        for (int b = 0; b < LDW; ++b) {
            const double w = w_elem<double, double>(Q, n, a, b);
            W[(size_t)a * LDW + b] = w;
            if (b > a && b < n) bound += fabs(w);
        }
    }
    if (!(bound <= 1.0e300)) return set_err(QBX_ERANGE, "float Q with non-finite values or energy bound above 1e300");
    const double scale = q_scale(bound, n);
    long long qbound = 0;
    for (int a = 0; a < n; ++a) {
        dq[a] = (int32_t)rint(d[a] * scale);
        qbound += dq[a] < 0 ? -(long long)dq[a] : dq[a];
        for (int b = 0; b < LDW; ++b) {
            const int32_t q = (int32_t)rint(W[(size_t)a * LDW + b] * scale);
            Wq[(size_t)a * LDW + b] = q;
            if (b > a && b < n) qbound += q < 0 ? -(long long)q : q;
        }
    }
    if (qbound > 0x7ffffffeLL) return set_err(QBX_ERANGE, "quantized energy bound above 2^31 - 2");
    for (uint32_t y = 0; y < (1u << M); ++y) {
        dld[y] = delta_elem<double, double>(Q, n, M, y);
        dlq[y] = (int32_t)trie_from_table<int32_t, long long>(Wq.data(), M, y);
    }
    const double qm = q_margin(n);

    const auto t0 = std::chrono::steady_clock::now();
    const long long lanes = 1ll << t.H;
    std::atomic<double> shared(__builtin_inf());
    double bE = __builtin_inf();
    uint64_t bR = ~0ull;
#pragma omp parallel num_threads(threads > 0 ? threads : omp_get_max_threads())
    {
        AtomicBound src{&shared};
        uint32_t buf_c[QCAP];
        int32_t buf_m[QCAP];
        double E = __builtin_inf();
        uint64_t R = ~0ull;
        NoBound none;
#pragma omp for schedule(dynamic, 16) nowait
        for (long long lam = 0; lam < lanes; ++lam) {
            const uint64_t c0 = pre ? now_ns() : 0;
            const LaneOutF64 o =
                pre ? walk_f64_lane<M, OMAX>(Wq, dq, W, d, t, (uint64_t)lam, dlq, dld, scale, qm, none, buf_c, buf_m)
                    : walk_f64_lane<M, OMAX>(Wq, dq, W, d, t, (uint64_t)lam, dlq, dld, scale, qm, src, buf_c, buf_m);
            if (o.E < E || (o.E == E && o.rank < R)) { E = o.E; R = o.rank; }
            if (pre) pre->add((uint64_t)lam, key_of_f64(o.E), now_ns() - c0);
        }
#pragma omp critical(qbx_cpu_reduce_f)
        {
            if (E < bE || (E == bE && R < bR)) { bE = E; bR = R; }
        }
    }
    fill_result(t, out, ms_since(t0), QBX_PATH_FLOAT);
    out->E = bE;
    out->E_int = (int64_t)bE;
    out->rank = bR;
    out->state = rank_to_state(bR, B);
    return QBX_OK;
}

int run_i32(const int32_t* Q, const Tiling& t, int threads, qbx_result_t* out, Prefixes* pre) {
    switch (t.M) {
        case 7: return solve_i32_t<7, 13>(Q, t, threads, out, pre);
        case 6: return solve_i32_t<6, 14>(Q, t, threads, out, pre);
        case 4: return solve_i32_t<4, 16>(Q, t, threads, out, pre);
        default: return solve_i32_t<1, 19>(Q, t, threads, out, pre);
    }
}

int run_f64(const double* Q, const Tiling& t, int threads, qbx_result_t* out, Prefixes* pre) {
    switch (t.M) {
        case 7: return solve_f64_t<7, 13>(Q, t, threads, out, pre);
        case 6: return solve_f64_t<6, 14>(Q, t, threads, out, pre);
        case 4: return solve_f64_t<4, 16>(Q, t, threads, out, pre);
        default: return solve_f64_t<1, 19>(Q, t, threads, out, pre);
    }
}

void write_prefixes(const Prefixes& pre, bool f64, void* minima, double* wave_ms) {
    for (size_t i = 0; i < pre.key.size(); ++i) {
        const uint64_t k = pre.key[i];
        if (f64) {
            const uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
            double e;
            memcpy(&e, &u, sizeof e);
            ((double*)minima)[i] = e;
        } else {
            ((int64_t*)minima)[i] = (int64_t)(k ^ 0x8000000000000000ull);
        }
        if (wave_ms) wave_ms[i] = (double)pre.ns[i] * 1e-6;
    }
}

}  // namespace

extern "C" {

const char* qbx_cpu_last_error(void) { return g_err.c_str(); }

int qbx_cpu_solve_i32(const int32_t* Q, int n, const qbx_opts_t* opts, int threads, qbx_result_t* out) {
    if (!Q || !out) return set_err(QBX_EINVAL, "null argument");
    Tiling t;
    const int rc = plan(n, opts, t);
    return rc ? rc : run_i32(Q, t, threads, out, nullptr);
}

int qbx_cpu_solve_f64(const double* Q, int n, const qbx_opts_t* opts, int threads, qbx_result_t* out) {
    if (!Q || !out) return set_err(QBX_EINVAL, "null argument");
    Tiling t;
    const int rc = plan(n, opts, t);
    return rc ? rc : run_f64(Q, t, threads, out, nullptr);
}

int qbx_cpu_solve_minima_i32(const int32_t* Q, int n, const qbx_opts_t* opts, int threads, int64_t* minima,
                             double* wave_ms, qbx_result_t* out) {
    if (!Q || !out || !minima) return set_err(QBX_EINVAL, "null argument");
    Tiling t;
    int rc = plan(n, opts, t, true);
    if (rc) return rc;
    Prefixes pre;
    pre.init(t);
    rc = run_i32(Q, t, threads, out, &pre);
    if (rc == QBX_OK) write_prefixes(pre, false, minima, wave_ms);
    return rc;
}

int qbx_cpu_solve_minima_f64(const double* Q, int n, const qbx_opts_t* opts, int threads, double* minima,
                             double* wave_ms, qbx_result_t* out) {
    if (!Q || !out || !minima) return set_err(QBX_EINVAL, "null argument");
    Tiling t;
    int rc = plan(n, opts, t, true);
    if (rc) return rc;
    Prefixes pre;
    pre.init(t);
    rc = run_f64(Q, t, threads, out, &pre);
    if (rc == QBX_OK) write_prefixes(pre, true, minima, wave_ms);
    return rc;
}

}  // extern "C"
