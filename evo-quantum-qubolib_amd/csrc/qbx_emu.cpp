// qbx_emu.cpp -- host build of the device lane walk (qbx_walk.h), for the CPU test suite.
//
// Not part of the product path: libqbx.so never loads it.  It runs the very same
// walk_lane_* templates the gfx950 kernels run, lane by lane on the CPU, with the same
// bit-ordered tables the prep kernel builds, so the CPU tests can check the walk's
// indexing (lane prefixes, block keys, tie ranks, transitions) against the oracle for
// every (n, L, B, M) shape before any GPU time is spent.
#include <stdint.h>
#include <string.h>
#include <vector>
#include "qbx_walk.h"

using namespace qbx;

namespace {

template <int M, int OMAX>
int emu_i32(const int32_t* Q, int n, int B, int L, int64_t* E_out, uint64_t* rank_out) {
    if (L < M || L > M + OMAX || L > n) return -2;
    std::vector<int32_t> W((size_t)n * LDW, 0), d(n, 0), dl(1 << M, 0);
    for (int a = 0; a < n; ++a) {
        d[a] = d_elem<int32_t, int32_t>(Q, n, a);
        for (int b = 0; b < LDW; ++b) W[(size_t)a * LDW + b] = w_elem<int32_t, int32_t>(Q, n, a, b);
    }
    for (uint32_t t = 0; t < (1u << M); ++t) dl[t] = delta_elem<int32_t, int32_t>(Q, n, M, t);
    const int H = n - L;
    const uint64_t lanes = 1ull << H;
    // the device flow: every lane's best block key, the (E, lane rank, block key) winner
    // over all lanes, then one re-walk of the winning block
    int32_t bE = 0x7fffffff;
    uint64_t bLam = ~0ull;
    uint32_t bK = 0xffffffffu;
    for (uint64_t lam = 0; lam < lanes; ++lam) {
        const int64_t key = lane_best_block_i32<M, OMAX>(W.data(), d.data(), n, L, B, lam, dl.data());
        if (key_E(key) < bE) { bE = key_E(key); bLam = lam; bK = key_K(key); }
    }
    *E_out = bE;
    *rank_out = block_rank_hi(bLam, L, M, bK) |
                block_rank_lo_i32<M>(W.data(), d.data(), n, L, B, bLam, bE, bK, dl.data());
    // the per-lane composition must agree with it
    const LaneOutI32 o = walk_lane_i32<M, OMAX>(W.data(), d.data(), n, L, B, bLam, dl.data());
    if (o.E != bE || o.rank != *rank_out) return -3;
    return 0;
}

template <int M, int OMAX>
int emu_f64(const double* Q, int n, int B, int L, double* E_out, uint64_t* rank_out) {
    if (L < M || L > M + OMAX || L > n) return -2;
    std::vector<double> W((size_t)n * LDW, 0.0), d(n, 0.0), dld(1 << M, 0.0);
    std::vector<float> dls(1 << M, 0.0f);
    double bound = 0.0;
    for (int a = 0; a < n; ++a) {
        d[a] = d_elem<double, double>(Q, n, a);
        bound += d[a] < 0 ? -d[a] : d[a];
        for (int b = 0; b < LDW; ++b) {
            W[(size_t)a * LDW + b] = w_elem<double, double>(Q, n, a, b);
            if (b > a && b < n) bound += W[(size_t)a * LDW + b] < 0 ? -W[(size_t)a * LDW + b] : W[(size_t)a * LDW + b];
        }
    }
    for (uint32_t t = 0; t < (1u << M); ++t) {
        dld[t] = delta_elem<double, double>(Q, n, M, t);
        dls[t] = (float)dld[t];
    }
    const double margin = filter_margin(bound, M);
    const int H = n - L;
    const uint64_t lanes = 1ull << H;
    double bE = __builtin_inf();
    uint64_t bR = ~0ull;
    // lanes see the best energy found so far as their bound, as the device's waves do
    for (uint64_t lam = 0; lam < lanes; ++lam) {
        FixedBound fb{bE};
        LaneOutF64 o = walk_lane_f64<M, OMAX>(W.data(), d.data(), n, L, B, lam, dls.data(), dld.data(), margin, fb);
        if (o.E < bE || (o.E == bE && o.rank < bR)) { bE = o.E; bR = o.rank; }
    }
    *E_out = bE;
    *rank_out = bR;
    return 0;
}

}  // namespace

extern "C" {

// Returns 0 on success; -1 unsupported M; -2 unsupported L.
int qbx_emu_solve_i32(const int32_t* Q, int n, int B, int L, int M, int64_t* E, uint64_t* rank,
                      uint64_t* state) {
    int rc;
    switch (M) {
        case 1: rc = emu_i32<1, 19>(Q, n, B, L, E, rank); break;
        case 4: rc = emu_i32<4, 16>(Q, n, B, L, E, rank); break;
        case 6: rc = emu_i32<6, 14>(Q, n, B, L, E, rank); break;
        case 7: rc = emu_i32<7, 13>(Q, n, B, L, E, rank); break;
        default: return -1;
    }
    if (rc == 0) *state = rank_to_state(*rank, B);
    return rc;
}

int qbx_emu_solve_f64(const double* Q, int n, int B, int L, int M, double* E, uint64_t* rank,
                      uint64_t* state) {
    int rc;
    switch (M) {
        case 1: rc = emu_f64<1, 19>(Q, n, B, L, E, rank); break;
        case 4: rc = emu_f64<4, 16>(Q, n, B, L, E, rank); break;
        case 6: rc = emu_f64<6, 14>(Q, n, B, L, E, rank); break;
        default: return -1;
    }
    if (rc == 0) *state = rank_to_state(*rank, B);
    return rc;
}

uint64_t qbx_emu_state_to_rank(uint64_t s, int B) { return state_to_rank(s, B); }
uint64_t qbx_emu_rank_to_state(uint64_t r, int B) { return rank_to_state(r, B); }

}  // extern "C"
