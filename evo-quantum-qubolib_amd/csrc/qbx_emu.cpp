// qbx_emu.cpp -- host build of the device lane walk (qbx_walk.h), for the CPU test suite.
//
// Not part of the product path: libqbx.so never loads it.  It runs the very same
// walk_lane_* templates the gfx950 kernels run, lane by lane on the CPU, with the same
// bit-ordered tables the prep kernel builds, so the CPU tests can check the walk's
// indexing (lane prefixes, block keys, tie ranks, transitions) against the oracle for
// every (n, L, B, M) shape before any GPU time is spent.
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <vector>
#define QBX_EMU_STATS 1
#include "qbx_walk.h"
#include "qbx_gen.h"

using namespace qbx;

namespace qbx {
unsigned long long emu_stat_flushes = 0, emu_stat_rewalks = 0;
}
static unsigned long long emu_stat_pair_lanes = 0;
static unsigned long long emu_stat_rp_lanes = 0, emu_stat_rp_int32_groups = 0;
static int emu_rp_slack_override = -1;
static unsigned long long emu_stat_qrp_lanes = 0, emu_stat_qrp_int32_groups = 0;
static int emu_qrp_slack_override = -1, emu_qrp_shift_delta = 0;

namespace {

template <int M, int OMAX>
int emu_i32(const int32_t* Q, int n, int B, int L, int64_t* E_out, uint64_t* rank_out) {
    if (L < M || L > M + OMAX || L > n) return -2;
    std::vector<int32_t> W((size_t)n * LDW, 0), d(n, 0), dl(1 << M, 0);
    for (int a = 0; a < n; ++a) {
        d[a] = d_elem<int32_t, int32_t>(Q, n, a);
        for (int b = 0; b < LDW; ++b) W[(size_t)a * LDW + b] = w_elem<int32_t, int32_t>(Q, n, a, b);
    }
    for (uint32_t t = 0; t < (1u << M); ++t) dl[t] = trie_elem<int32_t, int32_t>(Q, n, M, t);
    // the packed tables and the bound that selects the pair walk, as qbx_prep builds them
    std::vector<uint32_t> Wp(W.size()), dlp(dl.size());
    for (size_t i = 0; i < W.size(); ++i) Wp[i] = pack_both((uint32_t)W[i]);
    for (size_t i = 0; i < dl.size(); ++i) dlp[i] = pack_both((uint32_t)dl[i]);
    long long bound = 0;
    for (int a = 0; a < n; ++a) {
        bound += d[a] < 0 ? -(long long)d[a] : d[a];
        for (int b = a + 1; b < n; ++b) {
            const long long w = W[(size_t)a * LDW + b];
            bound += w < 0 ? -w : w;
        }
    }
    const bool pair = bound <= PAIR_BOUND && L > M;
    // the rebased pair walk runs wherever its guard constant leaves room (the device also
    // asks rp_worth_it; here every eligible problem is checked), with the slack overridable
    // so that tests can force the int32 groups
    const int64_t D = L > M ? rp_guard_D(W.data(), M, L) : PAIR_BOUND + 1;
    const bool rp = L > M && D < PAIR_BOUND;
    const int32_t slack = emu_rp_slack_override >= 0 ? emu_rp_slack_override : (int32_t)(PAIR_BOUND - D);
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
        // the packed pair walk must give every lane the same key
        if (pair) {
            const int64_t pk = lane_best_block_pair<M, OMAX>(W.data(), Wp.data(), d.data(), n, L, B, lam,
                                                             dlp.data());
            if (pk != key) return -5;
            ++emu_stat_pair_lanes;
        }
        if (rp) {
            uint32_t fallbacks = 0;
            const int64_t rk = lane_best_block_rp<M, OMAX>(W.data(), Wp.data(), d.data(), n, L, B, lam, dlp.data(),
                                                           slack, &fallbacks);
            if (rk != key) return -6;
            ++emu_stat_rp_lanes;
            emu_stat_rp_int32_groups += fallbacks;
        }
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
    // the tables qbx_prep builds: FP64 W, d, delta, and their quantized int32 copies
    std::vector<double> W((size_t)n * LDW, 0.0), d(64, 0.0), dld(1 << M, 0.0);
    std::vector<int32_t> Wq((size_t)n * LDW, 0), dq(64, 0), dlq(1 << M, 0);
    double bound = 0.0;
    for (int a = 0; a < n; ++a) {
        d[a] = d_elem<double, double>(Q, n, a);
        bound += fabs(d[a]);
        for (int b = 0; b < LDW; ++b) {
            W[(size_t)a * LDW + b] = w_elem<double, double>(Q, n, a, b);
            if (b > a && b < n) bound += fabs(W[(size_t)a * LDW + b]);
        }
    }
    if (!(bound <= 1.0e300)) return -4;
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
    if (qbound > 0x7ffffffeLL) return -4;
    for (uint32_t t = 0; t < (1u << M); ++t) {
        dld[t] = delta_elem<double, double>(Q, n, M, t);
        dlq[t] = (int32_t)trie_from_table<int32_t, long long>(Wq.data(), M, t);
    }
    const double qm = q_margin(n);
    const int H = n - L;
    const uint64_t lanes = 1ull << H;
    double bE = __builtin_inf();
    uint64_t bR = ~0ull;
    // one bound shared by all lanes, as the device's waves share the problem's key
    HostBound hb{__builtin_inf()};
    uint32_t buf_c[QCAP];
    int32_t buf_m[QCAP];
    for (uint64_t lam = 0; lam < lanes; ++lam) {
        LaneOutF64 o = walk_lane_q32<M, OMAX>(Wq.data(), dq.data(), W.data(), d.data(), n, L, B, lam, dlq.data(),
                                              dld.data(), scale, qm, hb, buf_c, buf_m, 1);
        if (o.E < bE || (o.E == bE && o.rank < bR)) { bE = o.E; bR = o.rank; }
    }
    *E_out = bE;
    *rank_out = bR;
    // the packed filter, where qbx_prep would select it (shift and slack overridable so
    // that tests can force fine scales and int32 groups), must give the same answer
    int64_t D = 0;
    int sh = qrp_shift(Wq.data(), dq.data(), n, M, L, &D);
    if (sh >= 0 && emu_qrp_shift_delta != 0) {
        sh = sh + emu_qrp_shift_delta < 0 ? 0 : (sh + emu_qrp_shift_delta > 30 ? 30 : sh + emu_qrp_shift_delta);
        D = qrp_guard_D(Wq.data(), M, L, sh);
        if (D < 0) sh = -1;
    }
    if (sh >= 0) {
        int32_t slack = (int32_t)(PAIR_BOUND - D);
        if (emu_qrp_slack_override >= 0 && emu_qrp_slack_override < slack) slack = emu_qrp_slack_override;
        std::vector<uint32_t> Wpc(Wq.size()), dlpc(dlq.size());
        for (size_t i = 0; i < Wq.size(); ++i) Wpc[i] = pack_both((uint32_t)rnd_shr(Wq[i], sh));
        for (size_t i = 0; i < dlq.size(); ++i) dlpc[i] = pack_both((uint32_t)rnd_shr(dlq[i], sh));
        double pE = __builtin_inf();
        uint64_t pR = ~0ull;
        HostBound hp{__builtin_inf()};
        for (uint64_t lam = 0; lam < lanes; ++lam) {
            uint32_t fallbacks = 0;
            LaneOutF64 o = walk_lane_qrp<M, OMAX>(Wq.data(), dq.data(), Wpc.data(), W.data(), d.data(), n, L, B,
                                                  lam, dlpc.data(), dlq.data(), dld.data(), scale, qm, sh, slack, hp,
                                                  buf_c, buf_m, 1, &fallbacks);
            if (o.E < pE || (o.E == pE && o.rank < pR)) { pE = o.E; pR = o.rank; }
            ++emu_stat_qrp_lanes;
            emu_stat_qrp_int32_groups += fallbacks;
        }
        if (!(pE == bE && pR == bR)) return -7;
    }
    return 0;
}

}  // namespace

extern "C" {

// Returns 0 on success; -1 unsupported M; -2 unsupported L; -3 the two i32 flows disagree;
// -4 float energies out of range; -5 the packed pair walk's lane key differs from the
// int32 walk's (checked for every lane when the problem's bound selects the pair walk);
// -6 the rebased pair walk's lane key differs (checked for every lane when its guard
// constant D < 15359); -7 the packed float filter's answer differs from the int32
// filter's (checked whenever qbx_prep would select it).
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
        case 7: rc = emu_f64<7, 13>(Q, n, B, L, E, rank); break;
        default: return -1;
    }
    if (rc == 0) *state = rank_to_state(*rank, B);
    return rc;
}

// RandomQUBO matrix from a PCG64 state (state hi, lo, increment hi, lo), as the device
// generator draws it: int32 elements for integer dtypes, double for float dtypes.
int qbx_emu_random_qubo(const uint64_t* st, int n, int64_t low, int64_t high, int dtype, void* out) {
    if (!gen_valid(low, high, dtype)) return -1;
    Pcg64 g;
    g.hi = st[0]; g.lo = st[1]; g.inc_hi = st[2]; g.inc_lo = st[3];
    g.buf32 = 0; g.buf16 = 0; g.has32 = g.has16 = 0;
    if (dtype == GEN_INT16 || dtype == GEN_INT32) gen_random_qubo<int32_t>(g, n, low, high, dtype, (int32_t*)out);
    else gen_random_qubo<double>(g, n, low, high, dtype, (double*)out);
    return 0;
}

// TSP QUBO from a PCG64 state, as the device generator builds it: n_cities in 1..7,
// dtype GEN_FLOAT32 or GEN_FLOAT64 (the cities' type), out: (n_cities^2)^2 doubles.
int qbx_emu_random_tsp(const uint64_t* st, int nc, double low, double high, int dtype, double* out) {
    if (nc < 1 || nc > TSP_MAX_CITIES || (dtype != GEN_FLOAT32 && dtype != GEN_FLOAT64)) return -1;
    Pcg64 g;
    g.hi = st[0]; g.lo = st[1]; g.inc_hi = st[2]; g.inc_lo = st[3];
    g.buf32 = 0; g.buf16 = 0; g.has32 = g.has16 = 0;
    double d[TSP_MAX_PAIRS];
    if (dtype == GEN_FLOAT32) {
        tsp_distances<float>(g, nc, low, high, d);
        tsp_qubo<float>(nc, d, out);
    } else {
        tsp_distances<double>(g, nc, low, high, d);
        tsp_qubo<double>(nc, d, out);
    }
    return 0;
}

// float filter path counters since the last call (buffer overflows, FP64 block re-walks)
void qbx_emu_take_stats(uint64_t* flushes, uint64_t* rewalks) {
    *flushes = emu_stat_flushes;
    *rewalks = emu_stat_rewalks;
    emu_stat_flushes = emu_stat_rewalks = 0;
}

// packed float filter: lanes walked and groups walked in int32 since the last call
void qbx_emu_take_qrp_stats(uint64_t* lanes, uint64_t* int32_groups) {
    *lanes = emu_stat_qrp_lanes;
    *int32_groups = emu_stat_qrp_int32_groups;
    emu_stat_qrp_lanes = emu_stat_qrp_int32_groups = 0;
}

// packed float filter overrides: a smaller guard slack (-1: PAIR_BOUND - D_c, as the
// device uses) and a change of the chosen coarse shift (0: as the device chooses; a finer
// scale fails the guard more often, a coarser one passes more candidates)
void qbx_emu_set_qrp(int32_t slack, int32_t shift_delta) {
    emu_qrp_slack_override = slack;
    emu_qrp_shift_delta = shift_delta;
}

// lanes whose pair-walk key was checked since the last call
uint64_t qbx_emu_take_pair_lanes(void) {
    const uint64_t v = emu_stat_pair_lanes;
    emu_stat_pair_lanes = 0;
    return v;
}

// rebased pair walk: lanes checked and groups walked in int32 since the last call
void qbx_emu_take_rp_stats(uint64_t* lanes, uint64_t* int32_groups) {
    *lanes = emu_stat_rp_lanes;
    *int32_groups = emu_stat_rp_int32_groups;
    emu_stat_rp_lanes = emu_stat_rp_int32_groups = 0;
}

// guard slack used by the rebased pair walk (-1: PAIR_BOUND - D, as the device uses)
void qbx_emu_set_rp_slack(int32_t slack) { emu_rp_slack_override = slack; }

uint64_t qbx_emu_state_to_rank(uint64_t s, int B) { return state_to_rank(s, B); }
uint64_t qbx_emu_rank_to_state(uint64_t r, int B) { return rank_to_state(r, B); }

}  // extern "C"
