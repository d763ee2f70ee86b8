// qbx_walk.h -- the per-lane exhaustive walk shared by the gfx950 kernels and by the
// host-compiled emulation used in the CPU test suite.
//
// Problem: minimise E(x) = x^T Q x over x in {0,1}^n (src/qubo_lib/solvers/qbf.py:22,
// qbf.py:264).  The state integer s has x[0] as its most significant bit
// (qbf.py:373-393), so bit position a of s holds variable v = n-1-a.  In bit order:
//     d[a]    = Q[v][v]
//     W[a][b] = Q[va][vb] + Q[vb][va]   (a != b), W[a][a] = 0
//     E(s)    = sum_{a set} d[a] + sum_{a<b set} W[a][b]
//
// Tie rule.  The reference QBF walks reference prefixes P = s >> B in Gray order and
// keeps the strictly-first minimum (qbf.py:212-230, 395-412); inside a prefix it takes
// numpy's argmin, the lowest suffix (qbf.py:229).  Its answer is therefore the
// lexicographic minimum of (E, rank(s)) with
//     rank(s) = (invgray(s >> B) << B) | (s & (2^B - 1)),   B = min(n, 20) by default
// (qbf.py:443-446).  Every kernel here returns exactly that minimum.
//
// Decomposition of one solve (n bits):
//     [ lane bits: H ][ outer walked bits: O ][ inner bits: M ]      L = O + M = n - H
// A wave owns 64 consecutive lane ranks lambda; lane prefix p is the top-H-bit state
// whose ranks start at lambda << L, so lane ranks are ordered like states.  The lane
// walks its 2^L states as 2^O blocks of 2^M states:
//   * inner block: a compile-time unrolled Gray walk over the M low bits.  Each step is
//       E += (+-f_k) + delta_t      (one v_add3_u32; f_k / -f_k per lane, delta_t uniform)
//     and block minima are taken two states at a time (v_min3_i32): ~1.5 VALU ops/state.
//   * block transition: the outer bits advance in Gray order; flipping outer bit J
//     changes E0 (energy of the block base) by +-(u_J + W[J][J-1]) because, at that
//     moment, the only walked bit below J that is set is J-1.  Fields f_k of the inner
//     bits and "upper fields" u_q of the outer bits below J move by +-W[.][J].
//   * a block is re-walked (to find its lowest-rank minimiser) only when its minimum
//     beats the lane's best in (E, rank) order -- rare after the first few blocks.
//   * ordering: block_key() gives rank bits [M, L) of a block's states and rank_lo() the
//     inner bits [0, M), so (E, lane, key, rank_lo) orders states as (E, rank) does.
#pragma once
#include <stdint.h>
#include <utility>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define QBX_HD __host__ __device__ __forceinline__
#define QBX_HD_MEMBER __host__ __device__ __forceinline__
#else
#define QBX_HD static inline
#define QBX_HD_MEMBER inline
#endif

namespace qbx {

// ---------------------------------------------------------------------------- bits
QBX_HD uint64_t shr64(uint64_t v, int b) { return b >= 64 ? 0ull : (v >> b); }
QBX_HD uint64_t shl64(uint64_t v, int b) { return b >= 64 ? 0ull : (v << b); }
QBX_HD uint64_t lowmask64(int b) { return b >= 64 ? ~0ull : ((1ull << b) - 1ull); }
QBX_HD uint32_t lowmask32(int b) { return b >= 32 ? 0xffffffffu : ((1u << b) - 1u); }
QBX_HD uint64_t gray64(uint64_t v) { return v ^ (v >> 1); }
QBX_HD uint64_t invgray64(uint64_t g) {
    g ^= g >> 1; g ^= g >> 2; g ^= g >> 4; g ^= g >> 8; g ^= g >> 16; g ^= g >> 32;
    return g;
}
QBX_HD uint32_t invgray32(uint32_t g) {
    g ^= g >> 1; g ^= g >> 2; g ^= g >> 4; g ^= g >> 8; g ^= g >> 16;
    return g;
}
QBX_HD int parity64(uint64_t v) { return __builtin_popcountll(v) & 1; }
QBX_HD int parity32(uint32_t v) { return __builtin_popcount(v) & 1; }
QBX_HD int ctz32(uint32_t v) { return __builtin_ctz(v); }

// rank(s) of the reference tie rule and its inverse (both are bijections on n bits).
QBX_HD uint64_t state_to_rank(uint64_t s, int B) {
    return shl64(invgray64(shr64(s, B)), B) | (s & lowmask64(B));
}
QBX_HD uint64_t rank_to_state(uint64_t r, int B) {
    return shl64(gray64(shr64(r, B)), B) | (r & lowmask64(B));
}
// top H bits (lane prefix) of the states whose ranks start with lambda.
QBX_HD uint64_t lane_prefix(uint64_t lam, int L, int B) {
    return shr64(rank_to_state(shl64(lam, L), B), L);
}

// rank of inner configuration y (M bits) inside its block.  When B >= M the inner bits
// are plain suffix bits; otherwise the top M-B of them belong to the reference prefix and
// are ranked by inverse Gray code, complemented when the parity of every state bit above
// the block (par) is odd.
template <int M>
QBX_HD uint32_t rank_lo(uint32_t y, int B, int par) {
    if (B >= M) return y;
    const uint32_t hi = invgray32(y >> B) ^ (par ? lowmask32(M - B) : 0u);
    return (hi << B) | (y & lowmask32(B));
}

// block key: bits [M, L) of the rank of every state in block c of a lane.
//   SBk  = min(max(B - M, 0), O): number of key bits that are suffix bits (binary order)
//   pmaskS = (parity(lane prefix) ? lowmask(O - SBk) : 0) << SBk
QBX_HD uint32_t block_key(uint32_t c, int SBk, uint32_t pmaskS) {
    const uint32_t gc = c ^ (c >> 1);
    return (((c >> SBk) << SBk) | (gc & lowmask32(SBk))) ^ pmaskS;
}

// ------------------------------------------------------------ per-instance tables
// Leading dimension of the bit-ordered W table: rows are zero padded to 64 entries so
// that unrolled code may read W[J][0..63] for any J < n without bounds checks.
constexpr int LDW = 64;

// W[a][b] in bit order from the row-major, variable-ordered source Q (int32 or double).
template <class T, class Acc>
QBX_HD Acc w_elem(const T* Q, int n, int a, int b) {
    if (a == b || a >= n || b >= n) return Acc(0);
    const int i = n - 1 - a, j = n - 1 - b;
    return Acc(Q[(int64_t)i * n + j]) + Acc(Q[(int64_t)j * n + i]);
}
template <class T, class Acc>
QBX_HD Acc d_elem(const T* Q, int n, int a) {
    const int i = n - 1 - a;
    return Acc(Q[(int64_t)i * n + i]);
}
// inner-step constant delta_t (t = 1 .. 2^M - 1): the pairwise part of the energy change
// when step t of the inner Gray walk flips inner bit k = ctz(t).
template <class T, class Acc>
QBX_HD Acc delta_elem(const T* Q, int n, int M, uint32_t t) {
    if (t == 0 || t >= (1u << M)) return Acc(0);
    const int k = ctz32(t);
    const uint32_t g = t ^ (t >> 1);
    const uint32_t y = (t - 1) ^ ((t - 1) >> 1);          // inner state before the flip
    const bool up = ((g >> k) & 1u) != 0u;
    Acc s = Acc(0);
    for (int l = 0; l < M; ++l)
        if (l != k && ((y >> l) & 1u)) s += w_elem<T, Acc>(Q, n, k, l);
    return up ? s : Acc(0) - s;
}

// ------------------------------------------------------------------ int32 lane walk
template <int M, int OMAX>
struct LaneI32 {
    int32_t E0;          // energy of the block base (inner bits = 0)
    int32_t f[M];        // field of inner bit k from every non-inner bit
    int32_t nf[M];       // -f[k]
    int32_t u[OMAX];     // upper field of outer bit M+q: d + sum over set bits above it
};

struct LaneOutI32 {
    int32_t E;
    uint64_t rank;
};

// mn = min(min(mn, a), b): associated so that the backend folds each pair of states into
// one v_min3 (min(mn, min(a, b)) is regrouped into a tree that keeps every block energy
// live and costs 0.75 ops per state).
QBX_HD int32_t min3_i32(int32_t mn, int32_t a, int32_t b) {
    const int32_t m = mn < a ? mn : a;
    return m < b ? m : b;
}
QBX_HD float min3_f32(float mn, float a, float b) {
    return __builtin_fminf(__builtin_fminf(mn, a), b);
}
// Empty volatile asm: makes a value opaque to the optimiser (no instruction is emitted).
// Used only on the rare re-walk path, so that its energies are not merged with the
// first walk's (which would keep all 2^M of them live across the block loop).
#if defined(__HIP_DEVICE_COMPILE__)
#define QBX_OPAQUE(x) asm volatile("" : "+v"(x))
#else
#define QBX_OPAQUE(x) ((void)0)
#endif

// Minimum energy over one inner block.
template <int M>
QBX_HD int32_t block_min_i32(int32_t E0, const int32_t* f, const int32_t* nf, const int32_t* dl) {
    uint32_t E = (uint32_t)E0;
    int32_t mn = E0;
    int32_t odd = E0;
#pragma unroll
    for (int t = 1; t < (1 << M); ++t) {
        const int k = __builtin_ctz((unsigned)t);
        const bool up = ((((unsigned)t ^ ((unsigned)t >> 1)) >> k) & 1u) != 0u;
        E = E + (uint32_t)(up ? f[k] : nf[k]) + (uint32_t)dl[t];
        if (t & 1) odd = (int32_t)E;
        else mn = min3_i32(mn, odd, (int32_t)E);
    }
    return min3_i32(mn, odd, odd);   // t = 2^M - 1 is odd and has no partner
}

// Lowest inner rank among the states of the block whose energy equals mn.  The walk only
// records which steps hit mn (a 2^M-bit mask); ranks are computed afterwards for the set
// bits alone, so no per-step rank arithmetic exists for the compiler to hoist.
template <int M>
QBX_HD uint32_t block_argmin_i32(int32_t E0, const int32_t* f, const int32_t* nf, const int32_t* dl,
                                 int32_t mn, int B, int par) {
    int32_t fo[M], nfo[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        fo[k] = f[k];
        nfo[k] = nf[k];
        QBX_OPAQUE(fo[k]);
        QBX_OPAQUE(nfo[k]);
    }
    QBX_OPAQUE(E0);
    constexpr int NW = ((1 << M) + 31) / 32;
    uint32_t mask[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) mask[w] = 0u;
    uint32_t E = (uint32_t)E0;
    if (E0 == mn) mask[0] = 1u;
#pragma unroll
    for (int t = 1; t < (1 << M); ++t) {
        const int k = __builtin_ctz((unsigned)t);
        const bool up = ((((unsigned)t ^ ((unsigned)t >> 1)) >> k) & 1u) != 0u;
        E = E + (uint32_t)(up ? fo[k] : nfo[k]) + (uint32_t)dl[t];
        if ((int32_t)E == mn) mask[t >> 5] |= 1u << (t & 31);
    }
    uint32_t best = 0xffffffffu;
    for (int w = 0; w < NW; ++w) {
        uint32_t m = mask[w];
        while (m) {
            const uint32_t t = (uint32_t)(w * 32 + __builtin_ctz(m));
            m &= m - 1u;
            const uint32_t r = rank_lo<M>(t ^ (t >> 1), B, par);
            best = r < best ? r : best;
        }
    }
    return best;
}

// Outer transition flipping outer bit I (position J = M + I) with I known at compile time.
// `up` is a compile-time constant for I < 3 inside an unrolled group and a wave-uniform
// value for I == 3, so either way no divergent control flow is created.
template <int M, int OMAX, int I>
QBX_HD void trans_static_i32(LaneI32<M, OMAX>& s, const int32_t* __restrict__ WJ, bool up) {
    constexpr int J = M + I;
    const uint32_t field = (uint32_t)s.u[I] + (I > 0 ? (uint32_t)WJ[J - 1] : 0u);
    if (up) {
        s.E0 = (int32_t)((uint32_t)s.E0 + field);
#pragma unroll
        for (int k = 0; k < M; ++k) { s.f[k] += WJ[k]; s.nf[k] -= WJ[k]; }
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] += WJ[M + q];
    } else {
        s.E0 = (int32_t)((uint32_t)s.E0 - field);
#pragma unroll
        for (int k = 0; k < M; ++k) { s.f[k] -= WJ[k]; s.nf[k] += WJ[k]; }
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] -= WJ[M + q];
    }
}

// Same for a run-time bit index i >= GROUP_BITS (taken once per group of 2^GROUP_BITS blocks).
template <int M, int OMAX>
QBX_HD void trans_dyn_i32(LaneI32<M, OMAX>& s, const int32_t* __restrict__ WJ, int i, bool up) {
    const int J = M + i;
    uint32_t uJ = 0;
#pragma unroll
    for (int q = 0; q < OMAX; ++q) if (q == i) uJ = (uint32_t)s.u[q];
    const uint32_t field = uJ + (uint32_t)WJ[J - 1];   // i >= 1 here
    const int32_t sg = up ? 1 : -1;
    s.E0 = (int32_t)((uint32_t)s.E0 + (uint32_t)sg * field);
#pragma unroll
    for (int k = 0; k < M; ++k) { const int32_t w = sg * WJ[k]; s.f[k] += w; s.nf[k] -= w; }
#pragma unroll
    for (int q = 0; q < OMAX; ++q) if (q < i) s.u[q] += sg * WJ[M + q];
}

constexpr int GROUP_BITS = 4;   // blocks are walked in unrolled groups of 16

// E0 and inner fields of an arbitrary block, from scratch: hs = the state's bits above the
// inner block (bit j of hs is state bit M + j).  Used once per lane, for its best block.
template <int M>
QBX_HD void block_state_i32(const int32_t* __restrict__ W, const int32_t* __restrict__ d, int n,
                            uint64_t hs, int32_t& E0, int32_t* f) {
    uint32_t E = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) f[k] = d[k];
    for (int a = M; a < n; ++a) {
        if (!((hs >> (a - M)) & 1ull)) continue;
        const int32_t* Wa = W + a * LDW;
        E += (uint32_t)d[a];
        for (int b = a + 1; b < n; ++b)
            if ((hs >> (b - M)) & 1ull) E += (uint32_t)Wa[b];
#pragma unroll
        for (int k = 0; k < M; ++k) f[k] += Wa[k];
    }
    E0 = (int32_t)E;
}

// One block of an unrolled group: the transition into block c = c0 + J (J > 0; its outer
// bit index ctz(J) is a compile-time constant), the block's minimum, and the best-key update.
// Returns false when c is past the lane's last block.
template <int M, int OMAX, int J>
QBX_HD bool group_step_i32(LaneI32<M, OMAX>& s, const int32_t* __restrict__ W, const int32_t* dl,
                           uint32_t c0, uint32_t nblk, int SBk, uint32_t pmaskS, int64_t& best) {
    const uint32_t c = c0 + (uint32_t)J;
    if (J > 0) {
        if (c >= nblk) return false;
        constexpr int i = J > 0 ? __builtin_ctz((unsigned)J) : 0;
        // gray(c) bit i = c_i ^ c_{i+1}, and c_i = 1 at the flip
        const bool up = (i + 1 < GROUP_BITS) ? ((((unsigned)J >> (i + 1)) & 1u) == 0u)
                                             : (((c0 >> GROUP_BITS) & 1u) == 0u);
        trans_static_i32<M, OMAX, i>(s, W + (M + i) * LDW, up);
    }
    const int32_t mn = block_min_i32<M>(s.E0, s.f, s.nf, dl);
    const uint32_t K = block_key(c, SBk, pmaskS);
    const int64_t key = (int64_t)(((uint64_t)(uint32_t)mn << 32) | (uint64_t)K);
    best = key < best ? key : best;
    return true;
}

template <int M, int OMAX, int... Js>
QBX_HD void group_i32(LaneI32<M, OMAX>& s, const int32_t* __restrict__ W, const int32_t* dl,
                      uint32_t c0, uint32_t nblk, int SBk, uint32_t pmaskS, int64_t& best,
                      std::integer_sequence<int, Js...>) {
    (void)(group_step_i32<M, OMAX, Js>(s, W, dl, c0, nblk, SBk, pmaskS, best) && ...);
}

// Lane set-up: E0 and the fields of the lane's first block.  The lane prefix p splits into
// its low 6 bits (per lane) and the bits above them, which depend only on lam >> 6 and so
// are the same for the 64 lanes of a wave: their contributions are accumulated once, from
// wave-uniform values (scalar registers on the device), and only the low bits cost
// per-lane work.
template <int M, int OMAX>
QBX_HD void lane_setup_i32(LaneI32<M, OMAX>& s, const int32_t* __restrict__ W, const int32_t* __restrict__ d,
                           int n, int L, int B, uint64_t lam) {
    const int H = n - L;
    const int O = L - M;
    const int Hlo = H < 6 ? H : 6;
    const uint64_t p = lane_prefix(lam, L, B);
    const uint64_t p_hi = shr64(lane_prefix(lam & ~63ull, L, B), 6);   // wave-uniform
    const uint32_t p_lo = (uint32_t)(p & 63ull);

    // contributions of the set high lane bits (uniform)
    uint32_t Ec = 0;
    int32_t fc[M], uc[OMAX], lc[6];
#pragma unroll
    for (int k = 0; k < M; ++k) fc[k] = d[k];
#pragma unroll
    for (int q = 0; q < OMAX; ++q) uc[q] = (q < O) ? d[M + q] : 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) lc[i] = (i < Hlo) ? d[L + i] : 0;
    for (int j = Hlo; j < H; ++j) {
        if (!((p_hi >> (j - Hlo)) & 1ull)) continue;
        const int a = L + j;
        const int32_t* Wa = W + a * LDW;
        uint32_t e = (uint32_t)d[a];
        for (int j2 = j + 1; j2 < H; ++j2)
            if ((p_hi >> (j2 - Hlo)) & 1ull) e += (uint32_t)Wa[L + j2];
        Ec += e;
#pragma unroll
        for (int k = 0; k < M; ++k) fc[k] += Wa[k];
#pragma unroll
        for (int q = 0; q < OMAX; ++q) uc[q] += Wa[M + q];
#pragma unroll
        for (int i = 0; i < 6; ++i) lc[i] += (i < Hlo) ? Wa[L + i] : 0;
    }
    // per-lane part: the low lane bits (lc[i] already holds d + the high-bit couplings)
    uint32_t E0 = Ec;
#pragma unroll
    for (int k = 0; k < M; ++k) s.f[k] = fc[k];
#pragma unroll
    for (int q = 0; q < OMAX; ++q) s.u[q] = uc[q];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        if (i >= Hlo) break;
        const uint32_t mi = 0u - ((p_lo >> i) & 1u);
        const int a = L + i;
        const int32_t* Wa = W + a * LDW;
        uint32_t e = (uint32_t)lc[i];
#pragma unroll
        for (int i2 = i + 1; i2 < 6; ++i2)
            if (i2 < Hlo) e += (uint32_t)Wa[L + i2] & (0u - ((p_lo >> i2) & 1u));
        E0 += e & mi;
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] += (int32_t)((uint32_t)Wa[k] & mi);
#pragma unroll
        for (int q = 0; q < OMAX; ++q) s.u[q] += (int32_t)((uint32_t)Wa[M + q] & mi);
    }
    s.E0 = (int32_t)E0;
#pragma unroll
    for (int k = 0; k < M; ++k) s.nf[k] = -s.f[k];
}

// Lane key: (E, K) packed as one signed 64-bit value, E in the high word.  Ordering lane
// keys orders the lane's blocks by (block minimum, block rank).
QBX_HD int32_t key_E(int64_t key) { return (int32_t)(key >> 32); }
QBX_HD uint32_t key_K(int64_t key) { return (uint32_t)key; }
constexpr int64_t KEY_NONE = 0x7fffffffffffffffll;

// Walk one lane's 2^L states and return the key of its best block.
// W: bit-ordered, LDW columns; d: n; dl: delta_t for t < 2^M.
template <int M, int OMAX>
QBX_HD int64_t lane_best_block_i32(const int32_t* __restrict__ W, const int32_t* __restrict__ d,
                                   int n, int L, int B, uint64_t lam, const int32_t* dl) {
    static_assert(OMAX >= GROUP_BITS, "OMAX must cover the unrolled transition bits");
    const int O = L - M;
    const uint64_t p = lane_prefix(lam, L, B);
    LaneI32<M, OMAX> s;
    lane_setup_i32<M, OMAX>(s, W, d, n, L, B, lam);

    const int SBk = (B - M) < 0 ? 0 : ((B - M) > O ? O : (B - M));
    const uint32_t pmaskS = (parity64(p) && O > SBk) ? (lowmask32(O - SBk) << SBk) : 0u;
    int64_t best = KEY_NONE;
    const uint32_t nblk = 1u << O;
    for (uint32_t c0 = 0;;) {
        group_i32<M, OMAX>(s, W, dl, c0, nblk, SBk, pmaskS, best,
                           std::make_integer_sequence<int, (1 << GROUP_BITS)>{});
        c0 += 1u << GROUP_BITS;
        if (c0 >= nblk) break;
        const int i = ctz32(c0);
        const bool up = (((c0 ^ (c0 >> 1)) >> i) & 1u) != 0u;
        trans_dyn_i32<M, OMAX>(s, W + (M + i) * LDW, i, up);
    }
    return best;
}

// Lowest inner rank of the minimisers inside block K of lane lam (whose minimum is E):
// rebuild the block's state from scratch and re-walk it once.
template <int M>
QBX_HD uint32_t block_rank_lo_i32(const int32_t* __restrict__ W, const int32_t* __restrict__ d,
                                  int n, int L, int B, uint64_t lam, int32_t E, uint32_t K,
                                  const int32_t* dl) {
    const int O = L - M;
    const uint64_t p = lane_prefix(lam, L, B);
    const int SBk = (B - M) < 0 ? 0 : ((B - M) > O ? O : (B - M));
    const int parp = parity64(p);
    const uint32_t pmaskS = (parp && O > SBk) ? (lowmask32(O - SBk) << SBk) : 0u;
    // invert K = ((c >> SBk) << SBk) | (gray(c) & lowmask(SBk)), then XOR pmaskS
    const uint32_t k = K ^ pmaskS;
    uint32_t c = (k >> SBk) << SBk;
    for (int b = SBk - 1; b >= 0; --b) c |= (((k >> b) & 1u) ^ ((c >> (b + 1)) & 1u)) << b;
    const uint32_t gc = c ^ (c >> 1);
    const uint64_t hs = (p << O) | (uint64_t)gc;
    int32_t bE0;
    int32_t bf[M], bnf[M];
    block_state_i32<M>(W, d, n, hs, bE0, bf);
#pragma unroll
    for (int kk = 0; kk < M; ++kk) bnf[kk] = -bf[kk];
    return block_argmin_i32<M>(bE0, bf, bnf, dl, E, B, parp ^ parity32(gc));
}

// Rank of the best state of block K of lane lam.
QBX_HD uint64_t block_rank_hi(uint64_t lam, int L, int M, uint32_t K) {
    return shl64(lam, L) | ((uint64_t)K << M);
}

// Whole-lane answer (used by the host build): best block, then its re-walk.
template <int M, int OMAX>
QBX_HD LaneOutI32 walk_lane_i32(const int32_t* __restrict__ W, const int32_t* __restrict__ d,
                                int n, int L, int B, uint64_t lam, const int32_t* dl) {
    const int64_t key = lane_best_block_i32<M, OMAX>(W, d, n, L, B, lam, dl);
    LaneOutI32 out;
    out.E = key_E(key);
    out.rank = block_rank_hi(lam, L, M, key_K(key)) |
               block_rank_lo_i32<M>(W, d, n, L, B, lam, key_E(key), key_K(key), dl);
    return out;
}

// ------------------------------------------------------- float-filter / FP64 walk
// Slow state in FP64, block walk in FP32 as a filter; any block whose FP32 minimum is
// within `margin` of the lane's FP64 best is re-walked in FP64 and compared exactly.
template <int M, int OMAX>
struct LaneF64 {
    double E0;
    double f[M];
    double u[OMAX];
};

struct LaneOutF64 {
    double E;
    uint64_t rank;
};

template <int M>
QBX_HD float block_min_f32(float E0, const float* f, const float* dl) {
    float E = E0, mn = E0, odd = E0;
#pragma unroll
    for (int t = 1; t < (1 << M); ++t) {
        const int k = __builtin_ctz((unsigned)t);
        const bool up = ((((unsigned)t ^ ((unsigned)t >> 1)) >> k) & 1u) != 0u;
        E = up ? (E + f[k]) : (E - f[k]);
        E = E + dl[t];
        if (t & 1) odd = E;
        else mn = min3_f32(mn, odd, E);
    }
    return min3_f32(mn, odd, odd);
}

template <int M, int OMAX, int I>
QBX_HD void trans_static_f64(LaneF64<M, OMAX>& s, const double* __restrict__ WJ, bool up) {
    constexpr int J = M + I;
    const double field = s.u[I] + (I > 0 ? WJ[J - 1] : 0.0);
    if (up) {
        s.E0 += field;
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] += WJ[k];
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] += WJ[M + q];
    } else {
        s.E0 -= field;
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] -= WJ[k];
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] -= WJ[M + q];
    }
}

template <int M, int OMAX>
QBX_HD void trans_dyn_f64(LaneF64<M, OMAX>& s, const double* __restrict__ WJ, int i, bool up) {
    const int J = M + i;
    double uJ = 0.0;
#pragma unroll
    for (int q = 0; q < OMAX; ++q) if (q == i) uJ = s.u[q];
    const double field = uJ + WJ[J - 1];
    if (up) {
        s.E0 += field;
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] += WJ[k];
#pragma unroll
        for (int q = 0; q < OMAX; ++q) if (q < i) s.u[q] += WJ[M + q];
    } else {
        s.E0 -= field;
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] -= WJ[k];
#pragma unroll
        for (int q = 0; q < OMAX; ++q) if (q < i) s.u[q] -= WJ[M + q];
    }
}

// f32 value >= x (x finite or +inf): x rounded to float, bumped up one ulp if it fell below.
QBX_HD float f32_next_up(float r) {
    union { float f; uint32_t u; } v;
    v.f = r;
    if (r != r || v.u == 0x7f800000u) return r;
    if (v.u == 0x80000000u) { v.u = 1u; return v.f; }
    if (v.u & 0x80000000u) v.u -= 1u; else v.u += 1u;
    return v.f;
}
QBX_HD float f32_at_least(double x) {
    float r = (float)x;
    if ((double)r < x) r = f32_next_up(r);
    return r;
}

// Lane set-up in FP64 (same split as lane_setup_i32: high lane bits are wave-uniform).
template <int M, int OMAX>
QBX_HD void lane_setup_f64(LaneF64<M, OMAX>& s, const double* __restrict__ W, const double* __restrict__ d,
                           int n, int L, int B, uint64_t lam) {
    const int H = n - L;
    const int O = L - M;
    const int Hlo = H < 6 ? H : 6;
    const uint64_t p = lane_prefix(lam, L, B);
    const uint64_t p_hi = shr64(lane_prefix(lam & ~63ull, L, B), 6);
    const uint32_t p_lo = (uint32_t)(p & 63ull);
    double Ec = 0.0, fc[M], uc[OMAX], lc[6];
#pragma unroll
    for (int k = 0; k < M; ++k) fc[k] = d[k];
#pragma unroll
    for (int q = 0; q < OMAX; ++q) uc[q] = (q < O) ? d[M + q] : 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) lc[i] = (i < Hlo) ? d[L + i] : 0.0;
    for (int j = Hlo; j < H; ++j) {
        if (!((p_hi >> (j - Hlo)) & 1ull)) continue;
        const int a = L + j;
        const double* Wa = W + a * LDW;
        double e = d[a];
        for (int j2 = j + 1; j2 < H; ++j2)
            if ((p_hi >> (j2 - Hlo)) & 1ull) e += Wa[L + j2];
        Ec += e;
#pragma unroll
        for (int k = 0; k < M; ++k) fc[k] += Wa[k];
#pragma unroll
        for (int q = 0; q < OMAX; ++q) uc[q] += Wa[M + q];
#pragma unroll
        for (int i = 0; i < 6; ++i) lc[i] += (i < Hlo) ? Wa[L + i] : 0.0;
    }
    double E0 = Ec;
#pragma unroll
    for (int k = 0; k < M; ++k) s.f[k] = fc[k];
#pragma unroll
    for (int q = 0; q < OMAX; ++q) s.u[q] = uc[q];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        if (i >= Hlo) break;
        if (!((p_lo >> i) & 1u)) continue;
        const int a = L + i;
        const double* Wa = W + a * LDW;
        double e = lc[i];
#pragma unroll
        for (int i2 = i + 1; i2 < 6; ++i2)
            if (i2 < Hlo && ((p_lo >> i2) & 1u)) e += Wa[L + i2];
        E0 += e;
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] += Wa[k];
#pragma unroll
        for (int q = 0; q < OMAX; ++q) s.u[q] += Wa[M + q];
    }
    s.E0 = E0;
}

// E0 and inner fields of an arbitrary block from scratch (hs: bits above the block).
template <int M>
QBX_HD void block_state_f64(const double* __restrict__ W, const double* __restrict__ d, int n,
                            uint64_t hs, double& E0, double* f) {
    double E = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) f[k] = d[k];
    for (int a = M; a < n; ++a) {
        if (!((hs >> (a - M)) & 1ull)) continue;
        const double* Wa = W + a * LDW;
        E += d[a];
        for (int b = a + 1; b < n; ++b)
            if ((hs >> (b - M)) & 1ull) E += Wa[b];
#pragma unroll
        for (int k = 0; k < M; ++k) f[k] += Wa[k];
    }
    E0 = E;
}

// Exact FP64 walk of one block; folds every state into the lane's (bestE, bestR).
template <int M>
QBX_HD void block_rewalk_f64(double E0, const double* f, const double* __restrict__ dld,
                             uint64_t rblk, int B, int par, double& bestE, uint64_t& bestR) {
    double E = E0;
    {
        const uint64_t r = rblk | rank_lo<M>(0u, B, par);
        if (E < bestE || (E == bestE && r < bestR)) { bestE = E; bestR = r; }
    }
    for (int t = 1; t < (1 << M); ++t) {
        const int k = __builtin_ctz((unsigned)t);
        const uint32_t g = (uint32_t)t ^ ((uint32_t)t >> 1);
        const bool up = ((g >> k) & 1u) != 0u;
        double fk = 0.0;
#pragma unroll
        for (int kk = 0; kk < M; ++kk) if (kk == k) fk = f[kk];
        E = up ? (E + fk) : (E - fk);
        E = E + dld[t];
        if (E <= bestE) {
            const uint64_t r = rblk | rank_lo<M>(g, B, par);
            if (E < bestE || r < bestR) { bestE = E; bestR = r; }
        }
    }
}

// One block of an unrolled float group: static transition, FP32 block minimum from the FP64
// slow state, candidate bit when the minimum is within the lane's threshold.
template <int M, int OMAX, int J>
QBX_HD bool group_step_f64(LaneF64<M, OMAX>& s, const double* __restrict__ W, const float* dls,
                           uint32_t c0, uint32_t nblk, float thr, uint32_t& cand) {
    const uint32_t c = c0 + (uint32_t)J;
    if (J > 0) {
        if (c >= nblk) return false;
        constexpr int i = J > 0 ? __builtin_ctz((unsigned)J) : 0;
        const bool up = (i + 1 < GROUP_BITS) ? ((((unsigned)J >> (i + 1)) & 1u) == 0u)
                                             : (((c0 >> GROUP_BITS) & 1u) == 0u);
        trans_static_f64<M, OMAX, i>(s, W + (M + i) * LDW, up);
    }
    float fs[M];
#pragma unroll
    for (int k = 0; k < M; ++k) fs[k] = (float)s.f[k];
    const float mn = block_min_f32<M>((float)s.E0, fs, dls);
    if (mn <= thr) cand |= 1u << J;
    return true;
}

template <int M, int OMAX, int... Js>
QBX_HD void group_f64(LaneF64<M, OMAX>& s, const double* __restrict__ W, const float* dls,
                      uint32_t c0, uint32_t nblk, float thr, uint32_t& cand,
                      std::integer_sequence<int, Js...>) {
    (void)(group_step_f64<M, OMAX, Js>(s, W, dls, c0, nblk, thr, cand) && ...);
}

// dls: FP32 inner deltas (registers); dld: FP64 inner deltas (memory, read on re-walks).
// Source of an upper bound on the answer: refresh() returns an exact energy already reached
// by some state of the problem (+inf if none); publish(e) offers the lane's exact best.
// Blocks whose FP32 minimum exceeds min(lane best, bound) + margin cannot hold the answer.
// The host build uses FixedBound; the device shares a running bound across waves.
struct FixedBound {
    double b;
    QBX_HD_MEMBER double refresh() const { return b; }
    QBX_HD_MEMBER void publish(double) {}
};

template <int M, int OMAX, class Bound>
QBX_HD LaneOutF64 walk_lane_f64(const double* __restrict__ W, const double* __restrict__ d,
                                int n, int L, int B, uint64_t lam, const float* dls,
                                const double* __restrict__ dld, double margin, Bound& src) {
    double bound = src.refresh();
    static_assert(OMAX >= GROUP_BITS, "OMAX must cover the unrolled transition bits");
    const int O = L - M;
    const uint64_t p = lane_prefix(lam, L, B);
    LaneF64<M, OMAX> s;
    lane_setup_f64<M, OMAX>(s, W, d, n, L, B, lam);

    const int SBk = (B - M) < 0 ? 0 : ((B - M) > O ? O : (B - M));
    const int parp = parity64(p);
    const uint32_t pmaskS = (parp && O > SBk) ? (lowmask32(O - SBk) << SBk) : 0u;
    const uint64_t laneRank = shl64(lam, L);

    // the lane's first block is walked exactly at once, so the filter starts tight
    double bestE = __builtin_inf();
    uint64_t bestR = ~0ull;
    block_rewalk_f64<M>(s.E0, s.f, dld, laneRank | ((uint64_t)block_key(0u, SBk, pmaskS) << M), B, parp,
                        bestE, bestR);
    float thr = f32_at_least((bestE < bound ? bestE : bound) + margin);

    const uint32_t nblk = 1u << O;
    for (uint32_t c0 = 0;;) {
        uint32_t cand = 0u;
        group_f64<M, OMAX>(s, W, dls, c0, nblk, thr, cand, std::make_integer_sequence<int, (1 << GROUP_BITS)>{});
        if (c0 == 0) cand &= ~1u;   // block 0 is already exact
        while (cand) {               // rare: rebuild each candidate block and walk it in FP64
            const uint32_t j = (uint32_t)__builtin_ctz(cand);
            cand &= cand - 1u;
            const uint32_t c = c0 + j;
            const uint32_t gc = c ^ (c >> 1);
            double bE0, bf[M];
            block_state_f64<M>(W, d, n, (p << O) | (uint64_t)gc, bE0, bf);
            block_rewalk_f64<M>(bE0, bf, dld, laneRank | ((uint64_t)block_key(c, SBk, pmaskS) << M), B,
                                parp ^ parity32(gc), bestE, bestR);
            thr = f32_at_least((bestE < bound ? bestE : bound) + margin);
        }
        // share progress with the rest of the problem: offer the lane's best, take the
        // problem's running bound (both are exact energies of real states)
        src.publish(bestE);
        const double b2 = src.refresh();
        bound = b2 < bound ? b2 : bound;
        thr = f32_at_least((bestE < bound ? bestE : bound) + margin);
        c0 += 1u << GROUP_BITS;
        if (c0 >= nblk) break;
        const int i = ctz32(c0);
        const bool up = (((c0 ^ (c0 >> 1)) >> i) & 1u) != 0u;
        trans_dyn_f64<M, OMAX>(s, W + (M + i) * LDW, i, up);
    }
    LaneOutF64 out;
    out.E = bestE;
    out.rank = bestR;
    return out;
}

// FP32 filter margin: every FP32 block value differs from its FP64 counterpart by at most
// u * bound * (1 + 6 * 2^M), u = 2^-24 (one rounding for E0, f_k and delta_t each, and
// two additions per step with partial sums bounded by 2 * bound).  8 * 2^M is used.
QBX_HD double filter_margin(double bound, int M) {
    return bound * 5.9604644775390625e-08 /* 2^-24 */ * 8.0 * (double)(1 << M);
}

// Order-preserving 64-bit keys for the per-instance atomic minimum.
QBX_HD uint64_t key_of_i64(int64_t E) { return (uint64_t)E ^ 0x8000000000000000ull; }
QBX_HD uint64_t key_of_f64(double E) {
    union { double d; uint64_t u; } v;
    v.d = E;
    return (v.u & 0x8000000000000000ull) ? ~v.u : (v.u | 0x8000000000000000ull);
}

}  // namespace qbx
