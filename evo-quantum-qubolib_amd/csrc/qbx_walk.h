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
//   * inner block: a compile-time unrolled walk of the M low bits as a binary trie (each
//     state is its parent with the lowest set bit cleared, plus that bit):
//       E(y) = E(parent) + f_k + c_y   (one v_add3_u32; f_k per lane, c_y uniform)
//     and block minima are taken two states at a time (v_min3_i32): ~1.5 VALU ops/state.
//     Only +f_k is ever added, so no negated fields are kept or updated.
//   * block transition: the outer bits advance in Gray order; flipping outer bit J
//     changes E0 (energy of the block base) by +-(u_J + W[J][J-1]) because, at that
//     moment, the only walked bit below J that is set is J-1.  Fields f_k of the inner
//     bits and "upper fields" u_q of the outer bits below J move by +-W[.][J].
//   * a block is re-walked (to find its lowest-rank minimiser) only when its minimum
//     beats the lane's best in (E, rank) order -- rare after the first few blocks.
//   * ordering: block_key() gives rank bits [M, L) of a block's states and rank_lo() the
//     inner bits [0, M), so (E, lane, key, rank_lo) orders states as (E, rank) does.
//   * float Q: the same int32 walk over quantized tables is a filter with a proven error
//     bound; blocks that pass it are rebuilt and walked exactly in FP64 (walk_lane_q32).
#pragma once
#include <stdint.h>
#include <utility>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define QBX_HD __host__ __device__ __forceinline__
#define QBX_HD_MEMBER __host__ __device__ __forceinline__
#else
#include <math.h>
#define QBX_HD static inline
#define QBX_HD_MEMBER inline
#endif

namespace qbx {

// Tables read through the constant address space: they are written by an earlier launch
// and only read by the walks, so where the address is wave-uniform the device reads them
// with scalar loads into SGPRs (scalar cache, short latency, no VGPRs).
#if defined(__HIP_DEVICE_COMPILE__)
#define QBX_CONST_AS __attribute__((address_space(4)))
#else
#define QBX_CONST_AS
#endif

#if defined(__HIPCC__) || defined(__HIP__)
// v from lane (this lane XOR off), 0 < off < 64: one ds_bpermute per 32 bits.  __shfl_xor
// also tests that the partner lies inside its width, a per-lane bound that the compiler
// keeps live across the whole float walk (in scratch); in a wave of 64 the partner always
// exists.
template <class T>
__device__ __forceinline__ T wave_xor(T v, int off) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    const int a = (int)((__lane_id() ^ (unsigned)off) << 2);
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_ds_bpermute(a, __builtin_bit_cast(int, v)));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)(uint32_t)u);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)(uint32_t)(u >> 32));
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    }
}
// v of lane l, l wave-uniform
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
#endif

// Path counters of the host build (the CPU tests check that every path of the float
// filter is exercised); compiled out of the device build.
#ifdef QBX_EMU_STATS
extern unsigned long long emu_stat_flushes, emu_stat_rewalks;
#define QBX_STAT(x) (x)
#else
#define QBX_STAT(x) ((void)0)
#endif

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
constexpr int MAX_N = 56;   // variables of the largest problem

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
// when step t of a Gray walk over the inner bits flips inner bit k = ctz(t) (the FP64
// block re-walk of float problems walks its blocks in Gray order).
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

// Inner-block trie: every inner state y != 0 is its parent p = y & (y - 1) (lowest set bit
// cleared) plus bit k = ctz(y), so E(y) = E(p) + f_k + c_y with the per-problem constant
//     c_y = sum over set bits l of p of W[k][l]        (y = 1 .. 2^M - 1).
// Walking the block as this trie only ever adds a field (never -f_k as a Gray walk does).
template <class T, class Acc>
QBX_HD Acc trie_elem(const T* Q, int n, int M, uint32_t y) {
    if (y == 0 || y >= (1u << M)) return Acc(0);
    const int k = ctz32(y);
    const uint32_t p = y & (y - 1u);
    Acc s = Acc(0);
    for (int l = 0; l < M; ++l)
        if ((p >> l) & 1u) s += w_elem<T, Acc>(Q, n, k, l);
    return s;
}
// the same from a bit-ordered table (row k, column l = W[k][l])
template <class T, class Acc>
QBX_HD Acc trie_from_table(const T* W, int M, uint32_t y) {
    if (y == 0 || y >= (1u << M)) return Acc(0);
    const int k = ctz32(y);
    const uint32_t p = y & (y - 1u);
    Acc s = Acc(0);
    for (int l = 0; l < M; ++l)
        if ((p >> l) & 1u) s += Acc(W[k * LDW + l]);
    return s;
}

// ------------------------------------------------------------------ int32 lane walk
template <int M, int OMAX>
struct LaneI32 {
    int32_t E0;          // energy of the block base (inner bits = 0)
    int32_t f[M];        // field of inner bit k from every non-inner bit
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
// Empty volatile asm: makes a value opaque to the optimiser (no instruction is emitted).
// Used only on the rare re-walk path, so that its energies are not merged with the
// first walk's (which would keep all 2^M of them live across the block loop).
#if defined(__HIP_DEVICE_COMPILE__)
#define QBX_OPAQUE(x) asm volatile("" : "+v"(x))
#else
#define QBX_OPAQUE(x) ((void)0)
#endif

// Minimum energy over one inner block, walked as the trie in depth-first order: each even
// node P is paired with its leaf P | 1 (one v_add3_u32 each, one v_min3_i32 per pair), then
// P's even children P | 2^k (k < ctz(P), or any k for P = 0) are visited.  At most M - 1
// energies are live on the way down.
QBX_HD constexpr int ctz_c(unsigned v) { return v ? __builtin_ctz(v) : 32; }

template <int M, unsigned P>
QBX_HD void trie_even(uint32_t EP, const int32_t* f, const int32_t* cc, int32_t& mn);

template <int M, unsigned P, int K>
QBX_HD void trie_children(uint32_t EP, const int32_t* f, const int32_t* cc, int32_t& mn) {
    if constexpr (K < M && K < ctz_c(P)) {
        constexpr unsigned C = P | (1u << K);
        trie_even<M, C>(EP + (uint32_t)f[K] + (uint32_t)cc[C], f, cc, mn);
        trie_children<M, P, K + 1>(EP, f, cc, mn);
    }
}

template <int M, unsigned P>
QBX_HD void trie_even(uint32_t EP, const int32_t* f, const int32_t* cc, int32_t& mn) {
    const uint32_t Eodd = EP + (uint32_t)f[0] + (uint32_t)cc[P | 1u];
    mn = min3_i32(mn, (int32_t)EP, (int32_t)Eodd);
    trie_children<M, P, 1>(EP, f, cc, mn);
}

template <int M>
QBX_HD int32_t block_min_i32(int32_t E0, const int32_t* f, const int32_t* cc) {
    int32_t mn = E0;
    trie_even<M, 0u>((uint32_t)E0, f, cc, mn);
    return mn;
}

// Lowest inner rank among the states of the block whose energy equals mn.  The walk only
// records which states hit mn (a 2^M-bit mask, bit y for inner state y); ranks are
// computed afterwards for the set bits alone, so no per-state rank arithmetic exists for
// the compiler to hoist.
template <int M, unsigned P>
QBX_HD void trie_even_mask(uint32_t EP, const int32_t* f, const int32_t* cc, int32_t mn, uint32_t* mask);

template <int M, unsigned P, int K>
QBX_HD void trie_children_mask(uint32_t EP, const int32_t* f, const int32_t* cc, int32_t mn, uint32_t* mask) {
    if constexpr (K < M && K < ctz_c(P)) {
        constexpr unsigned C = P | (1u << K);
        trie_even_mask<M, C>(EP + (uint32_t)f[K] + (uint32_t)cc[C], f, cc, mn, mask);
        trie_children_mask<M, P, K + 1>(EP, f, cc, mn, mask);
    }
}

template <int M, unsigned P>
QBX_HD void trie_even_mask(uint32_t EP, const int32_t* f, const int32_t* cc, int32_t mn, uint32_t* mask) {
    const uint32_t Eodd = EP + (uint32_t)f[0] + (uint32_t)cc[P | 1u];
    if ((int32_t)EP == mn) mask[P >> 5] |= 1u << (P & 31u);
    if ((int32_t)Eodd == mn) mask[(P | 1u) >> 5] |= 1u << ((P | 1u) & 31u);
    trie_children_mask<M, P, 1>(EP, f, cc, mn, mask);
}

template <int M>
QBX_HD uint32_t block_argmin_i32(int32_t E0, const int32_t* f, const int32_t* cc, int32_t mn, int B, int par) {
    int32_t fo[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        fo[k] = f[k];
        QBX_OPAQUE(fo[k]);
    }
    QBX_OPAQUE(E0);
    constexpr int NW = ((1 << M) + 31) / 32;
    uint32_t mask[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) mask[w] = 0u;
    trie_even_mask<M, 0u>((uint32_t)E0, fo, cc, mn, mask);
    uint32_t best = 0xffffffffu;
    for (int w = 0; w < NW; ++w) {
        uint32_t m = mask[w];
        while (m) {
            const uint32_t y = (uint32_t)(w * 32 + __builtin_ctz(m));
            m &= m - 1u;
            const uint32_t r = rank_lo<M>(y, B, par);
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
        for (int k = 0; k < M; ++k) s.f[k] += WJ[k];
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] += WJ[M + q];
    } else {
        s.E0 = (int32_t)((uint32_t)s.E0 - field);
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] -= WJ[k];
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] -= WJ[M + q];
    }
}

// Row M+I of W (entries 0 .. M+I-1: the inner bits and the outer bits below I), loaded one
// block ahead of the static transition that uses it, so that its memory latency is hidden
// behind a block walk.
template <int M>
struct WRow {
    int32_t w[M + 4];
};

template <int M, int I>
QBX_HD void load_wrow(WRow<M>& r, const int32_t* __restrict__ W) {
    const int32_t* WJ = W + (M + I) * LDW;
#pragma unroll
    for (int k = 0; k < M + I; ++k) r.w[k] = WJ[k];
}

template <int M, int OMAX, int I>
QBX_HD void trans_static_row_i32(LaneI32<M, OMAX>& s, const WRow<M>& r, bool up) {
    const uint32_t field = (uint32_t)s.u[I] + (I > 0 ? (uint32_t)r.w[M + I - 1] : 0u);
    if (up) {
        s.E0 = (int32_t)((uint32_t)s.E0 + field);
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] += r.w[k];
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] += r.w[M + q];
    } else {
        s.E0 = (int32_t)((uint32_t)s.E0 - field);
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] -= r.w[k];
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] -= r.w[M + q];
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
    for (int k = 0; k < M; ++k) s.f[k] += sg * WJ[k];
#pragma unroll
    for (int q = 0; q < OMAX; ++q) if (q < i) s.u[q] += sg * WJ[M + q];
}

constexpr int GROUP_BITS = 4;   // blocks are walked in unrolled groups of 16

// Issue priority of the wave from its progress through the lane (c of n blocks): 3 in the
// first quarter down to 0 in the last.  Two waves share a SIMD and the sequencer favours
// the older one, so without this the younger wave of a pair runs behind and finishes its
// walk alone, latency-bound; lowering the priority of the wave that is ahead lets the other
// catch up, and the pair finishes together.  Used by the float filters, whose single solves
// run about one tile per wave (measured on MI355X: n = 34 float 0.63 -> 0.60 ms); the
// integer walks run without it (it cost the n = 40 rebased walk 3%).  Host build: nothing.
QBX_HD void wave_priority_by_progress(uint32_t c, uint32_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t level = (uint32_t)(((uint64_t)c * 4u) / (n ? n : 1u));   // wave-uniform
    if (level == 0) __builtin_amdgcn_s_setprio(3);
    else if (level == 1) __builtin_amdgcn_s_setprio(2);
    else if (level == 2) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
#else
    (void)c;
    (void)n;
#endif
}

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
// The lane's best block so far, (E, K) ordered lexicographically.
struct LaneBest {
    int32_t E;
    uint32_t K;
};

template <int M, int OMAX, int J>
QBX_HD bool group_step_i32(LaneI32<M, OMAX>& s, const int32_t* __restrict__ W, const int32_t* dl,
                           uint32_t c0, uint32_t nblk, int SBk, uint32_t pmaskS, LaneBest& best,
                           WRow<M>& row) {
    static_assert(GROUP_BITS == 4, "WRow holds rows of the outer bits below 4");
    const uint32_t c = c0 + (uint32_t)J;
    if (J > 0) {
        if (c >= nblk) return false;
        constexpr int i = J > 0 ? __builtin_ctz((unsigned)J) : 0;
        // gray(c) bit i = c_i ^ c_{i+1}, and c_i = 1 at the flip
        const bool up = (i + 1 < GROUP_BITS) ? ((((unsigned)J >> (i + 1)) & 1u) == 0u)
                                             : (((c0 >> GROUP_BITS) & 1u) == 0u);
        trans_static_row_i32<M, OMAX, i>(s, row, up);
    }
    // the next block's transition row, in flight while this block is walked
    if (J + 1 < (1 << GROUP_BITS)) load_wrow<M, __builtin_ctz((unsigned)(J + 1))>(row, W);
    const int32_t mn = block_min_i32<M>(s.E0, s.f, dl);
    // rare after the first blocks: one compare per block on the common path, the block key
    // only when the block ties or beats the lane's best
    if (__builtin_expect(mn <= best.E, 0)) {
        const uint32_t K = block_key(c, SBk, pmaskS);
        if (mn < best.E || K < best.K) { best.E = mn; best.K = K; }
    }
    return true;
}

template <int M, int OMAX, int... Js>
QBX_HD void group_i32(LaneI32<M, OMAX>& s, const int32_t* __restrict__ W, const int32_t* dl,
                      uint32_t c0, uint32_t nblk, int SBk, uint32_t pmaskS, LaneBest& best,
                      std::integer_sequence<int, Js...>) {
    WRow<M> row;
    (void)(group_step_i32<M, OMAX, Js>(s, W, dl, c0, nblk, SBk, pmaskS, best, row) && ...);
}

// Lane set-up: E0 and the fields of the lane's first block.  The lane prefix p splits into
// its low 6 bits (per lane) and the bits above them, which depend only on lam >> 6 and so
// are the same for the 64 lanes of a wave: their contributions are accumulated once, from
// wave-uniform values (scalar registers on the device), and only the low bits cost
// per-lane work.
// SROWS: read the low lane bits' rows through the constant address space (scalar loads on
// the device).  Measured on MI355X: the float filters gain from it (n = 34 float solve
// 0.550 -> 0.541 ms), the integer walks lose (n = 40: 24.66 -> 25.19 ms), so only the
// float walks ask for it.
template <int M, int OMAX, bool SROWS = false>
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
#if defined(__HIP_DEVICE_COMPILE__)
    if (H > Hlo) {
        // Wave-cooperative (every lane of the wave calls this when H > 6): lane c sums
        // column c of W over the set high bits a, in full for the columns below them (the
        // fields of the inner, outer and low lane bits) and over a < c for a high bit c
        // itself (its share of E0).  One coalesced row load per high bit, all independent,
        // instead of a serial chain of scalar loads.  Integer sums: the same values.
        const int c = (int)__lane_id();
        const int hcol = L + Hlo;
        uint32_t acc = 0u;
        // batches of 8 unconditional row loads (rows clamped into the table), then masked
        // adds: every load of a batch is in flight before the first is used
        for (int j0 = Hlo; j0 < H; j0 += 8) {
            uint32_t w[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int j = j0 + t < H ? j0 + t : H - 1;
                w[t] = (uint32_t)W[(L + j) * LDW + c];
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int j = j0 + t, a = L + j;
                const bool on = j < H && ((p_hi >> (j - Hlo)) & 1ull) && (c < hcol || a < c);
                acc += w[t] & (0u - (uint32_t)on);
            }
        }
        const uint32_t v = (c < n ? (uint32_t)d[c] : 0u) + acc;
        const bool in_hi = c >= hcol && c < n && ((p_hi >> (c - hcol)) & 1ull);
        uint32_t e = in_hi ? v : 0u;
        // (the float walks exchange without __shfl_xor's width test: see wave_xor; the
        // integer walks' register allocation is best left as it is with the test)
        for (int o = 32; o > 0; o >>= 1)
            e += SROWS ? (uint32_t)wave_xor((int)e, o) : (uint32_t)__shfl_xor((int)e, o);
        Ec = (uint32_t)__builtin_amdgcn_readfirstlane((int)e);
#pragma unroll
        for (int k = 0; k < M; ++k) fc[k] = __builtin_amdgcn_readlane((int)v, k);
#pragma unroll
        for (int q = 0; q < OMAX; ++q) uc[q] = (q < O) ? __builtin_amdgcn_readlane((int)v, M + q) : 0;
#pragma unroll
        for (int i = 0; i < 6; ++i) lc[i] = (i < Hlo) ? __builtin_amdgcn_readlane((int)v, L + i) : 0;
    } else {
#pragma unroll
        for (int k = 0; k < M; ++k) fc[k] = d[k];
#pragma unroll
        for (int q = 0; q < OMAX; ++q) uc[q] = (q < O) ? d[M + q] : 0;
#pragma unroll
        for (int i = 0; i < 6; ++i) lc[i] = (i < Hlo) ? d[L + i] : 0;
    }
#else
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
#endif
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
        // (the same row, read with scalar loads or not)
        const QBX_CONST_AS int32_t* Was = (const QBX_CONST_AS int32_t*)(uintptr_t)(W + a * LDW);
        const int32_t* Wav = W + a * LDW;
        auto Wa = [&](int k) -> int32_t { return SROWS ? Was[k] : Wav[k]; };
        uint32_t e = (uint32_t)lc[i];
#pragma unroll
        for (int i2 = i + 1; i2 < 6; ++i2)
            if (i2 < Hlo) e += (uint32_t)Wa(L + i2) & (0u - ((p_lo >> i2) & 1u));
        E0 += e & mi;
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] += (int32_t)((uint32_t)Wa(k) & mi);
#pragma unroll
        for (int q = 0; q < OMAX; ++q) s.u[q] += (int32_t)((uint32_t)Wa(M + q) & mi);
    }
    s.E0 = (int32_t)E0;
}

// Lane key: (E, K) packed as one signed 64-bit value, E in the high word.  Ordering lane
// keys orders the lane's blocks by (block minimum, block rank).
QBX_HD int32_t key_E(int64_t key) { return (int32_t)(key >> 32); }
QBX_HD uint32_t key_K(int64_t key) { return (uint32_t)key; }

// Walk one lane's 2^L states and return the key of its best block.
// W: bit-ordered, LDW columns; d: n; dl: the trie constants c_y for y < 2^M.
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
    LaneBest best{0x7fffffff, 0xffffffffu};
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
    return (int64_t)(((uint64_t)(uint32_t)best.E << 32) | (uint64_t)best.K);
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
    int32_t bf[M];
    block_state_i32<M>(W, d, n, hs, bE0, bf);
    return block_argmin_i32<M>(bE0, bf, dl, E, B, parp ^ parity32(gc));
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

// ------------------------------------------------ packed pair walk (energies in 16 bits)
// For integer problems whose energy bound (sum |d| + sum_{a<b} |W|) is at most PAIR_BOUND,
// every energy fits in 16 bits and a lane walks two blocks at once: the states whose pair
// bit (state bit M, the lowest outer bit) is 0 in the low half of a 32-bit word, those with
// it set in the high half.  A packed value is V = x0 + x1 2^16 (mod 2^32); the map is
// linear, so the int32 walk carries over unchanged onto packed values (fields
// f0 + f1 2^16, a coupling shared by both halves as w 65537) and one v_add3_u32 now
// advances two states.  Only energies of actual states are ever unpacked, and those fit:
// they are stored biased, E' = E + PAIR_BIAS in [0x0401, 0x7bff].  Block minima are taken
// per half by v_pk_minimum3_f16: those 16-bit patterns are positive normal halfs, whose
// order is their integer order, and the minimum of normal operands is one of them,
// unchanged.  The packed walk covers outer bits 1 .. O-1 in Gray order; the pair bit never
// flips, so its couplings are folded into the fields of the walked bits once, at set-up.
// Block minima, keys and the winning block's re-walk are those of the int32 walk, so the
// result is the int32 walk's (E, rank) minimum exactly.
constexpr int32_t PAIR_BIAS = 0x4000;
constexpr int64_t PAIR_BOUND = 0x4000 - 0x0401;   // 15359 (QBX_PAIR_BOUND)

QBX_HD uint32_t pack_both(uint32_t w) { return w + (w << 16); }

// per-half minimum of three packed values (every half a biased energy)
QBX_HD uint32_t pmin3(uint32_t m, uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 r = __builtin_elementwise_minimum(
        __builtin_elementwise_minimum(__builtin_bit_cast(h2, m), __builtin_bit_cast(h2, a)),
        __builtin_bit_cast(h2, b));
    return __builtin_bit_cast(uint32_t, r);
#else
    uint32_t lo = m & 0xffffu, hi = m >> 16;
    lo = (a & 0xffffu) < lo ? (a & 0xffffu) : lo;
    lo = (b & 0xffffu) < lo ? (b & 0xffffu) : lo;
    hi = (a >> 16) < hi ? (a >> 16) : hi;
    hi = (b >> 16) < hi ? (b >> 16) : hi;
    return lo | (hi << 16);
#endif
}

template <int M, int OMAX>
struct LanePair {
    uint32_t E0;         // packed energies of the two block bases
    uint32_t f[M];       // packed fields of the inner bits
    uint32_t u[OMAX];    // packed upper fields of outer bit q + 1 (the pair bit included)
};

template <int M, unsigned P>
QBX_HD void ptrie_even(uint32_t EP, const uint32_t* f, const uint32_t* cc, uint32_t& mn);

template <int M, unsigned P, int K>
QBX_HD void ptrie_children(uint32_t EP, const uint32_t* f, const uint32_t* cc, uint32_t& mn) {
    if constexpr (K < M && K < ctz_c(P)) {
        constexpr unsigned C = P | (1u << K);
        ptrie_even<M, C>(EP + f[K] + cc[C], f, cc, mn);
        ptrie_children<M, P, K + 1>(EP, f, cc, mn);
    }
}

template <int M, unsigned P>
QBX_HD void ptrie_even(uint32_t EP, const uint32_t* f, const uint32_t* cc, uint32_t& mn) {
    const uint32_t Eodd = EP + f[0] + cc[P | 1u];
    mn = pmin3(mn, EP, Eodd);
    ptrie_children<M, P, 1>(EP, f, cc, mn);
}

// The same trie as two chains in lockstep: node y of the trie over bits 0 .. M-2 and its
// twin y | T (T = 2^(M-1)), whose parent is the twin of y's parent and whose added field is
// y's, so E(y | T) = E(parent(y) | T) + f_k + c_{y|T}.  Each half keeps its own minimum:
// no instruction of one chain waits on the other's.  Measured on MI355X, the pair walk
// gains from it (MaxCut G(40, 1/2): 24.00 -> 23.56 ms), the rebased walk loses (n = 40
// int16: 24.4 -> 24.9 ms) and the packed float filter is unchanged, so only the pair walk
// uses it (TWIN).
template <int M, unsigned P>
QBX_HD void ptrie2_even(uint32_t EP, uint32_t ET, const uint32_t* f, const uint32_t* cc, uint32_t& mn,
                        uint32_t& mt);

template <int M, unsigned P, int K>
QBX_HD void ptrie2_children(uint32_t EP, uint32_t ET, const uint32_t* f, const uint32_t* cc, uint32_t& mn,
                            uint32_t& mt) {
    if constexpr (K < M - 1 && K < ctz_c(P)) {
        constexpr unsigned C = P | (1u << K);
        constexpr unsigned T = 1u << (M - 1);
        ptrie2_even<M, C>(EP + f[K] + cc[C], ET + f[K] + cc[C | T], f, cc, mn, mt);
        ptrie2_children<M, P, K + 1>(EP, ET, f, cc, mn, mt);
    }
}

template <int M, unsigned P>
QBX_HD void ptrie2_even(uint32_t EP, uint32_t ET, const uint32_t* f, const uint32_t* cc, uint32_t& mn,
                        uint32_t& mt) {
    constexpr unsigned T = 1u << (M - 1);
    const uint32_t Eodd = EP + f[0] + cc[P | 1u];
    const uint32_t Todd = ET + f[0] + cc[P | T | 1u];
    mn = pmin3(mn, EP, Eodd);
    mt = pmin3(mt, ET, Todd);
    ptrie2_children<M, P, 1>(EP, ET, f, cc, mn, mt);
}

template <int M, bool TWIN>
QBX_HD uint32_t block_min_pair(uint32_t E0, const uint32_t* f, const uint32_t* cc) {
    uint32_t mn = E0;
    if constexpr (TWIN && M >= 2) {
        const uint32_t ET = E0 + f[M - 1] + cc[1u << (M - 1)];
        uint32_t mt = ET;
        ptrie2_even<M, 0u>(E0, ET, f, cc, mn, mt);
        return pmin3(mn, mt, mt);
    } else {
        ptrie_even<M, 0u>(E0, f, cc, mn);
        return mn;
    }
}

// Row M+1+I of the packed couplings: the inner entries and the walked outer bits below I.
template <int M>
struct WRowPair {
    uint32_t w[M + 4];
};

template <int M, int I>
QBX_HD void load_wrow_pair(WRowPair<M>& r, const uint32_t* __restrict__ Wp) {
    const uint32_t* WJ = Wp + (M + 1 + I) * LDW;
#pragma unroll
    for (int k = 0; k < M; ++k) r.w[k] = WJ[k];
#pragma unroll
    for (int q = 0; q < I; ++q) r.w[M + 1 + q] = WJ[M + 1 + q];
}

template <int M, int OMAX, int I>
QBX_HD void trans_row_pair(LanePair<M, OMAX>& s, const WRowPair<M>& r, bool up) {
    constexpr int J = M + 1 + I;
    const uint32_t field = s.u[I] + (I > 0 ? r.w[J - 1] : 0u);
    if (up) {
        s.E0 += field;
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] += r.w[k];
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] += r.w[M + 1 + q];
    } else {
        s.E0 -= field;
#pragma unroll
        for (int k = 0; k < M; ++k) s.f[k] -= r.w[k];
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] -= r.w[M + 1 + q];
    }
}

// run-time walked bit i >= GROUP_BITS (state bit M + 1 + i)
template <int M, int OMAX>
QBX_HD void trans_dyn_pair(LanePair<M, OMAX>& s, const uint32_t* __restrict__ WJ, int i, bool up) {
    const int J = M + 1 + i;
    uint32_t uJ = 0;
#pragma unroll
    for (int q = 0; q < OMAX; ++q) if (q == i) uJ = s.u[q];
    const uint32_t sg = up ? 1u : 0xffffffffu;
    s.E0 += sg * (uJ + WJ[J - 1]);
#pragma unroll
    for (int k = 0; k < M; ++k) s.f[k] += sg * WJ[k];
#pragma unroll
    for (int q = 0; q < OMAX; ++q) if (q < i) s.u[q] += sg * WJ[M + 1 + q];
}

// Packed block c holds the outer configurations (gray(c) << 1) | h, h = the pair bit; in
// the int32 walk's numbering that is block (c << 1) | (h ^ (c & 1)).
template <int M, int OMAX, int J>
QBX_HD bool group_step_pair(LanePair<M, OMAX>& s, const uint32_t* __restrict__ Wp, const uint32_t* dl,
                            uint32_t c0, uint32_t nblk, int SBk, uint32_t pmaskS, LaneBest& best,
                            WRowPair<M>& row) {
    const uint32_t c = c0 + (uint32_t)J;
    if (J > 0) {
        if (c >= nblk) return false;
        constexpr int i = J > 0 ? __builtin_ctz((unsigned)J) : 0;
        const bool up = (i + 1 < GROUP_BITS) ? ((((unsigned)J >> (i + 1)) & 1u) == 0u)
                                             : (((c0 >> GROUP_BITS) & 1u) == 0u);
        trans_row_pair<M, OMAX, i>(s, row, up);
    }
    if (J + 1 < (1 << GROUP_BITS)) load_wrow_pair<M, __builtin_ctz((unsigned)(J + 1))>(row, Wp);
    const uint32_t mp = block_min_pair<M, true>(s.E0, s.f, dl);
    const uint32_t lo = mp & 0xffffu, hi = mp >> 16;
    const uint32_t m = lo < hi ? lo : hi;
    if (__builtin_expect((int32_t)m <= best.E, 0)) {
        const uint32_t c0f = (c << 1) | (c & 1u);
        if (lo == m) {
            const uint32_t K = block_key(c0f, SBk, pmaskS);
            if ((int32_t)m < best.E || K < best.K) { best.E = (int32_t)m; best.K = K; }
        }
        if (hi == m) {
            const uint32_t K = block_key(c0f ^ 1u, SBk, pmaskS);
            if ((int32_t)m < best.E || K < best.K) { best.E = (int32_t)m; best.K = K; }
        }
    }
    return true;
}

template <int M, int OMAX, int... Js>
QBX_HD void group_pair(LanePair<M, OMAX>& s, const uint32_t* __restrict__ Wp, const uint32_t* dl,
                       uint32_t c0, uint32_t nblk, int SBk, uint32_t pmaskS, LaneBest& best,
                       std::integer_sequence<int, Js...>) {
    WRowPair<M> row;
    (void)(group_step_pair<M, OMAX, Js>(s, Wp, dl, c0, nblk, SBk, pmaskS, best, row) && ...);
}

// Set-up: the int32 walk's first block (pair bit 0) in the low halves; the high halves are
// the same states with the pair bit set.
template <int M, int OMAX>
QBX_HD void lane_setup_pair(LanePair<M, OMAX>& s, const int32_t* __restrict__ W, const int32_t* __restrict__ d,
                            int n, int L, int B, uint64_t lam) {
    LaneI32<M, OMAX> a;
    lane_setup_i32<M, OMAX>(a, W, d, n, L, B, lam);
    const int32_t* WM = W + M * LDW;
    const uint32_t e0 = (uint32_t)a.E0 + (uint32_t)PAIR_BIAS;
    const uint32_t e1 = e0 + (uint32_t)a.u[0];
    s.E0 = e0 + (e1 << 16);
#pragma unroll
    for (int k = 0; k < M; ++k) s.f[k] = (uint32_t)a.f[k] + (((uint32_t)a.f[k] + (uint32_t)WM[k]) << 16);
#pragma unroll
    for (int q = 0; q + 1 < OMAX; ++q)
        s.u[q] = (uint32_t)a.u[q + 1] + (((uint32_t)a.u[q + 1] + (uint32_t)WM[M + 1 + q]) << 16);
    s.u[OMAX - 1] = 0u;
}

// The int32 walk's lane key (lane_best_block_i32), from the packed walk.  Needs L > M and
// bound <= PAIR_BOUND; Wp, dlp: the couplings and trie constants times 65537.
template <int M, int OMAX>
QBX_HD int64_t lane_best_block_pair(const int32_t* __restrict__ W, const uint32_t* __restrict__ Wp,
                                    const int32_t* __restrict__ d, int n, int L, int B, uint64_t lam,
                                    const uint32_t* dlp) {
    static_assert(OMAX > GROUP_BITS, "OMAX must cover the unrolled transition bits");
    const int O = L - M;
    const uint64_t p = lane_prefix(lam, L, B);
    LanePair<M, OMAX> s;
    lane_setup_pair<M, OMAX>(s, W, d, n, L, B, lam);
    const int SBk = (B - M) < 0 ? 0 : ((B - M) > O ? O : (B - M));
    const uint32_t pmaskS = (parity64(p) && O > SBk) ? (lowmask32(O - SBk) << SBk) : 0u;
    LaneBest best{0x7fffffff, 0xffffffffu};
    const uint32_t nblk = 1u << (O - 1);
    for (uint32_t c0 = 0;;) {
        group_pair<M, OMAX>(s, Wp, dlp, c0, nblk, SBk, pmaskS, best,
                            std::make_integer_sequence<int, (1 << GROUP_BITS)>{});
        c0 += 1u << GROUP_BITS;
        if (c0 >= nblk) break;
        const int i = ctz32(c0);
        const bool up = (((c0 ^ (c0 >> 1)) >> i) & 1u) != 0u;
        trans_dyn_pair<M, OMAX>(s, Wp + (M + 1 + i) * LDW, i, up);
    }
    const int32_t E = best.E - PAIR_BIAS;
    return (int64_t)(((uint64_t)(uint32_t)E << 32) | (uint64_t)best.K);
}

// ------------------------------------------- rebased pair walk (energies relative to blocks)
// Integer problems beyond PAIR_BOUND whose energy *differences inside a block* fit 16 bits
// (random int16 QUBOs, for instance) walk the same two blocks per 32-bit word, with each
// half holding E(y) - E(base of its block) + PAIR_BIAS; the int32 block bases E0 (pair bit
// clear) and E0 + up (pair bit set) are kept beside them.  Inside a block,
//     |E(y) - E(base)| <= sum_k |f_k| + sum_{k<l<M} |W_kl|,
// and within one unrolled group of 16 blocks (walked bits 1..4 vary, as does nothing else)
// every field stays within its value at the group's start plus the couplings to those 4 bits;
// the pair bit adds |W_Mk| to the high half's fields.  So with
//     D = sum_{k<l<M} |W_kl| + sum_k sum_{i<4} |W_{M+1+i,k}| + sum_k |W_Mk|   (per problem)
// a group may run packed when sum_k |f_k(group start)| <= PAIR_BOUND - D (the guard, one
// v_sad_u32 per field); every energy the packed group materialises then fits its half.
// A wave any of whose lanes fails the guard walks that group in int32 (both halves, one after
// the other, trie constants unpacked from the packed table) -- a rare, slower path with the
// same keys.  Fields are kept unpacked between groups (exact, any size) and packed for the
// group: F_k = f_k + (f_k + W_Mk) 2^16; at the group's end the guard has bounded every |f_k|
// below 2^15, so f_k is the low half of F_k, sign-extended.
template <int M, int OMAX>
struct LaneRP {
    int32_t E0;          // energy of the low block's base (pair bit clear)
    int32_t up;          // field of the pair bit: the high block's base is E0 + up
    int32_t f[M];        // inner fields of the low block (unpacked between groups)
    int32_t u[OMAX];     // upper fields of walked bit q (state bit M + 1 + q), low block
};

// Row J = M+1+I of a packed group transition: the packed inner couplings and the unpacked
// couplings to the pair bit (w[0]) and to walked bits 0 .. I-1 (w[1 + q]).
template <int M>
struct WRowRP {
    uint32_t p[M];
    int32_t w[4];
};

template <int M, int I>
QBX_HD void load_wrow_rp(WRowRP<M>& r, const int32_t* __restrict__ W, const uint32_t* __restrict__ Wp) {
    constexpr int J = M + 1 + I;
    const uint32_t* WPJ = Wp + J * LDW;
    const int32_t* WJ = W + J * LDW;
#pragma unroll
    for (int k = 0; k < M; ++k) r.p[k] = WPJ[k];
    r.w[0] = WJ[M];
#pragma unroll
    for (int q = 0; q < I; ++q) r.w[1 + q] = WJ[M + 1 + q];
}

// The same row read through the constant address space: the tables are written by an
// earlier launch and only read by the walks, and the address is wave-uniform, so on the
// device these become scalar loads into SGPRs (scalar cache, short latency, no VGPRs).
template <int M, int I>
QBX_HD void load_wrow_rp_uniform(WRowRP<M>& r, const int32_t* __restrict__ W, const uint32_t* __restrict__ Wp) {
    constexpr int J = M + 1 + I;
    const QBX_CONST_AS uint32_t* WPJ = (const QBX_CONST_AS uint32_t*)(uintptr_t)(Wp + J * LDW);
    const QBX_CONST_AS int32_t* WJ = (const QBX_CONST_AS int32_t*)(uintptr_t)(W + J * LDW);
#pragma unroll
    for (int k = 0; k < M; ++k) r.p[k] = WPJ[k];
    r.w[0] = WJ[M];
#pragma unroll
    for (int q = 0; q < I; ++q) r.w[1 + q] = WJ[M + 1 + q];
}

template <int M, int OMAX, int I>
QBX_HD void trans_row_rp(LaneRP<M, OMAX>& s, uint32_t* F, const WRowRP<M>& r, bool up) {
    const uint32_t field = (uint32_t)s.u[I] + (I > 0 ? (uint32_t)r.w[I] : 0u);   // w[I] = W[J][J-1]
    if (up) {
        s.E0 = (int32_t)((uint32_t)s.E0 + field);
        s.up = (int32_t)((uint32_t)s.up + (uint32_t)r.w[0]);
#pragma unroll
        for (int k = 0; k < M; ++k) F[k] += r.p[k];
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] = (int32_t)((uint32_t)s.u[q] + (uint32_t)r.w[1 + q]);
    } else {
        s.E0 = (int32_t)((uint32_t)s.E0 - field);
        s.up = (int32_t)((uint32_t)s.up - (uint32_t)r.w[0]);
#pragma unroll
        for (int k = 0; k < M; ++k) F[k] -= r.p[k];
#pragma unroll
        for (int q = 0; q < I; ++q) s.u[q] = (int32_t)((uint32_t)s.u[q] - (uint32_t)r.w[1 + q]);
    }
}

// unpacked transition on walked bit i (state bit J = M + 1 + i), any i (run-time)
template <int M, int OMAX>
QBX_HD void trans_dyn_rp(LaneRP<M, OMAX>& s, const int32_t* __restrict__ WJ, int i, bool up) {
    const int J = M + 1 + i;
    uint32_t uJ = 0;
#pragma unroll
    for (int q = 0; q < OMAX; ++q) if (q == i) uJ = (uint32_t)s.u[q];
    const uint32_t field = uJ + (i > 0 ? (uint32_t)WJ[J - 1] : 0u);
    const uint32_t sg = up ? 1u : 0xffffffffu;
    s.E0 = (int32_t)((uint32_t)s.E0 + sg * field);
    s.up = (int32_t)((uint32_t)s.up + sg * (uint32_t)WJ[M]);
#pragma unroll
    for (int k = 0; k < M; ++k) s.f[k] = (int32_t)((uint32_t)s.f[k] + sg * (uint32_t)WJ[k]);
#pragma unroll
    for (int q = 0; q < OMAX; ++q) if (q < i) s.u[q] = (int32_t)((uint32_t)s.u[q] + sg * (uint32_t)WJ[M + 1 + q]);
}

// unpacked transition on walked bit i < GROUP_BITS (inside a group; run-time i)
template <int M, int OMAX>
QBX_HD void trans_grp_rp(LaneRP<M, OMAX>& s, const int32_t* __restrict__ WJ, int i, bool up) {
    const int J = M + 1 + i;
    uint32_t uJ = 0;
#pragma unroll
    for (int q = 0; q < GROUP_BITS; ++q) if (q == i) uJ = (uint32_t)s.u[q];
    const uint32_t field = uJ + (i > 0 ? (uint32_t)WJ[J - 1] : 0u);
    const uint32_t sg = up ? 1u : 0xffffffffu;
    s.E0 = (int32_t)((uint32_t)s.E0 + sg * field);
    s.up = (int32_t)((uint32_t)s.up + sg * (uint32_t)WJ[M]);
#pragma unroll
    for (int k = 0; k < M; ++k) s.f[k] = (int32_t)((uint32_t)s.f[k] + sg * (uint32_t)WJ[k]);
#pragma unroll
    for (int q = 0; q + 1 < GROUP_BITS; ++q) if (q < i) s.u[q] = (int32_t)((uint32_t)s.u[q] + sg * (uint32_t)WJ[M + 1 + q]);
}

// Lane best-key update from the minima of the two blocks of packed block c (mA, mB: int32)
QBX_HD void rp_best_update(int32_t mA, int32_t mB, uint32_t c, int SBk, uint32_t pmaskS, LaneBest& best) {
    const int32_t m = mA < mB ? mA : mB;
    if (__builtin_expect(m <= best.E, 0)) {
        const uint32_t cf = (c << 1) | (c & 1u);   // int32-walk block of the low half
        if (mA == m) {
            const uint32_t K = block_key(cf, SBk, pmaskS);
            if (m < best.E || K < best.K) { best.E = m; best.K = K; }
        }
        if (mB == m) {
            const uint32_t K = block_key(cf ^ 1u, SBk, pmaskS);
            if (m < best.E || K < best.K) { best.E = m; best.K = K; }
        }
    }
}

template <int M, int OMAX, int J>
QBX_HD bool group_step_rp(LaneRP<M, OMAX>& s, uint32_t* F, const int32_t* __restrict__ W,
                          const uint32_t* __restrict__ Wp, const uint32_t* dl, uint32_t c0, uint32_t nblk,
                          int SBk, uint32_t pmaskS, LaneBest& best, WRowRP<M>& row) {
    const uint32_t c = c0 + (uint32_t)J;
    if (J > 0) {
        if (c >= nblk) return false;
        constexpr int i = J > 0 ? __builtin_ctz((unsigned)J) : 0;
        const bool up = (i + 1 < GROUP_BITS) ? ((((unsigned)J >> (i + 1)) & 1u) == 0u)
                                             : (((c0 >> GROUP_BITS) & 1u) == 0u);
        trans_row_rp<M, OMAX, i>(s, F, row, up);
    }
    if (J + 1 < (1 << GROUP_BITS)) load_wrow_rp<M, __builtin_ctz((unsigned)(J + 1))>(row, W, Wp);
    const uint32_t mp = block_min_pair<M, false>((uint32_t)PAIR_BIAS * 65537u, F, dl);
    const int32_t b = s.E0 - PAIR_BIAS;
    const int32_t mA = b + (int32_t)(mp & 0xffffu);
    const int32_t mB = b + s.up + (int32_t)(mp >> 16);
    rp_best_update(mA, mB, c, SBk, pmaskS, best);
    return true;
}

template <int M, int OMAX, int... Js>
QBX_HD void group_rp(LaneRP<M, OMAX>& s, uint32_t* F, const int32_t* __restrict__ W, const uint32_t* __restrict__ Wp,
                     const uint32_t* dl, uint32_t c0, uint32_t nblk, int SBk, uint32_t pmaskS, LaneBest& best,
                     std::integer_sequence<int, Js...>) {
    WRowRP<M> row;
    (void)(group_step_rp<M, OMAX, Js>(s, F, W, Wp, dl, c0, nblk, SBk, pmaskS, best, row) && ...);
}

// Unpacked int32 minimum of one block whose trie constants come packed (low half, exact:
// every |c_y| <= sum_{k<l<M} |W_kl| < 2^15 on this path).  z is an opaque zero added before
// each decode, so that the decoded constants are not hoisted out of the caller's loop (that
// would keep 2^M more values live).
QBX_HD int32_t unpack_lo(uint32_t v, uint32_t z) { return (int32_t)(int16_t)((v + z) & 0xffffu); }

template <int M, unsigned P>
QBX_HD void utrie_even(uint32_t EP, const int32_t* f, const uint32_t* cc, uint32_t z, int32_t& mn);

template <int M, unsigned P, int K>
QBX_HD void utrie_children(uint32_t EP, const int32_t* f, const uint32_t* cc, uint32_t z, int32_t& mn) {
    if constexpr (K < M && K < ctz_c(P)) {
        constexpr unsigned C = P | (1u << K);
        utrie_even<M, C>(EP + (uint32_t)f[K] + (uint32_t)unpack_lo(cc[C], z), f, cc, z, mn);
        utrie_children<M, P, K + 1>(EP, f, cc, z, mn);
    }
}

template <int M, unsigned P>
QBX_HD void utrie_even(uint32_t EP, const int32_t* f, const uint32_t* cc, uint32_t z, int32_t& mn) {
    const uint32_t Eodd = EP + (uint32_t)f[0] + (uint32_t)unpack_lo(cc[P | 1u], z);
    mn = min3_i32(mn, (int32_t)EP, (int32_t)Eodd);
    utrie_children<M, P, 1>(EP, f, cc, z, mn);
}

// The group at c0 walked in int32 (the rare path of a wave whose guard failed): the low
// blocks' state advances as in the packed group (and is left as the packed group leaves
// it), and each block's high twin is evaluated from it (fields + W_Mk, base + up); blocks
// and halves in run-time loops, so that the path stays small in code and registers.
template <int M, int OMAX>
QBX_HD void group_rp_int32(LaneRP<M, OMAX>& s, const int32_t* __restrict__ W, const uint32_t* dl,
                           uint32_t c0, uint32_t nblk, int SBk, uint32_t pmaskS, LaneBest& best) {
    const int32_t* WM = W + M * LDW;
    const uint32_t cend = (c0 + (1u << GROUP_BITS) < nblk) ? c0 + (1u << GROUP_BITS) : nblk;
#pragma unroll 1
    for (uint32_t c = c0; c < cend; ++c) {
        if (c > c0) {
            const int i = ctz32(c - c0);
            const bool up = (((c ^ (c >> 1)) >> i) & 1u) != 0u;
            trans_grp_rp<M, OMAX>(s, W + (M + 1 + i) * LDW, i, up);
        }
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            int32_t fh[M];
#pragma unroll
            for (int k = 0; k < M; ++k) fh[k] = (int32_t)((uint32_t)s.f[k] + (h ? (uint32_t)WM[k] : 0u));
            const int32_t E0h = (int32_t)((uint32_t)s.E0 + (h ? (uint32_t)s.up : 0u));
            int32_t mn = E0h;
            uint32_t z = 0u;
            QBX_OPAQUE(z);
            utrie_even<M, 0u>((uint32_t)E0h, fh, dl, z, mn);
            if (mn <= best.E) {
                const uint32_t K = block_key((c << 1) | ((c & 1u) ^ (uint32_t)h), SBk, pmaskS);
                if (mn < best.E || K < best.K) { best.E = mn; best.K = K; }
            }
        }
    }
}

// Set-up: the int32 walk's first block (pair bit clear) and the pair bit's field.
template <int M, int OMAX, bool SROWS = false>
QBX_HD void lane_setup_rp(LaneRP<M, OMAX>& s, const int32_t* __restrict__ W, const int32_t* __restrict__ d,
                          int n, int L, int B, uint64_t lam) {
    LaneI32<M, OMAX> a;
    lane_setup_i32<M, OMAX, SROWS>(a, W, d, n, L, B, lam);
    s.E0 = a.E0;
    s.up = a.u[0];
#pragma unroll
    for (int k = 0; k < M; ++k) s.f[k] = a.f[k];
#pragma unroll
    for (int q = 0; q + 1 < OMAX; ++q) s.u[q] = a.u[q + 1];
    s.u[OMAX - 1] = 0;
}

// The int32 walk's lane key (lane_best_block_i32) from the rebased pair walk.  Needs L > M,
// every |c_y| < 2^15 and slack = PAIR_BOUND - D >= 0 (D as above, for this L and M).
// groups_int32 (nullable) counts the groups walked in int32.
template <int M, int OMAX>
QBX_HD int64_t lane_best_block_rp(const int32_t* __restrict__ W, const uint32_t* __restrict__ Wp,
                                  const int32_t* __restrict__ d, int n, int L, int B, uint64_t lam,
                                  const uint32_t* dlp, int32_t slack, uint32_t* groups_int32) {
    static_assert(OMAX > GROUP_BITS, "OMAX must cover the unrolled transition bits");
    const int O = L - M;
    const uint64_t p = lane_prefix(lam, L, B);
    LaneRP<M, OMAX> s;
    lane_setup_rp<M, OMAX>(s, W, d, n, L, B, lam);
    const int SBk = (B - M) < 0 ? 0 : ((B - M) > O ? O : (B - M));
    const uint32_t pmaskS = (parity64(p) && O > SBk) ? (lowmask32(O - SBk) << SBk) : 0u;
    LaneBest best{0x7fffffff, 0xffffffffu};
    const uint32_t nblk = 1u << (O - 1);
    const int32_t* WM = W + M * LDW;
    for (uint32_t c0 = 0;;) {
        uint32_t g = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) g += (uint32_t)(s.f[k] < 0 ? -s.f[k] : s.f[k]);
        bool ok = g <= (uint32_t)slack;
#if defined(__HIP_DEVICE_COMPILE__)
        ok = __all(ok);
#endif
        if (__builtin_expect(ok, 1)) {
            uint32_t F[M];
#pragma unroll
            for (int k = 0; k < M; ++k) F[k] = (uint32_t)s.f[k] + (((uint32_t)s.f[k] + (uint32_t)WM[k]) << 16);
            group_rp<M, OMAX>(s, F, W, Wp, dlp, c0, nblk, SBk, pmaskS, best,
                              std::make_integer_sequence<int, (1 << GROUP_BITS)>{});
#pragma unroll
            for (int k = 0; k < M; ++k) s.f[k] = (int32_t)(int16_t)(F[k] & 0xffffu);
        } else {
            group_rp_int32<M, OMAX>(s, W, dlp, c0, nblk, SBk, pmaskS, best);
            if (groups_int32) ++*groups_int32;
        }
        c0 += 1u << GROUP_BITS;
        if (c0 >= nblk) break;
        const int i = ctz32(c0);
        const bool up = (((c0 ^ (c0 >> 1)) >> i) & 1u) != 0u;
        trans_dyn_rp<M, OMAX>(s, W + (M + 1 + i) * LDW, i, up);
    }
    return (int64_t)(((uint64_t)(uint32_t)best.E << 32) | (uint64_t)best.K);
}

// D of the guard for a problem (bit-ordered W), inner bits M, walk bits L.
QBX_HD int64_t rp_guard_D(const int32_t* W, int M, int L) {
    int64_t cin = 0, dg = 0, dp = 0;
    for (int k = 0; k < M; ++k)
        for (int l = k + 1; l < M; ++l) {
            const int64_t w = W[k * LDW + l];
            cin += w < 0 ? -w : w;
        }
    for (int k = 0; k < M; ++k) {
        const int64_t wm = W[M * LDW + k];
        dp += wm < 0 ? -wm : wm;
        for (int i = 0; i < GROUP_BITS; ++i) {
            const int J = M + 1 + i;
            if (J >= L) break;
            const int64_t w = W[J * LDW + k];
            dg += w < 0 ? -w : w;
        }
    }
    return cin + dg + dp;
}

// Whether the rebased pair walk is expected to pay off: the guard's room PAIR_BOUND - D
// against 1.5 x sum_k sqrt(E f_k^2), the fields' root mean square over the states the
// walk visits (bits M .. n-1 uniform).  Only the speed depends on this choice.
QBX_HD bool rp_worth_it(const int32_t* W, const int32_t* d, int n, int M, int L, int64_t D) {
    if (D >= PAIR_BOUND || L <= M) return false;
    double S = 0.0;
    for (int k = 0; k < M; ++k) {
        double mean = d[k], var = 0.0;
        for (int j = M; j < n; ++j) {
            const double w = W[k * LDW + j];
            mean += 0.5 * w;
            var += 0.25 * w * w;
        }
        S += sqrt(mean * mean + var);
    }
    return (double)D + 1.5 * S <= (double)PAIR_BOUND;
}

// ----------------------------------------------- float problems: int32 filter, FP64 walk
// The block filter runs on quantized tables Wq = rint(W 2^s), dq = rint(d 2^s) in exact
// int32 arithmetic (the int walk above, unchanged).  Each of the terms d_a, W_ab that a
// state x switches on is off by at most 1/2 after quantization, so with T(x) <= n(n+1)/2
// terms, |Eq(x) - 2^s E(x)| <= n(n+1)/4.  A block can therefore hold a state with
// E(x) <= b only if its int32 minimum is <= 2^s b + n(n+1)/4; the blocks that pass are
// rebuilt from scratch in FP64 and re-walked exactly.
struct LaneOutF64 {
    double E;
    uint64_t rank;
};

// 2^s: the largest power of two that keeps the quantized energy bound within 2^29
// (sum |rint(v 2^s)| <= bound 2^s + (number of terms) / 2 <= 2^29 - 2), so that the
// packed filter's differences of two quantized energies fit int32 with room to spare.
QBX_HD double q_scale(double bound, int n) {
    const double half_terms = 0.25 * (double)n * (double)(n + 1);
    const double room = 536870910.0 - half_terms - 1.0;
    if (!(bound > 0.0)) return 1.0;
    const double r = room / bound;
    if (!(r < 0x1p1000)) return 0x1p1000;   // tiny entries: any large scale is exact enough
    int e = 0;
    (void)frexp(r, &e);                     // r = m 2^e, m in [0.5, 1)
    return ldexp(1.0, e - 1);
}
// margin in quantized units: n(n+1)/4, plus one for the FP64 rounding of the energies
QBX_HD double q_margin(int n) { return 0.25 * (double)n * (double)(n + 1) + 1.0; }
// int32 threshold for "can hold a state with E <= b"
QBX_HD int32_t q_threshold(double b, double scale, double qm) {
    // no threshold for a NaN; +inf (the empty key's energy) gives t = +inf below
    if (b != b) return 0x7fffffff;
    const double t = floor(b * scale + qm);
    if (t >= 2147483647.0) return 0x7fffffff;
    if (t <= -2147483648.0) return (int32_t)0x80000000;
    return (int32_t)t;
}
// E0 and inner fields of an arbitrary block, from scratch, in FP64 (hs: bits above the block).
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

// Rebuild block c of the lane (prefix p, O outer bits) in FP64 and walk it exactly.
template <int M>
QBX_HD void block_exact_f64(const double* __restrict__ W, const double* __restrict__ d, int n, uint64_t p,
                            int O, uint32_t c, uint64_t laneRank, int SBk, uint32_t pmaskS, int B, int parp,
                            const double* __restrict__ dld, double& bestE, uint64_t& bestR) {
    const uint32_t gc = c ^ (c >> 1);
    double bE0, bf[M];
    block_state_f64<M>(W, d, n, (p << O) | (uint64_t)gc, bE0, bf);
    block_rewalk_f64<M>(bE0, bf, dld, laneRank | ((uint64_t)block_key(c, SBk, pmaskS) << M), B,
                        parp ^ parity32(gc), bestE, bestR);
}

// Deferred candidate blocks of one lane: (block index, int32 block minimum) records in a
// per-lane buffer (LDS on the device, entry k of a lane at k * stride).  A block that
// passes the filter is only recorded; it is rebuilt and walked in FP64 at the end of the
// lane, or at a group boundary when the buffer might not hold another group, and only if
// it still passes then (the lane minimum and the shared bound keep tightening, so most
// records never are).  A group adds at most 2^GROUP_BITS records, so nothing is walked
// in FP64 inside the unrolled group.
constexpr int QCAP = 32;
constexpr uint32_t QSYNC = 8;   // groups between exchanges of the shared bound
// cap on the packed float filter's thresholds: above every quantized energy (< 2^29), so
// that threshold - base fits int32
constexpr int32_t QRP_THR_MAX = 0x20100000;

struct QBuf {
    uint32_t* c;
    int32_t* m;
    int stride;
    int count;
};

// Lane state of the filtered walk besides the int32 block state.
struct QLane {
    int32_t laneMin;      // least int32 block minimum seen by the lane
    int32_t thrB;         // threshold from the exact / shared bound
    int32_t thr;          // min(thrB, laneMin + 2 qm): the filter
    int32_t two_qm;       // 2 * n(n+1)/4 + 2, in quantized units
};

QBX_HD int32_t sat_add_i32(int32_t a, int32_t b) {     // b >= 0
    return a > 0x7fffffff - b ? 0x7fffffff : a + b;
}
QBX_HD void q_update_thr(QLane& q) {
    const int32_t loc = sat_add_i32(q.laneMin, q.two_qm);
    q.thr = q.thrB < loc ? q.thrB : loc;
}

template <int M>
struct QCtx {    // what a flush needs to rebuild and walk a block exactly
    const double* W;
    const double* d;
    const double* dld;
    int n, L, B;
    uint64_t lam;    // the lane; its prefix, rank bits and block-key mask are derived per flush
    // packed float filter: the quantized tables, for the int32 filter's exact test of a
    // record before its FP64 walk
    const int32_t* Wq = nullptr;
    const int32_t* dq = nullptr;
    const int32_t* dlq = nullptr;
    QBX_HD_MEMBER void set(const double* W_, const double* d_, const double* dld_, int n_, int L_, int B_,
                           uint64_t lam_) {
        W = W_; d = d_; dld = dld_; n = n_; L = L_; B = B_; lam = lam_;
    }
};

// Drop the records the filter no longer passes; with `walk`, walk the rest in FP64 and
// empty the buffer.  PACKED records (the packed float filter) name a packed block c and
// the halves h that passed, as (c << 2) | mask; half h is block (c << 1) | (h ^ (c & 1))
// of the int32 walk's numbering.  Their minimum m is a coarse lower bound, so before its
// FP64 walk each half gets the int32 filter's own test: its exact int32 minimum on the
// quantized tables, rebuilt from scratch, against the threshold.  What is walked in FP64
// is then what the int32 filter would walk.
template <int M, bool PACKED = false>
QBX_HD void q_compact(QBuf& b, QLane& q, const QCtx<M>& x, bool walk, double& bestE, uint64_t& bestR) {
    const int O = x.L - M;
    const uint64_t p = lane_prefix(x.lam, x.L, x.B);
    const uint64_t laneRank = shl64(x.lam, x.L);
    const int SBk = (x.B - M) < 0 ? 0 : ((x.B - M) > O ? O : (x.B - M));
    const int parp = parity64(p);
    const uint32_t pmaskS = (parp && O > SBk) ? (lowmask32(O - SBk) << SBk) : 0u;
    int keep = 0;
    for (int k = 0; k < b.count; ++k) {
        const int32_t m = b.m[k * b.stride];
        if (m > q.thr) continue;
        const uint32_t c = b.c[k * b.stride];
        if (walk) {
            if (PACKED) {
                const uint32_t cp = c >> 2;
                for (uint32_t h = 0; h < 2u; ++h) {
                    if (!((c >> h) & 1u)) continue;
                    const uint32_t cf = (cp << 1) | (h ^ (cp & 1u));
                    int32_t E0, f[M];
                    block_state_i32<M>(x.Wq, x.dq, x.n, (p << O) | (uint64_t)(cf ^ (cf >> 1)), E0, f);
                    const int32_t mn = block_min_i32<M>(E0, f, x.dlq);
                    // an exact minimum is a state's quantized energy: the lane minimum may
                    // take it (the coarse records only gave it upper bounds)
                    if (mn < q.laneMin) {
                        q.laneMin = mn;
                        q_update_thr(q);
                        q.thr = q.thr < QRP_THR_MAX ? q.thr : QRP_THR_MAX;
                    }
                    if (mn > q.thr) continue;
                    QBX_STAT(++emu_stat_rewalks);
                    block_exact_f64<M>(x.W, x.d, x.n, p, O, cf, laneRank, SBk, pmaskS, x.B, parp, x.dld, bestE,
                                       bestR);
                }
            } else {
                QBX_STAT(++emu_stat_rewalks);
                block_exact_f64<M>(x.W, x.d, x.n, p, O, c, laneRank, SBk, pmaskS, x.B, parp, x.dld, bestE, bestR);
            }
        } else {
            b.c[keep * b.stride] = c;
            b.m[keep * b.stride] = m;
            ++keep;
        }
    }
    b.count = keep;
}

#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int l) {
    return ((uint64_t)readlane_u32((uint32_t)(v >> 32), l) << 32) | readlane_u32((uint32_t)v, l);
}
__device__ __forceinline__ int32_t wave_min_i32(int32_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        const int32_t w = wave_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}

// Lane j's column of the base state (inner bits 0) of the block whose bits above the inner
// block are hs: d_j plus W_aj over the set high bits a (a >= M), all of them for an inner
// column j < M (its field f_j), those below j for j >= M (bit j's share of E0 when set).
// Row a is one coalesced load of its LDW padded entries; no load past row n - 1.
template <class T, class Acc, int M>
__device__ __forceinline__ Acc block_column_wave(const T* __restrict__ W, const T* __restrict__ d, int n,
                                                 uint64_t hs, int lane) {
    QBX_OPAQUE(lane);   // the column offsets from here, not kept from the kernel's start
    Acc acc = lane < n ? (Acc)d[lane] : (Acc)0;
    const int lim = lane < M ? n : lane;
    // batches of 8 unconditional row loads (rows clamped into the table), then the adds
    for (int a0 = M; a0 < n; a0 += 8) {
        T w[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) w[t] = W[(a0 + t < n ? a0 + t : n - 1) * LDW + lane];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int a = a0 + t;
            const bool on = a < n && ((hs >> (a - M)) & 1ull) && a < lim;
            acc = on ? acc + (Acc)w[t] : acc;
        }
    }
    return acc;
}

// Both of a block's base columns at once, the quantized (int32) and the FP64 one, in one
// pass over the rows: every row load of the two tables is issued before the first is
// needed, so a record's exact int32 test and its FP64 walk share one memory latency.
// The FP64 sum is block_column_wave's, in the same order.
template <int M>
__device__ __forceinline__ void block_columns_wave(const int32_t* __restrict__ Wq, const int32_t* __restrict__ dq,
                                                   const double* __restrict__ W, const double* __restrict__ d,
                                                   int n, uint64_t hs, int lane, uint32_t& colq, double& col) {
    QBX_OPAQUE(lane);   // the column offsets from here, not kept from the kernel's start (in scratch)
    uint32_t aq = lane < n ? (uint32_t)dq[lane] : 0u;
    double ad = lane < n ? d[lane] : 0.0;
    const int lim = lane < M ? n : lane;
    // batches of 8 unconditional row loads (rows clamped into the tables), then masked adds
    for (int a0 = M; a0 < n; a0 += 8) {
        uint32_t wq[8];
        double wd[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int a = a0 + t < n ? a0 + t : n - 1;
            wq[t] = (uint32_t)Wq[a * LDW + lane];
            wd[t] = W[a * LDW + lane];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int a = a0 + t;
            const bool on = a < n && ((hs >> (a - M)) & 1ull) && a < lim;
            aq += wq[t] & (0u - (uint32_t)on);
            ad = on ? ad + wd[t] : ad;
        }
    }
    colq = aq;
    col = ad;
}

// Energy of inner state y of a block from its base energy E0, fields f and the inner
// couplings W_lk (l < k < M): E0 + sum over set bits k of (f_k + sum over set l < k of W_lk).
template <class T, class Acc, int M>
__device__ __forceinline__ Acc inner_energy(Acc E0, const Acc* f, const T* __restrict__ W, uint32_t y) {
    Acc E = E0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        Acc e = f[k];
#pragma unroll
        for (int l = 0; l < k; ++l) e = ((y >> l) & 1u) ? e + (Acc)W[l * LDW + k] : e;
        E = ((y >> k) & 1u) ? E + e : E;
    }
    return E;
}
#endif

// The lane-end flush of the filtered walks (q_compact with `walk`), done by the whole wave:
// every passing record of every lane is tested and walked by the 64 lanes together.  A
// lane's own flush walks its records one after another, its block rebuilt serially from
// scratch, while the wave's other lanes wait: one record took 0.1 ms, a sixth of a 2^34-state
// solve, in the wave holding the answer.  Here lane j sums column j of the tables over the
// block's set high bits, the 2^M inner states are split over the lanes and evaluated
// directly, and wave reductions give the block's exact int32 minimum (PACKED records' test
// before their FP64 walk) and its FP64 (E, rank) minimum, which every lane folds into its
// best.  Records are pruned by the least lane minimum and threshold over the wave (each is
// a bound from a real state, whichever lane found it).  Every lane must call it, and the
// lanes' results are pooled: not for walks that keep per-lane minima.  Host builds flush
// each lane's records in turn (q_compact).
template <int M, bool PACKED>
QBX_HD void q_flush_wave(QBuf& b, QLane& q, const QCtx<M>& x, double& bestE, uint64_t& bestR) {
#if defined(__HIP_DEVICE_COMPILE__)
    // latency-bound and short: top issue priority, so that the other wave of the SIMD,
    // still walking, does not hold it up (the flush of the wave holding the answer ends
    // the solve)
    __builtin_amdgcn_s_setprio(3);
    q_compact<M, PACKED>(b, q, x, false, bestE, bestR);    // drop what no longer passes
    const int lane = (int)__lane_id();
    const int O = x.L - M;
    int32_t laneMin = wave_min_i32(q.laneMin);
    const int32_t thrB = wave_min_i32(q.thrB);
    auto thr_of = [&](int32_t lm) {
        const int32_t loc = sat_add_i32(lm, q.two_qm);
        int32_t t = thrB < loc ? thrB : loc;
        if (PACKED) t = t < QRP_THR_MAX ? t : QRP_THR_MAX;
        return t;
    };
    int32_t thr = thr_of(laneMin);
    uint64_t pending = __ballot(b.count > 0);
    while (pending) {
        const int s = __builtin_ctzll(pending);
        pending &= pending - 1ull;
        const int cnt = (int)readlane_u32((uint32_t)b.count, s);
        const uint64_t lam = readlane_u64(x.lam, s);
        const uint64_t p = lane_prefix(lam, x.L, x.B);
        const uint64_t laneRank = shl64(lam, x.L);
        const int SBk = (x.B - M) < 0 ? 0 : ((x.B - M) > O ? O : (x.B - M));
        const int parp = parity64(p);
        const uint32_t pmaskS = (parp && O > SBk) ? (lowmask32(O - SBk) << SBk) : 0u;
        for (int k = 0; k < cnt; ++k) {
            const uint32_t c = readlane_u32(b.c[k * b.stride], s);
            const int32_t m = (int32_t)readlane_u32((uint32_t)b.m[k * b.stride], s);
            if (m > thr) continue;
            for (uint32_t h = 0; h < (PACKED ? 2u : 1u); ++h) {
                uint32_t cf = c;
                if (PACKED) {
                    const uint32_t cp = c >> 2;
                    if (!((c >> h) & 1u)) continue;
                    cf = (cp << 1) | (h ^ (cp & 1u));
                }
                const uint32_t gc = cf ^ (cf >> 1);
                const uint64_t hs = (p << O) | (uint64_t)gc;
                uint32_t colq = 0u;
                double col;
                if (PACKED) block_columns_wave<M>(x.Wq, x.dq, x.W, x.d, x.n, hs, lane, colq, col);
                else col = block_column_wave<double, double, M>(x.W, x.d, x.n, hs, lane);
                if (PACKED) {
                    // the int32 filter's exact test: the block's least quantized energy
                    const bool set = lane >= M && lane < x.n && ((hs >> (lane - M)) & 1ull);
                    uint32_t e0 = set ? colq : 0u;
                    for (int o = 32; o > 0; o >>= 1) e0 += (uint32_t)wave_xor((int32_t)e0, o);
                    uint32_t fq[M];
#pragma unroll
                    for (int j = 0; j < M; ++j) fq[j] = readlane_u32(colq, j);
                    int32_t mn = 0x7fffffff;
#pragma unroll
                    for (int t = 0; t < ((1 << M) + 63) / 64; ++t) {
                        const uint32_t y = (uint32_t)lane + 64u * (uint32_t)t;
                        if (y < (1u << M)) {
                            const int32_t E = (int32_t)inner_energy<int32_t, uint32_t, M>(e0, fq, x.Wq, y);
                            mn = E < mn ? E : mn;
                        }
                    }
                    mn = wave_min_i32(mn);
                    if (mn < laneMin) {
                        laneMin = mn;
                        thr = thr_of(laneMin);
                    }
                    if (mn > thr) continue;
                }
                // the block in FP64: its states' energies and ranks, the wave's least pair
                const bool set = lane >= M && lane < x.n && ((hs >> (lane - M)) & 1ull);
                double e0 = set ? col : 0.0;
                for (int o = 32; o > 0; o >>= 1) e0 += wave_xor(e0, o);
                double f[M];
#pragma unroll
                for (int j = 0; j < M; ++j) f[j] = readlane_f64(col, j);
                const uint64_t rblk = laneRank | ((uint64_t)block_key(cf, SBk, pmaskS) << M);
                const int par = parp ^ parity32(gc);
                double E = __builtin_inf();
                uint64_t r = ~0ull;
#pragma unroll
                for (int t = 0; t < ((1 << M) + 63) / 64; ++t) {
                    const uint32_t y = (uint32_t)lane + 64u * (uint32_t)t;
                    if (y < (1u << M)) {
                        const double Ey = inner_energy<double, double, M>(e0, f, x.W, y);
                        const uint64_t ry = rblk | rank_lo<M>(y, x.B, par);
                        if (Ey < E || (Ey == E && ry < r)) { E = Ey; r = ry; }
                    }
                }
                for (int o = 32; o > 0; o >>= 1) {
                    const double E2 = wave_xor(E, o);
                    const uint64_t r2 = wave_xor(r, o);
                    if (E2 < E || (E2 == E && r2 < r)) { E = E2; r = r2; }
                }
                if (E < bestE || (E == bestE && r < bestR)) { bestE = E; bestR = r; }
            }
        }
    }
    b.count = 0;
#else
    q_compact<M, PACKED>(b, q, x, true, bestE, bestR);
#endif
}

// One block of an unrolled group of the filtered walk: static transition of the int32
// state, int32 block minimum, lane minimum and filter update, candidate record.
template <int M, int OMAX, int J>
QBX_HD bool group_step_q32(LaneI32<M, OMAX>& s, const int32_t* __restrict__ Wq, const int32_t* dlq,
                           uint32_t c0, uint32_t nblk, QLane& q, QBuf& b) {
    const uint32_t c = c0 + (uint32_t)J;
    if (J > 0) {
        if (c >= nblk) return false;
        constexpr int i = J > 0 ? __builtin_ctz((unsigned)J) : 0;
        const bool up = (i + 1 < GROUP_BITS) ? ((((unsigned)J >> (i + 1)) & 1u) == 0u)
                                             : (((c0 >> GROUP_BITS) & 1u) == 0u);
        // (no row prefetch here: its registers would cost this kernel its second wave)
        trans_static_i32<M, OMAX, i>(s, Wq + (M + i) * LDW, up);
    }
    const int32_t mn = block_min_i32<M>(s.E0, s.f, dlq);
    if (mn <= q.thr) {           // rare after the first blocks: record the block
        q.laneMin = mn < q.laneMin ? mn : q.laneMin;
        q_update_thr(q);
        b.c[b.count * b.stride] = c;
        b.m[b.count * b.stride] = mn;
        ++b.count;
    }
    return true;
}

template <int M, int OMAX, int... Js>
QBX_HD void group_q32(LaneI32<M, OMAX>& s, const int32_t* __restrict__ Wq, const int32_t* dlq,
                      uint32_t c0, uint32_t nblk, QLane& q, QBuf& b, std::integer_sequence<int, Js...>) {
    (void)(group_step_q32<M, OMAX, Js>(s, Wq, dlq, c0, nblk, q, b) && ...);
}

// Source of an upper bound on the answer shared by the lanes of one problem: refresh()
// returns an upper bound on the energy of some state (+inf if none yet); publish(e)
// offers one.  The host build keeps it in a plain double shared by the lanes it walks in
// turn; the device shares it across waves through the problem's reduction key.
struct HostBound {
    double b;
    QBX_HD_MEMBER double refresh() const { return b; }
    QBX_HD_MEMBER void publish(double e) { b = e < b ? e : b; }
    QBX_HD_MEMBER void note(double) {}
};

// Walk one lane of a float problem: int32 filter over quantized tables (Wq, dq, dlq in
// registers), FP64 exact re-walk (W, d, dld) of the blocks that still pass at the end.
// Returns the lane's exact (E, rank) minimum over the blocks it re-walked, which include
// every block that can hold a state of energy <= the problem's final bound (+inf, ~0 if
// it re-walked none).
//
// Why the filter is safe: with qm0 = n(n+1)/4, |Eq(x) - s E(x)| <= qm0 for every state x.
//   * shared bound b (an upper bound on some state's energy, so >= the answer E*): a
//     block holding a state x with E(x) <= b has int32 minimum <= s b + qm0 = thrB;
//   * lane minimum mq = Eq(xq): a state x of the lane with E(x) <= E(xq) has
//     Eq(x) <= s E(x) + qm0 <= s E(xq) + qm0 <= mq + 2 qm0.
// Both thresholds only decrease during the walk, so a block dropped at any time fails the
// final filter too.  The lane publishes (mq + qm0 + 1) / s, an upper bound on E(xq).
template <int M, int OMAX, bool COOP = false, class Bound>
QBX_HD LaneOutF64 walk_lane_q32(const int32_t* __restrict__ Wq, const int32_t* __restrict__ dq,
                                const double* __restrict__ W, const double* __restrict__ d,
                                int n, int L, int B, uint64_t lam, const int32_t* dlq,
                                const double* __restrict__ dld, double scale, double qm, Bound& src,
                                uint32_t* buf_c, int32_t* buf_m, int buf_stride) {
    static_assert(OMAX >= GROUP_BITS, "OMAX must cover the unrolled transition bits");
    const int O = L - M;
    LaneI32<M, OMAX> s;
    wave_priority_by_progress(0u, 1u);   // the set-up at the first quarter's priority
    lane_setup_i32<M, OMAX, true>(s, Wq, dq, n, L, B, lam);

    QCtx<M> x;
    x.set(W, d, dld, n, L, B, lam);

    double bound = src.refresh();
    src.note(bound);
    double bestE = __builtin_inf();
    uint64_t bestR = ~0ull;
    QLane q;
    q.laneMin = 0x7fffffff;
    q.two_qm = (int32_t)(2.0 * qm);
    q.thrB = q_threshold(bound, scale, qm);
    q_update_thr(q);
    QBuf b{buf_c, buf_m, buf_stride, 0};

    const uint32_t nblk = 1u << O;
    for (uint32_t c0 = 0;;) {
        wave_priority_by_progress(c0, nblk);
        group_q32<M, OMAX>(s, Wq, dlq, c0, nblk, q, b, std::make_integer_sequence<int, (1 << GROUP_BITS)>{});
        if (b.count > QCAP - (1 << GROUP_BITS)) {   // room for one more group of records
            q_compact<M>(b, q, x, false, bestE, bestR);
            if (b.count > QCAP - (1 << GROUP_BITS)) {
                QBX_STAT(++emu_stat_flushes);
                q_compact<M>(b, q, x, true, bestE, bestR);
            }
        }
        // share progress with the rest of the problem and take its running bound: after
        // the first group (when it tightens the most), then every QSYNC groups
        if ((c0 & ((QSYNC << GROUP_BITS) - 1u)) == 0u) {
            double mine = bestE;
            if (q.laneMin != 0x7fffffff) {
                const double ub = ((double)q.laneMin + qm) / scale;     // qm = qm0 + 1
                mine = ub < mine ? ub : mine;
            }
            src.publish(mine);
            const double b2 = src.refresh();
            src.note(b2);
            bound = b2 < bound ? b2 : bound;
            const double bb = bestE < bound ? bestE : bound;
            q.thrB = q_threshold(bb, scale, qm);
            q_update_thr(q);
        }
        c0 += 1u << GROUP_BITS;
        if (c0 >= nblk) break;
        const int i = ctz32(c0);
        const bool up = (((c0 ^ (c0 >> 1)) >> i) & 1u) != 0u;
        trans_dyn_i32<M, OMAX>(s, Wq + (M + i) * LDW, i, up);
    }
    if constexpr (COOP) q_flush_wave<M, false>(b, q, x, bestE, bestR);
    else q_compact<M>(b, q, x, true, bestE, bestR);
    LaneOutF64 out;
    out.E = bestE;
    out.rank = bestR;
    return out;
}

// ------------------------------------ float problems: packed 16-bit filter (rebased pairs)
// The filter of walk_lane_q32 on the rebased pair walk.  The lane state is the exact int32
// state of the quantized tables (bases E0 and E0 + up of the two blocks of a packed word,
// inner fields f, upper fields u: LaneRP, advanced by the integer rebased walk's
// transitions on Wq).  For each group of 16 packed blocks the inner fields are rounded to
// a coarse scale, 2^-sh of the quantized one, and both blocks of each word are walked as
// 16-bit coarse energies relative to their bases (PAIR_BIAS added), with coarse rows
// pc = rnd(Wq / 2^sh) for the group's transitions and coarse trie constants
// cc = rnd(c / 2^sh) (rnd: to nearest, halves up).
//
// Error bound.  For a state y of a block with exact quantized relative energy
// R(y) = sum_{k in y} f_k + sum_{trie path of y} c, the coarse value R_c(y) differs from
// R(y) / 2^sh by at most
//   * 1/2 per field of y for its rounding at the group's start (the high block's field
//     f_k + Wq_Mk is rounded as a whole),
//   * 1/2 per group bit flipped since the group's start, per field (a bit flipped twice
//     added and took away the same coarse row): at most 4 x 1/2 = 2,
//   * 1/2 per trie constant on y's path (at most M of them),
// so |R(y) / 2^sh - R_c(y)| <= M (1/2 + 2) + M / 2 = 3M <= QRP_ERR.  A block whose coarse
// minimum is m (unbiased) therefore holds no state with Eq < E0 + (m - QRP_ERR) 2^sh (the
// record's bound, tested against the threshold of walk_lane_q32) and holds a state with
// Eq <= E0 + (m + QRP_ERR) 2^sh (what the lane minimum takes).  A record that still passes
// when it is due for its FP64 walk first gets the int32 filter's own test, its exact
// minimum rebuilt from scratch (q_compact): the coarse walk is a pre-filter of the int32
// filter, and nothing is walked in FP64 that the int32 filter would not walk.
//
// Range.  Every coarse energy a packed group materialises is a state's, and
//     |R_c(y)| <= sum_k |fc_k(group start)| + D_c,
//     D_c = ceil(sum_{k<l<M} |Wq_kl| / 2^sh) + 4 + sum_k sum_{i<4} |pc_{M+1+i,k}|
//           + sum_k (|rnd(Wq_Mk / 2^sh)| + 2),
// so the guard sum_k |fc_k| <= PAIR_BOUND - D_c keeps every half in [0x0401, 0x7bff]; a
// wave any of whose lanes fails it walks the group with coarse values in int32 (fields
// rounded per block: error <= M / 2 + M / 2).  Float problems are quantized for
// |Eq| < 2^29 (q_scale), and thresholds are capped at QRP_THR_MAX, so threshold - base
// fits int32.
constexpr int32_t QRP_ERR = 21;

// v / 2^sh rounded to nearest, halves up (|v| <= 2^30, 0 <= sh <= 30)
QBX_HD int32_t rnd_shr(int32_t v, int sh) { return sh == 0 ? v : ((v + (1 << (sh - 1))) >> sh); }

// D_c of the guard for shift sh, or -1 when a coarse table value could leave 15 bits.
QBX_HD int64_t qrp_guard_D(const int32_t* Wq, int M, int L, int sh) {
    int64_t cin = 0, dg = 0, dp = 0;
    for (int k = 0; k < M; ++k)
        for (int l = k + 1; l < M; ++l) {
            const int64_t w = Wq[k * LDW + l];
            cin += w < 0 ? -w : w;
        }
    cin = ((cin + ((1ll << sh) - 1)) >> sh) + 4;
    for (int k = 0; k < M; ++k) {
        const int64_t wm = rnd_shr(Wq[M * LDW + k], sh);
        dp += (wm < 0 ? -wm : wm) + 2;
        for (int i = 0; i < GROUP_BITS; ++i) {
            const int J = M + 1 + i;
            if (J >= L) break;
            const int64_t w = rnd_shr(Wq[J * LDW + k], sh);
            dg += w < 0 ? -w : w;
        }
    }
    const int64_t D = cin + dg + dp;
    return D < PAIR_BOUND ? D : -1;
}

// Shift of the packed float filter: the smallest sh (finest coarse scale) for which the
// guard constant leaves room for the expected inner fields (the estimate of rp_worth_it,
// on the quantized tables), or -1 (then the int32 filter runs).  *D_out: D_c at sh.
// Split up so that the device can compute the terms and test the candidate shifts in
// parallel: qrp_shift_term() gives inner bit k's share of the estimate (the root mean
// square of its field, and its coupling sum, an exact integer), qrp_shift_first() the
// first shift tried (just below the estimate), qrp_shift_ok() the test of one shift.
// Summing the terms in k order reproduces the serial sums exactly.
QBX_HD void qrp_shift_term(const int32_t* Wq, const int32_t* dq, int n, int M, int k, double* rms, double* din) {
    double mean = dq[k], var = 0.0, dn = 0.0;
    for (int j = M; j < n; ++j) {
        const double w = Wq[k * LDW + j];
        mean += 0.5 * w;
        var += 0.25 * w * w;
    }
    for (int l = k + 1; l < M; ++l) dn += fabs((double)Wq[k * LDW + l]);
    for (int i = 0; i <= GROUP_BITS && M + i < n; ++i) dn += fabs((double)Wq[(M + i) * LDW + k]);
    *rms = sqrt(mean * mean + var);
    *din = dn;
}

QBX_HD int qrp_shift_first(double S, double Din, int M) {
    // coarse units scale as 2^-sh: start just below the estimate and walk up
    const double need = (Din + 1.5 * S) / (double)(PAIR_BOUND - 64 - 4 * M);
    int sh = 0;
    while (sh < 30 && (double)(1ll << sh) < need) ++sh;
    return sh > 1 ? sh - 1 : 0;
}

// first shift tried (-1: the packed filter cannot run) and the estimate's S
QBX_HD int qrp_shift_start(const int32_t* Wq, const int32_t* dq, int n, int M, int L, double* S_out) {
    if (L <= M + 1) return -1;
    double S = 0.0, Din = 0.0;
    for (int k = 0; k < M; ++k) {
        double r, dn;
        qrp_shift_term(Wq, dq, n, M, k, &r, &dn);
        S += r;
        Din += dn;
    }
    *S_out = S;
    return qrp_shift_first(S, Din, M);
}

QBX_HD bool qrp_shift_ok(const int32_t* Wq, int M, int L, int sh, double S, int64_t* D_out) {
    const int64_t D = qrp_guard_D(Wq, M, L, sh);
    if (D < 0) return false;
    *D_out = D;
    return (double)D + 1.5 * S / (double)(1ll << sh) <= (double)PAIR_BOUND;
}

QBX_HD int qrp_shift(const int32_t* Wq, const int32_t* dq, int n, int M, int L, int64_t* D_out) {
    double S = 0.0;
    const int sh0 = qrp_shift_start(Wq, dq, n, M, L, &S);
    if (sh0 < 0) return -1;
    for (int sh = sh0; sh <= 30; ++sh) {
        int64_t D = 0;
        if (qrp_shift_ok(Wq, M, L, sh, S, &D)) {
            *D_out = D;
            return sh;
        }
    }
    return -1;
}

// Candidate record of packed block c from the coarse minima of its halves (rare path): the
// halves whose coarse bound passes, their least lower bound E0_h + (m_h - bias - QRP_ERR)
// 2^sh as the record's minimum, and the lane minimum from the upper bound
// E0_h + (m_h - bias + QRP_ERR) 2^sh of the state realising m_h.
QBX_HD int32_t clamp_i32(int64_t v) {
    return v > 0x7fffffffLL ? 0x7fffffff : (v < -0x80000000LL ? (int32_t)0x80000000 : (int32_t)v);
}

template <int M, int OMAX>
QBX_HD void qrp_record(const LaneRP<M, OMAX>& s, uint32_t c, const int32_t* m, const int32_t* t, int32_t bias,
                       int sh, QLane& q, QBuf& b) {
    // at least one half passes (the caller's test); lb starts from the first that does, so
    // that no 64-bit sentinel stays live across the walk (the compiler kept one in scratch)
    uint32_t mask = 0;
    int64_t lb = 0;
    for (int h = 0; h < 2; ++h) {
        if (m[h] > t[h]) continue;
        const int64_t base = (int64_t)s.E0 + (h ? (int64_t)s.up : 0);
        const int64_t lo = base + ((int64_t)(m[h] - bias - QRP_ERR) << sh);
        const int32_t ub = clamp_i32(base + ((int64_t)(m[h] - bias + QRP_ERR) << sh));
        lb = (mask == 0u || lo < lb) ? lo : lb;
        mask |= 1u << h;
        q.laneMin = ub < q.laneMin ? ub : q.laneMin;
    }
    q_update_thr(q);
    q.thr = q.thr < QRP_THR_MAX ? q.thr : QRP_THR_MAX;
    b.c[b.count * b.stride] = (c << 2) | mask;
    b.m[b.count * b.stride] = clamp_i32(lb);
    ++b.count;
}

template <int M, int OMAX, int J>
QBX_HD bool group_step_qrp(LaneRP<M, OMAX>& s, uint32_t* F, const int32_t* __restrict__ Wq,
                           const uint32_t* __restrict__ Wpc, const uint32_t* dl, uint32_t c0, uint32_t nblk,
                           int sh, QLane& q, QBuf& b) {
    const uint32_t c = c0 + (uint32_t)J;
    if (J > 0) {
        if (c >= nblk) return false;
        constexpr int i = J > 0 ? __builtin_ctz((unsigned)J) : 0;
        const bool up = (i + 1 < GROUP_BITS) ? ((((unsigned)J >> (i + 1)) & 1u) == 0u)
                                             : (((c0 >> GROUP_BITS) & 1u) == 0u);
        // the row through the scalar cache (no prefetch: its VGPRs would cost this kernel
        // its second wave)
        WRowRP<M> row;
        load_wrow_rp_uniform<M, i>(row, Wq, Wpc);
        trans_row_rp<M, OMAX, i>(s, F, row, up);
    }
    const uint32_t mp = block_min_pair<M, false>((uint32_t)PAIR_BIAS * 65537u, F, dl);
    // half h passes when E0_h + (m_h - PAIR_BIAS - QRP_ERR) 2^sh <= thr, i.e. when
    // m_h <= floor((thr - E0_h) / 2^sh) + PAIR_BIAS + QRP_ERR
    int32_t m[2], t[2];
    m[0] = (int32_t)(mp & 0xffffu);
    m[1] = (int32_t)(mp >> 16);
    t[0] = ((q.thr - s.E0) >> sh) + (PAIR_BIAS + QRP_ERR);
    t[1] = ((q.thr - (s.E0 + s.up)) >> sh) + (PAIR_BIAS + QRP_ERR);
    if (__builtin_expect(m[0] <= t[0] || m[1] <= t[1], 0)) qrp_record<M, OMAX>(s, c, m, t, PAIR_BIAS, sh, q, b);
    return true;
}

template <int M, int OMAX, int... Js>
QBX_HD void group_qrp(LaneRP<M, OMAX>& s, uint32_t* F, const int32_t* __restrict__ Wq,
                      const uint32_t* __restrict__ Wpc, const uint32_t* dl, uint32_t c0, uint32_t nblk, int sh,
                      QLane& q, QBuf& b, std::integer_sequence<int, Js...>) {
    (void)(group_step_qrp<M, OMAX, Js>(s, F, Wq, Wpc, dl, c0, nblk, sh, q, b) && ...);
}

// The group at c0 with coarse values in int32 (a wave whose guard failed): the fine state
// advances block by block as in the integer fallback, and each block's halves get their
// fields rounded from the fine ones.
template <int M, int OMAX>
QBX_HD void group_qrp_int32(LaneRP<M, OMAX>& s, const int32_t* __restrict__ Wq, const uint32_t* dl,
                            uint32_t c0, uint32_t nblk, int sh, QLane& q, QBuf& b) {
    const int32_t* WM = Wq + M * LDW;
    const uint32_t cend = (c0 + (1u << GROUP_BITS) < nblk) ? c0 + (1u << GROUP_BITS) : nblk;
#pragma unroll 1
    for (uint32_t c = c0; c < cend; ++c) {
        if (c > c0) {
            const int i = ctz32(c - c0);
            const bool up = (((c ^ (c >> 1)) >> i) & 1u) != 0u;
            trans_grp_rp<M, OMAX>(s, Wq + (M + 1 + i) * LDW, i, up);
        }
        int32_t m[2], t[2];
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            int32_t fh[M];
#pragma unroll
            for (int k = 0; k < M; ++k) fh[k] = rnd_shr((int32_t)((uint32_t)s.f[k] + (h ? (uint32_t)WM[k] : 0u)), sh);
            int32_t mn = 0;
            uint32_t z = 0u;
            QBX_OPAQUE(z);
            utrie_even<M, 0u>(0u, fh, dl, z, mn);
            const int32_t base = (int32_t)((uint32_t)s.E0 + (h ? (uint32_t)s.up : 0u));
            m[h] = mn;
            t[h] = ((q.thr - base) >> sh) + QRP_ERR;
        }
        if (m[0] <= t[0] || m[1] <= t[1]) qrp_record<M, OMAX>(s, c, m, t, 0, sh, q, b);
    }
}

// Walk one lane of a float problem with the packed filter (tables: Wq, dq and the coarse
// packed Wpc, dlpc of shift sh; slack = PAIR_BOUND - D_c); same contract and FP64 re-walks
// as walk_lane_q32.  groups_int32 (nullable) counts the groups walked in int32.
template <int M, int OMAX, bool COOP = false, class Bound>
QBX_HD LaneOutF64 walk_lane_qrp(const int32_t* __restrict__ Wq, const int32_t* __restrict__ dq,
                                const uint32_t* __restrict__ Wpc, const double* __restrict__ W,
                                const double* __restrict__ d, int n, int L, int B, uint64_t lam,
                                const uint32_t* dlpc, const int32_t* __restrict__ dlq,
                                const double* __restrict__ dld, double scale, double qm, int sh, int32_t slack,
                                Bound& src, uint32_t* buf_c, int32_t* buf_m, int buf_stride,
                                uint32_t* groups_int32) {
    static_assert(OMAX > GROUP_BITS, "OMAX must cover the unrolled transition bits");
    const int O = L - M;
    LaneRP<M, OMAX> s;
    wave_priority_by_progress(0u, 1u);   // the set-up at the first quarter's priority
    lane_setup_rp<M, OMAX, true>(s, Wq, dq, n, L, B, lam);

    QCtx<M> x;
    x.set(W, d, dld, n, L, B, lam);
    x.Wq = Wq;
    x.dq = dq;
    x.dlq = dlq;

    double bound = src.refresh();
    src.note(bound);
    double bestE = __builtin_inf();
    uint64_t bestR = ~0ull;
    QLane q;
    q.laneMin = 0x7fffffff;
    q.two_qm = (int32_t)(2.0 * qm);
    q.thrB = q_threshold(bound, scale, qm);
    q_update_thr(q);
    q.thr = q.thr < QRP_THR_MAX ? q.thr : QRP_THR_MAX;
    QBuf b{buf_c, buf_m, buf_stride, 0};

    const int32_t* WM = Wq + M * LDW;
    const int32_t* W3 = Wq + (M + GROUP_BITS) * LDW;    // row of group bit GROUP_BITS - 1
    const uint32_t nblk = 1u << (O - 1);
    for (uint32_t c0 = 0;;) {
        wave_priority_by_progress(c0, nblk);
        int32_t fc[M];
        uint32_t g = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            fc[k] = rnd_shr(s.f[k], sh);
            const uint32_t a = (uint32_t)(fc[k] < 0 ? -fc[k] : fc[k]);
            g += a < 0x10000u ? a : 0x10000u;
        }
        bool ok = g <= (uint32_t)slack;
#if defined(__HIP_DEVICE_COMPILE__)
        ok = __all(ok);
#endif
        if (__builtin_expect(ok, 1)) {
            uint32_t F[M];
#pragma unroll
            for (int k = 0; k < M; ++k)
                F[k] = (uint32_t)fc[k] + ((uint32_t)rnd_shr((int32_t)((uint32_t)s.f[k] + (uint32_t)WM[k]), sh) << 16);
            group_qrp<M, OMAX>(s, F, Wq, Wpc, dlpc, c0, nblk, sh, q, b,
                               std::make_integer_sequence<int, (1 << GROUP_BITS)>{});
            // the fine fields follow the group's net change: a full group flips group bit
            // GROUP_BITS - 1 once (every lower bit an even number of times); after a partial
            // group (the lane's last) they are not used again
            if (c0 + (1u << GROUP_BITS) <= nblk) {
                const bool up = ((c0 >> GROUP_BITS) & 1u) == 0u;
#pragma unroll
                for (int k = 0; k < M; ++k)
                    s.f[k] = (int32_t)(up ? (uint32_t)s.f[k] + (uint32_t)W3[k] : (uint32_t)s.f[k] - (uint32_t)W3[k]);
            }
        } else {
            group_qrp_int32<M, OMAX>(s, Wq, dlpc, c0, nblk, sh, q, b);
            if (groups_int32) ++*groups_int32;
        }
        if (b.count > QCAP - (1 << GROUP_BITS)) {   // room for one more group of records
            q_compact<M, true>(b, q, x, false, bestE, bestR);
            if (b.count > QCAP - (1 << GROUP_BITS)) {
                QBX_STAT(++emu_stat_flushes);
                q_compact<M, true>(b, q, x, true, bestE, bestR);
            }
        }
        if ((c0 & ((QSYNC << GROUP_BITS) - 1u)) == 0u) {
            double mine = bestE;
            if (q.laneMin != 0x7fffffff) {
                const double ub = ((double)q.laneMin + qm) / scale;
                mine = ub < mine ? ub : mine;
            }
            src.publish(mine);
            const double b2 = src.refresh();
            src.note(b2);
            bound = b2 < bound ? b2 : bound;
            const double bb = bestE < bound ? bestE : bound;
            q.thrB = q_threshold(bb, scale, qm);
            q_update_thr(q);
            q.thr = q.thr < QRP_THR_MAX ? q.thr : QRP_THR_MAX;
        }
        c0 += 1u << GROUP_BITS;
        if (c0 >= nblk) break;
        const int i = ctz32(c0);
        const bool up = (((c0 ^ (c0 >> 1)) >> i) & 1u) != 0u;
        trans_dyn_rp<M, OMAX>(s, Wq + (M + 1 + i) * LDW, i, up);
    }
    wave_priority_by_progress(nblk, nblk);
    if constexpr (COOP) q_flush_wave<M, true>(b, q, x, bestE, bestR);
    else q_compact<M, true>(b, q, x, true, bestE, bestR);
    LaneOutF64 out;
    out.E = bestE;
    out.rank = bestR;
    return out;
}

// Order-preserving 64-bit keys for the per-instance atomic minimum.
QBX_HD uint64_t key_of_i64(int64_t E) { return (uint64_t)E ^ 0x8000000000000000ull; }
QBX_HD uint64_t key_of_f64(double E) {
    union { double d; uint64_t u; } v;
    v.d = E;
    return (v.u & 0x8000000000000000ull) ? ~v.u : (v.u | 0x8000000000000000ull);
}

}  // namespace qbx
