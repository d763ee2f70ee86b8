// qbx_gen.h -- RandomQUBO matrices generated where they are solved, bit for bit as the
// reference generates them on the host (src/qubo_lib/problems/random_qubo.py:19-33:
// Generator(PCG64(seed)).integers(low, high, (N, N), dtype) for integer dtypes,
// low + random((N, N), dtype) * (high - low) for float dtypes).  Shared by the device
// kernel and the host build used by the CPU tests.
//
// The bit generator is PCG64 (128-bit LCG, XSL-RR output): each draw first advances
// s <- s * M + inc (mod 2^128), then outputs rotr64(hi(s) ^ lo(s), s >> 122).  The host
// seeds it (SeedSequence) and passes the 128-bit state and increment.  Draws narrower than
// 64 bits come from buffers, low half first: 32-bit values split one 64-bit output; 16-bit
// values (int16 bounds) split one 32-bit value.  Bounded integers use Lemire's
// multiply-and-reject method on the draw width (16 bits for int16, 32 bits for int32 and
// int64 with a range below 2^32); a range of one value draws nothing.  Uniform floats are
// (u32 >> 8) 2^-24 (float32) and (u64 >> 11) 2^-53 (float64), scaled with one rounded
// multiply and one rounded add (not a fused multiply-add), as numpy evaluates it.  All of this was
// checked against numpy 2.2's Generator on the CPU (tests/test_generators_device.py).
#pragma once
#include <math.h>
#include <stdint.h>

#include "qbx_walk.h"

namespace qbx {

enum { GEN_INT16 = 16, GEN_INT32 = 32, GEN_FLOAT32 = -32, GEN_FLOAT64 = -64 };

struct Pcg64 {
    uint64_t hi, lo, inc_hi, inc_lo;
    uint64_t buf32;   // pending upper half of a 64-bit output
    uint32_t buf16;   // pending upper half of a 32-bit value
    int has32, has16;
};

QBX_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

QBX_HD uint64_t pcg64_next(Pcg64& g) {
    const uint64_t Mhi = 0x2360ED051FC65DA4ull, Mlo = 0x4385DF649FCCF645ull;
    const uint64_t lo = g.lo * Mlo;
    const uint64_t hi = mulhi64(g.lo, Mlo) + g.lo * Mhi + g.hi * Mlo;
    const uint64_t nlo = lo + g.inc_lo;
    const uint64_t nhi = hi + g.inc_hi + (nlo < lo ? 1ull : 0ull);
    g.lo = nlo;
    g.hi = nhi;
    const uint64_t x = nhi ^ nlo;
    const unsigned r = (unsigned)(nhi >> 58);
    return (x >> r) | (x << ((64u - r) & 63u));
}

QBX_HD uint32_t pcg64_u32(Pcg64& g) {
    if (g.has32) {
        g.has32 = 0;
        return (uint32_t)g.buf32;
    }
    const uint64_t v = pcg64_next(g);
    g.buf32 = v >> 32;
    g.has32 = 1;
    return (uint32_t)v;
}

QBX_HD uint16_t pcg64_u16(Pcg64& g) {
    if (g.has16) {
        g.has16 = 0;
        return (uint16_t)g.buf16;
    }
    const uint32_t v = pcg64_u32(g);
    g.buf16 = v >> 16;
    g.has16 = 1;
    return (uint16_t)v;
}

// numpy's Generator(PCG64(seed)) for an integer seed 0 <= seed < 2^64, on the host: the
// SeedSequence hash of the seed's 32-bit words (numpy/random/bit_generator.pyx: pool of
// four words, hashmix / mix with its published constants, generate_state(4, uint64) as
// eight 32-bit words, low word first), then PCG64's seeding (pcg64.h pcg64_set_seed:
// state = 0, inc = (seq << 1) | 1, step, state += initstate, step) with
// initstate = val[0]:val[1] and seq = val[2]:val[3].  Pinned against numpy on many seeds.
inline uint32_t seedseq_hashmix(uint32_t v, uint32_t& h) {
    v ^= h;
    h *= 0x931e8875u;
    v *= h;
    return v ^ (v >> 16);
}
inline uint32_t seedseq_mix(uint32_t x, uint32_t y) {
    const uint32_t r = 0xca01f9ddu * x - 0x4973f715u * y;
    return r ^ (r >> 16);
}
inline void pcg64_step(Pcg64& g) {
    const uint64_t Mhi = 0x2360ED051FC65DA4ull, Mlo = 0x4385DF649FCCF645ull;
    const unsigned __int128 s = ((unsigned __int128)g.hi << 64) | g.lo;
    const unsigned __int128 inc = ((unsigned __int128)g.inc_hi << 64) | g.inc_lo;
    const unsigned __int128 t = s * (((unsigned __int128)Mhi << 64) | Mlo) + inc;
    g.hi = (uint64_t)(t >> 64);
    g.lo = (uint64_t)t;
}
inline void pcg64_from_seed(uint64_t seed, Pcg64& g) {
    uint32_t ent[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const int n_ent = (seed >> 32) ? 2 : 1;
    uint32_t pool[4];
    uint32_t h = 0x43b0d7e5u;
    for (int i = 0; i < 4; ++i) pool[i] = seedseq_hashmix(i < n_ent ? ent[i] : 0u, h);
    for (int src = 0; src < 4; ++src)
        for (int dst = 0; dst < 4; ++dst)
            if (src != dst) pool[dst] = seedseq_mix(pool[dst], seedseq_hashmix(pool[src], h));
    uint32_t w[8];
    uint32_t hb = 0x8b51f9ddu;
    for (int i = 0; i < 8; ++i) {
        uint32_t v = pool[i & 3] ^ hb;
        hb *= 0x58f38dedu;
        v *= hb;
        w[i] = v ^ (v >> 16);
    }
    uint64_t val[4];
    for (int k = 0; k < 4; ++k) val[k] = (uint64_t)w[2 * k] | ((uint64_t)w[2 * k + 1] << 32);
    g.hi = g.lo = 0;
    const unsigned __int128 seq = ((unsigned __int128)val[2] << 64) | val[3];
    const unsigned __int128 inc = (seq << 1) | 1u;
    g.inc_hi = (uint64_t)(inc >> 64);
    g.inc_lo = (uint64_t)inc;
    pcg64_step(g);
    const unsigned __int128 s = (((unsigned __int128)g.hi << 64) | g.lo) + (((unsigned __int128)val[0] << 64) | val[1]);
    g.hi = (uint64_t)(s >> 64);
    g.lo = (uint64_t)s;
    pcg64_step(g);
    g.buf32 = 0;
    g.buf16 = 0;
    g.has32 = g.has16 = 0;
}

// value in [0, rng] (rng = high - low - 1 >= 1)
QBX_HD uint32_t lemire16(Pcg64& g, uint32_t rng) {
    const uint32_t excl = rng + 1u;
    uint32_t m = (uint32_t)pcg64_u16(g) * excl;
    uint32_t left = m & 0xffffu;
    if (left < excl) {
        const uint32_t thr = (0x10000u - excl) % excl;
        while (left < thr) {
            m = (uint32_t)pcg64_u16(g) * excl;
            left = m & 0xffffu;
        }
    }
    return m >> 16;
}

QBX_HD uint64_t lemire32(Pcg64& g, uint64_t rng) {
    const uint64_t excl = rng + 1ull;
    uint64_t m = (uint64_t)pcg64_u32(g) * excl;
    uint64_t left = m & 0xffffffffull;
    if (left < excl) {
        const uint64_t thr = (0x100000000ull - excl) % excl;
        while (left < thr) {
            m = (uint64_t)pcg64_u32(g) * excl;
            left = m & 0xffffffffull;
        }
    }
    return m >> 32;
}

// low + r * span with two roundings, never fused into one multiply-add (which rounds once
// and differs from numpy in the last bit): contraction is switched off for these
// functions on the device compiler, and the host build uses -ffp-contract=off.
QBX_HD float gen_scale_f32(float lo, float r, float span) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const float t = r * span;
    return lo + t;
}
QBX_HD double gen_scale_f64(double lo, double r, double span) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const double t = r * span;
    return lo + t;
}

// True when (low, high, dtype) is a supported generator configuration.
QBX_HD bool gen_valid(int64_t low, int64_t high, int dtype) {
    if (dtype == GEN_INT16) return high > low && low >= -32768 && high <= 32768;
    if (dtype == GEN_INT32) return high > low && low >= -2147483648ll && high <= 2147483648ll;
    if (dtype == GEN_FLOAT32 || dtype == GEN_FLOAT64)
        return low >= -(1ll << 53) && high <= (1ll << 53) && low <= high;
    return false;
}

// One matrix, row-major, in the element type the solver takes (int32 for integer dtypes,
// double for float dtypes: float32 values convert exactly).
template <class T>
QBX_HD void gen_random_qubo(Pcg64 g, int n, int64_t low, int64_t high, int dtype, T* out) {
    g.has32 = 0;
    g.has16 = 0;
    const uint64_t cells = (uint64_t)n * (uint64_t)n;
    if (dtype == GEN_INT16 || dtype == GEN_INT32) {
        const uint64_t rng = (uint64_t)(high - low) - 1ull;
        for (uint64_t i = 0; i < cells; ++i) {
            uint64_t v = 0;
            if (rng == 0) v = 0;
            else if (dtype == GEN_INT16) v = lemire16(g, (uint32_t)rng);
            else v = lemire32(g, rng);
            out[i] = (T)(low + (int64_t)v);
        }
    } else if (dtype == GEN_FLOAT32) {
        const float lo = (float)low, span = (float)(high - low);
        for (uint64_t i = 0; i < cells; ++i) {
            const float r = (float)(pcg64_u32(g) >> 8) * 5.9604644775390625e-08f;   // 2^-24
            out[i] = (T)gen_scale_f32(lo, r, span);
        }
    } else {
        const double lo = (double)low, span = (double)(high - low);
        for (uint64_t i = 0; i < cells; ++i) {
            const double r = (double)(pcg64_next(g) >> 11) * 1.1102230246251565e-16;   // 2^-53
            out[i] = (T)gen_scale_f64(lo, r, span);
        }
    }
}

// ---------------------------------------------------------------------------- TSP
// TSP.init_random_instance(seed, n_cities, low, high, dtype).build_Q_matrix()
// (src/qubo_lib/problems/tsp.py:50-59, 67-75, 272-321), bit for bit, in two parts.
//
// tsp_distances (host): the cities and their distances.
//   * cities: rng.uniform(low, high, (n_cities, 2)) -- per city x then y, each
//     low + (high - low) * next_double, two roundings -- cast to dtype (C: float or double);
//   * distances in C, element by element as tsp.py:279 evaluates them on numpy scalars:
//     sqrt((x_i - x_j) ** 2 + (y_i - y_j) ** 2).  numpy's scalar `**` is the C library's
//     powf / pow, and pow(x, 2) is not always the correctly rounded x * x (glibc's powf and
//     pow are accurate to < 1 ulp, not correctly rounded): about one square in 1500 of
//     float32 coordinates in [0, 100) differs in the last bit.  So the squares are taken
//     with the C library's own pow, on the host, with the exponent read through a volatile
//     so that the compiler cannot turn pow(x, 2) into x * x; the sum is one rounding and
//     sqrtf / sqrt are correctly rounded (IEEE 754), as numpy's np.sqrt.
//   Stored as double, the upper triangle row by row (i < j), TSP_MAX_PAIRS per problem.
//
// tsp_qubo (host and device): Q in double over index a = city * n + time: tour cost
//   w[i][j] at ((i, p), (j, p + 1)), plus A * C with A = 2 max(w) + 1 and C = -1 on the
//   diagonal, +1 for the same time or the same city; folded to the upper triangle
//   (Q_ab + Q_ba above the diagonal, one rounding); cast to dtype and written as double
//   (the solver's element type).  The multiply and the add of A * C + cost are rounded
//   separately (no fused multiply-add), as numpy does them.
constexpr int TSP_MAX_CITIES = 7;
constexpr int TSP_MAX_PAIRS = TSP_MAX_CITIES * (TSP_MAX_CITIES - 1) / 2;

QBX_HD double gen_next_double(Pcg64& g) { return (double)(pcg64_next(g) >> 11) * 1.1102230246251565e-16; }

// host only (plain functions: the device pass parses them and emits nothing)
inline volatile float tsp_two_f = 2.0f;     // exponents the compiler cannot see: pow stays a call
inline volatile double tsp_two_d = 2.0;
inline float tsp_square(float x) { return powf(x, tsp_two_f); }
inline double tsp_square(double x) { return pow(x, tsp_two_d); }
inline float tsp_sqrt(float x) { return sqrtf(x); }
inline double tsp_sqrt(double x) { return sqrt(x); }

template <class C>
inline void tsp_distances(Pcg64 g, int nc, double low, double high, double* d) {
    C xs[TSP_MAX_CITIES], ys[TSP_MAX_CITIES];
    const double span = high - low;
    for (int i = 0; i < nc; ++i) {
        xs[i] = (C)gen_scale_f64(low, gen_next_double(g), span);
        ys[i] = (C)gen_scale_f64(low, gen_next_double(g), span);
    }
    int k = 0;
    for (int i = 0; i < nc; ++i)
        for (int j = i + 1; j < nc; ++j) {
            const C sx = tsp_square((C)(xs[i] - xs[j]));
            const C sy = tsp_square((C)(ys[i] - ys[j]));
            d[k++] = (double)tsp_sqrt((C)(sx + sy));
        }
}

QBX_HD double tsp_penalty_sum(double q, double A, int c) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const double t = A * (double)c;
    return q + t;
}

template <class C>
QBX_HD void tsp_qubo(int nc, const double* d, double* out) {
    double w[TSP_MAX_CITIES][TSP_MAX_CITIES];
    double wmax = 0.0;
    int k = 0;
    for (int i = 0; i < nc; ++i) {
        w[i][i] = 0.0;
        for (int j = i + 1; j < nc; ++j) {
            w[i][j] = w[j][i] = d[k];
            wmax = d[k] > wmax ? d[k] : wmax;
            ++k;
        }
    }
    const double A = 2.0 * wmax + 1.0;
    const int N = nc * nc;
    // full (unfolded) entry (a, b): tour cost plus A times the constraint matrix
    auto full = [&](int a, int b) -> double {
        const int ci = a / nc, pa = a - ci * nc, cj = b / nc, pb = b - cj * nc;
        const double cost = (pb == pa + 1) ? w[ci][cj] : 0.0;
        int c = 0;
        if (a == b) c = -1;
        else if (pa == pb || ci == cj) c = 1;
        return tsp_penalty_sum(cost, A, c);
    };
    for (int a = 0; a < N; ++a) {
        for (int b = 0; b < N; ++b) {
            double v = 0.0;
            if (a == b) v = full(a, a);
            else if (a < b) v = full(a, b) + full(b, a);
            out[(size_t)a * N + b] = (double)(C)v;
        }
    }
}

}  // namespace qbx
