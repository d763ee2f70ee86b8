/*
 * qbf_oracle.c -- CPU oracle for the exhaustive QUBO path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls this
 * file; it is loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, always as the checker, never as the thing measured or shipped.
 *
 * Two independent methods, each in an integer (int64, exact) and a double flavour:
 *
 *  oracle_qbf_*   restates the reference QBF algorithm step by step
 *                 (src/qubo_lib/solvers/qbf.py):
 *                   - A/B split: A = n - B                      qbf.py:429-446
 *                   - diag_BB = energies of Q_1B over 2^B states qbf.py:246-254, 516-548
 *                   - ABBA = (Q_0B + Q_1A^T) @ X_B^T             qbf.py:264-289
 *                   - Gray walk over prefixes, chunk +-= ABBA row qbf.py:212-217, 291-325
 *                   - argmin of the chunk (lowest suffix index)  qbf.py:229
 *                   - prefix_E = b^T Q_0A b, strict '<' against a
 *                     running glob_E that starts at 0 with (0,0)  qbf.py:395-412, 98-100
 *                   - correction every `correction_period` steps qbf.py:219-244
 *                 Accumulation is int64 (exact) for integer Q, where the reference
 *                 accumulates in Q.dtype (qbf.py:369, 281) and may wrap; the two agree
 *                 whenever the reference does not overflow.  The double flavour
 *                 accumulates in double throughout.
 *
 *  oracle_brute_* an independent brute force: every state is visited by a Gray walk
 *                 with full local-field updates, and the answer is the
 *                 lexicographic minimum of (E, rank(s)) where
 *                 rank(s) = (invgray(s >> B) << B) | (s & (2^B - 1)).
 *                 That key is the reference's tie rule (earliest prefix in Gray
 *                 visiting order, then lowest suffix index; qbf.py:229, 397-412).
 *
 * State integer convention (qbf.py:373-393): x[0] is the most significant bit of
 * the state integer s, x[n-1] the least significant.  Energy E(x) = x^T Q x over the
 * full, possibly non-symmetric Q (qbf.py:264, 499).
 *
 * Pinning (tests/test_oracle.py): both methods are checked against an
 * itertools brute force for n <= 12, against each other on random instances,
 * against the MaxCut 'test' known answer (E = -19, best_solutions.csv in the
 * reference's MaxCut.zip) and against the showcase.ipynb golden solutions.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static uint64_t invgray64(uint64_t g) {
    g ^= g >> 1; g ^= g >> 2; g ^= g >> 4; g ^= g >> 8; g ^= g >> 16; g ^= g >> 32;
    return g;
}
static uint64_t lowmask(int b) { return b >= 64 ? ~0ull : ((1ull << b) - 1ull); }
static uint64_t shr(uint64_t v, int b) { return b >= 64 ? 0 : v >> b; }
static uint64_t shl(uint64_t v, int b) { return b >= 64 ? 0 : v << b; }

uint64_t oracle_rank_of_state(uint64_t s, int B) {
    return shl(invgray64(shr(s, B)), B) | (s & lowmask(B));
}

/* ----------------------------------------------------------------------------
 * Restatement of QBF (qbf.py).  Memory: (A + 1) * 2^B int64.
 * Returns 0 on success, -1 on bad arguments / allocation failure.
 * ------------------------------------------------------------------------- */

/* energy of sub-matrix Q[off:off+m, off:off+m] for every m-bit state, indexed by
 * the state integer (MSB = first variable).  qbf.py:516-548 builds the same table
 * recursively; for integer Q its values are exact, so a direct Gray walk over the
 * state table is an equivalent restatement. */
static void energy_table_i64(const int64_t *Q, int n, int off, int m, int64_t *out) {
    uint64_t cnt = 1ull << m;
    int64_t *field = (int64_t *)calloc((size_t)m, sizeof(int64_t));
    /* field[v] = Q_vv + sum_{w != v} (Q_vw + Q_wv) x_w  for local variable v */
    for (int v = 0; v < m; ++v) field[v] = Q[(size_t)(off + v) * n + (off + v)];
    int64_t E = 0;
    uint64_t s = 0;
    out[0] = 0;
    for (uint64_t i = 1; i < cnt; ++i) {
        int bit = __builtin_ctzll(i);            /* integer bit that flips */
        int v = m - 1 - bit;                     /* variable index (MSB first) */
        uint64_t g = i ^ (i >> 1);
        int up = (int)((g >> bit) & 1ull);
        if (up) E += field[v]; else E -= field[v];
        for (int w = 0; w < m; ++w) {
            if (w == v) continue;
            int64_t c = Q[(size_t)(off + v) * n + (off + w)] + Q[(size_t)(off + w) * n + (off + v)];
            field[w] += up ? c : -c;
        }
        s = g;
        out[s] = E;
    }
    free(field);
}

int oracle_qbf_i64(const int64_t *Q, int n, int B, int correction_period,
                   int64_t *E_out, uint64_t *x_out) {
    if (n <= 0 || n > 40 || B < 0 || B > n || B > 26) return -1;
    (void)correction_period; /* exact integers: a correction recomputes identical values */
    const int A = n - B;
    const uint64_t NB = 1ull << B;
    int64_t *diag_BB = (int64_t *)malloc(sizeof(int64_t) * NB);
    int64_t *ABBA = (int64_t *)malloc(sizeof(int64_t) * NB * (A > 0 ? A : 1));
    int64_t *chunk = (int64_t *)malloc(sizeof(int64_t) * NB);
    if (!diag_BB || !ABBA || !chunk) { free(diag_BB); free(ABBA); free(chunk); return -1; }

    /* diag_BB: energies of Q_1B = Q[A:, A:]   (qbf.py:246-254) */
    energy_table_i64(Q, n, A, B, diag_BB);

    /* ABBA[i, s] = sum_j (Q_0B + Q_1A^T)[i, j] * X_B[s, j]   (qbf.py:264-289);
     * X_B[s, j] = bit (B-1-j) of s  (fill_all_state_matrix, qbf.py:373-384) */
    for (int i = 0; i < A; ++i) {
        for (uint64_t s = 0; s < NB; ++s) {
            int64_t acc = 0;
            for (int j = 0; j < B; ++j)
                if ((s >> (B - 1 - j)) & 1ull)
                    acc += Q[(size_t)i * n + (A + j)] + Q[(size_t)(A + j) * n + i];
            ABBA[(size_t)i * NB + s] = acc;
        }
    }

    memcpy(chunk, diag_BB, sizeof(int64_t) * NB);
    int64_t glob_E = 0;
    uint64_t glob_pref = 0, glob_X = 0;
    int64_t bit_array[64];
    memset(bit_array, 0, sizeof(bit_array));
    uint64_t prev_gray = 0;
    const uint64_t NA = 1ull << A;
    for (uint64_t i = 0; i < NA; ++i) {
        uint64_t gray = i ^ (i >> 1);
        if (i > 0) {
            /* gray_code_iterator (qbf.py:314-319) and the row update (qbf.py:213-217) */
            int sign = gray > prev_gray ? 1 : -1;
            int flipped = __builtin_ctzll(gray ^ prev_gray);
            int row = A - flipped - 1;
            bit_array[row] += sign;
            const int64_t *r = ABBA + (size_t)row * NB;
            if (sign == 1) for (uint64_t s = 0; s < NB; ++s) chunk[s] += r[s];
            else           for (uint64_t s = 0; s < NB; ++s) chunk[s] -= r[s];
            prev_gray = gray;
        }
        /* argmin: first index of the minimum (numpy argmin, qbf.py:229) */
        uint64_t am = 0;
        int64_t mn = chunk[0];
        for (uint64_t s = 1; s < NB; ++s) if (chunk[s] < mn) { mn = chunk[s]; am = s; }
        /* prefix_E = b^T Q_0A b   (qbf.py:409) */
        int64_t prefix_E = 0;
        for (int a = 0; a < A; ++a) if (bit_array[a])
            for (int b = 0; b < A; ++b) if (bit_array[b]) prefix_E += Q[(size_t)a * n + b];
        if (gray == 0) { glob_E = 0; glob_pref = 0; glob_X = 0; }   /* qbf.py:397-401 */
        if (mn + prefix_E < glob_E) { glob_E = mn + prefix_E; glob_pref = gray; glob_X = am; }
    }
    *E_out = glob_E;
    *x_out = shl(glob_pref, B) | glob_X;
    free(diag_BB); free(ABBA); free(chunk);
    return 0;
}

/* double flavour of the same restatement; the correction of qbf.py:232-244 is a
 * re-computation chunk = diag_BB + x @ ABBA, done here in double. */
int oracle_qbf_f64(const double *Q, int n, int B, int correction_period,
                   double *E_out, uint64_t *x_out) {
    if (n <= 0 || n > 40 || B < 0 || B > n || B > 26) return -1;
    if (correction_period <= 0) correction_period = 1024;
    const int A = n - B;
    const uint64_t NB = 1ull << B;
    double *diag_BB = (double *)malloc(sizeof(double) * NB);
    double *ABBA = (double *)malloc(sizeof(double) * NB * (A > 0 ? A : 1));
    double *chunk = (double *)malloc(sizeof(double) * NB);
    double *field = (double *)calloc((size_t)(B > 0 ? B : 1), sizeof(double));
    if (!diag_BB || !ABBA || !chunk || !field) { free(diag_BB); free(ABBA); free(chunk); free(field); return -1; }

    /* diag_BB by a Gray walk (exact re-derivation per 4096 steps keeps drift away) */
    {
        for (int v = 0; v < B; ++v) field[v] = Q[(size_t)(A + v) * n + (A + v)];
        double E = 0.0;
        diag_BB[0] = 0.0;
        for (uint64_t i = 1; i < NB; ++i) {
            int bit = __builtin_ctzll(i);
            int v = B - 1 - bit;
            uint64_t g = i ^ (i >> 1);
            int up = (int)((g >> bit) & 1ull);
            E += up ? field[v] : -field[v];
            for (int w = 0; w < B; ++w) {
                if (w == v) continue;
                double c = Q[(size_t)(A + v) * n + (A + w)] + Q[(size_t)(A + w) * n + (A + v)];
                field[w] += up ? c : -c;
            }
            diag_BB[g] = E;
        }
    }
    for (int i = 0; i < A; ++i)
        for (uint64_t s = 0; s < NB; ++s) {
            double acc = 0.0;
            for (int j = 0; j < B; ++j)
                if ((s >> (B - 1 - j)) & 1ull)
                    acc += Q[(size_t)i * n + (A + j)] + Q[(size_t)(A + j) * n + i];
            ABBA[(size_t)i * NB + s] = acc;
        }
    memcpy(chunk, diag_BB, sizeof(double) * NB);
    double glob_E = 0.0;
    uint64_t glob_pref = 0, glob_X = 0;
    double bit_array[64];
    memset(bit_array, 0, sizeof(bit_array));
    uint64_t prev_gray = 0;
    const uint64_t NA = 1ull << A;
    for (uint64_t i = 0; i < NA; ++i) {
        uint64_t gray = i ^ (i >> 1);
        if (i > 0) {
            int sign = gray > prev_gray ? 1 : -1;
            int flipped = __builtin_ctzll(gray ^ prev_gray);
            int row = A - flipped - 1;
            bit_array[row] += sign;
            const double *r = ABBA + (size_t)row * NB;
            if (sign == 1) for (uint64_t s = 0; s < NB; ++s) chunk[s] += r[s];
            else           for (uint64_t s = 0; s < NB; ++s) chunk[s] -= r[s];
            prev_gray = gray;
            if ((i + 1) % (uint64_t)correction_period == 0) {   /* qbf.py:219: (i+1) % period, i = enumerate index */
                for (uint64_t s = 0; s < NB; ++s) {
                    double acc = diag_BB[s];
                    for (int a = 0; a < A; ++a) if (bit_array[a] != 0.0) acc += bit_array[a] * ABBA[(size_t)a * NB + s];
                    chunk[s] = acc;
                }
            }
        }
        uint64_t am = 0;
        double mn = chunk[0];
        for (uint64_t s = 1; s < NB; ++s) if (chunk[s] < mn) { mn = chunk[s]; am = s; }
        double prefix_E = 0.0;
        for (int a = 0; a < A; ++a) if (bit_array[a] != 0.0)
            for (int b = 0; b < A; ++b) if (bit_array[b] != 0.0) prefix_E += Q[(size_t)a * n + b];
        if (gray == 0) { glob_E = 0.0; glob_pref = 0; glob_X = 0; }
        if (mn + prefix_E < glob_E) { glob_E = mn + prefix_E; glob_pref = gray; glob_X = am; }
    }
    *E_out = glob_E;
    *x_out = shl(glob_pref, B) | glob_X;
    free(diag_BB); free(ABBA); free(chunk); free(field);
    return 0;
}

/* ----------------------------------------------------------------------------
 * Independent brute force.  Work is split by the top T bits; each chunk is walked
 * in Gray order over its low bits with full O(n) local-field updates.
 * W[a][b] (bit order, a != b) = Q_ij + Q_ji with i = n-1-a, j = n-1-b; d[a] = Q_ii.
 * ------------------------------------------------------------------------- */
typedef struct { int64_t E; uint64_t rank; } best_i64_t;

int oracle_brute_i64(const int64_t *Q, int n, int B, int nthreads,
                     int64_t *E_out, uint64_t *x_out) {
    if (n <= 0 || n > 48 || B < 0 || B > n) return -1;
    int64_t *W = (int64_t *)malloc(sizeof(int64_t) * n * n);
    int64_t *d = (int64_t *)malloc(sizeof(int64_t) * n);
    if (!W || !d) { free(W); free(d); return -1; }
    for (int a = 0; a < n; ++a) {
        int i = n - 1 - a;
        d[a] = Q[(size_t)i * n + i];
        for (int b = 0; b < n; ++b) {
            int j = n - 1 - b;
            W[(size_t)a * n + b] = (a == b) ? 0 : Q[(size_t)i * n + j] + Q[(size_t)j * n + i];
        }
    }
    /* int32 fields when the energy bound allows it (vectorises twice as wide) */
    int64_t bound = 0;
    for (int a = 0; a < n; ++a) {
        bound += llabs(d[a]);
        for (int b = a + 1; b < n; ++b) bound += llabs(W[(size_t)a * n + b]);
    }
    const int T = n > 16 ? (n > 30 ? 14 : n - 16) : 0;   /* top bits split across threads */
    const int Lw = n - T;
    const uint64_t chunks = 1ull << T;
    best_i64_t gbest = { INT64_MAX, ~0ull };
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const int use32 = bound < (int64_t)1 << 30;
#pragma omp parallel
    {
        best_i64_t lbest = { INT64_MAX, ~0ull };
        int64_t f64[64];
        int32_t f32[64] __attribute__((aligned(64)));
        int32_t *W32 = NULL;
        if (use32) {
            W32 = (int32_t *)malloc(sizeof(int32_t) * n * n);
            for (int k = 0; k < n * n; ++k) W32[k] = (int32_t)W[k];
        }
#pragma omp for schedule(dynamic, 1)
        for (uint64_t h = 0; h < chunks; ++h) {
            uint64_t s = h << Lw;
            /* initial energy and fields at s */
            int64_t E = 0;
            for (int a = 0; a < n; ++a) if ((s >> a) & 1ull) {
                E += d[a];
                for (int b = a + 1; b < n; ++b) if ((s >> b) & 1ull) E += W[(size_t)a * n + b];
            }
            for (int a = 0; a < Lw; ++a) {
                int64_t f = d[a];
                for (int b = 0; b < n; ++b) if (b != a && ((s >> b) & 1ull)) f += W[(size_t)a * n + b];
                f64[a] = f;
                f32[a] = (int32_t)f;
            }
            /* visit s, then the Gray walk */
            {
                uint64_t r = oracle_rank_of_state(s, B);
                if (E < lbest.E || (E == lbest.E && r < lbest.rank)) { lbest.E = E; lbest.rank = r; }
            }
            const uint64_t cnt = 1ull << Lw;
            for (uint64_t i = 1; i < cnt; ++i) {
                int k = __builtin_ctzll(i);
                s ^= 1ull << k;
                int up = (int)((s >> k) & 1ull);
                if (use32) {
                    int32_t fk = f32[k];
                    E += up ? fk : -fk;
                    const int32_t *wk = W32 + (size_t)k * n;
                    if (up) { for (int a = 0; a < Lw; ++a) f32[a] += wk[a]; }
                    else    { for (int a = 0; a < Lw; ++a) f32[a] -= wk[a]; }
                    /* W[k][k] == 0 so f32[k] is unchanged, as it must be */
                } else {
                    int64_t fk = f64[k];
                    E += up ? fk : -fk;
                    const int64_t *wk = W + (size_t)k * n;
                    if (up) { for (int a = 0; a < Lw; ++a) f64[a] += wk[a]; }
                    else    { for (int a = 0; a < Lw; ++a) f64[a] -= wk[a]; }
                }
                if (E <= lbest.E) {
                    uint64_t r = oracle_rank_of_state(s, B);
                    if (E < lbest.E || r < lbest.rank) { lbest.E = E; lbest.rank = r; }
                }
            }
        }
        if (W32) free(W32);
#pragma omp critical
        {
            if (lbest.E < gbest.E || (lbest.E == gbest.E && lbest.rank < gbest.rank)) gbest = lbest;
        }
    }
    *E_out = gbest.E;
    /* state from rank */
    {
        uint64_t r = gbest.rank, p = shr(r, B);
        *x_out = shl(p ^ (p >> 1), B) | (r & lowmask(B));
    }
    free(W); free(d);
    return 0;
}

typedef struct { double E; uint64_t rank; } best_f64_t;

int oracle_brute_f64(const double *Q, int n, int B, int nthreads,
                     double *E_out, uint64_t *x_out) {
    if (n <= 0 || n > 48 || B < 0 || B > n) return -1;
    double *W = (double *)malloc(sizeof(double) * n * n);
    double *d = (double *)malloc(sizeof(double) * n);
    if (!W || !d) { free(W); free(d); return -1; }
    for (int a = 0; a < n; ++a) {
        int i = n - 1 - a;
        d[a] = Q[(size_t)i * n + i];
        for (int b = 0; b < n; ++b) {
            int j = n - 1 - b;
            W[(size_t)a * n + b] = (a == b) ? 0.0 : Q[(size_t)i * n + j] + Q[(size_t)j * n + i];
        }
    }
    const int T = n > 16 ? (n > 30 ? 14 : n - 16) : 0;
    const int Lw = n - T;
    const uint64_t chunks = 1ull << T;
    best_f64_t gbest = { INFINITY, ~0ull };
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    {
        best_f64_t lbest = { INFINITY, ~0ull };
        double f[64] __attribute__((aligned(64)));
#pragma omp for schedule(dynamic, 1)
        for (uint64_t h = 0; h < chunks; ++h) {
            uint64_t s = h << Lw;
            double E = 0.0;
            for (int a = 0; a < n; ++a) if ((s >> a) & 1ull) {
                E += d[a];
                for (int b = a + 1; b < n; ++b) if ((s >> b) & 1ull) E += W[(size_t)a * n + b];
            }
            for (int a = 0; a < Lw; ++a) {
                double fa = d[a];
                for (int b = 0; b < n; ++b) if (b != a && ((s >> b) & 1ull)) fa += W[(size_t)a * n + b];
                f[a] = fa;
            }
            {
                uint64_t r = oracle_rank_of_state(s, B);
                if (E < lbest.E || (E == lbest.E && r < lbest.rank)) { lbest.E = E; lbest.rank = r; }
            }
            const uint64_t cnt = 1ull << Lw;
            for (uint64_t i = 1; i < cnt; ++i) {
                int k = __builtin_ctzll(i);
                s ^= 1ull << k;
                int up = (int)((s >> k) & 1ull);
                double fk = f[k];
                E += up ? fk : -fk;
                const double *wk = W + (size_t)k * n;
                if (up) { for (int a = 0; a < Lw; ++a) f[a] += wk[a]; }
                else    { for (int a = 0; a < Lw; ++a) f[a] -= wk[a]; }
                if (E <= lbest.E) {
                    uint64_t r = oracle_rank_of_state(s, B);
                    if (E < lbest.E || r < lbest.rank) { lbest.E = E; lbest.rank = r; }
                }
            }
        }
#pragma omp critical
        {
            if (lbest.E < gbest.E || (lbest.E == gbest.E && lbest.rank < gbest.rank)) gbest = lbest;
        }
    }
    *E_out = gbest.E;
    {
        uint64_t r = gbest.rank, p = shr(r, B);
        *x_out = shl(p ^ (p >> 1), B) | (r & lowmask(B));
    }
    free(W); free(d);
    return 0;
}

/* direct energy x^T Q x of one state (state integer, MSB = x[0]) */
int64_t oracle_energy_i64(const int64_t *Q, int n, uint64_t s) {
    int64_t E = 0;
    for (int i = 0; i < n; ++i) if ((s >> (n - 1 - i)) & 1ull)
        for (int j = 0; j < n; ++j) if ((s >> (n - 1 - j)) & 1ull) E += Q[(size_t)i * n + j];
    return E;
}

double oracle_energy_f64(const double *Q, int n, uint64_t s) {
    double E = 0.0;
    for (int i = 0; i < n; ++i) if ((s >> (n - 1 - i)) & 1ull)
        for (int j = 0; j < n; ++j) if ((s >> (n - 1 - j)) & 1ull) E += Q[(size_t)i * n + j];
    return E;
}
