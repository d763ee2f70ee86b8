/* Plain C99 caller of libqbx.so through include/qbx.h: opens device 0, solves a small
 * integer QUBO, a batch, a solve in four parts merged by (E, rank), and a burst through the
 * server, and prints one line per result for tests/test_c_abi.py to compare with the
 * Python binding's answers. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "qbx.h"

#define N 12
#define N2 20

static void lcg_fill(int32_t *Q, int count, uint32_t x) {   /* reproducible from Python too */
    for (int i = 0; i < count; ++i) {
        x = x * 1103515245u + 12345u;
        Q[i] = (int32_t)((x >> 16) % 41u) - 20;
    }
}

int main(void) {
    qbx_ctx_t *ctx = NULL;
    if (qbx_open(0, &ctx) != QBX_OK) {
        fprintf(stderr, "open: %s\n", qbx_last_error());
        return 2;
    }
    int32_t Q[N * N];
    lcg_fill(Q, N * N, 12345u);
    qbx_opts_t o = {-1, -1, -1, 0};
    qbx_result_t r;
    if (qbx_solve_i32(ctx, Q, N, &o, &r) != QBX_OK) {
        fprintf(stderr, "solve: %s\n", qbx_last_error());
        return 3;
    }
    printf("solve %lld %llu\n", (long long)r.E_int, (unsigned long long)r.state);

    int64_t offsets[2] = {0, N * N};
    int32_t ns[2] = {N, N};
    int32_t packed[2 * N * N];
    for (int i = 0; i < N * N; ++i) { packed[i] = Q[i]; packed[N * N + i] = -Q[i]; }
    qbx_result_t rb[2];
    if (qbx_batch_i32(ctx, 2, packed, offsets, ns, &o, rb) != QBX_OK) {
        fprintf(stderr, "batch: %s\n", qbx_last_error());
        return 4;
    }
    printf("batch %lld %llu %lld %llu\n", (long long)rb[0].E_int, (unsigned long long)rb[0].state,
           (long long)rb[1].E_int, (unsigned long long)rb[1].state);

    if (qbx_solve_i32(ctx, Q, 57, &o, &r) != QBX_EINVAL) return 5;   /* n out of range */

    /* one solve as four parts (as four GPUs would run it), merged by (E, rank) */
    int32_t Q2[N2 * N2];
    lcg_fill(Q2, N2 * N2, 777u);
    qbx_opts_t o2 = {-1, 8, 7, 0};
    qbx_result_t best;
    for (int part = 0; part < 4; ++part) {
        qbx_result_t rp;
        if (qbx_solve_part_i32(ctx, Q2, N2, &o2, part, 4, &rp) != QBX_OK) {
            fprintf(stderr, "part: %s\n", qbx_last_error());
            return 6;
        }
        if (part == 0 || rp.E_int < best.E_int || (rp.E_int == best.E_int && rp.rank < best.rank)) best = rp;
    }
    printf("parts %lld %llu\n", (long long)best.E_int, (unsigned long long)best.state);

    /* a burst of two requests through the server */
    qbx_server_t *srv = NULL;
    if (qbx_server_start(ctx, NULL, &srv) != QBX_OK) {
        fprintf(stderr, "server: %s\n", qbx_last_error());
        return 7;
    }
    uint64_t ids[2];
    if (qbx_server_submit_many_i32(srv, 2, packed, offsets, ns, -1, ids) != QBX_OK) {
        fprintf(stderr, "submit: %s\n", qbx_last_error());
        return 8;
    }
    qbx_request_result_t done[2], got[2];
    int have = 0;
    for (int tries = 0; have < 2 && tries < 1000; ++tries) {
        int cnt = 0;
        if (qbx_server_poll(srv, done, 2, 10, &cnt) != QBX_OK) return 9;
        for (int i = 0; i < cnt; ++i) got[done[i].id == ids[0] ? 0 : 1] = done[i];
        have += cnt;
    }
    qbx_server_stop(srv);
    if (have != 2) return 10;
    printf("server %lld %llu %lld %llu\n", (long long)got[0].E_int, (unsigned long long)got[0].state,
           (long long)got[1].E_int, (unsigned long long)got[1].state);
    /* seeded TSP instances built on the device from integer seeds (numpy's PCG64(seed)) */
    uint64_t seeds[3] = {42, 7, 12345678901ULL};
    int32_t cities[3] = {5, 4, 5};
    qbx_pcg64_t st[3];
    if (qbx_pcg64_from_seeds(3, seeds, st) != QBX_OK) return 11;
    qbx_result_t rt[3];
    if (qbx_batch_random_tsp(ctx, 3, st, cities, 0.0, 100.0, QBX_GEN_FLOAT32, &o, rt) != QBX_OK) {
        fprintf(stderr, "tsp: %s\n", qbx_last_error());
        return 12;
    }
    printf("tsp %.17g %llu\n", rt[0].E, (unsigned long long)rt[0].state);

    /* a float problem with either float filter forced: the same answer; both at once is refused */
    double F[N2 * N2];
    for (int i = 0; i < N2 * N2; ++i) F[i] = (double)Q2[i] / 4.0;
    qbx_opts_t op = {-1, 12, -1, QBX_FLAG_FLOAT_PACKED_FILTER}, oi = {-1, 12, -1, QBX_FLAG_FLOAT_INT32_FILTER};
    qbx_result_t fp, fi;
    if (qbx_solve_f64(ctx, F, N2, &op, &fp) != QBX_OK || qbx_solve_f64(ctx, F, N2, &oi, &fi) != QBX_OK) {
        fprintf(stderr, "float: %s\n", qbx_last_error());
        return 13;
    }
    if (fp.E != fi.E || fp.state != fi.state || fp.path != QBX_PATH_FLOAT_PAIR16 || fi.path != QBX_PATH_FLOAT) return 14;
    qbx_opts_t both = {-1, 12, -1, QBX_FLAG_FLOAT_PACKED_FILTER | QBX_FLAG_FLOAT_INT32_FILTER};
    if (qbx_solve_f64(ctx, F, N2, &both, &fp) != QBX_EINVAL) return 15;
    printf("float %.17g %llu\n", fi.E, (unsigned long long)fi.state);
    qbx_close(ctx);
    return 0;
}
