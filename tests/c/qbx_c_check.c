/* Plain C99 caller of libqbx.so through include/qbx.h: opens device 0, solves a small
 * integer QUBO, a batch, and a device-drawn RandomQUBO batch, and prints one line per
 * result for tests/test_c_abi.py to compare with the Python binding's answers. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "qbx.h"

#define N 12

int main(void) {
    qbx_ctx_t *ctx = NULL;
    if (qbx_open(0, &ctx) != QBX_OK) {
        fprintf(stderr, "open: %s\n", qbx_last_error());
        return 2;
    }
    int32_t Q[N * N];
    uint32_t x = 12345u;
    for (int i = 0; i < N * N; ++i) {      /* small LCG: reproducible from Python too */
        x = x * 1103515245u + 12345u;
        Q[i] = (int32_t)((x >> 16) % 41u) - 20;
    }
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
    qbx_close(ctx);
    return 0;
}
