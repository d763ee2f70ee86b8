/*
 * qbx_cpu.h -- C ABI of libqbx_cpu.so: the exhaustive QUBO solver on the host CPU.
 *
 * Serves device='cpu', the reference QBF's default device (src/qubo_lib/abc/solver_abc.py:
 * 62-64), which there runs QBF's numpy walk (src/qubo_lib/solvers/qbf.py:48-70, 188-230).
 * This library runs the GPU kernels' own lane walk (the same source, compiled for the host)
 * with OpenMP threads over lanes, so its answers are the GPU's: the exact (E, rank) minimum
 * under the reference tie rule (see qbx.h), integer problems in exact int32 arithmetic,
 * float problems by the int32 filter over quantized tables plus FP64 re-walks (FP64-exact).
 *
 * Conventions are those of qbx.h: Q is n x n row-major, 1 <= n <= 56; every call returns
 * 0 or a negative QBX_E* code and sets a thread-local message (qbx_cpu_last_error);
 * pointers are caller-owned; `threads` <= 0 uses OpenMP's default thread count.  Needs no
 * GPU and no HIP runtime.  out->kernel_ms is the wall time of the walk.
 */
#ifndef QBX_CPU_H
#define QBX_CPU_H

#include <stdint.h>

#include "qbx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Exact int32 walk; requires sum_i |Q_ii| + sum_{i<j} |Q_ij + Q_ji| <= 2^31 - 2, else
 * QBX_ERANGE.  Replaces QBF's CPU prefix walk (qbf.py:188-230) and its caches
 * (qbf.py:246-289) for integer Q. */
int qbx_cpu_solve_i32(const int32_t *Q, int n, const qbx_opts_t *opts, int threads, qbx_result_t *out);
/* Float Q as doubles (finite, bound <= 1e300, else QBX_ERANGE); FP64-exact answer.  Same
 * reference counterpart, for float Q (there with float32 accumulation and periodic FP64
 * correction, qbf.py:214-217, 232-244). */
int qbx_cpu_solve_f64(const double *Q, int n, const qbx_opts_t *opts, int threads, qbx_result_t *out);
/* The solve with the reference's per-prefix metrics beside its answer, as
 * qbx_solve_minima_* (qbx.h; energy_arr and time_arr, qbf.py:342-351, 362-365, 409-410):
 * minima[i] = least energy over the states of reference prefix i ^ (i >> 1), wave_ms[i]
 * (nullable) = thread time spent on that prefix, ms.  Needs 1 <= B, n - B <= 28. */
int qbx_cpu_solve_minima_i32(const int32_t *Q, int n, const qbx_opts_t *opts, int threads, int64_t *minima,
                             double *wave_ms, qbx_result_t *out);
int qbx_cpu_solve_minima_f64(const double *Q, int n, const qbx_opts_t *opts, int threads, double *minima,
                             double *wave_ms, qbx_result_t *out);
const char *qbx_cpu_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QBX_CPU_H */
