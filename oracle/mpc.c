/* Transformation-handler collapses of the oracle (test infrastructure only):
 * restates _collapse_vector_by_mpc, _collapse_matrix_by_mpc and
 * _enforce_mpc_values (solver/run_case/helpers.mojo:3288-3301, :3345-3377,
 * :3380-3397).  Each DOF i expands as u_i = sum_t coeff[t] u[dof_pool[t]]
 * over t in [row_offsets[i], row_offsets[i + 1]). */
#include "oracle_internal.h"

/* helpers.mojo:3345-3377: out = T^T K T on dense n x n matrices, terms of a
 * zero K entry skipped, in the reference's loop order. */
void or_mpc_collapse_matrix(int n, const double *K, const int *row_offsets, const int *dof_pool,
                            const double *coeff_pool, double *out) {
    for (long k = 0; k < (long)n * n; ++k) out[k] = 0.0;
    for (int i = 0; i < n; ++i) {
        const int a0 = row_offsets[i], a1 = row_offsets[i + 1];
        if (a0 == a1) continue;
        for (int j = 0; j < n; ++j) {
            const double value = K[(long)i * n + j];
            if (value == 0.0) continue;
            const int b0 = row_offsets[j], b1 = row_offsets[j + 1];
            if (b0 == b1) continue;
            for (int ti = a0; ti < a1; ++ti) {
                const int ri = dof_pool[ti];
                const double ci = coeff_pool[ti];
                for (int tj = b0; tj < b1; ++tj) out[(long)ri * n + dof_pool[tj]] += ci * value * coeff_pool[tj];
            }
        }
    }
}

/* helpers.mojo:3288-3301: out = T^T F */
void or_mpc_collapse_vector(int n, const double *F, const int *row_offsets, const int *dof_pool,
                            const double *coeff_pool, double *out) {
    for (int k = 0; k < n; ++k) out[k] = 0.0;
    for (int i = 0; i < n; ++i)
        for (int t = row_offsets[i]; t < row_offsets[i + 1]; ++t) out[dof_pool[t]] += coeff_pool[t] * F[i];
}

/* helpers.mojo:3380-3397: slave DOFs from their retained DOFs, fixed DOFs 0 */
void or_mpc_enforce(int n, double *u, const int *constrained, const int *slave, const int *row_offsets,
                    const int *dof_pool, const double *coeff_pool) {
    for (int i = 0; i < n; ++i) {
        if (slave[i]) {
            double value = 0.0;
            for (int t = row_offsets[i]; t < row_offsets[i + 1]; ++t) value += coeff_pool[t] * u[dof_pool[t]];
            u[i] = value;
        } else if (constrained[i]) {
            u[i] = 0.0;
        }
    }
}
