/*
 * cpu_baseline.c -- CPU restatement of the reference solve (A10-A14), TEST/BASELINE
 * INFRASTRUCTURE ONLY.  Filled in below with CSR operators and the Schur driver.
 */
#include <stddef.h>
void orc_csr_spmv(long n, const long *ptr, const int *col, const double *val, const double *x, double *y)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        double s = 0.0;
        for (long k = ptr[i]; k < ptr[i + 1]; k++) s += val[k] * x[col[k]];
        y[i] = s;
    }
}
