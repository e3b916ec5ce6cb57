/*
 * Multi-threaded (OpenMP) variants of the oracle's assembly and PCG loops —
 * TEST / BASELINE INFRASTRUCTURE ONLY.
 *
 * They time the reference's CPU algorithms on all host cores (the reference
 * itself is single-threaded outside faer). Element integration calls the
 * same or_hex_stiffness as the oracle, so every Ke is bit-identical to the
 * single-threaded restatement; only the PCG dot products change summation
 * order (per-thread partial sums), so iteration counts may differ by one or
 * two from or_pcg.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "axia_oracle.h"

int or_omp_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Element integration for all elements into ke_out (nelem x (3nn)^2),
 * analysis.rs:574-640 parallelised over elements. nthreads <= 0 = default. */
int or_element_ke_batch(int nn, int64_t nnode, const double *xyz, int64_t nelem,
                        const int32_t *conn, const double *e, const double *nu, double *ke_out,
                        int nthreads) {
    if (nn != 8 && nn != 20) return OR_E_INVALID;
    int m = 3 * nn;
    int status = OR_OK;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t el = 0; el < nelem; el++) {
        double exyz[60], d[36], vol;
        int ok = 1;
        for (int a = 0; a < nn; a++) {
            int32_t ni = conn[el * nn + a];
            if (ni < 0 || ni >= nnode) {
                ok = 0;
                break;
            }
            for (int q = 0; q < 3; q++) exyz[3 * a + q] = xyz[3 * (int64_t)ni + q];
        }
        int st = ok ? or_d_iso_3d(e[el], nu[el], d) : OR_E_INVALID;
        if (!st) st = or_hex_stiffness(nn, exyz, d, ke_out + el * (int64_t)m * m, &vol);
        if (st) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            status = st;
        }
    }
    (void)nthreads;
    return status;
}

static void spmv_omp(int64_t n, const int64_t *indptr, const int64_t *indices, const double *data,
                     const double *x, double *y) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int64_t k = indptr[i]; k < indptr[i + 1]; k++) s += data[k] * x[indices[k]];
        y[i] = s;
    }
}

/* linalg.rs:194-259 with OpenMP SpMV and reductions. If max_iter_cap > 0 the
 * loop stops after that many iterations without an error (used to time a
 * bounded sample); *iters reports the iterations run. */
int or_pcg_omp(int64_t n, const int64_t *indptr, const int64_t *indices, const double *data,
               const double *b, double tol, int64_t max_iter, int64_t max_iter_cap, double *x,
               int64_t *iters_out, double *residual_out) {
    double *r = (double *)malloc(sizeof(double) * (size_t)n * 5 + 8);
    if (!r) return OR_E_NOMEM;
    double *z = r + n, *p = r + 2 * n, *ap = r + 3 * n, *dinv = r + 4 * n;
    double bnorm = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : bnorm) schedule(static)
#endif
    for (int64_t i = 0; i < n; i++) {
        x[i] = 0.0;
        r[i] = b[i];
        bnorm += b[i] * b[i];
    }
    bnorm = sqrt(bnorm);
    if (bnorm < 1e-30) {
        free(r);
        *iters_out = 0;
        *residual_out = 0.0;
        return OR_OK;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < n; i++) {
        double d = 0.0;
        for (int64_t k = indptr[i]; k < indptr[i + 1]; k++)
            if (indices[k] == i) {
                d = data[k];
                break;
            }
        dinv[i] = fabs(d) > 1e-30 ? d : 1.0;
    }
    double rz = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : rz) schedule(static)
#endif
    for (int64_t i = 0; i < n; i++) {
        z[i] = r[i] / dinv[i];
        p[i] = z[i];
        rz += r[i] * z[i];
    }
    int64_t iters = 0;
    double residual = bnorm;
    int64_t lim = (max_iter_cap > 0 && max_iter_cap < max_iter) ? max_iter_cap : max_iter;
    for (int64_t it = 0; it < lim; it++) {
        iters = it + 1;
        spmv_omp(n, indptr, indices, data, p, ap);
        double pap = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : pap) schedule(static)
#endif
        for (int64_t i = 0; i < n; i++) pap += p[i] * ap[i];
        if (fabs(pap) < 1e-30) {
            free(r);
            return OR_E_NOT_SPD;
        }
        double alpha = rz / pap, rr = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : rr) schedule(static)
#endif
        for (int64_t i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
            rr += r[i] * r[i];
        }
        residual = sqrt(rr);
        if (residual <= tol * bnorm) break;
        double rz_new = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : rz_new) schedule(static)
#endif
        for (int64_t i = 0; i < n; i++) {
            z[i] = r[i] / dinv[i];
            rz_new += r[i] * z[i];
        }
        double beta = rz_new / rz;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        rz = rz_new;
    }
    free(r);
    *iters_out = iters;
    *residual_out = residual;
    if (max_iter_cap <= 0 && residual > tol * bnorm * 10.0) return OR_E_NOT_CONVERGED;
    return OR_OK;
}

/* Thread count for the OpenMP variants (<= 0 restores the default). */
void or_set_threads(int n) {
#ifdef _OPENMP
    static int dflt = 0;
    if (!dflt) dflt = omp_get_max_threads();
    omp_set_num_threads(n > 0 ? n : dflt);
#else
    (void)n;
#endif
}
