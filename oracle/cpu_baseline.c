/*
 * cpu_baseline.c -- CPU restatement of the reference's Stokes solve (TEST AND BASELINE
 * INFRASTRUCTURE ONLY: never linked into or called by the product library).
 *
 * Algorithm, following the reference:
 *   - Schur reduction (BSSCR_DRIVER_auglag, LU/Solvers/KSPSolvers/src/BSSCR/auglag-driver-DGTGD.c:64-584):
 *     solve S p = G^T K^{-1} f - h with S = G^T K^{-1} G by FGMRES preconditioned by the
 *     diagonal approxS (Q22_pc_type "uw", preconditioner.c:138-222), then back-solve
 *     u = K^{-1} (f - G p) (:494-531);
 *   - every K^{-1} is FGMRES to A11 rtol (PY/systems/_bsscr.py:282) preconditioned by one
 *     multigrid V-cycle (mg.c:62-96) whose prolongation is SROpGenerator_Simple
 *     (SROpGenerator.c:642-1342: a fine node averages its 1, 2, 4 or 8 coarse parents, BC rows
 *     skipped) and whose coarse operators are Galerkin P^T A P;
 *   - level smoother Chebyshev(2)/Jacobi, dense Cholesky on the coarsest level.
 * Operators are assembled CSR matrices on the free (non-Dirichlet) DOFs, as the reference's
 * PETSc AIJ matrices with removeBCs=True.  OpenMP over rows/entries.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

void orc_csr_spmv(long n, const long *ptr, const int *col, const double *val, const double *x, double *y)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        double s = 0.0;
        for (long k = ptr[i]; k < ptr[i + 1]; k++) s += val[k] * x[col[k]];
        y[i] = s;
    }
}

typedef struct {
    long n, m; /* rows, cols */
    long *ptr;
    int *col;
    double *val;
} csr;

static void csr_free(csr *A)
{
    free(A->ptr);
    free(A->col);
    free(A->val);
    memset(A, 0, sizeof(*A));
}

static void spmv(const csr *A, const double *x, double *y) { orc_csr_spmv(A->n, A->ptr, A->col, A->val, x, y); }

/* y = A^T x */
static void spmv_t(const csr *A, const double *x, double *y)
{
    memset(y, 0, sizeof(double) * A->m);
    for (long i = 0; i < A->n; i++)
        for (long k = A->ptr[i]; k < A->ptr[i + 1]; k++) y[A->col[k]] += A->val[k] * x[i];
}

static csr transpose(const csr *A)
{
    csr T = {A->m, A->n, calloc(A->m + 1, sizeof(long)), malloc(sizeof(int) * (A->ptr[A->n] + 1)),
             malloc(sizeof(double) * (A->ptr[A->n] + 1))};
    for (long k = 0; k < A->ptr[A->n]; k++) T.ptr[A->col[k] + 1]++;
    for (long i = 0; i < A->m; i++) T.ptr[i + 1] += T.ptr[i];
    long *pos = malloc(sizeof(long) * (A->m + 1));
    memcpy(pos, T.ptr, sizeof(long) * (A->m + 1));
    for (long i = 0; i < A->n; i++)
        for (long k = A->ptr[i]; k < A->ptr[i + 1]; k++) {
            long d = pos[A->col[k]]++;
            T.col[d] = (int)i;
            T.val[d] = A->val[k];
        }
    free(pos);
    return T;
}

/* C = A B (row-wise accumulation with a dense marker per thread) */
static csr spgemm(const csr *A, const csr *B)
{
    csr C = {A->n, B->m, calloc(A->n + 1, sizeof(long)), NULL, NULL};
    int nt = omp_get_max_threads();
    long **mark = malloc(sizeof(long *) * nt);
    for (int t = 0; t < nt; t++) {
        mark[t] = malloc(sizeof(long) * B->m);
        for (long j = 0; j < B->m; j++) mark[t][j] = -1;
    }
    #pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < A->n; i++) {
        long *mk = mark[omp_get_thread_num()];
        long cnt = 0;
        for (long ka = A->ptr[i]; ka < A->ptr[i + 1]; ka++) {
            int k = A->col[ka];
            for (long kb = B->ptr[k]; kb < B->ptr[k + 1]; kb++)
                if (mk[B->col[kb]] != i) {
                    mk[B->col[kb]] = i;
                    cnt++;
                }
        }
        C.ptr[i + 1] = cnt;
    }
    for (long i = 0; i < A->n; i++) C.ptr[i + 1] += C.ptr[i];
    C.col = malloc(sizeof(int) * (C.ptr[A->n] + 1));
    C.val = malloc(sizeof(double) * (C.ptr[A->n] + 1));
    double **acc = malloc(sizeof(double *) * nt);
    for (int t = 0; t < nt; t++) {
        acc[t] = calloc(B->m, sizeof(double));
        for (long j = 0; j < B->m; j++) mark[t][j] = -1;
    }
    #pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < A->n; i++) {
        int t = omp_get_thread_num();
        long *mk = mark[t];
        double *ac = acc[t];
        long w = C.ptr[i];
        for (long ka = A->ptr[i]; ka < A->ptr[i + 1]; ka++) {
            int k = A->col[ka];
            double a = A->val[ka];
            for (long kb = B->ptr[k]; kb < B->ptr[k + 1]; kb++) {
                int j = B->col[kb];
                if (mk[j] != i) {
                    mk[j] = i;
                    ac[j] = 0.0;
                    C.col[w++] = j;
                }
                ac[j] += a * B->val[kb];
            }
        }
        for (long k = C.ptr[i]; k < C.ptr[i + 1]; k++) C.val[k] = ac[C.col[k]];
    }
    for (int t = 0; t < nt; t++) {
        free(mark[t]);
        free(acc[t]);
    }
    free(mark);
    free(acc);
    return C;
}

/* ---------------------------------------------------------------------------------- */
/* multigrid hierarchy                                                                  */
/* ---------------------------------------------------------------------------------- */
typedef struct {
    csr A, P, R;          /* P: this level <- next coarser level; R = P^T */
    double *dinv, *r, *b, *x, *t1, *t2;
    double lmax;
    long n;
} level_t;

typedef struct {
    int nl;
    level_t L[16];
    double *chol; /* coarsest dense Cholesky factor */
    long nc;
    long vcycles;
} mg_t;

/* Free-DOF numbering of a Q1 node grid with free-slip walls (normal component fixed). */
static long number_free(int dim, const int *nn, long *eq)
{
    long N = (long)nn[0] * nn[1] * (dim == 3 ? nn[2] : 1), c = 0;
    for (long a = 0; a < N; a++) {
        int ij[3] = {(int)(a % nn[0]), (int)((a / nn[0]) % nn[1]), dim == 3 ? (int)(a / ((long)nn[0] * nn[1])) : 0};
        for (int d = 0; d < dim; d++) {
            int wall = ij[d] == 0 || ij[d] == nn[d] - 1;
            eq[a * dim + d] = wall ? -1 : c++;
        }
    }
    return c;
}

/* SROpGenerator_Simple prolongation between node grids nf = 2 nc - 1 per axis: each fine node
   averages its coincident coarse node, or its 2, 4 or 8 surrounding coarse nodes, with weight
   1/npoints; rows and columns of fixed DOFs are dropped. */
static csr build_P(int dim, const int *nnf, const int *nnc, const long *eqf, long nf, const long *eqc, long nc)
{
    long Nf = (long)nnf[0] * nnf[1] * (dim == 3 ? nnf[2] : 1);
    csr P = {nf, nc, calloc(nf + 1, sizeof(long)), malloc(sizeof(int) * nf * 8), malloc(sizeof(double) * nf * 8)};
    long w = 0;
    for (long a = 0; a < Nf; a++) {
        int ij[3] = {(int)(a % nnf[0]), (int)((a / nnf[0]) % nnf[1]), dim == 3 ? (int)(a / ((long)nnf[0] * nnf[1])) : 0};
        int lo[3], cnt[3];
        for (int d = 0; d < 3; d++) {
            if (d >= dim) {
                lo[d] = 0;
                cnt[d] = 1;
                continue;
            }
            lo[d] = ij[d] / 2;
            cnt[d] = (ij[d] & 1) ? 2 : 1;
        }
        int np = cnt[0] * cnt[1] * cnt[2];
        for (int d = 0; d < dim; d++) {
            long row = eqf[a * dim + d];
            if (row < 0) continue;
            for (int z = 0; z < cnt[2]; z++)
                for (int y = 0; y < cnt[1]; y++)
                    for (int x = 0; x < cnt[0]; x++) {
                        long cn = (lo[0] + x) + (long)nnc[0] * ((lo[1] + y) + (long)nnc[1] * (lo[2] + z));
                        long colc = eqc[cn * dim + d];
                        if (colc < 0) continue;
                        P.col[w] = (int)colc;
                        P.val[w] = 1.0 / np;
                        w++;
                    }
            P.ptr[row + 1] = w;
        }
    }
    /* rows were written in increasing order of their free index */
    for (long i = 0; i < nf; i++)
        if (P.ptr[i + 1] < P.ptr[i]) P.ptr[i + 1] = P.ptr[i];
    return P;
}

static double dotp(long n, const double *a, const double *b)
{
    double s = 0.0;
    #pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

static void cheb(level_t *A, const double *b, double *x, int degree, int zero_guess)
{
    const long n = A->n;
    const double emax = 1.1 * A->lmax, emin = 0.12 * A->lmax;
    const double theta = 0.5 * (emax + emin), delta = 0.5 * (emax - emin), sigma = theta / delta;
    double rho = 1.0 / sigma;
    double *r = A->t1, *d = A->t2;
    if (zero_guess) {
        #pragma omp parallel for
        for (long i = 0; i < n; i++) r[i] = b[i], x[i] = 0.0;
    } else {
        spmv(&A->A, x, r);
        #pragma omp parallel for
        for (long i = 0; i < n; i++) r[i] = b[i] - r[i];
    }
    #pragma omp parallel for
    for (long i = 0; i < n; i++) {
        d[i] = A->dinv[i] * r[i] / theta;
        x[i] += d[i];
    }
    for (int k = 1; k < degree; k++) {
        spmv(&A->A, x, r);
        double rn = 1.0 / (2.0 * sigma - rho);
        #pragma omp parallel for
        for (long i = 0; i < n; i++) {
            d[i] = rn * rho * d[i] + 2.0 * rn / delta * A->dinv[i] * (b[i] - r[i]);
            x[i] += d[i];
        }
        rho = rn;
    }
}

static void coarse_solve(mg_t *M, const double *b, double *x)
{
    long n = M->nc;
    const double *Lf = M->chol;
    for (long i = 0; i < n; i++) {
        double s = b[i];
        for (long k = 0; k < i; k++) s -= Lf[i * n + k] * x[k];
        x[i] = s / Lf[i * n + i];
    }
    for (long i = n - 1; i >= 0; i--) {
        double s = x[i];
        for (long k = i + 1; k < n; k++) s -= Lf[k * n + i] * x[k];
        x[i] = s / Lf[i * n + i];
    }
}

static void vcycle(mg_t *M, int l, const double *b, double *x)
{
    level_t *A = &M->L[l];
    if (l == M->nl - 1) {
        coarse_solve(M, b, x);
        return;
    }
    level_t *C = &M->L[l + 1];
    cheb(A, b, x, 2, 1);
    spmv(&A->A, x, A->r);
    #pragma omp parallel for
    for (long i = 0; i < A->n; i++) A->r[i] = b[i] - A->r[i];
    spmv(&A->R, A->r, C->b);
    vcycle(M, l + 1, C->b, C->x);
    spmv(&A->P, C->x, A->r);
    #pragma omp parallel for
    for (long i = 0; i < A->n; i++) x[i] += A->r[i];
    cheb(A, b, x, 2, 0);
}

static void level_alloc(level_t *A)
{
    long n = A->n;
    A->dinv = malloc(sizeof(double) * n);
    A->r = malloc(sizeof(double) * n);
    A->b = malloc(sizeof(double) * n);
    A->x = malloc(sizeof(double) * n);
    A->t1 = malloc(sizeof(double) * n);
    A->t2 = malloc(sizeof(double) * n);
    for (long i = 0; i < n; i++) {
        double dg = 0.0;
        for (long k = A->A.ptr[i]; k < A->A.ptr[i + 1]; k++)
            if (A->A.col[k] == i) dg = A->A.val[k];
        A->dinv[i] = dg != 0.0 ? 1.0 / dg : 0.0;
    }
    /* largest eigenvalue of D^-1 A by power iteration (deterministic start) */
    double *v = A->x, *w = A->b;
    for (long i = 0; i < n; i++) v[i] = 1.0 + 0.1 * ((i * 7919) % 13);
    double lam = 1.0;
    for (int it = 0; it < 12; it++) {
        double nv = sqrt(dotp(n, v, v));
        for (long i = 0; i < n; i++) v[i] /= nv;
        spmv(&A->A, v, w);
        for (long i = 0; i < n; i++) w[i] *= A->dinv[i];
        lam = dotp(n, v, w);
        memcpy(v, w, sizeof(double) * n);
    }
    A->lmax = lam;
}

static void mg_free(mg_t *M)
{
    for (int l = 0; l < M->nl; l++) {
        level_t *A = &M->L[l];
        if (l > 0) csr_free(&A->A);
        csr_free(&A->P);
        csr_free(&A->R);
        free(A->dinv), free(A->r), free(A->b), free(A->x), free(A->t1), free(A->t2);
    }
    free(M->chol);
}

/* ---------------------------------------------------------------------------------- */
/* FGMRES (right preconditioned, modified Gram-Schmidt, restart m)                      */
/* ---------------------------------------------------------------------------------- */
typedef void (*op_fn)(void *ctx, const double *x, double *y);

static int fgmres(long n, op_fn A, void *actx, op_fn M, void *mctx, const double *b, double *x, double rtol,
                  int maxit, int m, int *its_out)
{
    double *V = malloc(sizeof(double) * n * (m + 1));
    double *Z = malloc(sizeof(double) * n * m);
    double *w = malloc(sizeof(double) * n);
    double *H = calloc((size_t)(m + 1) * m, sizeof(double));
    double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m), *g = malloc(sizeof(double) * (m + 1));
    double *y = malloc(sizeof(double) * m);
    int its = 0, conv = 0;
    double r0 = -1.0;
    while (its < maxit && !conv) {
        A(actx, x, w);
        #pragma omp parallel for
        for (long i = 0; i < n; i++) V[i] = b[i] - w[i];
        double beta = sqrt(dotp(n, V, V));
        if (r0 < 0) r0 = beta;
        if (beta <= rtol * r0 || beta == 0.0) {
            conv = 1;
            break;
        }
        for (long i = 0; i < n; i++) V[i] /= beta;
        memset(g, 0, sizeof(double) * (m + 1));
        g[0] = beta;
        int j;
        for (j = 0; j < m && its < maxit; j++, its++) {
            M(mctx, V + j * n, Z + j * n);
            A(actx, Z + j * n, w);
            for (int i = 0; i <= j; i++) {
                double h = dotp(n, w, V + i * n);
                H[i * m + j] = h;
                #pragma omp parallel for
                for (long k = 0; k < n; k++) w[k] -= h * V[i * n + k];
            }
            double hn = sqrt(dotp(n, w, w));
            H[(j + 1) * m + j] = hn;
            #pragma omp parallel for
            for (long k = 0; k < n; k++) V[(j + 1) * n + k] = hn > 0 ? w[k] / hn : 0.0;
            for (int i = 0; i < j; i++) {
                double a = H[i * m + j], c = H[(i + 1) * m + j];
                H[i * m + j] = cs[i] * a + sn[i] * c;
                H[(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
            }
            double a = H[j * m + j], c = H[(j + 1) * m + j];
            double den = sqrt(a * a + c * c);
            cs[j] = den > 0 ? a / den : 1.0;
            sn[j] = den > 0 ? c / den : 0.0;
            H[j * m + j] = den;
            H[(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            if (fabs(g[j + 1]) <= rtol * r0) {
                j++;
                its++;
                conv = 1;
                break;
            }
        }
        for (int i = j - 1; i >= 0; i--) {
            double s = g[i];
            for (int k = i + 1; k < j; k++) s -= H[i * m + k] * y[k];
            y[i] = s / H[i * m + i];
        }
        for (int i = 0; i < j; i++) {
            #pragma omp parallel for
            for (long k = 0; k < n; k++) x[k] += y[i] * Z[i * n + k];
        }
    }
    free(V), free(Z), free(w), free(H), free(cs), free(sn), free(g), free(y);
    if (its_out) *its_out = its;
    return conv;
}

/* ---------------------------------------------------------------------------------- */
/* Schur driver                                                                         */
/* ---------------------------------------------------------------------------------- */
typedef struct {
    mg_t *mg;
    const csr *G, *Gt;
    const double *approxS;
    double inner_rtol;
    long nu, np;
    double *t1, *t2, *t3;
    long inner_its;
} schur_ctx;

static void K_op(void *c, const double *x, double *y) { spmv(&((schur_ctx *)c)->mg->L[0].A, x, y); }
static void V_op(void *c, const double *x, double *y)
{
    schur_ctx *s = (schur_ctx *)c;
    s->mg->vcycles++;
    vcycle(s->mg, 0, x, y);
}

static void K_solve(schur_ctx *s, const double *b, double *x)
{
    int its = 0;
    memset(x, 0, sizeof(double) * s->nu);
    fgmres(s->nu, K_op, s, V_op, s, b, x, s->inner_rtol, 10000, 30, &its);
    s->inner_its += its;
}

/* S p = G^T K^{-1} G p */
static void S_op(void *c, const double *p, double *y)
{
    schur_ctx *s = (schur_ctx *)c;
    spmv(s->G, p, s->t1);
    K_solve(s, s->t1, s->t2);
    spmv(s->Gt, s->t2, y);
}
static void Sp_op(void *c, const double *r, double *z)
{
    schur_ctx *s = (schur_ctx *)c;
    for (long i = 0; i < s->np; i++) z[i] = r[i] / s->approxS[i];
}

/* Builds the multigrid hierarchy on the free DOFs of the fine grid (element resolution
   res, free-slip walls) and solves the Stokes system.  K (nu x nu) and G (nu x np) are CSR
   on free DOFs in the numbering of number_free().  Returns the outer iteration count;
   stats[0..3] = outer its, total inner its, V-cycles, levels; times[0..1] = setup, solve (s). */
int orc_bsscr_solve(int dim, const int *res, long nu, const long *Kptr, const int *Kcol, const double *Kval, long np,
                    const long *Gptr, const int *Gcol, const double *Gval, const double *approxS, const double *f,
                    double outer_rtol, double inner_rtol, int coarse_max, int max_outer, double *u, double *p,
                    long *stats, double *times)
{
    double t0 = omp_get_wtime();
    mg_t M;
    memset(&M, 0, sizeof(M));
    int nn[3] = {res[0] + 1, res[1] + 1, dim == 3 ? res[2] + 1 : 1};
    long N = (long)nn[0] * nn[1] * nn[2];
    long *eq = malloc(sizeof(long) * N * dim);
    long n0 = number_free(dim, nn, eq);
    if (n0 != nu) {
        free(eq);
        return -1;
    }
    M.L[0].A = (csr){nu, nu, (long *)Kptr, (int *)Kcol, (double *)Kval};
    M.L[0].n = nu;
    int l = 0;
    int cur[3] = {nn[0], nn[1], nn[2]};
    long *eqf = eq;
    while (1) {
        int even = 1;
        for (int d = 0; d < dim; d++) even &= ((cur[d] - 1) % 2 == 0) && (cur[d] - 1) >= 2;
        if (!even || M.L[l].n <= coarse_max || l == 15) break;
        int nxt[3] = {(cur[0] - 1) / 2 + 1, (cur[1] - 1) / 2 + 1, dim == 3 ? (cur[2] - 1) / 2 + 1 : 1};
        long Nc = (long)nxt[0] * nxt[1] * nxt[2];
        long *eqc = malloc(sizeof(long) * Nc * dim);
        long nc = number_free(dim, nxt, eqc);
        M.L[l].P = build_P(dim, cur, nxt, eqf, M.L[l].n, eqc, nc);
        M.L[l].R = transpose(&M.L[l].P);
        csr AP = spgemm(&M.L[l].A, &M.L[l].P);
        M.L[l + 1].A = spgemm(&M.L[l].R, &AP);
        csr_free(&AP);
        M.L[l + 1].n = nc;
        if (eqf != eq) free(eqf);
        eqf = eqc;
        memcpy(cur, nxt, sizeof(cur));
        l++;
    }
    if (eqf != eq) free(eqf);
    free(eq);
    M.nl = l + 1;
    for (int k = 0; k < M.nl; k++) level_alloc(&M.L[k]);
    /* dense Cholesky of the coarsest operator */
    {
        level_t *C = &M.L[M.nl - 1];
        long n = C->n;
        M.nc = n;
        M.chol = calloc((size_t)n * n, sizeof(double));
        for (long i = 0; i < n; i++)
            for (long k = C->A.ptr[i]; k < C->A.ptr[i + 1]; k++) M.chol[i * n + C->A.col[k]] = C->A.val[k];
        for (long j = 0; j < n; j++) {
            double s = M.chol[j * n + j];
            for (long k = 0; k < j; k++) s -= M.chol[j * n + k] * M.chol[j * n + k];
            M.chol[j * n + j] = sqrt(s);
            #pragma omp parallel for
            for (long i = j + 1; i < n; i++) {
                double t = M.chol[i * n + j];
                for (long k = 0; k < j; k++) t -= M.chol[i * n + k] * M.chol[j * n + k];
                M.chol[i * n + j] = t / M.chol[j * n + j];
            }
        }
    }
    csr G = {nu, np, (long *)Gptr, (int *)Gcol, (double *)Gval};
    csr Gt = transpose(&G);
    double t1 = omp_get_wtime();
    schur_ctx s = {&M, &G, &Gt, approxS, inner_rtol, nu, np, malloc(sizeof(double) * nu), malloc(sizeof(double) * nu),
                   malloc(sizeof(double) * nu), 0};
    /* h_hat = G^T K^{-1} f (h = 0: incompressible) */
    double *ustar = malloc(sizeof(double) * nu), *hh = malloc(sizeof(double) * np);
    K_solve(&s, f, ustar);
    spmv(&Gt, ustar, hh);
    memset(p, 0, sizeof(double) * np);
    int outer = 0;
    fgmres(np, S_op, &s, Sp_op, &s, hh, p, outer_rtol, max_outer, 30, &outer);
    /* back-solve u = K^{-1} (f - G p) */
    spmv(&G, p, s.t3);
    for (long i = 0; i < nu; i++) s.t3[i] = f[i] - s.t3[i];
    K_solve(&s, s.t3, u);
    double t2 = omp_get_wtime();
    stats[0] = outer;
    stats[1] = s.inner_its;
    stats[2] = M.vcycles;
    stats[3] = M.nl;
    times[0] = t1 - t0;
    times[1] = t2 - t1;
    free(ustar), free(hh), free(s.t1), free(s.t2), free(s.t3);
    csr_free(&Gt);
    M.L[0].A = (csr){0};
    mg_free(&M);
    return outer;
}

int orc_num_threads(void) { return omp_get_max_threads(); }
