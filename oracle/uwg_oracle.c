/*
 * uwg_oracle.c -- CPU restatement (FP64, scalar C) of Underworld2's particle-in-cell
 * Stokes hot path.  TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and
 * the timed CPU baseline.  Nothing in the product (underworld2_amd/, include/) links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg do.
 *
 * Every routine follows the reference operation order so that, compiled with
 * -ffp-contract=off, it reproduces the reference's IEEE results.  Citations are relative
 * to the reference tree; LU/ = src/underworld/libUnderworld/.
 *
 * Pinning (what checks each routine beyond retyping):
 *   orc_layout_random     PerCellRandomLayout docstring values (layouts.py:200-214)
 *   orc_layout_sobol      PerCellSpaceFillerLayout docstring values (layouts.py:160-172), the
 *                         Sobol sequence against its Gray-code closed form (XOR of the
 *                         direction numbers over the set bits of gray(n)) and against the
 *                         first terms of the standard Sobol sequence for these three axes
 *   orc_newton_local      closed form xi = 2(x-x0)/h - 1 on affine cells, rtol 1e-12
 *   orc_search_element    brute force on the integer rule incl. adversarial faces
 *   orc_dvc_cell          DVCWeightsSuite expected centroids/weights (3D and 2D), grid
 *                         connectivity fixture, and invariants (sum w = 2^d)
 *   orc_element_K_*       Gauss variant against testIsoviscousSiffness.matrix (PETSc
 *                         binary fixture); particle variant -> Gauss variant when the
 *                         particles are the Gauss points; symmetry and rigid modes
 *   orc_element_G_gauss   sum over elements telescopes to boundary terms only
 *   orc_advect_rk2        doctest [0.27856742,-0.12143258] (_timeintegration.py:131-147)
 *                         and exactness for affine velocity fields
 *   orc_update_owners     hand-worked swap-with-last/append sequences
 *   orc_q2_* / *_q2       Lagrange property, partition of unity, derivatives against
 *                         central differences, Newton closed form on affine cells, exact
 *                         interpolation and RK2 closed form for quadratic velocity fields
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#define ORC_EPS 1e-12   /* Num_Approx epsilon, LU/StGermain/Base/Foundation/src/Numerics.c:15 */

/* ------------------------------------------------------------------------------------ */
/* A0  mesh geometry.  LU/StgDomain/Mesh/src/CartesianGenerator.c:2001-2039             */
/* ------------------------------------------------------------------------------------ */
static double node_coord_1d(int i, int nnodes, double lo, double hi)
{
    /* inds <= contactDepth(0) -> crdMin ; inds >= sizes-1 -> crdMax ; else
       crdMin + 0.0 + ((double)i / (double)(n-1)) * (steps - 0.0*contactGeom) */
    if (i <= 0) return lo;
    if (i >= nnodes - 1) return hi;
    {
        double steps = hi - lo;
        return (lo + 0.0) + ((double)i / (double)(nnodes - 1)) * (steps - 0.0 * 0.0);
    }
}

void orc_mesh_coords(int dim, const int *res, const double *lo, const double *hi, double *out)
{
    int n[3] = {res[0] + 1, res[1] + 1, dim == 3 ? res[2] + 1 : 1};
    long idx = 0;
    for (int k = 0; k < n[2]; k++)
        for (int j = 0; j < n[1]; j++)
            for (int i = 0; i < n[0]; i++, idx++) {
                out[idx * dim + 0] = node_coord_1d(i, n[0], lo[0], hi[0]);
                out[idx * dim + 1] = node_coord_1d(j, n[1], lo[1], hi[1]);
                if (dim == 3) out[idx * dim + 2] = node_coord_1d(k, n[2], lo[2], hi[2]);
            }
}

/* element-local node order: lexicographic, x fastest (BilinearElementType.c:165-201,
   TrilinearElementType.c:216-275) */
static void element_nodes(int dim, const int *res, long e, long *nodes)
{
    long nx = res[0] + 1, ny = res[1] + 1;
    long i = e % res[0];
    long j = (e / res[0]) % res[1];
    long k = dim == 3 ? e / ((long)res[0] * res[1]) : 0;
    long base = i + nx * (j + ny * k);
    nodes[0] = base; nodes[1] = base + 1; nodes[2] = base + nx; nodes[3] = base + nx + 1;
    if (dim == 3) {
        long s = nx * ny;
        nodes[4] = base + s; nodes[5] = base + s + 1; nodes[6] = base + s + nx; nodes[7] = base + s + nx + 1;
    }
}

/* ------------------------------------------------------------------------------------ */
/* Shape functions, exactly as written in the element types                              */
/* ------------------------------------------------------------------------------------ */
static void shape(int dim, const double *xi, double *N)
{
    if (dim == 2) {   /* BilinearElementType.c:179-190 */
        double x = xi[0], y = xi[1];
        N[0] = 0.25 * (1.0 - x) * (1.0 - y);
        N[1] = 0.25 * (1.0 + x) * (1.0 - y);
        N[2] = 0.25 * (1.0 - x) * (1.0 + y);
        N[3] = 0.25 * (1.0 + x) * (1.0 + y);
    } else {          /* TrilinearElementType.c:216-233 */
        double x = xi[0], y = xi[1], z = xi[2];
        N[0] = 0.125 * (1.0 - x) * (1.0 - y) * (1.0 - z);
        N[2] = 0.125 * (1.0 - x) * (1.0 + y) * (1.0 - z);
        N[3] = 0.125 * (1.0 + x) * (1.0 + y) * (1.0 - z);
        N[1] = 0.125 * (1.0 + x) * (1.0 - y) * (1.0 - z);
        N[4] = 0.125 * (1.0 - x) * (1.0 - y) * (1.0 + z);
        N[6] = 0.125 * (1.0 - x) * (1.0 + y) * (1.0 + z);
        N[7] = 0.125 * (1.0 + x) * (1.0 + y) * (1.0 + z);
        N[5] = 0.125 * (1.0 + x) * (1.0 - y) * (1.0 + z);
    }
}

/* GNi[d][node]  (BilinearElementType.c:193-215, TrilinearElementType.c:236-275) */
static void shape_local_derivs(int dim, const double *xi, double GNi[3][8])
{
    if (dim == 2) {
        double x = xi[0], y = xi[1];
        GNi[0][0] = -0.25 * (1.0 - y);
        GNi[0][1] = 0.25 * (1.0 - y);
        GNi[0][3] = 0.25 * (1.0 + y);
        GNi[0][2] = -0.25 * (1.0 + y);
        GNi[1][0] = -0.25 * (1.0 - x);
        GNi[1][1] = -0.25 * (1.0 + x);
        GNi[1][3] = 0.25 * (1.0 + x);
        GNi[1][2] = 0.25 * (1.0 - x);
    } else {
        double x = xi[0], y = xi[1], z = xi[2];
        GNi[0][0] = -0.125 * (1.0 - y) * (1.0 - z);
        GNi[0][2] = -0.125 * (1.0 + y) * (1.0 - z);
        GNi[0][3] = 0.125 * (1.0 + y) * (1.0 - z);
        GNi[0][1] = 0.125 * (1.0 - y) * (1.0 - z);
        GNi[0][4] = -0.125 * (1.0 - y) * (1.0 + z);
        GNi[0][6] = -0.125 * (1.0 + y) * (1.0 + z);
        GNi[0][7] = 0.125 * (1.0 + y) * (1.0 + z);
        GNi[0][5] = 0.125 * (1.0 - y) * (1.0 + z);

        GNi[1][0] = -0.125 * (1.0 - x) * (1.0 - z);
        GNi[1][2] = 0.125 * (1.0 - x) * (1.0 - z);
        GNi[1][3] = 0.125 * (1.0 + x) * (1.0 - z);
        GNi[1][1] = -0.125 * (1.0 + x) * (1.0 - z);
        GNi[1][4] = -0.125 * (1.0 - x) * (1.0 + z);
        GNi[1][6] = 0.125 * (1.0 - x) * (1.0 + z);
        GNi[1][7] = 0.125 * (1.0 + x) * (1.0 + z);
        GNi[1][5] = -0.125 * (1.0 + x) * (1.0 + z);

        GNi[2][0] = -0.125 * (1.0 - x) * (1.0 - y);
        GNi[2][2] = -0.125 * (1.0 - x) * (1.0 + y);
        GNi[2][3] = -0.125 * (1.0 + x) * (1.0 + y);
        GNi[2][1] = -0.125 * (1.0 + x) * (1.0 - y);
        GNi[2][4] = 0.125 * (1.0 - x) * (1.0 - y);
        GNi[2][6] = 0.125 * (1.0 - x) * (1.0 + y);
        GNi[2][7] = 0.125 * (1.0 + x) * (1.0 + y);
        GNi[2][5] = 0.125 * (1.0 + x) * (1.0 - y);
    }
}

/* ------------------------------------------------------------------------------------ */
/* A1  Park-Miller minimal standard with Schrage's method. LU/StgDomain/Swarm/src/Random.c */
/* ------------------------------------------------------------------------------------ */
static double pm_next(long *s)
{
    const long q = 127773, r = 2836, a = 16807, m = 2147483647;
    const long tmp = *s / q;
    *s = a * (*s - tmp * q) - r * tmp;
    if (*s < 0) *s += m;
    return (double)*s;
}

static void local_to_global(int dim, const double *verts /*[8][dim]*/, const double *xi, double *x)
{
    /* FeMesh_CoordLocalToGlobal, LU/StgFEM/Discretisation/src/FeMesh.c:333-365 */
    double N[8];
    int nn = dim == 3 ? 8 : 4;
    shape(dim, xi, N);
    for (int d = 0; d < dim; d++) x[d] = 0.0;
    for (int a = 0; a < nn; a++)
        for (int d = 0; d < dim; d++) x[d] += N[a] * verts[a * dim + d];
}

static void element_verts(int dim, const int *res, const double *lo, const double *hi, long e, double *verts)
{
    long nodes[8];
    int nn = dim == 3 ? 8 : 4;
    long nx = res[0] + 1, ny = res[1] + 1;
    element_nodes(dim, res, e, nodes);
    for (int a = 0; a < nn; a++) {
        long n = nodes[a];
        long i = n % nx, j = (n / nx) % ny, k = n / (nx * ny);
        verts[a * dim + 0] = node_coord_1d((int)i, (int)nx, lo[0], hi[0]);
        verts[a * dim + 1] = node_coord_1d((int)j, (int)ny, lo[1], hi[1]);
        if (dim == 3) verts[a * dim + 2] = node_coord_1d((int)k, res[2] + 1, lo[2], hi[2]);
    }
}

/* PerCellRandomLayout (layouts.py:186-229) -> _PerCellParticleLayout_InitialiseParticles
   (PerCellParticleLayout.c:117-139) -> _MeshParticleLayout_InitialiseParticlesOfCell
   (MeshParticleLayout.c:194-229, filltype 1).  The seed is applied at layout init
   (MeshParticleLayout.c:74): state = seed ^ 123459876.  Particle index = cell*ppc + k. */
void orc_layout_random(int dim, const int *res, const double *lo, const double *hi,
                       int ppc, long seed, double *coords)
{
    long s = seed ^ 123459876L;
    long E = (long)res[0] * res[1] * (dim == 3 ? res[2] : 1);
    double verts[24];
    for (long e = 0; e < E; e++) {
        element_verts(dim, res, lo, hi, e, verts);
        for (int p = 0; p < ppc; p++) {
            double xi[3];
            for (int d = 0; d < dim; d++) {
                /* Swarm_Random_Random_WithMinMax(-1,1): (max-min)*r/RandMax + min */
                double r = pm_next(&s);
                xi[d] = (1.0 - -1.0) * r / (double)2147483647L + -1.0;
            }
            local_to_global(dim, verts, xi, coords + (e * ppc + p) * dim);
        }
    }
}

/* Sobol generator, one per axis (SobolGenerator.c; MeshParticleLayout.c:58-60 builds the
   table entries 0, 1, 2: degree 1/2/3, polynomial coefficient 0/1/1, initial direction numbers
   {1} / {1,1} / {1,3,7}).  The running value is an unsigned int (SobolGenerator.h:31; the
   0.5 assigned at init, SobolGenerator.c _Init, truncates to 0).  Direction number i >= degree:
   v_i = XOR over c = 0..degree-1 with c == degree-1 or bit (degree-2-c) of the coefficient
   set of 2^(c+1) v_(i-1-c), then XOR v_(i-degree) (SobolGenerator_GetDirectionalNumber).
   Draw: c = rightmost zero bit of the draw count, value ^= v_c << (30 - c - 1), returned as
   value / 2^30 (SobolGenerator_GetNextNumber, WORDLENGTH 30). */
typedef struct {
    unsigned degree, coef, count, value;
    unsigned v[32];
} sobol_t;

static void sobol_init(sobol_t *g, int axis)
{
    static const unsigned deg[3] = {1, 2, 3}, coef[3] = {0, 1, 1};
    static const unsigned init[3][3] = {{1, 0, 0}, {1, 1, 0}, {1, 3, 7}};
    g->degree = deg[axis];
    g->coef = coef[axis];
    g->count = 0;
    g->value = 0;
    for (unsigned i = 0; i < g->degree; i++) g->v[i] = init[axis][i];
    for (unsigned i = g->degree; i < 32; i++) {
        unsigned result = 0, dn = 0;
        for (unsigned c = 0; c < g->degree; c++) {
            if (c == g->degree - 1 || (g->coef & (1u << (g->degree - 2 - c)))) {
                dn = g->v[i - 1 - c];
                result ^= (1u << (c + 1)) * dn;
            }
        }
        result ^= dn;
        g->v[i] = result;
    }
}

static double sobol_next(sobol_t *g)
{
    unsigned bit = 0, n = g->count;
    while (n & 1u) {
        n >>= 1;
        bit++;
    }
    g->count++;
    g->value ^= g->v[bit] << (30 - bit - 1);
    return (double)g->value / (double)(1L << 30);
}

/* PerCellSpaceFillerLayout (layouts.py:146-181) -> _MeshParticleLayout_InitialiseParticlesOfCell
   (MeshParticleLayout.c:194-229, filltype 0): per particle one Park-Miller draw
   Swarm_Random_Random_WithMinMax(0, 1) (seeded at layout init with the default seed 13,
   MeshParticleLayout.c:153, :74); per axis, when that draw exceeds 0.9 the axis' Sobol
   generator skips one number; the local coordinate is (1 - -1) * next + -1. */
void orc_layout_sobol(int dim, const int *res, const double *lo, const double *hi, int ppc, long seed,
                      double *coords)
{
    long s = seed ^ 123459876L;
    long E = (long)res[0] * res[1] * (dim == 3 ? res[2] : 1);
    sobol_t gen[3];
    for (int d = 0; d < 3; d++) sobol_init(&gen[d], d);
    double verts[24];
    for (long e = 0; e < E; e++) {
        element_verts(dim, res, lo, hi, e, verts);
        for (int p = 0; p < ppc; p++) {
            double xi[3];
            double randnum = (1.0 - 0.0) * pm_next(&s) / (double)2147483647L + 0.0;
            for (int d = 0; d < dim; d++) {
                if (randnum > 0.9) (void)sobol_next(&gen[d]);
                xi[d] = (1.0 - -1.0) * sobol_next(&gen[d]) + -1.0;
            }
            local_to_global(dim, verts, xi, coords + (e * ppc + p) * dim);
        }
    }
}

/* Raw stream of one axis' Sobol generator (values in [0, 1)). */
void orc_sobol_stream(int axis, long n, double *out)
{
    sobol_t g;
    sobol_init(&g, axis);
    for (long i = 0; i < n; i++) out[i] = sobol_next(&g);
}

/* Raw Park-Miller stream (for jump-ahead tests of the device generator). */
void orc_rng_stream(long seed, long n, double *out)
{
    long s = seed ^ 123459876L;
    for (long i = 0; i < n; i++) out[i] = pm_next(&s);
}

/* ------------------------------------------------------------------------------------ */
/* A3  Newton global->local.  LU/StgFEM/Discretisation/src/ElementType.c:177-301 with    */
/*     Cramer's rule LU/StgDomain/Geometry/src/TensorMath.c:959-1020                     */
/* ------------------------------------------------------------------------------------ */
static void cramer(int dim, const double *J, double *s, const double *r)
{
    if (dim == 3) {
        double det = +J[0] * (J[4] * J[8] - J[7] * J[5])
                     - J[1] * (J[3] * J[8] - J[5] * J[6])
                     + J[2] * (J[3] * J[7] - J[4] * J[6]);
        s[0] = (r[0] * (J[4] * J[8] - J[7] * J[5])
                - J[1] * (r[1] * J[8] - J[5] * r[2])
                + J[2] * (r[1] * J[7] - J[4] * r[2])) / det;
        s[1] = (J[0] * (r[1] * J[8] - r[2] * J[5])
                - r[0] * (J[3] * J[8] - J[5] * J[6])
                + J[2] * (J[3] * r[2] - r[1] * J[6])) / det;
        s[2] = (J[0] * (J[4] * r[2] - J[7] * r[1])
                - J[1] * (J[3] * r[2] - r[1] * J[6])
                + r[0] * (J[3] * J[7] - J[4] * J[6])) / det;
    } else {
        double det = J[0] * J[3] - J[2] * J[1];
        s[0] = (r[0] * J[3] - J[1] * r[1]) / det;
        s[1] = (r[1] * J[0] - J[2] * r[0]) / det;
    }
}

/* returns number of iterations used; xi[0] = 1.1 if not converged */
int orc_newton_local(int dim, const double *verts, const double *x, double *xi)
{
    int nn = dim == 3 ? 8 : 4;
    for (int d = 0; d < dim; d++) xi[d] = 0.0;
    for (int it = 0; it < 100; it++) {
        double J[9] = {0}, rhs[3] = {0, 0, 0}, inc[3], N[8], GNi[3][8];
        shape(dim, xi, N);
        shape_local_derivs(dim, xi, GNi);
        for (int n = 0; n < nn; n++) {
            const double *c = verts + n * dim;
            double sf = N[n];
            J[0 * dim + 0] += GNi[0][n] * c[0];
            J[0 * dim + 1] += GNi[1][n] * c[0];
            J[1 * dim + 0] += GNi[0][n] * c[1];
            J[1 * dim + 1] += GNi[1][n] * c[1];
            rhs[0] -= sf * c[0];
            rhs[1] -= sf * c[1];
            if (dim == 3) {
                J[2] += GNi[2][n] * c[0];
                J[5] += GNi[2][n] * c[1];
                J[6] += GNi[0][n] * c[2];
                J[7] += GNi[1][n] * c[2];
                J[8] += GNi[2][n] * c[2];
                rhs[2] -= sf * c[2];
            }
        }
        rhs[0] += x[0];
        rhs[1] += x[1];
        if (dim == 3) rhs[2] += x[2];
        cramer(dim, J, inc, rhs);
        xi[0] += inc[0];
        xi[1] += inc[1];
        if (dim == 3) xi[2] += inc[2];
        {
            double m = fabs(inc[0]);
            if (m < fabs(inc[1])) m = fabs(inc[1]);
            if (dim == 3 && m < fabs(inc[2])) m = fabs(inc[2]);
            if (m < 0.0001) return it + 1;
        }
    }
    xi[0] = 1.1;
    return 100;
}

/* ------------------------------------------------------------------------------------ */
/* A17 regular search.  LU/StgDomain/Mesh/src/Mesh_RegularAlgorithms.c:174-206          */
/* sep computed once as (max-min)/nel (:153-154).  Returns 1 when found.               */
/* ------------------------------------------------------------------------------------ */
static int num_approx(double var, double val)
{
    return var >= val - ORC_EPS && var <= val + ORC_EPS;
}

int orc_search_element(int dim, const int *res, const double *lo, const double *hi,
                       const double *x, long *el)
{
    long inds[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) {
        double sep = (hi[d] - lo[d]) / res[d];
        if (num_approx(x[d] - hi[d], 0.0)) inds[d] = res[d] - 1;
        else if (x[d] < lo[d] || x[d] > hi[d]) return 0;
        else {
            double out = (x[d] - lo[d]) / sep, integer;
            double frac = modf(out, &integer);
            inds[d] = (long)(unsigned)integer;
            if (inds[d] > 0 && num_approx(frac, 0.0)) inds[d]--;
        }
    }
    *el = inds[0] + (long)res[0] * (inds[1] + (long)res[1] * inds[2]);
    return 1;
}

/* _ElementCellLayout_CellOf (ElementCellLayout.c:291-320): domain filter first, then
   search.  Returns E (the "outside" cell index) when not found. */
long orc_cell_of(int dim, const int *res, const double *lo, const double *hi, const double *x)
{
    long E = (long)res[0] * res[1] * (dim == 3 ? res[2] : 1), el;
    for (int d = 0; d < dim; d++)
        if (x[d] < lo[d] || x[d] > hi[d]) return E;
    if (!orc_search_element(dim, res, lo, hi, x, &el)) return E;
    return el;
}

/* A3 CoincidentMapper (CoincidentMapper.c:76-130): xi for every slot of every cell, in
   cell-table order.  cell_ptr[E+1], cell_pid[P] (CSR), coords[P][dim] -> xi_slot[P][dim] */
void orc_coincident_map(int dim, const int *res, const double *lo, const double *hi,
                        const long *cell_ptr, const int *cell_pid, const double *coords,
                        double *xi_slot)
{
    long E = (long)res[0] * res[1] * (dim == 3 ? res[2] : 1);
    double verts[24];
    for (long e = 0; e < E; e++) {
        element_verts(dim, res, lo, hi, e, verts);
        for (long s = cell_ptr[e]; s < cell_ptr[e + 1]; s++)
            orc_newton_local(dim, verts, coords + (long)cell_pid[s] * dim, xi_slot + s * dim);
    }
}

/* ------------------------------------------------------------------------------------ */
/* A4  Discrete Voronoi cell weights.  LU/PICellerator/Weights/src/DVCWeights.c          */
/* ------------------------------------------------------------------------------------ */
typedef struct { int nb[6]; double c[3]; int p, done; } dvc_cell;
typedef struct { int *claimed, nclaimed, *bound, nbound, cap; } dvc_chain;

/* _DVCWeights_ConstructGrid (:491-546) and _ConstructGrid2D (:551-587).  Neighbour slots
   are S,N,E,W,U,D; -2 marks "no neighbour".  Cell centres accumulate X = X + dx. */
static dvc_cell *dvc_grid(int dim, const int *num)
{
    int nx = num[0], ny = num[1], nz = dim == 3 ? num[2] : 1;
    int ncell = nx * ny * nz;
    dvc_cell *cells = (dvc_cell *)malloc(sizeof(dvc_cell) * ncell);
    double dx = (1.0 - -1.0) / nx, dy = (1.0 - -1.0) / ny, dz = (1.0 - -1.0) / nz;
    for (int i = 0; i < ncell; i++) {
        for (int k = 0; k < 6; k++) cells[i].nb[k] = -2;
        cells[i].p = -1;
        cells[i].done = 0;
    }
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx * (ny - 1); i++) {
            cells[i + k * nx * ny].nb[1] = i + k * nx * ny + nx;       /* N */
            cells[i + nx + k * nx * ny].nb[0] = i + k * nx * ny;       /* S */
        }
    if (dim == 3)
        for (int k = 0; k < nz - 1; k++)
            for (int i = 0; i < nx * ny; i++) {
                cells[i + k * nx * ny].nb[4] = i + k * nx * ny + nx * ny; /* U */
                cells[i + k * nx * ny + nx * ny].nb[5] = i + k * nx * ny; /* D */
            }
    {
        double Z = -1.0 - dz / 2.0;
        for (int k = 0; k < nz; k++) {
            double Y;
            if (dim == 3) Z = Z + dz;
            Y = -1.0 - dy / 2.0;
            for (int j = 0; j < ny; j++) {
                double X;
                Y = Y + dy;
                X = -1.0 - dx / 2.0;
                for (int i = 0; i < nx; i++) {
                    int id = i + nx * j + k * nx * ny;
                    X = X + dx;
                    cells[id].c[0] = X;
                    cells[id].c[1] = Y;
                    cells[id].c[2] = Z;
                    if (i != nx - 1) {
                        cells[id].nb[2] = id + 1;      /* E */
                        cells[id + 1].nb[3] = id;      /* W */
                    }
                }
            }
        }
    }
    return cells;
}

/* grid connectivity/centres dump for the ConstructGrid fixtures: out_nb[ncell][6]
   (N,S,E,W,U,D order as in the fixture), out_c[ncell][3] */
void orc_dvc_grid_dump(int dim, const int *num, int *out_nb, double *out_c)
{
    dvc_cell *cells = dvc_grid(dim, num);
    int ncell = num[0] * num[1] * (dim == 3 ? num[2] : 1);
    for (int i = 0; i < ncell; i++) {
        out_nb[i * 6 + 0] = cells[i].nb[1];
        out_nb[i * 6 + 1] = cells[i].nb[0];
        out_nb[i * 6 + 2] = cells[i].nb[2];
        out_nb[i * 6 + 3] = cells[i].nb[3];
        out_nb[i * 6 + 4] = cells[i].nb[4];
        out_nb[i * 6 + 5] = cells[i].nb[5];
        for (int d = 0; d < 3; d++) out_c[i * 3 + d] = cells[i].c[d];
    }
    free(cells);
}

static void chain_push(dvc_chain *ch, int which, int v)
{
    int *n = which ? &ch->nbound : &ch->nclaimed;
    int **a = which ? &ch->bound : &ch->claimed;
    (void)v;
    if (*n >= ch->cap) {
        ch->cap *= 2;
        ch->claimed = (int *)realloc(ch->claimed, sizeof(int) * ch->cap);
        ch->bound = (int *)realloc(ch->bound, sizeof(int) * ch->cap);
    }
    (*a)[(*n)++] = v;
}

/* _DVCWeights_UpdateBchain (:388-438) */
static void dvc_update_bchain(dvc_chain *ch, dvc_cell *cells, int pi, int nnb)
{
    ch->nbound = 0;
    for (int i = 0; i < ch->nclaimed; i++) {
        int c0 = ch->claimed[i];
        for (int k = 0; k < nnb; k++) {
            int c1 = cells[c0].nb[k];
            if (c1 != -2 && cells[c1].p != pi && cells[c1].done != 1) {
                chain_push(ch, 1, c1);
                cells[c1].done = 1;
            }
        }
    }
    for (int i = 0; i < ch->nbound; i++) cells[ch->bound[i]].done = 0;
}

/* _DVCWeights_DistanceTest (:598-600) and 2D (:595-597) */
static double dvc_dist_test(int dim, const double *c, const double *p1, const double *p2)
{
    double r = (p1[0] + p2[0] - c[0] - c[0]) * (p1[0] - p2[0]) + (p1[1] + p2[1] - c[1] - c[1]) * (p1[1] - p2[1]);
    if (dim == 3) r = r + (p1[2] + p2[2] - c[2] - c[2]) * (p1[2] - p2[2]);
    return r;
}

/* _DVCWeights_ClaimCells (:256-311) */
static void dvc_claim(dvc_chain *ch, dvc_cell *cells, const double *pos, int dim, int pi)
{
    ch->nclaimed = 0;
    for (int i = 0; i < ch->nbound; i++) {
        int c0 = ch->bound[i];
        if (cells[c0].p == -1) {
            chain_push(ch, 0, c0);
            cells[c0].p = pi;
        } else if (cells[c0].p != pi) {
            if (dvc_dist_test(dim, cells[c0].c, pos + cells[c0].p * 3, pos + pi * 3) > 0.0) {
                chain_push(ch, 0, c0);
                cells[c0].p = pi;
            }
        }
    }
}

/* One cell: _DVCWeights_Calculate{3D,2D} (:772-927) = CreateVoronoi (:676-765) +
   GetCentroids (:198-254).  xi_in/xi_out [nump][dim], w_out [nump].  num = grid res.
   Returns 0 on success, -1 if nump == 0 (the reference firewall, :796/:874). */
int orc_dvc_cell(int dim, const int *num, int nump, const double *xi_in, double *xi_out, double *w_out)
{
    int nx = num[0], ny = num[1], nz = dim == 3 ? num[2] : 1;
    int ncell = nx * ny * nz, nnb = dim == 3 ? 6 : 4;
    double dx = (1.0 - -1.0) / nx, dy = (1.0 - -1.0) / ny, dz = (1.0 - -1.0) / nz;
    double da = dim == 3 ? dx * dy * dz : dx * dy;
    dvc_cell *cells;
    dvc_chain *ch;
    double *pos;
    int *count;
    if (nump <= 0) return -1;
    cells = dvc_grid(dim, num);
    ch = (dvc_chain *)malloc(sizeof(dvc_chain) * nump);
    pos = (double *)calloc((size_t)nump * 3, sizeof(double));
    for (int i = 0; i < nump; i++) {
        ch[i].cap = 64;
        ch[i].claimed = (int *)malloc(sizeof(int) * 64);
        ch[i].bound = (int *)malloc(sizeof(int) * 64);
        ch[i].nclaimed = ch[i].nbound = 0;
        for (int d = 0; d < dim; d++) pos[i * 3 + d] = xi_in[i * dim + d];
    }
    for (int i = 0; i < nump; i++) {
        int k = (int)((pos[i * 3 + 0] - -1.0) / dx);
        int j = (int)((pos[i * 3 + 1] - -1.0) / dy);
        int l = dim == 3 ? (int)((pos[i * 3 + 2] - -1.0) / dz) : 0;
        int id;
        if (k == nx) k--;
        if (j == ny) j--;
        if (dim == 3 && l == nz) l--;
        /* memory guard only: local coordinates lie in [-1, 1] up to roundoff */
        k = k < 0 ? 0 : (k > nx - 1 ? nx - 1 : k);
        j = j < 0 ? 0 : (j > ny - 1 ? ny - 1 : j);
        l = l < 0 ? 0 : (l > nz - 1 ? nz - 1 : l);
        id = k + j * nx + l * nx * ny;
        cells[id].p = i;
        ch[i].nclaimed = 0;
        chain_push(&ch[i], 0, id);
        dvc_update_bchain(&ch[i], cells, i, nnb);
    }
    {
        int claimed = 1;
        while (claimed != 0) {
            claimed = 0;
            for (int i = 0; i < nump; i++) {
                dvc_claim(&ch[i], cells, pos, dim, i);
                claimed += ch[i].nclaimed;
                dvc_update_bchain(&ch[i], cells, i, nnb);
            }
        }
    }
    count = (int *)calloc(nump, sizeof(int));
    for (int i = 0; i < nump; i++)
        for (int d = 0; d < dim; d++) xi_out[i * dim + d] = 0.0;
    for (int c = 0; c < ncell; c++) {
        int p = cells[c].p;
        for (int d = 0; d < dim; d++) xi_out[p * dim + d] += cells[c].c[d];
        count[p]++;
    }
    for (int i = 0; i < nump; i++) {
        w_out[i] = count[i] * da;
        if (count[i] != 0)
            for (int d = 0; d < dim; d++) xi_out[i * dim + d] /= count[i];
    }
    for (int i = 0; i < nump; i++) { free(ch[i].claimed); free(ch[i].bound); }
    free(ch); free(pos); free(count); free(cells);
    return 0;
}

/* WeightsCalculator_CalculateAll (WeightsCalculator.c:114-150): every cell in order. */
int orc_dvc_all(int dim, const int *num, long E, const long *cell_ptr, const double *xi_in,
                double *xi_out, double *w_out)
{
    for (long e = 0; e < E; e++) {
        long s0 = cell_ptr[e], n = cell_ptr[e + 1] - s0;
        if (orc_dvc_cell(dim, num, (int)n, xi_in + s0 * dim, xi_out + s0 * dim, w_out + s0) != 0)
            return -1 - (int)(e & 0x3fffffff);
    }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Global derivatives.  ElementType_ShapeFunctionsGlobalDerivs, ElementType.c:305-447   */
/* ------------------------------------------------------------------------------------ */
static void global_derivs(int dim, const double *verts, const double *xi, double *detJ, double GNx[3][8])
{
    double GNi[3][8], jac[3][3], cof[3][3], D = 0.0;
    int nn = dim == 3 ? 8 : 4;
    shape_local_derivs(dim, xi, GNi);
    if (dim == 2) {
        jac[0][0] = jac[0][1] = jac[1][0] = jac[1][1] = 0.0;
        for (int n = 0; n < nn; n++) {
            const double *c = verts + n * 2;
            jac[0][0] = jac[0][0] + GNi[0][n] * c[0];
            jac[0][1] = jac[0][1] + GNi[0][n] * c[1];
            jac[1][0] = jac[1][0] + GNi[1][n] * c[0];
            jac[1][1] = jac[1][1] + GNi[1][n] * c[1];
        }
        D = jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0];
        {
            double tmp = jac[0][0];
            jac[0][0] = jac[1][1] / D;
            jac[1][1] = tmp / D;
            jac[0][1] = -jac[0][1] / D;
            jac[1][0] = -jac[1][0] / D;
        }
    } else {
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) jac[i][j] = 0.0;
        for (int n = 0; n < nn; n++) {
            const double *c = verts + n * 3;
            for (int i = 0; i < 3; i++) {
                jac[i][0] = jac[i][0] + GNi[i][n] * c[0];
                jac[i][1] = jac[i][1] + GNi[i][n] * c[1];
                jac[i][2] = jac[i][2] + GNi[i][n] * c[2];
            }
        }
        D = jac[0][0] * (jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1])
            - jac[0][1] * (jac[1][0] * jac[2][2] - jac[1][2] * jac[2][0])
            + jac[0][2] * (jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0]);
        cof[0][0] = jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1];
        cof[1][0] = -(jac[1][0] * jac[2][2] - jac[1][2] * jac[2][0]);
        cof[2][0] = jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0];
        cof[0][1] = -(jac[0][1] * jac[2][2] - jac[0][2] * jac[2][1]);
        cof[1][1] = jac[0][0] * jac[2][2] - jac[0][2] * jac[2][0];
        cof[2][1] = -(jac[0][0] * jac[2][1] - jac[0][1] * jac[2][0]);
        cof[0][2] = jac[0][1] * jac[1][2] - jac[0][2] * jac[1][1];
        cof[1][2] = -(jac[0][0] * jac[1][2] - jac[0][2] * jac[1][0]);
        cof[2][2] = jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0];
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) jac[i][j] = cof[i][j] / D;
    }
    *detJ = D;
    for (int dx = 0; dx < dim; dx++)
        for (int n = 0; n < nn; n++) {
            double v = 0.0;
            for (int dxi = 0; dxi < dim; dxi++) v = v + GNi[dxi][n] * jac[dx][dxi];
            GNx[dx][n] = v;
        }
}

/* ------------------------------------------------------------------------------------ */
/* A6  K_e by particle quadrature, isotropic D.                                          */
/*     _ConstitutiveMatrixCartesian_AssembleElement ConstitutiveMatrixCartesian.cpp:238-429 */
/*     D = eta*diag(2,2,(2),1,(1,1)) * (detJ*w), D*B (:573-666, isDiagonal branch).       */
/*     Ke is row-major [nd][nd]; Ke[col*d+i][row*d+j] as written in the reference.        */
/* ------------------------------------------------------------------------------------ */
void orc_element_K_particles(int dim, const double *verts, int np, const double *xi,
                             const double *w, const double *eta, double *Ke)
{
    int nn = dim == 3 ? 8 : 4, nd = nn * dim;
    for (int i = 0; i < nd * nd; i++) Ke[i] = 0.0;
    for (int p = 0; p < np; p++) {
        double GNx[3][8], detJ, D[6];
        global_derivs(dim, verts, xi + p * dim, &detJ, GNx);
        {
            double f = detJ * w[p], e = eta[p];
            if (dim == 2) {
                D[0] = 0.0 + 2.0 * e; D[1] = 0.0 + 2.0 * e; D[2] = 0.0 + e;
                for (int i = 0; i < 3; i++) D[i] *= f;
            } else {
                D[0] = 0.0 + 2.0 * e; D[1] = 0.0 + 2.0 * e; D[2] = 0.0 + 2.0 * e;
                D[3] = 0.0 + e; D[4] = 0.0 + e; D[5] = 0.0 + e;
                for (int i = 0; i < 6; i++) D[i] *= f;
            }
        }
        for (int a = 0; a < nn; a++) {
            int r = a * dim;
            if (dim == 2) {
                double ddx = GNx[0][a], ddy = GNx[1][a], DB[3][2];
                DB[0][0] = D[0] * ddx; DB[0][1] = 0.0;
                DB[1][0] = 0.0;        DB[1][1] = D[1] * ddy;
                DB[2][0] = D[2] * ddy; DB[2][1] = D[2] * ddx;
                for (int b = 0; b < nn; b++) {
                    int c = b * 2;
                    double bx = GNx[0][b], by = GNx[1][b];
                    Ke[(c) * nd + r] += bx * DB[0][0] + by * DB[2][0];
                    Ke[(c) * nd + r + 1] += bx * DB[0][1] + by * DB[2][1];
                    Ke[(c + 1) * nd + r] += by * DB[1][0] + bx * DB[2][0];
                    Ke[(c + 1) * nd + r + 1] += by * DB[1][1] + bx * DB[2][1];
                }
            } else {
                double ddx = GNx[0][a], ddy = GNx[1][a], ddz = GNx[2][a], DB[6][3];
                DB[0][0] = D[0] * ddx; DB[0][1] = 0.0; DB[0][2] = 0.0;
                DB[1][0] = 0.0; DB[1][1] = D[1] * ddy; DB[1][2] = 0.0;
                DB[2][0] = 0.0; DB[2][1] = 0.0; DB[2][2] = D[2] * ddz;
                DB[3][0] = D[3] * ddy; DB[3][1] = D[3] * ddx; DB[3][2] = 0.0;
                DB[4][0] = D[4] * ddz; DB[4][1] = 0.0; DB[4][2] = D[4] * ddx;
                DB[5][0] = 0.0; DB[5][1] = D[5] * ddz; DB[5][2] = D[5] * ddy;
                for (int b = 0; b < nn; b++) {
                    int c = b * 3;
                    double bx = GNx[0][b], by = GNx[1][b], bz = GNx[2][b];
                    for (int j = 0; j < 3; j++) {
                        Ke[(c) * nd + r + j] += bx * DB[0][j] + by * DB[3][j] + bz * DB[4][j];
                        Ke[(c + 1) * nd + r + j] += by * DB[1][j] + bx * DB[3][j] + bz * DB[5][j];
                        Ke[(c + 2) * nd + r + j] += bz * DB[2][j] + bx * DB[4][j] + by * DB[5][j];
                    }
                }
            }
        }
    }
}

/* 3-point Gauss-Legendre rule per dimension (GaussIntegrationSwarm, particleCount 3). */
static void gauss3(int dim, int *ng, double *gxi, double *gw)
{
    const double a = sqrt(3.0 / 5.0);
    const double x1[3] = {-a, 0.0, a}, w1[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    int n = 0;
    for (int k = 0; k < (dim == 3 ? 3 : 1); k++)
        for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++, n++) {
                gxi[n * dim + 0] = x1[i];
                gxi[n * dim + 1] = x1[j];
                if (dim == 3) gxi[n * dim + 2] = x1[k];
                gw[n] = w1[i] * w1[j] * (dim == 3 ? w1[k] : 1.0);
            }
    *ng = n;
}

/* K_e with 3^d Gauss points and a constant viscosity (the IsoviscousStiffness fixture). */
void orc_element_K_gauss(int dim, const double *verts, double eta, double *Ke)
{
    double gxi[27 * 3], gw[27], ge[27];
    int ng;
    gauss3(dim, &ng, gxi, gw);
    for (int g = 0; g < ng; g++) ge[g] = eta;
    orc_element_K_particles(dim, verts, ng, gxi, gw, ge, Ke);
}

/* A7  f_e.  _VectorAssemblyTerm_NA__Fn_AssembleElement VectorAssemblyTerm_NA__Fn.cpp:143-202.
       force[np][dim] is the body-force function value at each integration point. */
void orc_element_f(int dim, const double *verts, int np, const double *xi, const double *w,
                   const double *force, double *fe)
{
    int nn = dim == 3 ? 8 : 4;
    for (int i = 0; i < nn * dim; i++) fe[i] = 0.0;
    for (int p = 0; p < np; p++) {
        double GNx[3][8], detJ, N[8], factor;
        global_derivs(dim, verts, xi + p * dim, &detJ, GNx);
        shape(dim, xi + p * dim, N);
        factor = detJ * w[p];
        for (int a = 0; a < nn; a++)
            for (int i = 0; i < dim; i++) fe[a * dim + i] += factor * force[p * dim + i] * N[a];
    }
}

/* A8  G_e.  _GradientStiffnessMatrixTerm_AssembleElement GradientStiffnessMatrixTerm.c:129-221
       at 3^d Gauss points; dQ0 pressure shape function = 1. */
void orc_element_G_gauss(int dim, const double *verts, double *Ge)
{
    double gxi[27 * 3], gw[27];
    int ng, nn = dim == 3 ? 8 : 4;
    gauss3(dim, &ng, gxi, gw);
    for (int i = 0; i < nn * dim; i++) Ge[i] = 0.0;
    for (int g = 0; g < ng; g++) {
        double GNx[3][8], detJ;
        global_derivs(dim, verts, gxi + g * dim, &detJ, GNx);
        for (int a = 0; a < nn; a++)
            for (int i = 0; i < dim; i++)
                Ge[a * dim + i] += +gw[g] * (-detJ) * (GNx[i][a] * 1.0);
    }
}

/* A9  approxS_e = sum_v Ge[v] * (Ge[v] / Ke[v][v]).
       _UzawaPreconditionerTerm_AssembleElement UzawaPreconditionerTerm.c:124-209 */
double orc_element_approxS(int dim, const double *Ke, const double *Ge)
{
    int nd = (dim == 3 ? 8 : 4) * dim;
    double s = 0.0;
    for (int v = 0; v < nd; v++) s += Ge[v] * (Ge[v] / Ke[v * nd + v]);
    return s;
}

/* All-element driver for A6/A7/A8/A9 (the element loop of __StiffnessMatrix_NewAssemble,
   StiffnessMatrix.c:552-744, and ForceVector_GlobalAssembly_General, ForceVector.c:231-363).
   Integration data in slot order: xi_slot[P][d], w_slot[P]; eta_slot[P], force_slot[P][d]
   are the function values at each slot.  Outputs per element. */
void orc_assemble_elements(int dim, const int *res, const double *lo, const double *hi,
                           const long *cell_ptr, const double *xi_slot, const double *w_slot,
                           const double *eta_slot, const double *force_slot,
                           double *Ke_all, double *fe_all, double *Ge_all, double *S_all)
{
    long E = (long)res[0] * res[1] * (dim == 3 ? res[2] : 1);
    int nd = (dim == 3 ? 8 : 4) * dim;
    double verts[24];
    for (long e = 0; e < E; e++) {
        long s0 = cell_ptr[e];
        int np = (int)(cell_ptr[e + 1] - s0);
        double *Ke = Ke_all + e * nd * nd;
        element_verts(dim, res, lo, hi, e, verts);
        orc_element_K_particles(dim, verts, np, xi_slot + s0 * dim, w_slot + s0, eta_slot + s0, Ke);
        orc_element_f(dim, verts, np, xi_slot + s0 * dim, w_slot + s0, force_slot + s0 * dim, fe_all + e * nd);
        orc_element_G_gauss(dim, verts, Ge_all + e * nd);
        S_all[e] = orc_element_approxS(dim, Ke, Ge_all + e * nd);
    }
}

/* ------------------------------------------------------------------------------------ */
/* A15/A16  RK2 advection.  TimeIntegrand_SecondOrder TimeIntegrand.c:249-319 with the   */
/* first-order fallback :441-476 (allowFallbackToFirstOrder = True for SwarmAdvector,   */
/* PY/systems/_timeintegration.py:110); velocity lookup _SwarmAdvector_TimeDeriv         */
/* SwarmAdvector.c:238-312 -> _FeVariable_InterpolateValueAt FeVariable.c:389-403 ->     */
/* _FeVariable_InterpolateNodeValuesToElLocalCoord FeVariable.c:1007-1037.               */
/* ------------------------------------------------------------------------------------ */
static int velocity_at(int dim, const int *res, const double *lo, const double *hi,
                       const double *vel, const double *x, double *v)
{
    long el, nodes[8];
    double verts[24], xi[3], N[8];
    int nn = dim == 3 ? 8 : 4;
    if (!orc_search_element(dim, res, lo, hi, x, &el)) return 0;
    element_verts(dim, res, lo, hi, el, verts);
    orc_newton_local(dim, verts, x, xi);
    element_nodes(dim, res, el, nodes);
    shape(dim, xi, N);
    for (int d = 0; d < dim; d++) v[d] = 0.0;
    for (int a = 0; a < nn; a++)
        for (int d = 0; d < dim; d++) v[d] += vel[nodes[a] * dim + d] * N[a];
    for (int d = 0; d < dim; d++) if (isinf(v[d])) return 0;
    return 1;
}

/* coords[P][dim] updated in place.  Returns number of particles whose predictor lookup
   failed (the reference firewall); fallbacks counted in *nfallback. */
long orc_advect_rk2(int dim, const int *res, const double *lo, const double *hi,
                    const double *vel, long P, double *coords, double dt, long *nfallback)
{
    long bad = 0, fb = 0;
    for (long p = 0; p < P; p++) {
        double *x = coords + p * dim, start[3], v0[3], v1[3];
        for (int d = 0; d < dim; d++) start[d] = x[d];
        if (!velocity_at(dim, res, lo, hi, vel, x, v0)) { bad++; continue; }
        for (int d = 0; d < dim; d++) x[d] = start[d] + 0.5 * dt * v0[d];
        if (velocity_at(dim, res, lo, hi, vel, x, v1)) {
            for (int d = 0; d < dim; d++) x[d] = start[d] + dt * v1[d];
        } else {
            for (int d = 0; d < dim; d++) x[d] = start[d];
            velocity_at(dim, res, lo, hi, vel, x, v0);
            for (int d = 0; d < dim; d++) x[d] = start[d] + dt * v0[d];
            fb++;
        }
    }
    if (nfallback) *nfallback = fb;
    return bad;
}

/* Interpolate a nodal vector field at points (for tests). ok[P] = 1 if found. */
void orc_interpolate(int dim, const int *res, const double *lo, const double *hi,
                     const double *vel, long P, const double *x, double *v, int *ok)
{
    for (long p = 0; p < P; p++) ok[p] = velocity_at(dim, res, lo, hi, vel, x + p * dim, v + p * dim);
}

/* ------------------------------------------------------------------------------------ */
/* A15/A16 with a Q2 velocity field: the lookup through the Biquadratic / Triquadratic    */
/* element types (LU/StgFEM/Discretisation/src/Biquadratic.c:138-192, Triquadratic.c:    */
/* 193-374), the generic Newton ElementType.c:177-301 over the element's 3^d vertices     */
/* (C2Generator incidence order a = ax + 3 ay + 9 az on the vertex grid of 2 res + 1).    */
/* Pins (tests/test_q2_advect_oracle.py): Lagrange property and partition of unity of     */
/* the basis, derivatives against central differences, Newton = closed form on affine     */
/* cells, interpolation exact for quadratic fields, RK2 = closed form for quadratic        */
/* velocity fields.                                                                         */
/* ------------------------------------------------------------------------------------ */
/* 1D factors f(t) = (t-1, 1-t^2, t+1); prefactor names of Triquadratic_EvalBasis          */
static void q2_basis3(const double *xi, double *N)
{
    double x = xi[0], y = xi[1], z = xi[2];
    double f[3][3] = {{x - 1.0, 1.0 - x * x, x + 1.0}, {y - 1.0, 1.0 - y * y, y + 1.0},
                      {z - 1.0, 1.0 - z * z, z + 1.0}};
    double m0 = 0.5 * x, m1 = 0.5 * y, m2 = 0.5 * z;
    double m3 = 0.25 * x * y, m4 = 0.25 * x * z, m5 = 0.25 * y * z;
    double m6 = 0.125 * x * y * z;
    for (int az = 0; az < 3; az++)
        for (int ay = 0; ay < 3; ay++)
            for (int ax = 0; ax < 3; ax++) {
                int a = ax + 3 * ay + 9 * az;
                int mx = ax == 1, my = ay == 1, mz = az == 1;
                double pre;
                if (mx && my && mz) { N[a] = f[0][1] * f[1][1] * f[2][1]; continue; }
                if (mx && my) pre = m2;
                else if (mx && mz) pre = m1;
                else if (my && mz) pre = m0;
                else if (mx) pre = m5;
                else if (my) pre = m4;
                else if (mz) pre = m3;
                else pre = m6;
                N[a] = pre * f[0][ax] * f[1][ay] * f[2][az];
            }
}

static void q2_derivs3(const double *xi, double G[3][27])
{
    double x = xi[0], y = xi[1], z = xi[2];
    double f[3][3] = {{x - 1.0, 1.0 - x * x, x + 1.0}, {y - 1.0, 1.0 - y * y, y + 1.0},
                      {z - 1.0, 1.0 - z * z, z + 1.0}};
    double g[3][3] = {{x - 0.5, -2.0 * x, x + 0.5}, {y - 0.5, -2.0 * y, y + 0.5},
                      {z - 0.5, -2.0 * z, z + 0.5}};
    double m0 = 0.5 * x, m1 = 0.5 * y, m2 = 0.5 * z;
    double m3 = 0.25 * x * y, m4 = 0.25 * x * z, m5 = 0.25 * y * z;
    for (int az = 0; az < 3; az++)
        for (int ay = 0; ay < 3; ay++)
            for (int ax = 0; ax < 3; ax++) {
                int a = ax + 3 * ay + 9 * az;
                int mx = ax == 1, my = ay == 1, mz = az == 1;
                /* d/dxi: g_x(ax) * pre(y,z) * f_y * f_z */
                if (my && mz) G[0][a] = g[0][ax] * f[1][ay] * f[2][az];
                else G[0][a] = g[0][ax] * (my ? m2 : (mz ? m1 : m5)) * f[1][ay] * f[2][az];
                /* d/deta: g_y(ay) * pre(x,z) * f_x * f_z */
                if (mx && mz) G[1][a] = g[1][ay] * f[0][ax] * f[2][az];
                else G[1][a] = g[1][ay] * (mx ? m2 : (mz ? m0 : m4)) * f[0][ax] * f[2][az];
                /* d/dzeta: g_z(az) * pre(x,y) * f_x * f_y */
                if (mx && my) G[2][a] = g[2][az] * f[0][ax] * f[1][ay];
                else G[2][a] = g[2][az] * (mx ? m1 : (my ? m0 : m3)) * f[0][ax] * f[1][ay];
            }
}

static void q2_basis2(const double *xi, double *N)
{
    double x = xi[0], y = xi[1];
    double a0 = x - 1.0, b0 = y - 1.0, a1 = 1.0 - x * x, b1 = 1.0 - y * y, a2 = x + 1.0, b2 = y + 1.0;
    double m0 = 0.5 * x, m1 = 0.5 * y, m2 = 0.25 * x * y;
    N[0] = m2 * a0 * b0; N[1] = m1 * a1 * b0; N[2] = m2 * a2 * b0;
    N[3] = m0 * a0 * b1; N[4] = a1 * b1;      N[5] = m0 * a2 * b1;
    N[6] = m2 * a0 * b2; N[7] = m1 * a1 * b2; N[8] = m2 * a2 * b2;
}

static void q2_derivs2(const double *xi, double G[3][27])
{
    double x = xi[0], y = xi[1];
    double a0 = x - 1.0, b0 = y - 1.0, a1 = x + 1.0, b1 = y + 1.0;
    double a2 = 2.0 * x - 1.0, b2 = 2.0 * y - 1.0, a3 = 2.0 * x + 1.0, b3 = 2.0 * y + 1.0;
    double a4 = 1.0 - x * x, b4 = 1.0 - y * y;
    double m0 = 0.25 * x, m1 = 0.25 * y, m2 = -x * y;
    G[0][0] = m1 * a2 * b0; G[0][2] = m1 * a3 * b0; G[0][6] = m1 * a2 * b1; G[0][8] = m1 * a3 * b1;
    G[1][0] = m0 * a0 * b2; G[1][2] = m0 * a1 * b2; G[1][6] = m0 * a0 * b3; G[1][8] = m0 * a1 * b3;
    G[0][1] = m2 * b0; G[0][7] = m2 * b1; G[0][3] = 0.5 * a2 * b4; G[0][5] = 0.5 * a3 * b4;
    G[1][1] = 0.5 * a4 * b2; G[1][7] = 0.5 * a4 * b3; G[1][3] = m2 * a0; G[1][5] = m2 * a1;
    G[0][4] = -2.0 * x * b4; G[1][4] = -2.0 * y * a4;
}

/* basis and local derivatives for the tests */
void orc_q2_basis(int dim, const double *xi, double *N, double *G /*[3][27]*/)
{
    double GG[3][27];
    if (dim == 3) { q2_basis3(xi, N); q2_derivs3(xi, GG); }
    else { q2_basis2(xi, N); q2_derivs2(xi, GG); }
    for (int d = 0; d < dim; d++)
        for (int a = 0; a < (dim == 3 ? 27 : 9); a++) G[d * 27 + a] = GG[d][a];
}

static void q2_element_verts(int dim, const int *res, const double *lo, const double *hi, long e, double *v)
{
    long i = e % res[0], j = (e / res[0]) % res[1], k = dim == 3 ? e / ((long)res[0] * res[1]) : 0;
    int nn = dim == 3 ? 27 : 9;
    for (int a = 0; a < nn; a++) {
        v[a * dim + 0] = node_coord_1d((int)(2 * i + a % 3), 2 * res[0] + 1, lo[0], hi[0]);
        v[a * dim + 1] = node_coord_1d((int)(2 * j + (a / 3) % 3), 2 * res[1] + 1, lo[1], hi[1]);
        if (dim == 3) v[a * dim + 2] = node_coord_1d((int)(2 * k + a / 9), 2 * res[2] + 1, lo[2], hi[2]);
    }
}

/* ElementType.c:177-301 with the quadratic element's 9 / 27 nodes; returns iterations */
int orc_q2_newton_local(int dim, const double *verts, const double *x, double *xi)
{
    int nn = dim == 3 ? 27 : 9;
    for (int d = 0; d < dim; d++) xi[d] = 0.0;
    for (int it = 0; it < 100; it++) {
        double J[9] = {0}, rhs[3] = {0, 0, 0}, inc[3], N[27], G[3][27];
        if (dim == 3) { q2_basis3(xi, N); q2_derivs3(xi, G); }
        else { q2_basis2(xi, N); q2_derivs2(xi, G); }
        for (int n = 0; n < nn; n++) {
            const double *c = verts + n * dim;
            double sf = N[n];
            J[0 * dim + 0] += G[0][n] * c[0];
            J[0 * dim + 1] += G[1][n] * c[0];
            J[1 * dim + 0] += G[0][n] * c[1];
            J[1 * dim + 1] += G[1][n] * c[1];
            rhs[0] -= sf * c[0];
            rhs[1] -= sf * c[1];
            if (dim == 3) {
                J[2] += G[2][n] * c[0];
                J[5] += G[2][n] * c[1];
                J[6] += G[0][n] * c[2];
                J[7] += G[1][n] * c[2];
                J[8] += G[2][n] * c[2];
                rhs[2] -= sf * c[2];
            }
        }
        rhs[0] += x[0];
        rhs[1] += x[1];
        if (dim == 3) rhs[2] += x[2];
        cramer(dim, J, inc, rhs);
        xi[0] += inc[0];
        xi[1] += inc[1];
        if (dim == 3) xi[2] += inc[2];
        {
            double m = fabs(inc[0]);
            if (m < fabs(inc[1])) m = fabs(inc[1]);
            if (dim == 3 && m < fabs(inc[2])) m = fabs(inc[2]);
            if (m < 0.0001) return it + 1;
        }
    }
    xi[0] = 1.1;
    return 100;
}

static int velocity_at_q2(int dim, const int *res, const double *lo, const double *hi,
                          const double *vel, const double *x, double *v)
{
    long el;
    double verts[81], xi[3], N[27];
    int nn = dim == 3 ? 27 : 9;
    if (!orc_search_element(dim, res, lo, hi, x, &el)) return 0;
    q2_element_verts(dim, res, lo, hi, el, verts);
    orc_q2_newton_local(dim, verts, x, xi);
    if (dim == 3) q2_basis3(xi, N); else q2_basis2(xi, N);
    {
        long i = el % res[0], j = (el / res[0]) % res[1], k = dim == 3 ? el / ((long)res[0] * res[1]) : 0;
        long nx = 2L * res[0] + 1, ny = 2L * res[1] + 1;
        for (int d = 0; d < dim; d++) v[d] = 0.0;
        for (int a = 0; a < nn; a++) {
            long node = (2 * i + a % 3) + nx * ((2 * j + (a / 3) % 3) + ny * (dim == 3 ? 2 * k + a / 9 : 0));
            for (int d = 0; d < dim; d++) v[d] += vel[node * dim + d] * N[a];
        }
    }
    for (int d = 0; d < dim; d++) if (isinf(v[d])) return 0;
    return 1;
}

long orc_advect_rk2_q2(int dim, const int *res, const double *lo, const double *hi,
                       const double *vel, long P, double *coords, double dt, long *nfallback)
{
    long bad = 0, fb = 0;
    for (long p = 0; p < P; p++) {
        double *x = coords + p * dim, start[3], v0[3], v1[3];
        for (int d = 0; d < dim; d++) start[d] = x[d];
        if (!velocity_at_q2(dim, res, lo, hi, vel, x, v0)) { bad++; continue; }
        for (int d = 0; d < dim; d++) x[d] = start[d] + 0.5 * dt * v0[d];
        if (velocity_at_q2(dim, res, lo, hi, vel, x, v1)) {
            for (int d = 0; d < dim; d++) x[d] = start[d] + dt * v1[d];
        } else {
            for (int d = 0; d < dim; d++) x[d] = start[d];
            velocity_at_q2(dim, res, lo, hi, vel, x, v0);
            for (int d = 0; d < dim; d++) x[d] = start[d] + dt * v0[d];
            fb++;
        }
    }
    if (nfallback) *nfallback = fb;
    return bad;
}

void orc_interpolate_q2(int dim, const int *res, const double *lo, const double *hi,
                        const double *vel, long P, const double *x, double *v, int *ok)
{
    for (long p = 0; p < P; p++) ok[p] = velocity_at_q2(dim, res, lo, hi, vel, x + p * dim, v + p * dim);
}

/* Mesh_HexType_GetMinimumSeparationGeneral on the Q2 mesh: pairwise distances of every
   element's 3^d vertices */
double orc_q2_min_sep(int dim, const int *res, const double *lo, const double *hi)
{
    long E = (long)res[0] * res[1] * (dim == 3 ? res[2] : 1);
    int nn = dim == 3 ? 27 : 9;
    double ms = HUGE_VAL, v[81];
    for (long e = 0; e < E; e++) {
        q2_element_verts(dim, res, lo, hi, e, v);
        for (int a = 0; a < nn - 1; a++)
            for (int b = a + 1; b < nn; b++) {
                double m = 0.0, t;
                for (int d = 0; d < dim; d++) { t = v[a * dim + d] - v[b * dim + d]; m += t * t; }
                m = sqrt(m);
                if (m < ms) ms = m;
            }
    }
    return ms;
}

/* A18  SwarmAdvector_MaxDt (SwarmAdvector.c:370-385): 0.5*minSep/max_node|v|.
   minSep: Mesh_HexType_GetMinimumSeparationGeneral (Mesh_HexType.c:275-330), min over
   elements of all pairwise node distances.  |v|: FeVariable_GetScalarAtNode (:453-469). */
double orc_max_dt(int dim, const int *res, const double *lo, const double *hi, const double *vel)
{
    long E = (long)res[0] * res[1] * (dim == 3 ? res[2] : 1);
    long N = (long)(res[0] + 1) * (res[1] + 1) * (dim == 3 ? res[2] + 1 : 1);
    int nn = dim == 3 ? 8 : 4;
    double minSep = HUGE_VAL, vmax, verts[24];
    for (long e = 0; e < E; e++) {
        element_verts(dim, res, lo, hi, e, verts);
        for (int a = 0; a < nn - 1; a++)
            for (int b = a + 1; b < nn; b++) {
                double m = 0.0, t;
                for (int d = 0; d < dim; d++) { t = verts[a * dim + d] - verts[b * dim + d]; m += t * t; }
                m = sqrt(m);
                if (m < minSep) minSep = m;
            }
    }
    {
        double m = 0.0;
        for (int d = 0; d < dim; d++) m += vel[d] * vel[d];
        vmax = sqrt(m);
    }
    for (long n = 1; n < N; n++) {
        double m = 0.0, c;
        for (int d = 0; d < dim; d++) m += vel[n * dim + d] * vel[n * dim + d];
        c = sqrt(m);
        if (c > vmax) vmax = c;
    }
    return 0.5 * minSep / vmax;
}

/* ------------------------------------------------------------------------------------ */
/* A17 owner update with per-cell tables.  Swarm_UpdateAllParticleOwners SwarmClass.c:   */
/* 947-986, Swarm_UpdateParticleOwner :988-1014 (linear search for the in-cell index),   */
/* Swarm_RemoveParticleFromCell :1016-1030 (swap-with-last), Swarm_AddParticleToCell     */
/* (append).  Cell tables are in CSR with per-cell capacity: tbl[E][cap], cnt[E].        */
/* owner[p] == E marks a particle outside the domain.  Returns #particles outside.       */
/* ------------------------------------------------------------------------------------ */
long orc_update_owners(int dim, const int *res, const double *lo, const double *hi,
                       long P, const double *coords, int *owner, int *tbl, int *cnt, int cap,
                       int *overflow)
{
    long E = (long)res[0] * res[1] * (dim == 3 ? res[2] : 1), outside = 0;
    *overflow = 0;
    for (long p = 0; p < P; p++) {
        long nc = orc_cell_of(dim, res, lo, hi, coords + p * dim);
        if (nc != owner[p]) {
            if (owner[p] < E) {
                int c = owner[p], k;
                for (k = 0; k < cnt[c]; k++) if (tbl[(long)c * cap + k] == (int)p) break;
                tbl[(long)c * cap + k] = tbl[(long)c * cap + cnt[c] - 1];
                cnt[c]--;
            }
            if (nc == E) owner[p] = (int)E;
            else {
                if (cnt[nc] >= cap) { *overflow = 1; return -1; }
                tbl[nc * cap + cnt[nc]] = (int)p;
                cnt[nc]++;
                owner[p] = (int)nc;
            }
        }
        if (owner[p] == (int)E) outside++;
    }
    return outside;
}
