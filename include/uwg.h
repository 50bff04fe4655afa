/*
 * uwg.h -- C ABI of libuwg, the MI355X (gfx950) particle-in-cell Stokes engine.
 *
 * Plain pointers and sizes only.  Every function returns an int status:
 *   0  success
 *   >0 warning (UWG_WARN_NOT_CONVERGED: a Krylov solve hit its iteration limit;
 *      UWG_WARN_NONFINITE: a non-finite value was produced)
 *   <0 error; uwg_last_error() returns the message.
 * Host arrays are copied in/out; device state lives inside the opaque handles.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * reference tree; PY/ = src/underworld/, LU/ = src/underworld/libUnderworld/).
 */
#ifndef UWG_H
#define UWG_H
#ifdef __cplusplus
extern "C" {
#endif

#define UWG_OK 0
#define UWG_WARN_NOT_CONVERGED 1
#define UWG_WARN_NONFINITE 2
#define UWG_ERR_ARG -1
#define UWG_ERR_DEVICE -2
#define UWG_ERR_DOMAIN -3      /* particles outside the mesh / lookups failed */
#define UWG_ERR_UNDERRESOLVED -4 /* an element holds no particle (DVC firewall) */

typedef struct uwg_mesh uwg_mesh;
typedef struct uwg_field uwg_field;
typedef struct uwg_swarm uwg_swarm;
typedef struct uwg_stokes uwg_stokes;

const char *uwg_last_error(void);
/* device name / arch string, for logs */
int uwg_device_info(char *buf, int n);

/* ---- mesh: FeMesh_Cartesian('Q1/dQ0', res, min, max)  PY/mesh/_mesh.py:1046-1190,
 *      CartesianGenerator_CalcGeom LU/StgDomain/Mesh/src/CartesianGenerator.c:2001-2039 */
int uwg_mesh_create(int dim, const int *res, const double *minc, const double *maxc, uwg_mesh **out);
/* FeMesh_Cartesian('Q2/dPc1', res, min, max) (PY/mesh/_mesh.py:1105-1190): the same element
 * grid with (2 res + 1)^dim vertices (C2Generator_GenElementVertexInc
 * LU/StgFEM/Discretisation/src/C2Generator.c:147-260, element-local vertices lexicographic on
 * the 3^dim sub-grid) and a dPc1 pressure submesh of dim+1 values per element
 * (Inner2DGenerator.c:258-283); swarms live on the element grid */
int uwg_mesh_create_q2(int dim, const int *res, const double *minc, const double *maxc, uwg_mesh **out);
int uwg_mesh_destroy(uwg_mesh *m);
/* node coordinates [N][dim] (MeshVariable mesh.data) */
int uwg_mesh_coords(uwg_mesh *m, double *out);
/* minimum element separation (Mesh_HexType_GetMinimumSeparationGeneral, Mesh_HexType.c:275-330) */
int uwg_mesh_min_separation(uwg_mesh *m, double *out);

/* ---- fields: MeshVariable(mesh, nodeDofCount) PY/mesh/_meshvariable.py; nodal [N][ncomp]
 *      (N = vertices of the mesh), with on_elements=1 per element [E][ncomp] (the dQ0
 *      pressure submesh), with on_elements=2 [E][dim+1][ncomp] (the dPc1 submesh) */
int uwg_field_create(uwg_mesh *m, int ncomp, int on_elements, uwg_field **out);
int uwg_field_destroy(uwg_field *f);
int uwg_field_set(uwg_field *f, const double *host);
int uwg_field_get(uwg_field *f, double *host);

/* ---- swarm: Swarm(mesh) PY/swarm/_swarm.py; owner tables SwarmClass.c:947-1030 */
int uwg_swarm_create(uwg_mesh *m, uwg_swarm **out);
int uwg_swarm_destroy(uwg_swarm *s);
/* PerCellRandomLayout(swarm, ppc, seed) + populate_using_layout:
 * PY/swarm/layouts.py:186-229, MeshParticleLayout.c:194-229, Random.c:36-53 */
int uwg_swarm_populate_random(uwg_swarm *s, int ppc, long seed);
/* PerCellSpaceFillerLayout (PY/swarm/layouts.py:146-181 -> MeshParticleLayout.c:194-229,
 * filltype 0): per-axis Sobol sequences (SobolGenerator.c) with the Park-Miller skip draw per
 * particle; the reference's default seed is 13 */
int uwg_swarm_populate_sobol(uwg_swarm *s, int ppc, long seed);
/* add_particles_with_coordinates (PY/swarm/_swarm.py): appended in order, owners updated */
int uwg_swarm_add_particles(uwg_swarm *s, long n, const double *coords);
int uwg_swarm_count(uwg_swarm *s, long *P);
int uwg_swarm_get_coords(uwg_swarm *s, double *out);
/* overwrite coordinates (swarm.particleCoordinates.data[:] = ...) then update owners */
int uwg_swarm_set_coords(uwg_swarm *s, const double *in, long *outside);
int uwg_swarm_get_owners(uwg_swarm *s, int *out);
/* per-cell particle tables, cell_ptr[E+1], cell_pid[P] (cellParticleTbl order) */
int uwg_swarm_get_cell_tables(uwg_swarm *s, int *cell_ptr, int *cell_pid);
/* integer material index per particle (SwarmVariable 'int'), used by material tables */
int uwg_swarm_set_material(uwg_swarm *s, const int *mat);
/* Swarm.update_particle_owners (PY/swarm/_swarm.py:613-654) */
int uwg_swarm_update_owners(uwg_swarm *s, long *outside);
/* VoronoiIntegrationSwarm.repopulate (PY/swarm/_integration_swarm.py:118-146):
 * coincident map (CoincidentMapper.c:76-130) then, if use_dvc, DVC weights at resolution
 * dvc_res (DVCWeights.c:772-937, default 15); otherwise w = 2^dim / count. */
int uwg_swarm_integration_points(uwg_swarm *s, int use_dvc, int dvc_res);
/* integration points in slot order: xi_slot[P][dim], w_slot[P] */
int uwg_swarm_get_integration(uwg_swarm *s, double *xi_slot, double *w_slot);
int uwg_swarm_set_integration(uwg_swarm *s, const double *xi_slot, const double *w_slot);

/* ---- advection: SwarmAdvector(vel, swarm, order=2).integrate(dt, update_owners)
 * PY/systems/_timeintegration.py:115-157, TimeIntegrand.c:249-319, SwarmAdvector.c:238-312.
 * The velocity field may live on a Q1 or a Q2 mesh (Biquadratic / Triquadratic lookup). */
int uwg_advect_rk2(uwg_swarm *s, uwg_field *vel, double dt, int update_owners, long *outside);
/* SwarmAdvector.get_max_dt (SwarmAdvector.c:370-385) */
int uwg_advect_max_dt(uwg_field *vel, double *dt);
/* interpolate a nodal field of 1-9 components at points, v[n][ncomp]
 * (FeVariable_InterpolateValueAt, FeVariable.c:389-403) */
int uwg_field_interpolate(uwg_field *f, long n, const double *x, double *v, int *ok);
/* Allocation-free variant of uwg_field_interpolate (no counterpart in the reference): for
 * callers that cannot allocate device memory (code under graph capture, tight serving loops).
 * Points, values and flags are staged in a fixed device array of the library (one per GPU,
 * held until the process ends, about 6.5 MB) instead of per-call allocations; at most
 * UWG_NOALLOC_MAX_POINTS points per call (more is an argument error).  Calls that overlap on
 * one GPU share that array and race. */
#define UWG_NOALLOC_MAX_POINTS 65536
int uwg_field_interpolate_noalloc(uwg_field *f, long n, const double *x, double *v, int *ok);

/* ---- Stokes: uw.systems.Stokes(vel, p, fn_viscosity, fn_bodyforce, voronoi_swarm, conditions)
 * PY/systems/_stokes.py:106-317.  bcmask[N]: bit i set = component i Dirichlet; NULL means
 * free-slip on every wall.  The BC values are taken from vel at solve time. */
int uwg_stokes_create(uwg_field *vel, uwg_field *p, const unsigned char *bcmask, uwg_stokes **out);
int uwg_stokes_destroy(uwg_stokes *st);
/* material tables: eta[nmat], force[nmat][dim] (fn.branching.map over a material index) */
int uwg_stokes_set_materials(uwg_stokes *st, int nmat, const double *eta, const double *force);
/* per-particle function values (general fn_viscosity / fn_bodyforce evaluated at particles) */
int uwg_stokes_set_particle_values(uwg_stokes *st, long P, const double *eta, const double *force);
/* a swarm variable (SwarmVariable.data, [P][ncomp] doubles, ncomp 1-3) copied to the device
 * under a slot 0-7 for uwg_stokes_eval_particle_functions (SwarmVariableFn.cpp) */
int uwg_swarm_set_var(uwg_swarm *s, int slot, int ncomp, const double *data);
/* per-particle viscosity and body force evaluated on the device from a compiled function
 * graph (PY/function/_function.py, branching.py, math.py; LU/Underworld/Function/src
 * Binary.cpp, Relational.cpp, Map.cpp, Conditional.cpp, FeVariableFn.cpp): prog[nins][5] =
 * (op, dst, a, b, n) register instructions, consts[nconst], nodal fields[nfields] (1-3
 * components, interpolated at each particle); out_eta / out_force name the result registers
 * (-1: not produced).  Replaces uwg_stokes_set_particle_values without a host round trip. */
int uwg_stokes_eval_particle_functions(uwg_stokes *st, uwg_swarm *s, int nins, const int *prog, int nconst,
                                       const double *consts, int nfields, uwg_field *const *fields, int out_eta,
                                       int out_force);
/* the per-particle viscosity [P] and body force [P][dim] last set or evaluated */
int uwg_stokes_get_particle_values(uwg_stokes *st, long P, double *eta, double *force);
/* element assembly from the swarm's integration points: K (A6), f (A7), approxS (A9)
 * (SystemLinearEquations_MatrixSetup/VectorSetup, SystemLinearEquations.c:559-595) */
int uwg_stokes_assemble(uwg_stokes *st, uwg_swarm *s);
/* Gauss-quadrature assembly (no voronoi swarm: the GaussIntegrationSwarm of
 * PY/swarm/_integration_swarm.py:156-190, 3 points per dim) from function values at the
 * Gauss points: eta_g[E][3^dim], force_g[E][3^dim][dim] (NULL: zero force) */
int uwg_stokes_assemble_fields(uwg_stokes *st, const double *eta_g, const double *force_g);
/*      (Q2/dPc1 systems: the 5^dim Gauss points of the Q2 mesh, point index qx + 5 qy + 25 qz,
 *      PY/swarm/_integration_swarm.py:173-181) */

typedef struct {
  int method;        /* 0: Schur complement (BSSCR auglag driver, auglag-driver-DGTGD.c:64-584)
                        1: block-preconditioned FGMRES on the coupled system */
  int precision;     /* 0: FP64 throughout; 1: FP32 multigrid cycle inside FP64 Krylov */
  int learned;       /* 1: learned coarse correction (needs uwg_stokes_load_nn) */
  double outer_rtol; /* scr_ksp_rtol (default 1e-5, PY/systems/_options.py:33) */
  double inner_rtol; /* A11_ksp_rtol (default 1e-6, PY/systems/_bsscr.py:282) */
  int mg_levels;     /* 0: automatic */
  int max_outer_its;
  int max_inner_its;
  int restart;
  int smooth_degree; /* Chebyshev degree per pre/post smoothing */
  int verbose;
} uwg_solve_opts;

typedef struct {
  int pressure_its;
  int velocity_presolve_its;
  int velocity_pressuresolve_its;
  int velocity_backsolve_its;
  int velocity_total_its;
  int outer_its;     /* coupled FGMRES iterations (method 1) */
  int vcycles;
  int levels;
  int converged;
  int fallback_used; /* learned correction rejected in favour of multigrid */
  double setup_time, solve_time, total_time;
  double res_momentum, res_continuity, norm_f; /* FP64 true residuals after the solve */
} uwg_solve_stats;

void uwg_solve_opts_default(uwg_solve_opts *o);
/* StokesSolver.solve (PY/systems/_bsscr.py:322-502) including re-assembly when
 * reassemble != 0 and a swarm is attached */
int uwg_stokes_solve(uwg_stokes *st, const uwg_solve_opts *o, uwg_solve_stats *stats);
/* eqResiduals (PY/systems/_stokes.py:395-416; Stokes_SLE.c:204-256): |f-Ku-Gp|, |G^T u|, |f| */
int uwg_stokes_residuals(uwg_stokes *st, double *out3);

/* ---- operator hooks (tests and the learned-part trainer) */
/* which: 0 = K (level), 1 = G, 2 = G^T, 3 = one multigrid V-cycle on K (level 0), 4 = K as the
 * FP32 cycle applies it (FP32 vectors; bfloat16 weights on 3D levels when the cycle uses them),
 * 5 = K in FP32 with FP32 weights (4 and 5: Q1 levels; x rounded to FP32, y widened),
 * 6 = the multigrid residual kernel y = b - K x with b = x (FP64), 7 = the same kernel as the
 * FP32 cycle runs it (cycle weights; x rounded to FP32, y widened).
 * Q2/dPc1 systems: level 0 is the Q2 level (pressure vectors [E][dim+1]; which 0..3, and 4 =
 * its FP32 K), levels 1.. are the Q1 levels of the multigrid hierarchy below it (K only) */
int uwg_stokes_apply(uwg_stokes *st, int which, int level, const double *x, double *y);
int uwg_stokes_levels(uwg_stokes *st, int *nlev, int *res_out /*[nlev][3]*/);
/* per-element quadrature weights W[3^dim][E] of a level, element force fe[E][2^dim*dim],
 * approxS[E]; Q2/dPc1 level 0: W[E][5^dim], fe[E][3^dim][dim], approxS blocks
 * [E][dim+1][dim+1] (UzawaPreconditionerTerm.c:124-209) */
int uwg_stokes_get_assembly(uwg_stokes *st, int level, double *W, double *fe, double *approxS);
int uwg_stokes_get_rhs(uwg_stokes *st, double *f);
/* ---- learned coarse correction (no counterpart in the reference: a convolutional network
 * standing in for, or correcting, the multigrid coarse correction).  The network acts on the nodes of one multigrid
 * level; mode 0 replaces the coarse cycle below that level, mode 1 adds a learned
 * correction of the coarse cycle's remaining error.  FP32 cycles only (precision 1). */
int uwg_stokes_load_nn(uwg_stokes *st, int level, int nparams, const float *params, int channels, int depth);
int uwg_stokes_set_nn_mode(uwg_stokes *st, int mode);
/* fallback monitor: while the network is in use the FP64 outer residual must fall by at least
 * min_drop (fraction) over every `window` iterations, else the plain preconditioner takes over
 * (defaults 20 and 0.01) */
int uwg_stokes_set_nn_monitor(uwg_stokes *st, int window, double min_drop);
/* second fallback test (no counterpart in the reference): the plain preconditioner also takes
 * over when the first residual estimate after a Krylov restart (one iteration from the
 * recomputed FP64 residual) exceeds `factor` times the previous cycle's final estimate -- the
 * error left is then Arnoldi-basis rounding, which the diagonal Schur block removes faster
 * than the learned one (default 1.25; 0 disables the test) */
int uwg_stokes_set_nn_restart_rise(uwg_stokes *st, double factor);
/* diagnostic, no counterpart in the reference: with rtol > 0 the coupled FP64 solve (method 1,
 * precision 0) applies K^-1 in its block preconditioner by an inner FGMRES + V-cycle solve to
 * rtol instead of one V-cycle, so its outer count is the floor set by the Schur-block choice
 * (approxS or the learned stencil); 0 restores the V-cycle */
int uwg_stokes_set_velocity_block_solve(uwg_stokes *st, double rtol);
/* record up to max_samples vectors of `level` (every `every`-th visit) during multigrid
 * cycles: kind 0 the right-hand side restricted to the level, kind 1 the residual that the
 * cycle below the level leaves (b - K x after it); retrieved as float[count][N_level*dim] */
int uwg_stokes_capture(uwg_stokes *st, int level, int kind, int max_samples, int every);
int uwg_stokes_get_captured(uwg_stokes *st, float *out, int *count);
/* solve K_level x = b to rtol (FGMRES + cycles from that level, FP64) */
int uwg_stokes_level_solve(uwg_stokes *st, int level, const double *b, double *x, double rtol, int *its);
/* log10(eta / eta_ref) per node of a level (eta: element-mean viscosity averaged over the
 * elements around the node; eta_ref: geometric mean over the level), and its Dirichlet mask */
int uwg_stokes_level_feature(uwg_stokes *st, int level, float *feat, double *eta_ref);
int uwg_stokes_level_mask(uwg_stokes *st, int level, unsigned char *mask);

/* network training / evaluation on a node grid nx*ny*nz (inputs [B][4][N], outputs
 * [B][3][N], channel-major); loss = mean_b |m (y - t)|^2 / |m t|^2 */
typedef struct uwg_nn uwg_nn;
int uwg_nn_create(int nx, int ny, int nz, int channels, int depth, long seed, uwg_nn **out);
int uwg_nn_destroy(uwg_nn *h);
int uwg_nn_param_count(uwg_nn *h, long *n);
int uwg_nn_get_params(uwg_nn *h, float *out);
int uwg_nn_set_params(uwg_nn *h, const float *in);
int uwg_nn_forward(uwg_nn *h, int B, const float *x, float *y);
int uwg_nn_loss_grad(uwg_nn *h, int B, const float *x, const float *t, const float *mask, float *loss, float *grads);
int uwg_nn_set_data(uwg_nn *h, long S, const float *x, const float *t, const float *mask);
int uwg_nn_train(uwg_nn *h, int steps, int B, float lr, float weight_decay, long seed, float *losses);
int uwg_nn_eval(uwg_nn *h, int B, float *loss);

/* ---- learned Schur-complement stencil (engine extension, no counterpart in the reference):
 * a convolutional network on the element grid maps local features of approxS (viscosity
 * log-ratios to the face neighbours, wall flags) to a 27-point stencil C per element; the
 * coupled solver's Schur block then applies -D^-1/2 C D^-1/2 (D = approxS) instead of
 * -approxS^-1 (C = identity is the plain method).  Trained on (features, r^, e^) triples,
 * r^ = D^-1/2 S D^-1/2 e^ with S = G^T K^-1 G applied by the plain solver's operators. */
typedef struct uwg_snn uwg_snn;
int uwg_snn_create(int channels, int layers, long seed, uwg_snn **out);
int uwg_snn_destroy(uwg_snn *h);
int uwg_snn_param_count(uwg_snn *h, long *n);
int uwg_snn_get_params(uwg_snn *h, float *out);
int uwg_snn_set_params(uwg_snn *h, const float *in);
/* feat[9][nx*ny*nz] from approxS (host arrays) */
int uwg_snn_features(int nx, int ny, int nz, const double *approxS, float *feat);
/* y[B][27][N] = the network's stencil offsets (the stencil is identity + y) */
int uwg_snn_forward(uwg_snn *h, int nx, int ny, int nz, int B, const float *x, float *y);
/* batch mean of |C(x) r - t|^2 / |t|^2 and its parameter gradient (grads may be NULL) */
int uwg_snn_loss_grad(uwg_snn *h, int nx, int ny, int nz, int B, const float *x, const float *r, const float *t,
                      float *loss, float *grads);
/* training set held on the device: nsamples triples from an assembled 3D Stokes system */
int uwg_snn_add_samples(uwg_snn *h, uwg_stokes *st, int nsamples, long seed, double inner_rtol);
int uwg_snn_sample_count(uwg_snn *h, long *n);
int uwg_snn_clear_data(uwg_snn *h);
int uwg_snn_train(uwg_snn *h, int steps, int B, float lr, float weight_decay, long seed, long ntrain, float *losses);
int uwg_snn_eval(uwg_snn *h, long first, long count, float *loss);
/* attach a trained stencil network to a 3D Q1/dQ0 Stokes system (used by solves with learned
 * set, coupled method, precision 0 or 1; coefficients recomputed after every assembly) */
int uwg_stokes_load_schur_nn(uwg_stokes *st, int channels, int layers, long nparams, const float *params);
/* kernel-apply tuning knob: node levels per workgroup of the 3D K apply (kz3 > 0 fixes it on
 * every grid; kz3 < 0 restores the default, chosen per grid to fill the GPU's resident
 * workgroup slots evenly), element rows per block of the 2D K apply (kz2 > 0; default 32);
 * zero leaves a setting unchanged */
int uwg_set_apply_chunk(int kz3, int kz2);
/* 3D Q2 K apply kernel: sum-factorised (2, default), matrix cores (1) or the dense vector
 * kernel (0); uwg_set_q2_mfma(on) selects 1 or 0 */
int uwg_set_q2_apply_kernel(int kind);
int uwg_set_q2_mfma(int on);
/* element weights of the 3D FP32 multigrid cycle (precision 1): bfloat16 (1, default; the
 * element contractions take bf16 inputs, FP32 arithmetic and accumulation) or FP32 (0).
 * Process-wide; takes effect at the next solve. */
int uwg_set_cycle_bf16(int on);
/* Arnoldi basis of the coupled FP64 FGMRES at precision 1: FP32 storage (1, the default) or
 * FP64 (0).  Dot products, the Hessenberg problem, the preconditioned vectors, the iterate and
 * the stopping test stay FP64.  No counterpart in the reference (an engine extension).
 * Process-wide; takes effect at the next solve. */
int uwg_set_krylov_basis_fp32(int on);
/* STREAM-style device copy of n doubles (reads 8n bytes, writes 8n bytes); ms per copy.
 * Measures attainable HBM bandwidth and calibrates byte counters. */
int uwg_stream_copy(long n, int reps, double *ms);
/* time `reps` applies of K on a multigrid level; ms per apply.  precision 0: FP64; 1: the FP32
 * cycle's apply (bfloat16 weights on 3D levels when the cycle uses them); 2: FP32 weights */
int uwg_stokes_time_apply(uwg_stokes *st, int precision, int level, int reps, double *ms);

/* ---- population control: PopulationControl(swarm, deleteThreshold, splitThreshold,
 *      maxDeletions, maxSplits).repopulate() (PY/swarm/_population_control.py:13-140;
 *      LU/PICellerator/Utils/src/PCDVC.c:313-1085, non-aggressive path).  Splits particles whose
 *      DVC weight exceeds splitThreshold * 2^dim (at most maxSplits per cell, the copy placed at
 *      the centroid of the parent's Voronoi cell and appended), deletes those below
 *      deleteThreshold * 2^dim (at most maxDeletions per cell, swap-with-last).  src[i]
 *      (capacity src_cap >= P + E * maxSplits) receives the original index whose data particle i
 *      carries; *p_new the new particle count */
int uwg_swarm_population_control(uwg_swarm *s, double delete_threshold, double split_threshold, int max_deletions,
                                 int max_splits, long src_cap, long *src, long *p_new);

/* ---- SUPG advection-diffusion: AdvectionDiffusion(phiField, velocityField, fn_diffusivity,
 *      fn_sourceTerm, method="SUPG", conditions, phiDotField)
 *      (PY/systems/_advectiondiffusion.py:838-1010; AD/ = LU/Underworld/Utils/AdvectionDiffusion/src/).
 *      Q1 mesh; phi, phidot nodal scalars, vel the nodal velocity on the same mesh; bcmask: one
 *      byte per vertex, bit 0 set where phi is fixed (those vertices keep their values) */
typedef struct uwg_advdiff uwg_advdiff;
int uwg_advdiff_create(uwg_field *phi, uwg_field *phidot, uwg_field *vel, const unsigned char *bcmask,
                       uwg_advdiff **out);
int uwg_advdiff_destroy(uwg_advdiff *a);
/* diffusivity and source at the 3^dim Gauss points of every element, [E][3^dim], point index
 * qx + 3 qy + 9 qz (GaussIntegrationSwarm of the Q1 mesh); source may be NULL.  Values are
 * held until set again (fn_diffusivity / fn_sourceTerm evaluated by the caller) */
int uwg_advdiff_set_coefficients(uwg_advdiff *a, const double *kappa_g, const double *source_g);
/* AdvDiffMulticorrector gamma (0.5) and corrector passes (2) (Multicorrector.c:96-106), upwind
 * parameter function 0 exact / 1 doubly asymptotic / 2 critical (Residual.cpp:150-160, 557-606) */
int uwg_advdiff_set_scheme(uwg_advdiff *a, double gamma, int iterations, int upwind);
/* one step: AdvectionDiffusion.integrate(dt) -> _AdvDiffMulticorrector_Solve
 * (Multicorrector.c:158-200) with the residual of Residual.cpp:223-513 and the lumped mass of
 * LumpedMassMatrixForceTerm.c:88-129; updates phi and phidot in place */
int uwg_advdiff_integrate(uwg_advdiff *a, double dt);
/* get_max_dt(): AdvectionDiffusionSLE_CalculateDt (Timestep.c:36-110) */
int uwg_advdiff_max_dt(uwg_advdiff *a, double *dt);

#ifdef __cplusplus
}
#endif
#endif
