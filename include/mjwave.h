/*
 * mjwave.h -- C ABI of the MI355X (gfx950) explicit spectral-element wave
 * solver: continuous Gauss-Lobatto Q_r elements on quadrilaterals and
 * hexahedra, lumped (diagonal) mass, second-order leapfrog, for the scalar
 * acoustic and the isotropic elastic wave equations.
 *
 * The entry points are the operator / time-scheme plug-in surface that
 * Montjoie's leapfrog driver calls (paths relative to MONTJOIE/src of the
 * reference):
 *   - VirtualOdeSystem<T> operators (Instationary/TimeSchemes.hxx:35-151,
 *     implemented by VarInstationary_Base, Instationary/VarInstationaryInline.cxx:
 *     174-242): ApplyOperatorKh / Dh / DhMinusdtSh, SolveOperatorDh /
 *     DhPlusdtSh, AddScalarTimeSource, SetDirichletCondition;
 *   - VirtualTimeScheme<T> leapfrog (ModifiedEquationIterator,
 *     Instationary/ModifiedEquation.hxx:78-114, .cxx:99-124,295-345):
 *     SetInitialConditionS, Advance, GetIterate;
 *   - the matrix-free stiffness plug-in FemMatrixFreeClass_Base<T>::MltAddFree
 *     (Computation/FemMatrixFreeClass.hxx:281-287) with its setup
 *     (VarGeometryProblem.cxx:648-790, Elliptic/Helmholtz/VarHelmholtz.cxx:53-124,
 *     Hyperbolic/Acoustic/TimeAcoustic.cxx:889-1004).
 *
 * Conventions
 *   - Every function returns an int status: MJW_OK (0) or a negative code;
 *     the message is available from mjw_last_error(handle) (thread-local for
 *     handle == NULL).  The reference aborts instead (e.g. TimeAcoustic.cxx:
 *     243-249); the C++ adapter include/mjwave_adapter.hpp restores that.
 *   - Host pointers are caller-owned and never retained.  Functions with the
 *     suffix _dev take device pointers (hipMalloc / torch CUDA tensors) that
 *     live on the handle's device and are used on the handle's stream.
 *   - One handle owns one HIP stream; calls on one handle serialise, distinct
 *     handles are independent.
 *   - Global host vectors: node-major, components blocked as in Montjoie
 *     (u[m*N + g], Mesh/NumberMap.cxx:693).  With S > 1 shots each shot is
 *     one such vector, shots consecutive (u[(s*ncomp + m)*N + g]).  Device
 *     vectors use the solver's own node order and layout (mjw_state_dev).
 *   - Element-local nodes q = i + n*j (+ n*n*k), n = order+1, i along xi_1;
 *     1-D nodes are the Gauss-Lobatto points of [0,1].
 */
#ifndef MJWAVE_H
#define MJWAVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MJW_OK 0
#define MJW_ERR_ARG -1
#define MJW_ERR_HIP -2
#define MJW_ERR_STATE -3
#define MJW_ERR_UNSUPPORTED -4
#define MJW_ERR_NOMEM -5
#define MJW_ERR_GEOMETRY -6

/* precision of the bulk time stepping */
#define MJW_FP64 0      /* state, geometry, arithmetic in FP64                        */
#define MJW_FP32 1      /* FP32 state and arithmetic; geometry computed in FP64,
                           stored rounded to FP32                                     */
#define MJW_FP32_G64 2  /* FP32 state; geometry stored and applied in FP64             */

typedef struct mjw_problem mjw_problem;

/* ---------------------------------------------------------------- version */
int mjw_version(void);
const char *mjw_last_error(const mjw_problem *p);

/* --------------------------------------------------------------- lifecycle */
/*
 * Create a problem on `device`.  Replaces Montjoie's setup chain for the
 * explicit path: ComputeMeshAndFiniteElement + ComputeMassMatrix
 * (VarGeometryProblem.cxx:648-790 geometry, VarHelmholtz.cxx:53-124
 * coefficients, TimeAcoustic.cxx:889-1004 lumped mass) + ComputeStiffnessMatrix
 * (VarInstationary.cxx:1303-1420).
 *   dim 2|3, order 1..7, ncomp 1 (acoustic) | dim (isotropic elastic; in 2-D
 *   plane strain, Elliptic/Elastic/VarElastic.cxx:22-30);
 *   node_xyz[n_nodes*dim] positions of all global GL nodes;
 *   elem_nodes[n_elem*(order+1)^dim] element -> global node map;
 *   rho[n_elem], mu[n_elem] per element; lambda[n_elem] for elasticity
 *   (NULL otherwise);  dirichlet[n_nodes] 1 on homogeneous-Dirichlet nodes
 *   (NULL: none);  nshots >= 1 independent right-hand sides sharing the mesh;
 *   precision MJW_FP64 | MJW_FP32 | MJW_FP32_G64.
 * Geometric factors J w DF^{-1} mu DF^{-T} (acoustic) or J w, DF^{-1}
 * (elastic) are computed on the device in FP64 from node_xyz
 * (isoparametric map, exact for multilinear elements).
 */
int mjw_create(mjw_problem **out, int device, int dim, int order, int ncomp, int64_t n_nodes,
               const double *node_xyz, int64_t n_elem, const int32_t *elem_nodes,
               const double *rho, const double *mu, const double *lambda,
               const uint8_t *dirichlet, int nshots, int precision);
int mjw_destroy(mjw_problem *p);

/* problem sizes: [0]=dim [1]=order [2]=ncomp [3]=n_nodes [4]=n_elem [5]=ncolors
 * [6]=nshots [7]=precision [8]=device bytes held */
int mjw_info(const mjw_problem *p, int64_t *info9);

/* Tuning: launch shape of the time loop's stiffness kernel (0 = default):
 * for one shot the elements per workgroup (a column tile EPB, -EPB for one
 * thread per node, 1000 + EPB for the persistent column kernel); for 3-D
 * problems with several (interleaved) shots the shot lanes per column.  mjw_tile_choices lists the valid values.  Used by the autotuner
 * (montjoie_amd/tuner.py); every shape gives identical results. */
int mjw_set_tile(mjw_problem *p, int elems_per_block);
int mjw_get_tile(const mjw_problem *p, int *elems_per_block);
/* Block order of the stiffness launches: blocks that the hardware places on
 * one XCD (round-robin over nxcd dies) get one contiguous run of elements, so
 * neighbours share an L2.  nxcd = 1 disables it; default 8 (MI355X). */
int mjw_set_xcd_order(mjw_problem *p, int nxcd);
/* supported elements-per-block values (up to cap); returns their count */
int mjw_tile_choices(const mjw_problem *p, int *out, int cap);

/*
 * Geometry of the 3-D stiffness kernels.
 *   MJW_GEOM_STORED: the per-node factors (acoustic J w mu DF^{-1} DF^{-T},
 *     6 values; elastic sqrt(J w) DF^{-1}, dim^2 values) are computed once by
 *     mjw_create and read every step;
 *   MJW_GEOM_TRILINEAR: for meshes of straight-sided hexahedra (every node the
 *     image of the GL grid under the trilinear map of the element's 8
 *     vertices, checked by mjw_create to 1e-11 of the element diameter) the
 *     kernels rebuild DF of that map at each node from 21 coefficients per
 *     element -- the transformation Montjoie itself differentiates for
 *     non-curved elements (ComputeLocalMassMatrix, VarGeometryProblem.cxx:
 *     671-690) -- and read no per-node geometry.
 * mjw_create selects TRILINEAR whenever the mesh allows it (3-D), STORED
 * otherwise.  Results of the two modes agree to rounding.
 */
#define MJW_GEOM_STORED 0
#define MJW_GEOM_TRILINEAR 1
int mjw_set_geometry_mode(mjw_problem *p, int mode);
int mjw_get_geometry_mode(const mjw_problem *p, int *mode, int *trilinear_available);

/* ---------------------------------------------------------------- operators */
/*
 * Y = beta*Y + alpha*K*U with K = -A (stiffness), all nodes, one shot
 * (ApplyOperatorKh, VarInstationaryInline.cxx:174-179 -> MltAddScalarH1,
 * Computation/ProdMatVectScalarH1.cxx:679-727; elastic MltAddVectorH1,
 * ProdMatVectH1.cxx:1228-1276).  Vectors of length n_nodes*ncomp, FP64.
 */
int mjw_apply_kh(mjw_problem *p, double alpha, const double *U, double beta, double *Y);
/* same on device vectors in the problem's state precision (double or float),
 * caller's node order, all nshots shots (nshots*ncomp*n_nodes values) */
int mjw_apply_kh_dev(mjw_problem *p, double alpha, const void *U, double beta, void *Y);
/*
 * The device keeps its vectors in a solver node order: nodes grouped by the
 * set of element colours that touch them, then by first touch in the
 * kernels' element order, so that each colour pass of the stiffness kernel
 * reads whole groups and a cache line carries only nodes of that pass.
 * Every host-vector entry point converts to and from the caller's order;
 * solver_index[g] (n_nodes values) = position of the caller's node g in the
 * device vectors returned by mjw_state_dev.  No counterpart in the reference
 * (Montjoie's vectors follow its own numbering, Mesh/NumberMap.cxx:661-777).
 */
int mjw_node_order(const mjw_problem *p, int32_t *solver_index);

/*
 * Diagonal (lumped) mass operators on one host vector of n_nodes*ncomp FP64
 * values, caller's numbering (ContinuousUnsteadyMassMatrix.cxx:1309-1420,
 * matrices built by ContinuousMassMatrixAcoustic::Init,
 * Hyperbolic/Acoustic/TimeAcoustic.cxx:929-1004).  D = lumped mass with
 * Dirichlet rows 1; S = lumped damping sigma J w (mjw_set_damping; 0 without
 * damping and on Dirichlet rows).  beta = 0 overwrites Y.
 *   mjw_apply_dh:             Y = beta*Y + alpha*D*U          (ApplyOperatorDh)
 *   mjw_apply_sh:             Y = beta*Y + alpha*S*U          (ApplyOperatorSh)
 *   mjw_apply_dh_minus_dt_sh: Y = beta*Y + alpha*(D - dt/2 S)*U (ApplyOperatorDhMinusdtSh)
 *   mjw_solve_dh:             Y = D^{-1} Y                    (SolveOperatorDh)
 *   mjw_solve_dh_plus_dt_sh:  Y = (D + dt/2 S)^{-1} Y         (SolveOperatorDhPlusdtSh)
 * dt is the time step the reference bakes into DhMinusdtSh / DhPlusdtSh.
 */
int mjw_apply_dh(mjw_problem *p, double alpha, const double *U, double beta, double *Y);
int mjw_apply_sh(mjw_problem *p, double alpha, const double *U, double beta, double *Y);
int mjw_apply_dh_minus_dt_sh(mjw_problem *p, double dt, double alpha, const double *U,
                             double beta, double *Y);
int mjw_solve_dh(mjw_problem *p, double *Y);
int mjw_solve_dh_plus_dt_sh(mjw_problem *p, double dt, double *Y);
/* Homogeneous Dirichlet condition: Y = 0 on the Dirichlet rows of every
 * component (SetDirichletCondition / SetDirichletConditionSource with no
 * inhomogeneous Dirichlet data, VarInstationary.cxx:2466-2475,
 * VarInstationaryInline / ImposeNullDirichletCondition). */
int mjw_set_dirichlet_condition(mjw_problem *p, double *Y);
/* lumped mass diagonal (n_nodes doubles, Dirichlet rows 1) */
int mjw_get_mass(mjw_problem *p, double *D);

/* ------------------------------------------------------ sources, receivers */
/*
 * Point (Dirac) sources: for source k, shot index shot[k] (NULL: all 0),
 * position pts[k*dim], polarisation pol[k*ncomp] (NULL: 1).  Each point adds
 * pol phi_i(F^{-1}(x)) averaged over the elements containing x to the
 * source vector F of its shot; then F's Dirichlet rows and its entries with
 * |F_i| <= eps ||F|| are dropped (AddDiracSourceGen,
 * Source/DefineSourceElliptic.cxx:2439-2725; ComputeRightHandSide,
 * VarInstationary.cxx:620-675).  Points of one shot may share nodes.
 */
int mjw_set_sources(mjw_problem *p, int nsrc, const double *pts, const double *pol,
                    const int32_t *shot);
/*
 * Time factor f_n applied at leapfrog step n (the pulse evaluated at loop
 * time t_n + dt, ModifiedEquation.cxx:316; zero after 2.5/f0,
 * VarInstationary.cxx:2394).  table[nsteps] FP64, same for all shots.
 */
int mjw_set_pulse(mjw_problem *p, int nsteps, const double *table);
/*
 * Time order of the leapfrog (OrderTimeScheme = r LEAP_FROG,
 * All_TimeScheme::SetInputData, Instationary/TimeSchemes.cxx:760-767, and
 * ModifiedEquationIterator::SetOrder, ModifiedEquation.cxx:71-79): 2 (the
 * default) or 4, the modified-equation scheme
 *   (D + dt/2 S) U^{n+1} = 2 D U^n - (D - dt/2 S) U^{n-1} + dt^2 (F^n - A U^n)
 *                          + dt^4/12 D d^4U/dt^4 + dt^4/6 S d^3U/dt^3
 * with the derivatives from the operator (Advance, ModifiedEquation.cxx:
 * 347-407), started by a fourth-order Taylor expansion from u0, du0
 * (:124-170).  Each step applies the stiffness twice (three times with
 * damping); its CFL limit is 2 sqrt(3) / sqrt(lambda_max) (GetCFL, :217-225).
 * The source needs f and its first two derivatives at t_n + dt: mjw_set_pulse
 * and mjw_set_pulse_derivatives (zero where not given).  Call before
 * mjw_leapfrog_init.  PML layers, batches, the adjoint, mjw_energy and
 * mjw_profile refuse order 4.
 */
int mjw_set_time_order(mjw_problem *p, int order);
/* f'(t_n + dt) and f''(t_n + dt), n < nsteps: mjw_temporal_source_table with
 * nb_deriv 2 and 3 (the fourth-order scheme's d^3U/dt^3 and d^4U/dt^4 terms) */
int mjw_set_pulse_derivatives(mjw_problem *p, int nsteps, const double *d1, const double *d2);
/*
 * Time function of the sources for the operator-level entry below
 * (GetNewTimeSource + InitSource, VarInstationary.cxx:409-420,2359-2365):
 * MJW_TIME_RICKER is the Ricker primitive interpolated by the degree-24
 * Gauss-Lobatto polynomial on windows of length 0.5/f0 from t_init, moving
 * forward with the evaluation time (TimeSourceHyperbolic, Source/
 * TimeSource.cxx:776-880); eps_time is the window tolerance, Montjoie's
 * (final_time - initial_time)*1e-15 (VarInstationary.cxx:993); the source is
 * cut off from t = 2.5/f0 on.  MJW_TIME_NONE removes it.
 */
#define MJW_TIME_NONE 0
#define MJW_TIME_RICKER 1
int mjw_set_time_source(mjw_problem *p, int kind, double f0, double t_init, double eps_time);
/*
 * Sources given as whole vectors (a volume source f b, TypeSource =
 * SRC_VOLUME, DefineSourceElliptic.cxx:470-560): F[nshots][ncomp*n_nodes],
 * caller's numbering, component-blocked; replaces the sources of every shot
 * (as mjw_set_sources does), Dirichlet rows and entries below
 * eps*||F_shot|| dropped (VarInstationary.cxx:637-658).
 */
int mjw_set_source_vectors(mjw_problem *p, const double *F);
/*
 * out[g] = integral of f phi_g over the mesh with the elements' own
 * Gauss-Lobatto quadrature (the nodes): sum over the elements holding g of
 * f(x_g) J w.  f_q[n_elem][(order+1)^dim] gives f at each element's
 * quadrature points (caller's element order, local node order), so f may
 * jump between elements; out has n_nodes entries, caller's numbering.  This
 * is AddVolumetricSource's load of a volume source
 * (DefineSourceElliptic.cxx:856-960).
 */
int mjw_lumped_integral(mjw_problem *p, const double *f_q, double *out);
/*
 * The same for every TemporalSource of the explicit path (GetNewTimeSource,
 * VarInstationary.cxx:409-481; primitives Source/TimeSource.cxx), through the
 * same interpolator and windows of 0.5/f0, with `param` the keyword's value:
 *   MJW_TIME_RICKER            (no parameter)      cut-off 2.5/f0
 *   MJW_TIME_DERIVATIVE_RICKER (no parameter)      cut-off 2.5/f0
 *   MJW_TIME_MODIFIED_RICKER   centre time t0      cut-off (2.5 + t0)/(sqrt(2) f0)
 *   MJW_TIME_GAUSSIAN          plateau width >= 0  cut-off 2.5/f0 + width
 *   MJW_TIME_HARMONIC          final time tf       cut-off tf
 *   MJW_TIME_SINUS_GAUSSIAN    alpha > 0           cut-off 2 sqrt(ln(1e6)/alpha)
 * mjw_add_scalar_time_source then adds nothing from the cut-off on.
 */
#define MJW_TIME_DERIVATIVE_RICKER 2
#define MJW_TIME_MODIFIED_RICKER 3
#define MJW_TIME_GAUSSIAN 4
#define MJW_TIME_HARMONIC 5
#define MJW_TIME_SINUS_GAUSSIAN 6
int mjw_set_temporal_source(mjw_problem *p, int kind, double f0, double param, double t_init,
                            double eps_time);
/*
 * Y += alpha * f^(nb_deriv)(t) * F_shot on one host vector (n_nodes*ncomp,
 * caller's numbering), F_shot = the sparse source vector of that shot
 * (mjw_set_sources); nothing from the kind's cut-off on (2.5/f0 for the
 * Ricker; AddScalarTimeSource ->
 * AddScalarSourceAtTime, VarInstationary.cxx:2391-2420,2509-2513).
 * nb_deriv = 0 is the primitive, 1 the pulse (what Advance adds), up to 9.
 * Calls must come with non-decreasing t, as in the reference (the windows
 * move forward only).
 */
int mjw_add_scalar_time_source(mjw_problem *p, int shot, double alpha, double t, int nb_deriv,
                               double *Y);
/* Gauss-Lobatto rule on [0,1] of `order` (order+1 points, weights sum 1):
 * Globatto::ConstructQuadrature(r, QUADRATURE_LOBATTO)
 * (Quadrature/GaussLobatto.cxx:29-75).  Host only, FP64, order 1..24. */
int mjw_gauss_lobatto(int order, double *points, double *weights);
/*
 * Ricker pulse table through Montjoie's degree-24 Gauss-Lobatto interpolator
 * of the Ricker primitive (TimeSourceHyperbolic, Source/TimeSource.cxx:
 * 33-51,818-880; window eps VarInstationary.cxx:993; cut-off :416-420):
 * out[n] = pulse(t_begin + n*dt + dt), n < nsteps.  Host only, FP64.
 */
int mjw_ricker_pulse_table(double f0, double t_begin, double dt, int nsteps, double *out);
/*
 * Receivers: u(x_r) = sum_i phi_i(x_r) u_i on component comp[r] (NULL: 0)
 * (Output/GridInterpolation.cxx SismoPoint / SismoCircle, :836-850,
 * :2721-2761).  Traces are recorded for every shot.
 */
int mjw_set_receivers(mjw_problem *p, int nrec, const double *pts, const int32_t *comp);
/*
 * Receivers of values or first derivatives: deriv[r] = -1 records u(x_r),
 * deriv[r] = a records d u / d x_a (x_r) = sum_i (DF^{-T} grad phi_i)_a u_i in
 * the first element holding x_r.  The elastic SismoPoint output interpolates
 * grad u at the point and writes the stress C grad u
 * (Output/GridInterpolation.cxx :836-850 with compute_grad;
 * Elliptic/Elastic/VarElastic.cxx:257-282).  elem (may be NULL) receives the
 * element of each receiver, in the caller's element numbering, so the caller
 * can apply that element's medium.
 */
int mjw_set_receivers_grad(mjw_problem *p, int nrec, const double *pts, const int32_t *comp,
                           const int32_t *deriv, int64_t *elem);

/* ---------------------------------------------------------------- leapfrog */
/*
 * Set dt and initial state (SetInitialConditionS, ModifiedEquation.cxx:99-124):
 * U^0 = u0, U^1 = U^0 + dt du0 + dt^2/2 D^{-1}(-A U^0), Dirichlet nodes 0.
 * u0 / du0: host FP64 [nshots][n_nodes*ncomp] or NULL (zero).  The step
 * counter is reset to 0 and t_n = t_begin + n*dt.
 */
int mjw_leapfrog_init(mjw_problem *p, double t_begin, double dt, const double *u0,
                      const double *du0);
/*
 * Advance nsteps steps (Advance, ModifiedEquation.cxx:295-345):
 *   U^{n+1} = 2U^n - U^{n-1} + dt^2 D^{-1}(f_n b - A U^n), Dirichlet rows 0,
 * recording R U^n before each step (GiveIterate, ModifiedEquation.cxx:44).
 * traces: host FP64 [nsteps][nshots][nrec] or NULL.  Step loops are captured
 * in a HIP graph when use_graph != 0.
 */
int mjw_leapfrog_advance(mjw_problem *p, int nsteps, double *traces, int use_graph);
/*
 * mjw_leapfrog_advance without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same steps, kernels and
 * results as mjw_leapfrog_advance(p, nsteps, traces, 0); the trace rows are
 * staged in one static device buffer per GPU of 2^20 doubles (8 MiB, held
 * until the process ends), so (nsteps + 1) * nshots * nrec must not exceed
 * 1048576 (MJW_ERR_ARG otherwise: advance in shorter calls).  Every problem
 * on one GPU shares that buffer: calls that overlap on one GPU (several host
 * threads) race.  No graph and no timing (mjw_last_timing is not updated).
 */
int mjw_leapfrog_advance_static(mjw_problem *p, int nsteps, double *traces);
/* traces of the current iterate only (GiveFinalIterate): [nshots][nrec] */
int mjw_record(mjw_problem *p, double *traces);
/*
 * Probes: the current iterate (the one mjw_record samples) at arbitrary
 * points, every component, for snapshot grids (SismoPlane / SismoLine
 * outputs, Output/GridInterpolation.cxx:4975-5166 builds the points,
 * GenerateGridPoints; WriteSnapshot, Instationary/VarInstationary.cxx:798-853,
 * samples them).  mjw_set_probes locates the points (inside[i] = 0 for a
 * point outside the mesh, whose value is then 0, as TranslatePoint leaves it);
 * mjw_probe writes values[(s*npts + i)*ncomp + m] in FP64.
 */
int mjw_set_probes(mjw_problem *p, int npts, const double *pts, uint8_t *inside);
/* probes of values (deriv[i] = -1) or first derivatives (deriv[i] = a:
 * d u / d x_a, as mjw_set_receivers_grad) for the elastic snapshot files,
 * which hold each displacement component and each stress component
 * (<name>%04d_U<m>.dat, <name>%04d_dU<k>.dat; the grid output of
 * Output/OutputHarmonic.cxx:3009-3027 with the gradient turned into the
 * stress by Elliptic/Elastic/VarElastic.cxx:257-282); elem (may be NULL)
 * receives the element holding each point, -1 outside the mesh. */
int mjw_set_probes_grad(mjw_problem *p, int npts, const double *pts, const int32_t *deriv,
                        uint8_t *inside, int64_t *elem);
int mjw_probe(mjw_problem *p, double *values);
/* current pair (U^n, U^{n+1}) as FP64 host vectors [nshots][n_nodes*ncomp];
 * either may be NULL (GetIterate / GetIterate(k)). */
int mjw_get_state(mjw_problem *p, double *Un, double *Un1);
/* device pointers of U^n and U^{n+1} in the solver's layout (state
 * precision, solver node order: see mjw_node_order): shots consecutive, except
 * 3-D problems with nshots > 1, whose shots are interleaved
 * (u[(m*n_nodes + g)*nshots + s]: one gather serves every shot of a node) */
int mjw_state_dev(mjw_problem *p, void **Un, void **Un1);
/* current step index n */
int mjw_get_step(const mjw_problem *p, int64_t *n);

/*
 * Discrete leapfrog energy per shot, FP64 accumulation:
 * E^{n+1/2} = 1/2 ||(U^{n+1}-U^n)/dt||_D^2 + 1/2 (U^{n+1})^T A U^n,
 * conserved by the scheme without source.  Non-finite values are reported
 * as NaN (blow-up detection, cf. NormeMaxSolution VarInstationary.cxx:362-373).
 * energy[nshots].
 */
int mjw_energy(mjw_problem *p, double *energy);

/*
 * lambda_max(D^{-1} A) by Lanczos (FP64, device), dt_max = 2/sqrt(lambda)
 * (EvaluateCFL_Generic, Instationary/EvaluationCflWave.cxx:717-760; the
 * ARPACK call is replaced by Lanczos).
 */
int mjw_lambda_max(mjw_problem *p, int iters, uint64_t seed, double *lambda_max);

/* ------------------------------------------------------- adjoint gradient */
/*
 * Gradient of a receiver misfit with respect to per-element mu
 * (m_e = mu_e = c_e^2 with rho = 1).  No counterpart in the reference (it has
 * no time-domain adjoint); built on the same kernels.
 *   mjw_adjoint_forward: from a zero initial state (mjw_leapfrog_init with
 *   u0 = du0 = NULL), computes U^0..U^nsteps keeping nsnap checkpoint pairs
 *   (U^k, U^{k+1}) and returns the traces [nsteps+1][nrec] (rows 0..nsteps).
 *   mjw_adjoint_backward: given dJ/d(trace) with the same shape, runs the
 *   discrete adjoint leapfrog backward; the forward iterates are rebuilt in
 *   reverse time between checkpoints (the leapfrog is time reversible) in the
 *   same fused kernel as the adjoint, and grad[n_elem] = dJ/dmu_e is
 *   accumulated in FP64 (original element order).  Single-shot acoustic.
 */
int mjw_adjoint_forward(mjw_problem *p, int nsteps, int nsnap, double *traces);
int mjw_adjoint_backward(mjw_problem *p, const double *dJ_dtrace, double *grad);
/* device time (ms) of the last adjoint forward [0] and backward [1] */
int mjw_adjoint_timing(const mjw_problem *p, double *ms2);
/* Replace the per-element mu of an acoustic problem (rescales the stored
 * geometric factors; rho and the mass are unchanged).  Media update used by
 * the gradient / inversion loop; no counterpart in the reference (Montjoie
 * rebuilds the problem from the data file). */
int mjw_set_mu(mjw_problem *p, const double *mu);

/* Media varying inside the elements: rho_q and mu_q hold rho and mu
 * at every quadrature point, [n_elem][(r+1)^dim] in the caller's element order
 * and the local node order of mjw_create.  Replaces ScalarPhysicalIndice /
 * TensorPhysicalIndice with a varying field (PhysicalMedia = rho|mu ref
 * ISOTROPE SINUS|MESH|..., Elliptic/PhysicalProperty.cxx:3614-3710), whose
 * values at the quadrature points VarGeometryProblem.cxx:1395-1560 stores
 * and VarHelmholtz.cxx:73-117 reads for the mass rho J w and the stiffness
 * tensor J w mu DF^{-1} DF^{-T}.  Rebuilds the geometric factors and the lumped
 * mass; the problem then uses stored geometry (MJW_GEOM_TRILINEAR is refused),
 * mjw_set_mu and the adjoint are refused, and mjw_leapfrog_init must be called
 * again.  Elastic problems take rho_q only (mu_q = NULL): the density enters
 * the mass rho J w alone (TimeElastic, ElasticEquation::GetTensorMass,
 * VarElastic.cxx:59-71); a varying stiffness tensor C is not supported. */
int mjw_set_media_nodal(mjw_problem *p, const double *rho_q, const double *mu_q);
/*
 * Damping sigma >= 0 per element (acoustic: rho u'' + sigma u' - div(mu grad u)
 * = f, the sigma of MateriauDielec = ref ISOTROPE rho mu sigma,
 * Elliptic/Helmholtz/VarHelmholtz.cxx:64-66,108-109; elastic:
 * rho u'' + sigma u' - div(C : grad u) = f, the fourth value of
 * MateriauDielec = ref ISOTROPE rho lambda mu sigma,
 * Elliptic/Elastic/VarElastic.cxx:59-71).  S_h = sigma J w per node (the same
 * for every component), lumped like the mass; the leapfrog then solves
 * (D + dt/2 S) U^{n+1} = 2 D U^n - (D - dt/2 S) U^{n-1} + dt^2 (F - A U^n)
 * (Advance, ModifiedEquation.cxx:295-345, with DhMinusdtSh / DhPlusdtSh,
 * Hyperbolic/Acoustic/TimeAcoustic.cxx:970-1004) and starts from
 * U^1 = U^0 + dt du0 + dt^2/2 D^{-1}(-A U^0 - S du0) (:99-124).  Call before
 * mjw_leapfrog_init; all-zero sigma removes the damping.  Adjoint, batch and
 * mjw_profile refuse damped problems.
 */
int mjw_set_damping(mjw_problem *p, const double *sigma);
/* damping sigma at every quadrature point, [n_elem][nd] in local order (a
 * PhysicalMedia sigma field, Elliptic/Helmholtz/VarHelmholtz.cxx:1219-1226) */
int mjw_set_damping_nodal(mjw_problem *p, const double *sigma_q);

/* ------------------------------------------------- PML layers (LEAP_FROG_PML) */
/*
 * Perfectly matched layers with Montjoie's first-order split formulation and
 * its leapfrog (OrderTimeScheme = 2 LEAP_FROG_PML: ModifiedEquationSystemIterator,
 * Instationary/ModifiedEquation.hxx:123-220, .cxx:441-481,511-560,773-812).
 * The mesh given to mjw_create already holds the layers (AddPML extrudes them,
 * Mesh/MeshBoundaries.cxx:114-120 -> Mesh2DBoundaries.cxx:104-192); pml[e] != 0
 * marks the elements inside them and tau[(e*nd + q)*dim + a] >= 0 is the
 * damping sigma_a at their quadrature points (the GL nodes, caller's element
 * order; GetDampingTauPML, Harmonic/BoundaryConditionHarmonic.cxx:1639-1760;
 * ignored outside the layers).  The unknowns become u on every node, its split
 * copies u* = u_x - u_y (and u** = u_x - u_z in 3-D) on the nodes of PML
 * elements, and the vector V at every quadrature point:
 *   D dU/dt = R^S V - S U + f(t) b,   B dV/dt = R^V U - S^V V
 * (element matrices Elliptic/Helmholtz/VarHelmholtz.cxx:133-230,296-378,
 * 620-660,3277-3720; block-diagonal D - dt/2 S, (D + dt/2 S)^{-1},
 * Hyperbolic/Acoustic/TimeAcoustic.cxx:1079-1235).  mjw_leapfrog_init then
 * starts this scheme from zero data (u0 = du0 = NULL) and mjw_leapfrog_advance
 * runs it: traces row n = R u^n; the pulse table of mjw_set_pulse holds the
 * source's primitive at t_n + dt/2 (mjw_time_source_table with nb_deriv = 0,
 * offset = dt/2, ModifiedEquation.cxx:784-785).  Acoustic, one shot, per-element
 * media, no physical damping; energy, adjoint, batches and mjw_profile refuse
 * such problems, and mjw_get_state returns u^n only.  Call once.
 */
int mjw_set_pml(mjw_problem *p, const uint8_t *pml, const double *tau);
/*
 * out[n] = f^(nb_deriv)(t_begin + n*dt + offset), n < nsteps, for the Ricker
 * source through Montjoie's degree-24 Gauss-Lobatto interpolator (window
 * tolerance nsteps*dt*1e-15, zero from 2.5/f0 on); nb_deriv 0 (primitive) to 9.
 * mjw_ricker_pulse_table is the case nb_deriv = 1, offset = dt.  Host only.
 */
int mjw_time_source_table(double f0, double t_begin, double dt, int nsteps, int nb_deriv,
                          double offset, double *out);
/* The same table for any MJW_TIME_* kind and its parameter (see
 * mjw_set_temporal_source), zero from the kind's cut-off on: the leapfrog's
 * mjw_set_pulse table is nb_deriv = 1, offset = dt.  Host only. */
int mjw_temporal_source_table(int kind, double f0, double param, double t_begin, double dt,
                              int nsteps, int nb_deriv, double offset, double *out);

/* --------------------------------------------------------- synthetic meshes */
/*
 * Regular Q_r mesh of [0,1]^dim with nx elements per direction (Montjoie's
 * REGULAR mesh, Mesh/Mesh3D.cxx:8406-8486), interior vertices displaced by
 * at most pert*h (seeded splitmix64; pert = 0: affine mesh).  Node positions
 * are the multilinear images of the GL nodes.  Sizes from mjw_mesh_regular_size.
 * Outputs: xyz[N*dim], elem_nodes[E*(r+1)^dim], dirichlet[N] (boundary of the
 * unit square/cube).
 */
int mjw_mesh_regular_size(int dim, int nx, int order, int64_t *n_nodes, int64_t *n_elem);
int mjw_mesh_regular(int dim, int nx, int order, double pert, uint64_t seed, double *xyz,
                     int32_t *elem_nodes, uint8_t *dirichlet);

/* ------------------------------------------------------------- batching */
/*
 * Many independent problems (different meshes, orders, sizes) advanced
 * together: each step runs every problem's colour passes as one launch per
 * colour over a flattened work list, the whole step loop captured in one HIP
 * graph.  All problems must share the precision; each keeps its own dt,
 * sources, receivers and state (initialise them with mjw_leapfrog_init).
 */
typedef struct mjw_batch mjw_batch;
/* single-shot 3-D acoustic problems, orders 1..7, same device and precision */
int mjw_batch_create(mjw_batch **out, int nprob, mjw_problem *const *probs);
/* advance every problem nsteps (all at the same step parity); traces of each
 * problem are kept on the device, rows 0..nsteps-1 = R U^n before each step */
int mjw_batch_advance(mjw_batch *b, int nsteps, int use_graph);
/* traces [nsteps][nrec] of problem iprob from the last mjw_batch_advance */
int mjw_batch_traces(mjw_batch *b, int iprob, double *traces);
/* tuning: order_mode 0 = one launch per colour, work items by decreasing
 * element order (default), 1 = the same launch problem by problem; 2, 3, 4 =
 * one launch per (colour, element order), blocks of about 256, 512 or 128
 * threads with LDS sized for that order; every mode gives the same bits;
 * nxcd reserved */
int mjw_batch_config(mjw_batch *b, int nxcd, int order_mode);
/* device ms per step of the last mjw_batch_advance */
int mjw_batch_timing(const mjw_batch *b, double *ms);
const char *mjw_batch_last_error(const mjw_batch *b);
int mjw_batch_destroy(mjw_batch *b);

/* ------------------------------------------------------------- timing */
/* Average device duration (ms) of the most recent leapfrog_advance's step,
 * measured with HIP events on the problem's stream; [0]=per step,
 * [1]=stiffness kernels per step.  For benchmarks. */
int mjw_last_timing(const mjw_problem *p, double *ms2);
/* Advance nsteps eager steps with HIP events on the problem's stream before
 * and after each step's colour passes (launched back to back) and after its
 * tail kernel: out[0] = mean stiffness time per colour pass ms (step's
 * passes / ncolors), out[1] = mean tail-kernel ms, out[2] = mean step ms,
 * out[3] = launches per step.  Advances the state like mjw_leapfrog_advance
 * (no traces). */
int mjw_profile(mjw_problem *p, int nsteps, double *out4);

#ifdef __cplusplus
}
#endif
#endif
