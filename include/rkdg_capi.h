/*
 * rkdg_capi.h — C-ABI of the MI355X-native explicit RKDG right-hand side.
 *
 * The entry points replace the reference's hot-path interfaces (vkorchagova/RKDG3D):
 *
 *   rkdg_create / rkdg_destroy   FE_Evolution ctor + UpdateInverseMassMatrix +
 *                                UpdateAfluxPointer after a re-mesh
 *                                (src/fe_evolution/fe_evolution.hpp:66-81,
 *                                 main.cpp:148-164, 610-611)
 *   rkdg_rhs                     FE_Evolution::Mult(const Vector&, Vector&)
 *                                (src/fe_evolution/fe_evolution.hpp:75,
 *                                 fe_evolution.cpp:102-216); the lambda_max
 *                                out-parameter replaces the global
 *                                max_char_speed (fe_evolution.hpp:10)
 *   rkdg_face_lambda             A.Mult(sol, z) used for the first dt
 *                                (main.cpp:467-484)
 *   rkdg_limit                   Limiter::update(Vector&) (src/limiter/limiter.hpp:111,
 *                                limiter.cpp:22-94)
 *   rkdg_step                    ExplicitRKLimitedSolver::Step(Vector&, double&, double&)
 *                                (src/ode_solver/rk_explicit_limited.hpp:24-60,
 *                                 rk_explicit_limited.cpp:30-82)
 *   rkdg_hmin                    minimum GetElementSize(i,1) (main.cpp:191-201)
 *   rkdg_means                   Averager::computeMeanValues (src/averager/averager.cpp:84-109)
 *
 * Plug-in families (RiemannSolver, BoundaryIntegrator::computeRightState,
 * Indicator, Limiter, EoS globals) are selected by the enums of
 * rkdg_scheme_desc; see src/riemannsolvers/rs_basic.hpp:18-51,
 * src/bc/boundary_integrator.hpp:23-64, src/indicator/indicator.hpp:19-84,
 * src/limiter/limiter.hpp:24-116, src/physics/physics.hpp:17-60.
 *
 * State layout is the reference's byNODES layout of the vector space:
 * x[v*NE*nd + e*nd + i] for conserved variable v, element e, local node i
 * (tensor elements: lexicographic nodes, x fastest).
 *
 * Pointers named x, y, k are DEVICE pointers (caller-owned). Calls on one
 * context are ordered on the context's stream and are not re-entrant.
 * Host out-parameters (lambda_max, hmin) synchronize the stream.
 * Return value: RKDG_OK or a negative error code; rkdg_last_error() gives text.
 */
#ifndef RKDG_CAPI_H
#define RKDG_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RKDG_OK 0
#define RKDG_E_ARG (-1)
#define RKDG_E_UNPHYSICAL (-2) /* a Riemann solver saw an unphysical state (reference: exit(1)) */
#define RKDG_E_HIP (-3)
#define RKDG_E_UNSUPPORTED (-4)
/* a Riemann solver's input held a NaN or an infinity (reference: StateIsPhysicalSay
 * fails and the integrator calls exit(1), face_integrator.cpp:214-218) */
#define RKDG_E_NAN (-5)

#define RKDG_GEOM_QUAD 0
#define RKDG_GEOM_TRI 1
#define RKDG_GEOM_HEX 2

#define RKDG_RS_RUSANOV 0
#define RKDG_RS_LLF 1
#define RKDG_RS_HLL 2
#define RKDG_RS_HLLC 3

#define RKDG_WS_TORO 0
#define RKDG_WS_EINFELDT 1

#define RKDG_BC_SLIP 0
#define RKDG_BC_OUTLET 1
#define RKDG_BC_SUPERSONIC_INLET 2
#define RKDG_BC_SUBSONIC_INLET_TOTAL_PRESSURE 3
#define RKDG_BC_SUBSONIC_INLET_FIXED_PRESSURE 4
#define RKDG_BC_FIXED_PRESSURE_OUTLET 5
#define RKDG_BC_TOTAL_PRESSURE_OUTLET 6
/* charOutlet (boundary_integrator_char_outlet.cpp:14-235): characteristic ghost state,
 * physical flux of the ghost (no Riemann solver); params (rhoI, UI[0], UI[1], UI[2], pI) */
#define RKDG_BC_CHAR_OUTLET 7

#define RKDG_IND_NOWHERE 0
#define RKDG_IND_EVERYWHERE 1
#define RKDG_IND_BJ 2
/* Shu's indicator on neighbour-extrapolated means (indicator_shu.cpp:13-74,
 * averager.cpp:152-273); it sets no multiplier, so it pairs with FinDiff only */
#define RKDG_IND_SHU 3

#define RKDG_PHI_MICHALAK 0
#define RKDG_PHI_BJ 1
#define RKDG_PHI_VENKAT 2

#define RKDG_LIM_NONE 0
#define RKDG_LIM_FINDIFF 1
#define RKDG_LIM_MULTI 2

#define RKDG_MAX_BC 16

typedef struct rkdg_ctx rkdg_ctx;

/* Mesh after any refinement, flattened by the host (the analogue of an MFEM
 * mesh with discontinuous nodes). All pointers are HOST pointers, copied at
 * create time. The element map is the vertices' multilinear map: affine
 * triangles; quadrilaterals (parallelograms on the tensor kernels, bilinear
 * ones on the dense kernels); hexahedra, parallelepipeds or trilinear (curved:
 * per-element records of adj(J) at the nodes, the inverse mass matrix and
 * per-point face normals, orders 1-3). */
typedef struct {
  int geom;                  /* RKDG_GEOM_* */
  int order;                 /* polynomial order p (DG_FECollection(p, dim)) */
  int64_t num_elements;
  const double* elem_coords; /* [NE][nv][dim] vertex coordinates, MFEM vertex order */
  const int64_t* elem_verts; /* [NE][nv] topological vertex ids (identified for periodic) */
  int64_t num_bdr;
  const int64_t* bdr_verts;  /* [num_bdr][nvf] */
  const int32_t* bdr_attr;   /* [num_bdr] 1-based attribute */
} rkdg_mesh_desc;

/* Scheme: RiemannSolver, wave speeds, EoS, boundary integrators, indicator, limiter. */
typedef struct {
  int riemann;      /* RKDG_RS_* (case_manager.cpp:523-543) */
  int wave_speeds;  /* RKDG_WS_* (physics.cpp:269-350) */
  int hllc_abs;     /* 1: HLLC reports max(|SL|,|SR|); 0: signed max(SL,SR) (rs_hllc.cpp:185) */
  double gamma, gas_constant, covolume;  /* specific_heat_ratio, gas_constant, covolume_constant */
  int num_bc;
  int bc_attr[RKDG_MAX_BC];
  int bc_type[RKDG_MAX_BC];
  double bc_param[RKDG_MAX_BC][8];
  int indicator;    /* RKDG_IND_* (case_manager.cpp:574-594) */
  int phi;          /* RKDG_PHI_* limiter function of the BJ-family indicator */
  double venkat_k;  /* Venkatakrishnan K (eps^2 = (K h)^3); 0 gives the K-free form */
  int limiter;      /* RKDG_LIM_* (case_manager.cpp:596-609) */
  int last_hope;    /* haveLastHope */
  int pow_nd;       /* 1: multiplier applied nd times (limiter_multiplier.cpp:24-38) */
  int linearize;    /* Multi: replace the element by its linear reconstruction first
                       (limiter_multiplier.cpp:101-178); p = 1 quadrilaterals and hexahedra */
} rkdg_scheme_desc;

typedef struct {
  int dim, num_equations, nodes_per_element, faces_per_element;
  int64_t num_elements, num_dofs;      /* num_dofs = NE*nd (scalar space) */
  int64_t state_size;                  /* num_equations*num_dofs */
  int64_t num_interior_faces, num_boundary_faces;
  int affine;                          /* 1 if every element is affine */
  int two_colour;                      /* 1 if the right-hand side runs as a solver/receiver pair of kernels */
} rkdg_info;

int rkdg_create(const rkdg_mesh_desc* mesh, const rkdg_scheme_desc* scheme, rkdg_ctx** out);
int rkdg_destroy(rkdg_ctx* ctx);
const char* rkdg_last_error(void);
int rkdg_get_info(rkdg_ctx* ctx, rkdg_info* info);
/* hipStream_t as void*; NULL selects the null stream. */
int rkdg_set_stream(rkdg_ctx* ctx, void* stream);

/* y = M^-1 [ face terms + volume terms ](x).  lambda_max: HOST pointer or NULL. */
int rkdg_rhs(rkdg_ctx* ctx, const double* x, double* y, double* lambda_max);
/* Max wave speed over face quadrature points only (first-dt path). */
int rkdg_face_lambda(rkdg_ctx* ctx, const double* x, double* lambda_max);
/* In-place limiter. troubled: DEVICE int8[NE] or NULL; bit0 = indicator < 1, bit1 = last hope. */
int rkdg_limit(rkdg_ctx* ctx, double* x, int8_t* troubled);
/* One RK step, Butcher tableau consumed with the packed a[l++] rule. s <= 4. */
int rkdg_step(rkdg_ctx* ctx, double* x, double t, double dt, int s, const double* a,
              const double* b, const double* c, double* lambda_max);
/* SSP-RK3 in the Shu-Osher form with the limiter applied to each stage value
 * (the classical limited RKDG step): u1 = L(x + dt R(x)),
 * u2 = L(3/4 x + 1/4 (u1 + dt R(u1))), x = L(1/3 x + 2/3 (u2 + dt R(u2))), with R
 * the right-hand side and L the limiter; 8 vector passes per step. No
 * counterpart in the reference, whose ExplicitRKLimitedSolver consumes Butcher
 * tableaus: with the limiter inactive this equals rkdg_step with the SSP-RK3
 * tableau to rounding; with it active the forms differ (the Butcher form limits
 * stage inputs it does not reuse). lambda_max: HOST pointer or NULL. */
int rkdg_step_ssprk3(rkdg_ctx* ctx, double* x, double t, double dt, double* lambda_max);
/* rkdg_step that allocates nothing. No counterpart in the reference: it is
 * for callers that cannot allocate device or pinned memory during a step (a
 * loop under HIP graph capture, a tight serving loop). The stage vectors live
 * in a static per-device scratch array of 128 MiB (2^24 doubles), so the
 * state may hold at most 2^22 doubles; larger contexts get RKDG_E_ARG. The
 * last stage's wave speed is copied to lambda_dev, a DEVICE double (nullable),
 * and there is no host synchronisation: deferred errors are read with
 * rkdg_check_error. Calls that overlap on one GPU share the array and race. */
int rkdg_step_noalloc(rkdg_ctx* ctx, double* x, double t, double dt, int s, const double* a,
                      const double* b, const double* c, double* lambda_dev);
/* Element means, layout [v*NE + e] (DEVICE pointer). */
int rkdg_means(rkdg_ctx* ctx, const double* x, double* means);
int rkdg_hmin(rkdg_ctx* ctx, double* hmin);
/* Post-step diagnostics. rkdg_totals: HOST totals[num_equations], the integral
 * over the mesh of each conserved variable (check_total_energy.cpp:3-63 forms
 * the energy's with an order-(p+1) rule at the nodes, exact for Q1 only; here
 * sum_e |e| mean_e, exact for every order). rkdg_primitives: the nodal fields
 * rho, u (dim components), p, T of main.cpp:388-451 into DEVICE
 * prim[(dim+3)][num_dofs] (nullable) and acc += weight * the same (nullable;
 * the time integrals UMean, pMean, TMean, rhoMean of main.cpp:705-730 with
 * weight dt). rkdg_interpolate: the state at npts reference points of every
 * element (HOST ref_pts[npts][dim]), DEVICE out[num_equations][NE][npts]
 * (ParaView output at the vertices, probes). */
int rkdg_totals(rkdg_ctx* ctx, const double* x, double* totals);
int rkdg_primitives(rkdg_ctx* ctx, const double* x, double* prim, double* acc, double weight);
int rkdg_interpolate(rkdg_ctx* ctx, const double* x, int npts, const double* ref_pts, double* out);
/* Host-pointer variants of rkdg_rhs, rkdg_limit and rkdg_step (x, y and
 * troubled are HOST arrays): the call copies the state to device buffers of its
 * own, runs the operator, copies the result back and synchronizes, reporting
 * a deferred device error as rkdg_check_error does. For tests and small
 * callers; they allocate per call. */
int rkdg_rhs_host(rkdg_ctx* ctx, const double* x, double* y, double* lambda_max);
int rkdg_limit_host(rkdg_ctx* ctx, double* x, int8_t* troubled);
int rkdg_step_host(rkdg_ctx* ctx, double* x, double t, double dt, int s, const double* a,
                   const double* b, const double* c, double* lambda_max);
/* The physics plug-ins on the host, the same code the kernels run.
 * rkdg_riemann_eval: RiemannSolver::Eval (rs_basic.hpp:49-50) -- the numerical
 * flux per unit normal through nor, which is normalized here (the reference's
 * face integrator normalizes CalcOrtho's normal before Eval,
 * face_integrator.cpp:152-162); max_speed = the solver's wave speed or -1 for
 * an unphysical input (then RKDG_E_UNPHYSICAL). rkdg_bc_right_state:
 * BoundaryIntegrator::computeRightState (boundary_integrator.hpp:63) of the
 * scheme's boundary slot, nor the outward normal (normalized here). States hold
 * dim+2 conserved variables. */
int rkdg_riemann_eval(const rkdg_scheme_desc* scheme, int dim, const double* state1, const double* state2,
                      const double* nor, double* flux, double* max_speed);
int rkdg_bc_right_state(const rkdg_scheme_desc* scheme, int dim, int slot, const double* state1,
                        const double* nor, double* state2);
/* Physical coordinates of the nodes, HOST [NE][nd][dim] (for initial conditions). */
int rkdg_node_coords(rkdg_ctx* ctx, double* xyz);
/* Synchronize and report a deferred device error (RKDG_E_UNPHYSICAL, or
 * RKDG_E_NAN when an input of the failing flux was not finite, with the
 * element pair in info[0], info[1]); clears it. */
int rkdg_check_error(rkdg_ctx* ctx, int64_t* info);
/* The element pair of the last device error reported by any entry point of
 * this context ({-1, -1} before the first): an entry point that returns
 * RKDG_E_UNPHYSICAL or RKDG_E_NAN itself has already cleared the error, and
 * this is where its pair stays (the reference prints it,
 * face_integrator.cpp:214-218). */
int rkdg_error_elements(rkdg_ctx* ctx, int64_t* pair);

/* Instrumentation, no counterpart in the reference (which only has a wall-clock
 * StopWatch per step, main.cpp:458-459): when enabled, HIP events on the
 * context's stream bracket every right-hand-side kernel (plain or RK stage) and
 * every limiter kernel launch. Enabling resets the counters; reading
 * synchronizes and returns the summed kernel milliseconds and launch counts. */
int rkdg_profile_enable(rkdg_ctx* ctx, int enable);
int rkdg_profile_read(rkdg_ctx* ctx, double* rhs_ms, int64_t* rhs_launches, double* limit_ms,
                      int64_t* limit_launches);

/* rkdg_create with flags. RKDG_CREATE_DENSE runs quadrilateral meshes on the
 * dense bilinear-element kernels even where every element is a parallelogram
 * (the kernels a batch uses, so a single instance can be compared with its
 * batch bit for bit). Same counterpart as rkdg_create. */
#define RKDG_CREATE_DENSE 1
/* RKDG_CREATE_TWO_COLOUR: on tensor meshes whose face graph is two-colourable
 * and has no boundary faces (affine periodic meshes of even extents), run the
 * right-hand side as a solver/receiver pair of kernels that solves each face
 * once; the receivers' share of the face fluxes goes through a buffer of the
 * size of one state vector. Other meshes keep the one-kernel form. */
#define RKDG_CREATE_TWO_COLOUR 2
int rkdg_create_ex(const rkdg_mesh_desc* mesh, const rkdg_scheme_desc* scheme, int flags, rkdg_ctx** out);

/* Non-conforming faces of an adaptively refined mesh: MFEM NCMesh's slave
 * faces, which the reference's face loop integrates as interior faces with
 * Elem1 = the fine (slave) element and Elem2 = the coarser (master) element
 * (face_integrator.cpp:11-276), and which its limiter puts in the stencil
 * (limiter.cpp:166-217, the infoNC branch). Record k: face slave_face[k]
 * (MFEM local numbering) of element slave[k] covers part of face
 * master_face[k] of element master[k]; param[(k*nvf + c)*(dim-1) + j] is
 * coordinate j of the slave face's corner c (the slave's local face order) in
 * the master face's parametrisation (its local face vertex 0 at the origin,
 * vertex 1 at 1 (2D) or (1,0), vertex 3 at (0,1) (3D)): the slave's point
 * matrix. The element kernels take affine quadrilaterals and affine or
 * trilinear hexahedra. */
typedef struct {
  int64_t count;
  const int64_t* slave;
  const int32_t* slave_face;
  const int64_t* master;
  const int32_t* master_face;
  const double* param;
} rkdg_nc_desc;
/* rkdg_create_ex on a mesh with non-conforming faces (after each refinement:
 * the reference's UpdateAndRebalance + FE_Evolution updates, amr.cpp:3-133,
 * main.cpp:280-311). The BJ-family limiters take the non-conforming
 * neighbours into their stencil; the Shu indicator is not built for them. */
int rkdg_create_nc(const rkdg_mesh_desc* mesh, const rkdg_scheme_desc* scheme, int flags, const rkdg_nc_desc* nc,
                   rkdg_ctx** out);

/* Batched ensembles. No counterpart in the reference, which runs one case per
 * process through main.cpp's time loop (main.cpp:467-553, 701, 733); a batch
 * runs that loop for `count` independent instances at once. Instances share
 * element type, order and scheme (only bc_param may differ per instance); the
 * context holds their union mesh (instance i owns elements
 * [offsets[i], offsets[i+1]) and the matching slices of every variable of the
 * byNODES state), so rkdg_rhs, rkdg_limit, rkdg_means and rkdg_node_coords
 * work on the whole ensemble. Quadrilaterals and triangles only. */
int rkdg_batch_create(const rkdg_mesh_desc* meshes, const rkdg_scheme_desc* schemes, int count,
                      rkdg_ctx** out);
/* [count+1] element offsets (HOST). */
int rkdg_batch_offsets(rkdg_ctx* ctx, int64_t* elem_offsets);
/* Start every instance's time loop: limit x, first dt from the face wave
 * speeds (main.cpp:265, 467-484). t0, t_final: HOST [count]. max_steps > 0
 * also stops an instance after that many steps. */
int rkdg_batch_init(rkdg_ctx* ctx, double* x, const double* t0, const double* t_final, double cfl,
                    int64_t max_steps);
/* Up to nsteps steps of every active instance, each with its own dt
 * (dt_real = min(dt, t_final - t); next dt = cfl*hmin/lambda/(2p+1); t += dt;
 * stop at t >= t_final - 1e-8 dt), entirely on the device: a captured HIP graph
 * per step. n_active (HOST, nullable): instances still running. 2 <= s <= 4. */
int rkdg_batch_advance(rkdg_ctx* ctx, double* x, int s, const double* a, const double* b, const double* c,
                       int64_t nsteps, int* n_active);
/* HOST [count] copies of t, next dt, steps taken and the active flags (each nullable). */
int rkdg_batch_status(rkdg_ctx* ctx, double* t, double* dt, int64_t* steps, int* active);
/* use_graph 0/1 (batches); grid_cap: persistent grid of the dense
 * right-hand-side kernel in workgroups (0: one per work block; -1: keep). Any
 * context accepts it; the tensor kernels ignore grid_cap. */
int rkdg_batch_set_options(rkdg_ctx* ctx, int use_graph, int grid_cap);
/* Autotuner (any context): times the right-hand side of x (unchanged) for
 * grid_cap 0 and each candidate, keeps the fastest. ms: [ncand+1] per-launch
 * milliseconds (nullable). */
int rkdg_tune_grid(rkdg_ctx* ctx, const double* x, const int* candidates, int ncand, int reps, int* best,
                   double* ms);

/* Discrete adjoint (dense 2D elements: triangles, and quadrilaterals on the
 * dense kernels; boundary types slip, outlet and supersonicInlet). No
 * counterpart in the reference, which has no adjoint: these are the backward
 * operators of rkdg_rhs, rkdg_limit and rkdg_step, with the forward branch
 * taken at every min/max and threshold switch. DEVICE pointers unless noted. */
/* Replace the 8 parameters of boundary slot `slot` (HOST array). */
int rkdg_set_bc_params(rkdg_ctx* ctx, int slot, const double* params);
/* xbar = addend + (d rkdg_rhs / dx)^T kbar; bc_param_bar [RKDG_MAX_BC*8]
 * accumulates the adjoint of the boundary parameters (both nullable). */
int rkdg_rhs_adjoint(rkdg_ctx* ctx, const double* x, const double* kbar, double* xbar, const double* addend,
                     double* bc_param_bar);
/* xbar = (d rkdg_limit / dx)^T ubar at the pre-limit state x. */
int rkdg_limit_adjoint(rkdg_ctx* ctx, const double* x, const double* ubar, double* xbar);
/* *J += coef * sum over the faces of boundary slot bc_slot of the integral of
 * the pressure of the interior trace; xbar += coef * its gradient (J, xbar
 * nullable, both DEVICE). */
int rkdg_objective(rkdg_ctx* ctx, const double* x, int bc_slot, double coef, double* J, double* xbar);
/* xbar = (d rkdg_step(x, dt) / dx)^T ybar; bc_param_bar: HOST [RKDG_MAX_BC*8] (nullable). */
int rkdg_step_adjoint(rkdg_ctx* ctx, const double* x, double dt, int s, const double* a, const double* b,
                      const double* c, const double* ybar, double* xbar, double* bc_param_bar);
/* nsteps steps with the HOST dt schedule dts; J (HOST, nullable) =
 * sum_n dts[n] * P(x_{n+1}) over the faces of boundary slot obj_slot (< 0: none). */
int rkdg_forward_solve(rkdg_ctx* ctx, double* x, int64_t nsteps, const double* dts, int s, const double* a,
                       const double* b, const double* c, int obj_slot, double* J);
/* Gradient of that J with respect to the initial state x0 (x0bar) and the
 * boundary parameters (bc_param_bar, HOST [RKDG_MAX_BC*8]) for the fixed dt
 * schedule, by sqrt(N) checkpointing: checkpoints every `every` steps (0:
 * ceil(sqrt(nsteps))), each segment recomputed once with its stage states
 * recorded. J: HOST; xN (nullable): final state. */
int rkdg_adjoint_solve(rkdg_ctx* ctx, const double* x0, int64_t nsteps, const double* dts, int s,
                       const double* a, const double* b, const double* c, int obj_slot, int64_t every,
                       double* J, double* x0bar, double* bc_param_bar, double* xN);

#ifdef __cplusplus
}
#endif
#endif
