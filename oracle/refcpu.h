/*
 * refcpu — CPU restatement of vkorchagova/RKDG3D's explicit RKDG right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity oracle: only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may load it, and only as a
 * checker or a timed CPU baseline. The product path (rkdg_amd/) never links it.
 *
 * It restates, with the reference's own data structures (CSR Aflux volume
 * matrix, dense per-element inverse mass, face loop in face-id order with shape
 * functions evaluated per quadrature point), these reference files:
 *   src/fe_evolution/fe_evolution.cpp       FE_Evolution::Mult, GetFlux, Me_inv
 *   src/eq_terms/domain_integrator.cpp      Aflux element matrix
 *   src/eq_terms/face_integrator.cpp        interior-face numerical flux
 *   src/bc/boundary_integrator*.cpp         boundary faces and ghost states
 *   src/physics/physics.cpp                 EoS, fluxes, wave speeds
 *   src/riemannsolvers/rs_*.cpp             Rusanov, LLF, HLL, HLLC
 *   src/limiter/limiter*.cpp, src/indicator/indicator*  limiter driver, multiplier (with linearize),
 *                                           FinDiff, BJ-family, Shu
 *   src/averager/averager.cpp:84-119        cell means
 *   src/averager/averager.cpp:152-273       neighbour-extrapolated means (Shu)
 *   src/ode_solver/rk_explicit_limited.cpp  Butcher RK with limiting
 * and the MFEM behaviour those files rely on (Gauss-Legendre nodal bases,
 * IntRules point counts, face Loc1/Loc2 maps, CalcOrtho normals, MassIntegrator,
 * GetElementAverages, GetElementSize).
 *
 * The same source builds a second library with REF_REAL = long double
 * (librefcpu_ld.so): the operator evaluated with a 64-bit significand, against
 * which the rounding error of the double oracle and of the GPU path are
 * measured on ill-conditioned workloads (a fluid at rest, whose RHS is a small
 * difference of large flux terms).
 *
 * The reference cannot be built here (MFEM, hypre, METIS and rapidyaml are
 * absent) and holds no test vectors, so the restatement is pinned by closed
 * forms, exact solutions, invariants and symmetries (tests/test_oracle_pins.py).
 * Parity unpinned: the MFEM behaviour the reference relies on and that only
 * MFEM itself could confirm (restated from its published algorithms):
 *   - the node positions of L2_TriangleElement for p >= 2 (only the Lagrange
 *     property at its own nodes is checked);
 *   - the face and volume quadrature point counts derived from OrderW;
 *   - Vector::Norml2 rounding beyond vectors with one non-zero entry;
 *   - MFEM's face numbering and Elem1 choice (results depend on them only
 *     through rounding, which the orientation-invariance pin bounds);
 *   - ElementTransformation::TransformBack for points outside the element
 *     (the Shu indicator's extrapolation): restated as the exact inverse map;
 *     MFEM's inverse-transformation solver may project such points onto the
 *     element instead, which only MFEM itself could settle;
 *   - the order of the points of IntRules.Get(geom, p + 1) that linearize
 *     pairs with the nodes (restated as x fastest, as the tensor rules are).
 * Quadrilaterals are mapped bilinearly everywhere, their Jacobian written on
 * edge vectors: the same map as MFEM's PointMat * dshape, without the
 * rounding that form leaves on elements small against their coordinates.
 */
#ifndef RKDG_REFCPU_H
#define RKDG_REFCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* geometry codes (shared meaning with the product C-ABI) */
#define REF_GEOM_QUAD 0
#define REF_GEOM_TRI 1
#define REF_GEOM_HEX 2

#define REF_RS_RUSANOV 0
#define REF_RS_LLF 1
#define REF_RS_HLL 2
#define REF_RS_HLLC 3

#define REF_WS_TORO 0
#define REF_WS_EINFELDT 1

#define REF_BC_SLIP 0
#define REF_BC_OUTLET 1
#define REF_BC_SUPERSONIC_INLET 2
#define REF_BC_SUBSONIC_INLET_TOTAL_PRESSURE 3
#define REF_BC_SUBSONIC_INLET_FIXED_PRESSURE 4
#define REF_BC_FIXED_PRESSURE_OUTLET 5
#define REF_BC_TOTAL_PRESSURE_OUTLET 6
#define REF_BC_CHAR_OUTLET 7

#define REF_IND_NOWHERE 0
#define REF_IND_EVERYWHERE 1
#define REF_IND_BJ 2
#define REF_IND_SHU 3

#define REF_PHI_MICHALAK 0
#define REF_PHI_BJ 1
#define REF_PHI_VENKAT 2

#define REF_LIM_NONE 0
#define REF_LIM_FINDIFF 1
#define REF_LIM_MULTI 2

#define REF_MAX_BC 16

typedef struct {
  int riemann;      /* REF_RS_* */
  int wave_speeds;  /* REF_WS_* */
  int hllc_abs;     /* 1: HLLC returns max(|SL|,|SR|); 0: signed max(SL,SR) as in rs_hllc.cpp:185 */
  double gamma, gas_constant, covolume;
  int num_bc;
  int bc_attr[REF_MAX_BC]; /* boundary attribute (1-based) */
  int bc_type[REF_MAX_BC];
  double bc_param[REF_MAX_BC][8];
  int indicator;    /* REF_IND_* */
  int phi;          /* REF_PHI_* */
  double venkat_k;  /* >0: eps^2=(K h)^3 added to the Venkatakrishnan form */
  int limiter;      /* REF_LIM_* */
  int last_hope;
  int pow_nd;       /* 1: multiplier applied nd times (limiter_multiplier.cpp:24-38) */
  int linearize;    /* multiplier: linear reconstruction first (limiter_multiplier.cpp:101-178; p = 1) */
} ref_scheme;

void* ref_create(int geom, int order, int64_t num_elements, const double* elem_coords,
                 const int64_t* elem_verts, int64_t num_bdr, const int64_t* bdr_verts,
                 const int* bdr_attr, const ref_scheme* scheme, char* err, int errlen);
/* ref_create with non-conforming faces (MFEM NCMesh slaves): face nc_slave_face[k]
 * of element nc_slave[k] covers part of face nc_master_face[k] of the coarser
 * element nc_master[k]; nc_param[(k*nvf + c)*(dim-1) + j] is coordinate j of the
 * slave face's corner c in the master face's parametrisation (the point matrix). */
void* ref_create_nc(int geom, int order, int64_t num_elements, const double* elem_coords,
                    const int64_t* elem_verts, int64_t num_bdr, const int64_t* bdr_verts,
                    const int* bdr_attr, const ref_scheme* scheme, int64_t num_nc,
                    const int64_t* nc_slave, const int* nc_slave_face, const int64_t* nc_master,
                    const int* nc_master_face, const double* nc_param, char* err, int errlen);
void ref_destroy(void* h);
int64_t ref_num_dofs(void* h);
int ref_nodes_per_element(void* h);
int64_t ref_num_interior_faces(void* h);
void ref_node_coords(void* h, double* xyz);
/* FE_Evolution::Mult. Returns 0, or -1 if a Riemann solver reported an unphysical state
 * (the reference exits); bad[0], bad[1] receive the element pair. */
int ref_mult(void* h, const double* x, double* y, double* max_char_speed, int64_t* bad);
/* A.Mult(x, z): face and boundary terms only (main.cpp:473-484 first-dt path) */
int ref_face_mult(void* h, const double* x, double* z, double* max_char_speed, int64_t* bad);
void ref_means(void* h, const double* x, double* means);
void ref_limiter_update(void* h, double* x, double* ind_values, double* min_values);
int ref_step(void* h, double* x, double t, double dt, int s, const double* a, const double* b,
             const double* c, double* max_char_speed, int64_t* bad);
double ref_hmin(void* h);
void ref_element_mass_inverse(void* h, int64_t e, double* out);
void ref_set_threads(int n);
/* single-point plug-ins for unit tests */
double ref_riemann_eval(const ref_scheme* s, int dim, const double* s1, const double* s2,
                        const double* nor, double* flux);
void ref_bc_state(const ref_scheme* s, int dim, int bc_index, const double* s1, const double* nor,
                  double* s2);
void ref_gl_rule(int n, double* x, double* w);
void ref_shape(int geom, int order, const double* xi, double* phi, double* dphi);

#ifdef __cplusplus
}
#endif
#endif
