/*
 * iceicle_amd.h -- C-ABI of the MI355X DG residual/Jacobian engine.
 *
 * The reference (ICEicle) has no FFI: its boundary is the duck-typed C++
 * template `form_residual(FESpace&, disc&, fespan u, fespan res)`
 * (src/solvers/include/iceicle/form_residual.hpp:56-268) together with the
 * ConservationLawDDG discretization (src/FiniteElement/include/iceicle/disc/
 * conservation_law.hpp:92-1289).  Every entry point below replaces one piece
 * of that C++ interface; each declaration cites the piece it replaces.
 *
 * Conventions
 *  - Plain C types only: pointers, sizes, ints, doubles.  Device pointers are
 *    caller-owned HIP device allocations; `stream` is a hipStream_t passed as
 *    void* (NULL = the legacy default stream).
 *  - Status: 0 (ICE_OK) on success, a negative ICE_ERR_* class otherwise.
 *    ice_last_error() returns the message of the last failure on the calling
 *    thread.  Handles are not thread-safe; distinct handles are independent.
 *  - Solution layout: the reference `fe_layout_right` order, global index
 *    (e * nbasis + i) * neq + v (layout_right.hpp:200-214, dglayout.hpp:113).
 *  - Basis/quadrature/mesh ordering follow the reference exactly: tensor
 *    indices with dimension 0 slowest (tensor_product.hpp:321-380,
 *    HypercubeGaussLegendre.hpp:17-67), nodes and elements x-fastest and the
 *    face order of the uniform AbstractMesh constructor (mesh.hpp:553-913).
 */
#ifndef ICEICLE_AMD_H
#define ICEICLE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ICE_API_VERSION 1
#define ICE_MAX_DIM 3
#define ICE_MAX_ORDER 4
#define ICE_MAX_BCFLAG 4

/* ---- status codes ---------------------------------------------------- */
#define ICE_OK 0
#define ICE_ERR_ARG (-1)
#define ICE_ERR_UNSUPPORTED (-2)
#define ICE_ERR_HIP (-3)
#define ICE_ERR_ALLOC (-4)
#define ICE_ERR_STATE (-5)
#define ICE_ERR_NUMERIC (-6)

/* ---- boundary conditions: same codes as the reference enum
 *      BOUNDARY_CONDITIONS (geometry/face.hpp:23-103) ------------------- */
#define ICE_BC_PERIODIC 0
#define ICE_BC_PARALLEL_COM 1
#define ICE_BC_NEUMANN 2
#define ICE_BC_DIRICHLET 3
#define ICE_BC_EXTRAPOLATION 4
#define ICE_BC_RIEMANN 5
#define ICE_BC_NO_SLIP_ISOTHERMAL 6
#define ICE_BC_SLIP_WALL 7
#define ICE_BC_WALL_GENERAL 8
#define ICE_BC_INLET 9
#define ICE_BC_OUTLET 10
#define ICE_BC_SPACETIME_PAST 11
#define ICE_BC_SPACETIME_FUTURE 12
#define ICE_BC_INTERIOR 13

/* ---- periodic handling ----------------------------------------------
 * REFERENCE: periodic faces stay boundary faces whose bcflag is the partner
 *   face index (mesh.hpp:857-866); boundaryIntegral then takes its `default`
 *   branch (conservation_law.hpp:1043-1051), which adds no flux for Burgers.
 * COUPLED: each partner pair is replaced by one interior face (textbook
 *   periodic coupling); h_ddg uses each side's own centroid-to-face distance. */
#define ICE_PERIODIC_REFERENCE 0
#define ICE_PERIODIC_COUPLED 1

/* ---- flux families (disc/burgers.hpp, disc/navier_stokes.hpp) ---------- */
#define ICE_FLUX_BURGERS 0        /* BurgersFlux + BurgersUpwind + BurgersDiffusionFlux */
#define ICE_FLUX_NAVIER_STOKES 1  /* Flux<full_ns> + VanLeer + DiffusionFlux<full_ns> */
#define ICE_FLUX_EULER 2          /* Flux<euler> + VanLeer + DiffusionFlux<euler> (zero) */

#define ICE_VISC_CONSTANT 0               /* constant_viscosity (navier_stokes.hpp:231-242) */
#define ICE_VISC_DIMENSIONLESS_SUTHERLAND 1 /* DimensionlessSutherlands (:271-301) */

/* ---- analytic fields: Dirichlet/Neumann values, sources, initial states.
 * The reference takes std::function callbacks (conservation_law.hpp:152-170);
 * across a C-ABI they are named analytic fields with parameters. */
#define ICE_FIELD_NONE 0
#define ICE_FIELD_CONSTANT 1      /* out[v] = p[v] */
#define ICE_FIELD_SINE 2          /* out[0] = p[0] * sin(p[1] * x + p[2]) (1D heat pin) */
#define ICE_FIELD_SINUSOIDS 3     /* out[0] = sum_k A_k sin(2 pi (m_k x + n_k y + l_k z));
                                     p[0] = K, then (A, m, n, l) per k, K <= 7 */
#define ICE_FIELD_NS_MMS_STATE 4  /* manufactured 2D NS state (conservative vars) */
#define ICE_FIELD_NS_MMS_SOURCE 5 /* its source term, sign convention of user_source */
#define ICE_FIELD_ISENTROPIC_VORTEX 6 /* isentropic vortex state, p = (x0, y0, strength, rho_inf, u_inf, v_inf) */
#define ICE_FIELD_EXP_DECAY_SINE 7 /* out[0] = exp(-p[0] * p[1]) * sin(x): exact 1D heat solution at t = p[1] */

typedef struct ice_field {
    int kind;
    double p[32];
} ice_field;

/* ---- mesh: uniform hypercube box (AbstractMesh uniform constructor,
 *      mesh.hpp:553-913, geometry order 1), optional moved node coordinates. */
typedef struct ice_mesh_desc {
    int ndim;                 /* 1, 2 or 3 */
    int nelem[3];             /* elements per direction */
    double xmin[3];
    double xmax[3];
    int bctype[6];            /* per domain face, hypercube numbering: face k is
                                 normal to axis k % ndim, negative side if k < ndim */
    int bcflag[6];
    const double* coord;      /* optional host array [n_nodes][ndim] of node
                                 coordinates in the generator's node order
                                 (x fastest); NULL keeps the uniform nodes */
    int periodic_mode;        /* ICE_PERIODIC_* */
} ice_mesh_desc;

typedef struct ice_space_desc {
    int order;                /* basis order p, 1 .. ICE_MAX_ORDER */
    int geometry_order;       /* must be 1 (see mesh.hpp:259-264 vs :584) */
} ice_space_desc;

typedef struct ice_mesh_counts {
    int ndim;
    int nelem;
    int nnode;
    int nvert_el;             /* 2^ndim */
    int nfacenode;            /* 2^(ndim-1) */
    int nface;                /* total faces (interior first, then boundary) */
    int nface_interior;
} ice_mesh_counts;

/* Face record exported by ice_mesh_export, one per face in reference order:
 * elemL, elemR, face_info_l, face_info_r, bctype, bcflag, nodes[nfacenode]
 * with face_info = face_nr * 512 + orientation (face.hpp:207-209). */
#define ICE_FACE_REC_FIXED 6

typedef struct ice_space_info {
    int ndim;
    int order;
    int nbasis;               /* (p+1)^ndim */
    int nqp;                  /* volume quadrature points (p+1)^ndim */
    int nfqp;                 /* face quadrature points (p+1)^(ndim-1) */
    int nelem;
    int nface;
    int nface_interior;
    int nface_boundary;
    int64_t ndof;             /* nelem * nbasis (per component) */
} ice_space_info;

/* Discretization parameters: ConservationLawDDG public members
 * (conservation_law.hpp:132-170) plus the flux functor coefficients. */
/* DDGIC interface-correction forms.  SOURCE follows conservation_law.hpp
 * as written: sigma_ic/2 times each side's test gradient at the face point
 * on interior faces (:580-610) and sigma_ic on Dirichlet faces (:750-780).
 * PUBLISHED is the form that reproduces the DDGIC table of
 * doc/sine_diffusion_1d_study.md:45-55 (P1-P4) to the digits of its
 * standard-DDG table: sigma_ic on each side of interior faces, no boundary
 * term, and each side's test gradient averaged between the face point and
 * its mirror through the element along the face normal (in 1D the mean of
 * the derivative at the element's two ends). */
#define ICE_DDGIC_SOURCE 0
#define ICE_DDGIC_PUBLISHED 1

typedef struct ice_disc_params {
    int flux;                 /* ICE_FLUX_* */
    /* BurgersCoefficients (burgers.hpp:14-40) */
    double mu;
    double a[3];
    double b[3];
    /* Navier-Stokes physics (navier_stokes.hpp:321-770) */
    double gamma;             /* default 1.4 */
    double Rgas;              /* default 287.052874 */
    double Pr;                /* default 0.72 */
    double ref[8];            /* ReferenceParameters t,l,rho,u,e,p,T,mu (default all 1) */
    int viscosity_kind;       /* ICE_VISC_* */
    double visc_mu;           /* constant viscosity value */
    double isothermal_T[ICE_MAX_BCFLAG];
    double freestream[6];     /* rho_inf, u_inf, dir x,y,z, T_inf (RIEMANN bc) */
    int ns_face_energy_textbook; /* 0: face energy flux as navier_stokes.hpp:1277-1279
                                    (mu*gamma/Pr * grad E . n); 1: consistent with the
                                    volume heat flux (grad e) */
    /* ConservationLawDDG */
    double sigma_ic;          /* DDGIC multiplier */
    int interior_penalty;     /* beta1 = 0 */
    int ddgic_form;           /* ICE_DDGIC_*: which reading of the interface
                                 correction term sigma_ic scales */
    ice_field dirichlet[ICE_MAX_BCFLAG];
    ice_field neumann[ICE_MAX_BCFLAG];
    ice_field source;         /* user_source; kind NONE = no source */
} ice_disc_params;

/* Fill an ice_disc_params with the reference defaults for a flux family. */
int ice_disc_params_default(int flux, ice_disc_params* out);

/* ---- opaque handles -------------------------------------------------- */
typedef struct ice_space ice_space;
typedef struct ice_disc ice_disc;

/* Library version and last error message of the calling thread. */
int ice_version(void);
const char* ice_last_error(void);

/* Host-only mesh build (no device use).  Replaces the uniform AbstractMesh
 * constructor (mesh.hpp:553-913) as an inspection path: with coord/conn/faces
 * NULL it only fills `counts`; otherwise it writes node coordinates
 * [nnode][ndim], element connectivity [nelem][2^ndim] (element-local tensor
 * order, mesh.hpp:637-649) and face records [nface][6 + nfacenode]. */
int ice_mesh_export(const ice_mesh_desc* mesh, ice_mesh_counts* counts,
                    double* coord, int* conn, int* faces);

/* FESpace construction (fespace.hpp:200-416): mesh, reference tables and the
 * trace maps are built on the host and uploaded to the current HIP device. */
int ice_space_create(const ice_mesh_desc* mesh, const ice_space_desc* space,
                     ice_space** out);

/* Batched space: nmesh independent meshes (same ndim and periodic mode; with
 * coupled periodic faces, the same box extents) concatenated into one space,
 * the elements of mesh k following those of meshes 0..k-1 and no face
 * connecting two meshes.  One residual launch then evaluates every instance;
 * with ice_disc_set_element_viscosity each instance keeps its own viscosity.
 * No reference counterpart (the reference solves one mesh per process).
 * The multigrid and line preconditioners reject batched spaces. */
int ice_space_create_batch(const ice_mesh_desc* meshes, int nmesh, const ice_space_desc* space,
                           ice_space** out);
int ice_space_destroy(ice_space* space);
int ice_space_query(const ice_space* space, ice_space_info* info);

/* Physical coordinates of the volume quadrature points, host array
 * [nelem][nqp][ndim] (FiniteElement::transform at getQP, finite_element.hpp:406-408). */
int ice_space_qp_coords(const ice_space* space, double* xq);

/* Discretization (ConservationLawDDG constructor + public members). */
int ice_disc_create(ice_space* space, const ice_disc_params* params, ice_disc** out);
int ice_disc_destroy(ice_disc* disc);

/* Host-only query: the number of conserved components (nv_comp,
 * conservation_law.hpp:129-130) and the length in doubles of a state or
 * residual vector, nelem * nbasis * neq (fe_layout_right size,
 * layout_right.hpp:91-215); either pointer may be null. */
int ice_disc_query(const ice_disc* disc, int* neq, int64_t* size);

/* Replace the parameters of an existing discretization in place (flux
 * coefficients, boundary data, DDG options) and rebuild its boundary-value
 * and source tables, so that many problems that differ only in parameters
 * reuse one space, one discretization and one solver.  The flux family and
 * whether DDGIC and the source term are on must stay as at creation
 * (ICE_ERR_ARG otherwise).  No reference counterpart: the reference builds a
 * new ConservationLawDDG per problem (conservation_law.cpp:287-346). */
int ice_disc_update(ice_disc* disc, const ice_disc_params* params);

/* Burgers only: a viscosity per element (host array [nelem]) replacing
 * params.mu in the volume and face fluxes, so that the instances of a
 * batched space keep their own viscosity; null restores params.mu.  The
 * device Jacobian is unavailable while it is set.  No reference counterpart. */
int ice_disc_set_element_viscosity(ice_disc* disc, const double* mu);

/* Element-kernel workgroup shape of this discretization's residual (its p,
 * number of variables and dimension are fixed): tile 0 = whole elements in
 * 256 threads (default), 1 = 128, 2 = 64; results are identical, only speed
 * changes.  Tile 3 selects, for the mixed-precision residual of 3D P3 Euler,
 * the element kernel whose tensor contractions run on the matrix cores
 * (v_mfma_f32_4x4x1f32; same FP32 arithmetic, its own summation order), and
 * the default shape elsewhere.  ice_disc_set_tile sets the shape of both the
 * FP64 and the mixed-precision residual.  ice_disc_autotune times `reps` FP64
 * residuals at u for each tile 0-2 on the stream, keeps the fastest (*best)
 * and writes the per-residual milliseconds to ms[3] (either may be null);
 * then, in 2D and 3D, times the mixed-precision residual the same way (tile 3
 * included where it applies) and keeps its fastest tile for that residual.
 * r is overwritten.  Synchronous.  No reference counterpart. */
int ice_disc_set_tile(ice_disc* disc, int tile);
int ice_disc_autotune(ice_disc* disc, const double* u_dev, double* r_dev, int reps, int* best, double* ms,
                      void* stream);

/* form_residual (form_residual.hpp:56-268): res <- R(u), FP64, device pointers
 * in fe_layout_right order, both nelem*nbasis*neq doubles. */
int ice_residual(ice_disc* disc, const double* u_dev, double* res_dev, void* stream);

/* Mixed-precision residual (no reference counterpart; the reference is FP64
 * throughout): the same R(u) with the trace interpolation, the physical and
 * numerical fluxes, the volume contractions and the face data in FP32, and
 * the geometric Jacobians, source load, face lifting sums and R in FP64.
 * ndim = 2 and 3 (ICE_ERR_UNSUPPORTED for ndim = 1).  Expect ~1e-7 of the
 * largest flux terms in absolute error. */
int ice_residual_mixed(ice_disc* disc, const double* u_dev, double* r_dev, void* stream);

/* Allocation-free FP64 residual.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same result as
 * ice_residual (same kernels, same summation order), but the per-call
 * buffers (face traces, face fluxes, DDGIC lifting vectors, wave-speed
 * partials) live in one fixed array of ICE_NOALLOC_SCRATCH_MB MiB of device
 * memory per GPU, held until the process ends; the call itself allocates
 * nothing.  A discretization needing more scratch than that (e.g. 2D NS P3
 * beyond ~190 x 190 elements) gets ICE_ERR_ARG: use ice_residual.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams or threads) race; calls ordered on one stream do not. */
#define ICE_NOALLOC_SCRATCH_MB 64
int ice_residual_noalloc(ice_disc* disc, const double* u_dev, double* res_dev, void* stream);
/* Scratch bytes ice_residual_noalloc needs for this discretization and the
 * fixed maximum (either pointer may be null; a null disc reports the maximum). */
int ice_noalloc_scratch_bytes(const ice_disc* disc, int64_t* need_bytes, int64_t* max_bytes);

/* Inverse mass matrix per element and component: out = M_e^-1 r.  Replaces
 * calculate_mass_matrix + decompose_lu + sub_lu of the explicit steppers
 * (finite_element.hpp:586-611, ssp_rk3.hpp:142-163); singular element mass
 * matrices fall back to 1e-8 r as tvd_rk3.hpp:190-204. */
int ice_mass_inverse_apply(ice_disc* disc, const double* r_dev, double* out_dev, void* stream);

/* L2 projection of an analytic field onto the space (Projection::domain_integral
 * + ElementLinearSolver, projection.hpp:83-110, element_linear_solve.hpp:37-90,
 * as used by projection_initialization, initialization.hpp:20-30).  Synchronous. */
int ice_project(ice_disc* disc, const ice_field* f, double* u_dev, void* stream);

/* Callback data given as values at points (the reference's std::function
 * callbacks: initial conditions, initialization.hpp:37-104; Dirichlet and
 * Neumann callbacks and the user source, conservation_law.hpp:700-730,
 * 300-312; set from Lua functions by solvers_lua_interface.hpp).  A host
 * caller evaluates its function at the points these report and hands the
 * values over on the device:
 *  - ice_project_values: L2 projection of values vq_dev[nelem][nqp][neq] at
 *    the volume quadrature points (ice_space_qp_coords order); synchronous;
 *  - ice_disc_set_source_values: the source term from the same kind of
 *    values (the load is formed once, with max(0, detJ) as the volume
 *    integral's source term); synchronous;
 *  - ice_space_bface_coords: per device boundary face b (ice_space_info's
 *    nface_boundary faces, in device order) its boundary type and flag, and the physical coordinates xb[b][q][ndim] of its quadrature
 *    points (any of the three outputs may be NULL);
 *  - ice_disc_set_boundary_values: bv_dev[b][q][neq], the Dirichlet state
 *    on Dirichlet faces and the Neumann data on Neumann faces (other faces
 *    ignore it); synchronous.
 * Values set this way replace the analytic fields' and are kept when the
 * geometry or the parameters are refreshed (ice_disc_update,
 * ice_mdg_form_residual).  No reference counterpart as entry points. */
int ice_project_values(ice_disc* disc, const double* vq_dev, double* u_dev, void* stream);
int ice_disc_set_source_values(ice_disc* disc, const double* vq_dev, void* stream);
int ice_space_bface_coords(const ice_space* space, double* xb, int* bctype, int* bcflag);
int ice_disc_set_boundary_values(ice_disc* disc, const double* bv_dev, void* stream);

/* L2 error of u against an analytic field (l2_error.hpp:37-94).  Synchronous. */
int ice_l2_error(ice_disc* disc, const ice_field* f, const double* u_dev, double* err, void* stream);

/* ---- element-block Jacobian (form_petsc_jacobian_fd DG part,
 *      form_petsc_jacobian.hpp:55-317, with the element block of
 *      ConservationLawDDG::domain_integral_jacobian, conservation_law.hpp:316-432)
 *
 * Block-sparse rows, one per element, nslots = 2*ndim + 1 blocks per row:
 * slot 0 is the diagonal block, slot 1+j couples the element with the element
 * across its local face j.  Blocks are bs x bs, bs = nbasis*neq, row-major
 * [i*neq+v][j*neq+w] like the reference's dense element blocks; rows and
 * columns are the fe_layout_right offsets of the two elements.  A slot whose
 * local face is a boundary face has column -1 and an all-zero block.  On a
 * periodic direction with fewer than three elements two slots of one row can
 * name the same column; their blocks then add. */
int ice_jacobian_layout(const ice_disc* disc, int* nslots, int* block_size, int* cols /* [nelem][nslots], or NULL */);
/* Finite-difference blocks at u (device) into blocks_dev[nelem][nslots][bs][bs]
 * (FP64), and the residual into r_dev unless it is NULL.  The face blocks
 * use the reference's eps = max(sqrt(eps_mach), sqrt(eps_mach) * max over the
 * face's element residual norms), the element blocks the per-quadrature-point
 * eps of domain_integral_jacobian; with sigma_ic != 0 the face blocks
 * include the DDGIC lifting (both forms).  1D and 2D at P1-P4, 3D at P1 and
 * P2 (ICE_ERR_UNSUPPORTED above P2 in 3D: the face kernel keeps its
 * per-column flux differences in LDS). */
int ice_jacobian(ice_disc* disc, const double* u_dev, double* r_dev, double* blocks_dev, void* stream);
/* y = J x for blocks in the layout above (FP64 blocks, FP64 vectors). */
int ice_jacobian_apply(const ice_disc* disc, const double* blocks_dev, const double* x_dev, double* y_dev, void* stream);

/* ---- Newton-linesearch steady solve (PetscNewton::solve,
 *      petsc_newton.hpp:205-329; ConvergenceCriteria and linesearches,
 *      nonlinear_solver_utils.hpp:23-268)
 *
 * Each step assembles the finite-difference block Jacobian (ice_jacobian),
 * solves J du = R with restarted GMRES right-preconditioned by block Jacobi
 * (the reference hands this to PETSc KSP, default GMRES(30) with SOR,
 * petsc_newton.hpp:174-183), picks alpha by the linesearch on
 * phi(alpha) = ||R(u - alpha du)||_2 (non-finite -> 1e9), sets u -= alpha du,
 * and stops when ||R||_2 <= tau_abs + tau_rel * ||R(u_0)||_2 or after kmax
 * steps.  The preconditioner is block Jacobi, one multicolour block
 * symmetric Gauss-Seidel sweep (elements coloured so that face neighbours
 * differ), or a V-cycle whose coarse levels are the Galerkin projections of
 * the Jacobian onto element-wise constants and their 2^ndim agglomerates
 * (the reference's default is PETSc's point SOR).  With ICE_PREC_FP32
 * the Krylov operator and the block-Jacobi factors
 * are FP32 copies of the FP64 blocks (FP64 vectors and accumulation) and the
 * correction is refined against the FP64 blocks: r = R - J du in FP64 until
 * ||r|| <= linear_rtol ||R|| or refine_max passes. */
#define ICE_LS_NONE 0
#define ICE_LS_WOLFE 1
#define ICE_LS_CORRIGAN 2
#define ICE_PREC_FP64 0
#define ICE_PREC_FP32 1
#define ICE_PREC_FP16 2
#define ICE_PC_BLOCK_JACOBI 0
#define ICE_PC_BLOCK_SGS 1
#define ICE_PC_MULTIGRID 2
#define ICE_PC_LINE 3
#define ICE_FORCING_FIXED 0
#define ICE_FORCING_EW 1
#define ICE_NEWTON_MAX_HIST 64

typedef struct {
    double tau_abs, tau_rel;  /* ConvergenceCriteria (defaults 1e-8, 1e-8) */
    int kmax;                 /* Newton steps (default 100) */
    int linesearch;           /* ICE_LS_* (default none) */
    int ls_kmax;              /* linesearch max_it (20) */
    double ls_alpha_initial;  /* 1 */
    double ls_alpha_max;      /* 10 */
    double ls_c1, ls_c2;      /* Wolfe constants (1e-4, 0.9) */
    double ls_alpha_min;      /* Corrigan alpha_min (0) */
    int jacobian_precision;   /* ICE_PREC_* (default FP32) */
    int gmres_restart;        /* 30 */
    int gmres_maxit;          /* 10000 */
    double linear_rtol;       /* 1e-5 */
    int refine_max;           /* FP64 refinement passes (3) */
    int verbose;              /* 1: one line per Newton step on stdout */
    int preconditioner;       /* ICE_PC_*: block Jacobi, multicolour block
                                 symmetric Gauss-Seidel (default), or a
                                 multigrid V-cycle over element agglomerates
                                 (P0 Galerkin coarse levels, SGS smoothing), or
                                 the exact block-tridiagonal (line) solve of a
                                 non-periodic 1D mesh (FP64 blocks) */
    int forcing;              /* ICE_FORCING_*: linear tolerance linear_rtol at
                                 every step (default, as the reference's KSP),
                                 or Eisenstat-Walker choice 2 */
    int smoother_precision;   /* ICE_PREC_FP32 (default) or ICE_PREC_FP16: the
                                 fine-level blocks the block-SGS smoother and
                                 the multigrid's fine residuals read (a fixed
                                 preconditioner either way; the Krylov operator
                                 keeps jacobian_precision); with FP32 blocks */
    int gmres_cgs_refine;     /* GMRES orthogonalisation by classical
                                 Gram-Schmidt: 0 (default) one pass, PETSc's
                                 GMRES default (KSP_GMRES_CGS_REFINE_NEVER)
                                 that the reference's KSP runs; 1 with one
                                 reorthogonalisation pass */
} ice_newton_opts;

typedef struct {
    int iterations;           /* Newton steps taken */
    int converged;
    double r0, rnorm;         /* ||R|| at u_0 and at the returned u */
    long long linear_iterations;
    int refinements;
    int singular_blocks;      /* diagonal blocks replaced by the identity in block Jacobi */
    double seconds, jacobian_seconds, linear_seconds, linesearch_seconds;
    int nhist;
    double rhist[ICE_NEWTON_MAX_HIST];   /* ||R|| after step k */
    double alpha[ICE_NEWTON_MAX_HIST];   /* linesearch multiplier of step k */
    int linear_its[ICE_NEWTON_MAX_HIST]; /* GMRES iterations of step k */
} ice_newton_report;

typedef struct ice_newton ice_newton;

/* Reference defaults (see the field comments). */
int ice_newton_opts_default(ice_newton_opts* opts);
/* Allocates the Jacobian (FP64 blocks, FP32 copies when asked), block-Jacobi
 * factors and the Krylov basis for this discretization. */
int ice_newton_create(ice_disc* disc, const ice_newton_opts* opts, ice_newton** out);
int ice_newton_destroy(ice_newton* solver);
/* Solve R(u) = 0 from the initial guess in u (device, fe_layout_right).
 * Synchronous; returns ICE_OK whether or not it converged (see report). */
int ice_newton_solve(ice_newton* solver, double* u_dev, ice_newton_report* report, void* stream);
/* The linear solve alone, J x = b for the Jacobian last assembled by the
 * solver at u (assembled here first when u_dev is not NULL); x and b device. */
int ice_newton_linear_solve(ice_newton* solver, const double* u_dev, const double* b_dev, double* x_dev,
                            int* iterations, double* rel_residual, void* stream);

/* ---- explicit time integration (RK3SSP ssp_rk3.hpp:106-233, RK3TVD
 *      tvd_rk3.hpp:107-270, ExplicitEuler explicit_euler.hpp:87-160) ----- */
#define ICE_STEP_RK3_SSP 0
#define ICE_STEP_RK3_TVD 1
#define ICE_STEP_EXPLICIT_EULER 2

typedef struct ice_stepper_desc {
    int kind;                 /* ICE_STEP_* */
    int reference_compat;     /* RK3_SSP only: 1 evaluates every stage residual at u^n
                                 as ssp_rk3.hpp:137 does; 0 = textbook SSP-RK3 */
    int use_cfl;              /* 1: CFLTimestep (explicit_utils.hpp:61-90), 0: FixedTimestep */
    double dt;                /* fixed step */
    double cfl;               /* CFL number */
    int use_tfinal;           /* 1: TfinalTermination, 0: TimestepTermination */
    double tfinal;
    long long ntime;
} ice_stepper_desc;

typedef struct ice_stepper ice_stepper;

int ice_stepper_create(ice_disc* disc, const ice_stepper_desc* desc, ice_stepper** out);
int ice_stepper_destroy(ice_stepper* stepper);

/* Per-element step sizes (host array [nelem], positive) for a fixed-step
 * stepper, e.g. each batched instance's own CFL step; null restores the
 * descriptor's dt.  No reference counterpart. */
int ice_stepper_set_element_dt(ice_stepper* stepper, const double* dt);

/* One HIP graph holding nsteps steps of every stepper, each advancing its own
 * device state u[k]: a single graph launch then runs all of them.  Fixed,
 * per-element and CFL step sizes all stay on the device (a CFL stepper's
 * wave-speed maxima are reset by its step kernels), and the termination
 * condition is a device flag, so the steps need no host synchronisation;
 * steps past the termination change nothing.  The state pointers are baked
 * into the graph and the executable graph is uploaded at creation.  No
 * reference counterpart. */
typedef struct ice_step_graph ice_step_graph;
int ice_step_graph_create(ice_stepper* const* steppers, double* const* u_dev, int n, int nsteps,
                          ice_step_graph** out);
/* The same for FP64 (state_fp32 = 0, u_dev[k] double*) or FP32 states
 * (state_fp32 = 1, u_dev[k] float*, stepped as ice_stepper_step_f32 does).
 * With concurrent = 1 and every stepper on its own discretization and state,
 * each stepper's steps are an independent branch of the graph, so steppers of
 * different sizes (e.g. one per polynomial order of a batch) can share the
 * device; otherwise the steps form one chain.  No reference counterpart. */
int ice_step_graph_create_ex(ice_stepper* const* steppers, void* const* u_dev, int state_fp32, int n, int nsteps,
                             int concurrent, ice_step_graph** out);
int ice_step_graph_launch(ice_step_graph* graph, void* stream);
int ice_step_graph_destroy(ice_step_graph* graph);
/* One step of u (device, fe_layout_right) on `stream`; time step, time and
 * step count live on the device, so consecutive steps need no host sync. */
int ice_stepper_step(ice_stepper* stepper, double* u_dev, void* stream);
/* One step of an FP32 state u (device, fe_layout_right, 2D and 3D; textbook
 * RK3-SSP or explicit Euler): stage residuals are the mixed-precision
 * residual of the FP32 coefficients (FP64 output), increments M^-1 R are kept
 * in FP32, each stage combination is formed in FP64 and rounded once to FP32.
 * The first call allocates the stepper's FP32 stage buffers.  No reference
 * counterpart (the reference's states are FP64). */
int ice_stepper_step_f32(ice_stepper* stepper, float* u_dev, void* stream);
/* Residual of an FP32 state: check_fp64 = 0 gives the mixed-precision residual
 * the FP32 steps use; check_fp64 = 1 widens the state to FP64 (a device copy
 * the discretization allocates on first use) and forms the FP64 residual, the
 * check of an FP32 solution against the FP64 discretization.  r is FP64. */
int ice_residual_f32state(ice_disc* disc, const float* u_dev, double* r_dev, int check_fp64, void* stream);
/* solve(): one residual for the wave speeds, then steps until termination. */
int ice_stepper_solve(ice_stepper* stepper, double* u_dev, void* stream, long long* nsteps, double* time);
/* Synchronous read of the device-side time, last step size and step count. */
int ice_stepper_status(ice_stepper* stepper, double* time, double* dt, long long* itime, int* done);

/* Kernel timing (no reference counterpart; the reference has no
 * instrumentation).  Between begin and end, every residual on this handle
 * records HIP events around its three launches on the caller's stream; end
 * waits for them and writes the summed milliseconds to stage_ms[0] (element
 * face traces), stage_ms[1] (face fluxes) and stage_ms[2] (element volume
 * terms and lifting), and the number of residuals to *ncalls. */
int ice_profile_begin(ice_disc* disc, int max_calls);
int ice_profile_end(ice_disc* disc, double* stage_ms, int* ncalls);

/* Host-side evaluation of an analytic field at npts points x[npts][ndim]
 * (the callbacks of conservation_law.hpp:146-170 and of the initial
 * conditions, initialization.hpp:37-104). */
int ice_field_eval_points(const ice_field* f, int ndim, int neq, int npts, const double* x, double* out);

/* Physical coordinates of the Lagrange nodes of every element, host array
 * [nelem][nbasis][ndim] (the nodal points of HypercubeLagrangeBasis mapped by
 * the element transformation). */
int ice_space_node_coords(const ice_space* space, double* xn);

/* Reference-element point of every face quadrature point for face number
 * `face_nr` seen with orientation `orientation` (TraceSpace transform_xiL /
 * transform_xiR, hypercube_face.hpp:230-252): xi[nfqp][ndim], host-only. */
int ice_trace_points(int ndim, int order, int face_nr, int orientation, double* xi);

/* ---- warm-start network (no reference counterpart) ----------------------
 * A fully connected MLP, tanh on hidden layers and identity on the output,
 * trained and evaluated on the device by this library's own kernels
 * (FP32 matrix-core GEMMs).  It predicts DG coefficients of a steady solution
 * from problem parameters, as a Newton initial guess; Newton's FP64 residual
 * decides whether the guess is used.  Arrays are FP32, row-major:
 * inputs [batch][widths[0]], outputs [batch][widths[nlayers]]; parameters are
 * packed layer by layer as W_l [widths[l]][widths[l+1]] then b_l.  Loss: mean
 * squared error over batch x outputs.  A handle is not thread-safe. */
typedef struct ice_mlp ice_mlp;
typedef struct {
    int batch;         /* minibatch rows (rows beyond the last full batch of an epoch are skipped) */
    int epochs;
    double lr;         /* Adam step, multiplied by lr_decay after every epoch */
    double lr_decay;
    double beta1, beta2, eps;
    unsigned long long seed;  /* minibatch shuffling */
} ice_mlp_train_opts;

/* ---- MDG interface conservation (moving discontinuous Galerkin) ----------
 * form_mdg_residual over a nodeset_dof_map (form_residual.hpp:270-327,
 * node_set_layout.hpp:46-83) with ConservationLawDDG::interface_conservation
 * (conservation_law.hpp:1178-1288).  A selection is a list of trace indices
 * in the reference's FESpace::traces order (interior faces, then boundary
 * faces, mesh face order); its degrees of freedom are the nodes of those
 * traces minus every node of a boundary face, in ascending node order.
 * Single meshes without coupled periodic faces.  ice_mdg_residual writes
 * res[node][neq] (FP64, device) for the current u and mesh coordinates. */
typedef struct ice_mdg_sel ice_mdg_sel;
int ice_mdg_select(const ice_space* space, const int* traces, int ntraces, ice_mdg_sel** out);
/* number of selected nodes, and their global node ids when nodes != NULL */
int ice_mdg_query(const ice_mdg_sel* sel, int* nnodes, int* nodes);
int ice_mdg_residual(ice_disc* disc, const ice_mdg_sel* sel, const double* u_dev, double* res_dev, void* stream);
/* The combined residual over the selection (form_residual with a
 * nodeset_dof_map, form_residual.hpp:380-424): u_dev = [DG coefficients
 * (fe_layout_right) | coordinates of the selected nodes [nnodes][ndim]],
 * res_dev = [DG residual | MDG residual [nnodes][neq]].  The selected nodes
 * are first moved to the given coordinates in the space's device geometry
 * (element vertices, face geometry table, source load), which keeps them for
 * every later call on the space, as the reference's mesh keeps them; host-side
 * quantities taken at creation (a stepper's CFL reference length) are not
 * updated. */
int ice_mdg_form_residual(ice_disc* disc, const ice_mdg_sel* sel, const double* u_dev, double* res_dev,
                          void* stream);
int ice_mdg_destroy(ice_mdg_sel* sel);
/* The interface error (ic_error, l2_error.hpp:279-410): the sum over interior
 * faces, face points and equations of |jump of the normal physical flux|
 * times w sqrt(g), gradients kept at every order.  Synchronous; *err host. */
int ice_interface_error(ice_disc* disc, const double* u_dev, double* err, void* stream);

/* Glorot-uniform weights (seeded), zero biases.  widths has nlayers+1 entries. */
int ice_mlp_create(int nlayers, const int* widths, unsigned long long seed, ice_mlp** out);
int ice_mlp_destroy(ice_mlp* mlp);
int ice_mlp_num_params(ice_mlp* mlp, long long* n);
/* host copies of the packed parameters; set also resets the Adam moments */
int ice_mlp_get_params(ice_mlp* mlp, float* host);
int ice_mlp_set_params(ice_mlp* mlp, const float* host);
/* y_dev <- network(x_dev), device pointers */
int ice_mlp_forward(ice_mlp* mlp, const float* x_dev, int batch, float* y_dev, void* stream);
/* loss (host, may be null) and the gradient of the loss in the packed
 * parameter layout (device, may be null) at targets t_dev; synchronous when
 * loss is requested */
int ice_mlp_gradient(ice_mlp* mlp, const float* x_dev, const float* t_dev, int batch, double* loss, float* grad_dev,
                     void* stream);
/* Adam training over n rows x_dev/t_dev (device); loss_hist (host, may be
 * null) receives the mean minibatch loss of each epoch.  Synchronous. */
int ice_mlp_train(ice_mlp* mlp, const float* x_dev, const float* t_dev, int n, const ice_mlp_train_opts* opts,
                  double* loss_hist, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ICEICLE_AMD_H */
