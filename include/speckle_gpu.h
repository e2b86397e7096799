/*
 * speckle_gpu.h -- C-ABI of the MI355X SIE solve loop (libspeckle_gpu.so).
 *
 * Drop-in boundary for the reference module API of SpeckleSim's
 * triangle MLFMM-GMRES path.  Every entry point takes plain pointers and
 * sizes; complex vectors are interleaved (re, im) doubles, i.e. Fortran
 * complex(8) / C double _Complex layout, in the reference's unknown order
 * x = [xe(1:N); xh(1:N)] (src/lib_sie_tri_mlfmm_interface.f90:300-311).
 * Indices are 0-based unless stated; mesh triangles are 1-based as the
 * reference reads them.  All functions return SPK_OK (0) or a negative
 * error code; spk_last_error() gives the message of the calling thread.
 * A handle is used by one host thread at a time (the reference calls the
 * matvec from its single GMRES reverse-communication loop).
 */
#ifndef SPECKLE_GPU_H
#define SPECKLE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPK_OK 0
#define SPK_ERR_ARG -1
#define SPK_ERR_HIP -2
#define SPK_ERR_STATE -3
#define SPK_ERR_MEM -4
#define SPK_ERR_MESH -5

/* Formulations, src/lib_sie_tri_calculation_mod.f90:463-644. */
#define SPK_PMCHWT 1
#define SPK_MCTF 2
#define SPK_ICTF 3
#define SPK_MCTF2 4
#define SPK_PMCHWT_S 5
#define SPK_JMCFIE 6
#define SPK_CNF 7
#define SPK_MNMF 8

/* Arithmetic of the MLFMM far field / inner solve. */
#define SPK_FP64 0
#define SPK_MIXED 1 /* complex64 operator + inner GMRES, FP64 refinement (no counterpart in the reference) */

/* Orthogonalisation, lib_math_solver_GMRES.f90 (icntl(5)). */
#define SPK_ORTH_MGS 0
#define SPK_ORTH_IMGS 1
#define SPK_ORTH_CGS 2
#define SPK_ORTH_ICGS 3

typedef struct spk_handle spk_handle;

/* Solver parameters: the reference's module globals set by main.f90
 * (set_calculation_parameters, initalize_gmres) and
 * set_local_material_parameters_tri (src/lib_sie_tri_data_container.f90:212-224). */
typedef struct spk_params {
  double lambda;                 /* vacuum wavelength [m] */
  double eps_r1_re, eps_r1_im;   /* exterior relative permittivity */
  double eps_r2_re, eps_r2_im;   /* interior relative permittivity */
  int formulation;               /* SPK_PMCHWT ... */
  int normal_convention;         /* calc_p(7): 1 or 2 */
  int quad_points;               /* ng: 1,3,4,6,7,12 (reference: 3) */
  int tree_s_opt;                /* tree_s_opt (reference: 250) */
  int truncation_number;         /* K at every level (reference: 19) */
  int precision;                 /* SPK_FP64 or SPK_MIXED */
  int device;                    /* HIP device ordinal */
} spk_params;

/* GMRES knobs: lib_sie_solver_gmres_get_parameter
 * (src/lib_sie_solver_gmres_interface.f90:102-121). */
typedef struct spk_gmres_params {
  int max_iterations;            /* icntl(7) */
  int restart;                   /* m (<=0: = max_iterations) */
  double tolerance;              /* cntl(1), Arnoldi backward error */
  int orthogonalization;         /* SPK_ORTH_* */
  int use_initial_guess;         /* icntl(6) */
  int refinement_steps;          /* SPK_MIXED: maximum outer refinement steps (0: 30) */
  double refinement_tolerance;   /* SPK_MIXED: relative residual of each inner complex64
                                    solve (0: 3e-4), relaxed to tolerance / (2 |r|/|b|) for
                                    the later solves; tolerance is the FP64 target */
  int preconditioning;           /* icntl(4): SPK_PRECOND_NONE or SPK_PRECOND_LEFT_JACOBI
                                    (precondition_left, src/lib_sie_solver_gmres_interface.f90:115;
                                    diagonal of the self blocks, lib_sie_tri_mlfmm_interface.f90:94-109);
                                    FP64 only.  SPK_PRECOND_NEAR_FIELD (no counterpart in the
                                    reference): flexible GMRES on the FP64 operator, right-
                                    preconditioned by inner_iterations GMRES steps on the
                                    near-field operator (complex64 vectors and fp8 blocks on
                                    an SPK_MIXED handle, where the outer products also switch
                                    to the complex64 operator once the residual estimate is
                                    at or below 200 x tolerance; the true residual that ends
                                    the solve is FP64 in every case).  SPK_PRECOND_BLOCK_LU (no
                                    counterpart in the reference): the same flexible GMRES,
                                    right-preconditioned by the inverse of each group's dense
                                    near-field matrix (groups = the leaf boxes under one box a
                                    level up; rocSOLVER getrf/getri once per operator, one
                                    batched GEMV per step); for systems the inner near-field
                                    GMRES does not precondition enough (the quadrilateral
                                    system); set restart (default 100) to the steps it needs */
  int inner_iterations;          /* SPK_PRECOND_NEAR_FIELD: inner steps per outer step (0: 30) */
} spk_gmres_params;
#define SPK_PRECOND_NONE 0
#define SPK_PRECOND_LEFT_JACOBI 1
#define SPK_PRECOND_NEAR_FIELD 2
#define SPK_PRECOND_BLOCK_LU 3

/* Per-level truncation numbers for the next spk_setup_tri: K[l] for levels
 * l < n (0: spk_params.truncation_number), non-increasing from the
 * coarsest level to the leaves.  The reference computes per-level values
 * in Truncation_number_calculation (src/lib_sie_mlfmm.f90:631-701) and then
 * sets every level to truncation_number_default; its SS and RR translations
 * (src/lib_sie_tri_mlfmm_interface.f90:466-528, 600-655) interpolate
 * between the levels' sample grids (Lagrange_Interpolation_New,
 * src/lib_sie_mlfmm.f90:268-390) where they differ.  SS interpolates a
 * child's coefficients to the parent's grid.  RR: SPK_RR_REFERENCE shifts
 * the parent's local coefficients and interpolates them to the child's
 * grid, as the reference's code does (an operator far from the dense one:
 * the local coefficients carry the translation's bandwidth and the
 * parent's quadrature weights); SPK_RR_ANTERPOLATE folds every level's
 * coefficients and applies the transpose of SS's interpolation (the
 * operator then stays within the interpolation error of the dense one).
 * FP64 handles only; the batched, transposed and complex64 products reject
 * such a handle.  No counterpart as a routine. */
#define SPK_RR_REFERENCE 0
#define SPK_RR_ANTERPOLATE 1

typedef struct spk_gmres_info {
  int info;                      /* info(1): 0 converged, -4 not converged */
  int iterations;                /* info(2) (sum over refinement sweeps) */
  double arnoldi_backward_error; /* last bea */
  double backward_error;         /* rinfo(1): ||M^-1 (b-Ax)|| / ||M^-1 b|| (M = I unpreconditioned) */
  int matvecs;
  double seconds;
  double true_backward_error;    /* rinfo(2): ||b-Ax|| / ||b|| */
} spk_gmres_info;

typedef struct spk_stats {
  int64_t n_edges;               /* m_pairs */
  int64_t near_blocks;           /* nnz of the near-field CSR */
  int l_min, l_max;
  int samples;                   /* S = 2K^2+2 */
  int64_t boxes_per_level[32];
  int64_t i4_pairs_per_level[32];
  int64_t singular_pairs;        /* triangle pairs on the singular branch */
  double setup_ms, near_fill_ms, patterns_ms, tables_ms, tree_ms;
  double bytes_near, bytes_patterns;
} spk_stats;

/* Create a context on the chosen device (no mesh yet). */
int spk_create(const spk_params *params, spk_handle **out);
void spk_destroy(spk_handle *h);
const char *spk_last_error(void);

/* lib_sie_ml_fmm_precalculation_tri (src/lib_sie_tri_mlfmm_interface.f90:59-92):
 * RWG topology, octree, radiation patterns, TL tables and the near-field
 * impedance fill on the GPU.  pp: 3*np node coordinates [m];
 * tt: 3*nt node indices, 1-based (read_pp_tt_file_tri). */
int spk_setup_tri(spk_handle *h, int np, const double *pp, int nt, const int *tt);

/* The MLFMM operator on an 8-node quadrilateral height-field surface (the
 * mesh of sorting_quad, src/lib_sie_quad_data_container.f90:321-403: Lx x Ly
 * (m), nx x ny elements, z[(2nx+1)(2ny+1)] node heights, x index fastest)
 * with the 10 edge functions per element of Sorting_edge
 * (src/lib_sie_quad_calculation_mod.f90:624-698).  No counterpart in the
 * reference, whose quadrilateral solve is dense LU
 * (lib_sie_quad_solver_normal.f90:62-270) and whose MLFMM pattern slot for
 * quadrilaterals (preCalculated_RT_quad, src/lib_sie_data_container.f90:84)
 * is never filled.  Each edge function sits in the octree at the midpoint
 * of its node pair (find_midpoint_edge_quad, :210-255); the near blocks are
 * the dense system's entries (impedance_edge_versus_element and the
 * assembly of :165-210); the far part is the plane-wave form of the dense
 * fill's regular rule.  PMCHWT, MCTF or ICTF; uniform truncation number.
 * Afterwards the handle's products, right-hand sides, solves (one vector or
 * a batch, FP64 or mixed) and scattered fields (Integration_scattered_field_quad,
 * :1281-1323, ngp 5, on the element-local coefficients of
 * lib_sie_quad_solver_normal.f90:312-323) work on the quadrilateral system
 * of 2M unknowns; the transposed operator, the adjoint gradient and the
 * triangle probes refuse it. */
int spk_setup_quad(spk_handle *h, double Lx, double Ly, int nx, int ny, const double *z);

/* ------------------------------------------------------------------------ */
/* 8-node quadrilaterals with 10 Huber edge functions: the near/dense fill   */
/* ------------------------------------------------------------------------ */
typedef struct spk_quad spk_quad;
/* Handle for the quadrilateral fill: lambda, eps_r1/eps_r2 and formulation
 * (PMCHWT, MCTF or ICTF, the three the reference assembles) from params;
 * set_local_material_parameters_quad (src/lib_sie_quad_data_container.f90:118-129). */
int spk_quad_create(const spk_params *params, spk_quad **out);
void spk_quad_destroy(spk_quad *q);
/* The height-field surface of sorting_quad (src/lib_sie_quad_data_container.f90:321-403):
 * Lx x Ly (m), nx x ny 8-node elements, z[(2nx+1)*(2ny+1)] node heights in the
 * height file's order (x index fastest); then the edge/corner neighbours
 * (Neighbours_EdgeCorner_Full, src/lib_sie_quad_calculation_mod.f90:952-995) and
 * the numbering of the edge functions (Sorting_edge, :624-698). */
int spk_quad_setup(spk_quad *q, double Lx, double Ly, int nx, int ny, const double *z);
int spk_quad_sizes(const spk_quad *q, int *n_edges, int *n_elements, int *n_nodes);
/* nodes pp[3*n_nodes] and element vertices vert8[8*n_elements] (1-based) */
int spk_quad_get_mesh(const spk_quad *q, double *pp, int *vert8);
/* Edge_I(s)%edge_p(1:5) for all edge functions: element, local index,
 * neighbour element, its local index, orientation (1-based, 0 = none) */
int spk_quad_get_edges(const spk_quad *q, int *edge_p5);
/* impedance_edge_versus_element(s, t, Sum_aa(3, 10))
 * (src/lib_sie_quad_singularity.f90:22-98) for n pairs, 0-based edge
 * function s[i] and element t[i]; out[i*60 + (k*10 + n)*2 + {0,1}] */
int spk_quad_edge_element(spk_quad *q, int n, const int *s, const int *t, double *out);
/* The assembled 2M x 2M system of the formulation
 * (src/lib_sie_quad_solver_normal.f90:165-210), row-major complex: host or
 * device destination. */
int spk_quad_fill(spk_quad *q, double *MM);
int spk_quad_fill_device(spk_quad *q, void *d_MM);
/* get_vector_b_quad with plane-wave illumination (incident_int,
 * src/lib_sie_quad_calculation_mod.f90:91-138, 1207-1240): b[2M] complex,
 * [E part; H part] scaled by the formulation; host or device destination. */
int spk_quad_rhs_plane_wave(spk_quad *q, const double k_in[3], const double e_in[3], double *b);
int spk_quad_rhs_plane_wave_device(spk_quad *q, const double k_in[3], const double e_in[3],
                                   void *d_b);
/* The reference's quadrilateral solve, solver_LU (ZGESV,
 * src/lib_sie_math.f90:649-676) as lib_sie_quad_solver_normal_I calls it
 * (src/lib_sie_quad_solver_normal.f90:216-217), in FP64: rocSOLVER's LU with
 * partial pivoting, then blocked triangular solves (rocBLAS).  d_MM is
 * the row-major 2M x 2M system of spk_quad_fill_device and is overwritten by
 * its LU factors; d_b holds nrhs right-hand sides of 2M (column j at
 * d_b + 2M j) and is overwritten by the solutions.  An exactly zero pivot
 * returns SPK_ERR_STATE (where the reference stops). */
int spk_quad_solve_lu_device(spk_quad *q, void *d_MM, int nrhs, void *d_b);
/* Fill, plane-wave right-hand side and LU solve in one call (the device
 * matrix, 16 (2M)^2 bytes, lives for the call only): x[2M] complex, the
 * V_EH that lib_sie_quad_solver_normal_I writes out (:225-229). */
int spk_quad_solve_plane_wave(spk_quad *q, const double k_in[3], const double e_in[3], double *x);
/* lib_sie_quad_solver_normal_II (src/lib_sie_quad_solver_normal.f90:272-335):
 * the element-local coefficients S_EH of a solution x[2M] and the scattered
 * E, H (3 complex each per point) at npts points r[3*npts]
 * (lib_sie_quad_observation_field_calculation, :337-, five-point rule);
 * arguments in the order of spk_scattered_field. */
int spk_quad_scattered_field(spk_quad *q, int npts, const double *r, const double *x, double *E,
                             double *H);
/* Host-only (no GPU): mesh, sizes and Edge_I of spk_quad_setup; arrays may
 * be NULL to query the sizes first. */
int spk_quad_topology_probe(double Lx, double Ly, int nx, int ny, const double *z, int *n_edges,
                            int *n_elements, int *n_nodes, double *pp, int *vert8, int *edge_p5);

/* One surface realisation: np nodes pp[3*np] (m) and nt triangles tt[3*nt]
 * (1-based node numbers), as spk_setup_tri takes them. */
typedef struct spk_mesh {
  int np;
  const double *pp;
  int nt;
  const int *tt;
} spk_mesh;

/* A speckle ensemble: n_mesh realisations, each solved for n_angles plane
 * waves (k_in[3*a], e_in[3*a], as spk_rhs_plane_wave) with gp (in batches of
 * up to 16 right-hand sides, spk_gmres_batch_device), then the scattered
 * field's intensity sum_c |E_c|^2 at npts points r[3*npts] (m):
 * intensity[(m * n_angles + a) * npts + p], info[m * n_angles + a].
 * n_workers (1..16) host threads, each with its own handle created from
 * params and its own HIP stream, take the realisations in order, so one
 * realisation's host precalculation overlaps another's GPU work.  No
 * counterpart in the reference, whose angle loop (main.f90:450-463) solves
 * one realisation and one angle at a time; a failure stops the ensemble and
 * names the realisation.  gp->inner_iterations < 0 (near-field
 * preconditioned solves) first tunes on realisation 0 (spk_autotune with
 * its first min(16, n_angles) angles) and applies the result to every
 * worker's handle. */
int spk_ensemble_intensity(const spk_params *params, int n_mesh, const spk_mesh *meshes,
                           int n_angles, const double *k_in, const double *e_in, int npts,
                           const double *r, const spk_gmres_params *gp, int n_workers,
                           double *intensity, spk_gmres_info *info);

/* Mixed-precision solves of R <= 16 right-hand sides of one mesh (incidence
 * angles) on an SPK_MIXED handle: each runs the refinement of spk_gmres
 * (FP64 residuals, complex64 inner GMRES with its own basis); the inner
 * steps of all right-hand sides share one batched complex64 matvec.  d_b,
 * d_x: host arrays of R device pointers (2N complex128 each); info[R].  No
 * counterpart in the reference, which solves one incidence at a time. */
int spk_gmres_batch_device(spk_handle *h, int R, const void *const *d_b, void *const *d_x,
                           const spk_gmres_params *params, spk_gmres_info *info);

/* No counterpart in the reference (a probe of the near-field preconditioner
 * of SPK_PRECOND_NEAR_FIELD): the near-field part of the operator of an
 * SPK_MIXED handle on R <= 16 complex64 device vectors (2N interleaved float
 * pairs each), with the blocks in complex64 (SPK_NEAR_C64), in the bf16 copy
 * (SPK_NEAR_BF16) or in the fp8 E4M3FN copy with per-row component scales
 * (SPK_NEAR_FP8).  batched = 0: one single-vector launch per vector;
 * batched = 1: the batched launch the inner solves of a batch use (for fp8,
 * the matrix-core product from R = 9, complex64 blocks below). */
#define SPK_NEAR_C64 0
#define SPK_NEAR_BF16 1
#define SPK_NEAR_FP8 2
int spk_near_product_lowp(spk_handle *h, int format, int batched, int R, const void *const *d_x,
                          void *const *d_y);

/* Device memory for the *_device entry points (no counterpart in the
 * reference).  A caller that has its own HIP allocations may pass those
 * instead; they must come from the same HIP runtime as this library. */
#define SPK_COPY_H2D 1
#define SPK_COPY_D2H 2
#define SPK_COPY_D2D 3
int spk_device_alloc(int device, size_t bytes, void **d_ptr);
int spk_device_free(void *d_ptr);
/* Waits for all work on the current device (every handle's stream), then
 * copies. */
int spk_device_copy(void *dst, const void *src, size_t bytes, int direction);
int spk_device_synchronize(int device);

int spk_get_stats(const spk_handle *h, spk_stats *out);
int spk_num_edges(const spk_handle *h, int64_t *n_edges);

/* lib_sie_ml_fmm_calculate_vector_b_tri (:285-319): y = Z x on host
 * buffers of 2N complex. */
int spk_matvec(spk_handle *h, const double *x, double *y);
/* Same on device buffers (HBM-resident), stream 0 of the handle. */
int spk_matvec_device(spk_handle *h, const void *d_x, void *d_y);
/* The handle's operator becomes its transpose Z^T (transposed != 0) or Z
 * again (0): every matvec, solve and low-precision copy that follows uses
 * Z^T (the near blocks are transposed in HBM, the MLFMM far field runs its
 * stages transposed), except the batched products of 2..16 vectors, which
 * reject it.  For the adjoint system Z^T lambda = dJ/dI.  No counterpart in
 * the reference.  spk_near_blocks keeps evaluating Z's blocks. */
int spk_set_transposed(spk_handle *h, int transposed);
int spk_get_transposed(const spk_handle *h, int *transposed);
/* The near-field preconditioned solves capture their inner near-field solve
 * (inner_iterations steps of near-field product and Arnoldi update) once as
 * a HIP graph and replay it at every outer step (enable = 1, the default),
 * or launch its kernels one by one (0).  A capture in progress makes HIP
 * reject synchronous copies issued by other host threads, so a process that
 * drives several handles from several threads disables it on each of them
 * (spk_ensemble_intensity does).  The environment variable
 * SPK_INNER_GRAPH=0 disables it for every handle. */
int spk_set_inner_graph(spk_handle *h, int enable);
/* The complex64 operator of an SPK_MIXED handle (the inner operator of the
 * mixed solve): y = Z x with near blocks, patterns and tables rounded to
 * complex64; vectors of 2N interleaved float pairs.  No counterpart in the
 * reference (which is FP64 only). */
int spk_matvec_c64(spk_handle *h, const float *x, float *y);
int spk_matvec_c64_device(spk_handle *h, const void *d_x, void *d_y);
/* The matvec for R <= 16 right-hand sides at once (incidence angles of one
 * mesh; no counterpart in the reference, which applies
 * lib_sie_ml_fmm_calculate_vector_b_tri to one vector at a time): d_x, d_y
 * are host arrays of R device pointers to 2N-complex vectors.  The operator
 * is read once per call for all R vectors. */
int spk_matvec_batch_device(spk_handle *h, int R, const void *const *d_x, void *const *d_y);
/* Same for the complex64 operator of an SPK_MIXED handle (2N complex64 vectors). */
int spk_matvec_batch_c64_device(spk_handle *h, int R, const void *const *d_x, void *const *d_y);
/* Far field only / near field only (parity probes, timing splits). */
int spk_matvec_parts(spk_handle *h, const double *x, double *y_near, double *y_far);

/* get_vector_b_tri with plane-wave illumination
 * (src/lib_sie_tri_calculation_mod.f90:391-460, 939-1002). */
int spk_rhs_plane_wave(spk_handle *h, const double k_in[3], const double e_in[3], double *b);
/* Same, written to a device buffer of 2N complex (HBM-resident). */
int spk_rhs_plane_wave_device(spk_handle *h, const double k_in[3], const double e_in[3],
                              void *d_b);

/* Integration_scattered_field_tri (src/lib_sie_tri_calculation_mod.f90:2782-2825):
 * scattered E and H in the exterior medium at npts points r[3*npts] (m) from
 * the surface currents x = [J; M] (2N complex, reference order), with the
 * reference's 3-point rule; E, H: 3*npts complex each.  As in the reference,
 * the charge term of E is omitted (:2811 multiplies it by 0).  The BRDF and
 * RCS evaluations of src/lib_sie_tri_solver_normal.f90:325-358 are built on
 * this call. */
int spk_scattered_field(spk_handle *h, int npts, const double *r, const double *x, double *E,
                        double *H);
/* Same with x in device memory. */
int spk_scattered_field_device(spk_handle *h, int npts, const double *r, const void *d_x,
                               double *E, double *H);
/* spk_scattered_field_device without allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The points and the E/H staging live in a library-owned device array of
 * 1.9 MB per GPU, held until the process ends; npts above
 * SPK_SCATTERED_NOALLOC_MAX_POINTS is rejected with SPK_ERR_ARG.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (from several
 * host threads or handles) race.  Results equal spk_scattered_field_device's. */
#define SPK_SCATTERED_NOALLOC_MAX_POINTS 16384
int spk_scattered_field_device_noalloc(spk_handle *h, int npts, const double *r,
                                       const void *d_x, double *E, double *H);

/* Adjoint gradient of a far-field intensity with respect to mesh-node
 * heights.  J = sum_p w[p] |E(r_p)|^2, E the scattered field (as
 * spk_scattered_field) at npts points r[3*npts] of the solution of Z I = b,
 * b the plane wave (k_in, e_in); grad[k] = dJ/dz of node nodes[k] (0-based
 * indices into the pp of spk_setup_tri).  One solve of Z I = b and one of
 * Z^T nu = F^T W conj(E) with gp, then for every node the central
 * differences (node moved by +-dz along z) of the near blocks, radiation
 * patterns, right-hand-side entries and detector fields it changes,
 * contracted with nu, I and the leaf fields: dJ = 2 Re[E^H W dF I +
 * nu^T (db - dZ I)].  Near blocks keep the singular / regular branch of the
 * unmoved mesh.  J is returned too; terms (may be NULL): 4 * n_nodes
 * values, node k's gradient split into the detector-field, right-hand-side,
 * near-block and far-field shares (their sum is grad[k]); info[2] (may be
 * NULL): the two solves.  No counterpart in the reference. */
int spk_adjoint_height_gradient(spk_handle *h, const double k_in[3], const double e_in[3],
                                int npts, const double *r, const double *w,
                                const spk_gmres_params *gp, int n_nodes, const int *nodes,
                                double dz, double *J, double *grad, double *terms,
                                spk_gmres_info *info);

/* Run-time tuning of the near-field preconditioned solver on the handle's
 * current mesh (no counterpart in the reference).  Times the batched fp8
 * near product with 1, 2 and 4 row tiles per workgroup (SPK_MIXED handles;
 * 16 vectors, HIP events) and one batched solve of the R incidences
 * (k_in[3R], e_in[3R]) with gp for each inner length 20, 25, 30, 36, 40;
 * keeps the fastest of each in the handle (used while
 * spk_gmres_params.inner_iterations is 0, and across spk_setup_tri calls).
 * out reports the times (inner_ms < 0: that length did not converge). */
typedef struct spk_tuning {
  int tiles_per_group;   /* chosen: 1, 2 or 4 */
  int inner_iterations;  /* chosen inner near-field GMRES steps */
  double tile_us[3];     /* the batched fp8 product with 1, 2, 4 tiles (0: not timed) */
  int n_inner;           /* inner lengths tried */
  int inner_tried[8];
  double inner_ms[8];    /* one batched solve with each */
} spk_tuning;
int spk_autotune(spk_handle *h, int R, const double *k_in, const double *e_in,
                 const spk_gmres_params *gp, spk_tuning *out);
int spk_set_level_truncation(spk_handle *h, int n, const int *K, int rr_mode);
/* The interpolation between the K_m and K_n sample grids as the library
 * builds it, applied on the host (no GPU) to one current's four channels
 * (theta k1, theta k2, phi k1, phi k2) of 2K^2 + 2 complex samples each:
 * a check of the library's construction.  No counterpart as a routine. */
int spk_interp_probe(int Km, int Kn, const double *in, double *out);

/* Applies tuned values (e.g. spk_autotune's on another handle of a similar
 * mesh); 0 restores a default.  No counterpart in the reference. */
int spk_set_tuning(spk_handle *h, int tiles_per_group, int inner_iterations);

/* lib_math_solver_gmres_run (src/lib/libmath/src/math/lib_math_solver_GMRES.f90:159-379)
 * with the GPU matvec; x holds the initial guess on entry. */
int spk_gmres(spk_handle *h, const double *b, double *x, const spk_gmres_params *p,
              spk_gmres_info *info);
int spk_gmres_device(spk_handle *h, const void *d_b, void *d_x, const spk_gmres_params *p,
                     spk_gmres_info *info);
/* Arnoldi backward error per iteration of the last solve (host copy). */
int spk_gmres_history(const spk_handle *h, double *hist, int cap, int *count);

/* Parity probes. */
/* Pair%corner / Pair%element of every RWG edge (1-based, N x 2 each). */
int spk_get_rwg(const spk_handle *h, int *corner, int *element);
/* Leaf Morton index of every edge, tree scaling D and x_min. */
int spk_get_tree(const spk_handle *h, int64_t *leaf_uindex, int *l_min, int *l_max,
                 double *D, double *xmin3);
/* I4 list of box n at level l: returns count in *count (cap entries written). */
int spk_get_i4(const spk_handle *h, int level, int64_t n, int64_t *out, int cap, int *count);
/* Near-field CSR in the reference's layout (ia, ja 0-based, N+1 / nnz)
 * and values D11,D12,D21,D22 per entry (nnz*8 doubles); any pointer may be
 * NULL to skip it. */
int spk_get_near_csr(const spk_handle *h, int64_t *ia, int32_t *ja, double *abcd);
/* Blocks Z(m,n) for arbitrary pairs (lib_sie_tri_get_impedance_edges):
 * out = npairs*8 doubles. */
int spk_near_blocks(spk_handle *h, int npairs, const int *m, const int *n, double *out);
/* Radiation pattern of edges (PreCalculate_RT): out = count * 4 * S complex,
 * [edge][k1 theta, k1 phi, k2 theta, k2 phi][sample]. */
int spk_get_patterns(const spk_handle *h, int first, int count, double *out);

/* Stage timing of the device matvec with HIP events on the handle's stream:
 * reps matvecs on device buffers; ms[i] = mean duration of stage i per
 * matvec and bytes[i] = its algorithmic HBM bytes, stages
 * 0 aggregation, 1 SS (all levels), 2 M2L (all levels), 3 RR (all levels),
 * 4 fold, 5 disaggregation, 6 near-field SpMV, 7 whole matvec. */
#define SPK_NSTAGES 8
int spk_profile_matvec(spk_handle *h, int reps, double *ms, double *bytes);
/* No counterpart in the reference: the streaming-read bandwidth of a
 * device (best of reps reads of a bytes-long buffer, GB/s), the practical
 * ceiling for the HBM-bound kernels. */
int spk_probe_read_bandwidth(int device, size_t bytes, int reps, double *gbps);
/* Same for the complex64 operator of an SPK_MIXED handle (bytes at 8 B per
 * complex). */
int spk_profile_matvec_c64(spk_handle *h, int reps, double *ms, double *bytes);
/* No counterpart in the reference: timing of the near-field preconditioner's
 * SpMV (bf16 blocks, complex64 vectors; SPK_PRECOND_NEAR_FIELD on an
 * SPK_MIXED handle) with HIP events on the handle's stream: *ms = mean over
 * reps launches, *bytes = its algorithmic HBM bytes (16 B per 2x2 block,
 * x read and y written at 8 B per complex). */
int spk_profile_near_bf(spk_handle *h, int reps, double *ms, double *bytes);
/* No counterpart in the reference: the same timing for the single-vector
 * preconditioner's fp8 copy (OCP E4M3FN, 8 B per 2x2 block plus 16 B of
 * per-row component scales per edge). */
int spk_profile_near_f8(spk_handle *h, int reps, double *ms, double *bytes);

/* Host-only setup probe (no GPU needed): RWG topology and octree of a mesh
 * exactly as spk_setup_tri builds them.  corner/element: N x 2 (1-based),
 * leaf: N Morton indices at l_max.  Call with NULL outputs to get *n_edges. */
int spk_topology_probe(int np, const double *pp, int nt, const int *tt, int normal_convention,
                       int tree_s_opt, int64_t *n_edges, int *corner, int *element,
                       int64_t *leaf, int *l_min, int *l_max, double *D, double *xmin3);

#ifdef __cplusplus
}
#endif
#endif
