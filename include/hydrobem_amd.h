/* hydrobem_amd: the low-order frequency-domain solve of HydroBEM on one AMD MI355X (gfx950).
 *
 * A C ABI (plain pointers and sizes, no exceptions, caller-owned buffers) that a Rust engine binds
 * through a thin FFI crate. Conventions follow the reference's C ABI
 * (crates/hydrobem-c/include/hydrobem.h; docs/reference/c-api.md:49-146):
 *   - every fallible call returns an HbgStatus with the reference's HbStatus values
 *     (hydrobem.h:42-98: 0 OK, 1 VALUE, 2 TYPE, 3 UNSUPPORTED, 4 STATE, 5 CANCELLED, 6 INTERNAL,
 *     7 ARGUMENT, 8 NUMERICAL, 9 IO, 10 EXISTS);
 *   - the message and location of the last failure are kept per thread
 *     (hb_last_error_message / hb_last_error_location, hydrobem.h);
 *   - out-pointers are checked before anything changes; inputs are read during the call only;
 *   - arrays are row-major C arrays unless stated; complex values are interleaved (re, im) doubles.
 *
 * What the functions replace: the engine seam solver::run::<R: Real> (crates/hydrobem/src/
 * solver.rs:265-345) with its statics (lo_static.rs:254-417) and its per-frequency
 * solve_frequency (lo_frequency.rs:1565-1635) for wetted bodies in infinite depth without
 * symmetry planes; the public seam hb_model_solve + hb_results_quantity (hydrobem.h:646, 1164)
 * for the quantities added_mass, radiation_damping, excitation_diffraction,
 * excitation_froude_krylov, excitation_scattering, excitation_haskind
 * (solve_driver.rs:336-600); hb_results_write_wamit (hydrobem.h:1297) for the .1/.2/.3 files.
 */
#ifndef HYDROBEM_AMD_H
#define HYDROBEM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t HbgStatus;
enum {
  HBG_OK = 0,
  HBG_ERR_VALUE = 1,
  HBG_ERR_TYPE = 2,
  HBG_ERR_UNSUPPORTED = 3,
  HBG_ERR_STATE = 4,
  HBG_ERR_CANCELLED = 5,
  HBG_ERR_INTERNAL = 6,
  HBG_ERR_ARGUMENT = 7,
  HBG_ERR_NUMERICAL = 8,
  HBG_ERR_IO = 9,
  HBG_ERR_EXISTS = 10
};

/* Stages of the progress callback (progress.rs:425-480 Stage::{Statics, Assembly, Lu, Outputs}). */
enum { HBG_STAGE_STATICS = 0, HBG_STAGE_ASSEMBLY = 1, HBG_STAGE_LU = 2, HBG_STAGE_OUTPUTS = 3 };

/* Progress callback (hydrobem.h:481-499 HbProgressFn): called on the caller's thread between
 * stages with the frequency index done of total; returning 0 cancels the solve
 * (HBG_ERR_CANCELLED), as the reference's callback does. */
typedef int32_t (*HbgProgressFn)(void* user, int32_t stage, int32_t done, int32_t total);

/* Precision of a solve. HBG_PRECISION_FP64: complex128 assembly and LU on the GPU (the
 * reference's double mode, env_hooks.rs:51-56). HBG_PRECISION_MIXED (no counterpart in the
 * reference): the FP64-assembled system factored in complex64 on the matrix cores (FP32
 * accumulation), the solution refined in FP64 against the FP64 matrix until the normwise backward
 * error |b - A x| / (|A| |x| + |b|) (infinity norms) is at most 1e-12; a frequency whose complex64
 * factorization needs row interchanges or does not converge in 10 steps is solved in FP64. */
enum { HBG_PRECISION_FP64 = 0, HBG_PRECISION_MIXED = 1, HBG_PRECISION_MIXED32 = 2 };
/* HBG_PRECISION_MIXED32 (the precision addition; no counterpart in the reference, whose
 * production mode, env_hooks.rs:51-56, assembles with FP32 Rankine tiers and factors in complex64
 * without refinement): the matrices assembled in FP32 (the wave kernel from FP32 tables of lower
 * degree, FP32 statics records), the self and adjacent pairs (centroids closer than twice the sum
 * of the panels' radii) in FP64 (statics, wave kernel and the log correction), the complex64
 * system factored on the matrix cores and the solution refined in FP64 against that matrix (its
 * complex64 entries with the FP64 near entries) to the backward error 10^-10 (HBG_TUNE_MIXED32_TOL);
 * a frequency that fails is solved in FP64 as in HBG_PRECISION_MIXED. Symmetry planes, walls, the
 * surrogate, and frequencies with pairs beyond the FP32 tables (K R or K |z + zeta| >= 48) take
 * HBG_PRECISION_MIXED for the whole solve. */

typedef struct HbgOptions {
  int32_t precision;     /* HBG_PRECISION_FP64 or HBG_PRECISION_MIXED */
  int32_t logarithmic;   /* analytic log-singularity integral (document_schema.rs:218, default 1) */
  double log_tolerance;  /* document_schema.rs:219, default 1e-4 */
  int32_t surrogate;     /* 1: finite-depth evanescent remainder from the learned surrogate (see
                            hbg_surrogate_train) where 0.1 <= R/h <= 20 and 0.05 <= nu h <= 20; 0 (default) */
  int32_t reserved;      /* 0 */
} HbgOptions;

typedef struct HbgSession HbgSession;

/* Thread-local message and location of the last failure ("" if none). */
const char* hbg_last_error_message(void);
const char* hbg_last_error_location(void);

/* Library version string (static storage). */
const char* hbg_version(void);

/* A session on GPU `device`: owns the device buffers of one geometry and its statics.
 * Replaces the model handle of hb_model_new / hb_model_free (hydrobem.h). */
HbgStatus hbg_session_create(int32_t device, HbgSession** out);
HbgStatus hbg_session_free(HbgSession* s);

/* Environment (problem.rs env; ρ, g, depth): depth +inf (infinite depth) or finite (the
 * finite-depth Green function of green_fin.rs; every panel centroid between -depth and 0). */
HbgStatus hbg_set_environment(HbgSession* s, double density, double gravity, double depth);

/* The wetted panels: vertices [n_panels][4][3] as the mesh gives them (a triangle repeats a
 * vertex), normals into the fluid by the vertex order; body [n_panels] (NULL: all body 0);
 * rotation_centres [n_bodies][3] (NULL: the origin). Each panel passes the reference's panel rule
 * (geometry.rs:120-199) or the call fails with HBG_ERR_VALUE naming the panel. Replaces the
 * geometry build of lo_geometry.rs:1100-1113 (body frame; the symmetry planes and walls of
 * hbg_set_symmetry and hbg_set_walls add the images). Invalidates the statics. */
HbgStatus hbg_set_panels(HbgSession* s, const double* vertices, int32_t n_panels, const int32_t* body,
                         int32_t n_bodies, const double* rotation_centres);

/* The artificial interior free surface of the bodies (irregular-frequency removal; the
 * SurfaceKind::Interior panels of lo_geometry.rs:598-615, interior.rs:519-570): panels in z = 0
 * inside the waterlines, vertices [n_panels][4][3] in the stored symmetry form of the body panels
 * (z within 1e-9 of 0, set to 0; normals taken upward), body [n_panels] (NULL: body 0); they follow
 * the body panels as extra unknowns: free term -4 pi, columns K S (lo_frequency.rs:85-97,
 * low_order.rs:771-785), no velocity, no forces. n_panels 0 removes them; hbg_set_panels removes
 * them too. Field points and Kochin functions leave them out of their sums (lo_frequency.rs:
 * 1311-1313, 1443-1458); at the frequency limits their unknowns are not used (identity rows,
 * lo_frequency.rs:99-104); walls (their images mirror the surface) and hbg_solve_group take them
 * too. The surrogate with them is HBG_ERR_UNSUPPORTED; HBG_PRECISION_MIXED32 solves them as
 * HBG_PRECISION_MIXED.
 * Invalidates the statics. */
HbgStatus hbg_set_interior(HbgSession* s, const double* vertices, int32_t n_panels, const int32_t* body);

/* Damping lids (the reference's lids, SurfaceKind::Lid with a damping factor; lo_frequency.rs:85-97,
 * 263-268, 1150-1156; low_order.rs:811-819; solve_common.rs:52-54): panels in z = 0 on the outer
 * free surface (a moonpool, the gap between two hulls), vertices [n_panels][4][3] as for
 * hbg_set_interior, body [n_panels] (NULL: body 0), damping_factor [n_panels] (eps >= 0, NULL: 0).
 * Extra unknowns with free term 4 pi and columns -(K - kappa) S, kappa = w^2 / (g + i w eps); the
 * diffraction boundary value (K - kappa) phi_I; the Haskind sum takes (K - kappa) phi_i phi_I over
 * them. eps = 0 is the free surface itself (kappa = K: the lid decouples, the body's results are
 * those without it). No velocity, no forces; Kochin functions leave them out (lo_frequency.rs:
 * 1279-1372 takes no lid); at the frequency limits their unknowns are not used (identity rows,
 * lo_frequency.rs:99-104). They replace the lids set before and keep the interior free surface
 * (and hbg_set_interior keeps the lids); n_panels 0 removes them, hbg_set_panels too. Field points
 * with lids and the surrogate with them are HBG_ERR_UNSUPPORTED; HBG_PRECISION_MIXED32 solves them
 * as HBG_PRECISION_MIXED. HBG_ERR_VALUE "lid panel i: a vertex off z = 0" / "damping factor must be
 * finite and >= 0". Invalidates the statics. */
HbgStatus hbg_set_lids(HbgSession* s, const double* vertices, int32_t n_panels, const int32_t* body,
                       const double* damping_factor);

/* Symmetry planes of the wetted surface (the mesh file's ISX, ISY; mesh.rs symmetry "yz", "xz",
 * "xz_and_yz"): sym_x 1 reflects the panels about x = 0, sym_y 1 about y = 0 (each 0 or 1). The
 * panels given to hbg_set_panels are then one sector; the solve adds their images (geometry.rs:
 * 209-234; each image panel belongs to its original's body, so every body must itself be symmetric)
 * and splits each frequency into G = 1, 2 or 4 systems of the sector's order by the characters of
 * the reflection group (lo_geometry.rs:25-35, lo_frequency.rs:860-943); the results are those of
 * the whole body. Kept across hbg_set_panels; set before or after it. With walls on the other
 * axis the group of both is used (only the characters even under the walls are solved); a plane
 * on the axis of a wall is HBG_ERR_VALUE. The frequency limits omega = 0 and infinity are solved the
 * same way (the limit matrices' image blocks combined into the characters; at omega = 0 in finite
 * depth the net-flux rule takes the whole body); hbg_solve_group takes such sessions too. With
 * G > 1, hbg_statics and hbg_assemble are HBG_ERR_UNSUPPORTED (the surrogate takes the image
 * blocks as the evaluator does). */
HbgStatus hbg_set_symmetry(HbgSession* s, int32_t sym_x, int32_t sym_y);

/* Walls (the domain x_wall, y_wall, orthogonal_walls of model_codes.rs:15-31): wall_x 1 a
 * reflecting vertical wall at x = 0, wall_y 1 at y = 0 (each 0 or 1). The panels of
 * hbg_set_panels are the real bodies, in the fluid (every vertex x >= 0, resp. y >= 0, or
 * HBG_ERR_VALUE); the solve adds their images in the walls, which move as the mirror image of the
 * bodies (lo_geometry.rs:947-983), solves the character even under every wall (lo_geometry.rs:
 * 1010-1030: one system of order N per frequency), takes the incident wave with its reflections
 * (low_order.rs:1470-1486) and returns the quantities of the real bodies. Kept across
 * hbg_set_panels. With a symmetry plane on the other axis the panels are the body's half and the
 * group of both is used (lo_geometry.rs:863-945); a wall on the axis of a symmetry plane is
 * HBG_ERR_VALUE. The frequency limits omega = 0 and infinity are solved with the walls too (at
 * omega = 0 in finite depth the net-flux rule takes the real bodies); hbg_solve_group takes such
 * sessions too. With walls hbg_statics and hbg_assemble are HBG_ERR_UNSUPPORTED. */
HbgStatus hbg_set_walls(HbgSession* s, int32_t wall_x, int32_t wall_y);

/* Field points (lo_frequency.rs:1392-1495, `Output` frad/finc/fsca/ftot, low_order.rs:1490-1500):
 * points [n][3] on or below z = 0 (in the fluid, on the free surface, or inside a body), n = 0
 * clears them. Every later hbg_solve (infinite or finite depth, every point above the sea floor
 * or HBG_ERR_VALUE; with symmetry planes or walls too; no frequency limits: else
 * HBG_ERR_UNSUPPORTED)
 * also evaluates, per frequency and point, the radiation
 * elevation of each rigid mode frad = (w^2 / g) phi_j(x) (per unit amplitude of motion) and the
 * incident, scattered and total elevations eta = -(i w / g) phi of each heading. */
HbgStatus hbg_set_field_points(HbgSession* s, const double* points, int32_t n);

/* The field values of the last hbg_solve with field points: frad [F][n][6 nbodies], finc, fsca,
 * ftot [F][n][H], complex (re, im) doubles; any output may be NULL. The sizes are those of that
 * solve (hbg_field_shape); the caller's buffers must hold them. HBG_ERR_STATE before such a solve
 * and after a new geometry or new field points. */
HbgStatus hbg_field_values(HbgSession* s, double* frad, double* finc, double* fsca, double* ftot);

/* The sizes of hbg_field_values's arrays (no counterpart in the reference, whose results carry
 * their shapes): F, n, 6 nbodies and H of the last solve with field points; HBG_ERR_STATE as
 * hbg_field_values. */
HbgStatus hbg_field_shape(HbgSession* s, int32_t* n_omegas, int32_t* n_points, int32_t* n_rad, int32_t* n_headings);

/* Kochin functions (lo_frequency.rs:1279-1372, `Output` kr/krt/kd/kdt): n_angles angles
 * theta_t = 2 pi t / n_angles (0: none). Every later hbg_solve (infinite or finite depth, with
 * symmetry planes or walls too; no frequency limits: else HBG_ERR_UNSUPPORTED) evaluates per
 * frequency H(theta) = sum_j area_j (phi_j dPhi/dn - Phi u_j), Phi = Z(z) e^{i k (x cos + y sin)}
 * the far-field kernel (low_order.rs:1417-1433), for every radiation column and, with the total
 * potential and no source term, every heading, and the theta derivatives. */
HbgStatus hbg_set_kochin(HbgSession* s, int32_t n_angles);

/* The Kochin functions of the last hbg_solve with angles: kr, krt [F][n_angles][6 nbodies], kd,
 * kdt [F][n_angles][H], complex (re, im) doubles; any output may be NULL. The sizes are those of
 * that solve (hbg_kochin_shape). HBG_ERR_STATE before such a solve and after a new geometry or a
 * new angle count. */
HbgStatus hbg_kochin_values(HbgSession* s, double* kr, double* krt, double* kd, double* kdt);

/* The sizes of hbg_kochin_values's arrays: F, n_angles, 6 nbodies and H of the last solve with
 * angles (no counterpart in the reference); HBG_ERR_STATE as hbg_kochin_values. */
HbgStatus hbg_kochin_shape(HbgSession* s, int32_t* n_omegas, int32_t* n_angles, int32_t* n_rad, int32_t* n_headings);

/* The solve of n_omegas frequencies (rad/s) and n_headings headings (degrees), radiation of the
 * six rigid modes of every body plus diffraction (solver::run, lo_frequency.rs:1565-1635).
 * D = 6 n_bodies. Outputs (any may be NULL):
 *   added_mass, damping         [F][D][D]                      (A_ij, B_ij)
 *   excitation, froude_krylov,
 *   scattering                  [F][H][D] complex              (X_i = FK + SC)
 *   haskind                     [F][D][H] complex              (Haskind exciting force, Output.HK)
 * An omega of 0 or +infinity requests that radiation limit (lo_static.rs:84-96,
 * low_order.rs:635-637, 757-758, green_fin.rs FreqMode::Zero / Infinite: in infinite depth the
 * statics alone, with 1/r - 1/r' at infinity; in finite depth the statics with their six images
 * and the limit's regular part), solved in FP64 whatever the precision option: its added mass row
 * is written and its damping and diffraction rows are zero (the reference's
 * added_mass_zero_frequency / added_mass_infinite_frequency). At omega = 0 in finite depth the
 * added mass of a mode with a net volume flux through the body (heave of a surface-piercing body)
 * diverges and is not defined: its row and column are NaN (the reference refuses such a mode,
 * rules/coordinates.rs:15-75, at the same relative flux 1e-6).
 * One system holds at most 32 right-hand sides (D radiation columns and H headings): with more,
 * the headings are solved in groups of 32 - D, each with the radiation columns (the matrices
 * assembled and factored once per group), and the outputs are the same as for one solve; not with
 * field points or Kochin functions, and D > 31 (more than five bodies), HBG_ERR_UNSUPPORTED.
 * HBG_ERR_NUMERICAL on a zero pivot ("singular system matrix (zero pivot k)") or a non-finite
 * matrix or solution (linalg.rs:221-252); HBG_ERR_CANCELLED when the callback returns 0. */
HbgStatus hbg_solve(HbgSession* s, const double* omegas, int32_t n_omegas, const double* headings_deg,
                    int32_t n_headings, const HbgOptions* options, double* added_mass, double* damping,
                    double* excitation, double* froude_krylov, double* scattering, double* haskind,
                    HbgProgressFn progress, void* user);

/* The solve of several geometries, one session each (all on one device; uneven panel counts and
 * body counts allowed), at the same frequencies and headings, with hbg_solve's options: the
 * (geometry, frequency) systems share every launch of the batched LU. No counterpart in the
 * reference, which solves each model on its own (the multi-body case of a single geometry is
 * hbg_solve's). The systems are taken in decreasing order of N, in batches; a system of N panels
 * goes to the LU at the largest order of its batch, with the identity on the rows and columns
 * past N (its LU is block-diagonal, its solution unchanged) and the GEMM tiles of that identity
 * block skipped. A session with symmetry planes or walls brings its character systems (G or one
 * per frequency, of the sector's order), kept in one batch, and its results are those of its own
 * hbg_solve, with an interior free surface too; ordinary frequencies only (else
 * HBG_ERR_UNSUPPORTED); HBG_PRECISION_MIXED32 runs as HBG_PRECISION_MIXED.
 * Outputs: arrays of count pointers (arrays or pointers may be NULL), entry i laid out as
 * hbg_solve's for session i. All work runs on the first session's stream; no session may be in
 * use elsewhere during the call. */
HbgStatus hbg_solve_group(HbgSession* const* sessions, int32_t count, const double* omegas, int32_t n_omegas,
                          const double* headings_deg, int32_t n_headings, const HbgOptions* options,
                          double* const* added_mass, double* const* damping, double* const* excitation,
                          double* const* froude_krylov, double* const* scattering, double* const* haskind);

/* The motions of the bodies (RAOs) from a solve's outputs (motion.rs:121-158 solve_motions):
 * per frequency (C - w^2 (M + A) + i w (B + B_ext)) xi = X on the n_dof = 6 n_bodies modes, complex,
 * partial pivoting. mass, stiffness, external_damping (NULL: none): [D][D] row-major, about the
 * rotation centres (the reference's motion_matrices, motion.rs:85-115, are the caller's);
 * added_mass, damping [F][D][D] and excitation [F][H][D] complex as hbg_solve gives them (a body
 * whose excitation is the Haskind force passes those columns instead). rao [F][H][D] complex.
 * D <= 64 (more headings than 32 are taken in groups of 32). HBG_ERR_NUMERICAL "motion equation at omega = w: singular system matrix
 * (zero pivot k)". */
HbgStatus hbg_motions(HbgSession* s, int32_t n_omegas, const double* omegas, int32_t n_headings, int32_t n_dof,
                      const double* mass, const double* stiffness, const double* external_damping,
                      const double* added_mass, const double* damping, const double* excitation, double* rao);

/* ---- kernel-level entry points, at the granularity of the reference's component tests
 * (crates/hydrobem/tests/core.rs, lo_a.rs; linalg.rs:422-489): */

/* The infinite-depth wave kernel at n points (green_inf.rs:316-368 wave_kernel):
 * out [n][7] = Ft, Ft_X, Ft_Y, Gi, Gi_X, Gi_Y, exp(-Y) J0(X), for X = K R >= 0, Y = K|z+zeta| >= 0. */
HbgStatus hbg_wave_kernel(HbgSession* s, const double* X, const double* Y, int64_t n, double* out);

/* The finite-depth kernel at n point pairs (green_fin.rs:1007-1095 finite_kernel, ordinary
 * frequency K = w^2/g, depth h): horizontal distance R[i], field-point depth z[i], source depth
 * zeta[i]; out [n][8] = Wr, dWr/dR, dWr/dzeta, dWr/dz, Wi, dWi/dR, dWi/dzeta, dWi/dz of the regular
 * part Wr = Re G - sum_p 1/r_p, Wi = Im G. K = 0 and K = +infinity: the limits omega = 0 and
 * omega = infinity (FreqMode::Zero / Infinite, Rankine signs + and (+, +, -, -, -, -), Wi = 0). */
HbgStatus hbg_finite_kernel(HbgSession* s, double h, double K, int32_t n, const double* R, const double* z,
                            const double* zeta, double* out);

/* The FP32 evaluators of HBG_PRECISION_MIXED32 at points, for the parity tests against
 * hbg_wave_kernel and hbg_finite_kernel (no counterpart in the reference): infinite depth
 * out [n][4] = Ft, Ft_X, Gi, Gi_X at 0 <= X < 48, 0 < Y < 48; finite depth out [n][8] as
 * hbg_finite_kernel, at points inside the FP32 tables (else HBG_ERR_VALUE). */
HbgStatus hbg_wave_kernel32(HbgSession* s, const double* X, const double* Y, int32_t n, double* out);
HbgStatus hbg_finite_kernel32(HbgSession* s, double h, double K, int32_t n, const double* R, const double* z,
                              const double* zeta, double* out);

/* Dispersion roots (dispersion.rs:19-95): out[0] = x0 (x0 tanh x0 = H), out[i] = x_i
 * (x_i tan x_i = -H, (i - 1/2) pi < x_i < i pi) for i = 1..n. */
HbgStatus hbg_dispersion_roots(double H, int32_t n, double* out);

/* The statics of the current panels (lo_static.rs:59-119): out [N][N][4] = rs, rd, dl, dd for
 * field point k (row) and panel j (column), logarithmic / log_tolerance as in HbgOptions. */
HbgStatus hbg_statics(HbgSession* s, int32_t logarithmic, double log_tolerance, double* out);

/* The influence matrices of one frequency (lo_frequency.rs:399-537, G = 1): S and A = 2πI - D,
 * complex [N][N] row-major (S[k][n] = influence of panel n at field point k). With the precision
 * HBG_PRECISION_MIXED32, the matrices of the FP32 assembly as its solve uses them (the complex64
 * entries, the FP64 values at the near entries). */
HbgStatus hbg_assemble(HbgSession* s, double omega, const HbgOptions* options, double* S, double* A);

/* The number fields of the WAMIT files (wamit_files.rs:1718-1757 fortran_e, fortran_i): x as
 * Fortran E<width>.<digits> (one_p 0: 0.ddddddE+xx, 1: 1P form d.ddddddE+xx), n as I<width>,
 * right-aligned, into out (NUL-terminated, capacity bytes). HBG_ERR_VALUE when the value does not
 * fit the field (a non-finite x, a three-digit exponent, too many digits). Host only. */
HbgStatus hbg_fortran_e(double x, int32_t one_p, int32_t width, int32_t digits, char* out, int32_t capacity);
HbgStatus hbg_fortran_i(int64_t n, int32_t width, char* out, int32_t capacity);

/* A x = b on the GPU in FP64 (linalg.rs:321-328 lu_solve_c128): a [n][n] and b [nrhs][n]
 * complex, column-major (a[j][i] = A(i, j)), overwritten by the factors and the solution. */
HbgStatus hbg_lu_solve(HbgSession* s, int32_t n, double* a, int32_t nrhs, double* b);

/* hbg_lu_solve without any device allocation, for callers that cannot allocate (serving loops,
 * harnesses that forbid hipMalloc; no counterpart in the reference). The staged [A | b] and the
 * LU's workspace live in a fixed per-GPU array of the library (32 MiB, held by each GPU the
 * library runs on until the process ends): n (n + nrhs) 16 bytes plus the workspace of order n
 * must fit (n <= 1300 with 32 right-hand sides), else HBG_ERR_VALUE. Calls no hipMalloc, hipFree
 * or allocating library routine; the launches are eager on the session's stream. Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU race. Results and errors as
 * hbg_lu_solve. */
HbgStatus hbg_lu_solve_noalloc(HbgSession* s, int32_t n, double* a, int32_t nrhs, double* b);

/* Frequencies in flight in hbg_solve (no counterpart in the reference, whose frequency pools run
 * one CPU solve per worker, parallel.rs:63-89): each holds its N x (N + R) system on the device
 * and their LUs advance panel by panel in the same launches. 0 (the default): as many as half the
 * free device memory holds, at most 48, balanced over the batches of a solve. */
HbgStatus hbg_set_batch(HbgSession* s, int32_t max_batch);

/* ---- the learned surrogate (no counterpart in the reference; it stands beside the finite-depth
 * evaluator of green_fin.rs:1007-1095, which remains the fallback and judges it). The network
 * (4 -> 128 -> 128 -> 128 -> 4, SiLU; BF16 weights with FP32 accumulation on the matrix cores)
 * maps (R/h, z/h, zeta/h, nu h) to the evanescent remainder T of the real kernel, Wr = T - A0 E
 * Y0(k0 R), as (h T, h^2 T_R, h^2 T_zeta, h^2 T_z); the propagating mode and Im G stay analytic. */
typedef struct HbgSurrogateTraining {
  uint64_t seed;
  int32_t n_train;         /* samples drawn, in groups of 8192 with one nu h each */
  int32_t n_holdout;       /* held-out samples for the error report */
  int32_t epochs;
  int32_t batch;           /* minibatch (>= 64, <= n_train) */
  double learning_rate;    /* Adam, decayed by cosine to 1 % */
} HbgSurrogateTraining;
#define HBG_SURROGATE_REPORT 12
/* Parameters of the network (FP32, output-major weights W1 [128][4], b1, W2 [128][128], b2, W3, b3,
 * W4 [4][128], b4). */
int32_t hbg_surrogate_size(void);
/* Train on the GPU (sampling, the FP64 evaluator's targets, forward and backward passes, Adam) and
 * keep the network in the session. report [9]: final training loss, held-out RMS error (network
 * output units), held-out max |error| of the four outputs, data seconds, training seconds, steps,
 * non-finite targets, non-finite parameters after the training, 0. */
HbgStatus hbg_surrogate_train(HbgSession* s, const HbgSurrogateTraining* training, double* report);
HbgStatus hbg_surrogate_set(HbgSession* s, const float* params, int32_t n);
HbgStatus hbg_surrogate_get(HbgSession* s, float* params, int32_t n);
/* The network at points in [n][4] = (R/h, z/h, zeta/h, nu h): out [n][4] = (h T, h^2 T_R, h^2 T_zeta,
 * h^2 T_z) through the BF16 inference kernel; T_zeta is the derivative in the source depth, T_z in
 * the field depth (z the field point's depth, zeta the source's). */
HbgStatus hbg_surrogate_eval(HbgSession* s, int32_t n, const double* in, double* out);
/* The same four values from the FP64 evaluator (the network's targets and judge). */
HbgStatus hbg_finite_remainder(HbgSession* s, double h, double K, int32_t n, const double* R, const double* z,
                               const double* zeta, double* out);
/* Loss (mean squared error over batch x 4) and its gradient for the parameters, by the training's
 * forward and backward kernels: inputs [batch][4] (the network's scaled inputs), targets [batch][4]. */
HbgStatus hbg_surrogate_gradient(HbgSession* s, const float* params, int32_t batch, const float* inputs,
                                 const float* targets, float* gradient, double* loss);

/* Kernel variants (no counterpart in the reference), per process: HBG_TUNE_LU_GEMM selects the
 * tiles of the LU's complex GEMM on the matrix cores, 1: 64 x 64 (default), 2: 128 x 64, and for
 * the complex64 LU of the mixed precision also 3: 64 x 64, 4: 128 x 64, 5: 128 x 128 (8 waves),
 * 6: 128 x 128 (4 waves of 64 x 64, operands loaded straight into LDS) for the trailing updates
 * and 64 x 64 for the panels' strips, 7: as 6 with two LDS buffers (the next chunk of k loading
 * while the current one is multiplied), on the FP16 matrix cores with each FP32 operand split
 * into two FP16 numbers (three FP16 products per FP32 product; the complex128 LU takes 1 for 3, 6
 * and 7 and 2 for 4 and 5); HBG_TUNE_LU_DEPTH the number of 64-column panels per
 * trailing update (1..64, default 4). Both set the complex128 and the complex64 LU and read back
 * the complex128 one; HBG_TUNE_LU_GEMM_MIXED and HBG_TUNE_LU_DEPTH_MIXED set and read the
 * complex64 LU's (the mixed precision) alone. HBG_TUNE_LU_GRAPHS 1 (default): the batched LU's
 * launches are captured once per system shape and buffer into a hipGraph and replayed (not while
 * stage profiling is on), 0: launched one by one. */
enum {
  HBG_TUNE_LU_GEMM = 0,
  HBG_TUNE_LU_DEPTH = 1,
  HBG_TUNE_LU_GRAPHS = 2,
  HBG_TUNE_LU_GEMM_MIXED = 3,
  HBG_TUNE_LU_DEPTH_MIXED = 4,
  HBG_TUNE_MIXED32_TOL = 5, /* HBG_PRECISION_MIXED32 refines to backward error 10^-value (6..14, default 10) */
  HBG_TUNE_NN_RMIN = 6, /* per session: the surrogate takes the pairs with value / 100 <= R/h <= 20, the
                           table/series evaluator the others (10..2000, default 10: the training domain) */
  HBG_TUNE_LU_SUB = 7,      /* panels per sub-block of an LU block (0: none, default): the in-block updates
                               take the sub-block's panels, its end an L-shaped update (both arithmetics) */
  HBG_TUNE_LU_SUB_MIXED = 8, /* the same for the complex64 LU alone */
  HBG_TUNE_LU_SUB2 = 9,       /* panels per sub-block of a sub-block (0: none, default), both arithmetics */
  HBG_TUNE_LU_SUB2_MIXED = 10 /* the same for the complex64 LU alone */
};
HbgStatus hbg_set_tuning(HbgSession* s, int32_t key, int32_t value);
/* The project's autotuner (no counterpart in the reference): times the batched LU on synthetic
 * systems of order n with nrhs right-hand sides in the given precision, first for every GEMM tile
 * variant (1, 2; mixed precision also 3 to 7) and block depth (2, 4, 8, 12, 16) at the largest batch
 * of 4, 8, 16, 32, 48 that fits in a third of the free memory, then deeper blocks with nested
 * sub-blocks ((depth, sub, sub of sub) = (16, 8, 0), (24, 8, 0), (24, 12, 0), (32, 16, 0),
 * (32, 16, 8), (64, 32, 16)) at the fastest variant, then the smaller batches at the fastest of
 * all; keeps the fastest variant, depth and sub-blocks (HBG_TUNE_LU_GEMM, HBG_TUNE_LU_DEPTH,
 * HBG_TUNE_LU_SUB, HBG_TUNE_LU_SUB2; for the mixed precision their _MIXED keys) and sets the
 * session's batch (the fastest per system). report: up to max_rows rows of (variant, depth,
 * sub-block, its sub-block, batch, ms per system), rows the number written. */
HbgStatus hbg_autotune(HbgSession* s, int32_t n, int32_t nrhs, int32_t precision, double* report, int32_t max_rows,
                       int32_t* rows);
HbgStatus hbg_get_tuning(HbgSession* s, int32_t key, int32_t* value);

/* Stage timing (the reference's HB_PROFILE phase marks, low_order.rs:184-208), by HIP events on the
 * session's stream: enable (clears the counters) or disable. hbg_profile_read copies up to n of the
 * HBG_PROFILE_COUNTERS counters, accumulated over the solves since enabling:
 *   0 statics ms, 1 assembly ms (boundary values + matrices), 2 right-hand sides ms, 3 LU ms
 *   (factorization + substitution), 4 LU trailing-update GEMM ms, 5 their algorithmic flops
 *   (8 m n k each), 6 GEMM launches, 7 forces ms, 8 frequencies, 9 LU algorithmic flops
 *   (8 N^3 / 3 each), 10 assembly algorithmic bytes (N^2 (32 + 2 x 16) each), 11 statics runs,
 *   12 LU panels that needed row interchanges (the pivoting fallback), 13 refinement steps of the
 *   mixed-precision solves, 14 their largest final backward error, 15 frequencies of a
 *   mixed-precision solve solved in FP64 instead, 16 the mixed-precision refinement ms (FP64
 *   residuals, corrections with the complex64 factors, the FP64 solution stored), 17 the largest
 *   backward error of a mixed-precision solve's first solution (that of the complex64 LU, before
 *   any refinement step), 18-20 ms, algorithmic flops and launches of the GEMM launches on the
 *   128 x 128 tiles of tuning variants 6 and 7 (part of 4-6). Counter 10 counts N^2 (16 + 3 x 8)
 *   bytes per system of HBG_PRECISION_MIXED32 (FP32 statics read, S, A and A's copy written). */
#define HBG_PROFILE_COUNTERS 21
HbgStatus hbg_set_profiling(HbgSession* s, int32_t enable);
HbgStatus hbg_profile_read(HbgSession* s, double* out, int32_t n);

/* WAMIT .1 (added mass and damping), .2 (Haskind exciting force) and .3 (pressure-integrated
 * exciting force) files of a solve (wamit_files.rs:794-867, docs/reference/file-formats.md):
 * stem + ".1" etc., lengths normalised by reference_length; overwrite 0 fails with HBG_ERR_EXISTS
 * if a file exists. Written all or none. */
HbgStatus hbg_write_wamit(const char* stem, int32_t overwrite, double density, double gravity,
                          double reference_length, const double* omegas, int32_t n_omegas,
                          const double* headings_deg, int32_t n_headings, int32_t n_bodies,
                          const double* added_mass, const double* damping, const double* haskind,
                          const double* excitation, const char* stamp);
/* The same and, for the arrays given (NULL: the file is not written), .2fk (the Froude-Krylov force
 * by solved column), .2sc (Haskind - Froude-Krylov), .3fk, .3sc (the Froude-Krylov and scattering
 * parts of the pressure force) and .4 (the motions, hbg_motions's rao, xi L^ri)
 * (wamit_files.rs:836-866): froude_krylov, scattering, rao [F][H][D] complex as hbg_solve and
 * hbg_motions give them. The radiation limits (omega 0, +infinity) have rows in .1 only. */
HbgStatus hbg_write_wamit_files(const char* stem, int32_t overwrite, double density, double gravity,
                                double reference_length, const double* omegas, int32_t n_omegas,
                                const double* headings_deg, int32_t n_headings, int32_t n_bodies,
                                const double* added_mass, const double* damping, const double* haskind,
                                const double* excitation, const double* froude_krylov, const double* scattering,
                                const double* rao, const char* stamp);
/* The same and, with n_points field points (points [P][3] global, as hbg_set_field_points), .fpt
 * (point, x, y, z) and .6p (wamit_files.rs:938-984; file-formats.md section 2: per ordinary
 * frequency the radiation block, the points in order with (Re, Im) of eta / L^r_k for every
 * column, then the diffraction block, per heading the points with Re, Im of eta / A):
 * field_radiation [F][P][D] and field_diffraction [F][P][H] complex, hbg_field_values's frad and
 * ftot (either NULL: its block is left out; both NULL: no .6p). */
HbgStatus hbg_write_wamit_all(const char* stem, int32_t overwrite, double density, double gravity,
                              double reference_length, const double* omegas, int32_t n_omegas,
                              const double* headings_deg, int32_t n_headings, int32_t n_bodies,
                              const double* added_mass, const double* damping, const double* haskind,
                              const double* excitation, const double* froude_krylov, const double* scattering,
                              const double* rao, int32_t n_points, const double* points,
                              const double* field_radiation, const double* field_diffraction, const char* stamp);

#ifdef __cplusplus
}
#endif
#endif
