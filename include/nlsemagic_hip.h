/*
 * nlsemagic_hip.h -- C ABI of the MI355X NLSE RK4 integrators
 * (library: nlsemagic_amd/libnlsemagic_hip.so).
 *
 * Drop-in entry points
 * --------------------
 * NLSE{1,2,3}D_TAKE_STEPS_HIP[_F] take the arguments of the reference C ABI
 * (src/NLSEmagic1D/NLSEmagic.h:1-35) in the same order and with the same
 * meaning, and replace its CUDA entry points:
 *   U     = split planes Uoldr/Uoldi, MATLAB column-major:
 *           1D U[i] (i < N); 2D U[i*M + j] (N columns x M rows, M contiguous);
 *           3D U[(i*N + j)*M + k] (an M x N x L MATLAB array)
 *   V     = real potential, same layout
 *   s, a  = nonlinearity and Laplacian coefficients of
 *           i U_t + a Lap(U) + (s|U|^2 - V) U = 0  (the reference's code
 *           convention: NLSEmagic1D_CPU.c:27-30)
 *   h2    = h^2 (the grid step squared, not h)
 *   BC    = 1 Dirichlet, 2 MSD, 3 Lap0, 4 one-sided (1D only); any other
 *           value gives a zero boundary, as in the reference
 *   chunk_size = number of RK4 steps, k = time step
 *   method = 1 CD (2nd order), 2 2SHOC (4th order)
 * The result is written to Unewr/Unewi.  Differences from the reference:
 *   - the inputs are not modified (the reference CPU integrators overwrite U);
 *   - arguments are validated; the functions return 0 on success or a
 *     nonzero NLSE_E* code, and nlse_last_error() describes the failure
 *     (the reference returns void and checks nothing);
 *   - the double-precision results are bit-identical to the reference CPU
 *     integrators (NLSEmagic{1,2}D_CPU.c, NLSE3D_TAKE_STEPS_{CD,2SHOC}.c).
 * Device buffers are cached between calls (the four most recently used
 * shapes), so a caller looping over chunks (as the reference MATLAB drivers
 * do), or alternating between a few problems, does not pay an allocation per
 * call.  Calls are synchronous and thread-safe (serialised).
 *
 * Resident-state API
 * ------------------
 * nlse_create .. nlse_destroy keep U, V and the RK4 work arrays on the GPU
 * across chunks (the reference re-uploads the state on every call).
 */
#ifndef NLSEMAGIC_HIP_H
#define NLSEMAGIC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    NLSE_OK = 0,
    NLSE_EINVAL = 1, /* bad argument */
    NLSE_EHIP = 2,   /* HIP runtime error (no device, launch failure, ...) */
    NLSE_ENOMEM = 3  /* device allocation failed */
};

/* Replaces NLSE1D_TAKE_STEPS_CUDA (src/NLSEmagic1D/NLSEmagic.h:1-2;
 * defined at src/NLSEmagic1D/NLSEmagic1D_GPU.cu:663). */
int NLSE1D_TAKE_STEPS_HIP(const double *Uoldr, const double *Uoldi, const double *V,
                          double *Unewr, double *Unewi, double s, double a, double h2,
                          int BC, int chunk_size, double k, int N, int method);
/* Replaces NLSE1D_TAKE_STEPS_CUDA_F (NLSEmagic.h:4-5; NLSEmagic1D_GPU.cu:754). */
int NLSE1D_TAKE_STEPS_HIP_F(const float *Uoldr, const float *Uoldi, const float *V,
                            float *Unewr, float *Unewi, float s, float a, float h2,
                            int BC, int chunk_size, float k, int N, int method);
/* Replaces NLSE2D_TAKE_STEPS_CUDA (NLSEmagic.h:13-14; declared but never
 * defined in the reference: src/NLSEmagic2D/NLSEmagic2D_GPU.cu is a copy of
 * the 1D file).  N = columns (x), M = rows (contiguous). */
int NLSE2D_TAKE_STEPS_HIP(const double *Uoldr, const double *Uoldi, const double *V,
                          double *Unewr, double *Unewi, double s, double a, double h2,
                          int BC, int chunk_size, double k, int N, int M, int method);
/* Replaces NLSE2D_TAKE_STEPS_CUDA_F (NLSEmagic.h:16-17). */
int NLSE2D_TAKE_STEPS_HIP_F(const float *Uoldr, const float *Uoldi, const float *V,
                            float *Unewr, float *Unewi, float s, float a, float h2,
                            int BC, int chunk_size, float k, int N, int M, int method);
/* Replaces NLSE3D_TAKE_STEPS_CUDA (NLSEmagic.h:25-26; the reference ships
 * the 3D GPU integrators only as MEX files,
 * src/NLSEmagic3D/NLSE3D_TAKE_STEPS_{CD,2SHOC}_CUDA_D.cu). */
int NLSE3D_TAKE_STEPS_HIP(const double *Uoldr, const double *Uoldi, const double *V,
                          double *Unewr, double *Unewi, double s, double a, double h2,
                          int BC, int chunk_size, double k, int N, int M, int L,
                          int method);
/* Replaces NLSE3D_TAKE_STEPS_CUDA_F (NLSEmagic.h:28-29). */
int NLSE3D_TAKE_STEPS_HIP_F(const float *Uoldr, const float *Uoldi, const float *V,
                            float *Unewr, float *Unewi, float s, float a, float h2,
                            int BC, int chunk_size, float k, int N, int M, int L,
                            int method);
/* Replaces the rmc_vorticity3D MEX post-processor
 * (src/NLSEmagic3D/rmc_vorticity3D.c:24-354; [wx,wy,wz,W] =
 * rmc_vorticity3D(U, 1/(2h), veps)): host planes Ur, Ui (Ui may be NULL for
 * a real U) of M x N x L doubles in MATLAB order (M fastest), outputs of the
 * same size.  Edge and corner handling: see nlse_vorticity. */
int NLSE3D_VORTICITY_HIP(const double *Ur, const double *Ui, double lh2, double veps, int M, int N, int L,
                         double *wx, double *wy, double *wz, double *W);

/* ---- resident-state API ---- */
typedef struct nlse_solver nlse_solver;

/* precision: 2 = double, 1 = single (the reference's `precision` codes,
 * src/NLSEmagic1D/NLSEmagic1D.m:22-24), 3 = single-precision stages with the
 * state kept as a compensated hi + lo pair of floats (final RK4 combine by
 * TwoSum; no counterpart in the reference).  dims: 1D (N), 2D (N, M), 3D (N, M, L)
 * with unused sizes passed as 1. */
int nlse_create(nlse_solver **out, int dim, int N, int M, int L, int method, int precision,
                int BC, double s, double a, double h2, double k);
void nlse_destroy(nlse_solver *h);
/* host planes -> device state; V may be NULL to keep the previous potential */
int nlse_set_state(nlse_solver *h, const double *Ur, const double *Ui, const double *V);
int nlse_set_state_f(nlse_solver *h, const float *Ur, const float *Ui, const float *V);
/* device state -> host planes (synchronises the solver's stream) */
int nlse_get_state(nlse_solver *h, double *Ur, double *Ui);
int nlse_get_state_f(nlse_solver *h, float *Ur, float *Ui);
/* enqueue `steps` RK4 steps on the solver's stream (asynchronous) */
int nlse_step(nlse_solver *h, int steps);
/* enqueue `steps` steps between two HIP events on the solver's stream,
 * wait, and return the elapsed device time in milliseconds */
int nlse_step_timed(nlse_solver *h, int steps, double *ms);
/* waits like nlse_sync, the calling thread asleep rather than spinning (for
 * hosts that wait on many solvers from several threads) */
int nlse_wait(nlse_solver *h);
int nlse_sync(nlse_solver *h);
/* non-blocking: 1 when all work enqueued on the solver's stream is done,
 * 0 while it runs, a negative NLSE_E* code on error */
int nlse_query(nlse_solver *h);
/* diagnostics of the resident state, reduced on the GPU in double precision:
 * out3[0] = sum |U|^2 (mass / h^d), out3[1] = max Re U, out3[2] = number of
 * non-finite points (the reference drivers' blow-up test) */
int nlse_diagnostics(nlse_solver *h, double *out3);
/* discrete energy of the resident state, reduced on the GPU in double
 * precision (no counterpart in the reference, whose drivers report the mass
 * only: NLSEmagic1D_FRS.m:582-586).  With h^d the cell volume and nearest-
 * neighbour pairs (p, q) along every axis, boundary points included:
 *   out4[1] = h^d (a/h^2) sum_pairs |U_p - U_q|^2
 *   out4[2] = h^d sum_p (V_p |U_p|^2 - (s/2) |U_p|^4)
 *   out4[0] = out4[1] + out4[2]   (the energy E)
 *   out4[3] = h^d sum_p |U_p|^2   (the mass)
 * E is the Hamiltonian of the semi-discrete CD system (F = i[a Lap U +
 * (s|U|^2 - V) U] is i dE/d conj(U) / h^d): with Dirichlet boundaries it is
 * conserved up to the RK4 error.  Allocates nothing (the solver's workspace). */
int nlse_energy(nlse_solver *h, double *out4);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the reduction of
 * nlse_diagnostics over a state the caller holds on the GPU.
 *   U     n interleaved complex values on the current GPU (double for
 *         precision 2, float for 1 and 3)
 *   Ulo   low part of a compensated (precision 3) state, same layout, or NULL
 *   out3  device memory for 3 doubles: sum |U|^2, max Re U, non-finite count
 *   stream  a hipStream_t (NULL: the null stream); the call only enqueues
 * Allocates nothing: the block partials (3 x 1024 doubles) live in a device
 * array of the library, one copy per GPU, located on the first call on that
 * GPU.  Calls that overlap on one GPU share that array and race: serialise
 * them (one stream, or events).  n: 1 .. 2^31-1, else NLSE_EINVAL. */
int nlse_diagnostics_noalloc(const void *U, const void *Ulo, long long n, int precision, double *out3,
                             void *stream);
/* The reference drivers' full RK4 stability bound of the current state when
 * they pick k themselves (k = 0): NLSEmagic1D_FRS.m:233-275, NLSE2D.m:427-477,
 * NLSE3D.m:443-495.  With x = (h^2/a)(s|U|^2 - V) at every point and the
 * scheme's eigenvalue offsets G,
 *   kfull = (h^2/a) sqrt(8) / max(|Bmax|, max over G and points |x - G|),
 * Bmax = 0 (Dirichlet), the real part of the largest-magnitude
 * (h^2/(i a)) F/U on the first interior layer (MSD; F = the RHS, computed by
 * one stage-1 launch into the solver's stage buffers), or max x on the
 * boundary (Lap0).  out3 = {kfull, klin, k} with klin the linear bound and
 * k = 0.9 min(kfull, klin) in 1D and 2D, 0.6 min(kfull, klin) in 3D. */
int nlse_kfull(nlse_solver *h, double *out3);
/* change the time step k of a solver (e.g. to nlse_kfull's k); the other
 * parameters stay; no counterpart in the reference, whose drivers pass k on
 * every chunk call */
int nlse_set_step(nlse_solver *h, double k);
/* 3D vorticity of the resident state (replaces the rmc_vorticity3D MEX file,
 * src/NLSEmagic3D/rmc_vorticity3D.c:24-354, as called by NLSE3D.m:594):
 * wx, wy, wz = curl of the velocity Im(conj(U) grad U) / (2|U|^2 + veps) and
 * W = |w|, with lh2 = 1/(2h) and the reference's difference rules, in double
 * precision whatever the solver's precision (a compensated state is widened
 * hi + lo first).  Outputs are host arrays of n doubles in the state's
 * layout; any may be NULL.  The reference leaves the gradient unset on edges
 * and corners; here it is 0 there, so the result equals the reference's at
 * every point whose stencil avoids edges and corners, and is 0 on them.
 * Needs a 3D solver with >= 3 points along every axis; synchronous. */
int nlse_vorticity(nlse_solver *h, double lh2, double veps, double *wx, double *wy, double *wz, double *W);
/* the same into device arrays (n doubles each, any may be NULL), enqueued
 * on the solver's stream without allocating (one kernel, no scratch) */
int nlse_vorticity_device(nlse_solver *h, double lh2, double veps, double *wx, double *wy, double *wz,
                          double *W);
/* measurement helper: `reps` nlse_vorticity_device calls between two HIP
 * events on the solver's stream; mean device ms per call */
int nlse_vorticity_timed(nlse_solver *h, double lh2, double veps, double *wx, double *wy, double *wz,
                         double *W, int reps, double *ms);
/* ---- closed-form fields and initial conditions of the resident state ----
 * Grid coordinates follow the reference drivers' meshgrid: 1D x = xmin + i h;
 * 2D U[i*M + j] at x = xmin + i h, y = ymin + j h; 3D U[(i*N + j)*M + k] at
 * z = zmin + i h, x = xmin + j h, y = ymin + k h.  Evaluated on the GPU in
 * double precision (single-precision states are rounded on store; a
 * compensated state keeps hi + lo); synchronous.
 *
 * Exact solutions, kind = the reference's IC code
 * (src/NLSEmagic1D/NLSEmagic1D_FRS.m:201-225, 535-545; 2D/3D linear case
 * NLSE2D.m:1010, NLSE3D.m:1272), par = {a, s, OM, c, x0}:
 *   1 bright soliton sqrt(2 OM/s) sech(sqrt(OM/a)(x - x0 - c t)) e^{i((c/2a)x + (OM - c^2/4a)t)} (1D)
 *   2 dark soliton   sqrt(OM/s) tanh(sqrt(|OM/2a|)(x - x0 - c t)) e^{same phase}                   (1D)
 *   3 linear ground state e^{-|x|^2/2a} e^{-i d t} in d dimensions, with V = |x|^2/a
 *   4 grey soliton (vc = sqrt(-2 a OM), E = sqrt(a/-OM))                                          (1D)
 * nlse_set_exact writes the solution at time t into the state (kind 3 also
 * sets V; the others leave V).  nlse_exact_error compares the state with it:
 * out6 = RMS of Re, Im and |U|^2 differences (NLSE2D.m:1011-1013) followed by
 * their maxima (NLSEmagic1D_FRS.m:591-593; NaNs skipped as MATLAB's max). */
int nlse_set_exact(nlse_solver *h, int kind, const double *par, double t, double xmin, double ymin,
                   double zmin, double hgrid);
int nlse_exact_error(nlse_solver *h, int kind, const double *par, double t, double xmin, double ymin,
                     double zmin, double hgrid, double *out6);
/* U = re + i im at every point (V unchanged): the uniform background the
 * vortex generators are combined into (NLSE3D.m:405-407) */
int nlse_fill_state(nlse_solver *h, double re, double im);
/* Vortex initial conditions combined into the state as the reference's
 * generators do (s < 0: the normalised fields are multiplied and rescaled to
 * the new vortex's amplitude; s > 0: the new field is added).  The radial
 * profile solves the reference's discrete steady state
 * (NLSEmagic3D/NLS_STEADY_F.m:10-55) by Newton's method with the exact
 * tridiagonal Jacobian (the reference uses Newton-Krylov, nsoli.m, to the
 * same fixed point).
 * Replaces NLS_add_vortex(charge,om,x0,y0,vx,vy,phase,U,h,a,s,amp,X,Y,opt)
 * (src/NLSEmagic2D/NLSEmagic2D_Full_Research_Scripts/NLS_add_vortex.m:9-158);
 * a 2D solver. */
int nlse_add_vortex_2d(nlse_solver *h, int charge, double om, double x0, double y0, double vx, double vy,
                       double phase, double hgrid, double a, double s, double amp, double xmin, double ymin,
                       int opt);
/* Replaces NLSE_add_vortex_line(m,OM,x0,y0,z0,vx,vy,vz,phase,U,h,a,s,amp,X,Y,Z)
 * (src/NLSEmagic3D/NLSE_add_vortex_line.m:10-125; amp is unused there); a
 * 3D solver, the line runs along z through (x0, y0). */
int nlse_add_vortex_line(nlse_solver *h, int m, double OM, double x0, double y0, double z0, double vx,
                         double vy, double vz, double phase, double hgrid, double a, double s, double xmin,
                         double ymin, double zmin);
/* Replaces NLSE_add_vortex_ring(m,OM,x0,y0,z0,thetax,thetay,thetaz,vx,vy,vz,phase,U,h,a,s,X,Y,Z,d,opt_2D,ea,eb)
 * (src/NLSEmagic3D/NLSE_add_vortex_ring.m:10-534) for opt_2D = 0: the
 * profile on the (r, z) slice, the ring on the rotated lattice with
 * ellipticity ea, eb and velocity (vx, vy, vz), quarter turns by axis
 * permutation and other orientations by the reference's nearest-neighbour
 * lookup.  opt_2D > 0 (a Newton-Krylov refinement of the slice) is rejected. */
int nlse_add_vortex_ring(nlse_solver *h, int m, double OM, double x0, double y0, double z0, double thetax,
                         double thetay, double thetaz, double vx, double vy, double vz, double phase,
                         double hgrid, double a, double s, double xmin, double ymin, double zmin, double d,
                         int opt_2D, double ea, double eb);
/* the radial profile alone (host): f[q] at r = q*hr, q < n, f[0] = 0 and
 * f[n-1] the far-field value fixed; ansatz 0 dark sqrt(OM/s + a m^2/(s r^2)),
 * 1 dark sqrt(OM/s) tanh(sqrt(|OM/2a|) r), 2 bright B sech(C (r - R0));
 * optimize = 1 runs Newton to a residual of tol (1 + |F(ansatz)|) */
int nlse_vortex_profile(int m, double OM, double a, double s, double hr, int n, int ansatz, int optimize,
                        double tol, double *f);
/* run `steps` steps with a HIP event around every kernel launch and return
 * the mean device time of each launch of a step in ms_launch[0..nlaunch-1]
 * (nlaunch = 4 / stages per launch; per-kernel timing for roofline reports;
 * not available for 1D N <= 4096, which runs whole chunks in one kernel) */
int nlse_profile_launches(nlse_solver *h, int steps, double *ms_launch, int *nlaunch);
/* RK4 stages fused per kernel launch (the results are identical for every
 * choice): 2D 1, 2 or 4 (default 4, a whole step per launch); 3D 1 or 2
 * (default 2: two stages per plane-marching launch, CD and 2SHOC, with a
 * boundary other than MSD and planes of fewer than 2^31 points; MSD runs one
 * stage per launch) */
int nlse_set_fusion(nlse_solver *h, int stages_per_launch);
/* 2D: the stage kernel family and its rows per tile (results identical for
 * every choice):
 *   kind 0  LDS tiles of 64 columns x 8, 16 or 32 rows
 *   kind 1  row-sweep LDS tiles (one lane per column, DPP neighbours),
 *           8, 12 or 16 (default) rows
 *   kind 2  register-marching strips (one wave per strip, no LDS), segments
 *           of 3..256 rows, for 1, 2 or 4 stages per launch
 * 3D 2SHOC, one stage per launch: kind 1 (default) marches planes with D
 * kept on chip, tiles of 4 (default) or 8 rows x 60 columns; kind 0 runs a D
 * pass and an F pass per stage.  MSD boundaries always use the two passes.
 * Two stages per launch (nlse_set_fusion) march tiles of 4 rows x 56
 * columns (2SHOC) or 4 x 60 (CD) whatever the kind.
 * 1D: kind 1 (default) advances `rows` steps (1..80 CD, 1..32 2SHOC; default
 * 32 CD, 16 2SHOC) per launch on windows of 64-256 points with a halo, one per CU;
 * kind 0 runs a whole chunk in one workgroup (N <= 4096) or one launch per
 * stage. */
int nlse_set_kernel(nlse_solver *h, int kind, int rows);
/* nlse_set_kernel(h, 0, tile_rows) */
int nlse_set_tile(nlse_solver *h, int tile_rows);
/* the configuration in use: stages per launch, kernel kind and rows as set
 * by nlse_set_kernel (2D rows per tile or strip segment, 3D march rows, 1D
 * steps per launch); any pointer may be NULL */
int nlse_get_config(const nlse_solver *h, int *fusion, int *kind, int *rows);
/* 2D: time `steps` steps for every configuration (stages per launch x kernel
 * kind x rows) on a saved copy of the state, keep the fastest, and restore
 * the state; reports the stages per launch, the rows and the ms per step
 * (any pointer may be NULL; nlse_get_config gives the kind).  3D (not MSD):
 * the same between one and two stages per plane-marching launch.  1D and 3D
 * MSD: no-op that reports the stages per launch in use. */
int nlse_autotune(nlse_solver *h, int steps, int *best_fusion, int *best_tile, double *best_ms);
/* replay the chunk loop from a captured hipGraph of `steps_per_graph` steps
 * (0 = plain stream launches, the default) */
int nlse_set_graph(nlse_solver *h, int steps_per_graph);
/* device bytes held by a solver of this shape (admission control) */
long long nlse_footprint_bytes(int dim, int N, int M, int L, int method, int precision);

/* ---- batched 2D API ----
 * nprob independent 2D problems (sizes N[p] x M[p], each with its own s, a,
 * h^2, k and potential V, one boundary code for all) advanced together: one
 * persistent kernel launch per fused group of RK4 stages walks the tiles of
 * every problem.  precision: 2 = double (bit-identical to the reference CPU
 * integrator, problem by problem), 3 = single with a compensated hi + lo
 * state.  Problems are addressed by index p; states are double planes. */
typedef struct nlse_batch nlse_batch;
int nlse_batch_create(nlse_batch **out, int nprob, const int *N, const int *M, const double *s,
                      const double *a, const double *h2, const double *k, int method,
                      int precision, int BC);
void nlse_batch_destroy(nlse_batch *b);
long long nlse_batch_points(const nlse_batch *b);
int nlse_batch_set_state(nlse_batch *b, int p, const double *Ur, const double *Ui, const double *V);
int nlse_batch_get_state(nlse_batch *b, int p, double *Ur, double *Ui);
int nlse_batch_step(nlse_batch *b, int steps);
int nlse_batch_step_timed(nlse_batch *b, int steps, double *ms);
int nlse_batch_sync(nlse_batch *b);
/* per-problem diagnostics, out[3p..3p+2] as for nlse_diagnostics */
int nlse_batch_diagnostics(nlse_batch *b, double *out);
/* stage kernel family: 0 LDS tiles (rows 8, 16, 32), 1 row-sweep tiles
 * (rows 8 (default), 12, 16), 2 register strips (3..256-row segments, one
 * wave each); see nlse_set_kernel */
int nlse_batch_set_kernel(nlse_batch *b, int kind, int rows);
/* stages per launch, kernel kind, rows per tile, persistent grid */
int nlse_batch_get_config(const nlse_batch *b, int *fusion, int *kind, int *rows, int *grid);
/* stages per launch (1, 2, 4), rows per tile (strip segment) of the current kernel family,
 * persistent grid (multiple of 8 workgroups) and hipGraph steps per replay
 * (0 = off) */
int nlse_batch_configure(nlse_batch *b, int stages_per_launch, int tile_rows, int grid,
                         int steps_per_graph);
/* time `steps` steps for each (kernel family, stages per launch, tile rows,
 * grid) on a saved copy of the state, keep the fastest, restore the state
 * (nlse_batch_get_config gives the family) */
int nlse_batch_autotune(nlse_batch *b, int steps, int *best_fusion, int *best_tile, int *best_grid,
                        double *best_ms);

/* Groups of resident 2D solvers (no counterpart in the reference, which
 * solves one problem per call): the members (2D, one method and precision;
 * any sizes, parameters and boundaries) are advanced together by the batched
 * kernels, one persistent launch per RK4 step over the 16-row tiles of every
 * member, each in its own buffers.  Results are bit-identical to
 * stepping each member alone.  nlse_group_create waits for the members'
 * queued work; nlse_group_step enqueues on the group's own stream (returns
 * at once); nlse_group_done returns 1 when the group's work has finished, 0
 * while it runs (a negative NLSE_E* code on error); nlse_group_destroy waits
 * for it and hands the members back (their state advanced).  While a group
 * exists, its members must not be stepped, read or written on their own. */
typedef struct nlse_group nlse_group;
int nlse_group_create(nlse_group **g, nlse_solver *const *members, int n);
int nlse_group_step(nlse_group *g, int steps);
int nlse_group_done(nlse_group *g);
/* waits until the group's enqueued work has finished, the calling thread
 * asleep rather than spinning */
int nlse_group_sync(nlse_group *g);
void nlse_group_destroy(nlse_group *g);

/* measurement helper: mean device ms of a plain streaming copy of `bytes`
 * (multiple of 16) between two fresh device buffers, over `reps` launches;
 * gives the HBM counters' reading for a known byte count */
int nlse_copy_probe(long long bytes, int reps, double *ms);

/* description of the last error on this thread ("" if none) */
const char *nlse_last_error(void);
/* number of HIP devices visible (0 when no GPU) */
int nlse_device_count(void);
/* number of device allocations the library has made since it was loaded
 * (no counterpart in the reference): shows which calls allocate, e.g. that a
 * chunk loop with per-chunk diagnostics allocates nothing after its first
 * call */
long long nlse_device_allocations(void);

#ifdef __cplusplus
}
#endif
#endif
