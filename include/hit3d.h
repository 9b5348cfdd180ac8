/*
 * hit3d C ABI — MI355X (gfx950) kernels for hit3d's pseudo-spectral
 * Navier–Stokes step.
 *
 * Plain C: device pointers, sizes, a HIP stream passed as `void*` (NULL =
 * the default stream). No call allocates device memory: every entry point
 * that needs scratch takes a caller-owned workspace whose size comes from
 * hit3d_workspace_size(). All calls are asynchronous on `stream` and never
 * synchronise the host. On failure a call returns a non-zero status and
 * hit3d_last_error() (per host thread) holds the message; no C++ exception
 * crosses this boundary.
 *
 * Arrays (C order, interleaved complex):
 *   spectral state  complex[ncomp][nx][ny][nz/2+1]    (grid.py:7-11 layout)
 *   physical field  real[ncomp][nx][ny][nz]
 * with real = float (precision 4, complex64) or double (precision 8,
 * complex128). Grid sizes are powers of two, 4 <= n <= 1024 per axis.
 *
 * Each entry point names the reference symbol it replaces (file:line in
 * nfrank22-gh/3DHIT src/hit3d/).
 */
#ifndef HIT3D_H_
#define HIT3D_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HIT3D_OK 0
#define HIT3D_ERR_INVALID 1     /* bad shape, size, pointer or parameter */
#define HIT3D_ERR_HIP 2         /* HIP launch/runtime error */
#define HIT3D_ERR_WORKSPACE 3   /* workspace smaller than required */

/* Box description: replaces the static fields of `Grid` (grid.py:41-86). */
typedef struct hit3d_grid {
  int64_t nx, ny, nz;   /* grid points per axis */
  double lx, ly, lz;    /* box lengths (2*pi for the unit-wavenumber box) */
  int32_t precision;    /* 4: float32 state, 8: float64 state */
  int32_t reserved;     /* must be 0 */
} hit3d_grid;

/* Forcing plug-in selector (forcing.py:55-163, plus the constant-energy rule). */
#define HIT3D_FORCING_NONE 0      /* NoForcing      forcing.py:55-71   */
#define HIT3D_FORCING_LINEAR 1    /* LinearForcing  forcing.py:74-106: f = a u */
#define HIT3D_FORCING_BAND 2      /* BandForcing    forcing.py:109-163 */
#define HIT3D_FORCING_LUNDGREN 3  /* constant-energy f = A(u) u,
                                     A = (2 nu Omega - (E - energy)/tau) / (2E) */
typedef struct hit3d_forcing {
  int32_t kind;
  int32_t reserved;       /* must be 0 */
  double a;               /* LINEAR: coefficient A */
  double eps, kmin, kmax; /* BAND: injection rate and shell limits */
  double nu, energy, tau; /* LUNDGREN: viscosity in A(u), target E, relaxation
                             time (tau <= 0 or non-finite: no relaxation term) */
} hit3d_forcing;

/* Operation selectors for hit3d_workspace_size. */
#define HIT3D_OP_TO_PHYSICAL 0
#define HIT3D_OP_TO_SPECTRAL 1
#define HIT3D_OP_RHS 2
#define HIT3D_OP_RK4 3
#define HIT3D_OP_SPECTRUM 4
#define HIT3D_OP_REDUCE 5
#define HIT3D_OP_RHS_VJP 6
#define HIT3D_OP_RK4_VJP 7

/* Message of the last failed call on this host thread (always NUL-terminated);
   returns the full message length. */
int hit3d_last_error(char* buf, size_t len);

/* Library ABI version (1). */
int hit3d_version(void);

/* Bytes of workspace `op` needs. `ncomp` applies to the transforms; `dealiased`
   selects the pruned (1) or dense (0) step layout for RHS/RK4. */
int hit3d_workspace_size(const hit3d_grid* grid, int32_t op, int64_t ncomp, int32_t dealiased,
                         size_t* bytes);

/* Number of shell bins written by hit3d_shell_spectrum: smax + 1 with
   smax = ceil(sqrt(3) max(n) / 2) + 1 (diagnostics.py:78-80). */
int hit3d_shell_bins(const hit3d_grid* grid, int64_t* nbins);

/* Per-mode tables as the kernels compute them, full spectral grid
   [nx][ny][nz/2+1]: k2 and inv_k2 in the state precision, the 2/3 dealias
   mask (uint8), the shell index floor(sqrt(k2)/k0 + 1/2) (int32). Any output
   may be NULL. Replaces make_grid's k2/inv_k2/dealias (grid.py:109-120) and
   _shells (diagnostics.py:76-82). */
int hit3d_grid_tables(const hit3d_grid* grid, void* k2, void* inv_k2, uint8_t* mask,
                      int32_t* shell, void* stream);

/* irfftn over axes (1,2,3), 1/N normalisation: to_physical (grid.py:181-183). */
int hit3d_to_physical(const hit3d_grid* grid, const void* u_hat, void* u, int64_t ncomp,
                      void* work, size_t work_bytes, void* stream);

/* rfftn over axes (1,2,3), unnormalised: to_spectral (grid.py:186-188). */
int hit3d_to_spectral(const hit3d_grid* grid, const void* u, void* u_hat, int64_t ncomp,
                      void* work, size_t work_bytes, void* stream);

/* out = project(dealias(in)) (grid.py:166-178); either step may be skipped.
   do_project needs ncomp == 3. In-place allowed. */
int hit3d_dealias_project(const hit3d_grid* grid, const void* in, void* out, int64_t ncomp,
                          int32_t do_dealias, int32_t do_project, void* stream);

/* i kx u, i ky u, i kz u, -k2 u for op = 0..3 (grid.py:146-163). In-place allowed. */
int hit3d_spectral_derivative(const hit3d_grid* grid, const void* in, void* out, int64_t ncomp,
                              int32_t op, void* stream);

/* du = rhs(u_hat, params, t) of make_ns_rhs (navierstokes.py:62-97):
   P[dealias(F(u x w) + f)] - nu k2 u. `dealiased` = 1 promises that u_hat has
   no content outside the 2/3 box (|n_i| <= n_i/3; the k = 0 mode may be
   non-zero), which selects the pruned transforms; 0 runs dense transforms. */
int hit3d_rhs(const hit3d_grid* grid, const void* u_hat, void* du_hat, double nu,
              const hit3d_forcing* forcing, double t, int32_t dealiased, void* work,
              size_t work_bytes, void* stream);

/* In place, nsteps classic RK4 steps of size dt from t0: rollout
   (integrators.py:41-53) over rk4_step (integrators.py:32-38). If `diag` is
   not NULL it receives 4 doubles (device memory) for the final state:
   [forcing coefficient, 2E, 2 Omega, 2 E_band] (FP64 sums, compensated). */
int hit3d_rk4_steps(const hit3d_grid* grid, void* u_hat, double nu, double dt, int64_t nsteps,
                    double t0, const hit3d_forcing* forcing, int32_t dealiased, double* diag,
                    void* work, size_t work_bytes, void* stream);

/* hit3d_rk4_steps with the integrating-factor (Lawson) form of classic RK4:
   the viscous operator -nu k2 of linear_operator (navierstokes.py:100-103,
   the hook integrators.py:14-16 leaves for such schemes) is integrated
   exactly through exp(-nu k2 dt/2) and exp(-nu k2 dt), and RK4 advances the
   rest of the RHS. Same arguments, workspace and diag as hit3d_rk4_steps;
   nu >= 0, dt >= 0. */
int hit3d_rk4_if_steps(const hit3d_grid* grid, void* u_hat, double nu, double dt, int64_t nsteps,
                       double t0, const hit3d_forcing* forcing, int32_t dealiased, double* diag,
                       void* work, size_t work_bytes, void* stream);

/* FP64 shell spectrum (diagnostics.py:85-114): out[s] (device, nbins doubles)
   = sum over modes of shell s of w |u_c|^2 / (2 Ntot^2), summed over
   components (component = -1) or for one component (0..2). Bin 0 holds the
   k = 0 mode. Deterministic (fixed order), Neumaier-compensated.
   `dealiased` = 1 promises that u_hat has no content outside the 2/3 box
   (as for hit3d_rhs), so only the modes inside it are read (about 0.3 of
   the array); 0 reads every mode. */
int hit3d_shell_spectrum(const hit3d_grid* grid, const void* u_hat, int32_t component,
                         int32_t dealiased, double* out, void* work, size_t work_bytes,
                         void* stream);

/* FP64 reductions (diagnostics.py:56-70, forcing.py:48-52): out (device, 4
   doubles) = [E, eps = 2 nu Omega, Omega, E_band], E_band over
   kmin^2 <= k2 <= kmax^2. */
int hit3d_energy_dissipation(const hit3d_grid* grid, const void* u_hat, double nu, double kmin,
                             double kmax, double* out, void* work, size_t work_bytes,
                             void* stream);

/* hit3d_energy_dissipation without a workspace argument. No counterpart in
   the reference: for callers that cannot allocate device memory (code under
   HIP graph capture, serving loops, harnesses that forbid hipMalloc). Its
   scratch (the 1024 FP64 block partials and the totals, 24.8 KB) is a
   file-scope device array of the library, one copy per GPU, whose address is
   looked up once per device on the first call there; the call allocates
   nothing. Same kernels and summation order as hit3d_energy_dissipation, so
   the results are bit-identical. Every call on one GPU shares that GPU's
   scratch: calls that overlap on one GPU (concurrent streams or host
   threads) race and must be serialised by the caller. */
int hit3d_energy_dissipation_noalloc(const hit3d_grid* grid, const void* u_hat, double nu,
                                     double kmin, double kmax, double* out, void* stream);

/* out (device, 1 double) = sum of |u_hat|^2 over the modes outside the 2/3
   box: zero means the state can take the pruned (`dealiased` = 1) path. */
int hit3d_alias_energy(const hit3d_grid* grid, const void* u_hat, double* out, void* work,
                       size_t work_bytes, void* stream);

/* U[-0.5, 0.5) noise real[3][nx][ny][nz] from Philox4x64-10 with key (seed, 0):
   the draw of random_field (fields.py:45-46) with a counter-based generator. */
int hit3d_uniform_noise(const hit3d_grid* grid, uint64_t seed, void* u, void* stream);

/* u_hat[c][mode] *= factors[shell(mode)] for shell < nbins (fields.py:58-68). */
int hit3d_shell_scale(const hit3d_grid* grid, void* u_hat, int64_t ncomp, const double* factors,
                      int64_t nbins, void* stream);

/* Adjoint of the RHS: ubar = (d rhs/d u_hat at u_hat)^* v, the derivative
   jax.vjp takes of `rhs` (navierstokes.py:62-97; tests/test_grad.py:48-56),
   written out by hand. Gradient convention: for a real J, dJ = Re sum over
   all entries of conj(ubar) * du_hat (jax's cotangent is conj(ubar)).
   Includes the derivative of the state-dependent forcing coefficients
   (BAND, LUNDGREN). `dealiased` = 1 promises that u_hat has no content
   outside the 2/3 box (as for hit3d_rhs): u, K u and the projected cotangent
   then take the pruned inverse transforms (ubar is formed at every mode
   either way). ubar must not alias u_hat or v.
   If `param_grad` (device, 3 doubles) is not NULL, the derivatives of
   Re sum conj(v) rhs(u_hat) with respect to the physical scalars are ADDED
   to it: [d/dnu, d/dA (LINEAR forcing; else 0), d/deps (BAND; else 0)] —
   the gradients jax.grad takes through NSParams' leaves
   (navierstokes.py:40-59, tests/test_grad.py:43-45, 59-74). */
int hit3d_rhs_vjp(const hit3d_grid* grid, const void* u_hat, const void* v, void* ubar, double nu,
                  const hit3d_forcing* forcing, double t, int32_t dealiased, double* param_grad,
                  void* work, size_t work_bytes, void* stream);

/* Adjoint of one classic RK4 step (integrators.py:32-38): from the step's
   start state u_hat and the cotangent ubar_next of its result, writes the
   cotangent ubar of u_hat. Takes the stage inputs from `stages` (device, the
   three states y2, y3, y4 one after the other, as hit3d_rk4_steps_record
   keeps them) or, when `stages` is NULL, recomputes them; then runs the four
   stage adjoints in reverse. ubar may alias ubar_next, not u_hat or stages.
   `dealiased` = 1 promises that u_hat and the stage inputs have no content
   outside the 2/3 box (as for hit3d_rhs; hit3d_rk4_steps_record run with
   dealiased = 1 keeps such stages): the stage inputs are then recomputed
   with the pruned transforms and the stage adjoints transform u, K u and the
   projected cotangent with the pruned inverse passes. `param_grad` (device,
   3 doubles, or NULL) accumulates the step's parameter derivatives as in
   hit3d_rhs_vjp, summed over the four stages. */
int hit3d_rk4_step_vjp(const hit3d_grid* grid, const void* u_hat, const void* ubar_next,
                       void* ubar, double nu, double dt, double t, const hit3d_forcing* forcing,
                       int32_t dealiased, const void* stages, double* param_grad, void* work,
                       size_t work_bytes, void* stream);

/* hit3d_rk4_steps that also keeps what the step adjoint linearizes about
   (the forward half of checkpointed reverse mode; the reference gets it from
   jax's scan over rk4_step, integrators.py:41-53): for step n of the call,
   record[4n] = the state at the step's start and record[4n + 1 .. 4n + 3] =
   the inputs y2, y3, y4 of stages 2..4 (integrators.py:35-37), each a full
   state in the API layout (record holds 4 * nsteps states). Same arguments,
   workspace and time stepping as hit3d_rk4_steps, without diag. */
int hit3d_rk4_steps_record(const hit3d_grid* grid, void* u_hat, double nu, double dt, int64_t nsteps,
                           double t0, const hit3d_forcing* forcing, int32_t dealiased, void* record,
                           void* work, size_t work_bytes, void* stream);

/* The forcing plug-in's term(u_hat, grid, t) (forcing.py:5-20, 55-163):
   out = f(u_hat), the field the RHS adds before dealiasing and projection —
   0 (NONE), A u (LINEAR), (eps / 2E_band) [kmin^2 <= k2 <= kmax^2] u with the
   finfo(dtype).eps guard (BAND), A(u) u of the constant-energy forcing
   (LUNDGREN). The coefficients come from FP64 sums of u_hat reduced on the
   device (no host read). Workspace of HIT3D_OP_REDUCE. out must not alias
   u_hat. */
int hit3d_forcing_term(const hit3d_grid* grid, const void* u_hat, const hit3d_forcing* forcing,
                       void* out, void* work, size_t work_bytes, void* stream);

/* Measurement kernel (no reference counterpart): dst = src over `bytes`
   (a multiple of 16; both pointers 16-byte aligned) by a plain streaming
   copy, with non-temporal loads and stores when `nontemporal` is non-zero.
   Its rate is the attainable HBM bandwidth the pass kernels are held to. */
int hit3d_stream_copy(const void* src, void* dst, size_t bytes, int32_t nontemporal, void* stream);

/* Measurement form of hit3d_rk4_steps (no reference counterpart): runs the
   same nsteps with a HIP event after every launch on `stream`, waits for the
   last one, and returns per kernel class the summed milliseconds `ms[7]` and
   launch counts `counts[7]` (host arrays) for: 0 x-inverse + curl,
   1 y-inverse, 2 z c2r/cross/r2c, 3 y-forward, 4 x-forward + stage,
   5 FP64 finalize, 6 set-up. Synchronises the host. */
int hit3d_profile_rk4(const hit3d_grid* grid, void* u_hat, double nu, double dt, int64_t nsteps,
                      const hit3d_forcing* forcing, int32_t dealiased, double* ms, int64_t* counts,
                      void* work, size_t work_bytes, void* stream);

/* ---- Learned subgrid closure of the coarse LES ------------------------------
 * No reference counterpart: the closure, its training kernels and the
 * filtered-DNS pipeline are additions to hit3d's solver.
 *
 * A layer is a periodic 3x3x3 convolution over a d x h x w voxel grid
 * (v = (a*h + b)*w + c, c fastest, i.e. the physical layout real[nx][ny][nz]
 * with (d, h, w) = (nx, ny, nz)), computed as an implicit GEMM on BF16 MFMA
 * with FP32 accumulation. Activations are channels-last, x[v][cin_pad]
 * (bf16, channels >= cin zero); weights w[cout][27][cin] (fp32, offset
 * o = 9*(da+1) + 3*(db+1) + (dc+1)) are packed to bf16 by hit3d_conv_pack;
 * out[v][co] = bias[co] + sum_{o,ci} w[co][o][ci] x[v + offset(o)][ci]. */
typedef struct hit3d_conv {
  int32_t d, h, w;           /* voxels per axis: w % 16 == 0, d*h*w % 1024 == 0 */
  int32_t cin, cout;         /* channels (1 <= cin <= cin_pad, 1 <= cout <= cout_pad) */
  int32_t cin_pad, cout_pad; /* stored channels, 16 or 32 */
} hit3d_conv;

/* Packed weight widths kp (Wp[cout_pad][kp]) and kt (WT[cin_pad][kt]), and the
   workspace of hit3d_conv_backward. */
int hit3d_conv_sizes(const hit3d_conv* conv, int64_t* kp, int64_t* kt, size_t* work_bytes);

/* w fp32 [cout][27][cin] -> wp bf16 [cout_pad][kp], wt bf16 [cin_pad][kt]. */
int hit3d_conv_pack(const hit3d_conv* conv, const float* w, void* wp, void* wt, void* stream);

/* gelu = 1: pre fp32 [V][cout_pad] = conv(x) + bias, act bf16 [V][cout_pad] =
   GELU(pre) (exact erf form; zero beyond cout); pre may be NULL when no
   backward pass follows (it is then not written). gelu = 0: out fp32
   [V][cout]. */
int hit3d_conv_forward(const hit3d_conv* conv, const void* x, const void* wp, const float* bias,
                       int32_t gelu, float* pre, void* act, float* out, void* stream);

/* Gradients of one layer from grad_out (gelu = 1: fp32 [V][cout_pad] w.r.t.
   act; gelu = 0: fp32 [V][cout] w.r.t. out): grad_x fp32 [V][cin_pad] (NULL
   to skip), grad_w fp32 [cout][27][cin], grad_b fp32 [cout]. Deterministic
   (fixed-order reductions, no atomics). */
int hit3d_conv_backward(const hit3d_conv* conv, const void* x, const void* wt, const float* pre,
                        const float* grad_out, int32_t gelu, float* grad_x, float* grad_w,
                        float* grad_b, void* work, size_t work_bytes, void* stream);

/* One Adam step (Kingma & Ba) on n fp32 parameters, step count t >= 1. */
int hit3d_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
               float beta2, float eps, int64_t t, void* stream);

/* aos[v][ch] = bf16(soa[ch][v] * scale[ch]) for ch < c, 0 up to cpad. */
int hit3d_channels_pack(const float* soa, int32_t c, int64_t nvox, const float* scale, int32_t cpad,
                        void* aos, void* stream);

/* soa[ch][v] = aos[v*ld + ch] * scale[ch] for ch < c (fp32). */
int hit3d_channels_unpack(const float* aos, int32_t c, int32_t ld, int64_t nvox, const float* scale,
                          float* soa, void* stream);

/* Unpacks the net's 6 stress channels (aos fp32 [V][6], times scale[c]) into
   tau fp32 [6][V] with backscatter clipping: tau = 0 where the modelled
   transfer -tau_ij S_ij < 0, S_ij = (d_j u_i + d_i u_j)/2 from grad9 fp32
   [3 j + i][V]; keep[v] (uint8) records where tau was kept. */
int hit3d_stress_unpack_clip(const float* aos, int64_t nvox, const float* scale, const float* grad9,
                             float* tau, uint8_t* keep, void* stream);

/* Sharp spherical spectral filter onto a coarser grid of the same cubic box:
   u_coarse = (Ncoarse/Nfine) * u_fine on the modes with |k| <= kcut that the
   coarse grid holds without its Nyquist planes, 0 elsewhere. */
int hit3d_spectral_truncate(const hit3d_grid* fine, const void* u_fine, const hit3d_grid* coarse,
                            void* u_coarse, int64_t ncomp, double kcut, void* stream);

/* uu[6] = (u0u0, u0u1, u0u2, u1u1, u1u2, u2u2) from physical u[3]. */
int hit3d_velocity_products(const hit3d_grid* grid, const void* u, void* uu, void* stream);

/* adjoint = 0: out[3] = F_i = -sum_j i k_j tau_ij from the spectral stress
   in[6] (order of hit3d_velocity_products); adjoint = 1: its adjoint, in[3]
   -> out[6], in the convention dL = Re sum conj(g) dx. No aliasing. */
int hit3d_stress_divergence(const hit3d_grid* grid, const void* in, void* out, int32_t adjoint,
                            void* stream);

/* The LES right-hand side with its closure force, in one pass (float32
   grids; no reference counterpart): out = [k2 <= kcut^2] (du + gain S P M F)
   with F_i = -sum_j i k_j tau_ij from the spectral stress tau_hat[6] (order of
   hit3d_velocity_products), P M the 2/3 mask and Leray projection, S the
   real spectral shape `shape` [nx][ny][nz/2+1] (NULL: 1) and `gain` a device
   float; tau_hat NULL: no closure force (out = masked du). du_hat, out: [3]
   full layout; out may alias du_hat. */
int hit3d_les_combine(const hit3d_grid* grid, const void* du_hat, const void* tau_hat,
                      const void* shape, const float* gain, double kcut, void* out, void* stream);

/* Forcing that holds the resolved energy of an LES (no reference
   counterpart; the LES analogue of forcing.py's constant-energy forcing, with
   the closure's transfer counted in the balance): out = du + A u, with
   D = sum w Re(conj(u) du) / Ntot^2 (the rate of change of E that du causes),
   2E = sum w |u|^2 / Ntot^2 and A = (-D - (E - energy)/tau) / (2E), so that
   dE/dt = -(E - energy)/tau; tau <= 0 or infinite drops the relaxation term
   (dE/dt = 0), A = 0 when 2E <= finfo(dtype).eps. Full-layout states [3];
   coef (device, 3 doubles) receives [A, D, 2E]; FP64 fixed-order sums;
   workspace of HIT3D_OP_REDUCE. out may alias du_hat or u_hat; out NULL:
   only coef is written (hit3d_rk4_combine can apply the term). */
int hit3d_hold_energy(const hit3d_grid* grid, const void* u_hat, const void* du_hat, void* out,
                      double energy, double tau, double* coef, void* work, size_t work_bytes,
                      void* stream);

/* One combination step of classic RK4 over full-layout states [3]
   (integrators.py:32-38, whose stage sums the Python layer otherwise forms
   with separate tensor adds):
     stage 1: out = u + dt/2 k, acc = u + dt/6 k
     stage 2: out = u + dt/2 k, acc = acc + dt/3 k
     stage 3: out = u + dt k,   acc = acc + dt/3 k
     stage 4: out = acc + dt/6 k (u_hat unused, may be NULL)
   k_hat is the stage's derivative, out the next stage's input (stage 4: the
   new state). With hold_coef (device, as hit3d_hold_energy writes it) and
   hold_us (the stage's input state) the derivative is k_hat + hold_coef[0]
   hold_us: the energy hold's term applied here instead of by its own pass
   (both NULL: k_hat alone). Elementwise, so out may alias u_hat, k_hat or
   hold_us; acc must not alias out. */
int hit3d_rk4_combine(const hit3d_grid* grid, int32_t stage, double dt, const void* u_hat,
                      const void* k_hat, const double* hold_coef, const void* hold_us, void* acc,
                      void* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* HIT3D_H_ */
