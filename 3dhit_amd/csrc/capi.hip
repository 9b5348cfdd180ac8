// C ABI of the hit3d HIP library: argument checks, workspace layouts and the
// launch sequences of the RHS / RK4 step (integrators.py:32-53,
// navierstokes.py:62-97) and the diagnostics (diagnostics.py:45-114).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hit3d.h"
#include "closure.hpp"
#include "misc.hpp"
#include "passes.hpp"

using h3::cpx;
using h3::Geo;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define H3_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(HIT3D_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Per-pass HIP-event timing for hit3d_profile_rk4: one event after every
// launch on the stream, so consecutive events bracket one kernel each.
enum PassId { kP1 = 0, kP2, kP3, kP4, kP5, kFinalize, kOther, kNumPass };

struct PassTimer {
  std::vector<hipEvent_t> ev;
  std::vector<int> id;
  hipStream_t s = nullptr;
  bool ok = true;
  void record(int pass) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { ok = false; return; }
    if (hipEventRecord(e, s) != hipSuccess) { ok = false; hipEventDestroy(e); return; }
    ev.push_back(e);
    id.push_back(pass);
  }
  ~PassTimer() {
    for (hipEvent_t e : ev) hipEventDestroy(e);
  }
};

inline void mark(PassTimer* t, int pass) {
  if (t) t->record(pass);
}

bool pow2_ok(int64_t n) { return n >= 4 && n <= 1024 && (n & (n - 1)) == 0; }

int check_grid(const hit3d_grid* gd) {
  if (!gd) return fail(HIT3D_ERR_INVALID, "grid is NULL");
  if (!pow2_ok(gd->nx) || !pow2_ok(gd->ny) || !pow2_ok(gd->nz))
    return fail(HIT3D_ERR_INVALID, "grid sizes must be powers of two in [4, 1024], got " +
                                       std::to_string(gd->nx) + "x" + std::to_string(gd->ny) +
                                       "x" + std::to_string(gd->nz));
  if (gd->precision != 4 && gd->precision != 8)
    return fail(HIT3D_ERR_INVALID, "precision must be 4 (float32) or 8 (float64), got " +
                                       std::to_string(gd->precision));
  if (!(gd->lx > 0 && gd->ly > 0 && gd->lz > 0) || !std::isfinite(gd->lx) ||
      !std::isfinite(gd->ly) || !std::isfinite(gd->lz))
    return fail(HIT3D_ERR_INVALID, "box lengths must be positive and finite");
  if (gd->reserved != 0) return fail(HIT3D_ERR_INVALID, "grid.reserved must be 0");
  return HIT3D_OK;
}

int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// pruned: keep only |n_i| <= N_i/3 (compact rows padded to 16 elements);
// dense: every mode, compact layout == API layout.
Geo make_geo(const hit3d_grid& gd, bool pruned) {
  Geo g{};
  g.nx = int(gd.nx);
  g.ny = int(gd.ny);
  g.nz = int(gd.nz);
  g.nzh = g.nz / 2 + 1;
  g.cx = g.nx / 3;
  g.cy = g.ny / 3;
  g.cz = g.nz / 3;
  if (pruned) {
    g.kx = 2 * g.cx + 1;
    g.ky = 2 * g.cy + 1;
    g.kz = g.cz + 1;
    // rows on 128-byte lines when that pads by at most 5 %, else on 64-byte
    // halves (86 -> 88 at N = 256, 171 -> 176 at 512, 342 -> 352 at 1024)
    g.kzs = int(round_up(g.kz, 16));
    if (20 * (g.kzs - g.kz) > g.kz) g.kzs = int(round_up(g.kz, 8));
    g.jp = int(round_up(int64_t(g.ky) * g.kzs, 16));
  } else {
    g.kx = g.nx;
    g.ky = g.ny;
    g.kz = g.nzh;
    g.kzs = g.kz;
    g.jp = g.ny * g.nzh;
  }
  g.px = (g.kx + 1) / 2;
  g.py = (g.ky + 1) / 2;
  g.k0x = 2 * M_PI / gd.lx;
  g.k0y = 2 * M_PI / gd.ly;
  g.k0z = 2 * M_PI / gd.lz;
  return g;
}

int64_t shell_bins(const hit3d_grid& gd) {
  const int64_t n = std::max(gd.nx, std::max(gd.ny, gd.nz));
  return int64_t(std::ceil(std::sqrt(3.0) * double(n) / 2.0)) + 1 + 1;
}

// Workspace carving: fixed order, 256-byte aligned slices.
struct Carve {
  size_t off = 0;
  char* base = nullptr;
  void* take(size_t bytes) {
    off = (off + 255) & ~size_t(255);
    void* p = base ? base + off : nullptr;
    off += bytes;
    return p;
  }
};

struct StepBuffers {
  void* u_c = nullptr;  // compact step-start state (pruned only)
  void* us = nullptr;   // stage input / rhs output
  void* acc = nullptr;  // RK4 accumulator
  void* bufA = nullptr; // [5][nx][jp]
  void* bufB = nullptr; // [6][nx][ny][kz]
  double* partials = nullptr;
  double* scalars = nullptr;
};

size_t cbytes(const hit3d_grid& gd) { return size_t(2) * gd.precision; }

StepBuffers carve_step(const hit3d_grid& gd, const Geo& g, bool pruned, bool rk4, void* work,
                       size_t* total) {
  Carve c;
  c.base = static_cast<char*>(work);
  const size_t cb = cbytes(gd);
  const size_t st = size_t(3) * g.kx * g.jp * cb;
  StepBuffers b;
  if (pruned) b.u_c = c.take(st);
  b.us = c.take(st);
  if (rk4) b.acc = c.take(st);
  b.bufA = c.take(size_t(5) * g.nx * g.jp * cb);
  // pencil buffer: six fields, x-major or y-major with padded rows
  const int yrow = gd.precision == 4 ? h3::pencil_row<float>(g) : h3::pencil_row<double>(g);
  b.bufB = c.take(size_t(6) * std::max<size_t>(size_t(g.nx) * g.ny * g.kzs, size_t(g.ny) * yrow) * cb);
  const int nb = std::max(h3::x_stage_blocks(g.nx, g.ky * g.kzs), h3::kReduceBlocks);
  b.partials = static_cast<double*>(c.take(size_t(3) * nb * sizeof(double)));
  b.scalars = static_cast<double*>(c.take(16 * sizeof(double)));
  *total = c.off;
  return b;
}

size_t reduce_bytes(const hit3d_grid& gd, double** partials, double** scalars, void* work) {
  Carve c;
  c.base = static_cast<char*>(work);
  double* p = static_cast<double*>(c.take(size_t(3) * h3::kReduceBlocks * sizeof(double)));
  double* s = static_cast<double*>(c.take(16 * sizeof(double)));
  if (partials) *partials = p;
  if (scalars) *scalars = s;
  (void)gd;
  return c.off;
}

size_t spectrum_bytes(const hit3d_grid& gd, double** partials, double** mid, void* work) {
  Carve c;
  c.base = static_cast<char*>(work);
  const Geo g = make_geo(gd, false);
  const int nbins = int(shell_bins(gd));
  double* p = static_cast<double*>(c.take(size_t(h3::spectrum_blocks(g)) * nbins * sizeof(double)));
  double* m = static_cast<double*>(c.take(h3::spectrum_mid_doubles(nbins) * sizeof(double)));
  if (partials) *partials = p;
  if (mid) *mid = m;
  return c.off;
}

size_t transform_bytes(const hit3d_grid& gd, int64_t ncomp) {
  return size_t(ncomp) * gd.nx * gd.ny * (gd.nz / 2 + 1) * cbytes(gd);
}

// ---- FP64 finalize: partial sums -> totals and the forcing coefficient ----

using h3::FinalizeArgs;

__global__ __launch_bounds__(256) void k_finalize(const double* __restrict__ partials, int nblocks,
                                                  FinalizeArgs fa, double* __restrict__ scalars) {
  __shared__ double sh[3][256];
  h3::KSum s[3];
  for (int i = threadIdx.x; i < nblocks; i += 256)
    for (int c = 0; c < 3; ++c) s[c].add(partials[3 * i + c]);
  for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] = s[c].value();
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) h3::finalize_write(fa, sh[0][0], sh[1][0], sh[2][0], scalars);
}

// coef = [A, D, 2E] from k_hold_partials' sums: D = dE/dt of du, 2E = sum w|u|^2 / Ntot^2,
// A = (-D - (E - energy)/tau) / (2E) (no relaxation term when tau <= 0 or infinite), 0 when
// 2E <= tiny. Same fixed-order Neumaier reduction as k_finalize.
__global__ __launch_bounds__(256) void k_hold_coef(const double* __restrict__ partials, int nblocks,
                                                   double inv_ntot2, double energy, double tau,
                                                   double tiny, double* __restrict__ coef) {
  __shared__ double sh[2][256];
  h3::KSum s[2];
  for (int i = threadIdx.x; i < nblocks; i += 256)
    for (int c = 0; c < 2; ++c) s[c].add(partials[3 * i + c]);
  for (int c = 0; c < 2; ++c) sh[c][threadIdx.x] = s[c].value();
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int c = 0; c < 2; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double D = sh[0][0] * inv_ntot2, twoE = sh[1][0] * inv_ntot2;
    double a = 0.0;
    if (twoE > tiny) {
      double rate = -D;
      if (tau > 0.0 && isfinite(tau)) rate -= (0.5 * twoE - energy) / tau;
      a = rate / twoE;
    }
    coef[0] = a;
    coef[1] = D;
    coef[2] = twoE;
  }
}

__global__ void k_energy_out(const double* __restrict__ scalars, double nu,
                             double* __restrict__ out) {
  // [E, eps, Omega, E_band]
  out[0] = 0.5 * scalars[1];
  out[1] = nu * scalars[2];
  out[2] = 0.5 * scalars[2];
  out[3] = 0.5 * scalars[3];
}

double ntot2_inv(const hit3d_grid& gd) {
  const double n = double(gd.nx) * double(gd.ny) * double(gd.nz);
  return 1.0 / (n * n);
}

int check_forcing(const hit3d_forcing* f) {
  if (!f) return HIT3D_OK;
  if (f->reserved != 0) return fail(HIT3D_ERR_INVALID, "forcing.reserved must be 0");
  if (f->kind < HIT3D_FORCING_NONE || f->kind > HIT3D_FORCING_LUNDGREN)
    return fail(HIT3D_ERR_INVALID, "unknown forcing kind " + std::to_string(f->kind));
  return HIT3D_OK;
}

FinalizeArgs finalize_args(const hit3d_grid& gd, const hit3d_forcing* f) {
  FinalizeArgs fa{};
  fa.kind = f ? f->kind : HIT3D_FORCING_NONE;
  if (f) {
    fa.eps = f->eps;
    fa.nu = f->nu;
    fa.energy = f->energy;
    fa.tau = f->tau;
  }
  fa.tiny = gd.precision == 4 ? double(FLT_EPSILON) : DBL_EPSILON;
  fa.inv_ntot2 = ntot2_inv(gd);
  return fa;
}

bool needs_sums(const hit3d_forcing* f) {
  return f && (f->kind == HIT3D_FORCING_BAND || f->kind == HIT3D_FORCING_LUNDGREN);
}

// P2-P4 run over nch chunks of the x range, each a sub-grid of Nx/nch
// x-planes through a pencil buffer that the chunks reuse. When the six-field
// pencil buffer exceeds 192 MB (no longer fits the 256 MB Infinity Cache
// beside the pass's other streams), the fewest chunks (2, 4, 8 or 16) of at
// most 160 MB each: a chunk's intermediates then stay in that cache between
// P2, P3 and P4. Measured (RK4 steps/s, one box): 256^3 float32 (277 MB)
// 702 -> 742 with 2 chunks (3: 714, 4: 657, as the passes' grids shrink);
// 512^3 float32 (2.2 GB) 85.4 -> 87.7 with 16 (8: 82.1, 4: 84.9); 1024^3
// would need ~110 chunks and runs whole.
int yz_chunks(const hit3d_grid& gd, const Geo& g) {
  const double mb = double(1 << 20);
  const double pencil_bytes = 6.0 * g.nx * g.ny * g.kzs * double(cbytes(gd));
  if (pencil_bytes <= 192.0 * mb) return 1;
  for (int nch = 2; nch <= 16; nch *= 2)
    if (g.nx % (2 * nch) == 0 && g.nx / nch >= 16 && pencil_bytes / nch <= 160.0 * mb) return nch;
  return 1;
}

// P1-P4 of one RHS evaluation on the compact stage input `us`.
template <class T>
int rhs_passes(const hit3d_grid& gd, const Geo& g, const cpx<T>* us, cpx<T>* bufA, cpx<T>* bufB,
               hipStream_t s, PassTimer* tm = nullptr, const h3::FinalizeJob* job = nullptr) {
  const int64_t st_fs = int64_t(g.kx) * g.jp;
  const int64_t a_fs = int64_t(g.nx) * g.jp;
  const int nch = yz_chunks(gd, g);
  auto x_at = [&](int ch) { return ch * g.nx / nch; };   // 2 nch divides nx
  Geo gm = g;
  gm.nx = g.nx / nch;
  const int yrow = h3::pencil_row<T>(gm);   // 0: x-major pencil buffer
  const int64_t b_fs = yrow ? int64_t(g.ny) * yrow : int64_t(gm.nx) * g.ny * g.kzs;
  const double n = double(gd.nx) * double(gd.ny) * double(gd.nz);
  H3_HIP(h3::launch_x_inverse_curl<T>(g.nx, us, st_fs, bufA, a_fs, g, s, job));
  mark(tm, kP1);
  for (int ch = 0; ch < nch; ++ch) {
    const int x0 = x_at(ch);
    Geo gc = g;
    gc.nx = x_at(ch + 1) - x0;
    cpx<T>* const bufAc = bufA + int64_t(x0) * g.jp;
    H3_HIP(h3::launch_y_inverse_curl<T>(g.ny, bufAc, a_fs, bufB, b_fs, gc, yrow, s));
    mark(tm, kP2);
    H3_HIP(h3::launch_z_fused<T>(g.nz, bufB, b_fs, gc.nx, g.ny, g.kz, g.kzs, yrow, T(1.0 / (n * n)), s));
    mark(tm, kP3);
    h3::StridedArgs<T> y{};
    y.in = bufB + h3::kZOut * b_fs;
    y.out = bufAc;
    y.in_fs = b_fs;
    y.out_fs = a_fs;
    if (yrow) {   // column j = x kzs + iz, rows yrow apart
      y.in_slab = g.kzs;
      y.in_rs = yrow;
    } else {
      y.in_slab = int64_t(g.ny) * g.kzs;
    }
    y.out_slab = g.jp;
    y.B = g.kzs;
    y.J = gc.nx * g.kzs;
    y.nf = 3;
    y.in_k = g.ny;
    y.in_p = (g.ny + 1) / 2;
    y.out_k = g.ky;
    y.out_p = g.py;
    H3_HIP(h3::launch_strided<T>(g.ny, -1, y, s));
    mark(tm, kP4);
  }
  return HIT3D_OK;
}

template <class T>
h3::StageArgs<T> stage_args(const hit3d_grid& gd, const Geo& g, const hit3d_forcing* f, double nu,
                            double dt, const cpx<T>* bufA, double* scalars) {
  h3::StageArgs<T> a{};
  a.nl = bufA;
  a.nl_fs = int64_t(g.nx) * g.jp;
  a.st_fs = int64_t(g.kx) * g.jp;
  a.nu = T(nu);
  a.dt = T(dt);
  a.fterm = h3::kTermNone;
  a.coef_const = T(0);
  a.coef_dev = nullptr;
  a.kmin2 = T(1);
  a.kmax2 = T(0);
  if (f) {
    switch (f->kind) {
      case HIT3D_FORCING_LINEAR:
        a.fterm = h3::kTermUniform;
        a.coef_const = T(f->a);
        break;
      case HIT3D_FORCING_BAND:
        a.fterm = h3::kTermBand;
        a.coef_dev = scalars;
        a.kmin2 = T(f->kmin * f->kmin);
        a.kmax2 = T(f->kmax * f->kmax);
        break;
      case HIT3D_FORCING_LUNDGREN:
        a.fterm = h3::kTermUniform;
        a.coef_dev = scalars;
        break;
      default:
        break;
    }
  }
  a.g = g;
  (void)gd;
  return a;
}

template <class T>
int reduce_state(const hit3d_grid& gd, const Geo& g, const cpx<T>* st, const hit3d_forcing* f,
                 double* partials, double* scalars, hipStream_t s) {
  T kmin2 = T(1), kmax2 = T(0);
  if (f && f->kind == HIT3D_FORCING_BAND) {
    kmin2 = T(f->kmin * f->kmin);
    kmax2 = T(f->kmax * f->kmax);
  }
  H3_HIP(h3::launch_quad_partials<T>(st, int64_t(g.kx) * g.jp, g, kmin2, kmax2, partials, s));
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, partials, h3::kReduceBlocks,
                     finalize_args(gd, f), scalars);
  H3_HIP(hipGetLastError());
  return HIT3D_OK;
}

template <class T>
int rhs_impl(const hit3d_grid& gd, const cpx<T>* u_hat, cpx<T>* du_hat, double nu,
             const hit3d_forcing* f, bool pruned, void* work, hipStream_t s) {
  const Geo g = make_geo(gd, pruned);
  size_t total;
  StepBuffers b = carve_step(gd, g, pruned, false, work, &total);
  const int64_t st_fs = int64_t(g.kx) * g.jp;
  const int64_t full_fs = int64_t(g.nx) * g.ny * g.nzh;
  const cpx<T>* us = u_hat;
  cpx<T>* out = du_hat;
  if (pruned) {
    H3_HIP(h3::launch_gather<T>(u_hat, full_fs, static_cast<cpx<T>*>(b.u_c), st_fs, 3, g, s));
    us = static_cast<cpx<T>*>(b.u_c);
    out = static_cast<cpx<T>*>(b.us);
  }
  if (needs_sums(f)) {
    int rc = reduce_state<T>(gd, g, us, f, b.partials, b.scalars, s);
    if (rc) return rc;
  }
  int rc = rhs_passes<T>(gd, g, us, static_cast<cpx<T>*>(b.bufA), static_cast<cpx<T>*>(b.bufB), s);
  if (rc) return rc;
  h3::StageArgs<T> a = stage_args<T>(gd, g, f, nu, 0.0, static_cast<cpx<T>*>(b.bufA), b.scalars);
  a.us = us;
  a.out = out;
  H3_HIP(h3::launch_x_forward_stage<T>(g.nx, h3::kStageRhs, a, s, nullptr));
  if (pruned) H3_HIP(h3::launch_scatter<T>(out, st_fs, du_hat, full_fs, 3, g, s));
  return HIT3D_OK;
}

template <class T>
int rk4_impl(const hit3d_grid& gd, cpx<T>* u_hat, double nu, double dt, int64_t nsteps,
             const hit3d_forcing* f, bool pruned, double* diag, void* work, hipStream_t s,
             PassTimer* tm = nullptr, bool integrating_factor = false, cpx<T>* record = nullptr) {
  const Geo g = make_geo(gd, pruned);
  size_t total;
  StepBuffers b = carve_step(gd, g, pruned, true, work, &total);
  const int64_t st_fs = int64_t(g.kx) * g.jp;
  const int64_t full_fs = int64_t(g.nx) * g.ny * g.nzh;
  cpx<T>* u = u_hat;
  cpx<T>* us = static_cast<cpx<T>*>(b.us);
  cpx<T>* acc = static_cast<cpx<T>*>(b.acc);
  cpx<T>* bufA = static_cast<cpx<T>*>(b.bufA);
  cpx<T>* bufB = static_cast<cpx<T>*>(b.bufB);
  if (pruned) {
    u = static_cast<cpx<T>*>(b.u_c);
    H3_HIP(h3::launch_gather<T>(u_hat, full_fs, u, st_fs, 3, g, s));
  }
  const bool sums = needs_sums(f);
  if (sums) {
    int rc = reduce_state<T>(gd, g, u, f, b.partials, b.scalars, s);
    if (rc) return rc;
  }
  mark(tm, kOther);
  h3::StageArgs<T> a = stage_args<T>(gd, g, f, nu, dt, bufA, b.scalars);
  a.partials = sums ? b.partials : nullptr;
  a.acc = acc;
  a.u = u;
  // The sums a stage kernel writes set the forcing coefficient of the next
  // stage's P5; block 0 of the next stage's P1 reduces them (FinalizeJob),
  // and only the last stage's sums need a launch of their own.
  h3::FinalizeJob job{b.partials, 0, finalize_args(gd, f), b.scalars};
  bool pending = false;
  // record: per step, the state at its start and the inputs of stages 2..4,
  // in the API layout (the linearization points of the step's adjoint)
  auto keep = [&](const cpx<T>* src, int64_t n, int slot) -> hipError_t {
    cpx<T>* dst = record + (4 * n + slot) * 3 * full_fs;
    if (pruned) return h3::launch_scatter<T>(src, st_fs, dst, full_fs, 3, g, s);
    return hipMemcpyAsync(dst, src, size_t(3 * full_fs) * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s);
  };
  for (int64_t n = 0; n < nsteps; ++n) {
    if (record) H3_HIP(keep(u, n, 0));
    for (int stage = 1; stage <= 4; ++stage) {
      const cpx<T>* us_in = stage == 1 ? u : us;
      int rc = rhs_passes<T>(gd, g, us_in, bufA, bufB, s, tm, pending ? &job : nullptr);
      if (rc) return rc;
      pending = false;
      a.us = us_in;
      a.out = stage == 4 ? u : us;
      int nb = 0;
      H3_HIP(h3::launch_x_forward_stage<T>(g.nx, integrating_factor ? stage + 4 : stage, a, s, &nb));
      mark(tm, kP5);
      if (record && stage < 4) H3_HIP(keep(us, n, stage));
      if (sums) {
        job.nblocks = nb;
        pending = true;
        mark(tm, kFinalize);
      }
    }
  }
  if (pending) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, b.partials, job.nblocks, job.fa,
                       b.scalars);
    H3_HIP(hipGetLastError());
  }
  if (diag) {
    int rc = reduce_state<T>(gd, g, u, f, b.partials, b.scalars, s);
    if (rc) return rc;
    H3_HIP(hipMemcpyAsync(diag, b.scalars, 4 * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  if (pruned) H3_HIP(h3::launch_scatter<T>(u, st_fs, u_hat, full_fs, 3, g, s));
  return HIT3D_OK;
}

// x- then y-inverse of nf dense fields (in -> work, then in place)
template <class T>
int xy_inverse_impl(const hit3d_grid& gd, const cpx<T>* in, int nf, cpx<T>* work, hipStream_t s) {
  const Geo g = make_geo(gd, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  h3::StridedArgs<T> x{};
  x.in = in;
  x.out = work;
  x.in_fs = x.out_fs = fs;
  x.in_slab = x.out_slab = 0;
  x.B = g.ny * g.nzh;
  x.J = x.B;
  x.nf = nf;
  x.in_k = x.out_k = g.nx;
  x.in_p = x.out_p = (g.nx + 1) / 2;
  H3_HIP(h3::launch_strided<T>(g.nx, +1, x, s));
  h3::StridedArgs<T> y{};
  y.in = work;
  y.out = work;
  y.in_fs = y.out_fs = fs;
  y.in_slab = y.out_slab = int64_t(g.ny) * g.nzh;
  y.B = g.nzh;
  y.J = g.nx * g.nzh;
  y.nf = nf;
  y.in_k = y.out_k = g.ny;
  y.in_p = y.out_p = (g.ny + 1) / 2;
  H3_HIP(h3::launch_strided<T>(g.ny, +1, y, s));
  return HIT3D_OK;
}

// y- then x-forward of nf dense fields (work in place, then work -> out)
template <class T>
int yx_forward_impl(const hit3d_grid& gd, cpx<T>* work, cpx<T>* out, int nf, hipStream_t s) {
  const Geo g = make_geo(gd, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  h3::StridedArgs<T> y{};
  y.in = work;
  y.out = work;
  y.in_fs = y.out_fs = fs;
  y.in_slab = y.out_slab = int64_t(g.ny) * g.nzh;
  y.B = g.nzh;
  y.J = g.nx * g.nzh;
  y.nf = nf;
  y.in_k = y.out_k = g.ny;
  y.in_p = y.out_p = (g.ny + 1) / 2;
  H3_HIP(h3::launch_strided<T>(g.ny, -1, y, s));
  h3::StridedArgs<T> x{};
  x.in = work;
  x.out = out;
  x.in_fs = x.out_fs = fs;
  x.in_slab = x.out_slab = 0;
  x.B = g.ny * g.nzh;
  x.J = x.B;
  x.nf = nf;
  x.in_k = x.out_k = g.nx;
  x.in_p = x.out_p = (g.nx + 1) / 2;
  H3_HIP(h3::launch_strided<T>(g.nx, -1, x, s));
  return HIT3D_OK;
}

template <class T>
int to_physical_impl(const hit3d_grid& gd, const cpx<T>* in, T* out, int nf, cpx<T>* work,
                     hipStream_t s) {
  const Geo g = make_geo(gd, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  int rc = xy_inverse_impl<T>(gd, in, nf, work, s);
  if (rc) return rc;
  const double n = double(gd.nx) * double(gd.ny) * double(gd.nz);
  H3_HIP(h3::launch_z_c2r<T>(g.nz, work, fs, out, int64_t(g.nx) * g.ny * g.nz, nf,
                             int64_t(g.nx) * g.ny, g.nzh, T(1.0 / n), s));
  return HIT3D_OK;
}

template <class T>
int to_spectral_impl(const hit3d_grid& gd, const T* in, cpx<T>* out, int nf, cpx<T>* work,
                     hipStream_t s) {
  const Geo g = make_geo(gd, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  H3_HIP(h3::launch_z_r2c<T>(g.nz, in, int64_t(g.nx) * g.ny * g.nz, work, fs, nf,
                             int64_t(g.nx) * g.ny, g.nzh, s));
  return yx_forward_impl<T>(gd, work, out, nf, s);
}

// ---- adjoint (VJP) of the RHS and of one RK4 step --------------------------

struct VjpBuffers {
  void* S = nullptr;     // [9] transform inputs, later [6] transformed cross products
  void* tw = nullptr;    // [9] transform workspace (the z pass works in place)
  double* partials = nullptr;
  double* scalars = nullptr;
  double* dot = nullptr;
};

VjpBuffers carve_vjp(const hit3d_grid& gd, void* work, size_t* total) {
  Carve c;
  c.base = static_cast<char*>(work);
  const size_t cb = cbytes(gd);
  const size_t fs = size_t(gd.nx) * gd.ny * (gd.nz / 2 + 1);
  // the pruned path's compact, x-inverse and y-inverse layouts (row padding
  // can make them larger than a dense field on small grids)
  const Geo gp = make_geo(gd, true);
  const size_t c_fs = size_t(gp.kx) * gp.jp, a_fs = size_t(gp.nx) * gp.jp;
  const size_t b_fs = size_t(gp.nx) * gp.ny * gp.kzs;
  VjpBuffers b;
  b.S = c.take(9 * std::max({fs, c_fs, b_fs}) * cb);
  b.tw = c.take(9 * std::max(fs, a_fs) * cb);
  b.partials = static_cast<double*>(c.take(size_t(3) * h3::kReduceBlocks * sizeof(double)));
  b.scalars = static_cast<double*>(c.take(16 * sizeof(double)));
  b.dot = static_cast<double*>(c.take(16 * sizeof(double)));
  *total = c.off;
  return b;
}

// The RHS VJP at u for the cotangent v into `out`, or through the epilogue
// `ep` (accumulated, and the next RK4 stage's cotangent formed) when given.
// `pruned`: u has no content outside the 2/3 box, so u, K u and P M v go
// through the pruned inverse transforms (compact prep, x and y passes over
// the retained rows, z pass on kz-mode pencils); the forward transforms of
// the cross products stay dense (ubar has content at every mode).
template <class T>
int rhs_vjp_impl(const hit3d_grid& gd, const cpx<T>* u, const cpx<T>* v, cpx<T>* out, double nu,
                 const hit3d_forcing* f, void* work, hipStream_t s, double* pgrad = nullptr,
                 const h3::AdjEpilogue<T>* ep = nullptr, bool pruned = false) {
  size_t total;
  VjpBuffers b = carve_vjp(gd, work, &total);
  const Geo g = make_geo(gd, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  const double ntot = double(gd.nx) * double(gd.ny) * double(gd.nz);
  cpx<T>* S = static_cast<cpx<T>*>(b.S);
  cpx<T>* tw = static_cast<cpx<T>*>(b.tw);
  const int64_t npencils = int64_t(g.nx) * g.ny;
  int rc = 0;
  if (pruned) {
    // S: compact prep [9][Kx][jp], then the y-inverse output [9][Nx][Ny][kzs];
    // tw: the x-inverse output [9][Nx][jp], then the z pass output (dense)
    const Geo gp = make_geo(gd, true);
    const int64_t c_fs = int64_t(gp.kx) * gp.jp, a_fs = int64_t(gp.nx) * gp.jp;
    const int64_t b_fs = int64_t(gp.nx) * gp.ny * gp.kzs;
    H3_HIP(h3::launch_adj_prep_compact<T>(u, v, fs, S, c_fs, gp, ntot, s));
    h3::StridedArgs<T> x{};
    x.in = S;
    x.out = tw;
    x.in_fs = c_fs;
    x.out_fs = a_fs;
    x.in_slab = x.out_slab = 0;
    x.B = gp.jp;
    x.J = gp.jp;
    x.nf = 9;
    x.in_k = gp.kx;
    x.in_p = gp.px;
    x.out_k = gp.nx;
    x.out_p = (gp.nx + 1) / 2;
    H3_HIP(h3::launch_strided<T>(gp.nx, +1, x, s));
    h3::StridedArgs<T> y{};
    y.in = tw;
    y.out = S;
    y.in_fs = a_fs;
    y.out_fs = b_fs;
    y.in_slab = gp.jp;
    y.out_slab = int64_t(gp.ny) * gp.kzs;
    y.B = gp.kzs;
    y.J = gp.nx * gp.kzs;
    y.nf = 9;
    y.in_k = gp.ky;
    y.in_p = gp.py;
    y.out_k = gp.ny;
    y.out_p = (gp.ny + 1) / 2;
    H3_HIP(h3::launch_strided<T>(gp.ny, +1, y, s));
    H3_HIP(h3::launch_z_adj<T>(g.nz, S, b_fs, gp.kz, gp.kzs, tw, fs, g.nzh, npencils, T(1.0 / ntot),
                               true, s));
  } else {
    // u, w = K u, g to physical space (x, y, then z fused with w x g, g x u
    // and the z-forward), in place in tw
    H3_HIP(h3::launch_adj_prep<T>(u, v, S, fs, g, ntot, s));
    if ((rc = xy_inverse_impl<T>(gd, S, 9, tw, s))) return rc;
    H3_HIP(h3::launch_z_adj<T>(g.nz, tw, fs, g.nzh, g.nzh, tw, fs, g.nzh, npencils, T(1.0 / ntot),
                               false, s));
  }
  // the cross products back along y and x into S[0..5]
  if ((rc = yx_forward_impl<T>(gd, tw, S, 6, s))) return rc;
  T kmin2 = T(1), kmax2 = T(0);
  h3::AdjCombine c{};
  c.fkind = f ? f->kind : HIT3D_FORCING_NONE;
  c.nu = nu;
  c.inv_ntot = 1.0 / ntot;
  c.inv_ntot2 = ntot2_inv(gd);
  c.tiny = gd.precision == 4 ? double(FLT_EPSILON) : DBL_EPSILON;
  c.scalars = b.scalars;
  c.dot = b.dot + 1;
  if (f && f->kind == HIT3D_FORCING_LINEAR) c.a = f->a;
  if (f && (f->kind == HIT3D_FORCING_BAND || f->kind == HIT3D_FORCING_LUNDGREN)) {
    const bool band = f->kind == HIT3D_FORCING_BAND;
    if (band) {
      kmin2 = T(f->kmin * f->kmin);
      kmax2 = T(f->kmax * f->kmax);
    }
    c.fnu = f->nu;
    c.tau = f->tau;
    rc = reduce_state<T>(gd, g, u, f, b.partials, b.scalars, s);  // coef, S_E, S_Om, S_band
    if (rc) return rc;
    H3_HIP(h3::launch_dot_partials<T>(u, v, fs, g, kmin2, kmax2, band ? 1 : 0, b.partials, s));
    FinalizeArgs fa{};
    fa.kind = HIT3D_FORCING_NONE;
    fa.inv_ntot2 = 1.0;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, b.partials, h3::kReduceBlocks, fa,
                       b.dot);  // dot[1] = the raw sum
    H3_HIP(hipGetLastError());
  }
  // parameter derivatives first: the epilogue may overwrite v
  if (pgrad)  // b.scalars[3] = 2E_band of u (reduce_state above, band forcing)
    H3_HIP(h3::launch_param_grad<T>(v, u, fs, g, kmin2, kmax2, c.fkind, b.scalars + 3, c.tiny,
                                    b.partials, pgrad, s));
  H3_HIP(h3::launch_adj_combine<T>(S, v, u, out, fs, g, kmin2, kmax2, c,
                                   ep ? *ep : h3::AdjEpilogue<T>{}, s));
  return HIT3D_OK;
}

size_t state_bytes(const hit3d_grid& gd) {
  return size_t(3) * gd.nx * gd.ny * (gd.nz / 2 + 1) * cbytes(gd);
}

struct StepVjpBuffers {
  void *ub, *k, *y2, *y3, *y4, *kb, *rhs_ws, *vjp_ws;
  size_t rhs_bytes, vjp_bytes;
};

StepVjpBuffers carve_step_vjp(const hit3d_grid& gd, void* work, size_t* total) {
  Carve c;
  c.base = static_cast<char*>(work);
  StepVjpBuffers b{};
  const size_t st = state_bytes(gd);
  b.ub = c.take(st);
  b.k = c.take(st);
  b.y2 = c.take(st);
  b.y3 = c.take(st);
  b.y4 = c.take(st);
  b.kb = c.take(st);
  size_t dense = 0, pruned = 0;  // the stage recomputation runs either layout
  carve_step(gd, make_geo(gd, false), false, false, nullptr, &dense);
  carve_step(gd, make_geo(gd, true), true, false, nullptr, &pruned);
  b.rhs_bytes = std::max(dense, pruned);
  carve_vjp(gd, nullptr, &b.vjp_bytes);
  b.rhs_ws = c.take(b.rhs_bytes);
  b.vjp_ws = c.take(b.vjp_bytes);
  *total = c.off;
  return b;
}

template <class T>
int rk4_step_vjp_impl(const hit3d_grid& gd, const cpx<T>* u, const cpx<T>* ub_in, cpx<T>* out,
                      double nu, double dt, const hit3d_forcing* f, void* work, hipStream_t s,
                      double* pgrad, bool pruned, const cpx<T>* stages = nullptr) {
  size_t total;
  StepVjpBuffers b = carve_step_vjp(gd, work, &total);
  const int64_t n = int64_t(3) * gd.nx * gd.ny * (gd.nz / 2 + 1);
  auto P = [](void* p) { return static_cast<cpx<T>*>(p); };
  cpx<T>*ub = P(b.ub), *k = P(b.k), *kb = P(b.kb);
  const cpx<T>*y2 = P(b.y2), *y3 = P(b.y3), *y4 = P(b.y4);
  const T hdt = T(0.5 * dt), fdt = T(dt), sdt = T(dt / 6.0), tdt = T(dt / 3.0);
  H3_HIP(hipMemcpyAsync(ub, ub_in, n * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
  int rc = 0;
  if (stages) {  // the caller kept the stage inputs of this step
    y2 = stages;
    y3 = stages + n;
    y4 = stages + 2 * n;
  } else {
    // stage inputs (integrators.py:34-37); the pruned transforms when u has no
    // content outside the 2/3 box (then neither have the stage inputs)
    if ((rc = rhs_impl<T>(gd, u, k, nu, f, pruned, b.rhs_ws, s))) return rc;
    H3_HIP(h3::launch_axpby<T>(P(b.y2), T(1), u, hdt, k, n, s));
    if ((rc = rhs_impl<T>(gd, y2, k, nu, f, pruned, b.rhs_ws, s))) return rc;
    H3_HIP(h3::launch_axpby<T>(P(b.y3), T(1), u, hdt, k, n, s));
    if ((rc = rhs_impl<T>(gd, y3, k, nu, f, pruned, b.rhs_ws, s))) return rc;
    H3_HIP(h3::launch_axpby<T>(P(b.y4), T(1), u, fdt, k, n, s));
  }
  // reverse sweep: kbar_i from the combination weights, then stage adjoints
  // 4..1, each accumulating its value into `out` and forming the next
  // kbar in its epilogue (kbar_3 = dt/3 ub + dt ybar_4, kbar_2 = dt/3 ub +
  // dt/2 ybar_3, kbar_1 = dt/6 ub + dt/2 ybar_2)
  H3_HIP(h3::launch_axpby<T>(kb, sdt, ub, T(0), ub, n, s));
  h3::AdjEpilogue<T> ep{};
  ep.acc = out;
  ep.acc_base = ub;
  ep.kbn = kb;
  ep.ub = ub;
  ep.alpha = tdt;
  ep.beta = fdt;
  if ((rc = rhs_vjp_impl<T>(gd, y4, kb, nullptr, nu, f, b.vjp_ws, s, pgrad, &ep, pruned))) return rc;
  ep.acc_base = out;
  ep.beta = hdt;
  if ((rc = rhs_vjp_impl<T>(gd, y3, kb, nullptr, nu, f, b.vjp_ws, s, pgrad, &ep, pruned))) return rc;
  ep.alpha = sdt;
  if ((rc = rhs_vjp_impl<T>(gd, y2, kb, nullptr, nu, f, b.vjp_ws, s, pgrad, &ep, pruned))) return rc;
  ep.kbn = nullptr;
  if ((rc = rhs_vjp_impl<T>(gd, u, kb, nullptr, nu, f, b.vjp_ws, s, pgrad, &ep, pruned))) return rc;
  return HIT3D_OK;
}

int check_work(void* work, size_t have, size_t need) {
  if (need > have)
    return fail(HIT3D_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) +
                                         " bytes, got " + std::to_string(have));
  if (need > 0 && !work) return fail(HIT3D_ERR_INVALID, "workspace is NULL");
  if (reinterpret_cast<uintptr_t>(work) % 256 != 0)
    return fail(HIT3D_ERR_INVALID, "workspace must be 256-byte aligned");
  return HIT3D_OK;
}

int check_ptr(const void* p, const char* name) {
  if (!p) return fail(HIT3D_ERR_INVALID, std::string(name) + " is NULL");
  if (reinterpret_cast<uintptr_t>(p) % 16 != 0)
    return fail(HIT3D_ERR_INVALID, std::string(name) + " must be 16-byte aligned");
  return HIT3D_OK;
}

#define H3_TRY(expr)       \
  do {                     \
    int rc_ = (expr);      \
    if (rc_) return rc_;   \
  } while (0)

// [E, eps, Omega, E_band] of one state from the 1024 fixed-block FP64 partial
// sums: `partials` (3 * kReduceBlocks doubles, every entry written by the
// partials kernel) and `scalars` (16 doubles) are device scratch.
int energy_dissipation_impl(const hit3d_grid& gd, const void* u_hat, double nu, double kmin,
                            double kmax, double* out, double* partials, double* scalars,
                            hipStream_t s) {
  const Geo g = make_geo(gd, false);
  hit3d_forcing band{};
  band.kind = HIT3D_FORCING_BAND;
  band.kmin = kmin;
  band.kmax = kmax;
  if (gd.precision == 4)
    H3_TRY(reduce_state<float>(gd, g, static_cast<const cpx<float>*>(u_hat), &band, partials,
                               scalars, s));
  else
    H3_TRY(reduce_state<double>(gd, g, static_cast<const cpx<double>*>(u_hat), &band, partials,
                                scalars, s));
  hipLaunchKernelGGL(k_energy_out, dim3(1), dim3(1), 0, s, scalars, nu, out);
  H3_HIP(hipGetLastError());
  return HIT3D_OK;
}

// ---- device-resident scratch of hit3d_energy_dissipation_noalloc ------------
// The partial sums and totals live in this file-scope array instead of a
// caller workspace. Each GPU holds its own copy for the life of the process;
// host code reaches it only through the address hipGetSymbolAddress returns
// on that device (the array's name is not a device address in host code).
// Every call on one GPU shares that GPU's copy: calls that overlap on one
// GPU (two streams, or two host threads) race on it.
constexpr size_t kScratchDoubles = 3 * size_t(h3::kReduceBlocks) + 32;
static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_dev;  // per device ordinal, null until first use

int scratch_on_current_device(double** out) {
  int dev = 0;
  H3_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (size_t(dev) >= g_scratch_dev.size()) g_scratch_dev.resize(size_t(dev) + 1, nullptr);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    H3_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(HIT3D_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return HIT3D_OK;
}

}  // namespace

extern "C" {

int hit3d_last_error(char* buf, size_t len) {
  if (buf && len) {
    const size_t n = std::min(len - 1, g_err.size());
    std::memcpy(buf, g_err.data(), n);
    buf[n] = '\0';
  }
  return int(g_err.size());
}

int hit3d_version(void) { return 1; }

int hit3d_workspace_size(const hit3d_grid* grid, int32_t op, int64_t ncomp, int32_t dealiased,
                         size_t* bytes) {
  H3_TRY(check_grid(grid));
  if (!bytes) return fail(HIT3D_ERR_INVALID, "bytes is NULL");
  switch (op) {
    case HIT3D_OP_TO_PHYSICAL:
    case HIT3D_OP_TO_SPECTRAL:
      if (ncomp < 1) return fail(HIT3D_ERR_INVALID, "ncomp must be >= 1");
      *bytes = transform_bytes(*grid, ncomp);
      return HIT3D_OK;
    case HIT3D_OP_RHS:
    case HIT3D_OP_RK4: {
      const bool pruned = dealiased != 0;
      const Geo g = make_geo(*grid, pruned);
      carve_step(*grid, g, pruned, op == HIT3D_OP_RK4, nullptr, bytes);
      return HIT3D_OK;
    }
    case HIT3D_OP_SPECTRUM:
      *bytes = spectrum_bytes(*grid, nullptr, nullptr, nullptr);
      return HIT3D_OK;
    case HIT3D_OP_REDUCE:
      *bytes = reduce_bytes(*grid, nullptr, nullptr, nullptr);
      return HIT3D_OK;
    case HIT3D_OP_RHS_VJP:
      carve_vjp(*grid, nullptr, bytes);
      return HIT3D_OK;
    case HIT3D_OP_RK4_VJP:
      carve_step_vjp(*grid, nullptr, bytes);
      return HIT3D_OK;
    default:
      return fail(HIT3D_ERR_INVALID, "unknown op " + std::to_string(op));
  }
}

int hit3d_shell_bins(const hit3d_grid* grid, int64_t* nbins) {
  H3_TRY(check_grid(grid));
  if (!nbins) return fail(HIT3D_ERR_INVALID, "nbins is NULL");
  *nbins = shell_bins(*grid);
  return HIT3D_OK;
}

int hit3d_grid_tables(const hit3d_grid* grid, void* k2, void* inv_k2, uint8_t* mask,
                      int32_t* shell, void* stream) {
  H3_TRY(check_grid(grid));
  const Geo g = make_geo(*grid, false);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    H3_HIP(h3::launch_tables<float>(g, static_cast<float*>(k2), static_cast<float*>(inv_k2), mask,
                                    shell, s));
  else
    H3_HIP(h3::launch_tables<double>(g, static_cast<double*>(k2), static_cast<double*>(inv_k2),
                                     mask, shell, s));
  return HIT3D_OK;
}

int hit3d_to_physical(const hit3d_grid* grid, const void* u_hat, void* u, int64_t ncomp,
                      void* work, size_t work_bytes, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(u, "u"));
  if (ncomp < 1 || ncomp > 64) return fail(HIT3D_ERR_INVALID, "ncomp must be in [1, 64]");
  H3_TRY(check_work(work, work_bytes, transform_bytes(*grid, ncomp)));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    return to_physical_impl<float>(*grid, static_cast<const cpx<float>*>(u_hat),
                                   static_cast<float*>(u), int(ncomp),
                                   static_cast<cpx<float>*>(work), s);
  return to_physical_impl<double>(*grid, static_cast<const cpx<double>*>(u_hat),
                                  static_cast<double*>(u), int(ncomp),
                                  static_cast<cpx<double>*>(work), s);
}

int hit3d_to_spectral(const hit3d_grid* grid, const void* u, void* u_hat, int64_t ncomp,
                      void* work, size_t work_bytes, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(u, "u"));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  if (ncomp < 1 || ncomp > 64) return fail(HIT3D_ERR_INVALID, "ncomp must be in [1, 64]");
  H3_TRY(check_work(work, work_bytes, transform_bytes(*grid, ncomp)));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    return to_spectral_impl<float>(*grid, static_cast<const float*>(u),
                                   static_cast<cpx<float>*>(u_hat), int(ncomp),
                                   static_cast<cpx<float>*>(work), s);
  return to_spectral_impl<double>(*grid, static_cast<const double*>(u),
                                  static_cast<cpx<double>*>(u_hat), int(ncomp),
                                  static_cast<cpx<double>*>(work), s);
}

int hit3d_dealias_project(const hit3d_grid* grid, const void* in, void* out, int64_t ncomp,
                          int32_t do_dealias, int32_t do_project, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(in, "in"));
  H3_TRY(check_ptr(out, "out"));
  if (ncomp < 1 || ncomp > 64) return fail(HIT3D_ERR_INVALID, "ncomp must be in [1, 64]");
  if (do_project && ncomp != 3) return fail(HIT3D_ERR_INVALID, "project needs ncomp == 3");
  const Geo g = make_geo(*grid, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    H3_HIP(h3::launch_dealias_project<float>(static_cast<const cpx<float>*>(in),
                                             static_cast<cpx<float>*>(out), fs, int(ncomp), g,
                                             do_dealias, do_project, s));
  else
    H3_HIP(h3::launch_dealias_project<double>(static_cast<const cpx<double>*>(in),
                                              static_cast<cpx<double>*>(out), fs, int(ncomp), g,
                                              do_dealias, do_project, s));
  return HIT3D_OK;
}

int hit3d_spectral_derivative(const hit3d_grid* grid, const void* in, void* out, int64_t ncomp,
                              int32_t op, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(in, "in"));
  H3_TRY(check_ptr(out, "out"));
  if (ncomp < 1 || ncomp > 64) return fail(HIT3D_ERR_INVALID, "ncomp must be in [1, 64]");
  if (op < 0 || op > 3) return fail(HIT3D_ERR_INVALID, "op must be 0..3");
  const Geo g = make_geo(*grid, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    H3_HIP(h3::launch_spectral_op<float>(static_cast<const cpx<float>*>(in),
                                         static_cast<cpx<float>*>(out), fs, int(ncomp), g, op, s));
  else
    H3_HIP(h3::launch_spectral_op<double>(static_cast<const cpx<double>*>(in),
                                          static_cast<cpx<double>*>(out), fs, int(ncomp), g, op,
                                          s));
  return HIT3D_OK;
}

int hit3d_rhs(const hit3d_grid* grid, const void* u_hat, void* du_hat, double nu,
              const hit3d_forcing* forcing, double t, int32_t dealiased, void* work,
              size_t work_bytes, void* stream) {
  (void)t;  // no forcing of the reference depends on t
  H3_TRY(check_grid(grid));
  H3_TRY(check_forcing(forcing));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(du_hat, "du_hat"));
  if (u_hat == du_hat) return fail(HIT3D_ERR_INVALID, "du_hat must not alias u_hat");
  const bool pruned = dealiased != 0;
  size_t need;
  carve_step(*grid, make_geo(*grid, pruned), pruned, false, nullptr, &need);
  H3_TRY(check_work(work, work_bytes, need));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    return rhs_impl<float>(*grid, static_cast<const cpx<float>*>(u_hat),
                           static_cast<cpx<float>*>(du_hat), nu, forcing, pruned, work, s);
  return rhs_impl<double>(*grid, static_cast<const cpx<double>*>(u_hat),
                          static_cast<cpx<double>*>(du_hat), nu, forcing, pruned, work, s);
}

int hit3d_rk4_steps(const hit3d_grid* grid, void* u_hat, double nu, double dt, int64_t nsteps,
                    double t0, const hit3d_forcing* forcing, int32_t dealiased, double* diag,
                    void* work, size_t work_bytes, void* stream) {
  (void)t0;
  H3_TRY(check_grid(grid));
  H3_TRY(check_forcing(forcing));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  if (nsteps < 0) return fail(HIT3D_ERR_INVALID, "nsteps must be >= 0");
  if (!std::isfinite(dt) || !std::isfinite(nu))
    return fail(HIT3D_ERR_INVALID, "dt and nu must be finite");
  const bool pruned = dealiased != 0;
  size_t need;
  carve_step(*grid, make_geo(*grid, pruned), pruned, true, nullptr, &need);
  H3_TRY(check_work(work, work_bytes, need));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    return rk4_impl<float>(*grid, static_cast<cpx<float>*>(u_hat), nu, dt, nsteps, forcing,
                           pruned, diag, work, s);
  return rk4_impl<double>(*grid, static_cast<cpx<double>*>(u_hat), nu, dt, nsteps, forcing,
                          pruned, diag, work, s);
}

int hit3d_rk4_steps_record(const hit3d_grid* grid, void* u_hat, double nu, double dt, int64_t nsteps,
                           double t0, const hit3d_forcing* forcing, int32_t dealiased, void* record,
                           void* work, size_t work_bytes, void* stream) {
  (void)t0;
  H3_TRY(check_grid(grid));
  H3_TRY(check_forcing(forcing));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(record, "record"));
  if (record == u_hat) return fail(HIT3D_ERR_INVALID, "record must not alias u_hat");
  if (nsteps < 0) return fail(HIT3D_ERR_INVALID, "nsteps must be >= 0");
  if (!std::isfinite(dt) || !std::isfinite(nu))
    return fail(HIT3D_ERR_INVALID, "dt and nu must be finite");
  const bool pruned = dealiased != 0;
  size_t need;
  carve_step(*grid, make_geo(*grid, pruned), pruned, true, nullptr, &need);
  H3_TRY(check_work(work, work_bytes, need));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    return rk4_impl<float>(*grid, static_cast<cpx<float>*>(u_hat), nu, dt, nsteps, forcing, pruned,
                           nullptr, work, s, nullptr, false, static_cast<cpx<float>*>(record));
  return rk4_impl<double>(*grid, static_cast<cpx<double>*>(u_hat), nu, dt, nsteps, forcing, pruned,
                          nullptr, work, s, nullptr, false, static_cast<cpx<double>*>(record));
}

int hit3d_rk4_if_steps(const hit3d_grid* grid, void* u_hat, double nu, double dt, int64_t nsteps,
                       double t0, const hit3d_forcing* forcing, int32_t dealiased, double* diag,
                       void* work, size_t work_bytes, void* stream) {
  (void)t0;
  H3_TRY(check_grid(grid));
  H3_TRY(check_forcing(forcing));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  if (nsteps < 0) return fail(HIT3D_ERR_INVALID, "nsteps must be >= 0");
  if (!std::isfinite(dt) || !std::isfinite(nu) || nu < 0.0 || dt < 0.0)
    return fail(HIT3D_ERR_INVALID, "dt and nu must be finite and non-negative");
  const bool pruned = dealiased != 0;
  size_t need;
  carve_step(*grid, make_geo(*grid, pruned), pruned, true, nullptr, &need);
  H3_TRY(check_work(work, work_bytes, need));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    return rk4_impl<float>(*grid, static_cast<cpx<float>*>(u_hat), nu, dt, nsteps, forcing, pruned,
                           diag, work, s, nullptr, true);
  return rk4_impl<double>(*grid, static_cast<cpx<double>*>(u_hat), nu, dt, nsteps, forcing, pruned,
                          diag, work, s, nullptr, true);
}

int hit3d_shell_spectrum(const hit3d_grid* grid, const void* u_hat, int32_t component,
                         int32_t dealiased, double* out, void* work, size_t work_bytes,
                         void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(out, "out"));
  if (component < -1 || component > 2) return fail(HIT3D_ERR_INVALID, "component must be -1..2");
  H3_TRY(check_work(work, work_bytes, spectrum_bytes(*grid, nullptr, nullptr, nullptr)));
  double *partials, *mid;
  spectrum_bytes(*grid, &partials, &mid, work);
  const Geo g = make_geo(*grid, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  const int nbins = int(shell_bins(*grid));
  const int box = dealiased != 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int nb = 0;
  if (grid->precision == 4)
    H3_HIP(h3::launch_spectrum_partials<float>(static_cast<const cpx<float>*>(u_hat), fs,
                                               component, g, box, nbins, partials, &nb, s));
  else
    H3_HIP(h3::launch_spectrum_partials<double>(static_cast<const cpx<double>*>(u_hat), fs,
                                                component, g, box, nbins, partials, &nb, s));
  H3_HIP(h3::launch_spectrum_merge(partials, nb, nbins, mid, 0.5 * ntot2_inv(*grid), out, s));
  return HIT3D_OK;
}

int hit3d_energy_dissipation(const hit3d_grid* grid, const void* u_hat, double nu, double kmin,
                             double kmax, double* out, void* work, size_t work_bytes,
                             void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(out, "out"));
  H3_TRY(check_work(work, work_bytes, reduce_bytes(*grid, nullptr, nullptr, nullptr)));
  double *partials, *scalars;
  reduce_bytes(*grid, &partials, &scalars, work);
  return energy_dissipation_impl(*grid, u_hat, nu, kmin, kmax, out, partials, scalars,
                                 static_cast<hipStream_t>(stream));
}

int hit3d_forcing_term(const hit3d_grid* grid, const void* u_hat, const hit3d_forcing* forcing,
                       void* out, void* work, size_t work_bytes, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_forcing(forcing));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(out, "out"));
  if (out == u_hat) return fail(HIT3D_ERR_INVALID, "out must not alias u_hat");
  H3_TRY(check_work(work, work_bytes, reduce_bytes(*grid, nullptr, nullptr, nullptr)));
  double *partials, *scalars;
  reduce_bytes(*grid, &partials, &scalars, work);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const Geo g = make_geo(*grid, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  const int kind = forcing ? forcing->kind : HIT3D_FORCING_NONE;
  const bool band = kind == HIT3D_FORCING_BAND;
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    const cpx<T>* u = static_cast<const cpx<T>*>(u_hat);
    // the coefficient of BAND / LUNDGREN from the FP64 sums of u, on the device
    if (needs_sums(forcing)) H3_TRY(reduce_state<T>(*grid, g, u, forcing, partials, scalars, s));
    const T c = kind == HIT3D_FORCING_LINEAR ? T(forcing->a) : T(0);
    H3_HIP(h3::launch_forcing_term<T>(u, static_cast<cpx<T>*>(out), fs, g, band ? 1 : 0, c,
                                      needs_sums(forcing) ? scalars : nullptr,
                                      band ? T(forcing->kmin * forcing->kmin) : T(1),
                                      band ? T(forcing->kmax * forcing->kmax) : T(0), s));
    return HIT3D_OK;
  };
  if (grid->precision == 4) return run(float{});
  return run(double{});
}

int hit3d_energy_dissipation_noalloc(const hit3d_grid* grid, const void* u_hat, double nu,
                                     double kmin, double kmax, double* out, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(out, "out"));
  const size_t need = reduce_bytes(*grid, nullptr, nullptr, nullptr);
  if (need > sizeof(double) * kScratchDoubles)
    return fail(HIT3D_ERR_WORKSPACE, "hit3d_energy_dissipation_noalloc needs " + std::to_string(need) +
                                         " bytes of scratch, its device array holds " +
                                         std::to_string(sizeof(double) * kScratchDoubles));
  double* base = nullptr;
  H3_TRY(scratch_on_current_device(&base));
  double *partials, *scalars;
  reduce_bytes(*grid, &partials, &scalars, base);  // same carving as the workspace form
  return energy_dissipation_impl(*grid, u_hat, nu, kmin, kmax, out, partials, scalars,
                                 static_cast<hipStream_t>(stream));
}

int hit3d_alias_energy(const hit3d_grid* grid, const void* u_hat, double* out, void* work,
                       size_t work_bytes, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(out, "out"));
  H3_TRY(check_work(work, work_bytes, reduce_bytes(*grid, nullptr, nullptr, nullptr)));
  double *partials, *scalars;
  reduce_bytes(*grid, &partials, &scalars, work);
  const Geo g = make_geo(*grid, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    H3_HIP(h3::launch_alias_partials<float>(static_cast<const cpx<float>*>(u_hat), fs, g, partials,
                                            s));
  else
    H3_HIP(h3::launch_alias_partials<double>(static_cast<const cpx<double>*>(u_hat), fs, g,
                                             partials, s));
  FinalizeArgs fa{};
  fa.kind = HIT3D_FORCING_NONE;
  fa.inv_ntot2 = 1.0;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, partials, h3::kReduceBlocks, fa,
                     scalars);
  H3_HIP(hipGetLastError());
  H3_HIP(hipMemcpyAsync(out, scalars + 1, sizeof(double), hipMemcpyDeviceToDevice, s));
  return HIT3D_OK;
}

int hit3d_stream_copy(const void* src, void* dst, size_t bytes, int32_t nontemporal,
                      void* stream) {
  H3_TRY(check_ptr(src, "src"));
  H3_TRY(check_ptr(dst, "dst"));
  if (bytes % 16 != 0) return fail(HIT3D_ERR_INVALID, "bytes must be a multiple of 16");
  H3_HIP(h3::launch_stream_copy(src, dst, int64_t(bytes / 16), nontemporal,
                                static_cast<hipStream_t>(stream)));
  return HIT3D_OK;
}

int hit3d_uniform_noise(const hit3d_grid* grid, uint64_t seed, void* u, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(u, "u"));
  const int64_t n = 3 * grid->nx * grid->ny * grid->nz;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    H3_HIP(h3::launch_philox_uniform<float>(static_cast<float*>(u), n, seed, s));
  else
    H3_HIP(h3::launch_philox_uniform<double>(static_cast<double*>(u), n, seed, s));
  return HIT3D_OK;
}

int hit3d_shell_scale(const hit3d_grid* grid, void* u_hat, int64_t ncomp, const double* factors,
                      int64_t nbins, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  if (!factors) return fail(HIT3D_ERR_INVALID, "factors is NULL");
  if (ncomp < 1 || ncomp > 64) return fail(HIT3D_ERR_INVALID, "ncomp must be in [1, 64]");
  if (nbins < 1 || nbins > shell_bins(*grid))
    return fail(HIT3D_ERR_INVALID, "nbins must be in [1, hit3d_shell_bins]");
  const Geo g = make_geo(*grid, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    H3_HIP(h3::launch_shell_scale<float>(static_cast<cpx<float>*>(u_hat), fs, int(ncomp), g,
                                         factors, int(nbins), s));
  else
    H3_HIP(h3::launch_shell_scale<double>(static_cast<cpx<double>*>(u_hat), fs, int(ncomp), g,
                                          factors, int(nbins), s));
  return HIT3D_OK;
}

int hit3d_profile_rk4(const hit3d_grid* grid, void* u_hat, double nu, double dt, int64_t nsteps,
                      const hit3d_forcing* forcing, int32_t dealiased, double* ms, int64_t* counts,
                      void* work, size_t work_bytes, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_forcing(forcing));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  if (!ms || !counts) return fail(HIT3D_ERR_INVALID, "ms and counts must not be NULL");
  if (nsteps < 1) return fail(HIT3D_ERR_INVALID, "nsteps must be >= 1");
  const bool pruned = dealiased != 0;
  size_t need;
  carve_step(*grid, make_geo(*grid, pruned), pruned, true, nullptr, &need);
  H3_TRY(check_work(work, work_bytes, need));
  hipStream_t s = static_cast<hipStream_t>(stream);
  PassTimer tm;
  tm.s = s;
  tm.record(-1);
  int rc = grid->precision == 4
               ? rk4_impl<float>(*grid, static_cast<cpx<float>*>(u_hat), nu, dt, nsteps, forcing,
                                 pruned, nullptr, work, s, &tm)
               : rk4_impl<double>(*grid, static_cast<cpx<double>*>(u_hat), nu, dt, nsteps,
                                  forcing, pruned, nullptr, work, s, &tm);
  if (rc) return rc;
  if (!tm.ok) return fail(HIT3D_ERR_HIP, "event creation/record failed");
  H3_HIP(hipEventSynchronize(tm.ev.back()));
  for (int i = 0; i < kNumPass; ++i) {
    ms[i] = 0.0;
    counts[i] = 0;
  }
  for (size_t i = 1; i < tm.ev.size(); ++i) {
    float t = 0.f;
    H3_HIP(hipEventElapsedTime(&t, tm.ev[i - 1], tm.ev[i]));
    ms[tm.id[i]] += t;
    counts[tm.id[i]] += 1;
  }
  return HIT3D_OK;
}

int hit3d_rhs_vjp(const hit3d_grid* grid, const void* u_hat, const void* v, void* ubar, double nu,
                  const hit3d_forcing* forcing, double t, int32_t dealiased, double* param_grad,
                  void* work, size_t work_bytes, void* stream) {
  (void)t;
  H3_TRY(check_grid(grid));
  H3_TRY(check_forcing(forcing));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(v, "v"));
  H3_TRY(check_ptr(ubar, "ubar"));
  if (ubar == u_hat || ubar == v) return fail(HIT3D_ERR_INVALID, "ubar must not alias u_hat or v");
  size_t need;
  carve_vjp(*grid, nullptr, &need);
  H3_TRY(check_work(work, work_bytes, need));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    return rhs_vjp_impl<float>(*grid, static_cast<const cpx<float>*>(u_hat),
                               static_cast<const cpx<float>*>(v), static_cast<cpx<float>*>(ubar),
                               nu, forcing, work, s, param_grad, nullptr, dealiased != 0);
  return rhs_vjp_impl<double>(*grid, static_cast<const cpx<double>*>(u_hat),
                              static_cast<const cpx<double>*>(v), static_cast<cpx<double>*>(ubar),
                              nu, forcing, work, s, param_grad, nullptr, dealiased != 0);
}

int hit3d_rk4_step_vjp(const hit3d_grid* grid, const void* u_hat, const void* ubar_next,
                       void* ubar, double nu, double dt, double t, const hit3d_forcing* forcing,
                       int32_t dealiased, const void* stages, double* param_grad, void* work,
                       size_t work_bytes, void* stream) {
  (void)t;
  H3_TRY(check_grid(grid));
  H3_TRY(check_forcing(forcing));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(ubar_next, "ubar_next"));
  H3_TRY(check_ptr(ubar, "ubar"));
  if (stages) H3_TRY(check_ptr(stages, "stages"));
  if (ubar == u_hat) return fail(HIT3D_ERR_INVALID, "ubar must not alias u_hat");
  if (stages) {
    const char* lo = static_cast<const char*>(stages);
    const char* o = static_cast<const char*>(ubar);
    if (o + state_bytes(*grid) > lo && o < lo + 3 * state_bytes(*grid))
      return fail(HIT3D_ERR_INVALID, "ubar must not overlap stages");
  }
  if (!std::isfinite(dt) || !std::isfinite(nu))
    return fail(HIT3D_ERR_INVALID, "dt and nu must be finite");
  size_t need;
  carve_step_vjp(*grid, nullptr, &need);
  H3_TRY(check_work(work, work_bytes, need));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    return rk4_step_vjp_impl<float>(*grid, static_cast<const cpx<float>*>(u_hat),
                                    static_cast<const cpx<float>*>(ubar_next),
                                    static_cast<cpx<float>*>(ubar), nu, dt, forcing, work, s,
                                    param_grad, dealiased != 0, static_cast<const cpx<float>*>(stages));
  return rk4_step_vjp_impl<double>(*grid, static_cast<const cpx<double>*>(u_hat),
                                   static_cast<const cpx<double>*>(ubar_next),
                                   static_cast<cpx<double>*>(ubar), nu, dt, forcing, work, s,
                                   param_grad, dealiased != 0, static_cast<const cpx<double>*>(stages));
}

}  // extern "C"

// ---- learned subgrid closure (no reference counterpart) --------------------

namespace {

constexpr int kConvChunk = 1024;  // voxels per backward-weight partial

int check_conv(const hit3d_conv* c) {
  if (!c) return fail(HIT3D_ERR_INVALID, "conv is NULL");
  if (c->d < 1 || c->h < 1 || c->w < 16 || c->w % 16 != 0)
    return fail(HIT3D_ERR_INVALID, "conv grid needs w a multiple of 16 and d, h >= 1");
  const int64_t V = int64_t(c->d) * c->h * c->w;
  if (V % kConvChunk != 0) return fail(HIT3D_ERR_INVALID, "conv voxel count must be a multiple of 1024");
  if ((c->cin_pad != 16 && c->cin_pad != 32) || (c->cout_pad != 16 && c->cout_pad != 32))
    return fail(HIT3D_ERR_INVALID, "stored channel counts must be 16 or 32");
  if (c->cin < 1 || c->cin > c->cin_pad || c->cout < 1 || c->cout > c->cout_pad)
    return fail(HIT3D_ERR_INVALID, "channel counts must be in [1, stored channels]");
  return HIT3D_OK;
}

h3::ConvGeom conv_geom(const hit3d_conv* c) { return h3::ConvGeom{c->d, c->h, c->w}; }

size_t conv_work_bytes(const hit3d_conv* c) {
  const int64_t V = int64_t(c->d) * c->h * c->w, nchunk = V / kConvChunk;
  size_t b = round_up(V * c->cout_pad * 2, 256);                               // dP (bf16)
  b += round_up(nchunk * 27 * c->cout_pad * c->cin_pad * int64_t(sizeof(float)), 256);  // partials
  b += round_up(nchunk * 4 * c->cout_pad * int64_t(sizeof(float)), 256);              // bias partials
  return b;
}

}  // namespace

extern "C" {

int hit3d_conv_sizes(const hit3d_conv* conv, int64_t* kp, int64_t* kt, size_t* work_bytes) {
  H3_TRY(check_conv(conv));
  if (kp) *kp = h3::conv_kp(conv->cin_pad);
  if (kt) *kt = h3::conv_kt(conv->cout_pad);
  if (work_bytes) *work_bytes = conv_work_bytes(conv);
  return HIT3D_OK;
}

int hit3d_conv_pack(const hit3d_conv* conv, const float* w, void* wp, void* wt, void* stream) {
  H3_TRY(check_conv(conv));
  H3_TRY(check_ptr(w, "w"));
  H3_TRY(check_ptr(wp, "wp"));
  H3_TRY(check_ptr(wt, "wt"));
  H3_HIP(h3::launch_pack_weights(w, conv->cout, conv->cin, conv->cout_pad, conv->cin_pad,
                                 h3::conv_kp(conv->cin_pad), h3::conv_kt(conv->cout_pad),
                                 static_cast<__bf16*>(wp), static_cast<__bf16*>(wt),
                                 static_cast<hipStream_t>(stream)));
  return HIT3D_OK;
}

int hit3d_conv_forward(const hit3d_conv* conv, const void* x, const void* wp, const float* bias,
                       int32_t gelu, float* pre, void* act, float* out, void* stream) {
  H3_TRY(check_conv(conv));
  H3_TRY(check_ptr(x, "x"));
  H3_TRY(check_ptr(wp, "wp"));
  H3_TRY(check_ptr(bias, "bias"));
  if (gelu) {
    H3_TRY(check_ptr(act, "act"));   // pre may be NULL: not kept
  } else {
    H3_TRY(check_ptr(out, "out"));
  }
  H3_HIP(h3::launch_conv_fwd(conv->cin_pad, conv->cout_pad, static_cast<const __bf16*>(x),
                             static_cast<const __bf16*>(wp), h3::conv_kp(conv->cin_pad), bias,
                             conv->cout, conv_geom(conv), gelu ? 0 : 1, pre,
                             static_cast<__bf16*>(act), out, static_cast<hipStream_t>(stream)));
  return HIT3D_OK;
}

int hit3d_conv_backward(const hit3d_conv* conv, const void* x, const void* wt, const float* pre,
                        const float* grad_out, int32_t gelu, float* grad_x, float* grad_w,
                        float* grad_b, void* work, size_t work_bytes, void* stream) {
  H3_TRY(check_conv(conv));
  H3_TRY(check_ptr(x, "x"));
  H3_TRY(check_ptr(wt, "wt"));
  H3_TRY(check_ptr(grad_out, "grad_out"));
  H3_TRY(check_ptr(grad_w, "grad_w"));
  H3_TRY(check_ptr(grad_b, "grad_b"));
  if (gelu) H3_TRY(check_ptr(pre, "pre"));
  H3_TRY(check_work(work, work_bytes, conv_work_bytes(conv)));
  const int64_t V = int64_t(conv->d) * conv->h * conv->w, nchunk = V / kConvChunk;
  char* base = static_cast<char*>(work);
  __bf16* dP = reinterpret_cast<__bf16*>(base);
  float* part = reinterpret_cast<float*>(base + round_up(V * conv->cout_pad * 2, 256));
  float* pb = part + round_up(nchunk * 27 * conv->cout_pad * conv->cin_pad * 4, 256) / 4;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const h3::ConvGeom g = conv_geom(conv);
  if (gelu)
    H3_HIP(h3::launch_gelu_bwd(grad_out, pre, V * conv->cout_pad, conv->cout_pad, conv->cout, dP, s));
  else
    H3_HIP(h3::launch_to_bf16_padded(grad_out, V, conv->cout, conv->cout_pad, dP, s));
  if (grad_x)
    H3_HIP(h3::launch_conv_bwd_data(conv->cout_pad, conv->cin_pad, dP, static_cast<const __bf16*>(wt),
                                    h3::conv_kt(conv->cout_pad), g, grad_x, s));
  H3_HIP(h3::launch_conv_bwd_weight(conv->cout_pad, conv->cin_pad, dP, static_cast<const __bf16*>(x), g,
                                    kConvChunk, part, pb, s));
  H3_HIP(h3::launch_conv_reduce(part, pb, int(nchunk), conv->cout_pad, conv->cin_pad, conv->cout,
                                conv->cin, grad_w, grad_b, s));
  return HIT3D_OK;
}

int hit3d_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
               float beta2, float eps, int64_t t, void* stream) {
  H3_TRY(check_ptr(p, "p"));
  H3_TRY(check_ptr(g, "g"));
  H3_TRY(check_ptr(m, "m"));
  H3_TRY(check_ptr(v, "v"));
  if (n < 1 || t < 1) return fail(HIT3D_ERR_INVALID, "adam needs n >= 1 and step t >= 1");
  const float c1 = float(1.0 - std::pow(double(beta1), double(t)));
  const float c2 = float(1.0 - std::pow(double(beta2), double(t)));
  H3_HIP(h3::launch_adam(p, g, m, v, n, lr, beta1, beta2, eps, c1, c2, static_cast<hipStream_t>(stream)));
  return HIT3D_OK;
}

int hit3d_channels_pack(const float* soa, int32_t c, int64_t nvox, const float* scale, int32_t cpad,
                        void* aos, void* stream) {
  H3_TRY(check_ptr(soa, "soa"));
  H3_TRY(check_ptr(scale, "scale"));
  H3_TRY(check_ptr(aos, "aos"));
  if (c < 1 || c > cpad || nvox < 1) return fail(HIT3D_ERR_INVALID, "need 1 <= c <= cpad, nvox >= 1");
  H3_HIP(h3::launch_soa_aos(soa, c, nvox, scale, cpad, 0, static_cast<__bf16*>(aos), nullptr, nullptr,
                            static_cast<hipStream_t>(stream)));
  return HIT3D_OK;
}

int hit3d_channels_unpack(const float* aos, int32_t c, int32_t ld, int64_t nvox, const float* scale,
                          float* soa, void* stream) {
  H3_TRY(check_ptr(aos, "aos"));
  H3_TRY(check_ptr(scale, "scale"));
  H3_TRY(check_ptr(soa, "soa"));
  if (c < 1 || c > ld || nvox < 1) return fail(HIT3D_ERR_INVALID, "need 1 <= c <= ld, nvox >= 1");
  H3_HIP(h3::launch_soa_aos(nullptr, c, nvox, scale, ld, 1, nullptr, aos, soa,
                            static_cast<hipStream_t>(stream)));
  return HIT3D_OK;
}

int hit3d_stress_unpack_clip(const float* aos, int64_t nvox, const float* scale, const float* grad9,
                             float* tau, uint8_t* keep, void* stream) {
  H3_TRY(check_ptr(aos, "aos"));
  H3_TRY(check_ptr(scale, "scale"));
  H3_TRY(check_ptr(grad9, "grad9"));
  H3_TRY(check_ptr(tau, "tau"));
  if (!keep) return fail(HIT3D_ERR_INVALID, "keep is NULL");
  if (nvox < 1) return fail(HIT3D_ERR_INVALID, "nvox must be >= 1");
  H3_HIP(h3::launch_unpack_clip(aos, nvox, scale, grad9, tau, keep, static_cast<hipStream_t>(stream)));
  return HIT3D_OK;
}

int hit3d_spectral_truncate(const hit3d_grid* fine, const void* u_fine, const hit3d_grid* coarse,
                            void* u_coarse, int64_t ncomp, double kcut, void* stream) {
  H3_TRY(check_grid(fine));
  H3_TRY(check_grid(coarse));
  H3_TRY(check_ptr(u_fine, "u_fine"));
  H3_TRY(check_ptr(u_coarse, "u_coarse"));
  if (fine->precision != coarse->precision) return fail(HIT3D_ERR_INVALID, "precisions differ");
  if (coarse->nx > fine->nx || coarse->ny > fine->ny || coarse->nz > fine->nz)
    return fail(HIT3D_ERR_INVALID, "the coarse grid must not exceed the fine grid");
  if (fine->lx != coarse->lx || fine->ly != coarse->ly || fine->lz != coarse->lz)
    return fail(HIT3D_ERR_INVALID, "fine and coarse boxes differ");
  if (fine->lx != fine->ly || fine->lx != fine->lz)
    return fail(HIT3D_ERR_INVALID, "the spherical filter needs a cubic box");
  if (ncomp < 1 || ncomp > 64 || !(kcut >= 0)) return fail(HIT3D_ERR_INVALID, "bad ncomp or kcut");
  // integer frequencies n kept when |n| k0 <= kcut
  const double k0 = 2.0 * M_PI / fine->lx;
  const double kc2 = (kcut / k0) * (kcut / k0);
  const double scale = double(coarse->nx) * coarse->ny * coarse->nz / (double(fine->nx) * fine->ny * fine->nz);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (fine->precision == 4)
    H3_HIP(h3::launch_truncate<float>(static_cast<const cpx<float>*>(u_fine), int(fine->nx), int(fine->ny),
                                      int(fine->nz), static_cast<cpx<float>*>(u_coarse), int(coarse->nx),
                                      int(coarse->ny), int(coarse->nz), int(ncomp), kc2, float(scale), s));
  else
    H3_HIP(h3::launch_truncate<double>(static_cast<const cpx<double>*>(u_fine), int(fine->nx), int(fine->ny),
                                       int(fine->nz), static_cast<cpx<double>*>(u_coarse), int(coarse->nx),
                                       int(coarse->ny), int(coarse->nz), int(ncomp), kc2, scale, s));
  return HIT3D_OK;
}

int hit3d_velocity_products(const hit3d_grid* grid, const void* u, void* uu, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(u, "u"));
  H3_TRY(check_ptr(uu, "uu"));
  const int64_t V = grid->nx * grid->ny * grid->nz;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    H3_HIP(h3::launch_products<float>(static_cast<const float*>(u), V, static_cast<float*>(uu), s));
  else
    H3_HIP(h3::launch_products<double>(static_cast<const double*>(u), V, static_cast<double*>(uu), s));
  return HIT3D_OK;
}

int hit3d_les_combine(const hit3d_grid* grid, const void* du_hat, const void* tau_hat,
                      const void* shape, const float* gain, double kcut, void* out, void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(du_hat, "du_hat"));
  H3_TRY(check_ptr(out, "out"));
  if (tau_hat) {
    H3_TRY(check_ptr(tau_hat, "tau_hat"));
    if (!gain) return fail(HIT3D_ERR_INVALID, "gain is NULL");
  }
  if (grid->precision != 4) return fail(HIT3D_ERR_INVALID, "the closure runs on float32 grids");
  if (!(kcut >= 0.0)) return fail(HIT3D_ERR_INVALID, "kcut must be >= 0");
  const Geo g = make_geo(*grid, false);
  H3_HIP(h3::launch_les_combine<float>(static_cast<const cpx<float>*>(du_hat),
                                       static_cast<const cpx<float>*>(tau_hat),
                                       static_cast<const float*>(shape), gain, float(kcut * kcut),
                                       static_cast<cpx<float>*>(out), g,
                                       static_cast<hipStream_t>(stream)));
  return HIT3D_OK;
}

int hit3d_stress_divergence(const hit3d_grid* grid, const void* in, void* out, int32_t adjoint,
                            void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(in, "in"));
  H3_TRY(check_ptr(out, "out"));
  if (in == out) return fail(HIT3D_ERR_INVALID, "in and out must not alias");
  const Geo g = make_geo(*grid, false);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    H3_HIP(h3::launch_stress_div<float>(static_cast<const cpx<float>*>(in), static_cast<cpx<float>*>(out), g,
                                        adjoint ? 1 : 0, s));
  else
    H3_HIP(h3::launch_stress_div<double>(static_cast<const cpx<double>*>(in), static_cast<cpx<double>*>(out),
                                         g, adjoint ? 1 : 0, s));
  return HIT3D_OK;
}

int hit3d_rk4_combine(const hit3d_grid* grid, int32_t stage, double dt, const void* u_hat,
                      const void* k_hat, const double* hold_coef, const void* hold_us, void* acc,
                      void* out, void* stream) {
  H3_TRY(check_grid(grid));
  if (stage < 1 || stage > 4) return fail(HIT3D_ERR_INVALID, "stage must be 1, 2, 3 or 4");
  if (!std::isfinite(dt)) return fail(HIT3D_ERR_INVALID, "dt must be finite");
  if (stage < 4) H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(k_hat, "k_hat"));
  H3_TRY(check_ptr(acc, "acc"));
  H3_TRY(check_ptr(out, "out"));
  if (acc == out) return fail(HIT3D_ERR_INVALID, "acc and out must be different arrays");
  if ((hold_coef == nullptr) != (hold_us == nullptr))
    return fail(HIT3D_ERR_INVALID, "hold_coef and hold_us must both be given or both be NULL");
  const int64_t n = 3 * int64_t(grid->nx) * grid->ny * (grid->nz / 2 + 1);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    H3_HIP(h3::launch_rk4_combine<float>(static_cast<const cpx<float>*>(u_hat),
                                         static_cast<const cpx<float>*>(k_hat),
                                         static_cast<cpx<float>*>(acc), static_cast<cpx<float>*>(out),
                                         n, stage, dt, hold_coef, static_cast<const cpx<float>*>(hold_us),
                                         s));
  else
    H3_HIP(h3::launch_rk4_combine<double>(static_cast<const cpx<double>*>(u_hat),
                                          static_cast<const cpx<double>*>(k_hat),
                                          static_cast<cpx<double>*>(acc),
                                          static_cast<cpx<double>*>(out), n, stage, dt, hold_coef,
                                          static_cast<const cpx<double>*>(hold_us), s));
  return HIT3D_OK;
}

int hit3d_hold_energy(const hit3d_grid* grid, const void* u_hat, const void* du_hat, void* out,
                      double energy, double tau, double* coef, void* work, size_t work_bytes,
                      void* stream) {
  H3_TRY(check_grid(grid));
  H3_TRY(check_ptr(u_hat, "u_hat"));
  H3_TRY(check_ptr(du_hat, "du_hat"));
  H3_TRY(check_ptr(coef, "coef"));
  if (!(energy >= 0.0) || std::isnan(tau))
    return fail(HIT3D_ERR_INVALID, "energy must be >= 0 and tau a number");
  H3_TRY(check_work(work, work_bytes, reduce_bytes(*grid, nullptr, nullptr, nullptr)));
  double* partials;
  reduce_bytes(*grid, &partials, nullptr, work);
  const Geo g = make_geo(*grid, false);
  const int64_t fs = int64_t(g.nx) * g.ny * g.nzh;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (grid->precision == 4)
    H3_HIP(h3::launch_hold_partials<float>(static_cast<const cpx<float>*>(u_hat),
                                           static_cast<const cpx<float>*>(du_hat), fs, g, partials,
                                           s));
  else
    H3_HIP(h3::launch_hold_partials<double>(static_cast<const cpx<double>*>(u_hat),
                                            static_cast<const cpx<double>*>(du_hat), fs, g,
                                            partials, s));
  const double tiny = grid->precision == 4 ? double(FLT_EPSILON) : DBL_EPSILON;
  hipLaunchKernelGGL(k_hold_coef, dim3(1), dim3(256), 0, s, partials, h3::kReduceBlocks,
                     ntot2_inv(*grid), energy, tau, tiny, coef);
  H3_HIP(hipGetLastError());
  if (out == nullptr) return HIT3D_OK;   // coefficient only
  if (grid->precision == 4)
    H3_HIP(h3::launch_hold_apply<float>(static_cast<const cpx<float>*>(u_hat),
                                        static_cast<const cpx<float>*>(du_hat),
                                        static_cast<cpx<float>*>(out), 3 * fs, coef, s));
  else
    H3_HIP(h3::launch_hold_apply<double>(static_cast<const cpx<double>*>(u_hat),
                                         static_cast<const cpx<double>*>(du_hat),
                                         static_cast<cpx<double>*>(out), 3 * fs, coef, s));
  return HIT3D_OK;
}

}  // extern "C"

