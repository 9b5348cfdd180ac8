// Kernel argument blocks and host launcher declarations shared by the HIP
// translation units and the C ABI.
#pragma once
#include "../../include/hit3d.h"
#include "common.hpp"

namespace h3 {

constexpr int kThreads = 256;   // strided passes: 4 waves
constexpr int kZThreads = 64;   // z passes: one wave, so the pencil exchanges need no cross-wave barrier
// first of the three pencil-buffer fields the fused z pass writes (u x w
// transformed, in place over consumed inputs)
constexpr int kZOut = 2;

// 1 selects the register-lean 1024-point float x-inverse and plain strided
// passes (passes.hip) wherever their 32-bit offsets fit
#ifndef H3_LEAN1024
#define H3_LEAN1024 1
#endif

// Minimum waves per SIMD the compiler must allow: the 256-point float
// strided passes run 8-wave workgroups, which fit twice per CU only at 4
// waves per SIMD (<= 128 VGPRs); double needs twice the registers.
template <int N, class T>
struct MinWaves {
  static constexpr int value = (N == 256 && sizeof(T) == 4) ? 4 : 1;
};

// Column-tiled passes: TC columns of Q threads per workgroup. At 128 points
// the plain strided and stage kernels take 8 columns (twice the workgroups
// of a 128^3 grid's few tiles), the x/y-inverse passes 16 (they split their
// outputs over gridDim.y instead).
template <int N, int TCW = (N == 256 ? 16 : N == 128 ? 8 : 0)>
struct StridedShape {
  using F = FftFor<N>;
  static constexpr int R = F::R, Q = F::Q;
  static constexpr int TC = TCW ? TCW : ((kThreads / Q) > 8 ? kThreads / Q : 8);  // columns per workgroup
  static constexpr int THREADS = TC * Q;
  static constexpr int LDS = TC * F::kSlots;
};
template <int N>
using CurlShape = StridedShape<N, (N == 256 || N == 128) ? 16 : 0>;
// Workgroup shape of the register-lean 1024-point passes (passes.hip): 16
// columns (one 128-byte line of complex64 per row) of Q threads.
template <int N>
struct LeanShape {
  using F = FftFor<N>;
  static constexpr int R = F::R, Q = F::Q, TC = 16;
  static constexpr int THREADS = TC * Q;
  static constexpr int LDS = TC * F::kSlots;
};
template <int N>
struct PencilBlock {
  using F = FftZFor<N>;
  static constexpr int R = F::R, Q = F::Q;
  static constexpr int PW = (kZThreads / Q) > 1 ? kZThreads / Q : 1;  // pencils per workgroup
  static constexpr int THREADS = PW * Q;
  static constexpr int LDS = PW * F::kSlots;
};

inline int ceil_div_i(int64_t a, int64_t b) { return int((a + b - 1) / b); }

inline int x_stage_blocks(int N, int J) {
  switch (N) {
    case 4: return ceil_div_i(J, StridedShape<4>::TC);
    case 8: return ceil_div_i(J, StridedShape<8>::TC);
    case 16: return ceil_div_i(J, StridedShape<16>::TC);
    case 32: return ceil_div_i(J, StridedShape<32>::TC);
    case 64: return ceil_div_i(J, StridedShape<64>::TC);
    case 128: return ceil_div_i(J, StridedShape<128>::TC);
    case 256: return ceil_div_i(J, StridedShape<256>::TC);
    case 512: return ceil_div_i(J, StridedShape<512>::TC);
    case 1024: return ceil_div_i(J, StridedShape<1024>::TC);
    default: return 0;
  }
}



// Strided pencil pass over [F][outer a][rows][B]: column j -> (a = j / B,
// b = j % B); element (f, a, row, b) at f*fs + a*slab + row*rs + b, the row
// stride rs = B unless in_rs / out_rs say otherwise (the y-forward reads
// y-major pencil buffers: slab B, row stride nx*B).  Input rows retained:
// in_k of N (compact map with in_p positives); output rows retained: out_k
// of N.  in_k == N / out_k == N is the identity map.
template <class T>
struct StridedArgs {
  const cpx<T>* in;
  cpx<T>* out;
  int64_t in_fs, out_fs, in_slab, out_slab;
  int B, J, nf;
  int in_k, in_p, out_k, out_p;
  int64_t in_rs, out_rs;   // row strides; 0: B
  __host__ __device__ int64_t in_row() const { return in_rs ? in_rs : B; }
  __host__ __device__ int64_t out_row() const { return out_rs ? out_rs : B; }
};

// Fused x-forward + spectral operator + RK4 stage (pass P5).
// Stage combinations of classic RK4 (1-4) and of the integrating-factor
// (Lawson) RK4 (5-8), whose stages apply exp(-nu k2 dt/2) and exp(-nu k2 dt)
// to the state and leave the viscous term out of the stage derivative.
enum StageMode { kStageRhs = 0, kStage1 = 1, kStage2 = 2, kStage3 = 3, kStage4 = 4,
                 kStageIf1 = 5, kStageIf2 = 6, kStageIf3 = 7, kStageIf4 = 8 };
enum ForcingTerm { kTermNone = 0, kTermUniform = 1, kTermBand = 2 };

template <class T>
struct StageArgs {
  const cpx<T>* nl;  // [3][Nx][jp], x-forward input (y- and z-transformed u x w)
  int64_t nl_fs;
  const cpx<T>* us;  // stage input state, compact [3][Kx][jp]
  const cpx<T>* u;   // step-start state (stages 1-3)
  cpx<T>* acc;       // RK4 accumulator (stages 1-4)
  cpx<T>* out;       // rhs: du; stages 1-3: next stage input; stage 4: new state
  int64_t st_fs;
  T nu, dt;
  int fterm;              // ForcingTerm
  T coef_const;           // used when coef_dev == nullptr
  const double* coef_dev; // device scalar written by the finalize kernel
  T kmin2, kmax2;         // band limits (dtype, as the reference compares them)
  double* partials;       // [gridDim.x][3] FP64 sums of the written state, or null
  Geo g;
};

// FP64 finalize: per-block partial sums [nblocks][3] of (w|u|^2, w k2|u|^2,
// band w|u|^2) -> scalars [coef, 2E, 2 Omega, 2 E_band] (normalised by
// 1/Ntot^2) and the forcing coefficient of the summed state.
struct FinalizeArgs {
  int kind;        // HIT3D_FORCING_*
  double eps;      // band injection
  double nu, energy, tau;
  double tiny;     // finfo(dtype).eps
  double inv_ntot2;
};

// coefficient rule of the forcings (forcing.py:109-163 band, forcing.py:5-20
// constant energy) from the three totals
__device__ __forceinline__ void finalize_write(const FinalizeArgs& fa, double sE, double sO,
                                               double sB, double* __restrict__ scalars) {
  const double twoE = sE * fa.inv_ntot2, twoO = sO * fa.inv_ntot2, twoB = sB * fa.inv_ntot2;
  double coef = 0.0;
  if (fa.kind == HIT3D_FORCING_BAND) {
    coef = twoB > fa.tiny ? fa.eps / twoB : 0.0;
  } else if (fa.kind == HIT3D_FORCING_LUNDGREN) {
    if (twoE > fa.tiny) {
      double rate = fa.nu * twoO;
      if (fa.tau > 0.0 && isfinite(fa.tau)) rate -= (0.5 * twoE - fa.energy) / fa.tau;
      coef = rate / twoE;
    }
  }
  scalars[0] = coef;
  scalars[1] = twoE;
  scalars[2] = twoO;
  scalars[3] = twoB;
}

// A finalize carried out by block 0 of the next pass (P1 of the next RK4
// stage), which starts after the stage kernel that wrote the partials has
// finished: no separate one-block launch between the passes.
struct FinalizeJob {
  const double* partials;  // null: nothing to do
  int nblocks;
  FinalizeArgs fa;
  double* scalars;
};

// Whole-block fixed-order reduction (blockDim a multiple of 64, at most 1024):
// Neumaier sums per thread over a strided set of blocks, a shuffle tree per
// wave, then the waves in order.
__device__ __forceinline__ void finalize_in_block(const FinalizeJob& j) {
  __shared__ double wsum[3][16];
  KSum acc[3];
  for (int i = threadIdx.x; i < j.nblocks; i += blockDim.x)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c].add(j.partials[3 * i + c]);
  double v[3] = {acc[0].value(), acc[1].value(), acc[2].value()};
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] += __shfl_xor(v[c], off);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c) wsum[c][threadIdx.x / 64] = v[c];
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[3] = {0.0, 0.0, 0.0};
    for (int w = 0; w < int(blockDim.x / 64); ++w)
      for (int c = 0; c < 3; ++c) t[c] += wsum[c][w];
    finalize_write(j.fa, t[0], t[1], t[2], j.scalars);
  }
}

template <class T>
hipError_t launch_strided(int N, int dir, const StridedArgs<T>& a, hipStream_t s);
template <class T>
hipError_t launch_x_inverse_curl(int N, const cpx<T>* us, int64_t us_fs, cpx<T>* out,
                                 int64_t out_fs, const Geo& g, hipStream_t s,
                                 const FinalizeJob* job = nullptr);
// out: the pencil buffer in the layout launch_z_fused reads
template <class T>
hipError_t launch_y_inverse_curl(int N, const cpx<T>* in, int64_t in_fs, cpx<T>* out,
                                 int64_t out_fs, const Geo& g, int yrow, hipStream_t s);
// The forward RHS's pencil buffer (y-inverse out, z pass in/out, y-forward
// in) is x-major, pencil (x, y) at (x ny + y) kzs, or, where pencil_row
// returns a row stride yrow > 0, y-major, pencil (x, y) at y yrow + x kzs
// (yrow = nx kzs + a pad of 32 values, so that rows yrow apart do not all
// fall in the same few cache sets and memory channels): a y-pass column j = x kzs + iz then has its rows yrow
// apart, and the 16 columns of a workgroup tile cover whole 128-byte lines of
// every row even when kzs complex values are not a whole number of lines
// (kzs = 88 at N = 256). A y-major field spans ny yrow elements.
template <class T>
int pencil_row(const Geo& g);
template <class T>
hipError_t launch_z_fused(int N, cpx<T>* buf, int64_t fs, int nx, int ny, int kz, int kzs,
                          int yrow, T scale, hipStream_t s);
template <class T>
hipError_t launch_z_c2r(int N, const cpx<T>* in, int64_t in_fs, T* out, int64_t out_fs, int nf,
                        int64_t npencils, int kz, T scale, hipStream_t s);
template <class T>
hipError_t launch_z_r2c(int N, const T* in, int64_t in_fs, cpx<T>* out, int64_t out_fs, int nf,
                        int64_t npencils, int kz, hipStream_t s);
// adjoint z pass: nine fields in (pencils of kz modes kzs apart), c2r,
// w x g and g x u, r2c into out fields 0..5 (pencils of nzh modes)
template <class T>
hipError_t launch_z_adj(int N, const cpx<T>* in, int64_t in_fs, int kz, int kzs, cpx<T>* out,
                        int64_t out_fs, int nzh, int64_t npencils, T scale, bool pruned,
                        hipStream_t s);
template <class T>
hipError_t launch_x_forward_stage(int N, int mode, const StageArgs<T>& a, hipStream_t s,
                                  int* nblocks_out);

}  // namespace h3
