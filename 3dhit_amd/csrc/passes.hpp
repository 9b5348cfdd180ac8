// Kernel argument blocks and host launcher declarations shared by the HIP
// translation units and the C ABI.
#pragma once
#include "common.hpp"

namespace h3 {

constexpr int kThreads = 256;   // strided passes: 4 waves
constexpr int kZThreads = 64;   // z passes: one wave, so the pencil exchanges need no cross-wave barrier

// Minimum waves per SIMD the compiler must allow: the 256-point float
// strided passes run 8-wave workgroups, which fit twice per CU only at 4
// waves per SIMD (<= 128 VGPRs); double needs twice the registers.
template <int N, class T>
struct MinWaves {
  static constexpr int value = (N == 256 && sizeof(T) == 4) ? 4 : 1;
};

template <int N>
struct StridedShape {
  using F = FftFor<N>;
  static constexpr int R = F::R, Q = F::Q;
  static constexpr int TC = N == 256 ? 16 : ((kThreads / Q) > 8 ? kThreads / Q : 8);  // columns per workgroup
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
// b = j % B); element (f, a, row, b) at f*fs + a*slab + row*B + b.  Input
// rows retained: in_k of N (compact map with in_p positives); output rows
// retained: out_k of N.  in_k == N / out_k == N is the identity map.
template <class T>
struct StridedArgs {
  const cpx<T>* in;
  cpx<T>* out;
  int64_t in_fs, out_fs, in_slab, out_slab;
  int B, J, nf;
  int in_k, in_p, out_k, out_p;
};

// Fused x-forward + spectral operator + RK4 stage (pass P5).
enum StageMode { kStageRhs = 0, kStage1 = 1, kStage2 = 2, kStage3 = 3, kStage4 = 4 };
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

template <class T>
hipError_t launch_strided(int N, int dir, const StridedArgs<T>& a, hipStream_t s);
template <class T>
hipError_t launch_x_inverse_curl(int N, const cpx<T>* us, int64_t us_fs, cpx<T>* out,
                                 int64_t out_fs, const Geo& g, hipStream_t s);
template <class T>
hipError_t launch_y_inverse_curl(int N, const cpx<T>* in, int64_t in_fs, cpx<T>* out,
                                 int64_t out_fs, const Geo& g, hipStream_t s);
template <class T>
hipError_t launch_z_fused(int N, cpx<T>* buf, int64_t fs, int64_t npencils, int kz, T scale,
                          hipStream_t s);
template <class T>
hipError_t launch_z_c2r(int N, const cpx<T>* in, int64_t in_fs, T* out, int64_t out_fs, int nf,
                        int64_t npencils, int kz, T scale, hipStream_t s);
template <class T>
hipError_t launch_z_r2c(int N, const T* in, int64_t in_fs, cpx<T>* out, int64_t out_fs, int nf,
                        int64_t npencils, int kz, hipStream_t s);
template <class T>
hipError_t launch_x_forward_stage(int N, int mode, const StageArgs<T>& a, hipStream_t s,
                                  int* nblocks_out);

}  // namespace h3
