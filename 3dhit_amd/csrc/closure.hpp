// Launchers of the learned-closure kernels (closure.hip).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace h3 {

struct ConvGeom {
  int D, H, W;  // voxels per axis; v = (a*H + b)*W + c; W % 16 == 0 and D*H*W % 64 == 0
};

// padded K extents of the packed weights
inline int conv_kp(int cin) { return ((27 * cin + 31) / 32) * 32; }
inline int conv_kt(int cout) { return ((27 * cout + 31) / 32) * 32; }

hipError_t launch_conv_fwd(int cin, int cout, const __bf16* X, const __bf16* Wp, int KP,
                           const float* bias, int cout_real, const ConvGeom& g, int mode, float* pre,
                           __bf16* act, float* out, hipStream_t s);
hipError_t launch_conv_bwd_data(int cout, int cin, const __bf16* dP, const __bf16* WT, int KT,
                                const ConvGeom& g, float* dX, hipStream_t s);
hipError_t launch_conv_bwd_weight(int cout, int cin, const __bf16* dP, const __bf16* X,
                                  const ConvGeom& g, int chunk, float* part, float* pb,
                                  hipStream_t s);
hipError_t launch_conv_reduce(const float* part, const float* pb, int nchunk, int cout, int cin,
                              int cout_real, int cin_real, float* dW, float* db, hipStream_t s);
hipError_t launch_gelu_bwd(const float* dA, const float* pre, int64_t n, int C, int creal, __bf16* dP,
                           hipStream_t s);
hipError_t launch_to_bf16_padded(const float* src, int64_t nvox, int creal, int C, __bf16* dst,
                                 hipStream_t s);
hipError_t launch_pack_weights(const float* w, int cout_real, int cin_real, int cout, int cin, int KP,
                               int KT, __bf16* Wp, __bf16* WT, hipStream_t s);
hipError_t launch_adam(float* p, const float* gr, float* m, float* v, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t s);
hipError_t launch_soa_aos(const float* soa, int C, int64_t V, const float* scale, int CP, int dir,
                          __bf16* aos_bf, const float* aos_f, float* soa_out, hipStream_t s);
hipError_t launch_unpack_clip(const float* aos, int64_t V, const float* scale, const float* g, float* tau,
                              uint8_t* keep, hipStream_t s);
template <class T>
hipError_t launch_truncate(const cpx<T>* fine, int fx, int fy, int fz, cpx<T>* coarse, int cx, int cy,
                           int cz, int ncomp, double kc2, T scale, hipStream_t s);
template <class T>
hipError_t launch_products(const T* u, int64_t V, T* uu, hipStream_t s);
template <class T>
hipError_t launch_stress_div(const cpx<T>* in, cpx<T>* out, const Geo& g, int dir, hipStream_t s);
template <class T>
hipError_t launch_les_combine(const cpx<T>* du, const cpx<T>* tau, const T* shape, const float* gain,
                              T kcut2, cpx<T>* out, const Geo& g, hipStream_t s);

}  // namespace h3
