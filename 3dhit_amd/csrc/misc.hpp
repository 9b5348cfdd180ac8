// Launchers for the non-FFT kernels: layout changes, FP64 reductions,
// shell spectrum, initial-condition noise, elementwise spectral operators.
#pragma once
#include "common.hpp"

namespace h3 {

constexpr int kReduceBlocks = 1024;  // fixed partition => deterministic sums

enum SpectralOp { kOpDdx = 0, kOpDdy = 1, kOpDdz = 2, kOpLaplacian = 3 };

template <class T>
hipError_t launch_gather(const cpx<T>* full, int64_t full_fs, cpx<T>* comp, int64_t comp_fs,
                         int nf, const Geo& g, hipStream_t s);
template <class T>
hipError_t launch_scatter(const cpx<T>* comp, int64_t comp_fs, cpx<T>* full, int64_t full_fs,
                          int nf, const Geo& g, hipStream_t s);
// partials[kReduceBlocks][3] of (sum w|u|^2, sum w k2|u|^2, sum over band of w|u|^2)
template <class T>
hipError_t launch_quad_partials(const cpx<T>* st, int64_t fs, const Geo& g, T kmin2, T kmax2,
                                double* partials, hipStream_t s);
// partials[kReduceBlocks][3]: (sum of |u|^2 outside the 2/3 box, 0, 0) over a full-layout state
template <class T>
hipError_t launch_alias_partials(const cpx<T>* full, int64_t fs, const Geo& g, double* partials,
                                 hipStream_t s);
// partials[nblocks][nbins] of a full-layout state (box = 1: only the modes
// inside the 2/3 box are read); returns nblocks through *nblocks
template <class T>
hipError_t launch_spectrum_partials(const cpx<T>* st, int64_t fs, int component, const Geo& g,
                                    int box, int nbins, double* partials, int* nblocks,
                                    hipStream_t s);
int spectrum_blocks(const Geo& g);  // the largest count either read takes
// out[b] = scale * (fixed-order sum over the nrows partial rows of bin b);
// mid holds spectrum_mid_doubles(nbins) doubles
size_t spectrum_mid_doubles(int nbins);
hipError_t launch_spectrum_merge(const double* partials, int nrows, int nbins, double* mid,
                                 double scale, double* out, hipStream_t s);
// dst = src, n16 16-byte words (both 16-byte aligned); nontemporal: streaming hints
hipError_t launch_stream_copy(const void* src, void* dst, int64_t n16, int nontemporal,
                              hipStream_t s);
template <class T>
hipError_t launch_philox_uniform(T* out, int64_t n, uint64_t seed, hipStream_t s);
template <class T>
hipError_t launch_dealias_project(const cpx<T>* in, cpx<T>* out, int64_t fs, int nf,
                                  const Geo& g, int do_dealias, int do_project, hipStream_t s);
template <class T>
hipError_t launch_shell_scale(cpx<T>* st, int64_t fs, int nf, const Geo& g, const double* c,
                              int nbins, hipStream_t s);
template <class T>
hipError_t launch_spectral_op(const cpx<T>* in, cpx<T>* out, int64_t fs, int nf, const Geo& g,
                              int op, hipStream_t s);
template <class T>
hipError_t launch_tables(const Geo& g, T* k2, T* inv_k2, uint8_t* mask, int32_t* shell,
                         hipStream_t s);

// ---- adjoint building blocks (full layout, dense Geo) ----
// S[0..2] = u, S[3..5] = i k x u, S[6..8] = (P M v) * (ntot / w)
template <class T>
hipError_t launch_adj_prep(const cpx<T>* u, const cpx<T>* v, cpx<T>* S, int64_t fs, const Geo& g,
                           double ntot, hipStream_t s);
// the same nine fields in the compact layout of pruned Geo g (u without
// content outside the 2/3 box)
template <class T>
hipError_t launch_adj_prep_compact(const cpx<T>* u, const cpx<T>* v, int64_t full_fs, cpx<T>* S,
                                   int64_t comp_fs, const Geo& g, double ntot, hipStream_t s);
// partials[kReduceBlocks][3]: (sum Re(conj(x) P M y) [over the band if use_band], 0, 0)
template <class T>
hipError_t launch_dot_partials(const cpx<T>* x, const cpx<T>* y, int64_t fs, const Geo& g,
                               T kmin2, T kmax2, int use_band, double* partials, hipStream_t s);
// pgrad[0..2] (device) += d/dnu, d/dA (linear forcing), d/deps (band forcing)
// of Re sum conj(kb) rhs(y), both full-layout [3] states; two_e_band (device)
// holds y's band 2E (read for the band forcing only); partials: kReduceBlocks*3
template <class T>
hipError_t launch_param_grad(const cpx<T>* kb, const cpx<T>* y, int64_t fs, const Geo& g, T kmin2,
                             T kmax2, int kind, const double* two_e_band, double tiny,
                             double* partials, double* pgrad, hipStream_t s);
// partials[kReduceBlocks][3]: (sum w Re(conj(u) du), sum w |u|^2, 0) over full-layout states,
// w the Hermitian weight of the mode's kz column
template <class T>
hipError_t launch_hold_partials(const cpx<T>* u, const cpx<T>* du, int64_t fs, const Geo& g,
                                double* partials, hipStream_t s);
// out = du + coef[0] u over n complex values (coef on the device)
template <class T>
hipError_t launch_hold_apply(const cpx<T>* u, const cpx<T>* du, cpx<T>* out, int64_t n,
                             const double* coef, hipStream_t s);
// one classic-RK4 combination step over n complex values (stages 1..4)
template <class T>
hipError_t launch_rk4_combine(const cpx<T>* u, const cpx<T>* k, cpx<T>* acc, cpx<T>* out, int64_t n,
                              int stage, double dt, const double* hold_coef, const cpx<T>* hold_us,
                              hipStream_t s);
struct AdjCombine {
  int fkind;          // HIT3D_FORCING_*
  double a;           // LINEAR coefficient
  double nu;          // viscosity of the rhs
  double fnu, tau;    // LUNDGREN parameters
  double inv_ntot, inv_ntot2;
  double tiny;
  const double* scalars;  // [coef, S_E, S_Omega, S_band] of the state (normalised by 1/Ntot^2)
  const double* dot;      // <band u, vp> (BAND) or <vp, u> (LUNDGREN)
};
// Where the combined value goes: with acc set, acc = acc_base + value (else
// out = value); with kbn set also kbn = alpha ub + beta value (the next RK4
// stage's cotangent). kbn may alias v, acc may alias acc_base.
template <class T>
struct AdjEpilogue {
  cpx<T>* acc = nullptr;
  const cpx<T>* acc_base = nullptr;
  cpx<T>* kbn = nullptr;
  const cpx<T>* ub = nullptr;
  T alpha = T(0), beta = T(0);
};
// value = (w/Ntot)(A + i k x B) - nu k2 v + forcing adjoint (P M v formed per mode)
template <class T>
hipError_t launch_adj_combine(const cpx<T>* AB, const cpx<T>* v, const cpx<T>* u, cpx<T>* out,
                              int64_t fs, const Geo& g, T kmin2, T kmax2, const AdjCombine& c,
                              const AdjEpilogue<T>& ep, hipStream_t s);
// forcing term over the full layout: out = c u, or c [kmin2 <= k2 <= kmax2] u
// when band; c = *coef_dev when coef_dev is set, else coef_const
template <class T>
hipError_t launch_forcing_term(const cpx<T>* u, cpx<T>* out, int64_t fs, const Geo& g, int band,
                               T coef_const, const double* coef_dev, T kmin2, T kmax2, hipStream_t s);
// out = a x + b y over n complex values
template <class T>
hipError_t launch_axpby(cpx<T>* out, T a, const cpx<T>* x, T b, const cpx<T>* y, int64_t n,
                        hipStream_t s);

}  // namespace h3
