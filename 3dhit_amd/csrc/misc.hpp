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
// partials[nblocks][nbins]; returns nblocks through *nblocks
template <class T>
hipError_t launch_spectrum_partials(const cpx<T>* st, int64_t fs, int component, const Geo& g,
                                    int nbins, double* partials, int* nblocks, hipStream_t s);
int spectrum_blocks(const Geo& g);
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

}  // namespace h3
