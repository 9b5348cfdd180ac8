// Non-FFT kernels. Full-layout kernels take a dense Geo (kx = nx, ky = ny,
// kz = nz/2+1, jp = ny*kz), under which the compact layout is the API layout.
#include <cmath>

#include "../../include/hit3d.h"
#include "misc.hpp"

namespace h3 {

constexpr int kBlock = 256;

static inline unsigned nblocks_for(int64_t n, int per = kBlock) {
  return unsigned((n + per - 1) / per);
}

// Deterministic block sum of three doubles (blockDim == 256).
__device__ __forceinline__ void block_sum3m(double a, double b, double c, double* dst) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
    c += __shfl_xor(c, off);
  }
  __shared__ double sh[3][kBlock / 64];
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = a;
    sh[1][w] = b;
    sh[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < 3; ++i) dst[i] = ((sh[i][0] + sh[i][1]) + sh[i][2]) + sh[i][3];
  }
}

// ---- full <-> compact ------------------------------------------------------

template <class T>
__global__ __launch_bounds__(kBlock) void k_gather(const cpx<T>* __restrict__ full, int64_t full_fs,
                                                   cpx<T>* __restrict__ comp, int64_t comp_fs,
                                                   int nf, Geo g) {
  // every compact slot, the padding (iz >= kz) included: padding gets zeros
  const int64_t n = int64_t(g.kx) * g.jp;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  const int ix = int(idx / g.jp);
  const int j = int(idx - int64_t(ix) * g.jp);
  const int iy = j / g.kzs, iz = j - iy * g.kzs;
  const bool mode = iy < g.ky && iz < g.kz;   // not a padding slot
  const int x = row_of_compact(ix, g.nx, g.kx, g.px), y = row_of_compact(iy, g.ny, g.ky, g.py);
  const int64_t src = mode ? (int64_t(x) * g.ny + y) * g.nzh + iz : 0;   // in range either way
  for (int f = 0; f < nf; ++f) comp[f * comp_fs + idx] = mode ? full[f * full_fs + src] : czero<T>();
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_scatter(const cpx<T>* __restrict__ comp, int64_t comp_fs,
                                                    cpx<T>* __restrict__ full, int64_t full_fs,
                                                    int nf, Geo g) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  const int iz = int(idx % g.nzh);
  const int64_t xy = idx / g.nzh;
  const int y = int(xy % g.ny), x = int(xy / g.ny);
  const int ix = compact_of_row(x, g.nx, g.kx, g.px), iy = compact_of_row(y, g.ny, g.ky, g.py);
  const bool kept = ix >= 0 && iy >= 0 && iz < g.kz;
  const int64_t src = int64_t(ix) * g.jp + int64_t(iy) * g.kzs + iz;
  for (int f = 0; f < nf; ++f) full[f * full_fs + idx] = kept ? comp[f * comp_fs + src] : czero<T>();
}

// ---- FP64 quadratic sums ---------------------------------------------------

template <class T>
__global__ __launch_bounds__(kBlock) void k_quad_partials(const cpx<T>* __restrict__ st, int64_t fs,
                                                          Geo g, T kmin2, T kmax2,
                                                          double* __restrict__ partials) {
  const int64_t J = int64_t(g.ky) * g.kz;
  const int64_t n = int64_t(g.kx) * J;
  double sE = 0.0, sO = 0.0, sB = 0.0;
  for (int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * kBlock) {
    const int ix = int(idx / J);
    const int j = int(idx - int64_t(ix) * J);
    const int iy = j / g.kz, iz = j - iy * g.kz;
    const int64_t o = int64_t(ix) * g.jp + int64_t(iy) * g.kzs + iz;
    double e = 0.0;
    for (int f = 0; f < 3; ++f) {
      const cpx<T> v = st[f * fs + o];
      e += double(v.x) * v.x + double(v.y) * v.y;
    }
    const T k2 = k2_of(wavenumber<T>(g.k0x, freq_of(ix, g.kx, g.px)),
                       wavenumber<T>(g.k0y, freq_of(iy, g.ky, g.py)), wavenumber<T>(g.k0z, iz));
    const double we = hermitian_weight(iz, g.nz) * e;
    sE += we;
    sO += we * double(k2);
    if (kmin2 <= k2 && k2 <= kmax2) sB += we;
  }
  block_sum3m(sE, sO, sB, partials + 3 * blockIdx.x);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_alias_partials(const cpx<T>* __restrict__ full,
                                                           int64_t fs, Geo g,
                                                           double* __restrict__ partials) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  double s = 0.0;
  for (int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * kBlock) {
    const int iz = int(idx % g.nzh);
    const int64_t xy = idx / g.nzh;
    const int y = int(xy % g.ny), x = int(xy / g.ny);
    const int nx = x < (g.nx + 1) / 2 ? x : x - g.nx;
    const int ny = y < (g.ny + 1) / 2 ? y : y - g.ny;
    if (abs(nx) <= g.cx && abs(ny) <= g.cy && iz <= g.cz) continue;
    for (int f = 0; f < 3; ++f) {
      const cpx<T> v = full[f * fs + idx];
      s += double(v.x) * v.x + double(v.y) * v.y;
    }
  }
  block_sum3m(s, 0.0, 0.0, partials + 3 * blockIdx.x);
}

// ---- shell spectrum ----------------------------------------------------------
// Each wave walks whole kz lines of the full-layout state. All loads of a
// line (up to kSpecChunks chunks of 64 modes, every component) are issued
// before any of them is used, so a wave keeps a whole line in flight. Along
// a line the shell index is non-decreasing in kz, so a segmented suffix sum
// across the 64 lanes gives each shell's run total in its head lane, which
// adds it to the wave's own LDS histogram (Neumaier sum and compensation):
// one writer per bin at a time, in a fixed order, so the result is
// bit-reproducible. The waves of a block merge in order into one row of
// partials; two fixed-order merge kernels reduce the rows.
//
// `box` = 1 reads only the modes inside the per-axis 2/3 box (the caller
// promises that every other mode is zero: the solver's dealiased states),
// about 0.3 of the full array.

constexpr int kSpecWaves = kBlock / 64;
constexpr int kSpecChunks = 9;        // ceil((1024/2 + 1) / 64)
constexpr int kSpecGroups = 32;       // row groups of the first merge level

struct SpecLines {
  int lx, ly, lz;   // lines along x and y, modes per line
  int px, py;       // positive-frequency counts of the retained x / y rows
};

__host__ __device__ inline SpecLines spec_lines(const Geo& g, int box) {
  SpecLines L;
  if (box) {
    L.lx = 2 * g.cx + 1;
    L.ly = 2 * g.cy + 1;
    L.lz = g.cz + 1;
  } else {
    L.lx = g.nx;
    L.ly = g.ny;
    L.lz = g.nzh;
  }
  L.px = (L.lx + 1) / 2;
  L.py = (L.ly + 1) / 2;
  return L;
}

int spectrum_blocks(const Geo& g, int box) {
  const SpecLines L = spec_lines(g, box);
  const int64_t lines = int64_t(L.lx) * L.ly;
  const int64_t b = (lines + kSpecWaves - 1) / kSpecWaves;
  return int(b < kReduceBlocks ? b : kReduceBlocks);
}

int spectrum_blocks(const Geo& g) { return spectrum_blocks(g, 0); }

template <class T>
__global__ __launch_bounds__(kBlock) void k_spectrum_partials(const cpx<T>* __restrict__ st,
                                                              int64_t fs, int component, Geo g,
                                                              int box, int nbins, int nrow,
                                                              double* __restrict__ partials) {
  // [kSpecWaves][nbins][2] (sum, compensation): nbins covers every shell the
  // read modes reach; the partial rows are nrow >= nbins long (zero beyond)
  extern __shared__ double hist[];
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < kSpecWaves * nbins * 2; i += kBlock) hist[i] = 0.0;
  __syncthreads();
  double* h = hist + wave * nbins * 2;
  const SpecLines L = spec_lines(g, box);
  const int64_t lines = int64_t(L.lx) * L.ly;
  const int f0 = component < 0 ? 0 : component, nf = component < 0 ? 3 : 1;
  const int nch = (L.lz + 63) / 64;
  for (int64_t l = int64_t(blockIdx.x) * kSpecWaves + wave; l < lines;
       l += int64_t(gridDim.x) * kSpecWaves) {
    const int ix = int(l / L.ly), iy = int(l - int64_t(ix) * L.ly);
    const int x = box ? row_of_compact(ix, g.nx, L.lx, L.px) : ix;
    const int y = box ? row_of_compact(iy, g.ny, L.ly, L.py) : iy;
    const T kxv = wavenumber<T>(g.k0x, x < (g.nx + 1) / 2 ? x : x - g.nx);
    const T kyv = wavenumber<T>(g.k0y, y < (g.ny + 1) / 2 ? y : y - g.ny);
    const cpx<T>* line = st + (int64_t(x) * g.ny + y) * g.nzh + int64_t(f0) * fs;
    // every load of the line first
    cpx<T> u[kSpecChunks][3];
#pragma unroll
    for (int c = 0; c < kSpecChunks; ++c) {
      const int iz = 64 * c + lane;
      const bool ok = c < nch && iz < L.lz;
#pragma unroll
      for (int f = 0; f < 3; ++f)
        u[c][f] = (ok && f < nf) ? line[int64_t(f) * fs + iz] : czero<T>();
    }
#pragma unroll
    for (int c = 0; c < kSpecChunks; ++c) {
      if (c < nch) {  // uniform across the wave
        const int iz = 64 * c + lane;
        double v = 0.0;
        int key = 0x7fffffff;
        if (iz < L.lz) {
#pragma unroll
          for (int f = 0; f < 3; ++f) v += double(u[c][f].x) * u[c][f].x + double(u[c][f].y) * u[c][f].y;
          v *= hermitian_weight(iz, g.nz);
          key = shell_of<T>(k2_of(kxv, kyv, wavenumber<T>(g.k0z, iz)), g.k0x);
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const double ov = __shfl_down(v, off);
          const int ok = __shfl_down(key, off);
          if (lane + off < 64 && ok == key) v += ov;
        }
        const int prev = __shfl_up(key, 1);
        const bool head = lane == 0 || prev != key;
        if (head && key < nbins) {
          // Neumaier update of bin `key`
          const double s = h[2 * key], t = s + v;
          h[2 * key + 1] += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
          h[2 * key] = t;
        }
      }
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nrow; b += kBlock) {
    KSum acc;
    if (b < nbins)
      for (int w = 0; w < kSpecWaves; ++w) {
        acc.add(hist[(w * nbins + b) * 2]);
        acc.add(hist[(w * nbins + b) * 2 + 1]);
      }
    partials[int64_t(blockIdx.x) * nrow + b] = acc.value();
  }
}

// First merge level: block (bx, gy) sums rows gy, gy + kSpecGroups, ... of
// bins [64 bx, 64 bx + 64): each wave takes every kSpecWaves-th of those
// rows (Neumaier), the waves combine in order, and the group's (sum,
// compensation) pair goes to mid[gy][bin].
__global__ __launch_bounds__(kBlock) void k_spectrum_merge_rows(const double* __restrict__ partials,
                                                                int nrows, int nbins,
                                                                double* __restrict__ mid) {
  __shared__ double sh[kSpecWaves][64][2];
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  const int b = blockIdx.x * 64 + lane;
  const int gy = blockIdx.y;
  KSum acc;
  if (b < nbins)
    for (int r = gy + kSpecGroups * wave; r < nrows; r += kSpecGroups * kSpecWaves)
      acc.add(partials[int64_t(r) * nbins + b]);
  sh[wave][lane][0] = acc.s;
  sh[wave][lane][1] = acc.c;
  __syncthreads();
  if (wave == 0 && b < nbins) {
    KSum t;
    for (int w = 0; w < kSpecWaves; ++w) {
      t.add(sh[w][lane][0]);
      t.add(sh[w][lane][1]);
    }
    mid[(int64_t(gy) * nbins + b) * 2] = t.s;
    mid[(int64_t(gy) * nbins + b) * 2 + 1] = t.c;
  }
}

// Second level: one thread per bin adds the kSpecGroups pairs in order.
__global__ __launch_bounds__(kBlock) void k_spectrum_merge_groups(const double* __restrict__ mid,
                                                                  int nbins, double scale,
                                                                  double* __restrict__ out) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= nbins) return;
  KSum t;
  for (int gy = 0; gy < kSpecGroups; ++gy) {
    t.add(mid[(int64_t(gy) * nbins + b) * 2]);
    t.add(mid[(int64_t(gy) * nbins + b) * 2 + 1]);
  }
  out[b] = t.value() * scale;
}

size_t spectrum_mid_doubles(int nbins) { return size_t(kSpecGroups) * nbins * 2; }

hipError_t launch_spectrum_merge(const double* partials, int nrows, int nbins, double* mid,
                                 double scale, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_spectrum_merge_rows, dim3((nbins + 63) / 64, kSpecGroups), dim3(kBlock), 0, s,
                     partials, nrows, nbins, mid);
  hipLaunchKernelGGL(k_spectrum_merge_groups, dim3((nbins + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     mid, nbins, scale, out);
  return hipGetLastError();
}

// ---- streaming copy (bandwidth reference) --------------------------------------
// dst = src for n16 16-byte words, U words per thread with all loads before
// the stores; NT marks both as non-temporal (streaming). The measured rate
// is the attainable HBM read + write bandwidth the pass kernels compare to.

typedef unsigned int StreamWord __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ __launch_bounds__(kBlock) void k_stream_copy(const StreamWord* __restrict__ src,
                                                        StreamWord* __restrict__ dst, int64_t n16) {
  const int64_t base = int64_t(blockIdx.x) * U * kBlock + threadIdx.x;
  StreamWord v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i = base + int64_t(u) * kBlock;
    if (i < n16) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i = base + int64_t(u) * kBlock;
    if (i < n16) {
      if constexpr (NT) __builtin_nontemporal_store(v[u], dst + i);
      else dst[i] = v[u];
    }
  }
}

hipError_t launch_stream_copy(const void* src, void* dst, int64_t n16, int nontemporal,
                              hipStream_t s) {
  if (n16 <= 0) return hipSuccess;
  const auto* in = static_cast<const StreamWord*>(src);
  auto* out = static_cast<StreamWord*>(dst);
  if (nontemporal) {
    const int64_t nb = (n16 + 4 * kBlock - 1) / (4 * kBlock);
    hipLaunchKernelGGL((k_stream_copy<4, true>), dim3(unsigned(nb)), dim3(kBlock), 0, s, in, out, n16);
  } else {
    const int64_t nb = (n16 + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((k_stream_copy<1, false>), dim3(unsigned(nb)), dim3(kBlock), 0, s, in, out, n16);
  }
  return hipGetLastError();
}

// ---- Philox4x64-10 uniform noise -------------------------------------------
// value i = word i%4 of Philox4x64-10(counter = (i/4 + 1, 0, 0, 0), key = (seed, 0)),
// mapped to U[-0.5, 0.5): (x >> 40) 2^-24 - 0.5 (float) or (x >> 11) 2^-53 - 0.5 (double).

template <class T>
__global__ __launch_bounds__(kBlock) void k_philox_uniform(T* __restrict__ out, int64_t n,
                                                           uint64_t seed) {
  const int64_t b = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (4 * b >= n) return;
  uint64_t c0 = uint64_t(b) + 1, c1 = 0, c2 = 0, c3 = 0, k0 = seed, k1 = 0;
  constexpr uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  constexpr uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t lo0 = M0 * c0, hi0 = __umul64hi(M0, c0);
    const uint64_t lo1 = M1 * c2, hi1 = __umul64hi(M1, c2);
    const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  const uint64_t words[4] = {c0, c1, c2, c3};
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const int64_t i = 4 * b + w;
    if (i < n) {
      if constexpr (sizeof(T) == 4) {
        out[i] = float(uint32_t(words[w] >> 40)) * 0x1p-24f - 0.5f;
      } else {
        out[i] = double(words[w] >> 11) * 0x1p-53 - 0.5;
      }
    }
  }
}

// ---- elementwise spectral operators (full layout) ----------------------------

struct ModeIdx {
  int nx, ny, nz;
};
__device__ __forceinline__ ModeIdx mode_of(int64_t idx, const Geo& g) {
  const int iz = int(idx % g.nzh);
  const int64_t xy = idx / g.nzh;
  const int y = int(xy % g.ny), x = int(xy / g.ny);
  return {x < (g.nx + 1) / 2 ? x : x - g.nx, y < (g.ny + 1) / 2 ? y : y - g.ny, iz};
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_dealias_project(const cpx<T>* in, cpx<T>* out,
                                                            int64_t fs, int nf, Geo g,
                                                            int do_dealias, int do_project) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  const ModeIdx m = mode_of(idx, g);
  const bool keep = !do_dealias || in_box(m.nx, m.ny, m.nz, g);
  if (do_project) {  // nf == 3 (checked by the caller)
    cpx<T> u0 = keep ? in[idx] : czero<T>(), u1 = keep ? in[fs + idx] : czero<T>(),
           u2 = keep ? in[2 * fs + idx] : czero<T>();
    const T kx = wavenumber<T>(g.k0x, m.nx), ky = wavenumber<T>(g.k0y, m.ny),
            kz = wavenumber<T>(g.k0z, m.nz);
    const T ik2 = inv_k2_of(k2_of(kx, ky, kz));
    const cpx<T> div = ik2 * (kx * u0 + ky * u1 + kz * u2);
    out[idx] = u0 - kx * div;
    out[fs + idx] = u1 - ky * div;
    out[2 * fs + idx] = u2 - kz * div;
  } else {
    for (int f = 0; f < nf; ++f) out[f * fs + idx] = keep ? in[f * fs + idx] : czero<T>();
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_shell_scale(cpx<T>* st, int64_t fs, int nf, Geo g,
                                                        const double* __restrict__ c, int nbins) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  const ModeIdx m = mode_of(idx, g);
  const T k2 = k2_of(wavenumber<T>(g.k0x, m.nx), wavenumber<T>(g.k0y, m.ny),
                     wavenumber<T>(g.k0z, m.nz));
  const int s = shell_of<T>(k2, g.k0x);
  const T f = s < nbins ? T(c[s]) : T(1);
  for (int i = 0; i < nf; ++i) st[i * fs + idx] = f * st[i * fs + idx];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_spectral_op(const cpx<T>* in, cpx<T>* out, int64_t fs,
                                                        int nf, Geo g, int op) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  const ModeIdx m = mode_of(idx, g);
  const T kx = wavenumber<T>(g.k0x, m.nx), ky = wavenumber<T>(g.k0y, m.ny),
          kz = wavenumber<T>(g.k0z, m.nz);
  for (int f = 0; f < nf; ++f) {
    const cpx<T> u = in[f * fs + idx];
    cpx<T> r;
    switch (op) {
      case kOpDdx: r = mul_i(kx * u); break;
      case kOpDdy: r = mul_i(ky * u); break;
      case kOpDdz: r = mul_i(kz * u); break;
      default: r = -k2_of(kx, ky, kz) * u; break;
    }
    out[f * fs + idx] = r;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_tables(Geo g, T* k2o, T* ik2o, uint8_t* masko,
                                                   int32_t* shello) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  const ModeIdx m = mode_of(idx, g);
  const T k2 = k2_of(wavenumber<T>(g.k0x, m.nx), wavenumber<T>(g.k0y, m.ny),
                     wavenumber<T>(g.k0z, m.nz));
  if (k2o) k2o[idx] = k2;
  if (ik2o) ik2o[idx] = inv_k2_of(k2);
  if (masko) masko[idx] = in_box(m.nx, m.ny, m.nz, g) ? 1 : 0;
  if (shello) shello[idx] = shell_of<T>(k2, g.k0x);
}

// ---- adjoint of the RHS: elementwise pieces ----------------------------------
// The VJP of rhs at u for a cotangent v (vp = P M v) is
//   ubar = (w/Ntot) F[w_p x g] + K (w/Ntot) F[g x u_p] - nu k2 v + forcing part,
// with g = Ntot irfftn(vp / w) the adjoint of rfftn, K = i k x (self-adjoint)
// and w the Hermitian weight of the mode's kz plane.

// P M v at one mode: the 2/3 mask, then the Leray projection, with the
// arithmetic of k_dealias_project (so both give the same bits)
template <class T>
__device__ __forceinline__ void masked_projection(const cpx<T>* __restrict__ v, int64_t fs, int64_t idx,
                                                  const ModeIdx& m, const Geo& g, T kx, T ky, T kz,
                                                  cpx<T> out[3]) {
  const bool keep = in_box(m.nx, m.ny, m.nz, g);
  const cpx<T> u0 = keep ? v[idx] : czero<T>(), u1 = keep ? v[fs + idx] : czero<T>(),
               u2 = keep ? v[2 * fs + idx] : czero<T>();
  const T ik2 = inv_k2_of(k2_of(kx, ky, kz));
  const cpx<T> div = ik2 * (kx * u0 + ky * u1 + kz * u2);
  out[0] = u0 - kx * div;
  out[1] = u1 - ky * div;
  out[2] = u2 - kz * div;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_adj_prep(const cpx<T>* __restrict__ u,
                                                     const cpx<T>* __restrict__ v,
                                                     cpx<T>* __restrict__ S, int64_t fs, Geo g,
                                                     double ntot) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  const ModeIdx m = mode_of(idx, g);
  const T kx = wavenumber<T>(g.k0x, m.nx), ky = wavenumber<T>(g.k0y, m.ny),
          kz = wavenumber<T>(g.k0z, m.nz);
  const cpx<T> u0 = u[idx], u1 = u[fs + idx], u2 = u[2 * fs + idx];
  S[idx] = u0;
  S[fs + idx] = u1;
  S[2 * fs + idx] = u2;
  S[3 * fs + idx] = mul_i(ky * u2 - kz * u1);
  S[4 * fs + idx] = mul_i(kz * u0 - kx * u2);
  S[5 * fs + idx] = mul_i(kx * u1 - ky * u0);
  const T sc = T(ntot / hermitian_weight(m.nz, g.nz));
  cpx<T> vp[3];
  masked_projection(v, fs, idx, m, g, kx, ky, kz, vp);
  for (int c = 0; c < 3; ++c) S[(6 + c) * fs + idx] = sc * vp[c];
}

// k_adj_prep into the compact layout of the pruned transforms ([Kx][jp] per
// field, padding slots zero): for a state with no content outside the 2/3
// box, u, i k x u and P M v have none either
template <class T>
__global__ __launch_bounds__(kBlock) void k_adj_prep_compact(const cpx<T>* __restrict__ u,
                                                             const cpx<T>* __restrict__ v,
                                                             int64_t full_fs, cpx<T>* __restrict__ S,
                                                             int64_t comp_fs, Geo g, double ntot) {
  const int64_t n = int64_t(g.kx) * g.jp;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  const int ix = int(idx / g.jp);
  const int j = int(idx - int64_t(ix) * g.jp);
  const int iy = j / g.kzs, iz = j - iy * g.kzs;
  if (!(iy < g.ky && iz < g.kz)) {  // padding slot
    for (int f = 0; f < 9; ++f) S[f * comp_fs + idx] = czero<T>();
    return;
  }
  const int x = row_of_compact(ix, g.nx, g.kx, g.px), y = row_of_compact(iy, g.ny, g.ky, g.py);
  const int64_t src = (int64_t(x) * g.ny + y) * g.nzh + iz;
  const ModeIdx m = mode_of(src, g);
  const T kx = wavenumber<T>(g.k0x, m.nx), ky = wavenumber<T>(g.k0y, m.ny),
          kz = wavenumber<T>(g.k0z, m.nz);
  const cpx<T> u0 = u[src], u1 = u[full_fs + src], u2 = u[2 * full_fs + src];
  S[idx] = u0;
  S[comp_fs + idx] = u1;
  S[2 * comp_fs + idx] = u2;
  S[3 * comp_fs + idx] = mul_i(ky * u2 - kz * u1);
  S[4 * comp_fs + idx] = mul_i(kz * u0 - kx * u2);
  S[5 * comp_fs + idx] = mul_i(kx * u1 - ky * u0);
  const T sc = T(ntot / hermitian_weight(m.nz, g.nz));
  cpx<T> vp[3];
  masked_projection(v, full_fs, src, m, g, kx, ky, kz, vp);
  for (int c = 0; c < 3; ++c) S[(6 + c) * comp_fs + idx] = sc * vp[c];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_dot_partials(const cpx<T>* __restrict__ x,
                                                         const cpx<T>* __restrict__ v, int64_t fs,
                                                         Geo g, T kmin2, T kmax2, int use_band,
                                                         double* __restrict__ partials) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  double s = 0.0;
  for (int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * kBlock) {
    const ModeIdx m = mode_of(idx, g);
    const T kx = wavenumber<T>(g.k0x, m.nx), ky = wavenumber<T>(g.k0y, m.ny),
            kz = wavenumber<T>(g.k0z, m.nz);
    if (use_band) {
      const T k2 = k2_of(kx, ky, kz);
      if (!(kmin2 <= k2 && k2 <= kmax2)) continue;
    }
    cpx<T> vp[3];
    masked_projection(v, fs, idx, m, g, kx, ky, kz, vp);
    for (int c = 0; c < 3; ++c) {
      const cpx<T> a = x[c * fs + idx];
      s += double(a.x) * vp[c].x + double(a.y) * vp[c].y;
    }
  }
  block_sum3m(s, 0.0, 0.0, partials + 3 * blockIdx.x);
}

// Pairings of a cotangent kb with the parameter derivatives of the RHS at
// the state y (navierstokes.py:89-95, forcing.py:74-163), over all entries of
// the full-layout arrays (the gradient convention dJ = Re sum conj(g) dx):
//   (sum Re conj(kb) (-k2 y),  sum Re conj(P M kb) y,  same over kmin2 <= k2 <= kmax2),
// i.e. d/dnu, d/dA of a linear forcing term A y and, once divided by the
// band's 2E, d/deps of the band forcing term (eps / 2E_band) y on the band.
// P M (projection after the 2/3 mask) is self-adjoint, so it acts on kb.
template <class T>
__global__ __launch_bounds__(kBlock) void k_param_partials(const cpx<T>* __restrict__ kb,
                                                           const cpx<T>* __restrict__ y, int64_t fs,
                                                           Geo g, T kmin2, T kmax2,
                                                           double* __restrict__ partials) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  double s_nu = 0.0, s_a = 0.0, s_band = 0.0;
  for (int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * kBlock) {
    const ModeIdx m = mode_of(idx, g);
    const T kx = wavenumber<T>(g.k0x, m.nx), ky = wavenumber<T>(g.k0y, m.ny),
            kz = wavenumber<T>(g.k0z, m.nz);
    const T k2 = k2_of(kx, ky, kz);
    const cpx<T> a0 = kb[idx], a1 = kb[fs + idx], a2 = kb[2 * fs + idx];
    const cpx<T> b0 = y[idx], b1 = y[fs + idx], b2 = y[2 * fs + idx];
    const double ab = double(a0.x) * b0.x + double(a0.y) * b0.y + double(a1.x) * b1.x +
                      double(a1.y) * b1.y + double(a2.x) * b2.x + double(a2.y) * b2.y;
    s_nu -= double(k2) * ab;
    if (in_box(m.nx, m.ny, m.nz, g)) {
      const cpx<T> div = inv_k2_of(k2) * (kx * a0 + ky * a1 + kz * a2);
      const cpx<T> p0 = a0 - kx * div, p1 = a1 - ky * div, p2 = a2 - kz * div;
      const double pb = double(p0.x) * b0.x + double(p0.y) * b0.y + double(p1.x) * b1.x +
                        double(p1.y) * b1.y + double(p2.x) * b2.x + double(p2.y) * b2.y;
      s_a += pb;
      if (kmin2 <= k2 && k2 <= kmax2) s_band += pb;
    }
  }
  block_sum3m(s_nu, s_a, s_band, partials + 3 * blockIdx.x);
}

// pgrad[0..2] += (S_nu, S_A when the forcing is linear, S_band / 2E_band when
// it is the band forcing and 2E_band > tiny): fixed-order Neumaier sums of the
// block partials; twoE_band (device) comes from the FP64 sums of the same state.
__global__ __launch_bounds__(kBlock) void k_param_accumulate(const double* __restrict__ partials,
                                                             int nblocks, int kind,
                                                             const double* __restrict__ two_e_band,
                                                             double tiny, double* __restrict__ pgrad) {
  __shared__ double sh[3][kBlock];
  KSum acc[3];
  for (int i = threadIdx.x; i < nblocks; i += kBlock)
    for (int c = 0; c < 3; ++c) acc[c].add(partials[3 * i + c]);
  for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] = acc[c].value();
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pgrad[0] += sh[0][0];
    if (kind == HIT3D_FORCING_LINEAR) pgrad[1] += sh[1][0];
    if (kind == HIT3D_FORCING_BAND) {
      const double tb = *two_e_band;
      if (tb > tiny) pgrad[2] += sh[2][0] / tb;
    }
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_hold_partials(const cpx<T>* __restrict__ u,
                                                          const cpx<T>* __restrict__ du,
                                                          int64_t fs, Geo g,
                                                          double* __restrict__ partials) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  double s1 = 0.0, s2 = 0.0;
  for (int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * kBlock) {
    double a = 0.0, b = 0.0;
    for (int c = 0; c < 3; ++c) {
      const cpx<T> x = u[c * fs + idx], y = du[c * fs + idx];
      a += double(x.x) * y.x + double(x.y) * y.y;
      b += double(x.x) * x.x + double(x.y) * x.y;
    }
    const int w = hermitian_weight(int(idx % g.nzh), g.nz);
    s1 += w * a;
    s2 += w * b;
  }
  block_sum3m(s1, s2, 0.0, partials + 3 * blockIdx.x);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_hold_apply(const cpx<T>* __restrict__ u,
                                                       const cpx<T>* __restrict__ du,
                                                       cpx<T>* __restrict__ out, int64_t n,
                                                       const double* __restrict__ coef) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const T a = T(coef[0]);
  const cpx<T> x = u[i], y = du[i];
  cpx<T> r;
  r.x = y.x + a * x.x;
  r.y = y.y + a * x.y;
  out[i] = r;
}

// One RK4 combination step over n complex values (integrators.py:32-38's
// stage sums, with the stage derivative k read once):
//   stage 1: out = u + dt/2 k, acc = u + dt/6 k
//   stage 2: out = u + dt/2 k, acc += dt/3 k
//   stage 3: out = u + dt k,   acc += dt/3 k
//   stage 4: out = acc + dt/6 k (u unused)
// c_out and c_acc are the stage's coefficients in the state's precision.
// With hold_coef, the stage derivative is k + hold_coef[0] us (the energy
// hold's term on the stage input us, added here instead of by its own pass).
template <class T>
__global__ __launch_bounds__(kBlock) void k_rk4_combine(const cpx<T>* u, const cpx<T>* k,
                                                        cpx<T>* __restrict__ acc, cpx<T>* out,
                                                        int64_t n, int stage, T c_out, T c_acc,
                                                        const double* __restrict__ hold_coef,
                                                        const cpx<T>* hold_us) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  cpx<T> kv = k[i];
  if (hold_coef) {   // uniform
    const T a = T(hold_coef[0]);
    const cpx<T> x = hold_us[i];
    kv = {kv.x + a * x.x, kv.y + a * x.y};
  }
  if (stage == 4) {
    const cpx<T> a = acc[i];
    out[i] = {a.x + c_acc * kv.x, a.y + c_acc * kv.y};
    return;
  }
  const cpx<T> x = u[i];
  const cpx<T> a = stage == 1 ? x : acc[i];
  acc[i] = {a.x + c_acc * kv.x, a.y + c_acc * kv.y};
  out[i] = {x.x + c_out * kv.x, x.y + c_out * kv.y};
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_adj_combine(const cpx<T>* __restrict__ AB,
                                                        const cpx<T>* v,
                                                        const cpx<T>* __restrict__ u,
                                                        cpx<T>* __restrict__ out, int64_t fs,
                                                        Geo g, T kmin2, T kmax2, AdjCombine c,
                                                        AdjEpilogue<T> ep) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  const ModeIdx m = mode_of(idx, g);
  const T kx = wavenumber<T>(g.k0x, m.nx), ky = wavenumber<T>(g.k0y, m.ny),
          kz = wavenumber<T>(g.k0z, m.nz);
  const T k2 = k2_of(kx, ky, kz);
  const int w = hermitian_weight(m.nz, g.nz);
  const T sc = T(w * c.inv_ntot);
  const cpx<T> b0 = AB[3 * fs + idx], b1 = AB[4 * fs + idx], b2 = AB[5 * fs + idx];
  cpx<T> r[3] = {AB[idx] + mul_i(ky * b2 - kz * b1), AB[fs + idx] + mul_i(kz * b0 - kx * b2),
                 AB[2 * fs + idx] + mul_i(kx * b1 - ky * b0)};
  const T nk2 = T(c.nu) * k2;
  // forcing part: coef * [band] * vp + dot * grad(coef) (grad along u)
  T fc = T(0), gu = T(0);
  if (c.fkind == 1) {  // LINEAR
    fc = T(c.a);
  } else if (c.fkind == 2) {  // BAND: coef = eps / S_B
    const double sB = c.scalars[3], coef = c.scalars[0];
    const bool in_band = kmin2 <= k2 && k2 <= kmax2;
    if (sB > c.tiny && in_band) {
      fc = T(coef);
      gu = T(-(coef / sB) * 2.0 * w * c.inv_ntot2 * *c.dot);
    }
  } else if (c.fkind == 3) {  // LUNDGREN: coef = (nu S_Om - (S_E/2 - E*)/tau) / S_E
    const double sE = c.scalars[1], coef = c.scalars[0];
    if (sE > c.tiny) {
      double dA_dE = -coef / sE;
      if (c.tau > 0.0 && isfinite(c.tau)) dA_dE -= 0.5 / c.tau / sE;
      fc = T(coef);
      gu = T((*c.dot) * 2.0 * w * c.inv_ntot2 * (c.fnu / sE * double(k2) + dA_dE));
    }
  }
  cpx<T> vp[3];
  masked_projection(v, fs, idx, m, g, kx, ky, kz, vp);
  for (int i = 0; i < 3; ++i) {
    const int64_t o = i * fs + idx;
    const cpx<T> val = sc * r[i] - nk2 * v[o] + fc * vp[i] + gu * u[o];
    // every read of entry o (v, ub, acc_base) precedes this thread's writes
    // of entry o, so kbn may alias v and acc may alias acc_base
    const cpx<T> ubo = ep.kbn ? ep.ub[o] : czero<T>();
    if (ep.acc)
      ep.acc[o] = ep.acc_base[o] + val;
    else
      out[o] = val;
    if (ep.kbn) ep.kbn[o] = ep.alpha * ubo + ep.beta * val;
  }
}

// The forcing term f(u) of the forcing plug-ins (forcing.py:74-163) over the
// full layout: out = c u (uniform: LINEAR with c = A, LUNDGREN with the
// device coefficient) or c [kmin2 <= k2 <= kmax2] u (BAND, device c).
template <class T>
__global__ __launch_bounds__(kBlock) void k_forcing_term(const cpx<T>* __restrict__ u,
                                                         cpx<T>* __restrict__ out, int64_t fs, Geo g,
                                                         int band, T coef_const,
                                                         const double* __restrict__ coef_dev, T kmin2,
                                                         T kmax2) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  const int64_t idx = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  T c = coef_dev ? T(*coef_dev) : coef_const;
  if (band) {
    const ModeIdx m = mode_of(idx, g);
    const T k2 = k2_of(wavenumber<T>(g.k0x, m.nx), wavenumber<T>(g.k0y, m.ny), wavenumber<T>(g.k0z, m.nz));
    if (!(kmin2 <= k2 && k2 <= kmax2)) c = T(0);
  }
  for (int f = 0; f < 3; ++f) out[f * fs + idx] = c * u[f * fs + idx];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_axpby(cpx<T>* out, T a, const cpx<T>* x, T b,
                                                  const cpx<T>* y, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}

// ---- launchers -------------------------------------------------------------

template <class T>
hipError_t launch_gather(const cpx<T>* full, int64_t full_fs, cpx<T>* comp, int64_t comp_fs,
                         int nf, const Geo& g, hipStream_t s) {
  const int64_t n = int64_t(g.kx) * g.jp;
  hipLaunchKernelGGL((k_gather<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, full, full_fs, comp,
                     comp_fs, nf, g);
  return hipGetLastError();
}

template <class T>
hipError_t launch_scatter(const cpx<T>* comp, int64_t comp_fs, cpx<T>* full, int64_t full_fs,
                          int nf, const Geo& g, hipStream_t s) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  hipLaunchKernelGGL((k_scatter<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, comp, comp_fs, full,
                     full_fs, nf, g);
  return hipGetLastError();
}

template <class T>
hipError_t launch_quad_partials(const cpx<T>* st, int64_t fs, const Geo& g, T kmin2, T kmax2,
                                double* partials, hipStream_t s) {
  hipLaunchKernelGGL((k_quad_partials<T>), dim3(kReduceBlocks), dim3(kBlock), 0, s, st, fs, g,
                     kmin2, kmax2, partials);
  return hipGetLastError();
}

template <class T>
hipError_t launch_alias_partials(const cpx<T>* full, int64_t fs, const Geo& g, double* partials,
                                 hipStream_t s) {
  hipLaunchKernelGGL((k_alias_partials<T>), dim3(kReduceBlocks), dim3(kBlock), 0, s, full, fs, g,
                     partials);
  return hipGetLastError();
}

template <class T>
hipError_t launch_spectrum_partials(const cpx<T>* st, int64_t fs, int component, const Geo& g,
                                    int box, int nbins, double* partials, int* nblocks,
                                    hipStream_t s) {
  if ((g.nz / 2 + 1 + 63) / 64 > kSpecChunks) return hipErrorInvalidValue;
  const int nb = spectrum_blocks(g, box);
  *nblocks = nb;
  // the box read reaches shells up to that of its corner (cx, cy, cz); two
  // bins of margin cover the rounding of the shell rule in the state precision
  int used = nbins;
  if (box) {
    const double kx = g.cx * g.k0x, ky = g.cy * g.k0y, kz = g.cz * g.k0z;
    const int corner = int(std::floor(std::sqrt(kx * kx + ky * ky + kz * kz) / g.k0x + 0.5));
    used = corner + 3 < nbins ? corner + 3 : nbins;
  }
  const size_t shmem = size_t(kSpecWaves) * used * 2 * sizeof(double);
  hipLaunchKernelGGL((k_spectrum_partials<T>), dim3(nb), dim3(kBlock), shmem, s, st, fs, component,
                     g, box, used, nbins, partials);
  return hipGetLastError();
}

template <class T>
hipError_t launch_philox_uniform(T* out, int64_t n, uint64_t seed, hipStream_t s) {
  const int64_t blocks4 = (n + 3) / 4;
  hipLaunchKernelGGL((k_philox_uniform<T>), dim3(nblocks_for(blocks4)), dim3(kBlock), 0, s, out, n,
                     seed);
  return hipGetLastError();
}

template <class T>
hipError_t launch_dealias_project(const cpx<T>* in, cpx<T>* out, int64_t fs, int nf,
                                  const Geo& g, int do_dealias, int do_project, hipStream_t s) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  hipLaunchKernelGGL((k_dealias_project<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, in, out, fs,
                     nf, g, do_dealias, do_project);
  return hipGetLastError();
}

template <class T>
hipError_t launch_shell_scale(cpx<T>* st, int64_t fs, int nf, const Geo& g, const double* c,
                              int nbins, hipStream_t s) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  hipLaunchKernelGGL((k_shell_scale<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, st, fs, nf, g,
                     c, nbins);
  return hipGetLastError();
}

template <class T>
hipError_t launch_spectral_op(const cpx<T>* in, cpx<T>* out, int64_t fs, int nf, const Geo& g,
                              int op, hipStream_t s) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  hipLaunchKernelGGL((k_spectral_op<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, in, out, fs, nf,
                     g, op);
  return hipGetLastError();
}

template <class T>
hipError_t launch_tables(const Geo& g, T* k2, T* inv_k2, uint8_t* mask, int32_t* shell,
                         hipStream_t s) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  hipLaunchKernelGGL((k_tables<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, g, k2, inv_k2, mask,
                     shell);
  return hipGetLastError();
}

template <class T>
hipError_t launch_adj_prep(const cpx<T>* u, const cpx<T>* v, cpx<T>* S, int64_t fs, const Geo& g,
                           double ntot, hipStream_t s) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  hipLaunchKernelGGL((k_adj_prep<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, u, v, S, fs, g,
                     ntot);
  return hipGetLastError();
}

template <class T>
hipError_t launch_adj_prep_compact(const cpx<T>* u, const cpx<T>* v, int64_t full_fs, cpx<T>* S,
                                   int64_t comp_fs, const Geo& g, double ntot, hipStream_t s) {
  const int64_t n = int64_t(g.kx) * g.jp;
  hipLaunchKernelGGL((k_adj_prep_compact<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, u, v, full_fs, S,
                     comp_fs, g, ntot);
  return hipGetLastError();
}

template <class T>
hipError_t launch_dot_partials(const cpx<T>* x, const cpx<T>* y, int64_t fs, const Geo& g,
                               T kmin2, T kmax2, int use_band, double* partials, hipStream_t s) {
  hipLaunchKernelGGL((k_dot_partials<T>), dim3(kReduceBlocks), dim3(kBlock), 0, s, x, y, fs, g,
                     kmin2, kmax2, use_band, partials);
  return hipGetLastError();
}

template <class T>
hipError_t launch_param_grad(const cpx<T>* kb, const cpx<T>* y, int64_t fs, const Geo& g, T kmin2,
                             T kmax2, int kind, const double* two_e_band, double tiny,
                             double* partials, double* pgrad, hipStream_t s) {
  hipLaunchKernelGGL((k_param_partials<T>), dim3(kReduceBlocks), dim3(kBlock), 0, s, kb, y, fs, g,
                     kmin2, kmax2, partials);
  hipLaunchKernelGGL(k_param_accumulate, dim3(1), dim3(kBlock), 0, s, partials, kReduceBlocks, kind,
                     two_e_band, tiny, pgrad);
  return hipGetLastError();
}

template <class T>
hipError_t launch_hold_partials(const cpx<T>* u, const cpx<T>* du, int64_t fs, const Geo& g,
                                double* partials, hipStream_t s) {
  hipLaunchKernelGGL((k_hold_partials<T>), dim3(kReduceBlocks), dim3(kBlock), 0, s, u, du, fs, g,
                     partials);
  return hipGetLastError();
}

template <class T>
hipError_t launch_hold_apply(const cpx<T>* u, const cpx<T>* du, cpx<T>* out, int64_t n,
                             const double* coef, hipStream_t s) {
  hipLaunchKernelGGL((k_hold_apply<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, u, du, out, n,
                     coef);
  return hipGetLastError();
}

template <class T>
hipError_t launch_rk4_combine(const cpx<T>* u, const cpx<T>* k, cpx<T>* acc, cpx<T>* out, int64_t n,
                              int stage, double dt, const double* hold_coef, const cpx<T>* hold_us,
                              hipStream_t s) {
  const double co = stage == 3 ? dt : 0.5 * dt;
  const double ca = (stage == 1 || stage == 4) ? dt / 6.0 : dt / 3.0;
  hipLaunchKernelGGL((k_rk4_combine<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, u, k, acc, out, n,
                     stage, T(co), T(ca), hold_coef, hold_us);
  return hipGetLastError();
}

template <class T>
hipError_t launch_adj_combine(const cpx<T>* AB, const cpx<T>* v, const cpx<T>* u, cpx<T>* out,
                              int64_t fs, const Geo& g, T kmin2, T kmax2, const AdjCombine& c,
                              const AdjEpilogue<T>& ep, hipStream_t s) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  hipLaunchKernelGGL((k_adj_combine<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, AB, v, u, out, fs, g,
                     kmin2, kmax2, c, ep);
  return hipGetLastError();
}

template <class T>
hipError_t launch_forcing_term(const cpx<T>* u, cpx<T>* out, int64_t fs, const Geo& g, int band,
                               T coef_const, const double* coef_dev, T kmin2, T kmax2, hipStream_t s) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  hipLaunchKernelGGL((k_forcing_term<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, u, out, fs, g, band,
                     coef_const, coef_dev, kmin2, kmax2);
  return hipGetLastError();
}

template <class T>
hipError_t launch_axpby(cpx<T>* out, T a, const cpx<T>* x, T b, const cpx<T>* y, int64_t n,
                        hipStream_t s) {
  hipLaunchKernelGGL((k_axpby<T>), dim3(nblocks_for(n)), dim3(kBlock), 0, s, out, a, x, b, y, n);
  return hipGetLastError();
}

#define H3_INSTANTIATE_ADJ(T)                                                                      \
  template hipError_t launch_adj_prep<T>(const cpx<T>*, const cpx<T>*, cpx<T>*, int64_t,          \
                                         const Geo&, double, hipStream_t);                         \
  template hipError_t launch_adj_prep_compact<T>(const cpx<T>*, const cpx<T>*, int64_t, cpx<T>*,  \
                                                 int64_t, const Geo&, double, hipStream_t);        \
  template hipError_t launch_dot_partials<T>(const cpx<T>*, const cpx<T>*, int64_t, const Geo&, T, \
                                             T, int, double*, hipStream_t);                        \
  template hipError_t launch_adj_combine<T>(const cpx<T>*, const cpx<T>*, const cpx<T>*, cpx<T>*, \
                                            int64_t, const Geo&, T, T, const AdjCombine&,          \
                                            const AdjEpilogue<T>&, hipStream_t);                   \
  template hipError_t launch_forcing_term<T>(const cpx<T>*, cpx<T>*, int64_t, const Geo&, int, T,  \
                                             const double*, T, T, hipStream_t);                    \
  template hipError_t launch_axpby<T>(cpx<T>*, T, const cpx<T>*, T, const cpx<T>*, int64_t,       \
                                      hipStream_t);                                                       \
  template hipError_t launch_hold_partials<T>(const cpx<T>*, const cpx<T>*, int64_t, const Geo&,  \
                                              double*, hipStream_t);                               \
  template hipError_t launch_param_grad<T>(const cpx<T>*, const cpx<T>*, int64_t, const Geo&, T, T, \
                                           int, const double*, double, double*, double*,          \
                                           hipStream_t);                                           \
  template hipError_t launch_hold_apply<T>(const cpx<T>*, const cpx<T>*, cpx<T>*, int64_t,        \
                                           const double*, hipStream_t);                            \
  template hipError_t launch_rk4_combine<T>(const cpx<T>*, const cpx<T>*, cpx<T>*, cpx<T>*,        \
                                            int64_t, int, double, const double*, const cpx<T>*,    \
                                            hipStream_t);

H3_INSTANTIATE_ADJ(float)
H3_INSTANTIATE_ADJ(double)

#define H3_INSTANTIATE_MISC(T)                                                                     \
  template hipError_t launch_gather<T>(const cpx<T>*, int64_t, cpx<T>*, int64_t, int, const Geo&, \
                                       hipStream_t);                                               \
  template hipError_t launch_scatter<T>(const cpx<T>*, int64_t, cpx<T>*, int64_t, int,            \
                                        const Geo&, hipStream_t);                                  \
  template hipError_t launch_quad_partials<T>(const cpx<T>*, int64_t, const Geo&, T, T, double*,  \
                                              hipStream_t);                                        \
  template hipError_t launch_alias_partials<T>(const cpx<T>*, int64_t, const Geo&, double*,       \
                                               hipStream_t);                                       \
  template hipError_t launch_spectrum_partials<T>(const cpx<T>*, int64_t, int, const Geo&, int,   \
                                                  int, double*, int*, hipStream_t);                \
  template hipError_t launch_philox_uniform<T>(T*, int64_t, uint64_t, hipStream_t);               \
  template hipError_t launch_dealias_project<T>(const cpx<T>*, cpx<T>*, int64_t, int, const Geo&, \
                                                int, int, hipStream_t);                            \
  template hipError_t launch_shell_scale<T>(cpx<T>*, int64_t, int, const Geo&, const double*, int, \
                                            hipStream_t);                                          \
  template hipError_t launch_spectral_op<T>(const cpx<T>*, cpx<T>*, int64_t, int, const Geo&, int, \
                                            hipStream_t);                                          \
  template hipError_t launch_tables<T>(const Geo&, T*, T*, uint8_t*, int32_t*, hipStream_t);

H3_INSTANTIATE_MISC(float)
H3_INSTANTIATE_MISC(double)

}  // namespace h3
