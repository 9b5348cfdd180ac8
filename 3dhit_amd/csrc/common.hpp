// Shared device code for the hit3d HIP kernels (gfx950, wave64).
//
// Layout conventions (all C order, complex interleaved re/im):
//   full spectral state   [3][Nx][Ny][Nz/2+1]          (the API layout)
//   compact state         [3][Kx][Jp]                  Jp >= Ky*Kz, column j = iy*Kz + iz
//   physical field        [F][Nx][Ny][Nz]              (real)
// A "retained" axis of K modes maps compact index i to the integer frequency
//   n(i) = i < P ? i : i - K,   P = (K + 1) / 2,
// and to the full-grid row  i < P ? i : i - K + N.  With K = N this is the
// fftfreq order itself (the dense layout); with K = 2*(N/3) + 1 it keeps only
// the modes inside the 2/3-rule box (|n| <= N/3) — the pruned layout.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace h3 {

template <class T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

}  // namespace h3

#include "twiddles.h"

namespace h3 {

template <class T> __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> __device__ __forceinline__ cpx<T> operator*(T s, cpx<T> a) { return {s * a.x, s * a.y}; }
template <class T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// i * a
template <class T> __device__ __forceinline__ cpx<T> mul_i(cpx<T> a) { return {-a.y, a.x}; }
template <class T> __device__ __forceinline__ cpx<T> czero() { return {T(0), T(0)}; }

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// ---------------------------------------------------------------------------
// In-register DFT of a compile-time power-of-two length R <= 64.
// DIR = -1: X[k] = sum_j x[j] e^{-2 pi i jk/R} (forward); DIR = +1: inverse
// (unnormalised). Recursive radix-2 decimation in time; every index is a
// compile-time constant after unrolling, so the data stay in VGPRs and the
// twiddles are immediates (exact 0/±1 and ±i cases cost no multiplies).
// ---------------------------------------------------------------------------

template <int R, int K, int DIR, class T>
__device__ __forceinline__ cpx<T> tw_const(cpx<T> a) {
  constexpr int k = ((K % R) + R) % R;
  if constexpr (k == 0) {
    return a;
  } else if constexpr (2 * k == R) {
    return {-a.x, -a.y};
  } else if constexpr (4 * k == R) {  // e^{DIR i pi/2} = DIR i
    if constexpr (DIR > 0) return {-a.y, a.x}; else return {a.y, -a.x};
  } else if constexpr (4 * k == 3 * R) {  // e^{DIR i 3pi/2} = -DIR i
    if constexpr (DIR > 0) return {a.y, -a.x}; else return {-a.y, a.x};
  } else {
    static_assert(64 % R == 0, "in-register DFT supports R <= 64");
    constexpr int j = k * (64 / R);
    constexpr T c = T(kCos64[j]);
    constexpr T s = T(DIR > 0 ? kSin64[j] : -kSin64[j]);
    return {a.x * c - a.y * s, a.x * s + a.y * c};
  }
}

template <int R, int DIR, int K = 0, class T>
__device__ __forceinline__ void apply_tw_const(cpx<T>* o) {
  if constexpr (K < R / 2) {
    o[K] = tw_const<R, K, DIR>(o[K]);
    apply_tw_const<R, DIR, K + 1>(o);
  }
}

template <int R, int DIR, class T>
__device__ __forceinline__ void dft(cpx<T>* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    cpx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else {
    cpx<T> e[R / 2], o[R / 2];
#pragma unroll
    for (int j = 0; j < R / 2; ++j) {
      e[j] = v[2 * j];
      o[j] = v[2 * j + 1];
    }
    dft<R / 2, DIR>(e);
    dft<R / 2, DIR>(o);
    apply_tw_const<R, DIR>(o);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      v[k] = e[k] + o[k];
      v[k + R / 2] = e[k] - o[k];
    }
  }
}

// exp(DIR 2 pi i j / N) for 0 <= j < N from the exact table.
template <int N, int DIR, class T>
__device__ __forceinline__ cpx<T> twiddle(int j) {
  static_assert(N <= HIT3D_TW_N, "twiddle table too small");
  const int idx = j * (HIT3D_TW_N / N);
  cpx<T> w;
  if constexpr (sizeof(T) == 4) {
    w = kTwF[idx];
  } else {
    w = kTwD[idx];
  }
  if constexpr (DIR > 0) w.y = -w.y;
  return w;
}

// ---------------------------------------------------------------------------
// Distributed pencil FFT of length N = R*Q over Q threads (one "column").
// Four-step split inside the workgroup: thread q holds x[q + Q r], r < R;
// it does the R-point DFT over r in registers, multiplies by w_N^{q k1},
// exchanges through LDS, then does R/Q Q-point DFTs.  On exit thread t holds
//     v[m*Q + k2] = X[(t + Q m) + R k2],   m < R/Q, k2 < Q.
// `Lds` maps (k1, q) of this thread's column to an LDS slot.
// The function starts with a barrier so the LDS buffer may be reused by
// consecutive calls; every thread of the block must call it.
// ---------------------------------------------------------------------------

template <int N>
struct PencilShape {
  static constexpr int LOG = ilog2(N);
  static constexpr int R = 1 << ((LOG + 1) / 2);
  static constexpr int Q = N / R;
  static constexpr int M = R / Q;  // 1 or 2
  static_assert(R * Q == N && (M == 1 || M == 2), "N must be a power of two");
};

// The inter-step twiddles w_N^{-q k1} (forward sign) of thread q, k1 < R,
// loaded once when a thread runs several transforms.
template <int N, class T>
__device__ __forceinline__ void load_twiddles(cpx<T>* tw, int q) {
#pragma unroll
  for (int k1 = 0; k1 < PencilShape<N>::R; ++k1) tw[k1] = twiddle<N, -1, T>(q * k1);
}

template <int N, int DIR, class T, class Lds>
__device__ __forceinline__ void pencil_fft_tw(cpx<T>* v, int q, Lds lds, const cpx<T>* tw) {
  constexpr int R = PencilShape<N>::R, Q = PencilShape<N>::Q, M = PencilShape<N>::M;
  dft<R, DIR>(v);
#pragma unroll
  for (int k1 = 1; k1 < R; ++k1) {
    cpx<T> w = tw[k1];
    if constexpr (DIR > 0) w.y = -w.y;
    v[k1] = cmul(v[k1], w);
  }
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < R; ++k1) *lds(k1, q) = v[k1];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < M; ++m) {
    cpx<T> u[Q];
#pragma unroll
    for (int qq = 0; qq < Q; ++qq) u[qq] = *lds(q + Q * m, qq);
    dft<Q, DIR>(u);
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) v[m * Q + k2] = u[k2];
  }
}

template <int N, int DIR, class T, class Lds>
__device__ __forceinline__ void pencil_fft(cpx<T>* v, int q, Lds lds) {
  constexpr int R = PencilShape<N>::R, Q = PencilShape<N>::Q, M = PencilShape<N>::M;
  dft<R, DIR>(v);
#pragma unroll
  for (int k1 = 1; k1 < R; ++k1) v[k1] = cmul(v[k1], twiddle<N, DIR, T>(q * k1));
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < R; ++k1) *lds(k1, q) = v[k1];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < M; ++m) {
    cpx<T> u[Q];
#pragma unroll
    for (int qq = 0; qq < Q; ++qq) u[qq] = *lds(q + Q * m, qq);
    dft<Q, DIR>(u);
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) v[m * Q + k2] = u[k2];
  }
}

// Strided passes tile columns across workgroups, and neighbouring tiles share
// the cache lines at their edges. Workgroups are dealt round-robin to the 8
// XCDs (each with its own L2), so map the workgroups one XCD receives to a
// contiguous run of tiles: a shared line is then fetched once, into one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  constexpr int X = 8;
  const int q = nb / X, r = nb % X;
  const int x = b % X, i = b / X;
  return x * q + min(x, r) + i;
}

// Output slot s = m*Q + k2 of pencil_fft holds frequency/position
//     (t + Q m) + R k2.
template <int N>
__device__ __forceinline__ int out_index(int t, int s) {
  constexpr int R = PencilShape<N>::R, Q = PencilShape<N>::Q;
  return t + Q * (s / Q) + R * (s % Q);
}

// pencil_fft's output slot s, seen as the input slot r of a following
// pencil_fft on the same thread (input position q + Q r).
template <int N>
constexpr int out_slot_as_input(int s) {
  constexpr int R = PencilShape<N>::R, Q = PencilShape<N>::Q;
  return (s / Q) + (R / Q) * (s % Q);
}

// LDS slot maps.  Strided passes: T columns interleaved, threads tid = c + T q.
template <int N, int T_COLS, class T>
struct LdsStrided {
  cpx<T>* base;  // already offset by column c
  __device__ __forceinline__ cpx<T>* operator()(int k1, int q) const {
    constexpr int Q = PencilShape<N>::Q;
    return base + (k1 * (Q + 1) + q) * T_COLS;
  }
};
// Contiguous (z) passes: one padded slab per pencil, threads tid = q + Q p.
template <int N, class T>
struct LdsPencil {
  cpx<T>* base;  // already offset by pencil p
  __device__ __forceinline__ cpx<T>* operator()(int k1, int q) const {
    constexpr int Q = PencilShape<N>::Q;
    return base + k1 * (Q + 1) + q;
  }
};
template <int N>
constexpr int lds_pencil_elems() {
  return PencilShape<N>::R * (PencilShape<N>::Q + 1);
}

// ---------------------------------------------------------------------------
// Grid geometry and the reference's wavenumber / mask / shell rules
// (grid.py:89-132, diagnostics.py:76-82), evaluated from indices.
// ---------------------------------------------------------------------------

struct Geo {
  int nx, ny, nz, nzh;  // grid points; nzh = nz/2 + 1
  int kx, ky, kz;       // retained modes per axis (compact extents)
  int px, py;           // positive-frequency counts (kx+1)/2, (ky+1)/2
  int cx, cy, cz;       // 2/3-rule cut N_i // 3
  int jp;               // row stride of x-pass arrays (>= ky*kz)
  double k0x, k0y, k0z; // 2 pi / L_i
};

// compact index -> integer frequency
__device__ __forceinline__ int freq_of(int i, int K, int P) { return i < P ? i : i - K; }
// full-grid row -> compact index, or -1 when the row is not retained
__device__ __forceinline__ int compact_of_row(int x, int N, int K, int P) {
  return x < P ? x : (x >= N - (K - P) ? x - N + K : -1);
}
// compact index -> full-grid row
__device__ __forceinline__ int row_of_compact(int i, int N, int K, int P) { return i < P ? i : i - K + N; }

// k = (2 pi / L * n).astype(dtype): product in double, rounded once.
template <class T>
__device__ __forceinline__ T wavenumber(double k0, int n) {
  return T(k0 * double(n));
}

// k2 = kx**2 + ky**2 + kz**2 in dtype, each operation rounded: contraction
// into FMA is switched off for this scope (an inlined x*x + y*y would
// otherwise fuse and round once, unlike the reference's NumPy tables).
template <class T>
__device__ __forceinline__ T k2_of(T kx, T ky, T kz) {
#pragma clang fp contract(off)
  const T a = kx * kx, b = ky * ky, c = kz * kz;
  return (a + b) + c;
}

// where(k2 == 0, 0, 1/k2), correctly rounded
template <class T>
__device__ __forceinline__ T inv_k2_of(T k2) {
  if constexpr (sizeof(T) == 4) {
    return k2 == 0.0f ? 0.0f : __fdiv_rn(1.0f, k2);
  } else {
    return k2 == 0.0 ? 0.0 : __ddiv_rn(1.0, k2);
  }
}

// floor(sqrt(k2) / k0 + 0.5) in dtype, as int32
template <class T>
__device__ __forceinline__ int shell_of(T k2, double k0) {
  if constexpr (sizeof(T) == 4) {
    return int(floorf(__fadd_rn(__fdiv_rn(__fsqrt_rn(k2), float(k0)), 0.5f)));
  } else {
    return int(floor(__dadd_rn(__ddiv_rn(__dsqrt_rn(k2), k0), 0.5)));
  }
}

// per-axis 2/3 box, False at k = 0
__device__ __forceinline__ bool in_box(int nx, int ny, int nz, const Geo& g) {
  return abs(nx) <= g.cx && abs(ny) <= g.cy && nz <= g.cz && (nx | ny | nz) != 0;
}

// Hermitian weight of kz index iz (1 at kz = 0 and the even-Nz Nyquist plane)
__device__ __forceinline__ int hermitian_weight(int iz, int nz) {
  return (iz == 0 || ((nz & 1) == 0 && 2 * iz == nz)) ? 1 : 2;
}

// Neumaier-compensated accumulation
struct KSum {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double v) {
    double t = s + v;
    c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

}  // namespace h3
