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
// Distributed pencil FFTs. A pencil of length N is spread over Q threads
// (its "column"), each holding R = N/Q values in registers; the threads
// exchange through LDS slots of the column (`slot(i)` maps slot i of this
// column to LDS). Two shapes, chosen by N (FftFor<N>):
//
// Fft2 (N <= 256): N = R*Q, R = 2^ceil(log2(N)/2), Q = N/R.  Thread q enters
//   with v[r] = x[q + Q r]; R-point DFT over r in registers, twiddle
//   w_N^{q k1}, one exchange, then R/Q Q-point DFTs.  Exit: thread t holds
//   v[m*Q + k2] = X[(t + Q m) + R k2].
// Fft3<N, Q> (Q threads, R = N/Q, needs R*R >= Q): thread q enters with
//   v[r] = x[q + Q r]; R-point DFT over r, twiddle w_N^{q k1}; the Q-point
//   DFT over q is split again as q = q1 + R q2 (Q2 = Q/R points over q2,
//   twiddle w_Q^{q1 j}, then R points over q1), with two exchanges.  Exit:
//   thread t holds v[s] = X[t + Q s].  Used with Q = 64 (one wave) for
//   N >= 512, and with Q = 32 for the 256-point z passes.
// In both, run() starts with a barrier (Sync) so consecutive calls may reuse
// the slots, and every thread of the block must call it.
// ---------------------------------------------------------------------------

// Barriers around the LDS exchanges. Multi-wave blocks need the workgroup
// barrier; a one-wave block only needs the compiler to keep its LDS accesses
// in program order, since the LDS executes one wave's operations in order.
struct BlockSync {
  __device__ static __forceinline__ void sync() { __syncthreads(); }
};
struct WaveSync {
  __device__ static __forceinline__ void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};

template <class T, int DIR>
__device__ __forceinline__ cpx<T> apply_tw(cpx<T> v, cpx<T> w) {  // w has the forward sign
  if constexpr (DIR > 0) w.y = -w.y;
  return cmul(v, w);
}

template <int N>
struct Fft2 {
  static constexpr int LOG = ilog2(N);
  static constexpr int R = 1 << ((LOG + 1) / 2);
  static constexpr int Q = N / R;
  static constexpr int M = R / Q;  // 1 or 2
  static constexpr int NPTS = N;
  static constexpr int kSlots = R * (Q + 1);
  static_assert(R * Q == N && (M == 1 || M == 2), "N must be a power of two");

  __host__ __device__ static constexpr int out_index(int t, int s) {
    return t + Q * (s / Q) + R * (s % Q);
  }
  // exit slot s, seen as the entry slot r of a following transform on the
  // same thread (entry position q + Q r)
  static constexpr int out_slot_as_input(int s) { return (s / Q) + (R / Q) * (s % Q); }

  template <class T>
  __device__ static __forceinline__ void load_tw(cpx<T>* tw, int q) {
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) tw[k1] = twiddle<N, -1, T>(q * k1);
  }

  template <int DIR, class Sync = BlockSync, class T, class Slot, class TW>
  __device__ static __forceinline__ void run(cpx<T>* v, int q, Slot slot, const TW& tw) {
    dft<R, DIR>(v);
#pragma unroll
    for (int k1 = 1; k1 < R; ++k1) v[k1] = apply_tw<T, DIR>(v[k1], tw[k1]);
    Sync::sync();
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) *slot.at(q, k1 * (Q + 1)) = v[k1];
    Sync::sync();
#pragma unroll
    for (int m = 0; m < M; ++m) {
      cpx<T> u[Q];
#pragma unroll
      for (int qq = 0; qq < Q; ++qq) u[qq] = *slot.at(q * (Q + 1), Q * m * (Q + 1) + qq);
      dft<Q, DIR>(u);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) v[m * Q + k2] = u[k2];
    }
  }
};

template <int N, int QT = 64>
struct Fft3 {
  static constexpr int Q = QT, R = N / QT;
  static constexpr int NPTS = N;
  static constexpr int Q2 = QT / R;       // points of the middle DFT (over q2)
  static constexpr int M1 = R / Q2;       // middle DFTs per thread
  static constexpr int E1 = QT + 1;       // k1 stride of the first exchange
  static constexpr int kSlots = (R * E1 > QT * (R + 1)) ? R * E1 : QT * (R + 1);
  static_assert(R * Q2 == QT && Q2 >= 1 && Q2 <= R && R <= 64, "Fft3 needs R >= Q/R");

  __host__ __device__ static constexpr int out_index(int t, int s) { return t + QT * s; }
  static constexpr int out_slot_as_input(int s) { return s; }

  template <class T>
  __device__ static __forceinline__ void load_tw(cpx<T>* tw, int q) {
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) tw[k1] = twiddle<N, -1, T>(q * k1);
  }

  template <int DIR, class Sync = BlockSync, class T, class Slot, class TW>
  __device__ static __forceinline__ void run(cpx<T>* v, int q, Slot slot, const TW& tw) {
    dft<R, DIR>(v);
#pragma unroll
    for (int k1 = 1; k1 < R; ++k1) v[k1] = apply_tw<T, DIR>(v[k1], tw[k1]);
    Sync::sync();
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) *slot.at(q, k1 * E1) = v[k1];
    Sync::sync();
    // middle DFTs: this thread takes k1 = q % R and q1 = q / R + Q2 m
    const int k1 = q % R, g = q / R;
    cpx<T> u[M1][Q2];
#pragma unroll
    for (int m = 0; m < M1; ++m)
#pragma unroll
      for (int q2 = 0; q2 < Q2; ++q2) u[m][q2] = *slot.at(k1 * E1 + g, Q2 * m + R * q2);
    Sync::sync();
#pragma unroll
    for (int m = 0; m < M1; ++m) {
      const int q1 = g + Q2 * m;
      dft<Q2, DIR>(u[m]);
#pragma unroll
      for (int j = 0; j < Q2; ++j) {
        const cpx<T> w = j ? apply_tw<T, DIR>(u[m][j], twiddle<QT, -1, T>(q1 * j)) : u[m][j];
        *slot.at(k1 * (R + 1) + g, R * j * (R + 1) + Q2 * m) = w;
      }
    }
    Sync::sync();
    // last DFT over q1: thread q = k1' + R j' holds X[q + Q s]
#pragma unroll
    for (int q1 = 0; q1 < R; ++q1) v[q1] = *slot.at(q * (R + 1), q1);
    dft<R, DIR>(v);
  }
};

// First-stage twiddles w_N^{q k1} of thread q: held in VGPRs (loaded once
// per kernel), or, for kernels that cannot spare the registers, read from the
// table at each use (the table stays in the vector L1).
template <class F, class T, bool LAZY = false>
struct Twiddles {
  cpx<T> w[F::R];
  __device__ __forceinline__ explicit Twiddles(int q) { F::load_tw(w, q); }
  __device__ __forceinline__ cpx<T> operator[](int k1) const { return w[k1]; }
};
template <class F, class T>
struct Twiddles<F, T, true> {
  int q;
  __device__ __forceinline__ explicit Twiddles(int q_) : q(q_) {}
  __device__ __forceinline__ cpx<T> operator[](int k1) const { return twiddle<F::NPTS, -1, T>(q * k1); }
};

// strided passes: Fft2 up to 256 points (16 threads per column keeps 16
// columns = 128-byte rows in a 256-thread block), one-wave Fft3 above
template <int N, bool THREE = (N >= 512)>
struct FftSel {
  using type = Fft2<N>;
};
template <int N>
struct FftSel<N, true> {
  using type = Fft3<N>;
};
// 256 points: the three-level transform over 32 threads (8 values per thread)
template <>
struct FftSel<256, false> {
  using type = Fft3<256, 32>;
};
// 128 points: the three-level transform over 16 threads (8 values per thread)
template <>
struct FftSel<128, false> {
  using type = Fft3<128, 16>;
};
template <int N>
using FftFor = typename FftSel<N>::type;

// z passes (contiguous pencils): fewer values per thread at 256 points
template <int N>
struct FftZSel {
  using type = FftFor<N>;
};
template <>
struct FftZSel<256> {
  using type = Fft3<256, 32>;
};
template <int N>
using FftZFor = typename FftZSel<N>::type;

// Slot maps. Strided passes interleave TC columns (threads tid = c + TC q);
// z passes give each pencil its own run of slots (threads tid = q + Q p).
template <int TC, class T>
struct SlotStrided {
  cpx<T>* base;  // already offset by column c
  __device__ __forceinline__ cpx<T>* operator()(int i) const { return base + i * TC; }
  // slot lane + c, c a compile-time constant after unrolling: the lane part
  // is one address register and c * TC * sizeof(cpx) the instruction's
  // immediate offset (a constant folded inside the product would instead be
  // materialised in a register per slot)
  __device__ __forceinline__ cpx<T>* at(int lane, int c) const { return (base + lane * TC) + c * TC; }
};
template <class T>
struct SlotPencil {
  cpx<T>* base;  // already offset by pencil p
  __device__ __forceinline__ cpx<T>* operator()(int i) const { return base + i; }
  __device__ __forceinline__ cpx<T>* at(int lane, int c) const { return (base + lane) + c; }
};

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

// The same runs, each walked from its end: a pass that follows one mapped by
// xcd_tile over the same index order starts on the tiles its producer wrote
// last, which are still in the 256 MB memory-side cache.
__device__ __forceinline__ int xcd_tile_rev(int b, int nb) {
  constexpr int X = 8;
  const int q = nb / X, r = nb % X;
  const int x = b % X, i = b / X;
  const int n = q + (x < r ? 1 : 0);
  return x * q + min(x, r) + (n - 1 - i);
}

// ---------------------------------------------------------------------------
// Grid geometry and the reference's wavenumber / mask / shell rules
// (grid.py:89-132, diagnostics.py:76-82), evaluated from indices.
// ---------------------------------------------------------------------------

struct Geo {
  int nx, ny, nz, nzh;  // grid points; nzh = nz/2 + 1
  int kx, ky, kz;       // retained modes per axis (compact extents)
  int kzs;              // row stride of the kz rows in the compact, x-pass and pencil arrays:
                        // kz rounded up to 16 (or 8, see make_geo) in the pruned layout, so
                        // rows start on 128- (64-) byte boundaries for complex64; the padding
                        // slots hold zeros. kz when dense
  int px, py;           // positive-frequency counts (kx+1)/2, (ky+1)/2
  int cx, cy, cz;       // 2/3-rule cut N_i // 3
  int jp;               // row stride of x-pass arrays (ky*kzs rounded up to 16)
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
