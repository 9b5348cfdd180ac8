// Pencil-pass kernels of the pseudo-spectral RK4 step (gfx950).
//
// One RHS evaluation = five passes over HBM (navierstokes.py:69-95):
//   P1 x-inverse of u and of the x-dependent curl parts      [3][Kx][Jp] -> [5][Nx][Jp]
//   P2 y-inverse, forming w0 = i(ky u2 - kz u1) on the fly   [5][Nx][Jp] -> [6][Nx][Ny][Kz]
//   P3 z c2r of (u_j + i w_j), u x w, z r2c                  [6][..][Kz] -> [3][..][Kz] in place
//   P4 y-forward                                             [3][Nx][Ny][Kz] -> [3][Nx][Jp]
//   P5 x-forward + forcing + dealias + projection + viscous + RK4 stage, FP64 sums
// The standalone to_physical / to_spectral use the plain strided pass and the
// z c2r / r2c kernels.
#include "passes.hpp"

#include <algorithm>
#include <type_traits>

namespace h3 {

// Rows a pruned axis never stores: the 2/3 box keeps K = 2(N/3) + 1 rows,
// the P = (K + 1)/2 non-negative frequencies first, so rows P .. N-(K-P)-1
// are absent. Bit r of kIn: entry slot r (row q + Q r) is absent for every
// thread q, so its value is a compile-time zero; bit s of kOut: exit slot s
// (row out_index(q, s)) likewise, so its value is never stored or used. At
// 1024 points 4 of a thread's 16 slots are absent either way, which frees
// the registers the lean kernels need.
template <class F, int N>
struct PrunedRows {
  static constexpr int K = 2 * (N / 3) + 1, P = (K + 1) / 2;
  static constexpr bool absent(int row) { return row >= P && row < N - (K - P); }
  static constexpr unsigned in_mask() {
    unsigned m = 0;
    for (int r = 0; r < F::R; ++r) {
      bool all = true;
      for (int q = 0; q < F::Q; ++q) all = all && absent(q + F::Q * r);
      if (all) m |= 1u << r;
    }
    return m;
  }
  static constexpr unsigned out_mask() {
    unsigned m = 0;
    for (int s = 0; s < F::R; ++s) {
      bool all = true;
      for (int q = 0; q < F::Q; ++q) all = all && absent(F::out_index(q, s));
      if (all) m |= 1u << s;
    }
    return m;
  }
  static constexpr unsigned kIn = in_mask(), kOut = out_mask();
};

// Point counts whose fat kernels have pruned-row instances (at 1024 the lean
// kernels take the pruned layout; below 128 no slot is absent for every thread)
template <int N>
constexpr bool kPrunedN = N == 128 || N == 256 || N == 512;

// ---------------------------------------------------------------------------
// Plain strided pass (any axis, either direction)
// ---------------------------------------------------------------------------
template <int N, int DIR, class T, bool PIN = false, bool POUT = false>
__global__ __launch_bounds__(StridedShape<N>::THREADS) __attribute__((amdgpu_waves_per_eu(MinWaves<N, T>::value))) void k_strided(StridedArgs<T> a) {
  using S = StridedShape<N>;
  using F = typename S::F;
  constexpr int R = S::R, Q = S::Q, TC = S::TC;
  constexpr unsigned kZeroIn = PIN ? PrunedRows<F, N>::kIn : 0u, kDeadOut = POUT ? PrunedRows<F, N>::kOut : 0u;
  __shared__ cpx<T> lds[S::LDS];
  const int c = threadIdx.x % TC, q = threadIdx.x / TC;
  const int j = xcd_tile(blockIdx.x, gridDim.x) * TC + c;
  const bool active = j < a.J;
  const int jj = active ? j : 0;
  const int ao = jj / a.B, bo = jj - ao * a.B;
  const cpx<T>* in0 = a.in + ao * a.in_slab + bo;
  cpx<T>* out0 = a.out + ao * a.out_slab + bo;
  SlotStrided<TC, T> L{lds + c};
  cpx<T> tw[R];
  F::load_tw(tw, q);
  int irow[R];
#pragma unroll
  for (int r = 0; r < R; ++r) irow[r] = active ? compact_of_row(q + Q * r, N, a.in_k, a.in_p) : -1;
  const int64_t irs = a.in_row(), ors = a.out_row();
  auto load = [&](int f, cpx<T>* v) {
    const cpx<T>* src = in0 + f * a.in_fs;
#pragma unroll
    for (int r = 0; r < R; ++r)
      v[r] = !((kZeroIn >> r) & 1u) && irow[r] >= 0 ? src[int64_t(irow[r]) * irs] : czero<T>();
  };
  // this block's fields [f0, f1): the launcher splits the fields over
  // gridDim.y when the column tiles alone would leave most CUs idle
  const int fpb = (a.nf + int(gridDim.y) - 1) / int(gridDim.y);
  const int f0 = int(blockIdx.y) * fpb, f1 = min(a.nf, f0 + fpb);
  if (f0 >= f1) return;   // uniform per block
  // field f+1 is loaded before field f's transform, whose barriers would
  // otherwise hold the loads back
  cpx<T> nxt[R];
  load(f0, nxt);
  for (int f = f0; f < f1; ++f) {
    cpx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = nxt[r];
    if (f + 1 < f1) load(f + 1, nxt);
    F::template run<DIR>(v, q, L, tw);
    cpx<T>* dst = out0 + f * a.out_fs;
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if ((kDeadOut >> s) & 1u) continue;
      const int o = compact_of_row(F::out_index(q, s), N, a.out_k, a.out_p);
      if (active && o >= 0) dst[int64_t(o) * ors] = v[s];
    }
  }
}

// ---------------------------------------------------------------------------
// P1: x-inverse of u0, u1, u2 and of the x-dependent vorticity parts
//   w1 = i(kz u0 - kx u2), w2 = i(kx u1 - ky u0)   (navierstokes.py:73-77)
// out fields: [u0, u1, u2, w1, w2]
// ---------------------------------------------------------------------------
template <int N, class T, bool PIN = false>
__global__ __launch_bounds__(CurlShape<N>::THREADS) __attribute__((amdgpu_waves_per_eu(MinWaves<N, T>::value))) void k_x_inverse_curl(const cpx<T>* __restrict__ us,
                                                            int64_t us_fs,
                                                            cpx<T>* __restrict__ out,
                                                            int64_t out_fs, Geo g,
                                                            FinalizeJob job) {
  using S = CurlShape<N>;
  using F = typename S::F;
  constexpr int R = S::R, Q = S::Q, TC = S::TC;
  constexpr unsigned kZeroIn = PIN ? PrunedRows<F, N>::kIn : 0u;
  static_assert(S::THREADS % 64 == 0 && S::THREADS <= 1024, "finalize_in_block needs whole waves");
  if (job.partials != nullptr && blockIdx.x == 0 && blockIdx.y == 0) finalize_in_block(job);   // uniform per block
  __shared__ cpx<T> lds[S::LDS];
  const int c = threadIdx.x % TC, q = threadIdx.x / TC;
  const int j = xcd_tile(blockIdx.x, gridDim.x) * TC + c;
  const bool active = j < g.ky * g.kzs;   // padding columns (iz >= kz) carry zeros
  const int jj = active ? j : 0;
  const int iy = jj / g.kzs, iz = jj - iy * g.kzs;
  const T kyv = wavenumber<T>(g.k0y, freq_of(iy, g.ky, g.py));
  const T kzv = wavenumber<T>(g.k0z, iz);
  SlotStrided<TC, T> L{lds + c};
  cpx<T> tw[R];
  F::load_tw(tw, q);
  int irow[R];
  T kxv[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    irow[r] = active ? compact_of_row(q + Q * r, N, g.kx, g.px) : -1;
    kxv[r] = wavenumber<T>(g.k0x, irow[r] >= 0 ? freq_of(irow[r], g.kx, g.px) : 0);
  }
  const cpx<T>* u0 = us + jj;
  const cpx<T>* u1 = u0 + us_fs;
  const cpx<T>* u2 = u1 + us_fs;
  auto load = [&](const cpx<T>* p, cpx<T>* v) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      v[r] = !((kZeroIn >> r) & 1u) && irow[r] >= 0 ? p[int64_t(irow[r]) * g.jp] : czero<T>();
  };
  auto store = [&](int f, const cpx<T>* v) {
    if (!active) return;
    cpx<T>* dst = out + f * out_fs + j;
#pragma unroll
    for (int s = 0; s < R; ++s) dst[int64_t(F::out_index(q, s)) * g.jp] = v[s];
  };
  cpx<T> a[R], b[R], c1[R], w[R];
  if (gridDim.y == 2) {   // uniform: the five outputs in two groups (u0 read by both)
    load(u0, a);
    if (blockIdx.y == 0) {   // w1 = i(kz u0 - kx u2), u2, u0
      load(u2, b);
#pragma unroll
      for (int r = 0; r < R; ++r) w[r] = mul_i(kzv * a[r] - kxv[r] * b[r]);
      F::template run<+1>(w, q, L, tw);
      store(3, w);
      F::template run<+1>(b, q, L, tw);
      store(2, b);
      F::template run<+1>(a, q, L, tw);
      store(0, a);
    } else {   // w2 = i(kx u1 - ky u0), u1
      load(u1, c1);
#pragma unroll
      for (int r = 0; r < R; ++r) w[r] = mul_i(kxv[r] * c1[r] - kyv * a[r]);
      F::template run<+1>(w, q, L, tw);
      store(4, w);
      F::template run<+1>(c1, q, L, tw);
      store(1, c1);
    }
    return;
  }
  // all three components are loaded up front (the barriers inside run()
  // would otherwise delay every later load until the previous transform ends)
  load(u0, a);
  load(u2, b);
  load(u1, c1);
#pragma unroll
  for (int r = 0; r < R; ++r) w[r] = mul_i(kzv * a[r] - kxv[r] * b[r]);
  F::template run<+1>(w, q, L, tw);
  store(3, w);
  F::template run<+1>(b, q, L, tw);
  store(2, b);
#pragma unroll
  for (int r = 0; r < R; ++r) w[r] = mul_i(kxv[r] * c1[r] - kyv * a[r]);
  F::template run<+1>(w, q, L, tw);
  store(4, w);
  F::template run<+1>(c1, q, L, tw);
  store(1, c1);
  F::template run<+1>(a, q, L, tw);
  store(0, a);
}

// ---------------------------------------------------------------------------
// P2: y-inverse; forms w0 = i(ky u2 - kz u1) before its transform.
// in  [5][Nx][jp] fields (u0, u1, u2, w1, w2), rows iy (stride Kz)
// out [6][Nx][Ny][Kz] fields (u0, u1, u2, w0, w1, w2)
// ---------------------------------------------------------------------------
template <int N, class T, bool PIN = false>
__global__ __launch_bounds__(CurlShape<N>::THREADS) __attribute__((amdgpu_waves_per_eu(MinWaves<N, T>::value))) void k_y_inverse_curl(const cpx<T>* __restrict__ in,
                                                            int64_t in_fs,
                                                            cpx<T>* __restrict__ out,
                                                            int64_t out_fs, Geo g, int yrow) {
  using S = CurlShape<N>;
  using F = typename S::F;
  constexpr int R = S::R, Q = S::Q, TC = S::TC;
  constexpr unsigned kZeroIn = PIN ? PrunedRows<F, N>::kIn : 0u;
  __shared__ cpx<T> lds[S::LDS];
  const int c = threadIdx.x % TC, q = threadIdx.x / TC;
  const int j = xcd_tile(blockIdx.x, gridDim.x) * TC + c;
  const bool active = j < g.nx * g.kzs;   // padding columns (iz >= kz) carry zeros
  const int jj = active ? j : 0;
  const int x = jj / g.kzs, iz = jj - x * g.kzs;
  const T kzv = wavenumber<T>(g.k0z, iz);
  SlotStrided<TC, T> L{lds + c};
  cpx<T> tw[R];
  F::load_tw(tw, q);
  int irow[R];
  T kyv[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    irow[r] = active ? compact_of_row(q + Q * r, N, g.ky, g.py) : -1;
    kyv[r] = wavenumber<T>(g.k0y, irow[r] >= 0 ? freq_of(irow[r], g.ky, g.py) : 0);
  }
  const cpx<T>* src = in + int64_t(x) * g.jp + iz;
  // pencil buffer rows: y-major (row y at y nx kzs, column j) or x-major
  // (pencil (x, y) at (x ny + y) kzs)
  cpx<T>* dst0 = out + (yrow ? int64_t(jj) : int64_t(x) * g.ny * g.kzs + iz);
  const int prs = yrow ? yrow : g.kzs;
  auto load = [&](int f, cpx<T>* v) {
    const cpx<T>* p = src + f * in_fs;
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = !((kZeroIn >> r) & 1u) && irow[r] >= 0 ? p[irow[r] * g.kzs] : czero<T>();
  };
  auto store = [&](int f, const cpx<T>* v) {
    if (!active) return;
    cpx<T>* p = dst0 + f * out_fs;
#pragma unroll
    for (int s = 0; s < R; ++s) p[F::out_index(q, s) * prs] = v[s];
  };
  // each field's loads are issued one transform ahead of its use (pf holds
  // the prefetched field), so the barriers inside run() hide their latency
  // The three-level transforms (N >= 256) need more registers, so field 0
  // is fetched after w's transform (three arrays live instead of four).
  constexpr bool kEarly = N < 256 || (PIN && N == 256);
  cpx<T> a[R], b[R], pf[R], w[R];
  if (gridDim.y == 2) {   // uniform: the six outputs in two groups of three
    if (blockIdx.y == 0) {   // u1, u2 and w0 = i(ky u2 - kz u1)
      load(1, a);
      load(2, b);
#pragma unroll
      for (int r = 0; r < R; ++r) w[r] = mul_i(kyv[r] * b[r] - kzv * a[r]);
      F::template run<+1>(w, q, L, tw);
      store(3, w);
      F::template run<+1>(b, q, L, tw);
      store(2, b);
      F::template run<+1>(a, q, L, tw);
      store(1, a);
    } else {   // u0, w1, w2
      load(0, pf);
      load(3, b);
      load(4, a);
      F::template run<+1>(pf, q, L, tw);
      store(0, pf);
      F::template run<+1>(b, q, L, tw);
      store(4, b);
      F::template run<+1>(a, q, L, tw);
      store(5, a);
    }
    return;
  }
  load(1, a);
  load(2, b);
  if constexpr (kEarly) load(0, pf);
#pragma unroll
  for (int r = 0; r < R; ++r) w[r] = mul_i(kyv[r] * b[r] - kzv * a[r]);
  F::template run<+1>(w, q, L, tw);
  store(3, w);
  if constexpr (!kEarly) load(0, pf);
  F::template run<+1>(b, q, L, tw);
  store(2, b);
  load(3, b);
  F::template run<+1>(a, q, L, tw);
  store(1, a);
  load(4, a);
  F::template run<+1>(pf, q, L, tw);
  store(0, pf);
  F::template run<+1>(b, q, L, tw);
  store(4, b);
  F::template run<+1>(a, q, L, tw);
  store(5, a);
}

// ---------------------------------------------------------------------------
// Register-lean forms of the 1024-point float x-inverse and plain strided
// passes. The forms above hold three or four fields' values per thread
// (~240 VGPRs, 2 waves per SIMD) on 8-column workgroups, so every row segment
// is half a 128-byte line and the other half belongs to another workgroup.
// These hold at most three transforms' values per thread (two in the plain
// pass) within 128 VGPRs, which allows 16-column workgroups of 16 waves
// (4 waves per SIMD, 136 KB of LDS): each row segment is a whole line.
// Addresses are a wave-uniform field base plus a 32-bit per-lane byte offset
// (the launchers take these kernels only when every offset fits), so no
// 64-bit per-lane address is held across the field sequence.
// ---------------------------------------------------------------------------
// A copy of x the compiler cannot see through (an empty asm statement): the
// lane-dependent values derived from it (row offsets, LDS addresses, middle-
// stage twiddles) are recomputed in each iteration of the field sequence
// instead of being hoisted out of it and held in registers.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

template <class T>
__device__ __forceinline__ cpx<T> ld_off(const cpx<T>* base, uint32_t off) {
  return *reinterpret_cast<const cpx<T>*>(reinterpret_cast<const char*>(base) + off);
}
template <class T>
__device__ __forceinline__ void st_off(cpx<T>* base, uint32_t off, cpx<T> v) {
  *reinterpret_cast<cpx<T>*>(reinterpret_cast<char*>(base) + off) = v;
}

// PIN / POUT: the input / output rows are the pruned layout's (compact index
// = row of the retained set, as compact_of_row maps them), so the absent
// slots above are compile-time.
template <int N, int DIR, class T, bool PIN, bool POUT>
__global__ __launch_bounds__(LeanShape<N>::THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_strided_lean(StridedArgs<T> a) {
  using S = LeanShape<N>;
  using F = typename S::F;
  using PR = PrunedRows<F, N>;
  constexpr int R = S::R, Q = S::Q, TC = S::TC;
  constexpr unsigned kZeroIn = PIN ? PR::kIn : 0u, kDeadOut = POUT ? PR::kOut : 0u;
  constexpr uint32_t E = sizeof(cpx<T>);
  __shared__ cpx<T> lds[S::LDS];
  const int c = threadIdx.x % TC, q = threadIdx.x / TC;
  const int j = xcd_tile(blockIdx.x, gridDim.x) * TC + c;
  const bool active = j < a.J;
  const int jj = active ? j : 0;
  const int ao = jj / a.B, bo = jj - ao * a.B;
  const uint32_t cin = uint32_t(int64_t(ao) * a.in_slab + bo) * E;
  const uint32_t cout = uint32_t(int64_t(ao) * a.out_slab + bo) * E;
  const uint32_t rs_in = uint32_t(a.in_row()) * E, rs_out = uint32_t(a.out_row()) * E;
  SlotStrided<TC, T> L{lds + c};
  const Twiddles<F, T> tw(q);
  for (int f = 0; f < a.nf; ++f) {
    const int qv = opaque(q);
    const cpx<T>* src = a.in + f * a.in_fs;
    cpx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int ir = compact_of_row(qv + Q * r, N, a.in_k, a.in_p);
      v[r] = ((kZeroIn >> r) & 1u) || !(active && ir >= 0) ? czero<T>() : ld_off(src, cin + uint32_t(ir) * rs_in);
    }
    F::template run<DIR>(v, qv, L, tw);
    cpx<T>* dst = a.out + f * a.out_fs;
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if ((kDeadOut >> s) & 1u) continue;
      const int o = compact_of_row(F::out_index(qv, s), N, a.out_k, a.out_p);
      if (active && o >= 0) st_off(dst, cout + uint32_t(o) * rs_out, v[s]);
    }
  }
}

// P1, lean: three arrays (u0 stays in registers for both curl parts; the
// twiddles come from the table).
template <int N, class T, bool PIN>
__global__ __launch_bounds__(LeanShape<N>::THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_x_inverse_curl_lean(
    const cpx<T>* __restrict__ us, int64_t us_fs, cpx<T>* __restrict__ out, int64_t out_fs, Geo g,
    FinalizeJob job) {
  using S = LeanShape<N>;
  using F = typename S::F;
  using PR = PrunedRows<F, N>;
  constexpr int R = S::R, Q = S::Q, TC = S::TC;
  constexpr unsigned kZeroIn = PIN ? PR::kIn : 0u;
  constexpr uint32_t E = sizeof(cpx<T>);
  if (job.partials != nullptr && blockIdx.x == 0) finalize_in_block(job);   // uniform per block
  __shared__ cpx<T> lds[S::LDS];
  const int c = threadIdx.x % TC, q = threadIdx.x / TC;
  const int j = xcd_tile(blockIdx.x, gridDim.x) * TC + c;
  const bool active = j < g.ky * g.kzs;   // padding columns (iz >= kz) carry zeros
  const int jj = active ? j : 0;
  const int iy = jj / g.kzs, iz = jj - iy * g.kzs;
  const T kyv = wavenumber<T>(g.k0y, freq_of(iy, g.ky, g.py));
  const T kzv = wavenumber<T>(g.k0z, iz);
  const uint32_t cj = uint32_t(jj) * E, rs = uint32_t(g.jp) * E;
  SlotStrided<TC, T> L{lds + c};
  // every step takes its own opaque copy of q (see opaque())
  auto row = [&](int qv, int r) { return compact_of_row(qv + Q * r, N, g.kx, g.px); };
  auto kx = [&](int qv, int r) {
    const int ir = row(qv, r);
    return wavenumber<T>(g.k0x, ir >= 0 ? freq_of(ir, g.kx, g.px) : 0);
  };
  auto load = [&](int f, cpx<T>* v) {
    const int qv = opaque(q);
    const cpx<T>* src = us + f * us_fs;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int ir = row(qv, r);
      v[r] = ((kZeroIn >> r) & 1u) || !(active && ir >= 0) ? czero<T>() : ld_off(src, cj + uint32_t(ir) * rs);
    }
  };
  auto store = [&](int f, const cpx<T>* v) {
    if (!active) return;
    const int qv = opaque(q);
    cpx<T>* dst = out + f * out_fs;
#pragma unroll
    for (int s = 0; s < R; ++s) st_off(dst, cj + uint32_t(F::out_index(qv, s)) * rs, v[s]);
  };
  auto fft = [&](cpx<T>* v) { const int qv = opaque(q); F::template run<+1>(v, qv, L, Twiddles<F, T, true>(qv)); };
  // pruned inputs (12 live values per field) leave room to issue u1's loads
  // one transform ahead of its use (before u2's transform); otherwise after
  constexpr bool kEarlyU1 = PIN;
  cpx<T> a[R], b[R], c1[R], w[R];
  load(0, a);
  load(2, b);
  {
    const int qv = opaque(q);
#pragma unroll
    for (int r = 0; r < R; ++r) w[r] = mul_i(kzv * a[r] - kx(qv, r) * b[r]);   // w1
  }
  fft(w);
  store(3, w);
  if constexpr (kEarlyU1) load(1, c1);
  fft(b);
  store(2, b);
  if constexpr (!kEarlyU1) load(1, c1);
  {
    const int qv = opaque(q);
#pragma unroll
    for (int r = 0; r < R; ++r) w[r] = mul_i(kx(qv, r) * c1[r] - kyv * a[r]);   // w2
  }
  fft(w);
  store(4, w);
  fft(c1);
  store(1, c1);
  fft(a);
  store(0, a);
}

// P2, lean: three arrays; u1 and u2 are loaded together (w0 needs both).
template <int N, class T, bool PIN>
__global__ __launch_bounds__(LeanShape<N>::THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_y_inverse_curl_lean(
    const cpx<T>* __restrict__ in, int64_t in_fs, cpx<T>* __restrict__ out, int64_t out_fs, Geo g) {
  using S = LeanShape<N>;
  using F = typename S::F;
  using PR = PrunedRows<F, N>;
  constexpr int R = S::R, Q = S::Q, TC = S::TC;
  constexpr unsigned kZeroIn = PIN ? PR::kIn : 0u;
  constexpr uint32_t E = sizeof(cpx<T>);
  __shared__ cpx<T> lds[S::LDS];
  const int c = threadIdx.x % TC, q = threadIdx.x / TC;
  const int j = xcd_tile(blockIdx.x, gridDim.x) * TC + c;
  const bool active = j < g.nx * g.kzs;   // padding columns (iz >= kz) carry zeros
  const int jj = active ? j : 0;
  const int x = jj / g.kzs, iz = jj - x * g.kzs;
  const T kzv = wavenumber<T>(g.k0z, iz);
  const uint32_t cin = uint32_t(int64_t(x) * g.jp + iz) * E;
  const uint32_t cout = uint32_t(int64_t(x) * g.ny * g.kzs + iz) * E;   // x-major pencil buffer
  const uint32_t rs = uint32_t(g.kzs) * E;
  SlotStrided<TC, T> L{lds + c};
  // every step takes its own opaque copy of q (see opaque())
  auto row = [&](int qv, int r) { return compact_of_row(qv + Q * r, N, g.ky, g.py); };
  auto load = [&](int f, cpx<T>* v) {
    const int qv = opaque(q);
    const cpx<T>* src = in + f * in_fs;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int ir = row(qv, r);
      v[r] = ((kZeroIn >> r) & 1u) || !(active && ir >= 0) ? czero<T>() : ld_off(src, cin + uint32_t(ir) * rs);
    }
  };
  auto store = [&](int f, const cpx<T>* v) {
    if (!active) return;
    const int qv = opaque(q);
    cpx<T>* dst = out + f * out_fs;
#pragma unroll
    for (int s = 0; s < R; ++s) st_off(dst, cout + uint32_t(F::out_index(qv, s)) * rs, v[s]);
  };
  auto fft = [&](cpx<T>* v) { const int qv = opaque(q); F::template run<+1>(v, qv, L, Twiddles<F, T, true>(qv)); };
  cpx<T> a[R], b[R], w[R];
  load(1, a);
  load(2, b);
  {
    const int qv = opaque(q);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int ir = row(qv, r);
      const T kyv = wavenumber<T>(g.k0y, ir >= 0 ? freq_of(ir, g.ky, g.py) : 0);
      w[r] = ((kZeroIn >> r) & 1u) ? czero<T>() : mul_i(kyv * b[r] - kzv * a[r]);   // w0
    }
  }
  fft(w);
  store(3, w);
  fft(b);
  store(2, b);
  fft(a);
  store(1, a);
  // the remaining fields are loaded just before their transforms: issuing
  // each one transform ahead made this pass 1.4 % faster at 1024^3 but the
  // whole step 0.7 % slower (the passes after it lost more)
  load(0, w);
  fft(w);
  store(0, w);
  load(3, b);
  fft(b);
  store(4, b);
  load(4, a);
  fft(a);
  store(5, a);
}

// ---------------------------------------------------------------------------
// z-axis helpers: Hermitian extension of two half spectra U, W into the
// spectrum of u + i w (c2r semantics: only Re of the kz = 0 and Nyquist bins).
// ---------------------------------------------------------------------------
// Branch-free: every lane loads (from index 0 when its mode is not stored,
// kz == 0 included) and the value is zeroed by a select, so a thread's loads
// for all its elements issue back to back and share one wait. A branch around
// each load pair would serialise them (one memory round trip per element).
template <int N, class T>
__device__ __forceinline__ cpx<T> herm_pair(const cpx<T>* U, const cpx<T>* W, int k, int kz) {
  const bool mirror = 2 * k > N;
  const int src = mirror ? N - k : k;
  const bool stored = src < kz;
  const int at = stored ? src : 0;
  cpx<T> u = U[at];
  cpx<T> w = W ? W[at] : czero<T>();
  const bool real_mode = src == 0 || 2 * src == N;
  const T uy = (stored && !real_mode) ? u.y : T(0), wy = (stored && !real_mode) ? w.y : T(0);
  const T ux = stored ? u.x : T(0), wx = stored ? w.x : T(0);
  // u + i w, or conj(u) + i conj(w) on the mirrored half
  return mirror ? cpx<T>{ux + wy, wx - uy} : cpx<T>{ux - wy, uy + wx};
}

// The same extension with the index work done once per element position and
// shared by several field pairs.
struct HermIdx {
  int at;          // stored source index (0 when not stored)
  bool mirror, stored, cplx;  // cplx: stored and not a real (kz = 0 / Nyquist) mode
};
template <int N>
__device__ __forceinline__ HermIdx herm_idx(int k, int kz) {
  HermIdx h;
  h.mirror = 2 * k > N;
  const int src = h.mirror ? N - k : k;
  h.stored = src < kz;
  h.at = h.stored ? src : 0;
  h.cplx = h.stored && !(src == 0 || 2 * src == N);
  return h;
}
template <class T>
__device__ __forceinline__ cpx<T> herm_pair_at(const cpx<T>* U, const cpx<T>* W, const HermIdx& h) {
  const cpx<T> u = U[h.at], w = W[h.at];
  const T uy = h.cplx ? u.y : T(0), wy = h.cplx ? w.y : T(0);
  const T ux = h.stored ? u.x : T(0), wx = h.stored ? w.x : T(0);
  return h.mirror ? cpx<T>{ux + wy, wx - uy} : cpx<T>{ux - wy, uy + wx};
}

// The 2/3 truncation of the pruned layout (kz = N/3 + 1 stored modes) leaves
// the modes kz <= k <= N - kz of the Hermitian-extended pencil zero. Bit r of
// kEntry: entry slot r (k = q + Q r) is in that band for every thread q, so it
// is a compile-time zero; bit s of kExit: exit slot s (k = out_index(q, s)) is
// in the band for every q, so the forward transforms' value there is never
// stored or read.
template <class F, int N>
struct ZeroBand {
  static constexpr int kz = N / 3 + 1;
  static constexpr bool in_band(int k) { return k >= kz && k <= N - kz; }
  static constexpr unsigned entry_mask() {
    unsigned m = 0;
    for (int r = 0; r < F::R; ++r) {
      bool all = true;
      for (int q = 0; q < F::Q; ++q) all = all && in_band(q + F::Q * r);
      if (all) m |= 1u << r;
    }
    return m;
  }
  static constexpr unsigned exit_mask() {
    unsigned m = 0;
    for (int s = 0; s < F::R; ++s) {
      bool all = true;
      for (int q = 0; q < F::Q; ++q) all = all && in_band(F::out_index(q, s));
      if (all) m |= 1u << s;
    }
    return m;
  }
  static constexpr unsigned kEntry = entry_mask(), kExit = exit_mask();
};

// Separate A = FFT(a + i b) into FFT(a)[k], FFT(b)[k] using A[N-k].
template <class T>
__device__ __forceinline__ void split_pair(cpx<T> ak, cpx<T> bk, cpx<T>& X, cpx<T>& Y) {
  const T h = T(0.5);
  X = {h * (ak.x + bk.x), h * (ak.y - bk.y)};
  Y = {h * (ak.y + bk.y), h * (bk.x - ak.x)};
}

// ---------------------------------------------------------------------------
// P3: per z pencil, c2r of (u_j + i w_j) for j = 0..2, u x w scaled by
// `scale` (= 1/Ntot^2, a power of two), then r2c. Pencils are kzs apart;
// their kz modes are followed by kzs - kz padding slots, written as zeros so
// that every line of the output is written whole. Outputs (nl0, nl1, nl2) go
// to fields kZOut, kZOut+1, kZOut+2 = 2, 3, 4, in place over inputs the
// pencil has consumed by then.
// Up to 512 points: z0 and z1 are transformed first, and nl2 = u0 w1 - u1 w0,
// which needs only them, is transformed and stored (over w1, field 4) before
// z2 is loaded, so at most two fields' physical values and one transform are
// live at a time (256 points: 1303 -> 1260 VALU instructions per wave, about
// 111 -> 106 us per launch at 256^3). At 1024 points all three fields are
// loaded before the first transform and both products formed after the third
// (the one-pencil waves there gain more from the early loads: 7.3 ms per
// launch at 1024^3 against 9.6 ms in the order above).
// ---------------------------------------------------------------------------
template <int N, class T, bool PRUNED>
__global__ __launch_bounds__(PencilBlock<N>::THREADS) void k_z_fused(cpx<T>* buf, int64_t fs, int nx, int ny,
                                                     int kz, int kzs, T scale, int yrow) {
  using B = PencilBlock<N>;
  using F = typename B::F;
  static_assert(B::THREADS == 64, "WaveSync exchanges need one-wave blocks");
  static_assert(kZOut == 2, "the in-place order below writes fields 2, 3, 4");
  constexpr int R = B::R, Q = B::Q, PW = B::PW;
  __shared__ cpx<T> lds[B::LDS];
  const int q = threadIdx.x % Q, p = threadIdx.x / Q;
  // Workgroups walk the pencils x-major (the order in which the y-inverse
  // pass wrote them), each XCD's run from its end. x-major storage: pencil
  // pp = x ny + y at pp kzs. y-major storage (yrow > 0, nx % PW == 0):
  // pencil (x, y) at y yrow + x kzs; the workgroup takes the PW x-adjacent
  // pencils (x0 + p, y), whose rows are contiguous, so a group of pencils
  // covers whole lines.
  const int64_t gi = xcd_tile_rev(blockIdx.x, gridDim.x);
  int64_t base;
  bool active;
  if (yrow) {
    active = gi < int64_t(nx / PW) * ny;
    // (walking the groups y-major, or in blocks of 4-16 x-groups per y, to
    // keep consecutive workgroups on adjacent memory measured slower: the
    // order the y-inverse pass wrote in matters more)
    const int ga = active ? int(gi) : 0, gx = ga / ny, gy = ga - gx * ny;
    base = int64_t(gy) * yrow + int64_t(gx * PW + p) * kzs;
  } else {
    const int64_t pp = gi * PW + p;
    active = pp < int64_t(nx) * ny;
    base = (active ? pp : 0) * kzs;
  }
  cpx<T>* lp = lds + p * F::kSlots;
  SlotPencil<T> L{lp};
  const int kzl = active ? kz : 0;  // inactive pencils load nothing
  cpx<T> tw[R];
  F::load_tw(tw, q);
  using Z = ZeroBand<F, N>;
  constexpr unsigned kZeroIn = PRUNED ? Z::kEntry : 0u, kUnused = PRUNED ? Z::kExit : 0u;
  // Pruned layout (the launcher guarantees padding, kzs > kz, whose slots hold
  // zeros): each element of the Hermitian-extended pencil is read through one
  // 32-bit byte offset per slot from the wave-uniform field bases, shared by
  // the six fields: the stored source index, or the last padding slot (a
  // zero) when the mode is not stored, so no select is needed except for the
  // kz = 0 mode's imaginary part (the Nyquist mode is never stored here).
  // Whether a slot lies on the mirrored half is a compile-time fact except in
  // the slot that holds N/2 (per lane there).
  // Dense layout: the index work per slot, then selects (herm_pair_at).
  auto mirrored = [&](int r) {
    return (2 * Q * r > N) ? true : ((2 * (Q * r + Q - 1) <= N) ? false : (2 * (q + Q * r) > N));
  };
  HermIdx hi[R];
  uint32_t off[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if constexpr (PRUNED) {
      if ((kZeroIn >> r) & 1u) continue;
      const int k = q + Q * r;
      const int src = mirrored(r) ? N - k : k;
      const int idx = src < kzl ? src : kzs - 1;
      off[r] = (uint32_t(base) + uint32_t(idx)) * uint32_t(sizeof(cpx<T>));
    } else {
      hi[r] = herm_idx<N>(q + Q * r, kzl);
    }
  }
  auto load = [&](int c, cpx<T>* z) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if ((kZeroIn >> r) & 1u) {
        z[r] = czero<T>();
      } else if constexpr (PRUNED) {
        cpx<T> u = ld_off(buf + c * fs, off[r]), w = ld_off(buf + (3 + c) * fs, off[r]);
        if (r == 0 && q == 0) u.y = w.y = T(0);   // kz = 0: c2r keeps the real part
        z[r] = mirrored(r) ? cpx<T>{u.x + w.y, w.x - u.y} : cpx<T>{u.x - w.y, u.y + w.x};
      } else {
        z[r] = herm_pair_at<T>(buf + c * fs + base, buf + (3 + c) * fs + base, hi[r]);
      }
    }
  };
  // nl2 (real) at exit slot s of the inverse transforms, placed at the entry
  // slot of the forward transform; then its retained modes and the padding
  auto nl2_of = [&](const cpx<T>* z0, const cpx<T>* z1, cpx<T>* Bv) {
#pragma unroll
    for (int s = 0; s < R; ++s)
      Bv[F::out_slot_as_input(s)] = {(z0[s].x * z1[s].y - z1[s].x * z0[s].y) * scale, T(0)};
  };
  // nl2 keeps only k < kz; in the pruned layout kz = N/3 + 1 is known here,
  // so the exit slots above it (5 of 8 at 256 points, 10 of 16 at 1024) are
  // never computed
  // output element k of this pencil in field f (pruned: uniform field base
  // and a 32-bit offset, as for the loads)
  auto put = [&](int f, int k, cpx<T> v) {
    if constexpr (PRUNED) st_off(buf + f * fs, (uint32_t(base) + uint32_t(k)) * uint32_t(sizeof(cpx<T>)), v);
    else buf[f * fs + base + k] = v;
  };
  auto store_nl2 = [&](const cpx<T>* Bv) {
    if (!active) return;
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if ((kUnused >> s) & 1u) continue;
      const int k = F::out_index(q, s);
      if (k < (PRUNED ? Z::kz : kz)) put(kZOut + 2, k, Bv[s]);
    }
    for (int k = kz + q; k < kzs; k += Q) put(kZOut + 2, k, czero<T>());
  };
  // physical space: z_c = u_c + i w_c at position out_index(q, s)
  cpx<T> A[R];
  if constexpr (N >= 1024) {
    cpx<T> z0[R], z1[R], z2[R];
    load(0, z0);
    load(1, z1);
    load(2, z2);
    F::template run<+1, WaveSync>(z0, q, L, tw);
    F::template run<+1, WaveSync>(z1, q, L, tw);
    F::template run<+1, WaveSync>(z2, q, L, tw);
    cpx<T> Bv[R];
    nl2_of(z0, z1, Bv);
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const T u0 = z0[s].x, u1 = z1[s].x, u2 = z2[s].x;
      const T w0 = z0[s].y, w1 = z1[s].y, w2 = z2[s].y;
      A[F::out_slot_as_input(s)] = {(u1 * w2 - u2 * w1) * scale, (u2 * w0 - u0 * w2) * scale};
    }
    F::template run<-1, WaveSync>(Bv, q, L, tw);
    store_nl2(Bv);
  } else {
    cpx<T> z0[R], z1[R];
    load(0, z0);
    F::template run<+1, WaveSync>(z0, q, L, tw);
    load(1, z1);
    F::template run<+1, WaveSync>(z1, q, L, tw);
    {
      cpx<T> Bv[R];
      nl2_of(z0, z1, Bv);
      F::template run<-1, WaveSync>(Bv, q, L, tw);
      store_nl2(Bv);
    }
    cpx<T> z2[R];
    load(2, z2);
    F::template run<+1, WaveSync>(z2, q, L, tw);
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const T u0 = z0[s].x, u1 = z1[s].x, u2 = z2[s].x;
      const T w0 = z0[s].y, w1 = z1[s].y, w2 = z2[s].y;
      A[F::out_slot_as_input(s)] = {(u1 * w2 - u2 * w1) * scale, (u2 * w0 - u0 * w2) * scale};
    }
  }
  F::template run<-1, WaveSync>(A, q, L, tw);
  WaveSync::sync();
#pragma unroll
  for (int s = 0; s < R; ++s)
    if (!((kUnused >> s) & 1u)) lp[F::out_index(q, s)] = A[s];
  WaveSync::sync();
  if (active) {
    for (int k = q; k < kzs; k += Q) {
      cpx<T> X = czero<T>(), Y = czero<T>();
      if (k < kz) split_pair(lp[k], lp[(N - k) & (N - 1)], X, Y);
      put(kZOut, k, X);
      put(kZOut + 1, k, Y);
    }
  }
}

// Fields [f0, f1) of the standalone z passes' block, whole pairs of fields
// (one complex transform each): the launchers split the pairs over
// gridDim.y when the pencils alone would leave most CUs idle.
struct FieldPairs {
  int f0, f1;
  __device__ __forceinline__ explicit FieldPairs(int nf) {
    const int pairs = (nf + 1) / 2, ppb = (pairs + int(gridDim.y) - 1) / int(gridDim.y);
    f0 = 2 * ppb * int(blockIdx.y);
    f1 = min(nf, f0 + 2 * ppb);
  }
};

template <int N, class T>
__global__ __launch_bounds__(PencilBlock<N>::THREADS) void k_z_c2r(const cpx<T>* __restrict__ in, int64_t in_fs,
                                                   T* __restrict__ out, int64_t out_fs, int nf,
                                                   int64_t npencils, int kz, T scale) {
  using B = PencilBlock<N>;
  using F = typename B::F;
  static_assert(B::THREADS == 64, "WaveSync exchanges need one-wave blocks");
  constexpr int R = B::R, Q = B::Q, PW = B::PW;
  __shared__ cpx<T> lds[B::LDS];
  const int q = threadIdx.x % Q, p = threadIdx.x / Q;
  const int64_t pp = int64_t(blockIdx.x) * PW + p;
  const bool active = pp < npencils;
  const int64_t ib = (active ? pp : 0) * kz, ob = (active ? pp : 0) * N;
  SlotPencil<T> L{lds + p * F::kSlots};
  const int kzl = active ? kz : 0;
  cpx<T> tw[R];
  F::load_tw(tw, q);
  const FieldPairs fp(nf);
  for (int f = fp.f0; f < fp.f1; f += 2) {
    const bool two = f + 1 < nf;
    cpx<T> z[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
      z[r] = herm_pair<N, T>(in + f * in_fs + ib, two ? in + (f + 1) * in_fs + ib : nullptr,
                             q + Q * r, kzl);
    F::template run<+1, WaveSync>(z, q, L, tw);
    if (active) {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        const int n = F::out_index(q, s);
        out[f * out_fs + ob + n] = z[s].x * scale;
        if (two) out[(f + 1) * out_fs + ob + n] = z[s].y * scale;
      }
    }
  }
}

// Standalone z r2c of nf real fields, two per complex transform.
template <int N, class T>
__global__ __launch_bounds__(PencilBlock<N>::THREADS) void k_z_r2c(const T* __restrict__ in, int64_t in_fs,
                                                   cpx<T>* __restrict__ out, int64_t out_fs,
                                                   int nf, int64_t npencils, int kz) {
  using B = PencilBlock<N>;
  using F = typename B::F;
  static_assert(B::THREADS == 64, "WaveSync exchanges need one-wave blocks");
  constexpr int R = B::R, Q = B::Q, PW = B::PW;
  __shared__ cpx<T> lds[B::LDS];
  const int q = threadIdx.x % Q, p = threadIdx.x / Q;
  const int64_t pp = int64_t(blockIdx.x) * PW + p;
  const bool active = pp < npencils;
  const int64_t ib = (active ? pp : 0) * N, ob = (active ? pp : 0) * kz;
  cpx<T>* lp = lds + p * F::kSlots;
  SlotPencil<T> L{lp};
  cpx<T> tw[R];
  F::load_tw(tw, q);
  const FieldPairs fp(nf);
  for (int f = fp.f0; f < fp.f1; f += 2) {
    const bool two = f + 1 < nf;
    cpx<T> z[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int n = q + Q * r;
      z[r] = {active ? in[f * in_fs + ib + n] : T(0),
              (active && two) ? in[(f + 1) * in_fs + ib + n] : T(0)};
    }
    F::template run<-1, WaveSync>(z, q, L, tw);
    WaveSync::sync();
#pragma unroll
    for (int s = 0; s < R; ++s) lp[F::out_index(q, s)] = z[s];
    WaveSync::sync();
    if (active) {
      for (int k = q; k < kz; k += Q) {
        cpx<T> X, Y;
        split_pair(lp[k], lp[(N - k) & (N - 1)], X, Y);
        out[f * out_fs + ob + k] = X;
        if (two) out[(f + 1) * out_fs + ob + k] = Y;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Adjoint z pass: per z pencil, c2r of the nine fields u, w = K u, g in the
// pairs (u0 + i u1), (u2 + i w0), (w1 + i w2), (g0 + i g1), g2, scaled by
// `scale` (1/Ntot), the cross products a = w x g and b = g x u, and r2c of
// (a0 + i a1), (a2 + i b0), (b1 + i b2) into out fields 0..5 (all nzh modes).
// Input pencils hold kz modes kzs apart (the pruned layout, PRUNED: entry
// slots in the 2/3 truncation's zero band are compile-time zeros) or nzh
// (dense; then out may equal in: each pencil's loads precede its stores).
// The same pairs, scaling and products as z c2r -> cross -> z r2c over
// physical buffers, without the 15 real fields' round trip.
// ---------------------------------------------------------------------------
template <int N, class T, bool PRUNED>
__global__ __launch_bounds__(PencilBlock<N>::THREADS) void k_z_adj(const cpx<T>* in, int64_t in_fs, int kz,
                                                   int kzs, cpx<T>* out, int64_t out_fs, int nzh,
                                                   int64_t npencils, T scale) {
  using B = PencilBlock<N>;
  using F = typename B::F;
  static_assert(B::THREADS == 64, "WaveSync exchanges need one-wave blocks");
  constexpr int R = B::R, Q = B::Q, PW = B::PW;
  __shared__ cpx<T> lds[B::LDS];
  const int q = threadIdx.x % Q, p = threadIdx.x / Q;
  const int64_t pp = int64_t(blockIdx.x) * PW + p;
  const bool active = pp < npencils;
  const int64_t ib = (active ? pp : 0) * kzs, ob = (active ? pp : 0) * nzh;
  cpx<T>* lp = lds + p * F::kSlots;
  SlotPencil<T> L{lp};
  const int kzl = active ? kz : 0;  // inactive pencils load nothing
  constexpr unsigned kZeroIn = PRUNED ? ZeroBand<F, N>::kEntry : 0u;
  cpx<T> tw[R];
  F::load_tw(tw, q);
  cpx<T> z[5][R];
#pragma unroll
  for (int c = 0; c < 5; ++c) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      z[c][r] = ((kZeroIn >> r) & 1u) ? czero<T>()
                                      : herm_pair<N, T>(in + 2 * c * in_fs + ib,
                                                        c < 4 ? in + (2 * c + 1) * in_fs + ib : nullptr,
                                                        q + Q * r, kzl);
    F::template run<+1, WaveSync>(z[c], q, L, tw);
  }
  cpx<T> X[3][R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const T u0 = z[0][s].x * scale, u1 = z[0][s].y * scale, u2 = z[1][s].x * scale;
    const T w0 = z[1][s].y * scale, w1 = z[2][s].x * scale, w2 = z[2][s].y * scale;
    const T g0 = z[3][s].x * scale, g1 = z[3][s].y * scale, g2 = z[4][s].x * scale;
    const int r = F::out_slot_as_input(s);
    X[0][r] = {w1 * g2 - w2 * g1, w2 * g0 - w0 * g2};
    X[1][r] = {w0 * g1 - w1 * g0, g1 * u2 - g2 * u1};
    X[2][r] = {g2 * u0 - g0 * u2, g0 * u1 - g1 * u0};
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    F::template run<-1, WaveSync>(X[c], q, L, tw);
    WaveSync::sync();
#pragma unroll
    for (int s = 0; s < R; ++s) lp[F::out_index(q, s)] = X[c][s];
    WaveSync::sync();
    if (active) {
      for (int k = q; k < nzh; k += Q) {
        cpx<T> a, b;
        split_pair(lp[k], lp[(N - k) & (N - 1)], a, b);
        out[2 * c * out_fs + ob + k] = a;
        out[(2 * c + 1) * out_fs + ob + k] = b;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Deterministic block reduction of three doubles (blockDim a multiple of 64, <= 1024):
// xor-butterfly inside each wave, then the four wave sums in fixed order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void block_sum3(double a, double b, double c, double* dst) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
    c += __shfl_xor(c, off);
  }
  __shared__ double sh[3][16];
  const int w = threadIdx.x / 64, nw = blockDim.x / 64;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = a;
    sh[1][w] = b;
    sh[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < 3; ++i) {
      double t = sh[i][0];
      for (int k = 1; k < nw; ++k) t += sh[i][k];
      dst[i] = t;
    }
  }
}

// exp(-x) in the state precision (x >= 0): the integrating factor
template <class T>
__device__ __forceinline__ T exp_neg(T x) {
  if constexpr (sizeof(T) == 4) return expf(-x); else return exp(-x);
}

// ---------------------------------------------------------------------------
// P5: x-forward of (nl0, nl1, nl2), then per retained mode
//   du = P[mask (NL + f)] - nu k2 us          (navierstokes.py:89-95)
// and the RK4 stage update (integrators.py:34-38).
// ---------------------------------------------------------------------------
// LEAN: the 1024-point float form for the pruned layout on 16-column
// workgroups within 128 VGPRs: rows that the compact state does not hold are
// compile-time dead (PrunedRows), so a transformed field keeps 12 of its 16
// values per thread; the fields are loaded one transform ahead, the twiddles
// read from the table, and the state is addressed by 32-bit offsets from
// uniform field bases (the launcher checks they fit).
template <int N, bool LEAN>
using StageShape = std::conditional_t<LEAN, LeanShape<N>, StridedShape<N>>;

// number of slots below s not set in DEAD
template <unsigned DEAD>
__host__ __device__ constexpr int live_rank(int s) {
  int r = 0;
  for (int i = 0; i < s; ++i) r += ((DEAD >> i) & 1u) ? 0 : 1;
  return r;
}

template <int N, int MODE, class T, bool LEAN = false, bool POUT = LEAN>
__global__ __launch_bounds__((StageShape<N, LEAN>::THREADS)) __attribute__((amdgpu_waves_per_eu(LEAN ? 4 : MinWaves<N, T>::value))) void k_x_forward_stage(StageArgs<T> a) {
  using S = StageShape<N, LEAN>;
  using F = typename S::F;
  constexpr int R = S::R, Q = S::Q, TC = S::TC;
  constexpr unsigned kDeadOut = POUT ? PrunedRows<F, N>::kOut : 0u;
  // lean: live slots per field, and how many of the second field's fit in
  // LDS beside the first's
  constexpr int kLive = live_rank<kDeadOut>(R);
  constexpr int kPark1 = LEAN ? (S::LDS / S::THREADS - kLive < kLive ? S::LDS / S::THREADS - kLive : kLive) : 0;
  __shared__ cpx<T> lds[S::LDS];
  const Geo& g = a.g;
  const int c = threadIdx.x % TC, q = threadIdx.x / TC;
  const int j = xcd_tile(blockIdx.x, gridDim.x) * TC + c;
  const bool active = j < g.ky * g.kzs;   // padding columns: in_box is false, zeros stay zeros
  const int jj = active ? j : 0;
  const int iy = jj / g.kzs, iz = jj - iy * g.kzs;
  const int ny = freq_of(iy, g.ky, g.py);
  const T kyv = wavenumber<T>(g.k0y, ny);
  const T kzv = wavenumber<T>(g.k0z, iz);
  SlotStrided<TC, T> L{lds + c};
  cpx<T> v0[R], v1[R], v2[R];
  cpx<T>* vv[3] = {v0, v1, v2};
  if constexpr (LEAN) {
    constexpr uint32_t E = sizeof(cpx<T>);
    auto load = [&](int f) {
      const int qv = opaque(q);
      const cpx<T>* src = a.nl + f * a.nl_fs;
#pragma unroll
      for (int r = 0; r < R; ++r)
        vv[f][r] = active ? ld_off(src, (uint32_t(qv + Q * r) * uint32_t(g.jp) + uint32_t(jj)) * E) : czero<T>();
    };
    auto fft = [&](int f) {
      const int qv = opaque(q);
      F::template run<-1>(vv[f], qv, L, Twiddles<F, T, true>(qv));
    };
    load(0);
    load(1);
    fft(0);
    load(2);
    fft(1);
    fft(2);
    // the transformed first field (its live slots) and as many of the
    // second's as fit wait in LDS while the epilogue runs, which frees their
    // registers; rows of THREADS values, thread-major: conflict-free
    static_assert(S::LDS >= (kLive + kPark1) * S::THREADS, "parking area");
    __syncthreads();
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if ((kDeadOut >> s) & 1u) continue;
      const int rk = live_rank<kDeadOut>(s);
      lds[rk * S::THREADS + threadIdx.x] = v0[s];
      if (rk < kPark1) lds[(kLive + rk) * S::THREADS + threadIdx.x] = v1[s];
    }
  } else {
    cpx<T> tw[R];
    F::load_tw(tw, q);
    // all three fields are loaded before the first transform (see k_z_fused)
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const cpx<T>* src = a.nl + f * a.nl_fs + jj;
#pragma unroll
      for (int r = 0; r < R; ++r) vv[f][r] = active ? src[int64_t(q + Q * r) * g.jp] : czero<T>();
    }
#pragma unroll
    for (int f = 0; f < 3; ++f) F::template run<-1>(vv[f], q, L, tw);
  }
  // state accesses of the epilogue: field f at mode offset o
  auto LD = [&](const cpx<T>* p, int f, int64_t o) -> cpx<T> {
    if constexpr (LEAN) return ld_off(p + f * a.st_fs, uint32_t(o) * uint32_t(sizeof(cpx<T>)));
    else return p[o + f * a.st_fs];
  };
  auto ST = [&](cpx<T>* p, int f, int64_t o, cpx<T> v) {
    if constexpr (LEAN) st_off(p + f * a.st_fs, uint32_t(o) * uint32_t(sizeof(cpx<T>)), v);
    else p[o + f * a.st_fs] = v;
  };
  T coef = a.coef_const;
  if (a.fterm != kTermNone && a.coef_dev) coef = T(*a.coef_dev);
  const int wgt = hermitian_weight(iz, g.nz);
  double sE = 0.0, sO = 0.0, sB = 0.0;
  const T hdt = T(0.5) * a.dt, sdt = a.dt / T(6), tdt = a.dt / T(3);
#pragma unroll
  for (int s = 0; s < R; ++s) {
    if ((kDeadOut >> s) & 1u) continue;
    // lean: an opaque index per mode keeps the compiler from issuing every
    // mode's state loads at once (which spills)
    const int qs = LEAN ? opaque(q) : q;
    const int ix = compact_of_row(F::out_index(qs, s), N, g.kx, g.px);
    if (!active || ix < 0) continue;
    const int nx = freq_of(ix, g.kx, g.px);
    const T kxv = wavenumber<T>(g.k0x, nx);
    const T k2 = k2_of(kxv, kyv, kzv);
    const T ik2 = inv_k2_of(k2);
    const bool keep = in_box(nx, ny, iz, g);
    const int64_t o = int64_t(ix) * g.jp + j;
    const cpx<T> s0 = LD(a.us, 0, o), s1 = LD(a.us, 1, o), s2 = LD(a.us, 2, o);
    T fc = T(0);
    if (a.fterm == kTermUniform) fc = coef;
    if (a.fterm == kTermBand) fc = (a.kmin2 <= k2 && k2 <= a.kmax2) ? coef : T(0);
    const int rk = live_rank<kDeadOut>(s);
    const cpx<T> x0 = LEAN ? lds[rk * S::THREADS + threadIdx.x] : v0[s];
    const cpx<T> x1 = LEAN && rk < kPark1 ? lds[(kLive + rk) * S::THREADS + threadIdx.x] : v1[s];
    cpx<T> d0 = keep ? x0 + fc * s0 : czero<T>();
    cpx<T> d1 = keep ? x1 + fc * s1 : czero<T>();
    cpx<T> d2 = keep ? v2[s] + fc * s2 : czero<T>();
    const cpx<T> div = ik2 * (kxv * d0 + kyv * d1 + kzv * d2);
    const T nk2 = a.nu * k2;
    constexpr bool IF = MODE >= kStageIf1;
    if constexpr (IF) {  // the integrating factor carries the viscous term
      d0 = d0 - kxv * div;
      d1 = d1 - kyv * div;
      d2 = d2 - kzv * div;
    } else {
      d0 = d0 - kxv * div - nk2 * s0;
      d1 = d1 - kyv * div - nk2 * s1;
      d2 = d2 - kzv * div - nk2 * s2;
    }
    cpx<T> o0, o1, o2;
    if constexpr (IF) {
      // Lawson RK4 with E_h = exp(-nu k2 dt/2), E_f = E_h^2 (u: step start):
      //   u2 = E_h (u + dt/2 k1),  u3 = E_h u + dt/2 k2,  u4 = E_f u + dt E_h k3,
      //   u_new = E_f u + dt/6 (E_f k1 + 2 E_h (k2 + k3)) + dt/6 k4
      const T eh = exp_neg<T>(T(0.5) * nk2 * a.dt), ef = eh * eh;
      if constexpr (MODE == kStageIf4) {
        o0 = LD(a.acc, 0, o) + sdt * d0;
        o1 = LD(a.acc, 1, o) + sdt * d1;
        o2 = LD(a.acc, 2, o) + sdt * d2;
      } else {
        // stage 1's input is the step-start state itself: its values are in hand
        const bool same = MODE == kStageIf1 && !LEAN && a.u == a.us;   // uniform (lean: more spill)
        const cpx<T> b0 = same ? s0 : LD(a.u, 0, o), b1 = same ? s1 : LD(a.u, 1, o),
                     b2 = same ? s2 : LD(a.u, 2, o);
        if constexpr (MODE == kStageIf1) {
          o0 = eh * (b0 + hdt * d0); o1 = eh * (b1 + hdt * d1); o2 = eh * (b2 + hdt * d2);
          ST(a.acc, 0, o, ef * (b0 + sdt * d0));
          ST(a.acc, 1, o, ef * (b1 + sdt * d1));
          ST(a.acc, 2, o, ef * (b2 + sdt * d2));
        } else {
          const T tde = tdt * eh;
          if constexpr (MODE == kStageIf2) {
            o0 = eh * b0 + hdt * d0; o1 = eh * b1 + hdt * d1; o2 = eh * b2 + hdt * d2;
          } else {
            const T fde = a.dt * eh;
            o0 = ef * b0 + fde * d0; o1 = ef * b1 + fde * d1; o2 = ef * b2 + fde * d2;
          }
          ST(a.acc, 0, o, LD(a.acc, 0, o) + tde * d0);
          ST(a.acc, 1, o, LD(a.acc, 1, o) + tde * d1);
          ST(a.acc, 2, o, LD(a.acc, 2, o) + tde * d2);
        }
      }
    } else if constexpr (MODE == kStageRhs) {
      o0 = d0; o1 = d1; o2 = d2;
    } else if constexpr (MODE == kStage4) {
      o0 = LD(a.acc, 0, o) + sdt * d0;
      o1 = LD(a.acc, 1, o) + sdt * d1;
      o2 = LD(a.acc, 2, o) + sdt * d2;
    } else {
      // stage 1's input is the step-start state itself: its values are in hand
      const bool same = MODE == kStage1 && !LEAN && a.u == a.us;   // uniform (lean: more spill)
      const cpx<T> b0 = same ? s0 : LD(a.u, 0, o), b1 = same ? s1 : LD(a.u, 1, o),
                   b2 = same ? s2 : LD(a.u, 2, o);
      const T h = MODE == kStage3 ? a.dt : hdt;
      o0 = b0 + h * d0; o1 = b1 + h * d1; o2 = b2 + h * d2;
      if constexpr (MODE == kStage1) {
        ST(a.acc, 0, o, b0 + sdt * d0);
        ST(a.acc, 1, o, b1 + sdt * d1);
        ST(a.acc, 2, o, b2 + sdt * d2);
      } else {
        ST(a.acc, 0, o, LD(a.acc, 0, o) + tdt * d0);
        ST(a.acc, 1, o, LD(a.acc, 1, o) + tdt * d1);
        ST(a.acc, 2, o, LD(a.acc, 2, o) + tdt * d2);
      }
    }
    ST(a.out, 0, o, o0);
    ST(a.out, 1, o, o1);
    ST(a.out, 2, o, o2);
    if (a.partials) {
      const double e = double(o0.x) * o0.x + double(o0.y) * o0.y + double(o1.x) * o1.x +
                       double(o1.y) * o1.y + double(o2.x) * o2.x + double(o2.y) * o2.y;
      const double we = wgt * e;
      sE += we;
      sO += we * double(k2);
      if (a.kmin2 <= k2 && k2 <= a.kmax2) sB += we;
    }
  }
  if (a.partials) block_sum3(sE, sO, sB, a.partials + 3 * blockIdx.x);
}

// ---------------------------------------------------------------------------
// Host launchers: runtime N -> template instance
// ---------------------------------------------------------------------------
#define H3_DISPATCH_N(N, CALL)                 \
  switch (N) {                                 \
    case 4: { constexpr int NN = 4; CALL; } break;       \
    case 8: { constexpr int NN = 8; CALL; } break;       \
    case 16: { constexpr int NN = 16; CALL; } break;     \
    case 32: { constexpr int NN = 32; CALL; } break;     \
    case 64: { constexpr int NN = 64; CALL; } break;     \
    case 128: { constexpr int NN = 128; CALL; } break;   \
    case 256: { constexpr int NN = 256; CALL; } break;   \
    case 512: { constexpr int NN = 512; CALL; } break;   \
    case 1024: { constexpr int NN = 1024; CALL; } break; \
    default: return hipErrorInvalidValue;      \
  }

static inline unsigned ceil_div(int64_t a, int64_t b) { return unsigned((a + b - 1) / b); }

// Blocks along y for a pass whose nb column (pencil) blocks each loop over
// `items` fields (field pairs): enough to reach about `want` blocks when nb
// is under half of it, with no y-block left empty.
static inline unsigned split_count(unsigned nb, int items, unsigned want) {
  if (items <= 1 || nb == 0 || nb >= want / 2) return 1;
  const unsigned k = std::min<unsigned>(unsigned(items), (want + nb - 1) / nb);
  const unsigned per = (unsigned(items) + k - 1) / k;
  return (unsigned(items) + per - 1) / per;
}

// The lean kernels address with 32-bit byte offsets inside a field.
static inline bool fits_u32(int64_t max_elem, size_t esize) {
  return max_elem >= 0 && uint64_t(max_elem + 1) * esize <= (uint64_t(1) << 32);
}
template <class T>
static constexpr bool lean_type() { return sizeof(T) == 4 && H3_LEAN1024; }

// the pruned layout's retained row count along an axis of N points
static inline bool pruned_rows(int N, int k, int p) { return k == 2 * (N / 3) + 1 && p == (k + 1) / 2 && k < N; }

template <int DIR, class T, bool PIN, bool POUT>
static void launch_strided_lean(const StridedArgs<T>& a, hipStream_t s) {
  constexpr int NN = 1024;
  const unsigned nb = ceil_div(a.J, LeanShape<NN>::TC);
  hipLaunchKernelGGL((k_strided_lean<NN, DIR, T, PIN, POUT>), dim3(nb), dim3(LeanShape<NN>::THREADS), 0, s, a);
}

template <class T>
hipError_t launch_strided(int N, int dir, const StridedArgs<T>& a, hipStream_t s) {
  if (a.J <= 0 || a.nf <= 0) return hipSuccess;
  if constexpr (lean_type<T>()) {
    const int64_t ao = (a.J - 1) / a.B;
    const bool fit = fits_u32(ao * a.in_slab + (a.B - 1) + int64_t(N - 1) * a.in_row(), sizeof(cpx<T>)) &&
                     fits_u32(ao * a.out_slab + (a.B - 1) + int64_t(N - 1) * a.out_row(), sizeof(cpx<T>));
    if (N == 1024 && fit) {
      const bool pin = pruned_rows(N, a.in_k, a.in_p), pout = pruned_rows(N, a.out_k, a.out_p);
      if (dir < 0) {
        if (pin) { if (pout) launch_strided_lean<-1, T, true, true>(a, s); else launch_strided_lean<-1, T, true, false>(a, s); }
        else { if (pout) launch_strided_lean<-1, T, false, true>(a, s); else launch_strided_lean<-1, T, false, false>(a, s); }
      } else {
        if (pin) { if (pout) launch_strided_lean<+1, T, true, true>(a, s); else launch_strided_lean<+1, T, true, false>(a, s); }
        else { if (pout) launch_strided_lean<+1, T, false, true>(a, s); else launch_strided_lean<+1, T, false, false>(a, s); }
      }
      return hipGetLastError();
    }
  }
  const bool pin = pruned_rows(N, a.in_k, a.in_p), pout = pruned_rows(N, a.out_k, a.out_p);
  H3_DISPATCH_N(N, {
    const unsigned nb = ceil_div(a.J, StridedShape<NN>::TC);
    const dim3 grid(nb, split_count(nb, a.nf, 1024));
    const dim3 blk(StridedShape<NN>::THREADS);
    if constexpr (kPrunedN<NN>) {
      if (pin && dir > 0) hipLaunchKernelGGL((k_strided<NN, +1, T, true, false>), grid, blk, 0, s, a);
      else if (pout && dir < 0) hipLaunchKernelGGL((k_strided<NN, -1, T, false, true>), grid, blk, 0, s, a);
      else if (dir < 0) hipLaunchKernelGGL((k_strided<NN, -1, T>), grid, blk, 0, s, a);
      else hipLaunchKernelGGL((k_strided<NN, +1, T>), grid, blk, 0, s, a);
    } else {
      if (dir < 0) hipLaunchKernelGGL((k_strided<NN, -1, T>), grid, blk, 0, s, a);
      else hipLaunchKernelGGL((k_strided<NN, +1, T>), grid, blk, 0, s, a);
    }
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_x_inverse_curl(int N, const cpx<T>* us, int64_t us_fs, cpx<T>* out,
                                 int64_t out_fs, const Geo& g, hipStream_t s,
                                 const FinalizeJob* job) {
  FinalizeJob j{};
  if (job) j = *job;
  if constexpr (lean_type<T>()) {
    if (N == 1024 && fits_u32(int64_t(g.kx) * g.jp, sizeof(cpx<T>)) &&
        fits_u32(int64_t(g.nx) * g.jp, sizeof(cpx<T>))) {
      constexpr int NN = 1024;
      const unsigned nb = ceil_div(int64_t(g.ky) * g.kzs, LeanShape<NN>::TC);
      if (pruned_rows(N, g.kx, g.px))
        hipLaunchKernelGGL((k_x_inverse_curl_lean<NN, T, true>), dim3(nb), dim3(LeanShape<NN>::THREADS), 0, s, us,
                           us_fs, out, out_fs, g, j);
      else
        hipLaunchKernelGGL((k_x_inverse_curl_lean<NN, T, false>), dim3(nb), dim3(LeanShape<NN>::THREADS), 0, s, us,
                           us_fs, out, out_fs, g, j);
      return hipGetLastError();
    }
  }
  const bool pin = pruned_rows(N, g.kx, g.px);
  H3_DISPATCH_N(N, {
    const unsigned nb = ceil_div(int64_t(g.ky) * g.kzs, CurlShape<NN>::TC);
    // under 512 column tiles (one round of two per CU): the outputs in two
    // groups over gridDim.y (128^3: 255 tiles, P1 12.9 -> 11.9 us, 4168 ->
    // 4255 RK4 steps/s; at 256^3's 941 tiles the groups measured slower)
    const dim3 grid(nb, nb < 512 ? 2 : 1);
    const dim3 blk(CurlShape<NN>::THREADS);
    if constexpr (kPrunedN<NN>) {
      if (pin) {
        hipLaunchKernelGGL((k_x_inverse_curl<NN, T, true>), grid, blk, 0, s, us, us_fs, out, out_fs, g, j);
        break;
      }
    }
    hipLaunchKernelGGL((k_x_inverse_curl<NN, T>), grid, blk, 0, s, us, us_fs, out, out_fs, g, j);
  });
  return hipGetLastError();
}

// y-major rows are padded by 32 complex values (256 bytes for complex64):
// measured at 256^3, pads of 0 / 16 / 32 / 48 / 64 / 128 / 256 values gave
// 698.6 / 702.8 / 705.4 / 697 / 700 / 700 / 698 RK4 steps/s
#ifndef H3_PENCIL_PAD
#define H3_PENCIL_PAD 32
#endif
template <class T>
int pencil_row(const Geo& g) {
  // rows off 128-byte lines, a field small enough that rows nx kzs apart
  // stay within few 2 MB pages, and whole groups of x-adjacent pencils
  const int64_t row_bytes = int64_t(g.kzs) * int64_t(sizeof(cpx<T>));
  const int64_t field_bytes = int64_t(g.nx) * g.ny * row_bytes;
  int pw = 0;
  switch (g.nz) {
#define H3_PW(NN) case NN: pw = PencilBlock<NN>::PW; break;
    H3_PW(4) H3_PW(8) H3_PW(16) H3_PW(32) H3_PW(64) H3_PW(128) H3_PW(256) H3_PW(512) H3_PW(1024)
#undef H3_PW
    default: return 0;
  }
  if (row_bytes % 128 == 0 || field_bytes > (int64_t(256) << 20) || g.nx % pw != 0) return 0;
  return g.nx * g.kzs + H3_PENCIL_PAD;
}

template <class T>
hipError_t launch_y_inverse_curl(int N, const cpx<T>* in, int64_t in_fs, cpx<T>* out,
                                 int64_t out_fs, const Geo& g, int yrow, hipStream_t s) {
  if constexpr (lean_type<T>()) {
    if (N == 1024 && !yrow && fits_u32(int64_t(g.nx) * g.jp, sizeof(cpx<T>)) &&
        fits_u32(int64_t(g.nx) * g.ny * g.kzs, sizeof(cpx<T>))) {
      constexpr int NN = 1024;
      const unsigned nb = ceil_div(int64_t(g.nx) * g.kzs, LeanShape<NN>::TC);
      if (pruned_rows(N, g.ky, g.py))
        hipLaunchKernelGGL((k_y_inverse_curl_lean<NN, T, true>), dim3(nb), dim3(LeanShape<NN>::THREADS), 0, s, in,
                           in_fs, out, out_fs, g);
      else
        hipLaunchKernelGGL((k_y_inverse_curl_lean<NN, T, false>), dim3(nb), dim3(LeanShape<NN>::THREADS), 0, s, in,
                           in_fs, out, out_fs, g);
      return hipGetLastError();
    }
  }
  const bool pin = pruned_rows(N, g.ky, g.py);
  H3_DISPATCH_N(N, {
    const unsigned nb = ceil_div(int64_t(g.nx) * g.kzs, CurlShape<NN>::TC);
    // under 512 column tiles (one round of two per CU): the outputs in two
    // groups of three over gridDim.y (128^3: 384 tiles, P2 16.8 -> 14.9 us,
    // 4068 -> 4195 RK4 steps/s; the 256^3 halves' 704 tiles: no change)
    const dim3 grid(nb, nb < 512 ? 2 : 1);
    const dim3 blk(CurlShape<NN>::THREADS);
    if constexpr (kPrunedN<NN>) {
      if (pin) {
        hipLaunchKernelGGL((k_y_inverse_curl<NN, T, true>), grid, blk, 0, s, in, in_fs, out, out_fs, g,
                           yrow);
        break;
      }
    }
    hipLaunchKernelGGL((k_y_inverse_curl<NN, T>), grid, blk, 0, s, in, in_fs, out, out_fs, g, yrow);
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_z_fused(int N, cpx<T>* buf, int64_t fs, int nx, int ny, int kz, int kzs,
                          int yrow, T scale, hipStream_t s) {
  const int64_t npencils = int64_t(nx) * ny;
  // the pruned layout stores exactly kz = N/3 + 1 modes; anything else (the
  // dense layout, kz = N/2 + 1) takes the kernel without compile-time zeros
  // (and padding slots beyond kz, read as zeros, with 32-bit offsets inside
  // a field)
  const int64_t span = yrow ? int64_t(ny) * yrow : npencils * kzs;   // elements of a field
  const bool pruned = kz == N / 3 + 1 && kz < N / 2 + 1 && kzs > kz && fits_u32(span, sizeof(cpx<T>));
  H3_DISPATCH_N(N, {
    const unsigned nb = ceil_div(npencils, PencilBlock<NN>::PW);
    if (pruned)
      hipLaunchKernelGGL((k_z_fused<NN, T, true>), dim3(nb), dim3(PencilBlock<NN>::THREADS), 0, s, buf, fs,
                         nx, ny, kz, kzs, scale, yrow);
    else
      hipLaunchKernelGGL((k_z_fused<NN, T, false>), dim3(nb), dim3(PencilBlock<NN>::THREADS), 0, s, buf, fs,
                         nx, ny, kz, kzs, scale, yrow);
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_z_c2r(int N, const cpx<T>* in, int64_t in_fs, T* out, int64_t out_fs, int nf,
                        int64_t npencils, int kz, T scale, hipStream_t s) {
  H3_DISPATCH_N(N, {
    const unsigned nb = ceil_div(npencils, PencilBlock<NN>::PW);
    hipLaunchKernelGGL((k_z_c2r<NN, T>), dim3(nb, split_count(nb, (nf + 1) / 2, 2048)), dim3(PencilBlock<NN>::THREADS), 0, s, in, in_fs, out, out_fs,
                       nf, npencils, kz, scale);
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_z_r2c(int N, const T* in, int64_t in_fs, cpx<T>* out, int64_t out_fs, int nf,
                        int64_t npencils, int kz, hipStream_t s) {
  H3_DISPATCH_N(N, {
    const unsigned nb = ceil_div(npencils, PencilBlock<NN>::PW);
    hipLaunchKernelGGL((k_z_r2c<NN, T>), dim3(nb, split_count(nb, (nf + 1) / 2, 2048)), dim3(PencilBlock<NN>::THREADS), 0, s, in, in_fs, out, out_fs,
                       nf, npencils, kz);
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_z_adj(int N, const cpx<T>* in, int64_t in_fs, int kz, int kzs, cpx<T>* out,
                        int64_t out_fs, int nzh, int64_t npencils, T scale, bool pruned,
                        hipStream_t s) {
  H3_DISPATCH_N(N, {
    const unsigned nb = ceil_div(npencils, PencilBlock<NN>::PW);
    if (pruned)
      hipLaunchKernelGGL((k_z_adj<NN, T, true>), dim3(nb), dim3(PencilBlock<NN>::THREADS), 0, s, in, in_fs, kz,
                         kzs, out, out_fs, nzh, npencils, scale);
    else
      hipLaunchKernelGGL((k_z_adj<NN, T, false>), dim3(nb), dim3(PencilBlock<NN>::THREADS), 0, s, in, in_fs,
                         kz, kzs, out, out_fs, nzh, npencils, scale);
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_x_forward_stage(int N, int mode, const StageArgs<T>& a, hipStream_t s,
                                  int* nblocks_out) {
  const int64_t J = int64_t(a.g.ky) * a.g.kzs;
  if constexpr (lean_type<T>()) {
    if (N == 1024 && pruned_rows(N, a.g.kx, a.g.px) && fits_u32(int64_t(a.g.nx) * a.g.jp, sizeof(cpx<T>)) &&
        fits_u32(int64_t(a.g.kx) * a.g.jp, sizeof(cpx<T>))) {
      constexpr int NN = 1024;
      const unsigned nb = ceil_div(J, LeanShape<NN>::TC);
      if (nblocks_out) *nblocks_out = int(nb);
      const dim3 blk(LeanShape<NN>::THREADS);
      switch (mode) {
        case kStageRhs: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageRhs, T, true>), dim3(nb), blk, 0, s, a); break;
        case kStage1: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage1, T, true>), dim3(nb), blk, 0, s, a); break;
        case kStage2: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage2, T, true>), dim3(nb), blk, 0, s, a); break;
        case kStage3: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage3, T, true>), dim3(nb), blk, 0, s, a); break;
        case kStage4: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage4, T, true>), dim3(nb), blk, 0, s, a); break;
        case kStageIf1: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf1, T, true>), dim3(nb), blk, 0, s, a); break;
        case kStageIf2: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf2, T, true>), dim3(nb), blk, 0, s, a); break;
        case kStageIf3: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf3, T, true>), dim3(nb), blk, 0, s, a); break;
        case kStageIf4: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf4, T, true>), dim3(nb), blk, 0, s, a); break;
        default: return hipErrorInvalidValue;
      }
      return hipGetLastError();
    }
  }
  const bool pout = pruned_rows(N, a.g.kx, a.g.px);
  H3_DISPATCH_N(N, {
    const unsigned nb = ceil_div(J, StridedShape<NN>::TC);
    if (nblocks_out) *nblocks_out = int(nb);
    const dim3 grid(nb);
    const dim3 blk(StridedShape<NN>::THREADS);
    if constexpr (kPrunedN<NN>) {
      if (pout) {
        switch (mode) {
          case kStageRhs: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageRhs, T, false, true>), grid, blk, 0, s, a); break;
          case kStage1: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage1, T, false, true>), grid, blk, 0, s, a); break;
          case kStage2: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage2, T, false, true>), grid, blk, 0, s, a); break;
          case kStage3: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage3, T, false, true>), grid, blk, 0, s, a); break;
          case kStage4: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage4, T, false, true>), grid, blk, 0, s, a); break;
          case kStageIf1: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf1, T, false, true>), grid, blk, 0, s, a); break;
          case kStageIf2: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf2, T, false, true>), grid, blk, 0, s, a); break;
          case kStageIf3: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf3, T, false, true>), grid, blk, 0, s, a); break;
          case kStageIf4: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf4, T, false, true>), grid, blk, 0, s, a); break;
          default: return hipErrorInvalidValue;
        }
        break;
      }
    }
    switch (mode) {
      case kStageRhs: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageRhs, T>), grid, blk, 0, s, a); break;
      case kStage1: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage1, T>), grid, blk, 0, s, a); break;
      case kStage2: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage2, T>), grid, blk, 0, s, a); break;
      case kStage3: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage3, T>), grid, blk, 0, s, a); break;
      case kStage4: hipLaunchKernelGGL((k_x_forward_stage<NN, kStage4, T>), grid, blk, 0, s, a); break;
      case kStageIf1: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf1, T>), grid, blk, 0, s, a); break;
      case kStageIf2: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf2, T>), grid, blk, 0, s, a); break;
      case kStageIf3: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf3, T>), grid, blk, 0, s, a); break;
      case kStageIf4: hipLaunchKernelGGL((k_x_forward_stage<NN, kStageIf4, T>), grid, blk, 0, s, a); break;
      default: return hipErrorInvalidValue;
    }
  });
  return hipGetLastError();
}

#define H3_INSTANTIATE(T)                                                                        \
  template hipError_t launch_strided<T>(int, int, const StridedArgs<T>&, hipStream_t);          \
  template hipError_t launch_x_inverse_curl<T>(int, const cpx<T>*, int64_t, cpx<T>*, int64_t,   \
                                               const Geo&, hipStream_t, const FinalizeJob*);     \
  template hipError_t launch_y_inverse_curl<T>(int, const cpx<T>*, int64_t, cpx<T>*, int64_t,   \
                                               const Geo&, int, hipStream_t);                         \
  template hipError_t launch_z_fused<T>(int, cpx<T>*, int64_t, int, int, int, int, int, T, hipStream_t); \
  template int pencil_row<T>(const Geo&); \
  template hipError_t launch_z_c2r<T>(int, const cpx<T>*, int64_t, T*, int64_t, int, int64_t,   \
                                      int, T, hipStream_t);                                      \
  template hipError_t launch_z_r2c<T>(int, const T*, int64_t, cpx<T>*, int64_t, int, int64_t,   \
                                      int, hipStream_t);                                         \
  template hipError_t launch_z_adj<T>(int, const cpx<T>*, int64_t, int, int, cpx<T>*, int64_t, int,   \
                                      int64_t, T, bool, hipStream_t);                              \
  template hipError_t launch_x_forward_stage<T>(int, int, const StageArgs<T>&, hipStream_t, int*);

H3_INSTANTIATE(H3_REAL)

}  // namespace h3
