// Learned subgrid closure for the coarse LES (gfx950): periodic 3x3x3
// convolutions as implicit GEMMs on BF16 MFMA with FP32 accumulation, their
// backward passes, the optimiser step, and the spectral/physical kernels of
// the filtered-DNS pipeline (spectral truncation, velocity products, the
// divergence of a stress tensor). The reference has no counterpart: the
// closure is an addition to hit3d's solver.
//
// Layouts. A field of V = D*H*W voxels (v = (a*H + b)*W + c, c fastest: the
// physical layout real[nx][ny][nz] of the solver) with C channels is stored
// channels-last, X[v][C]. A layer's weights w[co][o][ci] (o = 9*(dz+1) +
// 3*(dy+1) + (dx+1) over the 27 neighbour offsets) are packed for the MFMA A
// operand as Wp[COUT][KP] with k = o*CIN + ci (zero beyond 27*CIN), and
// transposed for the backward-data GEMM as WT[CIN][KT] with k = o*COUT + co.
//
// v_mfma_f32_16x16x32_bf16 lane layout (lane l, element e): A[i][k] at
// l = i + 16*(k/8), e = k%8; B[k][j] at l = j + 16*(k/8), e = k%8;
// D[i][j] at l = j + 16*(i/4), e = i%4.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "closure.hpp"

namespace h3 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// neighbour of voxel (a, b, c) at offset o
__device__ __forceinline__ int neighbour(int a, int b, int c, int o, const ConvGeom& g) {
  const int da = o / 9 - 1, db = (o / 3) % 3 - 1, dc = o % 3 - 1;
  return (wrap(a + da, g.D) * g.H + wrap(b + db, g.H)) * g.W + wrap(c + dc, g.W);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

// ---------------------------------------------------------------------------
// Forward: out[v][co] = bias[co] + sum_{o,ci} w[co][o][ci] X[nbr(v,o)][ci].
// One wave computes COUT output channels for 64 consecutive voxels (4 column
// blocks of 16, same row since W % 16 == 0).
// CIN: stored input channels (16: two offsets per K-step, 32: one).
// mode 0: pre (fp32 [V][COUT]) and act = GELU(pre) (bf16 [V][COUT]);
// mode 1: out (fp32 [V][cout_real]) only.
// ---------------------------------------------------------------------------
template <int CIN, int COUT>
__global__ __launch_bounds__(64) void k_conv_fwd(const __bf16* __restrict__ X,
                                                 const __bf16* __restrict__ Wp, int KP,
                                                 const float* __restrict__ bias, int cout_real,
                                                 ConvGeom g, int mode, float* __restrict__ pre,
                                                 __bf16* __restrict__ act, float* __restrict__ out) {
  constexpr int RB = COUT / 16;
  static_assert(CIN == 16 || CIN == 32, "CIN must be 16 or 32");
  const int l = threadIdx.x, li = l % 16, lg = l / 16;
  const int v0 = blockIdx.x * 64;
  int va[4], vb[4], vc[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int v = v0 + 16 * b + li;
    vc[b] = v % g.W;
    vb[b] = (v / g.W) % g.H;
    va[b] = v / (g.W * g.H);
  }
  f32x4 acc[RB][4];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[r][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  // K-steps are a compile-time count (KP == conv_kp(CIN)); the loop is
  // unrolled and double-buffered: step s + 1's operands are loaded before
  // step s's MFMAs issue
  constexpr int KS = (27 * CIN + 31) / 32;
  (void)KP;
  auto load = [&](int s, bf16x8* A, bf16x8* B) {
    const int k = 32 * s + 8 * lg;          // this lane's first k
    const int o = k / CIN, ci = k % CIN;    // offset and first channel
#pragma unroll
    for (int r = 0; r < RB; ++r)
      A[r] = *reinterpret_cast<const bf16x8*>(Wp + int64_t(16 * r + li) * (32 * KS) + k);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (o < 27) {
        const int n = neighbour(va[b], vb[b], vc[b], o, g);
        B[b] = *reinterpret_cast<const bf16x8*>(X + int64_t(n) * CIN + ci);
      } else {
        B[b] = bf16x8{};  // K padding (the packed weights are zero there too)
      }
    }
  };
  bf16x8 A0[RB], B0[4], A1[RB], B1[4];
  load(0, A0, B0);
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    bf16x8* A = (s & 1) ? A1 : A0;
    bf16x8* B = (s & 1) ? B1 : B0;
    if (s + 1 < KS) load(s + 1, (s & 1) ? A0 : A1, (s & 1) ? B0 : B1);
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[r][b] = mfma(A[r], B[b], acc[r][b]);
  }
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int v = v0 + 16 * b + li;
      const int co0 = 16 * r + 4 * lg;
      float y[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) y[e] = acc[r][b][e] + (co0 + e < cout_real ? bias[co0 + e] : 0.f);
      if (mode == 0) {
        if (pre != nullptr)   // kept for a backward pass only
          *reinterpret_cast<float4*>(pre + int64_t(v) * COUT + co0) = make_float4(y[0], y[1], y[2], y[3]);
        __bf16 h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) h[e] = (__bf16)(co0 + e < cout_real ? gelu(y[e]) : 0.f);
        *reinterpret_cast<uint2*>(act + int64_t(v) * COUT + co0) = *reinterpret_cast<uint2*>(h);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (co0 + e < cout_real) out[int64_t(v) * cout_real + co0 + e] = y[e];
      }
    }
}

// ---------------------------------------------------------------------------
// Forward on a tile staged in LDS: a workgroup of 8 waves computes a
// 4 x 8 x 16 block of voxels (wave w: plane a0 + w/2, rows b0 + 4(w%2) ..
// + 3, 16 voxels per row as in k_conv_fwd). Its periodic halo of 6 x 10 x 18
// voxels (all CIN channels) and the layer's packed weights are loaded into
// LDS once, so the K loop reads both MFMA operands from LDS: each input row
// is fetched from memory by one workgroup (plus the halo overlap) instead of
// by the waves of its 27 neighbours, and the weights once per workgroup
// instead of once per wave. Weight rows are padded by 16 bytes so that the
// 16 rows one A read touches fall in distinct banks, and so are the halo's
// voxel rows (CIN/8 + 1 chunks of 16 bytes): the 16 consecutive voxels of a
// B read then start 80 (48) bytes apart instead of 64 (32), which puts any 8
// of them in distinct banks. Needs D % 4 == 0,
// H % 8 == 0, W % 16 == 0. Same products and summation order as k_conv_fwd.
// ---------------------------------------------------------------------------
constexpr int kTileA = 4, kTileB = 8, kTileC = 16, kTileWaves = 8;
constexpr int kHaloA = kTileA + 2, kHaloB = kTileB + 2, kHaloC = kTileC + 2;

template <int CIN, int COUT>
__global__ __launch_bounds__(64 * kTileWaves) void k_conv_fwd_tiled(const __bf16* __restrict__ X,
                                                                    const __bf16* __restrict__ Wp,
                                                                    const float* __restrict__ bias,
                                                                    int cout_real, ConvGeom g, int mode,
                                                                    float* __restrict__ pre,
                                                                    __bf16* __restrict__ act,
                                                                    float* __restrict__ out) {
  constexpr int NT = 64 * kTileWaves;
  constexpr int RB = COUT / 16;
  constexpr int KS = (27 * CIN + 31) / 32;
  constexpr int CHUNKS = CIN / 8;                    // 16-byte chunks per voxel row
  constexpr int CSTR = CHUNKS + 1;                   // padded voxel row in LDS
  constexpr int HALO = kHaloA * kHaloB * kHaloC;
  constexpr int WROW = 4 * KS;                       // 16-byte chunks per weight row
  constexpr int WROWP = WROW + 1;                    // padded row
  constexpr int NHALO = HALO * CHUNKS, NW = COUT * WROW;
  constexpr int NLOAD = (NHALO + NW + NT - 1) / NT;
  static_assert(CIN == 16 || CIN == 32, "CIN must be 16 or 32");
  __shared__ bf16x8 tile[HALO * CSTR];
  __shared__ bf16x8 wts[COUT * WROWP];
  const int tid = threadIdx.x, wave = tid / 64, l = tid % 64, li = l % 16, lg = l / 16;
  const int nc = g.W / kTileC, nbk = g.H / kTileB;
  const int c0 = (blockIdx.x % nc) * kTileC;
  const int b0 = ((blockIdx.x / nc) % nbk) * kTileB;
  const int a0 = (blockIdx.x / (nc * nbk)) * kTileA;
  const int wa = wave / 2, wb = 4 * (wave % 2);      // this wave's plane and first row in the tile
  // stage halo and weights: all loads of a thread issue before its LDS stores
  bf16x8 stage[NLOAD];
#pragma unroll
  for (int j = 0; j < NLOAD; ++j) {
    const int i = tid + NT * j;
    if (i < NHALO) {
      const int t = i / CHUNKS, ch = i % CHUNKS;
      const int tc = t % kHaloC, tb = (t / kHaloC) % kHaloB, ta = t / (kHaloC * kHaloB);
      const int n = (wrap(a0 - 1 + ta, g.D) * g.H + wrap(b0 - 1 + tb, g.H)) * g.W + wrap(c0 - 1 + tc, g.W);
      stage[j] = *reinterpret_cast<const bf16x8*>(X + int64_t(n) * CIN + 8 * ch);
    } else if (i < NHALO + NW) {
      stage[j] = reinterpret_cast<const bf16x8*>(Wp)[i - NHALO];
    }
  }
#pragma unroll
  for (int j = 0; j < NLOAD; ++j) {
    const int i = tid + NT * j;
    if (i < NHALO) {
      tile[(i / CHUNKS) * CSTR + i % CHUNKS] = stage[j];
    } else if (i < NHALO + NW) {
      const int w = i - NHALO;
      wts[(w / WROW) * WROWP + w % WROW] = stage[j];
    }
  }
  __syncthreads();

  f32x4 acc[RB][4];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[r][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  auto load = [&](int s, bf16x8* A, bf16x8* B) {
    const int k = 32 * s + 8 * lg;
    const int o = k / CIN, ci = k % CIN;
#pragma unroll
    for (int r = 0; r < RB; ++r) A[r] = wts[(16 * r + li) * WROWP + 4 * s + lg];
    const int da = o / 9, db = (o / 3) % 3, dc = o % 3;   // offsets + 1
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (o < 27) {
        const int t = ((wa + da) * kHaloB + (wb + b + db)) * kHaloC + (li + dc);
        B[b] = tile[t * CSTR + ci / 8];
      } else {
        B[b] = bf16x8{};
      }
    }
  };
  bf16x8 A0[RB], B0[4], A1[RB], B1[4];
  load(0, A0, B0);
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    bf16x8* A = (s & 1) ? A1 : A0;
    bf16x8* B = (s & 1) ? B1 : B0;
    if (s + 1 < KS) load(s + 1, (s & 1) ? A0 : A1, (s & 1) ? B0 : B1);
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[r][b] = mfma(A[r], B[b], acc[r][b]);
  }
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int v = ((a0 + wa) * g.H + (b0 + wb + b)) * g.W + c0 + li;
      const int co0 = 16 * r + 4 * lg;
      float y[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) y[e] = acc[r][b][e] + (co0 + e < cout_real ? bias[co0 + e] : 0.f);
      if (mode == 0) {
        if (pre != nullptr)   // kept for a backward pass only
          *reinterpret_cast<float4*>(pre + int64_t(v) * COUT + co0) = make_float4(y[0], y[1], y[2], y[3]);
        __bf16 h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) h[e] = (__bf16)(co0 + e < cout_real ? gelu(y[e]) : 0.f);
        *reinterpret_cast<uint2*>(act + int64_t(v) * COUT + co0) = *reinterpret_cast<uint2*>(h);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (co0 + e < cout_real) out[int64_t(v) * cout_real + co0 + e] = y[e];
      }
    }
}

// ---------------------------------------------------------------------------
// Backward data: dX[u][ci] = sum_{o,co} w[co][o][ci] dP[u - o][co]
// (D[ci][u] = WT[ci][k] * dP[u - o(k)][co(k)], k = o*COUT + co).
// dP bf16 [V][COUT]; dX fp32 [V][CIN].
// ---------------------------------------------------------------------------
template <int COUT, int CIN>
__global__ __launch_bounds__(64) void k_conv_bwd_data(const __bf16* __restrict__ dP,
                                                      const __bf16* __restrict__ WT, int KT,
                                                      ConvGeom g, float* __restrict__ dX) {
  constexpr int RB = CIN / 16;
  static_assert(COUT == 16 || COUT == 32, "COUT must be 16 or 32");
  const int l = threadIdx.x, li = l % 16, lg = l / 16;
  const int u0 = blockIdx.x * 64;
  int ua[4], ub[4], uc[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int u = u0 + 16 * b + li;
    uc[b] = u % g.W;
    ub[b] = (u / g.W) % g.H;
    ua[b] = u / (g.W * g.H);
  }
  f32x4 acc[RB][4];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[r][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  constexpr int KS = (27 * COUT + 31) / 32;  // KT == conv_kt(COUT)
  (void)KT;
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int k = 32 * s + 8 * lg;
    const int o = k / COUT, co = k % COUT;
    bf16x8 A[RB], B[4];
#pragma unroll
    for (int r = 0; r < RB; ++r)
      A[r] = *reinterpret_cast<const bf16x8*>(WT + int64_t(16 * r + li) * (32 * KS) + k);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (o < 27) {
        const int n = neighbour(ua[b], ub[b], uc[b], 26 - o, g);  // u - o: the mirrored offset
        B[b] = *reinterpret_cast<const bf16x8*>(dP + int64_t(n) * COUT + co);
      } else {
        B[b] = bf16x8{};
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[r][b] = mfma(A[r], B[b], acc[r][b]);
  }
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int u = u0 + 16 * b + li;
      const int ci0 = 16 * r + 4 * lg;
      *reinterpret_cast<float4*>(dX + int64_t(u) * CIN + ci0) =
          make_float4(acc[r][b][0], acc[r][b][1], acc[r][b][2], acc[r][b][3]);
    }
}

// ---------------------------------------------------------------------------
// Backward weights, per voxel chunk and group of 9 offsets:
//   part[chunk][o][co][ci] = sum_{u in chunk} dP[u][co] X[u + o][ci]
// (D[co][ci] = dP^T[co][u] * Xo[u][ci], K = the chunk's voxels), and for the
// first group the bias partials pb[chunk][co] = sum_u dP[u][co].
// ---------------------------------------------------------------------------
template <int COUT, int CIN>
__global__ __launch_bounds__(64) void k_conv_bwd_weight(const __bf16* __restrict__ dP,
                                                        const __bf16* __restrict__ X, ConvGeom g,
                                                        int chunk, float* __restrict__ part,
                                                        float* __restrict__ pb) {
  constexpr int RB = COUT / 16, CB = CIN / 16, OG = 9;
  const int l = threadIdx.x, li = l % 16, lg = l / 16;
  const int c = blockIdx.x, og = blockIdx.y;
  f32x4 acc[OG][RB][CB];
#pragma unroll
  for (int o = 0; o < OG; ++o)
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int q = 0; q < CB; ++q) acc[o][r][q] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum[RB];
#pragma unroll
  for (int r = 0; r < RB; ++r) bsum[r] = 0.f;
  const int ubeg = c * chunk;
  for (int u0 = ubeg; u0 < ubeg + chunk; u0 += 32) {
    // A operand: dP[u][co] for u = u0 + 8*lg + e, co = 16 r + li
    bf16x8 A[RB];
    int ua[8], ub[8], uc[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int u = u0 + 8 * lg + e;
      uc[e] = u % g.W;
      ub[e] = (u / g.W) % g.H;
      ua[e] = u / (g.W * g.H);
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const __bf16 x = dP[int64_t(u) * COUT + 16 * r + li];
        A[r][e] = x;
        bsum[r] += float(x);
      }
    }
#pragma unroll
    for (int o = 0; o < OG; ++o) {
      bf16x8 B[CB];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int n = neighbour(ua[e], ub[e], uc[e], OG * og + o, g);
#pragma unroll
        for (int q = 0; q < CB; ++q) B[q][e] = X[int64_t(n) * CIN + 16 * q + li];
      }
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int q = 0; q < CB; ++q) acc[o][r][q] = mfma(A[r], B[q], acc[o][r][q]);
    }
  }
  // D[co][ci]: lane holds ci = 16 q + li, co = 16 r + 4 lg + e
#pragma unroll
  for (int o = 0; o < OG; ++o)
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int q = 0; q < CB; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int co = 16 * r + 4 * lg + e, ci = 16 * q + li;
          part[((int64_t(c) * 27 + OG * og + o) * COUT + co) * CIN + ci] = acc[o][r][q][e];
        }
  if (og == 0) {
    // lanes li, li+16, li+32, li+48 hold disjoint voxel sums of channel 16 r + li
#pragma unroll
    for (int r = 0; r < RB; ++r) pb[(int64_t(c) * 4 + lg) * COUT + 16 * r + li] = bsum[r];
  }
}

// dW[co][o][ci] (real sizes) and db[co] from the chunk partials, summed in
// chunk order (deterministic).
__global__ void k_conv_reduce(const float* __restrict__ part, const float* __restrict__ pb,
                              int nchunk, int COUT, int CIN, int cout_real, int cin_real,
                              float* __restrict__ dW, float* __restrict__ db) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nw = cout_real * 27 * cin_real;
  if (i < nw) {
    const int co = i / (27 * cin_real), o = (i / cin_real) % 27, ci = i % cin_real;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[((int64_t(c) * 27 + o) * COUT + co) * CIN + ci];
    dW[i] = s;
  } else if (i < nw + cout_real) {
    const int co = i - nw;
    float s = 0.f;
    for (int c = 0; c < nchunk * 4; ++c) s += pb[int64_t(c) * COUT + co];
    db[co] = s;
  }
}

// dP = dA * GELU'(pre) in bf16 (channels beyond cout_real are zero).
__global__ void k_gelu_bwd(const float* __restrict__ dA, const float* __restrict__ pre,
                           int64_t n, int C, int creal, __bf16* __restrict__ dP) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int ch = int(i % C);
    dP[i] = (__bf16)(ch < creal ? dA[i] * gelu_grad(pre[i]) : 0.f);
  }
}

// fp32 [V][C] -> bf16 [V][C] (the last layer's output gradient).
__global__ void k_to_bf16_padded(const float* __restrict__ src, int64_t nvox, int creal, int C,
                                 __bf16* __restrict__ dst) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nvox * C;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t v = i / C;
    const int ch = int(i % C);
    dst[i] = (__bf16)(ch < creal ? src[v * creal + ch] : 0.f);
  }
}

// Packs w[co][o][ci] (fp32, real sizes) into Wp[COUT][KP] and WT[CIN][KT].
__global__ void k_pack_weights(const float* __restrict__ w, int cout_real, int cin_real, int COUT,
                               int CIN, int KP, int KT, __bf16* __restrict__ Wp,
                               __bf16* __restrict__ WT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int np = COUT * KP, nt = CIN * KT;
  if (i < np) {
    const int co = i / KP, k = i % KP, o = k / CIN, ci = k % CIN;
    const float x = (co < cout_real && o < 27 && ci < cin_real)
                        ? w[(co * 27 + o) * cin_real + ci] : 0.f;
    Wp[i] = (__bf16)x;
  } else if (i < np + nt) {
    const int j = i - np, ci = j / KT, k = j % KT, o = k / COUT, co = k % COUT;
    const float x = (co < cout_real && o < 27 && ci < cin_real)
                        ? w[(co * 27 + o) * cin_real + ci] : 0.f;
    WT[j] = (__bf16)x;
  }
}

// Adam (Kingma & Ba 2015) with bias correction; c1 = 1 - beta1^t, c2 = 1 - beta2^t.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float gi = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// The pack (dir 0) runs one voxel per thread with 16-byte stores when the
// channel groups of 8 start on 16-byte boundaries.
__host__ __device__ inline bool pack_by_voxel(int dir, int CP, const __bf16* aos_bf) {
  return dir == 0 && (CP & 7) == 0 && (reinterpret_cast<uintptr_t>(aos_bf) & 15) == 0;
}

// SoA fp32 [C][V] (scaled by s[c]) <-> channels-last [V][CP]: dir 0 writes
// bf16 channels-last (zero channels >= C), dir 1 reads fp32 channels-last
// (ld = its channel stride) into SoA.

__global__ void k_soa_aos(const float* __restrict__ soa, int C, int64_t V, const float* __restrict__ s,
                          int CP, int dir, __bf16* __restrict__ aos_bf, const float* __restrict__ aos_f,
                          float* __restrict__ soa_out) {
  if (pack_by_voxel(dir, CP, aos_bf)) {
    // one voxel per thread: channel reads coalesce across the voxels, and
    // the voxel's CP bf16 values leave as 16-byte stores
    for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
         v += int64_t(gridDim.x) * blockDim.x) {
      for (int c8 = 0; c8 < CP; c8 += 8) {
        union {
          __bf16 h[8];
          uint4 u;
        } pk;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int ch = c8 + j;
          pk.h[j] = (__bf16)(ch < C ? soa[ch * V + v] * s[ch] : 0.f);
        }
        *reinterpret_cast<uint4*>(aos_bf + v * CP + c8) = pk.u;
      }
    }
    return;
  }
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < V * CP;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t v = i / CP;
    const int ch = int(i % CP);
    if (dir == 0) {
      aos_bf[i] = (__bf16)(ch < C ? soa[ch * V + v] * s[ch] : 0.f);
    } else if (ch < C) {
      soa_out[ch * V + v] = aos_f[i] * s[ch];
    }
  }
}

// tau[c][v] = aos[v][c] * s[c] (6 stress channels), set to 0 where the
// modelled transfer Pi = -tau_ij S_ij is negative, S_ij = (d_j u_i + d_i u_j)/2
// from the gradient channels g[3 j + i][v]; keep[v] = (Pi >= 0).
__global__ void k_unpack_clip(const float* __restrict__ aos, int64_t V, const float* __restrict__ s,
                              const float* __restrict__ g, float* __restrict__ tau,
                              uint8_t* __restrict__ keep) {
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
       v += int64_t(gridDim.x) * blockDim.x) {
    float t[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) t[c] = aos[v * 6 + c] * s[c];
    float d[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) d[c] = g[c * V + v];
    const float s00 = d[0], s11 = d[4], s22 = d[8];
    const float s01 = 0.5f * (d[3] + d[1]), s02 = 0.5f * (d[6] + d[2]), s12 = 0.5f * (d[7] + d[5]);
    const float pi = -(t[0] * s00 + 2.f * t[1] * s01 + 2.f * t[2] * s02 + t[3] * s11 + 2.f * t[4] * s12 +
                       t[5] * s22);
    const bool k = pi >= 0.f;
#pragma unroll
    for (int c = 0; c < 6; ++c) tau[c * V + v] = k ? t[c] : 0.f;
    keep[v] = k ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// Spectral pieces of the filtered-DNS pipeline and the closure force.
// ---------------------------------------------------------------------------

// coarse[c][ix][iy][iz] = scale * fine[c][...] for integer frequencies with
// n2 = nx^2 + ny^2 + nz^2 <= kc2 (sharp spherical filter), 0 elsewhere.
template <class T>
__global__ void k_truncate(const cpx<T>* __restrict__ fine, int fx, int fy, int fz,
                           cpx<T>* __restrict__ coarse, int cx, int cy, int cz, int ncomp,
                           double kc2, T scale) {
  const int czh = cz / 2 + 1, fzh = fz / 2 + 1;
  const int64_t n = int64_t(cx) * cy * czh;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n * ncomp;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int comp = int(i / n);
    const int64_t r = i % n;
    const int iz = int(r % czh), iy = int((r / czh) % cy), ix = int(r / (int64_t(czh) * cy));
    const int nx = ix < cx / 2 ? ix : ix - cx, ny = iy < cy / 2 ? iy : iy - cy, nz = iz;
    const bool inside = 2 * abs(nx) < cx && 2 * abs(ny) < cy && 2 * nz < cz &&
                        double(nx) * nx + double(ny) * ny + double(nz) * nz <= kc2;
    cpx<T> val{T(0), T(0)};
    if (inside) {
      const int jx = nx >= 0 ? nx : nx + fx, jy = ny >= 0 ? ny : ny + fy;
      const cpx<T> f = fine[int64_t(comp) * fx * fy * fzh + (int64_t(jx) * fy + jy) * fzh + nz];
      val = {f.x * scale, f.y * scale};
    }
    coarse[i] = val;
  }
}

// uu[6][V] = (u0u0, u0u1, u0u2, u1u1, u1u2, u2u2) from u[3][V].
template <class T>
__global__ void k_products(const T* __restrict__ u, int64_t V, T* __restrict__ uu) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < V;
       i += int64_t(gridDim.x) * blockDim.x) {
    const T a = u[i], b = u[V + i], c = u[2 * V + i];
    uu[i] = a * a;
    uu[V + i] = a * b;
    uu[2 * V + i] = a * c;
    uu[3 * V + i] = b * b;
    uu[4 * V + i] = b * c;
    uu[5 * V + i] = c * c;
  }
}

// Stress index of (i, j) in the (00, 01, 02, 11, 12, 22) order.
__device__ __forceinline__ int sidx(int i, int j) {
  const int a = i < j ? i : j, b = i < j ? j : i;
  return a == 0 ? b : (a == 1 ? 2 + b : 5);
}

// dir 0: F_i = -sum_j i k_j tau_ij (6 -> 3 fields);
// dir 1 (adjoint in the convention dL = Re sum conj(g) dx):
//   gtau_ij = i k_j gF_i + i k_i gF_j (i != j), i k_i gF_i (i == j).
template <class T>
__global__ void k_stress_div(const cpx<T>* __restrict__ in, cpx<T>* __restrict__ out, Geo g, int dir) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  for (int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; m < n;
       m += int64_t(gridDim.x) * blockDim.x) {
    const int iz = int(m % g.nzh), iy = int((m / g.nzh) % g.ny), ix = int(m / (int64_t(g.nzh) * g.ny));
    // integer frequencies as numpy.fft.fftfreq (index N/2 -> -N/2), the grid's convention
    const T k[3] = {wavenumber<T>(g.k0x, ix < g.nx / 2 ? ix : ix - g.nx),
                    wavenumber<T>(g.k0y, iy < g.ny / 2 ? iy : iy - g.ny),
                    wavenumber<T>(g.k0z, iz)};
    if (dir == 0) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        cpx<T> s{T(0), T(0)};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const cpx<T> t = in[sidx(i, j) * n + m];
          s = {s.x + k[j] * t.x, s.y + k[j] * t.y};
        }
        out[i * n + m] = {s.y, -s.x};  // -i * s
      }
    } else {
      cpx<T> gf[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) gf[i] = in[i * n + m];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
          cpx<T> s{k[j] * gf[i].x, k[j] * gf[i].y};
          if (j != i) s = {s.x + k[i] * gf[j].x, s.y + k[i] * gf[j].y};
          out[sidx(i, j) * n + m] = {-s.y, s.x};  // i * s
        }
    }
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline unsigned grid_for(int64_t n, int threads = 256) {
  int64_t b = (n + threads - 1) / threads;
  return unsigned(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

hipError_t launch_conv_fwd(int cin, int cout, const __bf16* X, const __bf16* Wp, int KP,
                           const float* bias, int cout_real, const ConvGeom& g, int mode, float* pre,
                           __bf16* act, float* out, hipStream_t s) {
  const unsigned nb = unsigned(int64_t(g.D) * g.H * g.W / 64);
  const bool tiled = g.D % kTileA == 0 && g.H % kTileB == 0 && g.W % kTileC == 0;
#define H3_CONV_FWD(CI, CO)                                                                                \
  do {                                                                                                     \
    if (tiled)                                                                                             \
      hipLaunchKernelGGL((k_conv_fwd_tiled<CI, CO>), dim3(nb / (kTileA * kTileB * kTileC / 64)),           \
                         dim3(64 * kTileWaves), 0, s, X, Wp, bias, cout_real,                              \
                         g, mode, pre, act, out);                                                          \
    else                                                                                                   \
      hipLaunchKernelGGL((k_conv_fwd<CI, CO>), dim3(nb), dim3(64), 0, s, X, Wp, KP, bias, cout_real, g,     \
                         mode, pre, act, out);                                                             \
  } while (0)
  if (cin == 16 && cout == 32) H3_CONV_FWD(16, 32);
  else if (cin == 32 && cout == 32) H3_CONV_FWD(32, 32);
  else if (cin == 32 && cout == 16) H3_CONV_FWD(32, 16);
  else if (cin == 16 && cout == 16) H3_CONV_FWD(16, 16);
  else return hipErrorInvalidValue;
#undef H3_CONV_FWD
  return hipGetLastError();
}

hipError_t launch_conv_bwd_data(int cout, int cin, const __bf16* dP, const __bf16* WT, int KT,
                                const ConvGeom& g, float* dX, hipStream_t s) {
  const unsigned nb = unsigned(int64_t(g.D) * g.H * g.W / 64);
#define H3_CONV_BD(CO, CI) hipLaunchKernelGGL((k_conv_bwd_data<CO, CI>), dim3(nb), dim3(64), 0, s, dP, WT, KT, g, dX)
  if (cout == 32 && cin == 16) H3_CONV_BD(32, 16);
  else if (cout == 32 && cin == 32) H3_CONV_BD(32, 32);
  else if (cout == 16 && cin == 32) H3_CONV_BD(16, 32);
  else if (cout == 16 && cin == 16) H3_CONV_BD(16, 16);
  else return hipErrorInvalidValue;
#undef H3_CONV_BD
  return hipGetLastError();
}

hipError_t launch_conv_bwd_weight(int cout, int cin, const __bf16* dP, const __bf16* X,
                                  const ConvGeom& g, int chunk, float* part, float* pb,
                                  hipStream_t s) {
  const int64_t V = int64_t(g.D) * g.H * g.W;
  const dim3 grid(unsigned(V / chunk), 3);
#define H3_CONV_BW(CO, CI) hipLaunchKernelGGL((k_conv_bwd_weight<CO, CI>), grid, dim3(64), 0, s, dP, X, g, chunk, part, pb)
  if (cout == 32 && cin == 16) H3_CONV_BW(32, 16);
  else if (cout == 32 && cin == 32) H3_CONV_BW(32, 32);
  else if (cout == 16 && cin == 32) H3_CONV_BW(16, 32);
  else if (cout == 16 && cin == 16) H3_CONV_BW(16, 16);
  else return hipErrorInvalidValue;
#undef H3_CONV_BW
  return hipGetLastError();
}

hipError_t launch_conv_reduce(const float* part, const float* pb, int nchunk, int cout, int cin,
                              int cout_real, int cin_real, float* dW, float* db, hipStream_t s) {
  const int n = cout_real * 27 * cin_real + cout_real;
  hipLaunchKernelGGL(k_conv_reduce, dim3((n + 255) / 256), dim3(256), 0, s, part, pb, nchunk, cout, cin,
                     cout_real, cin_real, dW, db);
  return hipGetLastError();
}

hipError_t launch_gelu_bwd(const float* dA, const float* pre, int64_t n, int C, int creal, __bf16* dP,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_gelu_bwd, dim3(grid_for(n)), dim3(256), 0, s, dA, pre, n, C, creal, dP);
  return hipGetLastError();
}

hipError_t launch_to_bf16_padded(const float* src, int64_t nvox, int creal, int C, __bf16* dst,
                                 hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16_padded, dim3(grid_for(nvox * C)), dim3(256), 0, s, src, nvox, creal, C, dst);
  return hipGetLastError();
}

hipError_t launch_pack_weights(const float* w, int cout_real, int cin_real, int cout, int cin, int KP,
                               int KT, __bf16* Wp, __bf16* WT, hipStream_t s) {
  const int n = cout * KP + cin * KT;
  hipLaunchKernelGGL(k_pack_weights, dim3((n + 255) / 256), dim3(256), 0, s, w, cout_real, cin_real, cout,
                     cin, KP, KT, Wp, WT);
  return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* gr, float* m, float* v, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0, s, p, gr, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t launch_soa_aos(const float* soa, int C, int64_t V, const float* scale, int CP, int dir,
                          __bf16* aos_bf, const float* aos_f, float* soa_out, hipStream_t s) {
  const int64_t n = pack_by_voxel(dir, CP, aos_bf) ? V : V * CP;   // threads the kernel's loop takes
  hipLaunchKernelGGL(k_soa_aos, dim3(grid_for(n)), dim3(256), 0, s, soa, C, V, scale, CP, dir, aos_bf,
                     aos_f, soa_out);
  return hipGetLastError();
}

hipError_t launch_unpack_clip(const float* aos, int64_t V, const float* scale, const float* g, float* tau,
                              uint8_t* keep, hipStream_t s) {
  hipLaunchKernelGGL(k_unpack_clip, dim3(grid_for(V)), dim3(256), 0, s, aos, V, scale, g, tau, keep);
  return hipGetLastError();
}

template <class T>
hipError_t launch_truncate(const cpx<T>* fine, int fx, int fy, int fz, cpx<T>* coarse, int cx, int cy,
                           int cz, int ncomp, double kc2, T scale, hipStream_t s) {
  const int64_t n = int64_t(cx) * cy * (cz / 2 + 1) * ncomp;
  hipLaunchKernelGGL((k_truncate<T>), dim3(grid_for(n)), dim3(256), 0, s, fine, fx, fy, fz, coarse, cx, cy,
                     cz, ncomp, kc2, scale);
  return hipGetLastError();
}

template <class T>
hipError_t launch_products(const T* u, int64_t V, T* uu, hipStream_t s) {
  hipLaunchKernelGGL((k_products<T>), dim3(grid_for(V)), dim3(256), 0, s, u, V, uu);
  return hipGetLastError();
}

// out = [k2 <= kcut2] (du + gain S(k) P M F), F_i = -i sum_j k_j tau_ij: the
// closure force assembled onto the resolved RHS in one pass over the modes
// (the stress divergence of k_stress_div, the 2/3 mask and Leray projection
// of the nonlinear term, the force's gain and spectral shape S, the sum and
// the LES's spherical mask). `shape` (real [nx][ny][nz/2+1]) may be null
// (S = 1); `gain` is a device scalar, so captured steps follow its updates;
// a null `tau` leaves only the mask (no closure force).
template <class T>
__global__ void k_les_combine(const cpx<T>* du, const cpx<T>* __restrict__ tau,
                              const T* __restrict__ shape, const float* __restrict__ gain, T kcut2,
                              cpx<T>* out, Geo g) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  const T gn = tau ? T(*gain) : T(0);
  for (int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; m < n;
       m += int64_t(gridDim.x) * blockDim.x) {
    const int iz = int(m % g.nzh), iy = int((m / g.nzh) % g.ny), ix = int(m / (int64_t(g.nzh) * g.ny));
    const int nx = ix < g.nx / 2 ? ix : ix - g.nx, ny = iy < g.ny / 2 ? iy : iy - g.ny;
    const T k[3] = {wavenumber<T>(g.k0x, nx), wavenumber<T>(g.k0y, ny), wavenumber<T>(g.k0z, iz)};
    const T k2 = k2_of(k[0], k[1], k[2]);
    if (!(k2 <= kcut2)) {
      for (int i = 0; i < 3; ++i) out[i * n + m] = czero<T>();
      continue;
    }
    cpx<T> f[3] = {czero<T>(), czero<T>(), czero<T>()};
    if (tau && in_box(nx, ny, iz, g)) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        cpx<T> s{T(0), T(0)};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const cpx<T> t = tau[sidx(i, j) * n + m];
          s = {s.x + k[j] * t.x, s.y + k[j] * t.y};
        }
        f[i] = {s.y, -s.x};  // -i s
      }
      const cpx<T> div = inv_k2_of(k2) * (k[0] * f[0] + k[1] * f[1] + k[2] * f[2]);
      const T c = gn * (shape ? shape[m] : T(1));
#pragma unroll
      for (int i = 0; i < 3; ++i) f[i] = c * (f[i] - k[i] * div);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i * n + m] = du[i * n + m] + f[i];
  }
}

template <class T>
hipError_t launch_les_combine(const cpx<T>* du, const cpx<T>* tau, const T* shape, const float* gain,
                              T kcut2, cpx<T>* out, const Geo& g, hipStream_t s) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  hipLaunchKernelGGL((k_les_combine<T>), dim3(grid_for(n)), dim3(256), 0, s, du, tau, shape, gain, kcut2,
                     out, g);
  return hipGetLastError();
}

template <class T>
hipError_t launch_stress_div(const cpx<T>* in, cpx<T>* out, const Geo& g, int dir, hipStream_t s) {
  const int64_t n = int64_t(g.nx) * g.ny * g.nzh;
  hipLaunchKernelGGL((k_stress_div<T>), dim3(grid_for(n)), dim3(256), 0, s, in, out, g, dir);
  return hipGetLastError();
}

#define H3_CLOSURE_INST(T)                                                                          \
  template hipError_t launch_truncate<T>(const cpx<T>*, int, int, int, cpx<T>*, int, int, int, int,  \
                                         double, T, hipStream_t);                                    \
  template hipError_t launch_products<T>(const T*, int64_t, T*, hipStream_t);                        \
  template hipError_t launch_stress_div<T>(const cpx<T>*, cpx<T>*, const Geo&, int, hipStream_t);    \
  template hipError_t launch_les_combine<T>(const cpx<T>*, const cpx<T>*, const T*, const float*, T, \
                                            cpx<T>*, const Geo&, hipStream_t);
H3_CLOSURE_INST(float)
H3_CLOSURE_INST(double)

}  // namespace h3
