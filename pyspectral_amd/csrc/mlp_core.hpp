// Shared pieces of the learned closure's network: the MFMA tile layout, parameter
// offsets, BF16 packing, tanh, input normalisation and the per-tile forward pass and
// closure decision, used by the network kernels and by the fused closure z pass.
//
// MFMA tiles: for D = A B + C with A 16x16 (M x K), B 16x16 (K x N), lane l (g = l/16,
// r = l%16) holds A[r][4g..4g+3], B[4g..4g+3][r] and D[4g..4g+3][r].  Points run along
// N and features along M/K, so a layer's D tiles (16 features x 16 points) are the next
// layer's B operands as they stand: the layers chain in registers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/psx.h"

namespace psx_mlp_impl {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NIN = 9, NH = 64, NOUT = 6;
// FP32 parameter block: W1[64][9] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[6][64] b4[6]
constexpr int OW1 = 0, OB1 = OW1 + NH * NIN, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NH, OW4 = OB3 + NH, OB4 = OW4 + NOUT * NH, NPARAM = OB4 + NOUT;
static_assert(NPARAM == PSX_MLP_NPARAM, "parameter count");
// packed BF16 operand tiles (64 lanes x 4 values each): forward A = W (M = out, K = in),
// backward A = W^T (M = in, K = out)
constexpr int T1 = 4 * 1, T2 = 4 * 4, T3 = 4 * 4, T4 = 1 * 4;  // forward tiles per layer
constexpr int U2 = 4 * 4, U3 = 4 * 4, U4 = 4 * 1;             // transposed tiles, layers 2..4
constexpr int NPACK = T1 + T2 + T3 + T4 + U2 + U3 + U4;       // tiles
// activations / deltas kept for the weight gradients, feature-major [feature][n]:
// a0 (16, inputs padded), a1 a2 a3 (64 each), d1 d2 d3 (64 each), d4 (16)
constexpr int FA0 = 0, FA1 = 16, FA2 = 80, FA3 = 144, FD1 = 208, FD2 = 272, FD3 = 336, FD4 = 400, NFEAT = 416;

__device__ __forceinline__ unsigned short f2bf(float x) {
  // round to nearest even
  unsigned int u = __float_as_uint(x);
  const unsigned int lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf2f(unsigned short h) {
#ifdef __HIP_DEVICE_COMPILE__
  return __uint_as_float(((unsigned int)h) << 16);
#else
  unsigned int u = ((unsigned int)h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
#endif
}
// four floats to BF16, round to nearest even, with the hardware's two-value
// conversion (v_cvt_pk_bf16_f32): element 0 of each pair in the low half
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef unsigned int u2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned int pack2(float a, float b) {
  f2v t;
  t.x = a;
  t.y = b;
  return __builtin_bit_cast(unsigned int, __builtin_convertvector(t, bf2v));
}
__device__ __forceinline__ bf4 pack4(f4 v) {
  u2v w;
  w.x = pack2(v[0], v[1]);
  w.y = pack2(v[2], v[3]);
  return __builtin_bit_cast(bf4, w);
}
__device__ __forceinline__ f4 mfma(bf4 a, bf4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
__device__ __forceinline__ float tanh_f(float x) {
  // tanh(x) = 2 / (1 + e) - 1, e = exp(-2x) = 2^(-2 log2(e) x), from the hardware
  // exp2 and reciprocal (about 1 ulp each): e -> inf gives -1, e -> 0 gives 1; the
  // absolute error (a few 1e-8) is far below the BF16 rounding the activations take
  const float e = __builtin_amdgcn_exp2f(x * -2.8853900817779268f);
  return fmaf(2.0f, __builtin_amdgcn_rcpf(1.0f + e), -1.0f);
}

struct Norm {
  // x_n = (x - mx) * isx with isx = 1/sx rounded to float, y = y_n sy + my
  float mx[NIN], isx[NIN], my[NOUT], sy[NOUT];
};

// Smagorinsky stress -2 c2 |S| S_ij from the gradients g[3i + j] = du_i/dx_j, with
// |S| = sqrt(2 S_ij S_ij); components o0 .. o0+3 of the order 11 22 33 12 13 23
__device__ __forceinline__ void strain(const float* g, float (&S)[6]) {
  S[0] = g[0];
  S[1] = g[4];
  S[2] = g[8];
  S[3] = 0.5f * (g[1] + g[3]);
  S[4] = 0.5f * (g[2] + g[6]);
  S[5] = 0.5f * (g[5] + g[7]);
}
__device__ __forceinline__ float strain_norm(const float (&S)[6]) {
  return sqrtf(2.f * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + 2.f * (S[3] * S[3] + S[4] * S[4] + S[5] * S[5])));
}
// component i (0..5, else 0) of a 6-vector, as selects (a run-time index into a
// per-thread array would place it in scratch memory)
__device__ __forceinline__ float sel6(int i, const float (&v)[6]) {
  return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : i == 3 ? v[3] : i == 4 ? v[4] : i == 5 ? v[5] : 0.f;
}
__device__ __forceinline__ f4 smag_stress(const float* g, float c2, int o0) {
  float S[6];
  strain(g, S);
  const float Smag = strain_norm(S);
  f4 t;
#pragma unroll
  for (int j = 0; j < 4; ++j) t[j] = -2.f * c2 * Smag * sel6(o0 + j, S);
  return t;
}

__device__ __forceinline__ bf4 tile_of(const short* __restrict__ pk, int tile, int lane) {
  const bf4* p = reinterpret_cast<const bf4*>(pk);
  return p[tile * 64 + lane];
}

// The forward weight tiles, biases and this lane's normalisation constants (features
// and outputs 4g .. 4g+3), held in registers by an inference loop.
struct Net {
  bf4 A[T1 + T2 + T3 + T4];
  f4 b1[4], b2[4], b3[4], b4;
  float mx[4], is[4], my[4], sy[4];
};

__device__ __forceinline__ void net_load(Net& R, const short* __restrict__ pk, const float* __restrict__ w,
                                         const Norm& nm, int lane) {
  const int g = lane >> 4;
#pragma unroll
  for (int t = 0; t < T1 + T2 + T3 + T4; ++t) R.A[t] = tile_of(pk, t, lane);
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      R.b1[t][j] = w[OB1 + 16 * t + 4 * g + j];
      R.b2[t][j] = w[OB2 + 16 * t + 4 * g + j];
      R.b3[t][j] = w[OB3 + 16 * t + 4 * g + j];
    }
#pragma unroll
  for (int j = 0; j < 4; ++j) R.b4[j] = (4 * g + j < NOUT) ? w[OB4 + 4 * g + j] : 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    R.mx[j] = R.is[j] = R.my[j] = R.sy[j] = 0.f;
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      const int f = 4 * gg + j;
      if (g == gg && f < NIN) {
        R.mx[j] = nm.mx[f];
        R.is[j] = nm.isx[f];
      }
      if (g == gg && f < NOUT) {
        R.my[j] = nm.my[f];
        R.sy[j] = nm.sy[f];
      }
    }
  }
}

// Normalised inputs of this lane (features 4g .. 4g+3, zero beyond the nine) for the
// point whose nine gradients start at gp; valid = false gives zeros.  Branch-free:
// every read stays inside the point's record, then a select.
__device__ __forceinline__ f4 net_input(const Net& R, const float* gp, bool valid, int g) {
  f4 xn;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int f = 4 * g + j, fc = f < NIN ? f : NIN - 1;
    const float v = (gp[fc] - R.mx[j]) * R.is[j];
    xn[j] = (valid && f < NIN) ? v : 0.f;
  }
  return xn;
}

// The four layers for TP tiles at once (independent tiles interleave: one tile's MFMAs
// overlap another's tanh work); out: the normalised outputs D[4g..4g+3][r].
template <int TP>
__device__ __forceinline__ void net_forward(const Net& R, const f4 (&xn)[TP], f4 (&out)[TP]) {
  bf4 B[TP][4];
#pragma unroll
  for (int q = 0; q < TP; ++q) B[q][0] = pack4(xn[q]);
  {
    f4 h[TP][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int q = 0; q < TP; ++q) h[q][t] = mfma(R.A[t], B[q][0], R.b1[t]);
#pragma unroll
    for (int q = 0; q < TP; ++q)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int j = 0; j < 4; ++j) h[q][t][j] = tanh_f(h[q][t][j]);
        B[q][t] = pack4(h[q][t]);
      }
  }
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const int base = layer == 0 ? T1 : T1 + T2;
    f4 h[TP][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int q = 0; q < TP; ++q) {
        f4 z = layer == 0 ? R.b2[t] : R.b3[t];
#pragma unroll
        for (int k = 0; k < 4; ++k) z = mfma(R.A[base + 4 * t + k], B[q][k], z);
        h[q][t] = z;
      }
#pragma unroll
    for (int q = 0; q < TP; ++q)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int j = 0; j < 4; ++j) h[q][t][j] = tanh_f(h[q][t][j]);
        B[q][t] = pack4(h[q][t]);
      }
  }
#pragma unroll
  for (int q = 0; q < TP; ++q) {
    f4 o = R.b4;
#pragma unroll
    for (int k = 0; k < 4; ++k) o = mfma(R.A[T1 + T2 + T3 + k], B[q][k], o);
    out[q] = o;
  }
}

// The closure's stress components 4g .. 4g+3 for one tile: the network's output, or
// the Smagorinsky stress -2 c2 |S| S_ij of the point's gradients gp where the largest
// |normalised input| exceeds bound or, with clip, where the network's stress is
// backscatter (tau_ij S_ij > 0, a negative subgrid dissipation).  Every lane of the
// wavefront takes part (cross-lane reductions over the four lane groups).
__device__ __forceinline__ f4 net_closure(const Net& R, f4 out, f4 xn, const float* gp, bool valid, int g, float c2,
                                          float bound, int clip) {
  float xm = 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j) xm = fmaxf(xm, fabsf(xn[j]));
  xm = fmaxf(xm, __shfl_xor(xm, 16));
  xm = fmaxf(xm, __shfl_xor(xm, 32));
  bool fallback = !(xm <= bound);
  f4 tv;
#pragma unroll
  for (int j = 0; j < 4; ++j) tv[j] = out[j] * R.sy[j] + R.my[j];
  if (clip) {
    float S[6];
    strain(gp, S);
    float part = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int o = 4 * g + j;
      part += (o < 3 ? 1.f : 2.f) * tv[j] * sel6(o, S);
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    fallback = fallback || part > 0.f;
  }
  // vector values, not arrays: a select between two arrays can become a select of
  // addresses into scratch memory
  return (fallback && valid) ? smag_stress(gp, c2, 4 * g) : tv;
}

}  // namespace psx_mlp_impl
