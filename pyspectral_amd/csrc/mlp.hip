// Learned subgrid closure: a pointwise MLP 9 -> 64 -> 64 -> 64 -> 6 with tanh hidden
// layers, BF16 operands and FP32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_16x16x16_bf16), FP32 master weights and Adam.  No counterpart in the
// reference.
// The tile layout and the per-tile forward pass are in mlp_core.hpp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/psx.h"
#include "mlp_core.hpp"

namespace psx_mlp_impl {

// Pack the FP32 parameters into MFMA A-operand tiles (one thread per tile value).
__global__ void pack_weights(const float* __restrict__ w, short* __restrict__ pk) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // over NPACK * 64 * 4 values
  if (e >= NPACK * 256) return;
  const int tile = e / 256, lane = (e / 4) % 64, j = e % 4;
  const int r = lane % 16, kk = 4 * (lane / 16) + j;
  float v = 0.f;
  int t = tile;
  if (t < T1) {  // W1: mt = t, kt = 0; rows out, cols in
    const int o = 16 * t + r, i = kk;
    v = i < NIN ? w[OW1 + o * NIN + i] : 0.f;
  } else if ((t -= T1) < T2) {
    const int mt = t / 4, kt = t % 4, o = 16 * mt + r, i = 16 * kt + kk;
    v = w[OW2 + o * NH + i];
  } else if ((t -= T2) < T3) {
    const int mt = t / 4, kt = t % 4, o = 16 * mt + r, i = 16 * kt + kk;
    v = w[OW3 + o * NH + i];
  } else if ((t -= T3) < T4) {  // W4: mt = 0, kt = t
    const int o = r, i = 16 * t + kk;
    v = o < NOUT ? w[OW4 + o * NH + i] : 0.f;
  } else if ((t -= T4) < U2) {  // W2^T: rows = in (mt), k = out (kt)
    const int mt = t / 4, kt = t % 4, i = 16 * mt + r, o = 16 * kt + kk;
    v = w[OW2 + o * NH + i];
  } else if ((t -= U2) < U3) {
    const int mt = t / 4, kt = t % 4, i = 16 * mt + r, o = 16 * kt + kk;
    v = w[OW3 + o * NH + i];
  } else {  // W4^T: rows = in (mt = t - ...), k = out (kt = 0)
    t -= U3;
    const int i = 16 * t + r, o = kk;
    v = o < NOUT ? w[OW4 + o * NH + i] : 0.f;
  }
  pk[e] = (short)f2bf(v);
}


// Training pass over tiles of 16 points, one wavefront per tile at a time: forward,
// MSE loss, backward-data.  Writes the loss partial of its block and the activations
// and deltas (feature-major, BF16) for the weight-gradient kernels.  The weight tiles
// are read from the packed array per tile (it stays in cache): holding them in
// registers next to the activations would spill.
__global__ __launch_bounds__(64) void mlp_train_tiles(const float* __restrict__ x, const float* __restrict__ target,
                                                      long long n, const float* __restrict__ w,
                                                      const short* __restrict__ pk, Norm nm, short* __restrict__ act,
                                                      float* __restrict__ loss_part, float gscale) {
  const int lane = threadIdx.x, g = lane >> 4, r = lane & 15;
  const short* pkt = pk;  // per-tile copy of pk the compiler cannot hoist loads through
  auto Af = [&](int t) -> bf4 { return tile_of(pkt, t, lane); };
  f4 b1[4], b2[4], b3[4], b4;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      b1[t][j] = w[OB1 + 16 * t + 4 * g + j];
      b2[t][j] = w[OB2 + 16 * t + 4 * g + j];
      b3[t][j] = w[OB3 + 16 * t + 4 * g + j];
    }
#pragma unroll
  for (int j = 0; j < 4; ++j) b4[j] = (4 * g + j < NOUT) ? w[OB4 + 4 * g + j] : 0.f;
  float lsum = 0.f;
  const long long ntiles = (n + 15) / 16;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    asm volatile("" : "+s"(pkt));  // an empty statement: keeps weight loads per tile
    const long long p = tile * 16 + r;
    const bool valid = p < n;
    // inputs: B0[k = 4g + j][point r], normalised, zero beyond the 9 features
    f4 x0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int f = 4 * g + j;
      x0[j] = (valid && f < NIN) ? (x[p * NIN + f] - nm.mx[f]) * nm.isx[f] : 0.f;
    }
    const bf4 B0 = pack4(x0);
    f4 a1[4], a2[4], a3[4];
    bf4 B1[4], B2[4], B3[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f4 z = mfma(Af(t), B0, b1[t]);
#pragma unroll
      for (int j = 0; j < 4; ++j) z[j] = tanh_f(z[j]);
      a1[t] = z;
      B1[t] = pack4(z);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f4 z = b2[t];
#pragma unroll
      for (int k = 0; k < 4; ++k) z = mfma(Af(T1 + 4 * t + k), B1[k], z);
#pragma unroll
      for (int j = 0; j < 4; ++j) z[j] = tanh_f(z[j]);
      a2[t] = z;
      B2[t] = pack4(z);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f4 z = b3[t];
#pragma unroll
      for (int k = 0; k < 4; ++k) z = mfma(Af(T1 + T2 + 4 * t + k), B2[k], z);
#pragma unroll
      for (int j = 0; j < 4; ++j) z[j] = tanh_f(z[j]);
      a3[t] = z;
      B3[t] = pack4(z);
    }
    f4 out = b4;
#pragma unroll
    for (int k = 0; k < 4; ++k) out = mfma(Af(T1 + T2 + T3 + k), B3[k], out);
    {
      // MSE on normalised outputs: d4 = gscale (y - t), gscale = 2 / (6 n)
      f4 d4;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int o = 4 * g + j;
        float e = 0.f;
        if (valid && o < NOUT) e = out[j] - (target[p * NOUT + o] - nm.my[o]) / nm.sy[o];
        lsum += e * e;
        d4[j] = gscale * e;
      }
      const bf4 D4 = pack4(d4);
      // backward data: d3 = (W4^T d4) * (1 - a3^2), then d2, d1
      f4 d3[4], d2[4], d1[4];
      bf4 E3[4], E2[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        f4 z = {0.f, 0.f, 0.f, 0.f};
        z = mfma(tile_of(pkt, T1 + T2 + T3 + T4 + U2 + U3 + t, lane), D4, z);
#pragma unroll
        for (int j = 0; j < 4; ++j) z[j] *= 1.f - a3[t][j] * a3[t][j];
        d3[t] = z;
        E3[t] = pack4(z);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        f4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 4; ++k) z = mfma(tile_of(pkt, T1 + T2 + T3 + T4 + U2 + 4 * t + k, lane), E3[k], z);
#pragma unroll
        for (int j = 0; j < 4; ++j) z[j] *= 1.f - a2[t][j] * a2[t][j];
        d2[t] = z;
        E2[t] = pack4(z);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        f4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 4; ++k) z = mfma(tile_of(pkt, T1 + T2 + T3 + T4 + 4 * t + k, lane), E2[k], z);
#pragma unroll
        for (int j = 0; j < 4; ++j) z[j] *= 1.f - a1[t][j] * a1[t][j];
        d1[t] = z;
      }
      // feature-major stores: feature 16t + 4g + j of point p
      if (valid) {
        auto put = [&](int base, int t, const f4& v) {
#pragma unroll
          for (int j = 0; j < 4; ++j) act[(long long)(base + 16 * t + 4 * g + j) * n + p] = (short)f2bf(v[j]);
        };
        put(FA0, 0, x0);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          put(FA1, t, a1[t]);
          put(FA2, t, a2[t]);
          put(FA3, t, a3[t]);
          put(FD1, t, d1[t]);
          put(FD2, t, d2[t]);
          put(FD3, t, d3[t]);
        }
        put(FD4, 0, d4);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
  if (lane == 0) loss_part[blockIdx.x] = lsum;
}

// Inference over TP tiles of 16 points per wavefront at a time: the TP tiles' matrix
// products and tanh evaluations are independent, so one tile's MFMAs overlap another's
// VALU work inside the wavefront.  Outputs y (denormalised); with the range guard a
// point whose largest |normalised input| exceeds fb_bound takes the Smagorinsky stress
// of its own gradients.
template <int TP>
__global__ __launch_bounds__(64) void mlp_infer(const float* __restrict__ x, float* __restrict__ y, long long n,
                                                const float* __restrict__ w, const short* __restrict__ pk, Norm nm,
                                                float fb_c2, float fb_bound, int clip) {
  const int lane = threadIdx.x, g = lane >> 4, r = lane & 15;
  Net R;
  net_load(R, pk, w, nm, lane);
  const long long ntiles = (n + 15) / 16;
  for (long long t0 = (long long)blockIdx.x * TP; t0 < ntiles; t0 += (long long)gridDim.x * TP) {
    long long p[TP];
    bool valid[TP];
    f4 xn[TP], out[TP];
#pragma unroll
    for (int q = 0; q < TP; ++q) {
      p[q] = (t0 + q) * 16 + r;
      valid[q] = p[q] < n;
      // clamped to an in-bounds record; its values are discarded for invalid points
      xn[q] = net_input(R, x + (valid[q] ? p[q] : n - 1) * NIN, valid[q], g);
    }
    net_forward<TP>(R, xn, out);
#pragma unroll
    for (int q = 0; q < TP; ++q) {
      const f4 res = net_closure(R, out[q], xn[q], x + (valid[q] ? p[q] : n - 1) * NIN, valid[q], g, fb_c2,
                                 fb_bound, clip);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int o = 4 * g + j;
        if (valid[q] && o < NOUT) y[p[q] * NOUT + o] = res[j];
      }
    }
  }
}

// Weight gradients dW[o][i] = sum_p d[o][p] a[i][p] for one (layer, 16x16 output tile)
// per blockIdx.y and a slice of points per blockIdx.x; partials [slice][tile][256].
struct GradTile {
  int dfeat, afeat;  // feature-major rows of the delta and activation blocks
  int o0, i0;        // output rows / input columns of the tile
};
__constant__ GradTile c_gtiles[64];

__global__ __launch_bounds__(64) void mlp_wgrad(const short* __restrict__ act, long long n, int kchunks_per_block,
                                                float* __restrict__ part) {
  const int lane = threadIdx.x, g = lane >> 4, r = lane & 15;
  const GradTile gt = c_gtiles[blockIdx.y];
  const long long nk = n / 16;  // n is a multiple of 16 (the caller pads)
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  const long long k0 = (long long)blockIdx.x * kchunks_per_block;
  for (long long kc = k0; kc < k0 + kchunks_per_block && kc < nk; ++kc) {
    const long long p = kc * 16 + 4 * g;
    bf4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a[j] = act[(long long)(gt.dfeat + gt.o0 + r) * n + p + j];  // A[o][k = point]
      b[j] = act[(long long)(gt.afeat + gt.i0 + r) * n + p + j];  // B[k = point][i]
    }
    acc = mfma(a, b, acc);
  }
  // D[o0 + 4g + j][i0 + r]
  float* dst = part + ((long long)blockIdx.x * gridDim.y + blockIdx.y) * 256;
#pragma unroll
  for (int j = 0; j < 4; ++j) dst[(4 * g + j) * 16 + r] = acc[j];
}

// Bias gradients db[o] = sum_p d[o][p] for the 208 delta rows (d1 d2 d3 d4), one row
// per block, fixed-order reduction.
__global__ __launch_bounds__(256) void mlp_bgrad(const short* __restrict__ act, long long n, float* __restrict__ db) {
  __shared__ float sh[256];
  const long long row = FD1 + blockIdx.x;
  float s = 0.f;
  for (long long p = threadIdx.x; p < n; p += 256) s += bf2f((unsigned short)act[row * n + p]);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[blockIdx.x] = sh[0];
}

// Sum the weight-gradient partials into the FP32 gradient block (layout of the
// parameters) and copy the bias gradients there.
__global__ void mlp_gather_grad(const float* __restrict__ part, int nslices, int ntiles, const float* __restrict__ db,
                                float* __restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NPARAM) return;
  auto tile_sum = [&](int tile, int ro, int ci) {
    float s = 0.f;
    for (int k = 0; k < nslices; ++k) s += part[((long long)k * ntiles + tile) * 256 + ro * 16 + ci];
    return s;
  };
  float v = 0.f;
  if (e < OB1) {  // W1[o][i], tiles 0..3: (o-tile, i-tile 0)
    const int o = e / NIN, i = e % NIN;
    v = tile_sum(o / 16, o % 16, i);
  } else if (e < OW2) {
    v = db[e - OB1];
  } else if (e < OB2) {
    const int q = e - OW2, o = q / NH, i = q % NH;
    v = tile_sum(4 + (o / 16) * 4 + i / 16, o % 16, i % 16);
  } else if (e < OW3) {
    v = db[64 + (e - OB2)];
  } else if (e < OB3) {
    const int q = e - OW3, o = q / NH, i = q % NH;
    v = tile_sum(20 + (o / 16) * 4 + i / 16, o % 16, i % 16);
  } else if (e < OW4) {
    v = db[128 + (e - OB3)];
  } else if (e < OB4) {
    const int q = e - OW4, o = q / NH, i = q % NH;
    v = tile_sum(36 + i / 16, o, i % 16);
  } else {
    v = db[192 + (e - OB4)];
  }
  grad[e] = v;
}

__global__ void mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                         const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NPARAM) return;
  const float gr = grad[e];
  const float mm = b1 * m[e] + (1.f - b1) * gr;
  const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
  m[e] = mm;
  v[e] = vv;
  w[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void mlp_loss_final(const float* __restrict__ part, int nparts, float scale, float* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < nparts; ++i) s += part[i];
    *out = s * scale;
  }
}

}  // namespace psx_mlp_impl

using namespace psx_mlp_impl;

struct psx_mlp {
  float* w = nullptr;      // FP32 master parameters
  float* m = nullptr;      // Adam moments
  float* v = nullptr;
  float* grad = nullptr;
  short* pk = nullptr;     // packed BF16 operand tiles
  short* act = nullptr;    // activations / deltas of the last training batch
  long long act_n = 0;
  float* part = nullptr;   // weight-gradient partials
  float* db = nullptr;
  float* lpart = nullptr;
  Norm nm{};
  int step = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
};

namespace {
constexpr int kFwdBlocks = 2048;
constexpr int kSlices = 64;  // point slices of the weight-gradient GEMMs
constexpr int kGradTiles = 4 + 16 + 16 + 4;

int hip_ok(psx_mlp* m, hipError_t e, const char* what) {
  if (e == hipSuccess) return PSX_OK;
  m->err = std::string(what) + ": " + hipGetErrorString(e);
  return PSX_E_HIP;
}

int repack(psx_mlp* m) {
  hipLaunchKernelGGL(pack_weights, dim3((NPACK * 256 + 255) / 256), dim3(256), 0, m->stream, m->w, m->pk);
  return hip_ok(m, hipGetLastError(), "pack_weights");
}

// inference: two tiles per wavefront per iteration, at most kFwdBlocks one-wave blocks
constexpr int kInferTiles = 2;
hipError_t launch_infer(const psx_mlp* m, const float* x, float* y, long long n, float c2, float bound, int clip,
                        hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long long groups = ((n + 15) / 16 + kInferTiles - 1) / kInferTiles;
  const unsigned blocks = (unsigned)(groups < kFwdBlocks ? groups : kFwdBlocks);
  hipLaunchKernelGGL((mlp_infer<kInferTiles>), dim3(blocks), dim3(64), 0, s, x, y, n, (const float*)m->w,
                     (const short*)m->pk, m->nm, c2, bound, clip);
  return hipGetLastError();
}

int upload_tiles() {
  GradTile t[kGradTiles];
  int k = 0;
  for (int mt = 0; mt < 4; ++mt) t[k++] = GradTile{FD1, FA0, 16 * mt, 0};
  for (int mt = 0; mt < 4; ++mt)
    for (int it = 0; it < 4; ++it) t[k++] = GradTile{FD2, FA1, 16 * mt, 16 * it};
  for (int mt = 0; mt < 4; ++mt)
    for (int it = 0; it < 4; ++it) t[k++] = GradTile{FD3, FA2, 16 * mt, 16 * it};
  for (int it = 0; it < 4; ++it) t[k++] = GradTile{FD4, FA3, 0, 16 * it};
  return hipMemcpyToSymbol(HIP_SYMBOL(c_gtiles), t, sizeof(t)) == hipSuccess ? PSX_OK : PSX_E_HIP;
}
}  // namespace

extern "C" {

int psx_mlp_create(psx_mlp** out, int device, void* stream, unsigned long long seed) {
  if (!out) return PSX_E_ARG;
  *out = nullptr;
  psx_mlp* m = new psx_mlp();
  m->device = device;
  m->stream = (hipStream_t)stream;
  int rc = hip_ok(m, hipSetDevice(device), "hipSetDevice");
  const size_t pb = sizeof(float) * NPARAM;
  if (rc == PSX_OK) rc = hip_ok(m, hipMalloc(&m->w, pb), "hipMalloc");
  if (rc == PSX_OK) rc = hip_ok(m, hipMalloc(&m->m, pb), "hipMalloc");
  if (rc == PSX_OK) rc = hip_ok(m, hipMalloc(&m->v, pb), "hipMalloc");
  if (rc == PSX_OK) rc = hip_ok(m, hipMalloc(&m->grad, pb), "hipMalloc");
  if (rc == PSX_OK) rc = hip_ok(m, hipMalloc(&m->pk, sizeof(short) * NPACK * 256), "hipMalloc");
  if (rc == PSX_OK) rc = hip_ok(m, hipMalloc(&m->part, sizeof(float) * kSlices * kGradTiles * 256), "hipMalloc");
  if (rc == PSX_OK) rc = hip_ok(m, hipMalloc(&m->db, sizeof(float) * 208), "hipMalloc");
  if (rc == PSX_OK) rc = hip_ok(m, hipMalloc(&m->lpart, sizeof(float) * kFwdBlocks), "hipMalloc");
  if (rc == PSX_OK) rc = upload_tiles();
  if (rc == PSX_OK) {
    // Glorot-uniform weights from a fixed-seed generator, zero biases, unit normalisation
    std::vector<float> h(NPARAM, 0.f);
    unsigned long long s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto uni = [&]() {
      s = s * 6364136223846793005ULL + 1442695040888963407ULL;
      return (float)((s >> 11) * (1.0 / 9007199254740992.0));
    };
    auto fill = [&](int off, int fo, int fi) {
      const float a = std::sqrt(6.f / (float)(fi + fo));
      for (int k = 0; k < fo * fi; ++k) h[off + k] = (2.f * uni() - 1.f) * a;
    };
    fill(OW1, NH, NIN);
    fill(OW2, NH, NH);
    fill(OW3, NH, NH);
    fill(OW4, NOUT, NH);
    rc = hip_ok(m, hipMemcpy(m->w, h.data(), pb, hipMemcpyHostToDevice), "hipMemcpy");
    std::vector<float> z(NPARAM, 0.f);
    if (rc == PSX_OK) rc = hip_ok(m, hipMemcpy(m->m, z.data(), pb, hipMemcpyHostToDevice), "hipMemcpy");
    if (rc == PSX_OK) rc = hip_ok(m, hipMemcpy(m->v, z.data(), pb, hipMemcpyHostToDevice), "hipMemcpy");
    for (int i = 0; i < NIN; ++i) m->nm.mx[i] = 0.f, m->nm.isx[i] = 1.f;
    for (int i = 0; i < NOUT; ++i) m->nm.my[i] = 0.f, m->nm.sy[i] = 1.f;
    if (rc == PSX_OK) rc = repack(m);
  }
  if (rc != PSX_OK) {
    psx_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return PSX_OK;
}

int psx_mlp_destroy(psx_mlp* m) {
  if (!m) return PSX_E_ARG;
  (void)hipFree(m->w);
  (void)hipFree(m->m);
  (void)hipFree(m->v);
  (void)hipFree(m->grad);
  (void)hipFree(m->pk);
  (void)hipFree(m->act);
  (void)hipFree(m->part);
  (void)hipFree(m->db);
  (void)hipFree(m->lpart);
  delete m;
  return PSX_OK;
}

const char* psx_mlp_last_error(const psx_mlp* m) { return m ? m->err.c_str() : "null network"; }

int psx_mlp_set_stream(psx_mlp* m, void* stream) {
  if (!m) return PSX_E_ARG;
  m->stream = (hipStream_t)stream;
  return PSX_OK;
}

int psx_mlp_get_params(psx_mlp* m, float* host) {
  if (!m || !host) return PSX_E_ARG;
  int rc = hip_ok(m, hipStreamSynchronize(m->stream), "hipStreamSynchronize");
  if (rc == PSX_OK) rc = hip_ok(m, hipMemcpy(host, m->w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost), "hipMemcpy");
  return rc;
}

int psx_mlp_set_params(psx_mlp* m, const float* host) {
  if (!m || !host) return PSX_E_ARG;
  int rc = hip_ok(m, hipStreamSynchronize(m->stream), "hipStreamSynchronize");
  if (rc == PSX_OK) rc = hip_ok(m, hipMemcpy(m->w, host, sizeof(float) * NPARAM, hipMemcpyHostToDevice), "hipMemcpy");
  if (rc == PSX_OK) rc = repack(m);
  return rc;
}

int psx_mlp_set_norm(psx_mlp* m, const float* mx, const float* sx, const float* my, const float* sy) {
  if (!m || !mx || !sx || !my || !sy) return PSX_E_ARG;
  for (int i = 0; i < NIN; ++i) {
    if (!(sx[i] > 0.f)) return PSX_E_ARG;
    m->nm.mx[i] = mx[i];
    m->nm.isx[i] = 1.f / sx[i];
  }
  for (int i = 0; i < NOUT; ++i) {
    if (!(sy[i] > 0.f)) return PSX_E_ARG;
    m->nm.my[i] = my[i];
    m->nm.sy[i] = sy[i];
  }
  return PSX_OK;
}

int psx_mlp_forward(psx_mlp* m, const float* x, float* y, long long n) {
  if (!m || !x || !y || n < 0) return PSX_E_ARG;
  if (n == 0) return PSX_OK;
  const long long tiles = (n + 15) / 16;
  const unsigned blocks = (unsigned)(tiles < kFwdBlocks ? tiles : kFwdBlocks);
  return hip_ok(m, launch_infer(m, x, y, n, 0.f, INFINITY, 0, m->stream), "mlp_infer");
}

int psx_mlp_closure(psx_mlp* m, const float* g, float* tau, long long n, double c2, double bound, int flags) {
  if (!m || !g || !tau || n < 0 || std::isnan(bound) || (flags & ~PSX_CLOSURE_CLIP_BACKSCATTER)) {
    if (m) m->err = "bad closure arguments";
    return PSX_E_ARG;
  }
  return hip_ok(m, launch_infer(m, g, tau, n, (float)c2, (float)bound, (flags & PSX_CLOSURE_CLIP_BACKSCATTER) ? 1 : 0,
                                m->stream),
                "mlp_infer(closure)");
}

int psx_mlp_grad(psx_mlp* m, const float* x, const float* target, long long n, float* loss_dev) {
  if (!m || !x || !target || n <= 0 || (n % 16) != 0) {
    if (m) m->err = "batch must be a positive multiple of 16";
    return PSX_E_ARG;
  }
  int rc = PSX_OK;
  if (m->act_n != n) {
    (void)hipFree(m->act);
    m->act = nullptr;
    m->act_n = 0;
    rc = hip_ok(m, hipMalloc(&m->act, sizeof(short) * NFEAT * (size_t)n), "hipMalloc(act)");
    if (rc != PSX_OK) return rc;
    m->act_n = n;
  }
  const long long tiles = n / 16;
  const unsigned blocks = (unsigned)(tiles < kFwdBlocks ? tiles : kFwdBlocks);
  const float gscale = 2.f / (6.f * (float)n);
  hipLaunchKernelGGL(mlp_train_tiles, dim3(blocks), dim3(64), 0, m->stream, x, target, n, (const float*)m->w,
                     (const short*)m->pk, m->nm, m->act, m->lpart, gscale);
  rc = hip_ok(m, hipGetLastError(), "mlp_train_tiles");
  if (rc != PSX_OK) return rc;
  const int kpb = (int)((tiles + kSlices - 1) / kSlices);
  hipLaunchKernelGGL(mlp_wgrad, dim3(kSlices, kGradTiles), dim3(64), 0, m->stream, (const short*)m->act, n, kpb,
                     m->part);
  rc = hip_ok(m, hipGetLastError(), "mlp_wgrad");
  if (rc != PSX_OK) return rc;
  hipLaunchKernelGGL(mlp_bgrad, dim3(208), dim3(256), 0, m->stream, (const short*)m->act, n, m->db);
  rc = hip_ok(m, hipGetLastError(), "mlp_bgrad");
  if (rc != PSX_OK) return rc;
  hipLaunchKernelGGL(mlp_gather_grad, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, (const float*)m->part,
                     kSlices, kGradTiles, (const float*)m->db, m->grad);
  rc = hip_ok(m, hipGetLastError(), "mlp_gather_grad");
  if (rc != PSX_OK) return rc;
  if (loss_dev) {
    hipLaunchKernelGGL(mlp_loss_final, dim3(1), dim3(1), 0, m->stream, (const float*)m->lpart, (int)blocks,
                       1.f / (6.f * (float)n), loss_dev);
    rc = hip_ok(m, hipGetLastError(), "mlp_loss_final");
  }
  return rc;
}

int psx_mlp_gradient(psx_mlp* m, float* grad_dev) {
  if (!m || !grad_dev) return PSX_E_ARG;
  return hip_ok(m, hipMemcpyAsync(grad_dev, m->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToDevice, m->stream),
                "hipMemcpyAsync");
}

int psx_mlp_train_step(psx_mlp* m, const float* x, const float* target, long long n, float lr, float* loss_dev) {
  int rc = psx_mlp_grad(m, x, target, n, loss_dev);
  if (rc != PSX_OK) return rc;
  m->step += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
  hipLaunchKernelGGL(mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->w, m->m, m->v,
                     (const float*)m->grad, lr, b1, b2, eps, c1, c2);
  rc = hip_ok(m, hipGetLastError(), "mlp_adam");
  if (rc == PSX_OK) rc = repack(m);
  return rc;
}

}  // extern "C"

namespace psx {
// What the fused closure z pass needs of a network: its packed BF16 tiles, its FP32
// parameters (biases) and its normalisation.
void mlp_closure_args(const psx_mlp* m, const short** pk, const float** w, psx_mlp_impl::Norm* nm) {
  *pk = m->pk;
  *w = m->w;
  *nm = m->nm;
}

// The network as the learned closure of an LES right-hand side: stresses tau[n][6] from
// the gradients g[n][9] on stream s; points outside the normalised input range
// [-bound, bound], and with clip points where the network's stress is backscatter, take
// the Smagorinsky stress with c2 = (Cs Delta)^2.
hipError_t mlp_closure(const psx_mlp* m, const float* g, float* tau, long long n, float c2, float bound, int clip,
                       hipStream_t s) {
  return launch_infer(m, g, tau, n, c2, bound, clip, s);
}
}  // namespace psx
