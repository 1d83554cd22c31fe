// Small elementwise / reduction kernels, used only by the host library TU.
#pragma once
#include "kernels.hpp"

namespace psx {

// ------------------------------------------------------------------ small kernels
// dt from the CFL maximum (src/Classes.py:687-698).  red holds max over the grid
// of |u|/dx+|v|/dy+|w|/dz in raw (unnormalised inverse) units.
__global__ void cfl_finalize(const unsigned long long* red, double* dt, double inv_sqrt_nt, double cfl,
                             double nu, double visc_lim) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double mv = __longlong_as_double((long long)*red) * inv_sqrt_nt;
    double d = cfl / (mv + 1e-10);
    if (nu > 0) d = fmin(d, visc_lim);
    *dt = d;
  }
}

__global__ void set_u64(unsigned long long* p, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}
__global__ void set_f64(double* p, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}
// Device-side fill (used instead of hipMemsetAsync so callers may graph-capture).
__global__ void fill_f64(double* p, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Copy a state with the Nyquist entries zeroed (the reference's unpad(pad(uhat))).
__global__ void copy_truncate(const double2* __restrict__ in, double2* __restrict__ out, int N1, int N2, int Nc,
                              int N3, int truncate) {
  const long long total = 3LL * N1 * N2 * Nc;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(e % Nc);
    const long long r = e / Nc;
    const int ky = (int)(r % N2);
    const int kx = (int)((r / N2) % N1);
    const bool nyq = kx == N1 / 2 || ky == N2 / 2 || kz == N3 / 2;
    out[e] = (truncate && nyq) ? make_double2(0.0, 0.0) : in[e];
  }
}

// Diagnostics partial sums: energy, enstrophy, resolved energy with the Hermitian
// half-spectrum weights (1 at k3 = 0, 2 for 0 < k3 < N3/2, 0 at k3 = N3/2),
// src/Classes.py:652-676,733-744.  One partial triple per block; a second pass
// sums them in a fixed order, so results are bitwise reproducible.
__global__ __launch_bounds__(NT) void stats_partial(const double2* __restrict__ u, double* __restrict__ part,
                                                    int N1, int N2, int N3, int Nc, double ks0, double ks1,
                                                    double ks2, int kc) {
  __shared__ double sh[3][NT];
  const long long plane = (long long)N2 * Nc;
  const long long fs = (long long)N1 * plane;
  const long long total = fs;
  double e = 0.0, o = 0.0, er = 0.0;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < total; i += (long long)gridDim.x * NT) {
    const int kz = (int)(i % Nc);
    const long long r = i / Nc;
    const int ky = (int)(r % N2), kx = (int)(r / N2);
    const double w = kz == 0 ? 1.0 : (kz < N3 / 2 ? 2.0 : 0.0);
    if (w == 0.0) continue;
    const double k1 = ks0 * signed_freq(kx, N1), k2 = ks1 * signed_freq(ky, N2), k3 = ks2 * kz;
    const double2 a = u[i], b = u[fs + i], c = u[2 * fs + i];
    const double ea = a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y + c.x * c.x + c.y * c.y;
    // omega = i k x u
    const double2 o1 = make_double2(-(k2 * c.y - k3 * b.y), k2 * c.x - k3 * b.x);
    const double2 o2 = make_double2(-(k3 * a.y - k1 * c.y), k3 * a.x - k1 * c.x);
    const double2 o3 = make_double2(-(k1 * b.y - k2 * a.y), k1 * b.x - k2 * a.x);
    const double oa = o1.x * o1.x + o1.y * o1.y + o2.x * o2.x + o2.y * o2.y + o3.x * o3.x + o3.y * o3.y;
    const bool in1 = kx < kc || kx > N1 - kc;
    const bool in2 = ky < kc || ky > N2 - kc;
    const bool gf = in1 && in2 && kz < kc;
    e += w * ea;
    o += w * oa;
    if (gf) er += w * ea;
  }
  sh[0][threadIdx.x] = e;
  sh[1][threadIdx.x] = o;
  sh[2][threadIdx.x] = er;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
      sh[2][threadIdx.x] += sh[2][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = sh[0][0];
    part[3 * blockIdx.x + 1] = sh[1][0];
    part[3 * blockIdx.x + 2] = sh[2][0];
  }
}

__global__ void stats_final(const double* __restrict__ part, int nblk, double* __restrict__ out, double scale) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double e = 0.0, o = 0.0, er = 0.0;
    for (int b = 0; b < nblk; ++b) {
      e += part[3 * b];
      o += part[3 * b + 1];
      er += part[3 * b + 2];
    }
    out[0] = 0.5 * e * scale;
    out[1] = 0.5 * o * scale;
    out[2] = 0.5 * er * scale;
  }
}

// Shell spectrum E_c(k) with bins rint(|k|) in integer wavenumbers
// (src/Classes.py:762-797); optional sharp cutoff Gf (src/Classes.py:825-844).
__global__ void spectrum_kernel(const double2* __restrict__ u, double* __restrict__ spec, int nbins, int N1,
                                int N2, int N3, int Nc, int kc, int resolved, double scale) {
  const long long plane = (long long)N2 * Nc;
  const long long fs = (long long)N1 * plane;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < fs; i += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(i % Nc);
    const long long r = i / Nc;
    const int ky = (int)(r % N2), kx = (int)(r / N2);
    const double w = kz == 0 ? 1.0 : (kz < N3 / 2 ? 2.0 : 0.0);
    if (w == 0.0) continue;
    if (resolved) {
      const bool in1 = kx < kc || kx > N1 - kc;
      const bool in2 = ky < kc || ky > N2 - kc;
      if (!(in1 && in2 && kz < kc)) continue;
    }
    const double a1 = signed_freq(kx, N1), a2 = signed_freq(ky, N2), a3 = kz;
    const int b = (int)rint(sqrt(a1 * a1 + a2 * a2 + a3 * a3));
    if (b >= nbins) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double2 v = u[c * fs + i];
      atomicAdd(&spec[3 * b + c], w * (v.x * v.x + v.y * v.y) * scale);
    }
  }
}

}  // namespace psx
