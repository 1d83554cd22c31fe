// Small elementwise / reduction kernels, used only by the host library TU.
#pragma once
#include "kernels.hpp"

namespace psx {

// ------------------------------------------------------------------ small kernels
// dt from the CFL maximum (src/Classes.py:687-698).  part holds per-block maxima of
// |u|/dx+|v|/dy+|w|/dz in raw (unnormalised inverse) units; one block of NT threads
// per ensemble member (block m reads part + m*nparts, writes dt[m], viscosity
// nu_m[m] when nu_m is set).  ksum = 1/dx^2 + 1/dy^2 + 1/dz^2.  As numpy's amax and
// minimum do, a NaN maximum gives dt = NaN and an infinite one dt = 0; either raises
// the context's non-finite flag (*flag = 1, never cleared here).
__global__ __launch_bounds__(NT) void cfl_finalize(const double* __restrict__ part, int nparts, double* dt,
                                                   double inv_sqrt_nt, double cfl, double nu_all,
                                                   const double* __restrict__ nu_m, double ksum, int* flag) {
  __shared__ double sh[NT];
  part += (long long)blockIdx.x * nparts;
  const double nu = nu_m ? nu_m[blockIdx.x] : nu_all;
  double m = 0.0;
  for (int i = threadIdx.x; i < nparts; i += NT) m = max_nan(m, part[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = max_nan(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double mv = sh[0] * inv_sqrt_nt;
    double d = cfl / (mv + 1e-10);
    if (nu > 0 && d == d) d = fmin(d, 0.634 / ksum / nu * cfl / 1.35);
    dt[blockIdx.x] = d;
    if (!isfinite(mv)) *flag = 1;
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

// A NaN or infinite energy or enstrophy raises the context's non-finite flag (when
// flag is set).  One block of NT threads: thread i sums partials i, i + NT, ... in
// order, then a fixed tree, so the result is bitwise reproducible.
__global__ __launch_bounds__(NT) void stats_final(const double* __restrict__ part, int nblk,
                                                  double* __restrict__ out, double scale, int* flag) {
  __shared__ double sh[3][NT];
  double e = 0.0, o = 0.0, er = 0.0;
  for (int b = threadIdx.x; b < nblk; b += NT) {
    e += part[3 * b];
    o += part[3 * b + 1];
    er += part[3 * b + 2];
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
    out[0] = 0.5 * sh[0][0] * scale;
    out[1] = 0.5 * sh[1][0] * scale;
    out[2] = 0.5 * sh[2][0] * scale;
    if (flag && (!isfinite(sh[0][0]) || !isfinite(sh[1][0]))) *flag = 1;
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

// Velocity as the t-model family sees it: Nyquist entries zeroed and, with
// kf > 0, the sharp test filter |k_i| < kf (src/Classes.py:132-150).
__global__ void filter_copy(const double2* __restrict__ in, double2* __restrict__ out, int N1, int N2, int N3,
                            int Nc, double ks0, double ks1, double ks2, double kf) {
  const long long total = 3LL * N1 * N2 * Nc;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(e % Nc);
    const long long r = e / Nc;
    const int ky = (int)(r % N2);
    const int kx = (int)((r / N2) % N1);
    bool keep = kx != N1 / 2 && ky != N2 / 2 && kz != N3 / 2;
    if (kf > 0) {
      keep = keep && fabs(ks0 * signed_freq(kx, N1)) < kf && fabs(ks1 * signed_freq(ky, N2)) < kf &&
             fabs(ks2 * kz) < kf;
    }
    out[e] = keep ? in[e] : make_double2(0.0, 0.0);
  }
}

// Dynamic tau from the three reductions (src/RHSfunctions.py:822-861):
// tau = |LE| / (|PLQLE_filt| - |filt_PLQLE| + 1e-5), used clamped at 0.
// scal[0] raw tau, scal[1] clamped tau, scal[2..7] LE, PLQLE_filt, filt_PLQLE (re, im).
// One block of NT threads: each thread sums a strided share of each partial array,
// then a fixed-order tree over the block (the same result for a given partial count).
__global__ __launch_bounds__(NT) void tau_finalize(const double* __restrict__ pLE, int nLE,
                                                   const double* __restrict__ pA, int nA,
                                                   const double* __restrict__ pB, int nB, double inv_nt,
                                                   double* __restrict__ scal) {
  __shared__ double sh[6][NT];
  const int tid = threadIdx.x;
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = tid; i < nLE; i += NT) { v[0] += pLE[2 * i]; v[1] += pLE[2 * i + 1]; }
  for (int i = tid; i < nA; i += NT) { v[2] += pA[2 * i]; v[3] += pA[2 * i + 1]; }
  for (int i = tid; i < nB; i += NT) { v[4] += pB[2 * i]; v[5] += pB[2 * i + 1]; }
#pragma unroll
  for (int k = 0; k < 6; ++k) sh[k][tid] = v[k];
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if (tid < st) {
#pragma unroll
      for (int k = 0; k < 6; ++k) sh[k][tid] += sh[k][tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    double le[2] = {sh[0][0], sh[1][0]}, a[2] = {sh[2][0], sh[3][0]}, b[2] = {sh[4][0], sh[5][0]};
    for (int j = 0; j < 2; ++j) { le[j] *= inv_nt; a[j] *= inv_nt; b[j] *= inv_nt; }
    const double tau = hypot(le[0], le[1]) / (hypot(a[0], a[1]) - hypot(b[0], b[1]) + 1.0e-5);
    scal[0] = tau;
    scal[1] = fmax(0.0, tau);
    scal[2] = le[0]; scal[3] = le[1];
    scal[4] = a[0]; scal[5] = a[1];
    scal[6] = b[0]; scal[7] = b[1];
  }
}

// out = q0 + dt*c*(PL + tau*PLQL + Coriolis) (or the RHS itself when q0 is null);
// w0 = tau*PLQL when requested.  pl/plql hold k3 < N3/2; the k3 = N3/2 plane is zero
// (unpad_2x drops it).  src/RHSfunctions.py:480-491.
__global__ void tau_update(const double2* __restrict__ pl, const double2* __restrict__ plql,
                           const double* __restrict__ tau, const double2* __restrict__ uin,
                           const double2* __restrict__ q0, double2* __restrict__ out, double2* __restrict__ w0,
                           const double* __restrict__ dt, double c, int N1, int N2, int N3, int Nc, double ks0,
                           double ks1, double ks2, int zero_nyq_out, int rotate, double om0, double om1, double om2,
                           double force) {
  const long long fs = (long long)N1 * N2 * Nc;
  const double t = *tau;
  const double dtc = q0 ? (*dt) * c : 0.0;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < fs; e += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(e % Nc);
    const long long r = e / Nc;
    const int ky = (int)(r % N2);
    const int kx = (int)(r / N2);
    const bool nyq = kx == N1 / 2 || ky == N2 / 2 || kz == N3 / 2;
    double2 rhs[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double2 w = make_double2(0.0, 0.0), p = make_double2(0.0, 0.0);
      if (kz < N3 / 2) {
        w = cscale(plql[i * fs + e], t);
        p = pl[i * fs + e];
      }
      if (w0) w0[i * fs + e] = w;
      rhs[i] = cadd(p, w);
    }
    if (rotate && !nyq) {
      const double k[3] = {ks0 * signed_freq(kx, N1), ks1 * signed_freq(ky, N2), ks2 * kz};
      const double kinv = 1.0 / (k[0] * k[0] + k[1] * k[1] + k[2] * k[2] + 1e-50);
      const double2 u0 = uin[e], u1 = uin[fs + e], u2 = uin[2 * fs + e];
      double2 cr[3];
      cr[0] = csub(cscale(u1, om2), cscale(u2, om1));
      cr[1] = csub(cscale(u2, om0), cscale(u0, om2));
      cr[2] = csub(cscale(u0, om1), cscale(u1, om0));
      const double2 X = cadd(cadd(cscale(cr[0], k[0]), cscale(cr[1], k[1])), cscale(cr[2], k[2]));
#pragma unroll
      for (int i = 0; i < 3; ++i) rhs[i] = cadd(rhs[i], csub(cscale(cr[i], 2.0), cscale(X, 2.0 * k[i] * kinv)));
    }
    // linear forcing A u of the truncated velocity (zero at the Nyquist entries)
    if (force != 0.0 && !nyq) {
#pragma unroll
      for (int i = 0; i < 3; ++i) rhs[i] = cadd(rhs[i], cscale(uin[i * fs + e], force));
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double2 o = q0 ? cadd(q0[i * fs + e], cscale(rhs[i], dtc)) : rhs[i];
      if (zero_nyq_out && nyq) o = make_double2(0.0, 0.0);
      out[i * fs + e] = o;
    }
  }
}

// Closure coefficient from the device step-start time: Ct*t (t-model,
// src/RHSfunctions.py:342) or max(0, tau_a + tau_b t) (static tau, :480-485).
__global__ void time_coef(const double* t, int tmodel, double Ct, double ta, double tb, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = tmodel ? Ct * (*t) : fmax(0.0, ta + tb * (*t));
}

// t += dt on the device (src/PySpec_3d.py:125).
__global__ void advance_time(double* t, const double* dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *t += *dt;
}

}  // namespace psx

namespace psx {

// out = G(k) in over nf fields of the state layout (filters: filter_gain)
__global__ void apply_filter(const double2* __restrict__ in, double2* __restrict__ out, int nf, int N1, int N2,
                             int Nc, int kind, double kc, double delta, double ks0, double ks1, double ks2) {
  const double ks[3] = {ks0, ks1, ks2};
  const long long total = (long long)nf * N1 * N2 * Nc;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(e % Nc);
    const long long r = e / Nc;
    const int ky = (int)(r % N2);
    const int kx = (int)((r / N2) % N1);
    out[e] = cscale(in[e], filter_gain(kind, signed_freq(kx, N1), signed_freq(ky, N2), kz, kc, delta, ks));
  }
}

// w0_i = -i k_j tau_ij + i k_i (i/k^2 k_m k_j tau_mj), the divergence/projection form
// of src/Classes.py:911-918 (tau order 11 22 33 12 13 23, state layout)
__global__ void sgs_w0(const double2* __restrict__ tau, double2* __restrict__ w0, int N1, int N2, int Nc,
                       double ks0, double ks1, double ks2) {
  const long long fs = (long long)N1 * N2 * Nc;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < fs;
       e += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(e % Nc);
    const long long r = e / Nc;
    const int ky = (int)(r % N2);
    const int kx = (int)(r / N2);
    const double k1 = ks0 * signed_freq(kx, N1), k2 = ks1 * signed_freq(ky, N2), k3 = ks2 * kz;
    const double2 t11 = tau[e], t22 = tau[fs + e], t33 = tau[2 * fs + e];
    const double2 t12 = tau[3 * fs + e], t13 = tau[4 * fs + e], t23 = tau[5 * fs + e];
    const double kinv = 1.0 / (k1 * k1 + k2 * k2 + k3 * k3 + 1e-50);
    double2 kk = cscale(t11, k1 * k1);
    kk = cadd(kk, cscale(t22, k2 * k2));
    kk = cadd(kk, cscale(t33, k3 * k3));
    kk = cadd(kk, cscale(t12, 2.0 * k1 * k2));
    kk = cadd(kk, cscale(t13, 2.0 * k1 * k3));
    kk = cadd(kk, cscale(t23, 2.0 * k2 * k3));
    // i k_i * (i kinv kk) = -k_i kinv kk
    const double2 p = cscale(kk, -kinv);
    const double2 d1 = cadd(cadd(cscale(t11, k1), cscale(t12, k2)), cscale(t13, k3));
    const double2 d2 = cadd(cadd(cscale(t12, k1), cscale(t22, k2)), cscale(t23, k3));
    const double2 d3 = cadd(cadd(cscale(t13, k1), cscale(t23, k2)), cscale(t33, k3));
    w0[e] = cadd(cmul_i<-1>(d1), cscale(p, k1));
    w0[fs + e] = cadd(cmul_i<-1>(d2), cscale(p, k2));
    w0[2 * fs + e] = cadd(cmul_i<-1>(d3), cscale(p, k3));
  }
}

}  // namespace psx

namespace psx {

// ---- dynamic Smagorinsky (src/RHSfunctions.py:1663-1862)
// The spectral store D[f][k1][k2][k3 < H] holds, for the velocity and its test-filtered
// copy, the products N_ij (0..5 / 12..17) and T_ij = -2 c2 FFT(|S| S_ij) (6..11 / 18..23),
// c2 = Delta^2 and alpha^2 Delta^2 respectively (index order 11 22 33 12 13 23).
// Cs^2 = mean(M_ij L_ij)/mean(M_ij M_ij) over the padded grid with
// M_ij = TF_ij - G T_ij and L_ij = G N_ij - NF_ij (G: the sharp test filter at kL); the
// fields are band-limited, so the means are the Hermitian inner products of the
// spectra (weight 2 for 0 < k3), computed here as per-block partial sums.
__global__ __launch_bounds__(NT) void dsmag_partials(const double2* __restrict__ D, double* __restrict__ part,
                                                     int N1, int N2, int H, double kL) {
  const long long fs = (long long)N1 * N2 * H;
  double num = 0.0, den = 0.0;
  for (long long e = (long long)blockIdx.x * NT + threadIdx.x; e < fs; e += (long long)gridDim.x * NT) {
    const int kz = (int)(e % H);
    const long long r = e / H;
    const int ky = (int)(r % N2), kx = (int)(r / N2);
    const double g = (abs(signed_freq(kx, N1)) < kL && abs(signed_freq(ky, N2)) < kL && kz < kL) ? 1.0 : 0.0;
    const double w = kz == 0 ? 1.0 : 2.0;
    double sn = 0.0, sd = 0.0;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const double2 Nv = D[m * fs + e], T = D[(6 + m) * fs + e];
      const double2 NF = D[(12 + m) * fs + e], TF = D[(18 + m) * fs + e];
      const double2 M = csub(TF, cscale(T, g));
      const double2 L = csub(cscale(Nv, g), NF);
      sn += M.x * L.x + M.y * L.y;
      sd += M.x * M.x + M.y * M.y;
    }
    num += w * sn;
    den += w * sd;
  }
  __shared__ double sh[2][NT];
  sh[0][threadIdx.x] = num;
  sh[1][threadIdx.x] = den;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0];
    part[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// cs2[0] = cs2[1] = sum num / sum den; one block of NT threads sums strided shares of the
// partials, then combines them in a fixed-order tree.
__global__ __launch_bounds__(NT) void dsmag_finalize(const double* __restrict__ part, int nparts,
                                                     double* __restrict__ cs2) {
  __shared__ double sh[2][NT];
  const int tid = threadIdx.x;
  double n = 0.0, d = 0.0;
  for (int i = tid; i < nparts; i += NT) {
    n += part[2 * i];
    d += part[2 * i + 1];
  }
  sh[0][tid] = n;
  sh[1][tid] = d;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if (tid < st) {
      sh[0][tid] += sh[0][tid + st];
      sh[1][tid] += sh[1][tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    cs2[0] = sh[0][0] / sh[1][0];
    cs2[1] = sh[0][0] / sh[1][0];
  }
}

struct DsmagArgs {
  const double2* D;      // spectral store (see dsmag_partials)
  const double2* uprev;  // velocity the RHS is evaluated at (state layout)
  const double2* q0;     // stage base (nullptr: write the RHS)
  double2* out;
  double2* w0;           // optional SGS term
  const double* cs2;     // device Cs^2
  const double* dt;
  double c, nu;
  double ks[3];
  int N1, N2, N3, Nc, H;
  int rotate;
  double Om[3];
  double force;  // linear forcing: RHS += force * u
};

// RHS_i = -i k_j N_ij - i k_i p - nu k^2 u_i + Cs^2 w0_i(T) (+ Coriolis), with the
// Smagorinsky projection form of w0 (src/RHSfunctions.py:1836-1862), then the
// low-storage stage update q0 + dt c RHS.
__global__ void dsmag_final(DsmagArgs a) {
  const long long fs = (long long)a.N1 * a.N2 * a.Nc, ds = (long long)a.N1 * a.N2 * a.H;
  const double cs2 = *a.cs2;
  const double dtc = a.q0 ? (*a.dt) * a.c : 0.0;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < fs;
       e += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(e % a.Nc);
    const long long r = e / a.Nc;
    const int ky = (int)(r % a.N2), kx = (int)(r / a.N2);
    const double k[3] = {a.ks[0] * signed_freq(kx, a.N1), a.ks[1] * signed_freq(ky, a.N2), a.ks[2] * kz};
    const double ksq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    const double kinv = 1.0 / (ksq + 1e-50);
    double2 Nm[6], T[6];
    const long long di = r * a.H + kz;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      Nm[m] = kz < a.H ? a.D[m * ds + di] : make_double2(0.0, 0.0);
      T[m] = kz < a.H ? a.D[(6 + m) * ds + di] : make_double2(0.0, 0.0);
    }
    double2 u[3];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) u[cc] = a.uprev[cc * fs + e];
    // symmetric components: t(i, j) with 11 22 33 12 13 23 at 0..5
    auto idx = [](int i, int j) { return i == j ? i : (i + j == 1 ? 3 : (i + j == 2 ? 4 : 5)); };
    double2 kkN = make_double2(0.0, 0.0), kkT = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        kkN = cadd(kkN, cscale(Nm[idx(i, j)], k[i] * k[j]));
        kkT = cadd(kkT, cscale(T[idx(i, j)], k[i] * k[j]));
      }
    double2 p = cscale(kkN, -kinv);
    double2 cr[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    if (a.rotate) {
      cr[0] = csub(cscale(u[1], a.Om[2]), cscale(u[2], a.Om[1]));
      cr[1] = csub(cscale(u[2], a.Om[0]), cscale(u[0], a.Om[2]));
      cr[2] = csub(cscale(u[0], a.Om[1]), cscale(u[1], a.Om[0]));
      // phat -= 2 ksqr_i 1j (k . (u x Om)) (src/RHSfunctions.py:1689-1692)
      const double2 X = cadd(cadd(cscale(cr[0], k[0]), cscale(cr[1], k[1])), cscale(cr[2], k[2]));
      p = csub(p, cmul_i<1>(cscale(X, 2.0 * kinv)));
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double2 divN = make_double2(0.0, 0.0), divT = make_double2(0.0, 0.0);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        divN = cadd(divN, cscale(Nm[idx(i, j)], k[j]));
        divT = cadd(divT, cscale(T[idx(i, j)], k[j]));
      }
      // w0_i = -i k_j tau_ij - k_i (k k tau)/k^2 with tau = Cs^2 T
      const double2 w = cscale(csub(cmul_i<-1>(divT), cscale(kkT, k[i] * kinv)), cs2);
      double2 rhs = cadd(cmul_i<-1>(divN), cmul_i<-1>(cscale(p, k[i])));
      rhs = cadd(rhs, cscale(u[i], a.force - a.nu * (ksq + 1e-50)));
      rhs = cadd(rhs, w);
      if (a.rotate) rhs = cadd(rhs, cscale(cr[i], 2.0));
      if (a.w0) a.w0[i * fs + e] = w;
      a.out[i * fs + e] = a.q0 ? cadd(a.q0[i * fs + e], cscale(rhs, dtc)) : rhs;
    }
  }
}

// Restriction of nf spectral fields of a finer grid (S1, S2, Sc) to this grid
// (N1, N2, Nc): the signed modes |k1| < N1/2, |k2| < N2/2, 0 <= k3 < Nc - 1 are copied
// times scale; this grid's Nyquist rows and plane are zero.
__global__ void restrict_modes(const double2* __restrict__ src, double2* __restrict__ dst, int nf, int N1, int N2,
                               int Nc, int S1, int S2, int Sc, double scale) {
  const long long per = (long long)N1 * N2 * Nc, total = per * nf;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(e / per);
    long long r = e - f * per;
    const int i1 = (int)(r / ((long long)N2 * Nc));
    r -= (long long)i1 * N2 * Nc;
    const int i2 = (int)(r / Nc), k3 = (int)(r - (long long)i2 * Nc);
    const int k1 = i1 < N1 / 2 ? i1 : i1 - N1, k2 = i2 < N2 / 2 ? i2 : i2 - N2;
    double2 v = make_double2(0.0, 0.0);
    if (k1 != -N1 / 2 && k2 != -N2 / 2 && k3 != Nc - 1) {
      const int s1 = k1 >= 0 ? k1 : k1 + S1, s2 = k2 >= 0 ? k2 : k2 + S2;
      const double2 w = src[(((long long)f * S1 + s1) * S2 + s2) * Sc + k3];
      v = make_double2(scale * w.x, scale * w.y);
    }
    dst[e] = v;
  }
}

}  // namespace psx
