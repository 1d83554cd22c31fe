// Fused z pass with one wavefront per position group, for z lengths L = 64 R0.
//
// Same contract as zpass (kernels.hpp): c2r of the real z-pencils of P positions,
// pointwise products / closure on the physical z line, r2c of the results with the
// truncation to Hout modes.  The difference is the data path.  A lane l owns the
// entries k = l + 64 m (m < R0) of every complex pencil, and a length-L transform
// x_n = sum_k Z_k w_L^{S n k} runs as
//   Y_a(l)        = w_L^{S a l} sum_m Z_{l+64m} w_R0^{S a m}     (registers)
//   x_{a + R0 b}  = sum_l Y_a(l) w_64^{S b l}                     (LDS, radix 8 x 8)
// with b = b1 + 8 b2 and l = j + 8 r.  The pencils' half spectra are read from
// global memory straight into the first (register) step, the pointwise phase
// feeds the forward transform's register step directly, and a workgroup is one
// wavefront, so there is no workgroup barrier and no index division anywhere.
#pragma once
#include "kernels.hpp"

namespace psx {


// Waves per SIMD the register allocation must allow: the Smagorinsky pass at L = 192
// with 64-entry half spectra fits 4 (128 VGPRs, no scratch); every other pass keeps
// 1.
template <int L, int MODE, int HK> constexpr int zw_min_waves() {
  return (MODE == Z_SMAG && L == 192 && HK == 64) ? 4 : 1;
}

template <int L> struct ZWave {
  static constexpr int R0 = L / 64;
  static constexpr bool ok = (L % 64 == 0) && (R0 == 1 || R0 == 2 || R0 == 3 || R0 == 4 || R0 == 6 ||
                                               R0 == 8 || R0 == 12);
  // complex pencils transformed together through the radix-8 passes
  static constexpr int G = R0 >= 8 ? 1 : 8 / R0;
  // LDS slot of one 64-point sub-transform: 64 entries, pass-1 output at b1*9 + j
  static constexpr int SUB = 72;
  // natural-order rows: one pad entry per 16, except at L = 192:
  // no pad inside the row and a row stride of 202 entries, which a bank model of the
  // second pass's stride-3 writes and the lanes' contiguous reads finds conflict-free
  // where the padded layout puts n and n + 16 (or 17) in the same banks
  static constexpr bool FLAT = L == 192;
  static constexpr int NATS = FLAT ? 202 : L + L / 16;
  static __device__ __forceinline__ int nat(int n) { return FLAT ? n : n + (n >> 4); }
};

// One workgroup = one wavefront.  The compiler drops the s_barrier of a
// single-wavefront workgroup; the fence keeps LDS accesses in program order.
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// zin without a divergent branch: the index is clamped into the row and the loaded
// bits are masked (an exact +0 outside the row), so the value is always used and
// the loads of a pencil group are issued together instead of inside per-lane ifs.
__device__ __forceinline__ double2 zw_mask(double2 v, bool keep) {
  const unsigned long long m = keep ? ~0ull : 0ull;
  return make_double2(__longlong_as_double(__double_as_longlong(v.x) & m),
                      __longlong_as_double(__double_as_longlong(v.y) & m));
}
// HK > 0: every input field has HK half-spectrum entries (a compile-time length).
template <int HK>
__device__ __forceinline__ double2 zw_in(const ZArgs& a, const ZMember& zm, int f, long long pos, int k) {
  // the field's buffer is chosen by (wave-uniform) selects, not a branch
  const bool first = f < a.nin0;
  const double2* base = first ? zm.in0 : zm.in1;
  const long long sf = first ? a.in_sf0 : a.in_sf1;
  const int H = HK > 0 ? HK : (first ? a.Hin0 : a.Hin1);
  const int g = first ? f : f - a.nin0;
  const double2 v = base[(long long)g * sf + pos * H + (k < H ? k : H - 1)];
  return zw_mask(v, k < H);
}

// Register chunk m (entries k = l + 64 m, l < 64) of a c2r input whose half spectra
// hold HK entries is zero in every lane when each of its k has min(k, L - k) >= HK:
// with the 3/2 rule (HK = L/3) that is the middle third of the chunks, the zero pad.
template <int L, int HK>
__host__ __device__ constexpr bool zw_zero_chunk(int m) {
  if (HK <= 0) return false;
  for (int l = 0; l < 64; ++l) {
    const int k = l + 64 * m, kk = k > L / 2 ? L - k : k;
    if (kk < HK) return false;
  }
  return true;
}

// The recipes that build the real z-pencils from the input fields (ZFields,
// kernels.hpp).  Every mode without a recipe here uses input field f as pencil f.
template <int MODE> __host__ __device__ constexpr bool zw_uses(int fr, int fi) {
  // gradient pencils A_ij = du_i/dx_j, row-major: ux uy uz vx vy vz wx wy wz from the
  // y-pass fields 0 u, 1 v, 2 w, 3 u_x, 4 v_x, 5 w_x, 6 u_y, 7 v_y, 8 w_y
  constexpr int src[9] = {3, 6, 0, 4, 7, 1, 5, 8, 2};
  if (MODE == Z_GRAD) return src[fr] == fi;
  if (MODE == Z_LRNF) return fr < 3 ? fr == fi : src[fr - 3] == fi;  // u v w, then the gradients
  if (MODE != Z_SMAG) return fr == fi;
  switch (fr) {
    case 0: return fi == 0;
    case 1: return fi == 1;
    case 2: return fi == 2;
    case 3: return fi == 3;
    case 4: return fi == 4;
    case 5: return fi == 2;
    case 6: return fi == 5;
    case 7: return fi == 0 || fi == 6;
    default: return fi == 1 || fi == 7;
  }
}
// bit fi of bits[g][p]: input field fi of position p is read by the pencils of
// complex rows [g*G, g*G + G) -- a table built at compile time
template <int MODE, int P, int NRI, int G> struct ZGroupUse {
  static constexpr int NG = ((NRI + 1) / 2 + G - 1) / G;
  unsigned bits[NG][P];
  constexpr ZGroupUse() : bits{} {
    for (int g = 0; g < NG; ++g)
      for (int r = 2 * g * G; r < 2 * (g + 1) * G && r < NRI; ++r)
        for (int fi = 0; fi < ZFields<MODE>::n; ++fi)
          if (zw_uses<MODE>(r / P, fi)) bits[g][r % P] |= 1u << fi;
  }
};
// pencil r's half-spectrum entry from the raw field entries of its position
template <int MODE>
__device__ __forceinline__ double2 zw_recipe(int fr, const double2* raw, double kz) {
  if constexpr (MODE == Z_SMAG) {
    switch (fr) {
      case 0: return raw[0];
      case 1: return raw[1];
      case 2: return raw[2];
      case 3: return raw[3];
      case 4: return raw[4];
      case 5: return cscale(cmul_i<1>(raw[2]), kz);
      case 6: return raw[5];
      case 7: return cscale(cadd(cscale(cmul_i<1>(raw[0]), kz), raw[6]), 0.5);
      default: return cscale(cadd(cscale(cmul_i<1>(raw[1]), kz), raw[7]), 0.5);
    }
  } else if constexpr (MODE == Z_GRAD || MODE == Z_LRNF) {
    if constexpr (MODE == Z_LRNF) {
      if (fr < 3) return raw[fr];
      fr -= 3;
    }
    switch (fr) {
      case 0: return raw[3];
      case 1: return raw[6];
      case 2: return cscale(cmul_i<1>(raw[0]), kz);
      case 3: return raw[4];
      case 4: return raw[7];
      case 5: return cscale(cmul_i<1>(raw[1]), kz);
      case 6: return raw[5];
      case 7: return raw[8];
      default: return cscale(cmul_i<1>(raw[2]), kz);
    }
  } else {
    return raw[fr];
  }
}

// Entry k of complex pencil q = (real pencils 2q, 2q+1) of the Hermitian pair
// Z = A + iB, A and B the half spectra at kk = min(k, L - k) (c2r semantics: real
// part only at k = 0 and L/2).  raw[p][fi]: input field fi of position p at kk.
template <int L, int MODE, int P, int NRI, int NF>
__device__ __forceinline__ double2 zw_hermitian(int q, int k, double kz, const double2 (&raw)[P][NF]) {
  const bool mirror = k > L / 2;
  const double2 A = zw_recipe<MODE>((2 * q) / P, raw[(2 * q) % P], kz);
  const double2 B = (2 * q + 1 < NRI) ? zw_recipe<MODE>((2 * q + 1) / P, raw[(2 * q + 1) % P], kz)
                                      : make_double2(0.0, 0.0);
  if (k == 0 || k == L / 2) return make_double2(A.x, B.x);
  return mirror ? make_double2(A.x + B.y, B.x - A.y) : make_double2(A.x - B.y, A.y + B.x);
}

// Row b2 of the forward transform's second radix-8 pass holds x_n for
// n = a + R0 (b1 + 8 b2), a < R0, b1 < 8, i.e. n in [8 R0 b2, 8 R0 (b2 + 1)).  With
// HK half-spectrum entries kept, the r2c split reads only n < HK and n > L - HK, so
// a row whose every n lies in [HK, L - HK] is never read.
template <int L, int HK>
__host__ __device__ constexpr bool zw_row_read(int b2) {
  if (HK <= 0) return true;
  constexpr int R0 = L / 64;
  const int lo = R0 * 8 * b2, hi = lo + R0 * 8 - 1;
  return lo < HK || hi > L - HK;
}

// The 64-point transforms of `nrows` pencils (R0 sub-transforms each) held in LDS
// slots f*SUB + l; the result x_{a + R0 b} of pencil g lands in natural order at
// g*NATS + nat(n) (for the forward transform with HK > 0, only the rows the r2c
// split reads).  w8 = w_64^{l mod 8}, the lane's one table twiddle, is read once per
// kernel by the caller: a table read here would wait on the vector memory counter,
// which retires loads in order, and so on the next group's prefetched loads too.
template <int L, int SIGN, int HK = 0>
__device__ __forceinline__ void zw_sub64(double2* buf, double2 w8, int l, int nrows) {
  using W = ZWave<L>;
  constexpr int R0 = W::R0, SUB = W::SUB, NB = W::G * R0 * 8, IT = (NB + 63) / 64;
  double2 y[IT][8];
  // pass 1: 8-point transforms over r (l = j + 8 r), twiddle w_64^{S j b1}
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int u = l + 64 * it, f = u >> 3, j = u & 7;
    if ((NB % 64 == 0 || u < NB) && f < nrows * R0) {
#pragma unroll
      for (int r = 0; r < 8; ++r) y[it][r] = buf[f * SUB + j + 8 * r];
      Dft<8, SIGN>::run(y[it]);
      // twiddle w^b1 as a running power of w = w_64^{S j}: two live values instead of
      // seven (a few ulp, like the table's own rounding)
      // j = l mod 8; the empty asm hides that w8 is the same every call, so the
      // compiler keeps computing the powers here instead of holding all seven live
      double2 w1 = w8;
      asm volatile("" : "+v"(w1.x), "+v"(w1.y));
      if constexpr (SIGN > 0) w1.y = -w1.y;
      double2 wc = w1;
#pragma unroll
      for (int b1 = 1; b1 < 8; ++b1) {
        y[it][b1] = cmul(y[it][b1], wc);
        if (b1 < 7) wc = cmul(wc, w1);
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int u = l + 64 * it, f = u >> 3, j = u & 7;
    if ((NB % 64 == 0 || u < NB) && f < nrows * R0) {
#pragma unroll
      for (int b1 = 0; b1 < 8; ++b1) buf[f * SUB + b1 * 9 + j] = y[it][b1];
    }
  }
  wave_sync();
  // pass 2: 8-point transforms over j for each b1
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int u = l + 64 * it, f = u >> 3, b1 = u & 7;
    if ((NB % 64 == 0 || u < NB) && f < nrows * R0) {
#pragma unroll
      for (int j = 0; j < 8; ++j) y[it][j] = buf[f * SUB + b1 * 9 + j];
      Dft<8, SIGN>::run(y[it]);
    }
  }
  wave_sync();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int u = l + 64 * it, f = u >> 3, b1 = u & 7;
    if ((NB % 64 == 0 || u < NB) && f < nrows * R0) {
      const int g = f / R0, av = f - g * R0;
      if constexpr (SIGN < 0 && HK > 0) {  // rows the r2c split never reads are skipped
#pragma unroll
        for (int b2 = 0; b2 < 8; ++b2)
          if (zw_row_read<L, HK>(b2)) buf[g * W::NATS + W::nat(av + R0 * (b1 + 8 * b2))] = y[it][b2];
      } else {
#pragma unroll
        for (int b2 = 0; b2 < 8; ++b2) buf[g * W::NATS + W::nat(av + R0 * (b1 + 8 * b2))] = y[it][b2];
      }
    }
  }
  wave_sync();
}

// HK > 0: the input and output half spectra hold HK entries (the launcher checks
// a.Hin0, a.Hin1 and a.Hout), so the zero-pad chunks are neither loaded nor held.
template <int L, int P, int MODE, int HK = 0>
__global__ __launch_bounds__(64, (zw_min_waves<L, MODE, HK>())) void zwave(ZArgs a, const double2* __restrict__ tw) {
  using W = ZWave<L>;
  static_assert(W::ok, "zwave: L must be 64 * {1,2,3,4,6,8,12}");
  static_assert(W::G * W::NATS <= W::G * W::R0 * W::SUB, "zwave: natural-order rows exceed the LDS buffer");
  constexpr int R0 = W::R0, G = W::G, SUB = W::SUB;
  constexpr int NIN = ZShape<MODE>::nin, NOUT = ZShape<MODE>::nout;
  constexpr int NRI = NIN * P, NRO = NOUT * P;
  constexpr int NCI = (NRI + 1) / 2, NCO = (NRO + 1) / 2;
  __shared__ double2 buf[G * R0 * SUB];
  const int l = threadIdx.x;
  const ZMember zm = zmember(a, (long long)blockIdx.x * P);
  const long long pos0 = zm.pos0;

  // register-step twiddles w_L^{-a l} (the inverse uses the conjugates)
  double2 t1[R0];
#pragma unroll
  for (int av = 0; av < R0; ++av) t1[av] = tw[av * l];
  const double2 w8 = tw[R0 * (l & 7)];  // w_64^{l mod 8}, for every 64-point pass

  // ---- inverse transforms; lane l keeps x_n, n = l + 64 m, of every pencil
  constexpr int NF = ZFields<MODE>::n;
  constexpr ZGroupUse<MODE, P, NRI, G> use{};
  // every raw entry a pencil group reads, loaded before any is used; with
  // the next group's loads are issued before this group's LDS passes
  auto load_group = [&](int g0, double2 (&raw)[R0][P][NF]) {
#pragma unroll
    for (int m = 0; m < R0; ++m) {
      if (zw_zero_chunk<L, HK>(m)) continue;
      const int k = l + 64 * m, kk = k > L / 2 ? L - k : k;
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int fi = 0; fi < NF; ++fi)
          if ((use.bits[g0 / G][p] >> fi) & 1u) raw[m][p][fi] = zw_in<HK>(a, zm, fi, pos0 + p, kk);
    }
    // keep the group's loads together: nothing is scheduled across this point
    __builtin_amdgcn_sched_barrier(0);
  };
  double2 x[NCI][R0];
  double2 raw[2][R0][P][NF];
  load_group(0, raw[0]);
#pragma unroll
  for (int g0 = 0; g0 < NCI; g0 += G) {
    const int cur = (g0 / G) & 1;
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      const int q = g0 + gi;
      if (q < NCI) {
        double2 v[R0];
#pragma unroll
        for (int m = 0; m < R0; ++m) {
          const int k = l + 64 * m, kk = k > L / 2 ? L - k : k;
          v[m] = zw_zero_chunk<L, HK>(m) ? make_double2(0.0, 0.0)
                                         : zw_hermitian<L, MODE, P, NRI, NF>(q, k, a.kzscale * kk, raw[cur][m]);
        }
        Dft<R0, +1>::run(v);
        buf[(gi * R0) * SUB + l] = v[0];
#pragma unroll
        for (int av = 1; av < R0; ++av) buf[(gi * R0 + av) * SUB + l] = cmul(v[av], cconj(t1[av]));
      }
    }
    if (g0 + G < NCI) load_group(g0 + G, raw[cur ^ 1]);
    wave_sync();
    zw_sub64<L, +1>(buf, w8, l, NCI - g0 < G ? NCI - g0 : G);
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      const int q = g0 + gi;
      if (q < NCI) {
#pragma unroll
        for (int m = 0; m < R0; ++m) x[q][m] = buf[gi * W::NATS + W::nat(l + 64 * m)];
      }
    }
    wave_sync();
  }

  if constexpr (MODE == Z_PHYS || MODE == Z_GRAD) {
#pragma unroll
    for (int m = 0; m < R0; ++m) {
      const int zz = l + 64 * m;
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const long long pos = pos0 + p;
#pragma unroll
        for (int f = 0; f < NIN; ++f) {
          const int r = f * P + p;
          const double v = ((r & 1) ? x[r >> 1][m].y : x[r >> 1][m].x) * a.out_scale;
          if constexpr (MODE == Z_PHYS) {
            if (f < a.nphys) a.phys[(long long)f * a.out_sf + pos * L + zz] = v;
          }
          else
            a.gout[(((long long)blockIdx.x * P + p) * L + zz) * 9 + f] = (float)v;
        }
      }
    }
    return;
  } else if constexpr (MODE == Z_DT) {
    double mx = 0.0;
#pragma unroll
    for (int m = 0; m < R0; ++m) {
#pragma unroll
      for (int p = 0; p < P; ++p) {
        double v[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          const int r = f * P + p;
          v[f] = (r & 1) ? x[r >> 1][m].y : x[r >> 1][m].x;
        }
        mx = max_nan(mx, fabs(v[0]) * a.dinv[0] + fabs(v[1]) * a.dinv[1] + fabs(v[2]) * a.dinv[2]);
      }
    }
    for (int off = 32; off > 0; off >>= 1) mx = max_nan(mx, __shfl_xor(mx, off));
    if (l == 0) a.red[blockIdx.x] = mx;
    return;
  } else {
    // ---- Z_LRNF: the closure network on this position's own gradients.  They are
    // staged through LDS as float point records of stride 16 (the nine gradients at
    // 0..8, then the six stresses at 9..14), so that 16-point MFMA tiles can be read
    // with the points along N; the stresses go back into the same records.
    [[maybe_unused]] float* rec = nullptr;
    if constexpr (MODE == Z_LRNF) {
      static_assert(P == 1, "fused closure: one position");
      __shared__ float recs[L * 16];
      rec = recs;
#pragma unroll
      for (int m = 0; m < R0; ++m) {
        const int zz = l + 64 * m;
#pragma unroll
        for (int f = 0; f < 9; ++f) {
          const int r = 3 + f;
          recs[zz * 16 + f] = (float)(((r & 1) ? x[r >> 1][m].y : x[r >> 1][m].x) * a.gscale);
        }
      }
      wave_sync();
      const int g = l >> 4, rr = l & 15;
      psx_mlp_impl::Net R;
      psx_mlp_impl::net_load(R, a.npk, a.nw, a.nnorm, l);
      constexpr int TP = 1;  // one 16-point tile at a time (two interleaved needed 256 VGPRs)
#pragma unroll 1
      for (int t0 = 0; t0 < L / 16; t0 += TP) {
        psx_mlp_impl::f4 xn[TP], out[TP];
#pragma unroll
        for (int q = 0; q < TP; ++q) xn[q] = psx_mlp_impl::net_input(R, &recs[((t0 + q) * 16 + rr) * 16], true, g);
        psx_mlp_impl::net_forward<TP>(R, xn, out);
#pragma unroll
        for (int q = 0; q < TP; ++q) {
          float* gp = &recs[((t0 + q) * 16 + rr) * 16];
          const psx_mlp_impl::f4 res =
              psx_mlp_impl::net_closure(R, out[q], xn[q], gp, true, g, a.nc2, a.nbound, a.nclip);
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (4 * g + j < 6) gp[9 + 4 * g + j] = res[j];
        }
      }
      wave_sync();
    }
    // ---- pointwise products / closure on the lane's R0 z points
    double2 xo[NCO][R0];
#pragma unroll
    for (int m = 0; m < R0; ++m) {
      double o[P][NOUT];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        double in[NIN];
#pragma unroll
        for (int f = 0; f < NIN; ++f) {
          const int r = f * P + p;
          in[f] = (r & 1) ? x[r >> 1][m].y : x[r >> 1][m].x;
        }
        if constexpr (MODE == Z_LRNF) {
          // the products, then the stresses joined as Z_LRN does
          zpoint<Z_DNS>(a, in, o[p]);
#pragma unroll
          for (int c = 0; c < 6; ++c) o[p][6 + c] = a.c2 * (double)rec[(l + 64 * m) * 16 + 9 + c];
        } else {
          zpoint<MODE>(a, in, o[p], ((long long)blockIdx.x * P + p) * L + l + 64 * m);
        }
        zpoint_fold<MODE>(a, o[p]);
      }
#pragma unroll
      for (int q = 0; q < NCO; ++q) {
        const int r0 = 2 * q, r1 = 2 * q + 1;
        xo[q][m].x = o[r0 % P][r0 / P];
        xo[q][m].y = r1 < NRO ? o[r1 % P][r1 / P] : 0.0;
      }
    }

    // ---- forward transforms, truncation to Hout modes, split of the pencil pairs
#pragma unroll
    for (int g0 = 0; g0 < NCO; g0 += G) {
#pragma unroll
      for (int gi = 0; gi < G; ++gi) {
        const int q = g0 + gi;
        if (q < NCO) {
          double2 v[R0];
#pragma unroll
          for (int m = 0; m < R0; ++m) v[m] = xo[q][m];
          Dft<R0, -1>::run(v);
          buf[(gi * R0) * SUB + l] = v[0];
#pragma unroll
          for (int av = 1; av < R0; ++av) buf[(gi * R0 + av) * SUB + l] = cmul(v[av], t1[av]);
        }
      }
      wave_sync();
      zw_sub64<L, -1, HK>(buf, w8, l, NCO - g0 < G ? NCO - g0 : G);
#pragma unroll
      for (int gi = 0; gi < G; ++gi) {
        const int q = g0 + gi;
        if (q < NCO) {
          const int r0 = 2 * q, r1 = 2 * q + 1;
          const int Ho = HK > 0 ? HK : a.Hout;
          for (int k = l; k < Ho; k += 64) {
            const double2 z1 = buf[gi * W::NATS + W::nat(k)];
            const double2 z2 = buf[gi * W::NATS + W::nat(k == 0 ? 0 : L - k)];
            store_nt(&zm.out[(long long)(r0 / P) * a.out_sf + (pos0 + r0 % P) * a.Hout + k],
                     make_double2(0.5 * (z1.x + z2.x), 0.5 * (z1.y - z2.y)));
            if (r1 < NRO)
              store_nt(&zm.out[(long long)(r1 / P) * a.out_sf + (pos0 + r1 % P) * a.Hout + k],
                       make_double2(0.5 * (z1.y + z2.y), 0.5 * (z2.x - z1.x)));
          }
        }
      }
      wave_sync();
    }
  }
}

}  // namespace psx
