// Fused z pass for the short z lengths of small grids: L = 3 W with W = 16, 24 or 32
// lanes per position and S = 64 NW / W positions per workgroup of NW wavefronts.
//
// Same contract as zwave (zwave.hpp) with one position per lane group, and the same
// data path scaled down.  Lane l < W of group s owns the entries k = l + W m (m < 3)
// of every complex pencil of position blockIdx.x * S + s, and a length-L transform
// x_n = sum_k Z_k w_L^{e n k} (e = +1 inverse, -1 forward) runs as a radix-3 register
// step and a W-point transform over the group's lanes in two LDS passes, W = A B with
// l = j + A r and b = b1 + B b2:
//   Y_a(l)                = w_L^{e a l} sum_m Z_{l + W m} w_3^{e a m}      (registers)
//   U_a(j, b1)            = w_W^{e j b1} sum_r Y_a(j + A r) w_B^{e b1 r}  (LDS pass 1)
//   x_{a + 3 (b1 + B b2)} = sum_j U_a(j, b1) w_A^{e b2 j}                 (LDS pass 2)
// The block-wide Stockham kernel (zpass) spends most of its time on barriers and idle
// threads at these lengths; one wavefront per S positions needs no workgroup barrier.
// At W = 24 the products modes run three wavefronts per workgroup (eight lane groups,
// no idle lanes) and pay a barrier per LDS pass; the fused closure runs one wavefront,
// whose lanes 48..63 form a third group that computes on the first group's position
// in LDS of its own and stores nothing.
#pragma once
#include "zwave.hpp"

namespace psx {

// NW: wavefronts per workgroup of the products modes
template <int L> struct ZSub { static constexpr bool ok = false; };
template <> struct ZSub<48> { static constexpr bool ok = true; static constexpr int W = 16, A = 4, B = 4, NW = 1; };
template <> struct ZSub<72> { static constexpr bool ok = true; static constexpr int W = 24, A = 8, B = 3, NW = 3; };
template <> struct ZSub<96> { static constexpr bool ok = true; static constexpr int W = 32, A = 8, B = 4, NW = 1; };
// the fused closure keeps one wavefront (its register use leaves room for one
// three-wavefront workgroup per CU)
template <int L, int MODE> constexpr int zs_waves() { return MODE == Z_LRNF ? 1 : ZSub<L>::NW; }

template <int L> struct ZSubShape {
  static constexpr int W = ZSub<L>::W, A = ZSub<L>::A, B = ZSub<L>::B;
  static constexpr int R0 = L / W;                      // 3
  static constexpr int G = 2;                           // complex pencils per round of LDS passes
  static constexpr int SUBW = B * (A + 1);              // sub-transform slot; pass-1 output at b1 (A + 1) + j
  static constexpr int NATS = L + L / 16 + 1;           // natural-order row, one pad entry per 16
  static constexpr int SP = R0 * SUBW > NATS ? R0 * SUBW : NATS;  // LDS entries of one pencil
  static constexpr int SG = G * SP;                     // LDS entries of one group
  static __device__ __forceinline__ int nat(int n) { return n + (n >> 4); }
  static_assert(W <= SUBW, "register-step slots");
};
// the lane groups of a workgroup of NW wavefronts
template <int L, int NW> struct ZSubGroups {
  static constexpr int W = ZSub<L>::W;
  static constexpr int S = 64 * NW / W;                 // positions per workgroup
  static constexpr int NG = (64 * NW + W - 1) / W;      // lane groups (S, or S + 1 idle)
};

// Entries k = l + W m (every l < W) of chunk m all have min(k, L - k) >= HK: the 3/2
// rule's zero pad, neither loaded nor transformed.
template <int L, int W, int HK>
__host__ __device__ constexpr bool zs_zero_chunk(int m) {
  if (HK <= 0) return false;
  for (int l = 0; l < W; ++l) {
    const int k = l + W * m, kk = k > L / 2 ? L - k : k;
    if (kk < HK) return false;
  }
  return true;
}

// Forward transforms store only the outputs n < L/3 and n > 2L/3 (the 3/2 rule's kept
// modes k and L - k, k < L/3); pass 2 writes x_n, n = a + R0 (b1 + B b2), for every a,
// b1 of a given b2 -- that write is skipped when all of those n fall in between.
template <int L, int SIGN>
__host__ __device__ constexpr bool zs_unread(int b2) {
  using Z = ZSubShape<L>;
  const int lo = Z::R0 * Z::B * b2, hi = lo + Z::R0 * Z::B - 1;
  return SIGN < 0 && lo >= L / 3 && hi <= L - L / 3;
}

// The W-point transforms of the R0 sub-transforms of np <= G pencils, pencil g's held
// at sb[g SP + a SUBW + l]; the result x_{a + R0 b} of pencil g lands at
// sb[g SP + nat(a + R0 b)].  Pass-1 twiddles w_W^{j b1} (forward sign) of the lane's
// column j = l mod A, which is the same for every butterfly the lane runs (W is a
// multiple of A): from the registers wj[b1] (WJ), or from the length-L table tw in LDS
// (the fused closure, which has no registers to spare).  The group's W
// lanes share the butterflies of all np pencils in each pass, so a round of passes
// (and its wave syncs) serves two pencils.
template <int L, int SIGN, bool WJ>
__device__ __forceinline__ void zs_subW(double2* sb, const double2 (&wj)[ZSub<L>::B], const double2* tw, int l,
                                        int np) {
  using Z = ZSubShape<L>;
  constexpr int A = Z::A, B = Z::B, R0 = Z::R0, SUBW = Z::SUBW, W = Z::W, SP = Z::SP, G = Z::G;
  // pass 1: butterfly u = (g, a, j), u = (g R0 + a) A + j; a B-point transform over r,
  // twiddle w_W^{e j b1}
  constexpr int N1 = G * R0 * A, IT1 = (N1 + W - 1) / W;
  double2 y[IT1][B];
#pragma unroll
  for (int it = 0; it < IT1; ++it) {
    const int u = l + W * it, ga = u / A, j = u - ga * A;
    if ((N1 % W == 0 || u < N1) && ga < np * R0) {
      const int g = ga / R0, a1 = ga - g * R0;
#pragma unroll
      for (int r = 0; r < B; ++r) y[it][r] = sb[g * SP + a1 * SUBW + j + A * r];
      Dft<B, SIGN>::run(y[it]);
#pragma unroll
      for (int b1 = 1; b1 < B; ++b1) {
        const double2 w = WJ ? wj[b1] : tw[R0 * j * b1];
        y[it][b1] = cmul(y[it][b1], SIGN > 0 ? cconj(w) : w);
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int it = 0; it < IT1; ++it) {
    const int u = l + W * it, ga = u / A, j = u - ga * A;
    if ((N1 % W == 0 || u < N1) && ga < np * R0) {
      const int g = ga / R0, a1 = ga - g * R0;
#pragma unroll
      for (int b1 = 0; b1 < B; ++b1) sb[g * SP + a1 * SUBW + b1 * (A + 1) + j] = y[it][b1];
    }
  }
  wave_sync();
  // pass 2: butterfly u = (g, a, b1), u = (g R0 + a) B + b1; an A-point transform over j
  constexpr int N2 = G * R0 * B, IT2 = (N2 + W - 1) / W;
  double2 z[IT2][A];
#pragma unroll
  for (int it = 0; it < IT2; ++it) {
    const int u = l + W * it, ga = u / B, b1 = u - ga * B;
    if ((N2 % W == 0 || u < N2) && ga < np * R0) {
      const int g = ga / R0, a2 = ga - g * R0;
#pragma unroll
      for (int jj = 0; jj < A; ++jj) z[it][jj] = sb[g * SP + a2 * SUBW + b1 * (A + 1) + jj];
      Dft<A, SIGN>::run(z[it]);
    }
  }
  wave_sync();
#pragma unroll
  for (int it = 0; it < IT2; ++it) {
    const int u = l + W * it, ga = u / B, b1 = u - ga * B;
    if ((N2 % W == 0 || u < N2) && ga < np * R0) {
      const int g = ga / R0, a2 = ga - g * R0;
#pragma unroll
      for (int b2 = 0; b2 < A; ++b2)
        if (!zs_unread<L, SIGN>(b2)) sb[g * SP + Z::nat(a2 + R0 * (b1 + B * b2))] = z[it][b2];
    }
  }
  wave_sync();
}

// HK: the input and output half spectra hold HK = L / 3 entries (the launcher checks).
template <int L, int MODE, int HK>
__global__ __launch_bounds__((64 * zs_waves<L, MODE>())) void zsub(ZArgs a, const double2* __restrict__ tw) {
  using Z = ZSubShape<L>;
  static_assert(ZSub<L>::ok, "zsub: L must be 48, 72 or 96");
  static_assert(MODE == Z_DNS || MODE == Z_SMAG || MODE == Z_LRNF, "zsub: product modes only");
  constexpr int NW = zs_waves<L, MODE>(), NTB = 64 * NW;
  using ZG = ZSubGroups<L, NW>;
  constexpr int W = Z::W, R0 = Z::R0, S = ZG::S, SUBW = Z::SUBW, SG = Z::SG;
  constexpr int NIN = ZShape<MODE>::nin, NOUT = ZShape<MODE>::nout;
  constexpr int NRI = NIN, NRO = NOUT;
  constexpr int NCI = (NRI + 1) / 2, NCO = (NRO + 1) / 2;
  constexpr int NF = ZFields<MODE>::n;
  __shared__ double2 buf[ZG::NG * SG];
  constexpr bool WJ = MODE != Z_LRNF;
  __shared__ double2 twl[WJ ? 1 : L];
  const int lane = threadIdx.x;
  const int s = lane / W, l = lane - (lane / W) * W;
  const bool act = s < S;
  double2* const sb = buf + s * SG;
  const ZMember zm = zmember(a, (long long)blockIdx.x * S + (act ? s : 0));
  const long long pos0 = zm.pos0;

  // register-step twiddles w_L^{-a l} and the W-point transforms' pass-1 twiddles
  // w_L^{-R0 j b1} = w_W^{-j b1}, j = l mod A (the inverse uses the conjugates)
  double2 t1[R0], wj[Z::B];
#pragma unroll
  for (int av = 0; av < R0; ++av) t1[av] = tw[av * l];
#pragma unroll
  for (int b1 = 0; b1 < Z::B; ++b1) wj[b1] = WJ ? tw[R0 * (l % Z::A) * b1] : make_double2(1.0, 0.0);
  static_assert(W % Z::A == 0 && R0 * (Z::A - 1) * (Z::B - 1) < L, "pass-1 twiddles");
  if constexpr (!WJ)
    for (int i = lane; i < L; i += NTB) twl[i] = tw[i];

  // ---- inverse transforms, G pencils per round; lane l keeps x_n, n = l + W m, of
  // every pencil
  constexpr int G = Z::G, SP = Z::SP;
  constexpr ZGroupUse<MODE, 1, NRI, G> use{};
  // every raw entry of a round's pencils, loaded before any is used; the next round's
  // loads are issued before this round's LDS passes
  auto load_round = [&](int q0, double2 (&raw)[R0][1][NF]) {
#pragma unroll
    for (int m = 0; m < R0; ++m) {
      if (zs_zero_chunk<L, W, HK>(m)) continue;
      const int k = l + W * m, kk = k > L / 2 ? L - k : k;
#pragma unroll
      for (int fi = 0; fi < NF; ++fi)
        if ((use.bits[q0 / G][0] >> fi) & 1u) raw[m][0][fi] = zw_in<HK>(a, zm, fi, pos0, kk);
    }
    __builtin_amdgcn_sched_barrier(0);
  };
  double2 x[NCI][R0];
  double2 raw[2][R0][1][NF];
  load_round(0, raw[0]);
  if constexpr (!WJ) wave_sync();  // the LDS twiddle table
#pragma unroll
  for (int q0 = 0; q0 < NCI; q0 += G) {
    const int cur = (q0 / G) & 1;
    const int np = NCI - q0 < G ? NCI - q0 : G;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int q = q0 + g;
      if (q < NCI) {
        double2 v[R0];
#pragma unroll
        for (int m = 0; m < R0; ++m) {
          const int k = l + W * m, kk = k > L / 2 ? L - k : k;
          v[m] = zs_zero_chunk<L, W, HK>(m) ? make_double2(0.0, 0.0)
                                            : zw_hermitian<L, MODE, 1, NRI, NF>(q, k, a.kzscale * kk, raw[cur][m]);
        }
        Dft<R0, +1>::run(v);
        sb[g * SP + l] = v[0];
#pragma unroll
        for (int av = 1; av < R0; ++av) sb[g * SP + av * SUBW + l] = cmul(v[av], cconj(t1[av]));
      }
    }
    if (q0 + G < NCI) load_round(q0 + G, raw[cur ^ 1]);
    wave_sync();
    zs_subW<L, +1, WJ>(sb, wj, twl, l, np);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int q = q0 + g;
      if (q < NCI) {
#pragma unroll
        for (int m = 0; m < R0; ++m) x[q][m] = sb[g * SP + Z::nat(l + W * m)];
      }
    }
    wave_sync();
  }

  // ---- Z_LRNF: the closure network on the positions' own gradients, as zwave runs
  // it: float point records of stride 16 in LDS (the nine gradients at 0..8, the six
  // stresses written back at 9..14), point p = s L + z of the wavefront's S positions,
  // 16-point MFMA tiles over all 64 lanes
  [[maybe_unused]] float* rec = nullptr;
  if constexpr (MODE == Z_LRNF) {
    static_assert(NW == 1, "the fused closure runs one wavefront per workgroup");
    __shared__ float recs[ZG::NG * L * 16];
    rec = recs + s * L * 16;
#pragma unroll
    for (int m = 0; m < R0; ++m) {
      const int zz = l + W * m;
#pragma unroll
      for (int f = 0; f < 9; ++f) {
        const int r = 3 + f;
        rec[zz * 16 + f] = (float)(((r & 1) ? x[r >> 1][m].y : x[r >> 1][m].x) * a.gscale);
      }
    }
    wave_sync();
    const int g = lane >> 4, rr = lane & 15;
    psx_mlp_impl::Net R;
    psx_mlp_impl::net_load(R, a.npk, a.nw, a.nnorm, lane);
    static_assert((S * L) % 16 == 0, "whole 16-point tiles");
#pragma unroll 1
    for (int t0 = 0; t0 < S * L / 16; ++t0) {
      float* gp = &recs[(t0 * 16 + rr) * 16];
      psx_mlp_impl::f4 xn[1], out[1];
      xn[0] = psx_mlp_impl::net_input(R, gp, true, g);
      psx_mlp_impl::net_forward<1>(R, xn, out);
      const psx_mlp_impl::f4 res = psx_mlp_impl::net_closure(R, out[0], xn[0], gp, true, g, a.nc2, a.nbound, a.nclip);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * g + j < 6) gp[9 + 4 * g + j] = res[j];
    }
    wave_sync();
  }

  // ---- pointwise products / closure on the lane's R0 z points
  double2 xo[NCO][R0];
#pragma unroll
  for (int m = 0; m < R0; ++m) {
    double in[NIN], o[NOUT];
#pragma unroll
    for (int f = 0; f < NIN; ++f) in[f] = (f & 1) ? x[f >> 1][m].y : x[f >> 1][m].x;
    if constexpr (MODE == Z_LRNF) {
      // the products, then the stresses joined as zwave's fused pass does
      zpoint<Z_DNS>(a, in, o);
#pragma unroll
      for (int c = 0; c < 6; ++c) o[6 + c] = a.c2 * (double)rec[(l + W * m) * 16 + 9 + c];
    } else {
      zpoint<MODE>(a, in, o);
    }
    zpoint_fold<MODE>(a, o);
#pragma unroll
    for (int q = 0; q < NCO; ++q) {
      xo[q][m].x = o[2 * q];
      xo[q][m].y = 2 * q + 1 < NRO ? o[2 * q + 1] : 0.0;
    }
  }

  // ---- forward transforms, G pencils per round, truncation to HK modes, split of the
  // pencil pairs
#pragma unroll
  for (int q0 = 0; q0 < NCO; q0 += G) {
    const int np = NCO - q0 < G ? NCO - q0 : G;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int q = q0 + g;
      if (q < NCO) {
        double2 v[R0];
#pragma unroll
        for (int m = 0; m < R0; ++m) v[m] = xo[q][m];
        Dft<R0, -1>::run(v);
        sb[g * SP + l] = v[0];
#pragma unroll
        for (int av = 1; av < R0; ++av) sb[g * SP + av * SUBW + l] = cmul(v[av], t1[av]);
      }
    }
    wave_sync();
    zs_subW<L, -1, WJ>(sb, wj, twl, l, np);
    if (act) {
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int q = q0 + g;
        if (q < NCO) {
          const int r0 = 2 * q, r1 = 2 * q + 1;
#pragma unroll
          for (int k = l; k < HK; k += W) {
            const double2 z1 = sb[g * SP + Z::nat(k)];
            const double2 z2 = sb[g * SP + Z::nat(k == 0 ? 0 : L - k)];
            store_nt(&zm.out[(long long)r0 * a.out_sf + pos0 * a.Hout + k],
                     make_double2(0.5 * (z1.x + z2.x), 0.5 * (z1.y - z2.y)));
            if (r1 < NRO)
              store_nt(&zm.out[(long long)r1 * a.out_sf + pos0 * a.Hout + k],
                       make_double2(0.5 * (z1.y + z2.y), 0.5 * (z2.x - z1.x)));
          }
        }
      }
    }
    wave_sync();
  }
}

}  // namespace psx
