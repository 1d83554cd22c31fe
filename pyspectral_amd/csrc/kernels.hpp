// Pseudo-spectral pipeline kernels for gfx950.
//
// Layouts (complex128 = double2, kz contiguous everywhere):
//   state  S[c][k1][k2][k3]       (3, N1, N2, N3/2+1)  -- the reference's per-field layout
//   X      X[f][x ][k2][k3<H]     after the x-inverse pass (x padded to M1)
//   Y      Y[f][x ][y ][k3<H]     after the y-inverse pass (y padded to M2)
//   Z      Z[f][x ][y ][k3<Ho]    after the fused z pass (products transformed back in z)
//   W      W[f][x ][k2][k3<Ho]    after the y-forward pass (truncated in k2)
// The x-forward pass is fused with the spectral epilogue (projection, closures,
// viscous term, RK stage update) and writes the state layout.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_core.hpp"
#include "mlp_core.hpp"

namespace psx {

// in-row LDS index of a kernel templated on transform length L
#define RW(i) (::psx::Row<L>::at(i))

constexpr int NT = 256;       // threads per block

// Twiddle tables are staged in LDS per block.  A table read from global memory inside
// a transform waits on the vector memory counter, which retires loads and stores in
// order, so it would also wait for loads the block has in flight for its next field
// (the epilogue's prefetch, the second input of a pencil job) and for the stores of its
// previous transform; reads from LDS wait only on the LDS counter.
#define PSX_STAGE_TW(NTH)                                              \
  __shared__ double2 tw_lds[L];                                        \
  for (int i_ = threadIdx.x; i_ < L; i_ += (NTH)) tw_lds[i_] = tw[i_]; \
  const double2* twp = tw_lds;
constexpr int MAXJOBS = 16;   // input fields per pencil pass
constexpr int MAXREC = 12;    // real pencils built by the z pass

// ------------------------------------------------------------------ small selects
// A select of one of three values, written as selects so no per-thread array is
// indexed at run time (which would place it in scratch memory).
__device__ __forceinline__ double sel3(int i, double a, double b, double c) { return i == 0 ? a : (i == 1 ? b : c); }

__device__ __forceinline__ void cfma(double2& x, double w, double2 v) { x = make_double2(x.x + w * v.x, x.y + w * v.y); }
// The k-weighted sums V_i = sum_j X_ij k_j of a symmetric tensor, one component at a
// time: X_mn (g = 0..5: mn = 11 22 33 12 13 23) adds k_n v to V_m and, off the diagonal,
// X_nm = X_mn adds k_m v to V_n.  g is the same for the whole block, so the branch is
// taken by the scalar unit and only those one or two complex multiply-adds run.
__device__ __forceinline__ void sym_accumulate(int g, const double* k, double2 v, double2* V) {
  switch (g) {
    case 0: cfma(V[0], k[0], v); break;
    case 1: cfma(V[1], k[1], v); break;
    case 2: cfma(V[2], k[2], v); break;
    case 3: cfma(V[0], k[1], v); cfma(V[1], k[0], v); break;
    case 4: cfma(V[0], k[2], v); cfma(V[2], k[0], v); break;
    default: cfma(V[1], k[2], v); cfma(V[2], k[1], v); break;
  }
}

// Maximum that propagates NaN, as numpy's amax does (fmax drops a NaN operand, so a
// NaN state would give a finite CFL limit where src/Classes.py:693 gives NaN).
__device__ __forceinline__ double max_nan(double m, double v) { return (v > m || v != v) ? v : m; }

// A copy of v the compiler cannot see through.  Index arithmetic derived from it
// inside a loop is recomputed there instead of being hoisted out and held in
// registers for the whole loop (the LDS addresses of every transform stage), which
// costs occupancy.  Measured: it pays in the PL pass (188 -> 166 VGPRs, 413 -> 397 us
// at 256 x 4); in the 3/2-rule pencil passes and the epilogue the recomputation
// cost more than the extra waves gained.
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// ------------------------------------------------------------------ index maps
// Signed frequency of FFT-ordered spectral index i on an n-point axis; index n/2
// is the -n/2 "oddball" mode, as in the reference's wavenumber arrays
// (src/Classes.py:565-566).
__device__ __forceinline__ int signed_freq(int i, int n) { return i < n / 2 ? i : i - n; }

// Slot of spectral index i (of n) in an L-point transform, or -1 when the index
// is dropped.  keep_nyq = 0 drops index n/2 (3/2 rule, src/padding.py:8-24);
// keep_nyq = 1 keeps it at L - n/2 (2x rule, src/padding.py:34-48) and, for L == n,
// gives the identity.
__device__ __forceinline__ int pad_slot(int i, int n, int L, int keep_nyq) {
  if (i < n / 2) return i;
  if (i == n / 2 && !keep_nyq) return -1;
  return L + i - n;
}

// ------------------------------------------------------------------ block order
// Workgroups are dispatched to the 8 XCDs round-robin (id mod 8), each XCD with its
// own L2.  xcd_block gives every XCD a contiguous range of logical block ids, so
// neighbouring column tiles -- which share the cache lines of rows that are not
// 128-byte aligned, like the state's (N3/2+1)-entry rows -- run on one L2.
// Measured: it helps the pencil passes that read the state (the CFL passes most).
// The x-forward epilogue uses its own order: one k2 row's column tiles per XCD with
// rows round-robin (traffic of the shared lines stays in one L2; measured 126 against
// 131 us in dispatch order and 255 us with contiguous ranges).
template <bool ON>
__device__ __forceinline__ int xcd_block() {
  if constexpr (ON) {
    const int w = blockIdx.x, G = gridDim.x;
    const int q = G >> 3, r = G & 7, x = w & 7;
    return x * q + (x < r ? x : r) + (w >> 3);
  } else {
    return blockIdx.x;
  }
}

// Logical block of a pencil pass of A outer rows with ct column tiles each, by block
// order: 0 = contiguous per-XCD ranges (xcd_block), 1 = dispatch order, 2 = the ct
// tiles of one outer row on one XCD, consecutive rows round-robin over the 8 XCDs (the
// epilogue's order; it needs A % 8 == 0, else order 0 is used).  With gridDim.x = A ct
// a multiple of 8, block b runs on XCD b & 7, which takes rows x, x + 8, ... in turn.
__device__ __forceinline__ int pencil_block(int order, int ct, int A) {
  if (order == 1) return blockIdx.x;
  if (order == 2 && (A & 7) == 0) {
    const int b = blockIdx.x, s = b >> 3;
    return (8 * (s / ct) + (b & 7)) * ct + s % ct;
  }
  return xcd_block<true>();
}

// ------------------------------------------------------------------ pencil passes
struct PencilArgs {
  const double2* in;
  double2* out;
  long long in_sf, in_sa, in_sl;     // strides (complex units): field, outer, transform axis
  long long out_sf, out_sa, out_sl;
  int A;         // outer extent
  int C;         // contiguous (k3) extent
  int Nspec;     // spectral entries along the transform axis
  int keep_nyq;  // see pad_slot
  double kscale; // 2*pi/L_axis, the derivative factor
  double scale;  // applied to every output
  int job_in[MAXJOBS];
  int job_plain[MAXJOBS];
  int job_deriv[MAXJOBS];
  // inverse passes: the derivative output is dcoef * (i k in + in2), in2 = field
  // job_in2 of the same input array (-1: none) -- e.g. (u_y + v_x) / 2 in one transform
  int job_in2[MAXJOBS];
  double job_dcoef[MAXJOBS];
  // ensembles: grid.y = jpm * members; member m = blockIdx.y / jpm reads at
  // in + m*in_ms and writes at out + m*out_ms (jpm = 0: one member)
  int jpm;
  long long in_ms, out_ms;
  // block order (pencil_block): 0 gives each XCD a contiguous range of blocks, 1 keeps
  // the dispatch order, 2 keeps each outer row's column tiles on one XCD (which is
  // fastest depends on the pass and the job count)
  int order;
};

// (member, job) of a pencil block
__device__ __forceinline__ int pencil_member(const PencilArgs& a) { return a.jpm ? (int)blockIdx.y / a.jpm : 0; }

// Inverse along a strided axis: Nspec spectral entries -> L physical points, with
// the zero-pad fused in; optionally a second output multiplied by i*k_axis.
// grid = (A * ceil(C/T), njobs)
template <int L, int T>
__global__ __launch_bounds__(NT) void pencil_inverse(PencilArgs a, const double2* __restrict__ tw) {
  constexpr int LS = Row<L>::LS;
  __shared__ double2 buf[T * LS];
  const int tid = threadIdx.x;
  PSX_STAGE_TW(NT)
  const int ct = (a.C + T - 1) / T;
  const int bid = pencil_block(a.order, ct, a.A);
  const int aidx = bid / ct;
  const int c0 = (bid - aidx * ct) * T;
  const int mem = pencil_member(a);
  const int job = blockIdx.y - mem * a.jpm;
  const double2* src = a.in + mem * a.in_ms + (long long)a.job_in[job] * a.in_sf + (long long)aidx * a.in_sa + c0;
  const int in2 = a.job_in2[job];
  const double2* src2 = a.in + mem * a.in_ms + (long long)(in2 < 0 ? 0 : in2) * a.in_sf + (long long)aidx * a.in_sa + c0;
  constexpr int IT = (L * T + NT - 1) / NT;
  double2 r[IT];
  const int nload = a.Nspec * T;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + it * NT;
    r[it] = make_double2(0.0, 0.0);
    if (e < nload) {
      const int l = e / T, c = e - (e / T) * T;
      if (c0 + c < a.C) r[it] = src[(long long)l * a.in_sl + c];
    }
  }
  for (int var = 0; var < 2; ++var) {
    const int fo = var == 0 ? a.job_plain[job] : a.job_deriv[job];
    if (fo < 0) continue;
    for (int e = tid; e < T * L; e += NT) buf[(e / L) * LS + RW((e % L))] = make_double2(0.0, 0.0);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      if (e < nload) {
        const int l = e / T, c = e - (e / T) * T;
        const int p = pad_slot(l, a.Nspec, L, a.keep_nyq);
        if (p >= 0) {
          double2 v = r[it];
          if (var == 1) v = cscale(cmul_i<1>(v), a.kscale * signed_freq(l, a.Nspec) * a.job_dcoef[job]);
          buf[c * LS + RW(p)] = v;
        }
      }
    }
    if (var == 1 && in2 >= 0) {
      // the second input, added by the thread that staged each slot (no barrier), after
      // the first input's registers are dead
      const double dc = a.job_dcoef[job];
#pragma unroll 3
      for (int it = 0; it < IT; ++it) {
        const int e = tid + it * NT;
        if (e < nload) {
          const int l = e / T, c = e - (e / T) * T;
          const int p = pad_slot(l, a.Nspec, L, a.keep_nyq);
          if (p >= 0 && c0 + c < a.C) buf[c * LS + RW(p)] = cadd(buf[c * LS + RW(p)], cscale(src2[(long long)l * a.in_sl + c], dc));
        }
      }
    }
    __syncthreads();
    fft_batch<L, T, NT, LS, +1>(buf, twp, tid);
    double2* dst = a.out + mem * a.out_ms + (long long)fo * a.out_sf + (long long)aidx * a.out_sa + c0;
    for (int e = tid; e < L * T; e += NT) {
      const int p = e / T, c = e - (e / T) * T;
      if (c0 + c < a.C) dst[(long long)p * a.out_sl + c] = cscale(buf[c * LS + RW(p)], a.scale);
    }
    __syncthreads();
  }
}

// A block's L x T tile of a strided pencil pass, loaded into registers as one batch
// of IT loads per thread (element e = l T + c: row l, column c; columns at or beyond
// ncol read as zero), then staged into the LDS rows (column c at buf[c LS ..]).
// Issuing every load before the first use keeps IT requests in flight per thread,
// where a load-then-store loop would wait for each one in turn.
template <int L, int T> struct PencilLoad {
  static constexpr int IT = (L * T + NT - 1) / NT;
  static __device__ __forceinline__ void load(double2* r, const double2* src, long long sl, int ncol, int tid) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      const int l = e / T, c = e - (e / T) * T;
      r[it] = make_double2(0.0, 0.0);
      if (((L * T) % NT == 0 || e < L * T) && c < ncol) r[it] = src[(long long)l * sl + c];
    }
  }
  static __device__ __forceinline__ void stage(const double2* r, double2* buf, int tid) {
    constexpr int LS = Row<L>::LS;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      const int l = e / T, c = e - (e / T) * T;
      if ((L * T) % NT == 0 || e < L * T) buf[c * LS + RW(l)] = r[it];
    }
  }
};

// Forward along a strided axis: L physical points -> Nspec spectral entries, with
// the truncation fused in.  grid = (A * ceil(C/T), njobs)
template <int L, int T>
__global__ __launch_bounds__(NT) void pencil_forward(PencilArgs a, const double2* __restrict__ tw) {
  constexpr int LS = Row<L>::LS;
  __shared__ double2 buf[T * LS];
  const int tid = threadIdx.x;
  PSX_STAGE_TW(NT)
  const int ct = (a.C + T - 1) / T;
  const int bid = pencil_block(a.order, ct, a.A);
  const int aidx = bid / ct;
  const int c0 = (bid - aidx * ct) * T;
  const int mem = pencil_member(a);
  const int job = blockIdx.y - mem * a.jpm;
  const double2* src = a.in + mem * a.in_ms + (long long)a.job_in[job] * a.in_sf + (long long)aidx * a.in_sa + c0;
  double2 r[PencilLoad<L, T>::IT];
  PencilLoad<L, T>::load(r, src, a.in_sl, a.C - c0, tid);
  const int fo = a.job_plain[job];
  double2* dst = a.out + mem * a.out_ms + (long long)fo * a.out_sf + (long long)aidx * a.out_sa + c0;
  // eight entries per thread (2x lengths 256, 512 at 8 and 4 columns): the loaded
  // entries n = t / T + (L / 8) e feed a radix-8 first stage and the radix-8 last
  // stage's outputs p = t / T + (L / 8) e go straight to their output rows (in-place
  // middle stages: the one LDS buffer keeps the block's occupancy)
  if constexpr (RegPlan<L, 8>::ok && L * T == 8 * NT) {
    fft_rows_reg<L, 8, T, NT, LS, -1, true>(r, buf, buf, twp, tid);
    const int b = tid % T, jj = tid / T;
    if (c0 + b < a.C) {
      const int half = a.Nspec / 2, gap = L - a.Nspec;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        // the output row whose slot is p (pad_slot's inverse), or -1 for a dropped slot
        const int p = jj + (L / 8) * e;
        const int i = p < half ? p : (p >= L - half ? p - gap : -1);
        if (i >= 0)
          dst[(long long)i * a.out_sl + b] =
              pad_slot(i, a.Nspec, L, a.keep_nyq) == p ? cscale(r[e], a.scale) : make_double2(0.0, 0.0);
      }
    }
    return;
  }
  PencilLoad<L, T>::stage(r, buf, tid);
  __syncthreads();
  fft_batch<L, T, NT, LS, -1>(buf, twp, tid);
  for (int e = tid; e < a.Nspec * T; e += NT) {
    const int i = e / T, c = e - (e / T) * T;
    if (c0 + c < a.C) {
      const int p = pad_slot(i, a.Nspec, L, a.keep_nyq);
      dst[(long long)i * a.out_sl + c] = p >= 0 ? cscale(buf[c * LS + RW(p)], a.scale) : make_double2(0.0, 0.0);
    }
  }
}

// Forward along a strided axis where an output field may be a weighted sum of up to
// three transformed inputs, the weight of term t at output index i (signed frequency
// s2 = signed_freq(i, Nspec)) and column c (k3 = ks3 (c0 + c)) being
// coef_t (ks2 s2)^p2_t k3^p3_t.  The Smagorinsky-form y-forward uses it to hand the
// epilogue A = tau_11, B = 2 k2 tau_12 + 2 k3 tau_13 and C = k2^2 tau_22 +
// 2 k2 k3 tau_23 + k3^2 tau_33 (3 fields) instead of the 6 stresses.  The transform is
// linear, so each term is transformed in turn and summed in registers at the output.
constexpr int MAXTERMS = 3;
struct SumArgs {
  PencilArgs p;                      // strides, extents, job_plain (output field), scale
  int nterm[MAXJOBS];
  int tf[MAXJOBS][MAXTERMS];         // input field of each term
  double tc[MAXJOBS][MAXTERMS];      // coefficient
  int tp2[MAXJOBS][MAXTERMS];        // power of k2 (0..2)
  int tp3[MAXJOBS][MAXTERMS];        // power of k3 (0..2)
  double ks2, ks3;                   // 2 pi / L2, 2 pi / L3
};

__device__ __forceinline__ double ipow2(double x, int p) { return p == 0 ? 1.0 : (p == 1 ? x : x * x); }

template <int L, int T>
__global__ __launch_bounds__(NT) void pencil_forward_sum(SumArgs sa, const double2* __restrict__ tw) {
  const PencilArgs& a = sa.p;
  // the 3/2 rule's truncation: Nspec = 2L/3 spectral outputs (the launcher checks)
  constexpr int NS = 2 * L / 3;
  constexpr int LS = Row<L>::LS;
  __shared__ double2 buf[T * LS];
  const int tid = threadIdx.x;
  PSX_STAGE_TW(NT)
  const int ct = (a.C + T - 1) / T;
  const int bid = pencil_block(a.order, ct, a.A);
  const int aidx = bid / ct;
  const int c0 = (bid - aidx * ct) * T;
  const int mem = pencil_member(a);
  const int job = blockIdx.y - mem * a.jpm;
  const double2* base = a.in + mem * a.in_ms + (long long)aidx * a.in_sa + c0;
  constexpr int IO = (NS * T + NT - 1) / NT;  // the outputs a thread stores
  double2 acc[IO];
#pragma unroll
  for (int it = 0; it < IO; ++it) acc[it] = make_double2(0.0, 0.0);
  const int nt = sa.nterm[job];
  // term t+1's pencils are loaded into registers while term t is transformed
  double2 r[PencilLoad<L, T>::IT];
  PencilLoad<L, T>::load(r, base + (long long)sa.tf[job][0] * a.in_sf, a.in_sl, a.C - c0, tid);
#pragma unroll 1
  for (int t = 0; t < nt; ++t) {
    if (t > 0) __syncthreads();  // every thread has read the previous term's results
    PencilLoad<L, T>::stage(r, buf, tid);
    __syncthreads();
    if (t + 1 < nt) PencilLoad<L, T>::load(r, base + (long long)sa.tf[job][t + 1] * a.in_sf, a.in_sl, a.C - c0, tid);
    fft_batch<L, T, NT, LS, -1>(buf, twp, tid);
    const double coef = sa.tc[job][t];
    const int p2 = sa.tp2[job][t], p3 = sa.tp3[job][t];
#pragma unroll
    for (int it = 0; it < IO; ++it) {
      const int e = tid + it * NT;
      if (e < NS * T) {
        const int i = e / T, c = e - (e / T) * T;
        const int p = pad_slot(i, NS, L, a.keep_nyq);
        if (p >= 0) {
          // recomputed per term (an empty asm keeps the factors from being hoisted and
          // held across the transform)
          double k2 = sa.ks2 * signed_freq(i, NS), k3 = sa.ks3 * (c0 + c);
          asm volatile("" : "+v"(k2), "+v"(k3));
          const double w = coef * ipow2(k2, p2) * ipow2(k3, p3);
          acc[it] = cadd(acc[it], cscale(buf[c * LS + RW(p)], w));
        }
      }
    }
  }
  const int fo = a.job_plain[job];
  double2* dst = a.out + mem * a.out_ms + (long long)fo * a.out_sf + (long long)aidx * a.out_sa + c0;
#pragma unroll
  for (int it = 0; it < IO; ++it) {
    const int e = tid + it * NT;
    if (e < NS * T) {
      const int i = e / T, c = e - (e / T) * T;
      if (c0 + c < a.C) dst[(long long)i * a.out_sl + c] = cscale(acc[it], a.scale);
    }
  }
}

// ------------------------------------------------------------------ fused z pass
enum ZMode { Z_DNS = 0, Z_SMAG = 1, Z_TM2 = 2, Z_DT = 3, Z_SGS = 4, Z_PHYS = 5, Z_GRAD = 6, Z_LRN = 7, Z_LRNF = 8 };

template <int MODE> struct ZShape;
template <> struct ZShape<Z_DNS> { static constexpr int nin = 3, nout = 6; };
template <> struct ZShape<Z_SMAG> { static constexpr int nin = 9, nout = 12; };
template <> struct ZShape<Z_TM2> { static constexpr int nin = 6, nout = 9; };
template <> struct ZShape<Z_DT> { static constexpr int nin = 3, nout = 0; };
template <> struct ZShape<Z_SGS> { static constexpr int nin = 6, nout = 12; };
// physical values of 3 fields; the 9 velocity gradients; products + a given stress
template <> struct ZShape<Z_PHYS> { static constexpr int nin = 3, nout = 0; };
template <> struct ZShape<Z_GRAD> { static constexpr int nin = 9, nout = 0; };
template <> struct ZShape<Z_LRN> { static constexpr int nin = 3, nout = 12; };
// the learned closure fused: velocity and its nine gradients in, products + stress out
template <> struct ZShape<Z_LRNF> { static constexpr int nin = 12, nout = 12; };

// Input fields of each mode (the pencils above are built from them by recipes).
// Smagorinsky: y-pass fields 0 u, 1 v, 2 w, 3 u_x, 4 v_y, 5 S12 = (u_y + v_x)/2, 6 w_x,
// 7 w_y; pencils u v w S11 S22 S33 S12 S13 S23 (src/RHSfunctions.py:174-179).  Gradient
// modes: 0 u, 1 v, 2 w, 3 u_x, 4 v_x, 5 w_x, 6 u_y, 7 v_y, 8 w_y.  Every other mode's
// pencil f is input field f.
template <int MODE> struct ZFields {
  static constexpr int n = MODE == Z_SMAG ? 8
                                          : ((MODE == Z_GRAD || MODE == Z_LRNF) ? 9
                                                                                : ((MODE == Z_TM2 || MODE == Z_SGS) ? 6 : 3));
};

// A real z-pencil's half spectrum is built as sum_t coef_t (i k3)^{dz_t} field_{f_t}.
struct Recipe {
  int nterm;
  int f[2];
  double coef[2];
  int dz[2];
};

struct ZArgs {
  const double2* in0;     // input fields f < nin0
  const double2* in1;     // input fields f >= nin0
  long long in_sf0, in_sf1;
  int Hin0, Hin1, nin0;
  double2* out;
  long long out_sf;
  int Hout;
  int npos;               // physical (x, y) positions
  Recipe rec[MAXREC];
  double kzscale;         // 2*pi/L3
  double out_scale;       // multiplies every product before the forward transform
  double c2;              // (Cs*Delta)^2 for the Smagorinsky stress
  double dinv[3];         // 1/dx, 1/dy, 1/dz (CFL reduction)
  double* red;            // CFL: per-block maxima, reduced by cfl_finalize
  // ensembles: npos_m positions per member; member m's fields start at in0 + m*ms_in0,
  // in1 + m*ms_in1, out + m*ms_out (npos_m = 0: one member)
  int npos_m;
  long long ms_in0, ms_in1, ms_out;
  // Z_PHYS: real output phys[f*out_sf + pos*L + z] (scaled by out_scale; one member);
  // Z_GRAD: the 9 gradients as float, g[(gpos*L + z)*9 + c] (scaled by out_scale);
  // Z_LRN: the stress tau[(gpos*L + z)*6 + c] times c2 joins the products.  gpos is the
  // position over all members, so the point arrays of the members follow one another.
  double* phys;
  float* gout;
  const float* tin;
  int nphys;  // Z_PHYS: fields written (<= 3)
  // Z_LRNF: the closure network (mlp_core.hpp) evaluated inside the pass on the
  // gradients times gscale; c2 then scales its stresses into the products' units as in
  // Z_LRN, and nc2 / nbound / nclip are the Smagorinsky fallback's (Cs Delta)^2, the
  // input range bound and the backscatter clip
  const short* npk;
  const float* nw;
  psx_mlp_impl::Norm nnorm;
  double gscale;
  float nc2, nbound;
  int nclip;
  // Smagorinsky-form modes: 1 writes D_mn = N_mn + tau_mn in place of the products
  // (fields 0..5), for the folded epilogue E_SMAGF
  int fold;
};

// A z-pass block's member and its data pointers
struct ZMember {
  const double2* in0;
  const double2* in1;
  double2* out;
  long long pos0;  // first position of the block within its member
};
__device__ __forceinline__ ZMember zmember(const ZArgs& a, long long gpos) {
  const long long m = a.npos_m ? gpos / a.npos_m : 0;
  return ZMember{a.in0 + m * a.ms_in0, a.in1 ? a.in1 + m * a.ms_in1 : nullptr, a.out ? a.out + m * a.ms_out : nullptr,
                 gpos - m * a.npos_m};
}

__device__ __forceinline__ double2 zin(const ZArgs& a, const ZMember& zm, int f, long long pos, int k) {
  if (f < a.nin0) {
    return k < a.Hin0 ? zm.in0[(long long)f * a.in_sf0 + pos * a.Hin0 + k] : make_double2(0.0, 0.0);
  }
  const int g = f - a.nin0;
  return k < a.Hin1 ? zm.in1[(long long)g * a.in_sf1 + pos * a.Hin1 + k] : make_double2(0.0, 0.0);
}

__device__ __forceinline__ double2 zspec(const ZArgs& a, const ZMember& zm, int r, long long pos, int k) {
  const Recipe& R = a.rec[r];
  double2 acc = make_double2(0.0, 0.0);
  for (int t = 0; t < R.nterm; ++t) {
    double2 v = zin(a, zm, R.f[t], pos, k);
    double s = R.coef[t];
    if (R.dz[t]) {
      v = cmul_i<1>(v);
      s *= a.kzscale * k;
    }
    acc.x += s * v.x;
    acc.y += s * v.y;
  }
  return acc;
}

template <int MODE>
__device__ __forceinline__ void zpoint(const ZArgs& a, const double* in, double* out, long long pt = 0) {
  if constexpr (MODE == Z_DNS || MODE == Z_LRN) {
    const double u = in[0], v = in[1], w = in[2], s = a.out_scale;
    out[0] = s * (u * u);
    out[1] = s * (v * v);
    out[2] = s * (w * w);
    out[3] = s * (u * v);
    out[4] = s * (u * w);
    out[5] = s * (v * w);
    if constexpr (MODE == Z_LRN) {
#pragma unroll
      for (int c = 0; c < 6; ++c) out[6 + c] = a.c2 * (double)a.tin[pt * 6 + c];
    }
  } else if constexpr (MODE == Z_SMAG) {
    // src/RHSfunctions.py:195-212: |S| = sqrt(2 S_ij S_ij), nu_t = (Cs Delta)^2 |S|,
    // tau_ij = -2 nu_t S_ij; the stress is quadratic in the velocity, so it takes
    // the same normalisation factor as the products.
    const double u = in[0], v = in[1], w = in[2], s = a.out_scale;
    const double S11 = in[3], S22 = in[4], S33 = in[5], S12 = in[6], S13 = in[7], S23 = in[8];
    const double Smag = sqrt(2.0 * (S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * S12 * S12 +
                                    2.0 * S13 * S13 + 2.0 * S23 * S23));
    const double m2nut = -2.0 * a.c2 * Smag * s;
    out[0] = s * (u * u);
    out[1] = s * (v * v);
    out[2] = s * (w * w);
    out[3] = s * (u * v);
    out[4] = s * (u * w);
    out[5] = s * (v * w);
    out[6] = m2nut * S11;
    out[7] = m2nut * S22;
    out[8] = m2nut * S33;
    out[9] = m2nut * S12;
    out[10] = m2nut * S13;
    out[11] = m2nut * S23;
  } else if constexpr (MODE == Z_SGS) {
    // u_i u_j of the velocity (0..5) and of the filtered velocity (6..11), order
    // 11 22 33 12 13 23 (src/Classes.py:886-909)
    const double s = a.out_scale;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const double u = in[3 * h], v = in[3 * h + 1], w = in[3 * h + 2];
      out[6 * h + 0] = s * (u * u);
      out[6 * h + 1] = s * (v * v);
      out[6 * h + 2] = s * (w * w);
      out[6 * h + 3] = s * (u * v);
      out[6 * h + 4] = s * (u * w);
      out[6 * h + 5] = s * (v * w);
    }
  } else if constexpr (MODE == Z_TM2) {
    // u_j * (PL_q)_i, stored at index 3 i + j (src/RHSfunctions.py:325-335)
    const double s = a.out_scale;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) out[3 * i + j] = s * (in[j] * in[3 + i]);
  }
}

// a.fold: the products leave as D_mn = N_mn + tau_mn (outputs 0..5), next to the
// stresses (6..11)
template <int MODE>
__device__ __forceinline__ void zpoint_fold(const ZArgs& a, double* out) {
  if constexpr (MODE == Z_SMAG || MODE == Z_LRN || MODE == Z_LRNF) {
    if (a.fold) {
#pragma unroll
      for (int g = 0; g < 6; ++g) out[g] += out[6 + g];
    }
  }
}

// One block = P consecutive (x, y) positions.  Real pencil r = f*P + p (field f,
// position p) lives in half (r & 1) of complex pencil r >> 1, so one complex FFT
// transforms two real pencils -- two positions of a field when P is even, two
// fields of the one position when P = 1.  Inverse: Z = A + iB with the Hermitian
// halves mirrored (Re only at k = 0 and L/2, as c2r does); forward:
// A = (Z_k + conj Z_{L-k})/2, B = (Z_k - conj Z_{L-k})/(2i).  NTZ threads; LDS rows
// unpadded (PAD = false measured faster for this kernel).
template <int L, int P, int MODE, int NTZ>
__global__ __launch_bounds__(NTZ) void zpass(ZArgs a, const double2* __restrict__ tw) {
  constexpr bool PADZ = false;  // unpadded rows measured faster in this kernel
  constexpr int NIN = ZShape<MODE>::nin;
  constexpr int NOUT = ZShape<MODE>::nout;
  constexpr int NRI = NIN * P, NRO = NOUT * P;  // real pencils in / out
  constexpr int NCI = (NRI + 1) / 2;
  constexpr int NCO = (NRO + 1) / 2;
  constexpr int NCMAX = NCI > NCO ? NCI : NCO;
  constexpr int LS = Row<L, PADZ>::LS;
  constexpr int L2 = L / 2;
  __shared__ double2 buf[NCMAX * LS];
  const int tid = threadIdx.x;
  PSX_STAGE_TW(NTZ)
  const ZMember zm = zmember(a, (long long)blockIdx.x * P);
  const long long pos0 = zm.pos0;
#define ZRW(i) (::psx::Row<L, PADZ>::at(i))

  // --- build the inverse spectra
  constexpr int HS = L2 + 1;  // staged k3 modes per real input field
  constexpr int NRS = ZFields<MODE>::n * P;  // staged rows: input fields x positions
  // stage the block's input spectra in LDS with one coalesced load pass when they fit
  constexpr bool STAGE = NRS * HS <= NCMAX * LS;
  if constexpr (STAGE) {
    // one coalesced pass of independent global loads into LDS, then the recipe
    // combinations from LDS into registers, then the Hermitian pencils
    for (int e = tid; e < NRS * HS; e += NTZ) {
      const int r = e / HS, k = e - r * HS;
      buf[e] = zin(a, zm, r / P, pos0 + r % P, k);
    }
    __syncthreads();
    constexpr int IB = (NCI * (L2 + 1) + NTZ - 1) / NTZ;
    double2 zA[IB], zB[IB];
#pragma unroll
    for (int ib = 0; ib < IB; ++ib) {
      const int e = tid + ib * NTZ;
      zA[ib] = zB[ib] = make_double2(0.0, 0.0);
      if (e < NCI * (L2 + 1)) {
        const int q = e / (L2 + 1), k = e - q * (L2 + 1);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int r = 2 * q + h;
          if (r < NRI) {
            const int p = r % P;
            const Recipe& R = a.rec[r / P];
            double2 acc = make_double2(0.0, 0.0);
            for (int t = 0; t < R.nterm; ++t) {
              double2 v = buf[(R.f[t] * P + p) * HS + k];
              double sc = R.coef[t];
              if (R.dz[t]) {
                v = cmul_i<1>(v);
                sc *= a.kzscale * k;
              }
              acc.x += sc * v.x;
              acc.y += sc * v.y;
            }
            if (h == 0) zA[ib] = acc; else zB[ib] = acc;
          }
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int ib = 0; ib < IB; ++ib) {
      const int e = tid + ib * NTZ;
      if (e < NCI * (L2 + 1)) {
        const int q = e / (L2 + 1), k = e - q * (L2 + 1);
        const double2 A = zA[ib], B = zB[ib];
        if (k == 0 || k == L2) {
          buf[q * LS + ZRW(k)] = make_double2(A.x, B.x);
        } else {
          buf[q * LS + ZRW(k)] = make_double2(A.x - B.y, A.y + B.x);
          buf[q * LS + ZRW(L - k)] = make_double2(A.x + B.y, B.x - A.y);
        }
      }
    }
  } else {
    for (int e = tid; e < NCI * (L2 + 1); e += NTZ) {
      const int q = e / (L2 + 1), k = e - q * (L2 + 1);
      const int r0 = 2 * q, r1 = 2 * q + 1;
      const double2 A = zspec(a, zm, r0 / P, pos0 + r0 % P, k);
      const double2 B = r1 < NRI ? zspec(a, zm, r1 / P, pos0 + r1 % P, k) : make_double2(0.0, 0.0);
      if (k == 0 || k == L2) {
        buf[q * LS + ZRW(k)] = make_double2(A.x, B.x);
      } else {
        buf[q * LS + ZRW(k)] = make_double2(A.x - B.y, A.y + B.x);
        buf[q * LS + ZRW(L - k)] = make_double2(A.x + B.y, B.x - A.y);
      }
    }
  }
  __syncthreads();
  fft_batch<L, NCI, NTZ, LS, +1, PADZ>(buf, twp, tid);

  if constexpr (MODE == Z_PHYS || MODE == Z_GRAD) {
    for (int z = tid; z < L; z += NTZ) {
      double2 w[NCI];
#pragma unroll
      for (int q = 0; q < NCI; ++q) w[q] = buf[q * LS + ZRW(z)];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const long long pos = pos0 + p;
#pragma unroll
        for (int f = 0; f < NIN; ++f) {
          const int r = f * P + p;
          const double v = ((r & 1) ? w[r >> 1].y : w[r >> 1].x) * a.out_scale;
          if constexpr (MODE == Z_PHYS) {
            if (f < a.nphys) a.phys[(long long)f * a.out_sf + pos * L + z] = v;
          }
          else
            a.gout[(((long long)blockIdx.x * P + p) * L + z) * 9 + f] = (float)v;
        }
      }
    }
    return;
  } else if constexpr (MODE == Z_DT) {
    double m = 0.0;
    for (int z = tid; z < L; z += NTZ) {
      double2 w[NCI];
#pragma unroll
      for (int q = 0; q < NCI; ++q) w[q] = buf[q * LS + ZRW(z)];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        double v[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          const int r = f * P + p;
          v[f] = (r & 1) ? w[r >> 1].y : w[r >> 1].x;
        }
        m = max_nan(m, fabs(v[0]) * a.dinv[0] + fabs(v[1]) * a.dinv[1] + fabs(v[2]) * a.dinv[2]);
      }
    }
    // wave reduction, then the block's maximum into its own slot (no atomics)
    for (int off = 32; off > 0; off >>= 1) m = max_nan(m, __shfl_xor(m, off));
    __shared__ double wmax[NTZ / 64];
    if ((tid & 63) == 0) wmax[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      double b = wmax[0];
      for (int w = 1; w < NTZ / 64; ++w) b = max_nan(b, wmax[w]);
      a.red[blockIdx.x] = b;
    }
    return;
  } else {
    // --- pointwise products / closure: a thread owns one z for all P positions and
    // moves whole complex pencil entries (16 B per lane, contiguous across lanes)
    static_assert(P <= 2, "pointwise phase handles P = 1 or 2");
    constexpr int IT = (L + NTZ - 1) / NTZ;
    double2 res[IT][NCO];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int z = tid + it * NTZ;
      if ((L % NTZ) == 0 || z < L) {
        double2 v[NCI];
#pragma unroll
        for (int q = 0; q < NCI; ++q) v[q] = buf[q * LS + ZRW(z)];
        double o[P][NOUT];
#pragma unroll
        for (int p = 0; p < P; ++p) {
          double in[NIN];
#pragma unroll
          for (int f = 0; f < NIN; ++f) {
            const int r = f * P + p;
            in[f] = (r & 1) ? v[r >> 1].y : v[r >> 1].x;
          }
          zpoint<MODE>(a, in, o[p], ((long long)blockIdx.x * P + p) * L + z);
          zpoint_fold<MODE>(a, o[p]);
        }
#pragma unroll
        for (int q = 0; q < NCO; ++q) {
          const int r0 = 2 * q, r1 = 2 * q + 1;
          res[it][q].x = o[r0 % P][r0 / P];
          res[it][q].y = r1 < NRO ? o[r1 % P][r1 / P] : 0.0;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int z = tid + it * NTZ;
      if ((L % NTZ) == 0 || z < L) {
#pragma unroll
        for (int q = 0; q < NCO; ++q) buf[q * LS + ZRW(z)] = res[it][q];
      }
    }
    __syncthreads();
    fft_batch<L, NCO, NTZ, LS, -1, PADZ>(buf, twp, tid);
    for (int e = tid; e < NCO * a.Hout; e += NTZ) {
      const int q = e / a.Hout, k = e - q * a.Hout;
      const double2 z1 = buf[q * LS + ZRW(k)];
      const double2 z2 = buf[q * LS + ZRW((k == 0 ? 0 : L - k))];
      const int r0 = 2 * q, r1 = 2 * q + 1;
      zm.out[(long long)(r0 / P) * a.out_sf + (pos0 + r0 % P) * a.Hout + k] =
          make_double2(0.5 * (z1.x + z2.x), 0.5 * (z1.y - z2.y));
      if (r1 < NRO)
        zm.out[(long long)(r1 / P) * a.out_sf + (pos0 + r1 % P) * a.Hout + k] =
            make_double2(0.5 * (z1.y + z2.y), 0.5 * (z2.x - z1.x));
    }
  }
#undef ZRW
}

// Sum a complex value over the block (deterministic tree order); uses >= NT
// entries of the scratch LDS buffer, which must be free.  Result valid in tid 0.
template <int NTH = NT>
__device__ __forceinline__ double2 block_sum2(double2 v, double2* scratch, int tid) {
  __syncthreads();
  scratch[tid] = v;
  __syncthreads();
  for (int s = NTH / 2; s > 0; s >>= 1) {
    if (tid < s) scratch[tid] = cadd(scratch[tid], scratch[tid + s]);
    __syncthreads();
  }
  return scratch[0];
}

// ------------------------------------------------------------------ x-forward + epilogue
enum EpiMode { E_DNS = 0, E_SMAG = 1, E_TAU2 = 2, E_SMAGF = 3, E_SMAG9 = 4 };
template <int EPI> struct EpiShape;
template <> struct EpiShape<E_DNS> { static constexpr int nf = 6; };
template <> struct EpiShape<E_SMAG> { static constexpr int nf = 12; };
// E_SMAGF: the Smagorinsky form with products and stresses folded.  The input holds
// D_mn = N_mn + tau_mn (fields 0..5, ZArgs::fold) and the stresses tau_mn (6..11, mn =
// 11 22 33 12 13 23); the epilogue transforms 9 fields: D_mn, A = tau_11,
// B = 2 k2 tau_12 + 2 k3 tau_13 and C = k2^2 tau_22 + 2 k2 k3 tau_23 + k3^2 tau_33,
// formed before its x transform (k2 and k3 are constants of a block column), so that
// sum_mn s_mn k_m k_n tau_mn = k1^2 A + k1 B + C (s = 1 on the diagonal, 2 off it).
template <> struct EpiShape<E_SMAGF> { static constexpr int nf = 9; };
// E_SMAG9: the same 9 fields with A, B and C already formed by the y-forward
// (pencil_forward_sum), so the epilogue reads 9 fields instead of 12.
template <> struct EpiShape<E_SMAG9> { static constexpr int nf = 9; };
// The epilogue reads its input one field per step.  E_SMAGF has 12 steps for its 9
// transformed fields: steps 0..6 are D_mn and A; steps 7, 8 (tau_12, tau_13) build B
// and steps 9, 10, 11 (tau_22, tau_23, tau_33) build C in LDS, and only the last step
// of each transforms.  Other modes: step s is field s.
template <int EPI> struct EpiSteps { static constexpr int n = EpiShape<EPI>::nf; };
template <> struct EpiSteps<E_SMAGF> { static constexpr int n = 12; };
__host__ __device__ constexpr int epi_step_field(int epi, int s) {
  return epi != E_SMAGF || s < 7 ? s : (s == 7 ? 9 : (s == 8 ? 10 : (s == 9 ? 7 : (s == 10 ? 11 : 8))));
}
template <> struct EpiShape<E_TAU2> { static constexpr int nf = 9; };

struct EpiArgs {
  const double2* in;          // W layout [f][x][k2][k3<H]
  long long in_sf;            // field stride
  int N1, N2, N3, Nc, H;
  int keep_nyq;               // x truncation rule (see pad_slot)
  const double2* uprev;       // state-layout velocity the RHS is evaluated at
  const double2* q0;          // stage base Q0 (nullptr: write the RHS itself)
  double2* out;               // state layout
  double2* w0;                // optional: the SGS term w0 (state layout), or nullptr
  const double2* pl;          // E_TAU2: truncated PL (state layout)
  const double* dt;           // device dt
  double c;                   // stage coefficient
  double nu;
  double ks[3];               // 2*pi/L_i
  int trunc_u;                // the model zeroes Nyquist entries of uhat before use
  int zero_nyq_out;           // zero Nyquist entries of the output
  double coef;                // E_TAU2: w0 coefficient (host value) ...
  const double* coef_dev;     // ... or a device scalar (overrides coef when set)
  int rotate;
  double Om[3];
  // E_TAU2 reductions for the dynamic tau-model: per-block partial (re, im) of
  // sum_w PLQL . conj(red_u) with the half-spectrum weights; store_only writes
  // PLQL (coefficient 1) to w0 (if set) and skips the state update.
  const double2* red_u;
  double* red_part;
  int store_only;
  // ensembles: grid.y = members; member m reads in + m*ms_in, the state-layout arrays
  // at + m*ms_state, dt + m, and its viscosity nu_m[m] (nu_m = nullptr: nu)
  long long ms_in, ms_state;
  const double* nu_m;
  int members;  // grid.y (0 or 1: one member)
  double force;  // linear forcing coefficient: RHS += force * u (u truncated where trunc_u)
};

// The epilogue's RHS is linear in its transformed fields, and every projection in it
// has the form -i V_i + i k_i (k.V)/k^2 of a vector V_i = sum_j X_ij k_j of a symmetric
// tensor X.  So each field only adds its k-weighted value to the sums of its tensor,
// and epi_project applies the projections once per output.  Sums per output:
//   V   -- X = the products N_mn (E_DNS, E_SMAG), D_mn = N_mn + tau_mn (E_SMAGF,
//          E_SMAG9), or A_ij + A_ji with A_ij = u_j (PL_q)_i (E_TAU2);
//   Vt  -- X = tau_mn (E_SMAG: the stresses keep their own projection);
//   St  -- k.tau.k = k1^2 A + k1 B + C (E_SMAGF, E_SMAG9: fields 6, 7, 8).
// Fields 0..5 (and E_SMAG's 6..11) hold the symmetric components mn = 11 22 33 12 13
// 23; E_TAU2's field 3 i + j holds A_ij.
template <int EPI>
__device__ __forceinline__ void epi_accumulate(int f, const double* k, double2 v, double2* V, double2* Vt,
                                               double2& St) {
  if constexpr (EPI == E_SMAGF || EPI == E_SMAG9) {
    if (f >= 6) {
      const double w = f == 6 ? k[0] * k[0] : (f == 7 ? k[0] : 1.0);
      St = make_double2(St.x + w * v.x, St.y + w * v.y);
      return;
    }
  }
  if constexpr (EPI == E_TAU2) {
    // A_ij adds k_j v to V_i and k_i v to V_j (2 k_i v on the diagonal); f is the same
    // for the whole block (a scalar branch)
    switch (f) {
      case 0: cfma(V[0], 2.0 * k[0], v); break;
      case 1: cfma(V[0], k[1], v); cfma(V[1], k[0], v); break;
      case 2: cfma(V[0], k[2], v); cfma(V[2], k[0], v); break;
      case 3: cfma(V[1], k[0], v); cfma(V[0], k[1], v); break;
      case 4: cfma(V[1], 2.0 * k[1], v); break;
      case 5: cfma(V[1], k[2], v); cfma(V[2], k[1], v); break;
      case 6: cfma(V[2], k[0], v); cfma(V[0], k[2], v); break;
      case 7: cfma(V[2], k[1], v); cfma(V[1], k[2], v); break;
      default: cfma(V[2], 2.0 * k[2], v); break;
    }
  } else {
    sym_accumulate(f % 6, k, v, (EPI == E_SMAG && f >= 6) ? Vt : V);
  }
}

// The RHS (or E_TAU2's PLQL) at wavenumber k from an output's sums, kinv = 1/(|k|^2 +
// 1e-50):
//   out_i = -i V_i + i k_i (k.V) kinv     the Leray projection of the products'
//                                         divergence (src/RHSfunctions.py:29-45), or
//                                         PLQL (src/RHSfunctions.py:338-352)
//   sgs_i = -i Vt_i - k_i (k.Vt) kinv     E_SMAG: the stress term with its 1j*1j as
//                                         written (src/RHSfunctions.py:215-222)
//   out_i -= (1 + i) k_i St kinv          E_SMAGF, E_SMAG9: with the products'
//                                         projection applied to D = N + tau, the
//                                         stress term as written leaves this
// E_SMAG's out includes sgs.
template <int EPI>
__device__ __forceinline__ void epi_project(const double* k, double kinv, const double2* V, const double2* Vt,
                                            double2 St, double2* out, double2* sgs) {
  const double2 kV = make_double2(k[0] * V[0].x + k[1] * V[1].x + k[2] * V[2].x,
                                  k[0] * V[0].y + k[1] * V[1].y + k[2] * V[2].y);
  const double2 w = cscale(kV, kinv);
#pragma unroll
  for (int i = 0; i < 3; ++i) out[i] = cmul_i<1>(csub(cscale(w, k[i]), V[i]));
  if constexpr (EPI == E_SMAG) {
    const double2 kT = make_double2(k[0] * Vt[0].x + k[1] * Vt[1].x + k[2] * Vt[2].x,
                                    k[0] * Vt[0].y + k[1] * Vt[1].y + k[2] * Vt[2].y);
    const double2 wt = cscale(kT, kinv);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      sgs[i] = csub(cmul_i<-1>(Vt[i]), cscale(wt, k[i]));
      out[i] = cadd(out[i], sgs[i]);
    }
  }
  if constexpr (EPI == E_SMAGF || EPI == E_SMAG9) {
    const double2 ws = cscale(St, kinv);
    const double2 t = make_double2(ws.x - ws.y, ws.x + ws.y);
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = csub(out[i], cscale(t, k[i]));
  }
}

// One block = (k2, k3-tile of T) running the x-forward FFT of every field and
// accumulating the spectral RHS in registers; blocks past N2*ceil(H/T) handle the
// k3 >= H planes (no transformed data there) elementwise.
template <int L, int NS1, int T, int EPI, bool W0>
__global__ __launch_bounds__(NT) void xforward_epilogue(EpiArgs a, const double2* __restrict__ tw) {
  constexpr int LS = Row<L>::LS;
  constexpr int NF = EpiShape<EPI>::nf;
  constexpr int IT = (NS1 * T + NT - 1) / NT;
  // four entries per thread and NS1 = L / 2 (the t-model family's 2x lengths 128, 256,
  // 512): the transforms keep their first and last stages in registers (fft_rows_reg),
  // fed by the loaded field entries; the kept outputs kx < NS1 / 2 and kx >= NS1 / 2
  // are the last stage's outputs 0 and 3 of the thread
  constexpr bool REG = RegPlan<L, 4>::ok && L * T == 4 * NT && 2 * NS1 == L && EPI != E_SMAGF;
  __shared__ double2 buf[(REG ? 2 : 1) * T * LS];
  const int tid = threadIdx.x;
  PSX_STAGE_TW(NT)
  const int ct = (a.H + T - 1) / T;
  const int nmain = a.N2 * ct;
  // the ct column tiles of one k2 row run on one XCD (they share the cache lines of
  // the state's unaligned rows) and consecutive rows go round-robin over the 8 XCDs.
  // (Dispatching the blocks of the k3 >= H planes first measured slower: 119 against
  // 115 us at 128^3.)
  const int bid = ((int)blockIdx.x < nmain && (a.N2 & 7) == 0)
                      ? (8 * (((int)blockIdx.x >> 3) / ct) + ((int)blockIdx.x & 7)) * ct + ((int)blockIdx.x >> 3) % ct
                      : (int)blockIdx.x;
  const int mem = blockIdx.y;
  const double2* const in_ = a.in + mem * a.ms_in;
  const double2* const uprev_ = a.uprev + mem * a.ms_state;
  const double2* const q0_ = a.q0 ? a.q0 + mem * a.ms_state : nullptr;
  double2* const out_ = a.out ? a.out + mem * a.ms_state : nullptr;
  double2* const w0_ = a.w0 ? a.w0 + mem * a.ms_state : nullptr;
  const double2* const pl_ = a.pl ? a.pl + mem * a.ms_state : nullptr;
  const double* const dt_ = a.dt ? a.dt + mem : nullptr;
  const double nu_ = a.nu_m ? a.nu_m[mem] : a.nu;
  const long long plane = (long long)a.N2 * a.Nc;
  const long long fstride = (long long)a.N1 * plane;  // component stride in the state layout

  if ((int)bid >= nmain) {
    // k3 in [H, Nc): elementwise (no transformed fields)
    const int nrest = a.Nc - a.H;
    const long long total = (long long)a.N1 * a.N2 * nrest;
    for (long long e = (long long)(bid - nmain) * NT + tid; e < total;
         e += (long long)(gridDim.x - nmain) * NT) {
      const int kz = a.H + (int)(e % nrest);
      const long long r = e / nrest;
      const int ky = (int)(r % a.N2), kx = (int)(r / a.N2);
      const long long idx = (long long)kx * plane + (long long)ky * a.Nc + kz;
      double k[3] = {a.ks[0] * signed_freq(kx, a.N1), a.ks[1] * signed_freq(ky, a.N2), a.ks[2] * kz};
      const double ksq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
      const bool nyq = (kx == a.N1 / 2) || (ky == a.N2 / 2) || (kz == a.N3 / 2);
      double2 u[3], rhs[3], qv[3];
      // every load first, then the arithmetic: one memory round trip
#pragma unroll
      for (int c = 0; c < 3; ++c) u[c] = uprev_[c * fstride + idx];
      if (q0_) {
#pragma unroll
        for (int c = 0; c < 3; ++c) qv[c] = q0_[c * fstride + idx];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (a.trunc_u && nyq) u[c] = make_double2(0.0, 0.0);
        rhs[c] = EPI == E_TAU2 ? make_double2(0.0, 0.0) : cscale(u[c], a.force - nu_ * (ksq + 1e-50));
        if (EPI == E_TAU2 && w0_) w0_[c * fstride + idx] = make_double2(0.0, 0.0);
        if (EPI == E_SMAG && w0_) w0_[c * fstride + idx] = make_double2(0.0, 0.0);
      }
      if (EPI == E_TAU2 && a.store_only) continue;
      if (a.rotate) {
        const double kinv = 1.0 / (ksq + 1e-50);
        double2 cr[3];
        cr[0] = csub(cscale(u[1], a.Om[2]), cscale(u[2], a.Om[1]));
        cr[1] = csub(cscale(u[2], a.Om[0]), cscale(u[0], a.Om[2]));
        cr[2] = csub(cscale(u[0], a.Om[1]), cscale(u[1], a.Om[0]));
        const double2 X = cadd(cadd(cscale(cr[0], k[0]), cscale(cr[1], k[1])), cscale(cr[2], k[2]));
#pragma unroll
        for (int c = 0; c < 3; ++c) rhs[c] = cadd(rhs[c], csub(cscale(cr[c], 2.0), cscale(X, 2.0 * k[c] * kinv)));
      }
      const double dtc = q0_ ? (*dt_) * a.c : 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double2 o = q0_ ? cadd(qv[c], cscale(rhs[c], dtc)) : rhs[c];
        if (a.zero_nyq_out && nyq) o = make_double2(0.0, 0.0);
        out_[c * fstride + idx] = o;
      }
    }
    if (EPI == E_TAU2 && a.red_part && tid == 0) {
      a.red_part[2 * bid] = 0.0;
      a.red_part[2 * bid + 1] = 0.0;
    }
    return;
  }

  const int ky = bid / ct;
  const int c0 = (bid - ky * ct) * T;
  const long long in_sl = (long long)a.N2 * a.H;  // x stride in W layout
  const double2* src0 = in_ + (long long)ky * a.H + c0;
  constexpr bool SGS = W0 && (EPI == E_SMAG);
  constexpr bool VT = EPI == E_SMAG;                      // the stresses' own sums
  constexpr bool ST = EPI == E_SMAGF || EPI == E_SMAG9;   // k.tau.k from A, B, C
  static_assert(EPI != E_TAU2 || T * LS >= NT, "block_sum2 needs NT LDS entries");
  // per output: the k-weighted sums of epi_accumulate
  double2 acc[IT][3];
  double2 accs[VT ? IT : 1][3];
  double2 acct[ST ? IT : 1];
#pragma unroll
  for (int it = 0; it < IT; ++it) acc[it][0] = acc[it][1] = acc[it][2] = make_double2(0.0, 0.0);
#pragma unroll
  for (int it = 0; it < (VT ? IT : 1); ++it) accs[it][0] = accs[it][1] = accs[it][2] = make_double2(0.0, 0.0);
#pragma unroll
  for (int it = 0; it < (ST ? IT : 1); ++it) acct[it] = make_double2(0.0, 0.0);

  const double k2 = a.ks[1] * signed_freq(ky, a.N2);
  // One step: stage input field epi_step_field(s) in LDS (from the registers `pre`,
  // or straight from memory without prefetch), weighted and added to what the
  // group's earlier steps staged (E_SMAGF's B and C), refill `pre` with step snext,
  // and after a group's last step transform and accumulate.  Between steps without a
  // transform each thread touches only its own staged elements, so no barrier.
  constexpr int LT = (L * T + NT - 1) / NT;
  constexpr int NSTEP = EpiSteps<EPI>::n;
  auto load_field = [&](int s, double2 (&pre)[LT]) {
    const double2* src = src0 + (long long)epi_step_field(EPI, s) * a.in_sf;
#pragma unroll
    for (int i = 0; i < LT; ++i) {
      const int e = tid + i * NT;
      const int l = e / T, c = e - (e / T) * T;
      pre[i] = ((L * T) % NT == 0 || e < L * T) && (c0 + c < a.H) ? src[(long long)l * in_sl + c]
                                                                    : make_double2(0.0, 0.0);
    }
  };
  // E_SMAGF step weights: B = 2 k2 tau_12 + 2 k3 tau_13, C = k2^2 tau_22 + 2 k2 k3 tau_23
  // + k3^2 tau_33 (k3 of the element's column)
  auto weight = [&](int s, int c) -> double {
    const double k3 = a.ks[2] * (c0 + c);
    return s < 7 ? 1.0
                 : (s == 7 ? 2.0 * k2
                           : (s == 8 ? 2.0 * k3 : (s == 9 ? k2 * k2 : (s == 10 ? 2.0 * k2 * k3 : k3 * k3))));
  };
  // The state entries this thread updates, uprev and q0 (when set), all loaded before
  // any is used: the last step issues the uprev loads before its transform and the q0
  // loads before its accumulation, so that the round trips overlap that work (not with
  // the SGS output, whose accumulators leave no room for them)
  constexpr bool PRE = !SGS;
  double2 su[PRE ? IT : 1][3], sq[PRE ? IT : 1][3];
  auto load_state = [&](int which) {
    if constexpr (PRE) {
      const double2* src = which == 0 ? uprev_ : q0_;
      if (src) {
#pragma unroll
        for (int it = 0; it < IT; ++it) {
          const int e = tid + it * NT;
          const int kx = e / T, c = e - (e / T) * T;
          const bool ok = e < NS1 * T && c0 + c < a.H;
          const long long idx = (long long)kx * plane + (long long)ky * a.Nc + c0 + c;
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            const double2 v = ok ? src[cc * fstride + idx] : make_double2(0.0, 0.0);
            if (which == 0)
              su[it][cc] = v;
            else
              sq[it][cc] = v;
          }
        }
      }
    }
  };
  // LAST: the final step, which issues the state loads in place of a next field's
  // (a compile-time flag, so that those registers are live in that step only)
  [[maybe_unused]] double2* xa = buf;
  [[maybe_unused]] double2* xb = buf + (REG ? T * LS : 0);
  auto run_step = [&](auto last, int s, double2 (&pre)[LT], int snext) {
    if constexpr (REG) {
      static_assert(LT == 4 && IT == 2, "four entries per thread, two kept");
      double2 v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = pre[i];
      if constexpr (decltype(last)::value)
        load_state(0);
      else if (snext < NSTEP)
        load_field(snext, pre);
      fft_rows_reg<L, 4, T, NT, LS, -1>(v, xa, xb, twp, tid);
      if constexpr (fft_rows_reg_swap<L, 4>()) {
        double2* t = xa;
        xa = xb;
        xb = t;
      }
      if constexpr (decltype(last)::value) load_state(1);
      const int c = tid % T;
      if (c0 + c < a.H) {
#pragma unroll
        for (int it = 0; it < 2; ++it) {
          const int kx = tid / T + (L / 4) * it;  // slot kx (it = 0) or L / 2 + kx (it = 1)
          if (pad_slot(kx, NS1, L, a.keep_nyq) >= 0) {
            const double k[3] = {a.ks[0] * signed_freq(kx, NS1), k2, a.ks[2] * (c0 + c)};
            epi_accumulate<EPI>(s, k, v[it == 0 ? 0 : 3], acc[it], accs[VT ? it : 0], acct[ST ? it : 0]);
          }
        }
      }
      return;
    }
    const bool fold = EPI == E_SMAGF && s >= 7;
    const bool add = fold && (s == 8 || s == 10 || s == 11);
    const bool transform = !fold || s == 8 || s == 11;
    const int f = fold ? (s <= 8 ? 7 : 8) : s;  // the transformed field this step feeds
#pragma unroll
    for (int i = 0; i < LT; ++i) {
      const int e = tid + i * NT;
      const int l = e / T, c = e - (e / T) * T;
      if ((L * T) % NT == 0 || e < L * T) {
        double2 v = pre[i];
        if (fold) v = cscale(v, weight(s, c));
        if (add) v = cadd(v, buf[c * LS + RW(l)]);
        buf[c * LS + RW(l)] = v;
      }
    }
    if constexpr (decltype(last)::value)
      load_state(0);
    else if (snext < NSTEP)
      load_field(snext, pre);
    if (!transform) return;
    __syncthreads();
    fft_batch<L, T, NT, LS, -1>(buf, twp, tid);
    if constexpr (decltype(last)::value) load_state(1);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      if (e < NS1 * T) {
        const int kx = e / T, c = e - (e / T) * T;
        const int p = pad_slot(kx, NS1, L, a.keep_nyq);
        if (p >= 0 && c0 + c < a.H) {
          const double k[3] = {a.ks[0] * signed_freq(kx, NS1), k2, a.ks[2] * (c0 + c)};
          epi_accumulate<EPI>(f, k, buf[c * LS + RW(p)], acc[it], accs[VT ? it : 0], acct[ST ? it : 0]);
        }
      }
    }
    __syncthreads();
  };
  // step s+1 loads while step s is transformed (a second buffer, two steps ahead,
  // measured no faster)
  double2 pa[LT];
  load_field(0, pa);
#pragma unroll 1
  for (int s = 0; s < NSTEP - 1; ++s) run_step(std::false_type{}, s, pa, s + 1);
  run_step(std::true_type{}, NSTEP - 1, pa, NSTEP);

  const double dtc = q0_ ? (*dt_) * a.c : 0.0;
  const double coef = (EPI == E_TAU2 && a.store_only) ? 1.0 : (a.coef_dev ? *a.coef_dev : a.coef);
  double2 red = make_double2(0.0, 0.0);
  // the t-model's PL and reduction velocity, loaded before any is used
  constexpr bool PRET = EPI == E_TAU2;
  double2 sp[PRET ? IT : 1][3], sr[PRET ? IT : 1][3];
  if constexpr (PRET) {
    const bool want_pl = pl_ && !a.store_only;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      const int kx = e / T, c = e - (e / T) * T;
      const bool ok = e < NS1 * T && c0 + c < a.H;
      const long long idx = (long long)kx * plane + (long long)ky * a.Nc + c0 + c;
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        sp[it][cc] = ok && want_pl ? pl_[cc * fstride + idx] : make_double2(0.0, 0.0);
        sr[it][cc] = ok && a.red_u ? a.red_u[cc * fstride + idx] : make_double2(0.0, 0.0);
      }
    }
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + it * NT;
    if (e < NS1 * T) {
      const int kx = e / T, c = e - (e / T) * T;
      const int kz = c0 + c;
      if (kz < a.H) {
        const long long idx = (long long)kx * plane + (long long)ky * a.Nc + kz;
        const double k[3] = {a.ks[0] * signed_freq(kx, NS1), k2, a.ks[2] * kz};
        const double ksq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        const double kinv = 1.0 / (ksq + 1e-50);
        const bool nyq = (kx == a.N1 / 2) || (ky == a.N2 / 2) || (kz == a.N3 / 2);
        double2 u[3], rhs[3], nl[3], sgs[3];
        epi_project<EPI>(k, kinv, acc[it], accs[VT ? it : 0], acct[ST ? it : 0], nl, sgs);
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          u[cc] = (a.trunc_u && nyq) ? make_double2(0.0, 0.0)
                                     : (PRE ? su[PRE ? it : 0][cc] : uprev_[cc * fstride + idx]);
          if constexpr (EPI == E_TAU2) {
            const double2 w = cscale(nl[cc], coef);
            if (w0_) w0_[cc * fstride + idx] = w;
            if (a.red_u) {
              const double wz = kz == 0 ? 1.0 : (kz < a.N3 / 2 ? 2.0 : 0.0);
              const double2 r = sr[PRET ? it : 0][cc];
              red.x += wz * (w.x * r.x + w.y * r.y);
              red.y += wz * (w.y * r.x - w.x * r.y);
            }
            rhs[cc] = a.store_only ? make_double2(0.0, 0.0)
                                   : cadd(cadd(sp[PRET ? it : 0][cc], w), cscale(u[cc], a.force));
          } else {
            rhs[cc] = cadd(nl[cc], cscale(u[cc], a.force - nu_ * (ksq + 1e-50)));
          }
        }
        if constexpr (SGS) {
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) w0_[cc * fstride + idx] = sgs[cc];
        }
        if (a.rotate) {
          double2 cr[3];
          cr[0] = csub(cscale(u[1], a.Om[2]), cscale(u[2], a.Om[1]));
          cr[1] = csub(cscale(u[2], a.Om[0]), cscale(u[0], a.Om[2]));
          cr[2] = csub(cscale(u[0], a.Om[1]), cscale(u[1], a.Om[0]));
          const double2 X = cadd(cadd(cscale(cr[0], k[0]), cscale(cr[1], k[1])), cscale(cr[2], k[2]));
#pragma unroll
          for (int cc = 0; cc < 3; ++cc)
            rhs[cc] = cadd(rhs[cc], csub(cscale(cr[cc], 2.0), cscale(X, 2.0 * k[cc] * kinv)));
        }
        if (!(EPI == E_TAU2 && a.store_only)) {
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            double2 o = q0_ ? cadd(PRE ? sq[PRE ? it : 0][cc] : q0_[cc * fstride + idx], cscale(rhs[cc], dtc))
                            : rhs[cc];
            if (a.zero_nyq_out && nyq) o = make_double2(0.0, 0.0);
            out_[cc * fstride + idx] = o;
          }
        }
      }
    }
  }
  if constexpr (EPI == E_TAU2) {
    if (a.red_part) {
      const double2 tot = block_sum2(red, buf, tid);
      if (tid == 0) {
        a.red_part[2 * bid] = tot.x;
        a.red_part[2 * bid + 1] = tot.y;
      }
    }
  }
}

// ------------------------------------------------------------------ filtered-DNS stress
// Filters on the N grid.  PSX_FILTER_SHARP is grid.Gf (src/Classes.py:584-588):
// 1 where the signed indices satisfy |i1|, |i2| < kc and i3 < kc.  PSX_FILTER_BOX is
// the top-hat of width delta: prod_i sin(k_i delta/2)/(k_i delta/2).
enum FilterKind { F_SHARP = 0, F_BOX = 1 };
__device__ __forceinline__ double sinc_half(double k, double delta) {
  const double x = 0.5 * k * delta;
  return x == 0.0 ? 1.0 : sin(x) / x;
}
__device__ __forceinline__ double filter_gain(int kind, int s1, int s2, int kz, double kc, double delta,
                                              const double* ks) {
  if (kind == F_SHARP) return (abs(s1) < kc && abs(s2) < kc && kz < kc) ? 1.0 : 0.0;
  return sinc_half(ks[0] * s1, delta) * sinc_half(ks[1] * s2, delta) * sinc_half(ks[2] * kz, delta);
}

struct SgsArgs {
  const double2* in;   // W layout: 12 fields, products of u (0..5) and of G u (6..11)
  long long in_sf;
  int N1, N2, N3, Nc, H;
  double2* tau;        // state layout, 6 fields (11 22 33 12 13 23)
  int filter;
  double kc, delta;
  double ks[3];
};

// x-forward of the twelve product fields and tau_hat_mn = G N_mn - N~_mn into the state
// layout (src/Classes.py:904-909); the 3/2 truncation zeroes index N1/2 and k3 >= H.
template <int L, int NS1, int T>
__global__ __launch_bounds__(NT) void xforward_sgs(SgsArgs a, const double2* __restrict__ tw) {
  constexpr int LS = Row<L>::LS;
  constexpr int IT = (NS1 * T + NT - 1) / NT;
  __shared__ double2 buf[T * LS];
  const int tid = threadIdx.x;
  const int ct = (a.H + T - 1) / T;
  const int nmain = a.N2 * ct;
  const long long plane = (long long)a.N2 * a.Nc;
  const long long fstride = (long long)a.N1 * plane;
  if ((int)blockIdx.x >= nmain) {
    const int nrest = a.Nc - a.H;
    const long long total = (long long)a.N1 * a.N2 * nrest;
    for (long long e = (long long)(blockIdx.x - nmain) * NT + tid; e < total;
         e += (long long)(gridDim.x - nmain) * NT) {
      const int kz = a.H + (int)(e % nrest);
      const long long r = e / nrest;
      const long long idx = (r / a.N2) * plane + (r % a.N2) * a.Nc + kz;
#pragma unroll
      for (int m = 0; m < 6; ++m) a.tau[m * fstride + idx] = make_double2(0.0, 0.0);
    }
    return;
  }
  const int ky = blockIdx.x / ct;
  const int c0 = (blockIdx.x - ky * ct) * T;
  const long long in_sl = (long long)a.N2 * a.H;
  const double2* src0 = a.in + (long long)ky * a.H + c0;
  const int s2 = signed_freq(ky, a.N2);
  double2 acc[IT][6];
#pragma unroll
  for (int it = 0; it < IT; ++it)
#pragma unroll
    for (int m = 0; m < 6; ++m) acc[it][m] = make_double2(0.0, 0.0);
#pragma unroll
  for (int f = 0; f < 12; ++f) {
    const double2* src = src0 + (long long)f * a.in_sf;
    for (int e = tid; e < L * T; e += NT) {
      const int l = e / T, c = e - (e / T) * T;
      buf[c * LS + RW(l)] = (c0 + c < a.H) ? src[(long long)l * in_sl + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_batch<L, T, NT, LS, -1>(buf, tw, tid);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NT;
      if (e < NS1 * T) {
        const int kx = e / T, c = e - (e / T) * T;
        const int p = pad_slot(kx, NS1, L, 0);
        if (p >= 0 && c0 + c < a.H) {
          const double2 v = buf[c * LS + RW(p)];
          const double g = f < 6 ? filter_gain(a.filter, signed_freq(kx, NS1), s2, c0 + c, a.kc, a.delta, a.ks)
                                 : -1.0;
          acc[it][f % 6] = cadd(acc[it][f % 6], cscale(v, g));
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + it * NT;
    if (e < NS1 * T) {
      const int kx = e / T, c = e - (e / T) * T;
      const int kz = c0 + c;
      if (kz < a.H) {
        const long long idx = (long long)kx * plane + (long long)ky * a.Nc + kz;
#pragma unroll
        for (int m = 0; m < 6; ++m) a.tau[m * fstride + idx] = acc[it][m];
      }
    }
  }
}

// ------------------------------------------------------------------ 2x-grid PL pass
// First round of the t-model family (src/RHSfunctions.py:269-323): the x-forward
// of the six 2x-grid product spectra (no truncation), the projected RHS
//   PL_i = -i k_j N_ij - i k_i p - nu k^2 pad_2x(u)_i,  p = -k_m k_n N_mn / k^2,
// its truncation to the base grid (unpad_2x, kept for the final sum), the q
// part PL_q = PL outside the resolved cube, and the x-inverse of PL_q for the
// second round -- all in one kernel per (k2, k3-tile).
struct PLArgs {
  const double2* in;      // [6][P1][P2][C] x physical, k2 and k3 spectral
  long long in_sf;
  int P1, P2, C;          // 2x-grid lengths; C = k3 modes carried (N3)
  int N1, N2, N3, Nc;     // base grid
  const double2* u;       // base-grid velocity for the viscous term (already truncated/filtered)
  double nu;
  double ks[3];
  int kpx, kpy, kpz;      // resolved cube: |s1| < kpx, |s2| < kpy, k3 < kpz (index units)
  double2* pln;           // out: PL on the base grid (state layout, k3 < N3/2), or nullptr
  double2* xq;            // out: [3][P1][P2][C] x-inverse of PL_q
  // dynamic tau-model Leonard term: G_i = -i k_j N_ij + i k_i p restricted to
  // |k_i| < kf (src/RHSfunctions.py:814-819)
  double2* g_out;         // store G (state layout) or nullptr
  const double2* g_ref;   // subtract this G and reduce against uf, or nullptr
  const double2* uf;      // filtered velocity for the LE reduction
  double kf;
  double* part;           // per-block (re, im) partial of sum_w (G - g_ref) . conj(uf)
};

template <int L, int T, bool DYN, int NTH>
__global__ __launch_bounds__(NTH) void pl_pass(PLArgs a, const double2* __restrict__ tw) {
  constexpr int LS = Row<L>::LS;
  constexpr int IT = (L * T + NTH - 1) / NTH;
  // four entries per thread (L = 128, 256, 512): the transforms keep their first and
  // last stages in registers (fft_rows_reg) on two LDS buffers, so the loaded field
  // entries feed the first stage and the last stage's outputs feed the accumulation,
  // with no staging pass in between
  constexpr bool REG = RegPlan<L, 4>::ok && L * T == 4 * NTH;
  // one buffer used in place where two would exceed 64 KiB (8 columns at length 256)
  constexpr bool INPL = REG && 2 * T * LS * 16 > 65536;
  static_assert(T * LS >= NTH, "block_sum2 needs NTH LDS entries");
  __shared__ double2 buf[(REG && !INPL ? 2 : 1) * T * LS];
  const int tid = threadIdx.x;
  PSX_STAGE_TW(NTH)
  const int ct = (a.C + T - 1) / T;
  // contiguous block ranges per XCD: the column tiles of one k2 row (T = 4 at L = 256:
  // 64-byte row segments, two tiles per 128-byte line) share one L2
  const int bid = xcd_block<true>();
  const int ky = bid / ct;
  const int c0 = (bid - ky * ct) * T;
  const long long in_sl = (long long)a.P2 * a.C;
  const double2* src0 = a.in + (long long)ky * a.C + c0;
  const int s2 = signed_freq(ky, a.P2);
  const double k2 = a.ks[1] * s2;
  // acc: V_i = sum_j N_ij k_j of the products; the divergence part of PL is -i V_i and
  // its projection part i k_i (k.V)/k^2 (see epi_accumulate)
  double2 acc[IT][3];
#pragma unroll
  for (int it = 0; it < IT; ++it)
#pragma unroll
    for (int i = 0; i < 3; ++i) acc[it][i] = make_double2(0.0, 0.0);
  // field f + 1 is loaded into registers while field f is transformed
  constexpr int LT = (L * T + NTH - 1) / NTH;
  double2 pre[LT];
  auto load_field = [&](int f, int tl) {
    const double2* src = src0 + (long long)f * a.in_sf;
#pragma unroll
    for (int i = 0; i < LT; ++i) {
      const int e = tl + i * NTH;
      const int l = e / T, c = e - (e / T) * T;
      pre[i] = ((L * T) % NTH == 0 || e < L * T) && (c0 + c < a.C) ? src[(long long)l * in_sl + c]
                                                                    : make_double2(0.0, 0.0);
    }
  };
  load_field(0, tid);
  [[maybe_unused]] double2* xa = buf;
  [[maybe_unused]] double2* xb = buf + (REG && !INPL ? T * LS : 0);
  if constexpr (REG) {
    static_assert(LT == 4 && IT == 4, "four entries per thread");
#pragma unroll 1
    for (int f = 0; f < 6; ++f) {
      const int tl = opaque(tid);  // addresses are recomputed for every field
      // thread tl holds entries kx = tl / T + (L / 4) it of column tl % T, as loaded
      double2 v[4];
#pragma unroll
      for (int it = 0; it < 4; ++it) v[it] = pre[it];
      if (f + 1 < 6) load_field(f + 1, tl);
      fft_rows_reg<L, 4, T, NTH, LS, -1, INPL>(v, xa, xb, twp, tl);
      if constexpr (!INPL && fft_rows_reg_swap<L, 4>()) {
        double2* t = xa;
        xa = xb;
        xb = t;
      }
      const int c = tl % T;
      if (c0 + c < a.C) {
#pragma unroll
        for (int it = 0; it < 4; ++it) {
          const int kx = tl / T + (L / 4) * it;
          const double k[3] = {a.ks[0] * signed_freq(kx, L), k2, a.ks[2] * (c0 + c)};
          sym_accumulate(f, k, v[it], acc[it]);
        }
      }
    }
  } else {
#pragma unroll 1
  for (int f = 0; f < 6; ++f) {
    const int tl = opaque(tid);  // addresses are recomputed for every field
#pragma unroll
    for (int i = 0; i < LT; ++i) {
      const int e = tl + i * NTH;
      if ((L * T) % NTH == 0 || e < L * T) buf[(e - (e / T) * T) * LS + RW(e / T)] = pre[i];
    }
    if (f + 1 < 6) load_field(f + 1, tl);
    __syncthreads();
    fft_batch<L, T, NTH, LS, -1>(buf, twp, tl);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tl + it * NTH;
      const int kx = e / T, c = e - (e / T) * T;
      if (e < L * T && c0 + c < a.C) {
        const double k[3] = {a.ks[0] * signed_freq(kx, L), k2, a.ks[2] * (c0 + c)};
        sym_accumulate(f, k, buf[c * LS + RW(kx)], acc[it]);
      }
    }
    __syncthreads();
  }
  }

  // PL = D + P - nu k^2 u (left in acc); base-grid copy; Leonard reduction
  const long long plane = (long long)a.N2 * a.Nc;
  const long long fs = (long long)a.N1 * plane;
  double2 red = make_double2(0.0, 0.0);
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + it * NTH;
    const int kx = e / T, c = e - (e / T) * T;
    const int kz = c0 + c;
    if (e < L * T && kz < a.C) {
      const int s1 = signed_freq(kx, L);
      const double k[3] = {a.ks[0] * s1, k2, a.ks[2] * kz};
      const double ksq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
      const bool in_base = (s1 >= -a.N1 / 2 && s1 < a.N1 / 2 && s2 >= -a.N2 / 2 && s2 < a.N2 / 2 && kz < a.N3 / 2);
      const long long bidx = in_base ? ((long long)((s1 + a.N1) % a.N1) * plane +
                                        (long long)((s2 + a.N2) % a.N2) * a.Nc + kz)
                                     : 0;
      const bool has_u = in_base && s1 > -a.N1 / 2 && s2 > -a.N2 / 2;
      // divergence part D_i = -i V_i and projection part P_i = i k_i (k.V)/k^2
      const double kinv = 1.0 / (ksq + 1e-50);
      const double2 kV = make_double2(k[0] * acc[it][0].x + k[1] * acc[it][1].x + k[2] * acc[it][2].x,
                                      k[0] * acc[it][0].y + k[1] * acc[it][1].y + k[2] * acc[it][2].y);
      double2 Dv[3], Pv[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        Dv[i] = cmul_i<-1>(acc[it][i]);
        Pv[i] = cmul_i<1>(cscale(kV, k[i] * kinv));
      }
      if constexpr (DYN) {
        const bool in_f = in_base && fabs(k[0]) < a.kf && fabs(k[1]) < a.kf && fabs(k[2]) < a.kf;
        if (in_f) {
          const double wz = kz == 0 ? 1.0 : (kz < a.N3 / 2 ? 2.0 : 0.0);
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            // G = -i k_j N_ij + i k_i p = D - P
            const double2 G = csub(Dv[i], Pv[i]);
            if (a.g_out) a.g_out[i * fs + bidx] = G;
            if (a.g_ref) {
              const double2 d = csub(G, a.g_ref[i * fs + bidx]);
              const double2 r = a.uf[i * fs + bidx];
              red.x += wz * (d.x * r.x + d.y * r.y);
              red.y += wz * (d.y * r.x - d.x * r.y);
            }
          }
        }
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double2 pl = cadd(Dv[i], Pv[i]);
        if (has_u) pl = csub(pl, cscale(a.u[i * fs + bidx], a.nu * (ksq + 1e-50)));
        if (in_base && a.pln) a.pln[i * fs + bidx] = pl;
        acc[it][i] = pl;
      }
    }
  }
  if constexpr (DYN) {
    if (a.part) {
      const double2 tot = block_sum2<NTH>(red, buf, tid);
      if (tid == 0) {
        a.part[2 * bid] = tot.x;
        a.part[2 * bid + 1] = tot.y;
      }
    }
  }
  // x-inverse of PL_q (PL outside the resolved cube), one component at a time
  // (unrolled: acc is indexed by the component)
  if constexpr (REG) {
    __syncthreads();  // block_sum2's last reads of buf
    const int b = tid % T;
    const bool valid = c0 + b < a.C;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double2 v[4];
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int s1 = signed_freq(tid / T + (L / 4) * it, L);
        const bool inside = (s1 > -a.kpx && s1 < a.kpx && s2 > -a.kpy && s2 < a.kpy && c0 + b < a.kpz);
        v[it] = (valid && !inside) ? acc[it][i] : make_double2(0.0, 0.0);
      }
      fft_rows_reg<L, 4, T, NTH, LS, +1, INPL>(v, xa, xb, twp, tid);
      if constexpr (!INPL && fft_rows_reg_swap<L, 4>()) {
        double2* t = xa;
        xa = xb;
        xb = t;
      }
      double2* dst = a.xq + (long long)i * a.P1 * a.P2 * a.C + (long long)ky * a.C + c0;
      if (valid) {
#pragma unroll
        for (int it = 0; it < 4; ++it) dst[(long long)(tid / T + (L / 4) * it) * in_sl + b] = v[it];
      }
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = tid + it * NTH;
      if (e < L * T) {
        const int kx = e / T, c = e - (e / T) * T;
        const int s1 = signed_freq(kx, L);
        const bool inside = (s1 > -a.kpx && s1 < a.kpx && s2 > -a.kpy && s2 < a.kpy && c0 + c < a.kpz);
        const bool valid = c0 + c < a.C;
        buf[c * LS + RW(kx)] = (valid && !inside) ? acc[it][i] : make_double2(0.0, 0.0);
      }
    }
    __syncthreads();
    fft_batch<L, T, NTH, LS, +1>(buf, twp, tid);
    double2* dst = a.xq + (long long)i * a.P1 * a.P2 * a.C + (long long)ky * a.C + c0;
    for (int e = tid; e < L * T; e += NTH) {
      const int x = e / T, c = e - (e / T) * T;
      if (c0 + c < a.C) dst[(long long)x * in_sl + c] = buf[c * LS + RW(x)];
    }
  }
}

}  // namespace psx
