// Batched FP64 complex FFTs staged in LDS, for gfx950 (wave64).
//
// Algorithm: mixed-radix Stockham autosort (Govindaraju et al., "High Performance
// Discrete Fourier Transforms on Graphics Processors", SC'08, Fig. 2): stage s
// with radix R reads butterfly j's inputs at j + r*L/R, applies the twiddles
// W_{Ns*R}^{(j mod Ns) r}, does an R-point DFT and writes to
// (j / Ns)*Ns*R + (j mod Ns) + r*Ns.  Output is in natural order.  Each stage
// reads all of a thread's butterflies into registers, barriers, and writes them
// back in place, so one LDS buffer of B*LS complex values holds the batch.
//
// Twiddles come from a precomputed table tw[m] = exp(-2*pi*i*m/L), m in [0, L),
// built on the host in double precision (no trigonometry in the kernels).
// SIGN = -1 is the forward transform exp(-i...), SIGN = +1 the unnormalised
// inverse exp(+i...).
#pragma once
#include <hip/hip_runtime.h>

namespace psx {

__host__ __device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// Store with the nontemporal hint (streaming: no allocation in the last-level cache).
// The z passes' outputs are read by the next kernel only after the whole pass, and
// measured faster downstream so stored (y-forward 131.5 -> 129.3 us at 128^3).
__device__ __forceinline__ void store_nt(double2* p, double2 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}
// multiply by +i (S = +1) or -i (S = -1)
template <int S>
__host__ __device__ __forceinline__ double2 cmul_i(double2 a) {
  return S > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// ---------------------------------------------------------------- radix plans
// The Stockham stages of each length, radices <= 8 (a family preferring radices 16
// and 12 -- fewer LDS passes, more registers per butterfly -- measured slower).
template <int L> struct Plan;
#define PSX_PLAN(LEN, ...)                                        \
  template <> struct Plan<LEN> {                                  \
    static constexpr int r[] = {__VA_ARGS__};                     \
    static constexpr int n = sizeof(r) / sizeof(int);             \
  };
PSX_PLAN(8, 8)
PSX_PLAN(12, 4, 3)
PSX_PLAN(16, 4, 4)
PSX_PLAN(24, 8, 3)
PSX_PLAN(32, 8, 4)
PSX_PLAN(36, 4, 3, 3)
PSX_PLAN(48, 4, 4, 3)
PSX_PLAN(64, 8, 8)
PSX_PLAN(72, 8, 3, 3)
PSX_PLAN(96, 8, 4, 3)
PSX_PLAN(128, 8, 4, 4)
PSX_PLAN(144, 4, 4, 3, 3)
PSX_PLAN(192, 8, 8, 3)
PSX_PLAN(256, 8, 8, 4)
PSX_PLAN(288, 8, 4, 3, 3)
PSX_PLAN(384, 8, 4, 4, 3)
PSX_PLAN(512, 8, 8, 8)
PSX_PLAN(576, 8, 8, 3, 3)
PSX_PLAN(768, 8, 8, 4, 3)
#undef PSX_PLAN

// ---------------------------------------------------------------- small DFTs
template <int R, int SIGN> struct Dft;

template <int SIGN> struct Dft<1, SIGN> {
  static __host__ __device__ __forceinline__ void run(double2*) {}
};

template <int SIGN> struct Dft<2, SIGN> {
  static __host__ __device__ __forceinline__ void run(double2* v) {
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int SIGN> struct Dft<3, SIGN> {
  static __host__ __device__ __forceinline__ void run(double2* v) {
    // exp(SIGN*2*pi*i/3) = -1/2 + SIGN*i*sqrt(3)/2
    constexpr double h = 0.86602540378443864676372317075294;  // sqrt(3)/2
    double2 t = cadd(v[1], v[2]);
    double2 d = csub(v[1], v[2]);
    double2 m = make_double2(v[0].x - 0.5 * t.x, v[0].y - 0.5 * t.y);
    double2 r = cmul_i<SIGN>(cscale(d, h));
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <int SIGN> struct Dft<4, SIGN> {
  static __host__ __device__ __forceinline__ void run(double2* v) {
    double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    double2 t2 = cadd(v[1], v[3]), t3 = cmul_i<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <int SIGN> struct Dft<8, SIGN> {
  static __host__ __device__ __forceinline__ void run(double2* v) {
    constexpr double s = 0.70710678118654752440084436210485;  // 1/sqrt(2)
    double2 e[4] = {v[0], v[2], v[4], v[6]};
    double2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, SIGN>::run(e);
    Dft<4, SIGN>::run(o);
    // W^k = exp(SIGN*2*pi*i*k/8)
    double2 w1 = make_double2(s * (o[1].x - SIGN * o[1].y), s * (o[1].y + SIGN * o[1].x));
    double2 w2 = cmul_i<SIGN>(o[2]);
    double2 w3 = make_double2(s * (-o[3].x - SIGN * o[3].y), s * (-o[3].y + SIGN * o[3].x));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
  }
};

// 6 = 2 x 3: n = 3 n1 + n2, k = k1 + 2 k2; DFT2 over n1, twiddle W6^(n2 k1), DFT3 over n2.
template <int SIGN> struct Dft<6, SIGN> {
  static __host__ __device__ __forceinline__ void run(double2* v) {
    constexpr double h = 0.86602540378443864676372317075294;  // sqrt(3)/2
    double2 b0[3], b1[3];
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
      b0[n2] = cadd(v[n2], v[n2 + 3]);
      b1[n2] = csub(v[n2], v[n2 + 3]);
    }
    // W6^1 = 1/2 + SIGN i sqrt(3)/2, W6^2 = -1/2 + SIGN i sqrt(3)/2
    b1[1] = cmul(b1[1], make_double2(0.5, SIGN * h));
    b1[2] = cmul(b1[2], make_double2(-0.5, SIGN * h));
    Dft<3, SIGN>::run(b0);
    Dft<3, SIGN>::run(b1);
    v[0] = b0[0];
    v[2] = b0[1];
    v[4] = b0[2];
    v[1] = b1[0];
    v[3] = b1[1];
    v[5] = b1[2];
  }
};

// 12 = 4 x 3: n = 3 n1 + n2, k = k1 + 4 k2; DFT4 over n1, twiddle W12^(n2 k1), DFT3 over n2.
template <int SIGN> struct Dft<12, SIGN> {
  static __host__ __device__ __forceinline__ void run(double2* v) {
    constexpr double h = 0.86602540378443864676372317075294;  // sqrt(3)/2
    double2 b[3][4];
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
      double2 t[4] = {v[n2], v[n2 + 3], v[n2 + 6], v[n2 + 9]};
      Dft<4, SIGN>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) b[n2][k1] = t[k1];
    }
    // W12^m, m = n2 k1 in {1, 2, 3} (n2 = 1) and {2, 4, 6} (n2 = 2)
    b[1][1] = cmul(b[1][1], make_double2(h, SIGN * 0.5));
    b[1][2] = cmul(b[1][2], make_double2(0.5, SIGN * h));
    b[1][3] = cmul_i<SIGN>(b[1][3]);
    b[2][1] = cmul(b[2][1], make_double2(0.5, SIGN * h));
    b[2][2] = cmul(b[2][2], make_double2(-0.5, SIGN * h));
    b[2][3] = make_double2(-b[2][3].x, -b[2][3].y);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      double2 t[3] = {b[0][k1], b[1][k1], b[2][k1]};
      Dft<3, SIGN>::run(t);
      v[k1] = t[0];
      v[k1 + 4] = t[1];
      v[k1 + 8] = t[2];
    }
  }
};


// ---------------------------------------------------------------- LDS row layout
// Rows of L complex values are stored with one pad slot after every 8 values (so
// the stride-R writes of a Stockham stage with Ns = 1 fall in distinct banks) and
// one extra slot per row (so column-wise accesses across rows do too); PAD = false
// (L + 1 entries per row) for kernels whose access pattern measured faster unpadded.
template <int L, bool PAD = true> struct Row {
  static constexpr int LS = PAD ? L + L / 8 + 1 : L + 1;
  static __device__ __forceinline__ int at(int i) { return PAD ? i + (i >> 3) : i; }
};

// ---------------------------------------------------------------- Stockham batch
// B transforms of length L, transform b at buf[b*LS .. b*LS+L), NT threads.
template <int L, int B, int NT, int LS, int SIGN, int S, int NS, bool PAD>
__device__ __forceinline__ void fft_stages(double2* buf, const double2* __restrict__ tw, int tid) {
  if constexpr (S < Plan<L>::n) {
    constexpr int R = Plan<L>::r[S];
    constexpr int NBF = L / R;
    constexpr int TOT = B * NBF;
    constexpr int IT = (TOT + NT - 1) / NT;
    constexpr int TWS = L / (NS * R);
    double2 v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int j = tid + it * NT;
      if ((TOT % NT) == 0 || j < TOT) {
        const int b = j / NBF;
        const int jj = j - b * NBF;
        const double2* src = buf + b * LS;
#pragma unroll
        for (int r = 0; r < R; ++r) v[it][r] = src[Row<L, PAD>::at(jj + r * NBF)];
        if constexpr (NS > 1) {
          const int k = jj % NS;
          // one table load per butterfly; the other twiddles are its running powers
          // (at most R - 2 products, a few ulp; two live values instead of R - 1)
          double2 w1 = tw[k * TWS];
          if constexpr (SIGN > 0) w1.y = -w1.y;
          double2 wc = w1;
#pragma unroll
          for (int r = 1; r < R; ++r) {
            v[it][r] = cmul(v[it][r], wc);
            if (r + 1 < R) wc = cmul(wc, w1);
          }
        }
        Dft<R, SIGN>::run(v[it]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int j = tid + it * NT;
      if ((TOT % NT) == 0 || j < TOT) {
        const int b = j / NBF;
        const int jj = j - b * NBF;
        const int k = jj % NS;
        double2* dst = buf + b * LS;
        const int base = (jj / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[Row<L, PAD>::at(base + r * NS)] = v[it][r];
      }
    }
    __syncthreads();
    fft_stages<L, B, NT, LS, SIGN, S + 1, NS * R, PAD>(buf, tw, tid);
  }
}

// Run the batch.  Caller must __syncthreads() before (data staged) — the routine
// ends with a barrier, so results are visible on return.
template <int L, int B, int NT, int LS, int SIGN, bool PAD = true>
__device__ __forceinline__ void fft_batch(double2* buf, const double2* __restrict__ tw, int tid) {
  fft_stages<L, B, NT, LS, SIGN, 0, 1, PAD>(buf, tw, tid);
}

// ------------------------------------------------------- register-resident batch
// Radix sequences whose first and last stages are radix E, for batches where each
// thread holds exactly E entries (B L = E NT).
template <int L, int E> struct RegPlan { static constexpr bool ok = false; };
#define PSX_REGPLAN(LEN, E, ...)                                  \
  template <> struct RegPlan<LEN, E> {                            \
    static constexpr bool ok = true;                              \
    static constexpr int r[] = {__VA_ARGS__};                     \
    static constexpr int n = sizeof(r) / sizeof(int);             \
  };
PSX_REGPLAN(128, 4, 4, 8, 4)
PSX_REGPLAN(256, 4, 4, 4, 4, 4)
PSX_REGPLAN(512, 4, 4, 8, 4, 4)
PSX_REGPLAN(256, 8, 8, 4, 8)
PSX_REGPLAN(512, 8, 8, 8, 8)
#undef PSX_REGPLAN

// One middle stage S of fft_rows_reg: read `in`, twiddle, radix R, write `out`, barrier;
// the next stage reads `out`.  In place (in == out), a barrier separates the reads from
// the writes.
template <int L, int E, int B, int NT, int LS, int SIGN, bool INPLACE, int S, int NS>
__device__ __forceinline__ void reg_mid(double2* in, double2* out, const double2* __restrict__ tw, int tid) {
  using P = RegPlan<L, E>;
  if constexpr (S < P::n - 1) {
    constexpr int R = P::r[S];
    constexpr int NBF = L / R, TOT = B * NBF, IT = (TOT + NT - 1) / NT, TWS = L / (NS * R);
    double2 v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int j = tid + it * NT;
      if ((TOT % NT) == 0 || j < TOT) {
        const int b = j / NBF, jj = j - b * NBF;
#pragma unroll
        for (int r = 0; r < R; ++r) v[it][r] = in[b * LS + Row<L>::at(jj + r * NBF)];
        double2 w1 = tw[(jj % NS) * TWS];
        if constexpr (SIGN > 0) w1.y = -w1.y;
        double2 wc = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[it][r] = cmul(v[it][r], wc);
          if (r + 1 < R) wc = cmul(wc, w1);
        }
        Dft<R, SIGN>::run(v[it]);
      }
    }
    if constexpr (INPLACE) __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int j = tid + it * NT;
      if ((TOT % NT) == 0 || j < TOT) {
        const int b = j / NBF, jj = j - b * NBF, k = jj % NS;
        const int base = (jj / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) out[b * LS + Row<L>::at(base + r * NS)] = v[it][r];
      }
    }
    __syncthreads();
    reg_mid<L, E, B, NT, LS, SIGN, INPLACE, S + 1, NS * R>(out, in, tw, tid);
  }
}

// Stockham transforms of B rows of length L (B L = E NT) with the data in registers on
// entry and exit: thread t owns row b = t % B and holds its entries
// n = t / B + (L / E) e, e < E -- on entry the inputs, on return the outputs at the
// same indices (the layout of a row-batched load of B contiguous columns, element
// t + e NT split as n = e / B, b = e % B).  The radix-E first and last stages run on
// the registers; the middle stages pass through LDS (B rows of stride LS each, Row<L>
// padding), one barrier per stage.  Two buffers x0, x1 are used in turn; no barrier is
// needed before the first stage, which writes x0, provided the previous call's last
// stage read x1 -- so consecutive calls pass the buffers swapped when the plan has an
// even number of stages (fft_rows_reg_swap).  INPLACE: one buffer (x1 = x0), a barrier
// between each middle stage's reads and writes, and one on entry.
template <int L, int E, int B, int NT, int LS, int SIGN, bool INPLACE = false>
__device__ __forceinline__ void fft_rows_reg(double2 (&v)[E], double2* x0, double2* x1,
                                             const double2* __restrict__ tw, int tid) {
  using P = RegPlan<L, E>;
  static_assert(P::ok && B * L == E * NT, "E entries per thread");
  static_assert(P::r[0] == E && P::r[P::n - 1] == E, "radix-E first and last stages");
  constexpr int Q = L / E;
  const int b = tid % B, jj = tid / B;
  if constexpr (INPLACE) {
    x1 = x0;
    __syncthreads();  // the previous call's last-stage reads
  }
  // first stage (NS = 1): no twiddle; outputs at E jj + r
  Dft<E, SIGN>::run(v);
#pragma unroll
  for (int r = 0; r < E; ++r) x0[b * LS + Row<L>::at(E * jj + r)] = v[r];
  __syncthreads();
  reg_mid<L, E, B, NT, LS, SIGN, INPLACE, 1, E>(x0, x1, tw, tid);
  // last stage (NS = Q): inputs jj + Q r from the buffer the last middle stage wrote,
  // twiddle w_L^{r jj}, outputs jj + Q r
  const double2* src = (P::n % 2 == 0) ? x0 : x1;
#pragma unroll
  for (int r = 0; r < E; ++r) v[r] = src[b * LS + Row<L>::at(jj + Q * r)];
  double2 w1 = tw[jj];
  if constexpr (SIGN > 0) w1.y = -w1.y;
  double2 wc = w1;
#pragma unroll
  for (int r = 1; r < E; ++r) {
    v[r] = cmul(v[r], wc);
    if (r + 1 < E) wc = cmul(wc, w1);
  }
  Dft<E, SIGN>::run(v);
}
// whether consecutive two-buffer fft_rows_reg calls swap x0 and x1 (the last stage read x0)
template <int L, int E> constexpr bool fft_rows_reg_swap() { return RegPlan<L, E>::n % 2 == 0; }

}  // namespace psx
