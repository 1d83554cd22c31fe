// Host-side check of the FFT building blocks (radix butterflies, plans, the
// Stockham stage order and twiddle indexing): the same Dft<R> code the kernels
// run, driven by a sequential restatement of one Stockham stage.  Built into
// libpsx_hosttest.so for the CPU test suite; no GPU involved.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fft_core.hpp"

namespace {
template <int L, int SIGN, int S, int NS>
void stages(std::vector<double2>& x, const std::vector<double2>& tw) {
  if constexpr (S < psx::Plan<L>::n) {
    constexpr int R = psx::Plan<L>::r[S];
    constexpr int NBF = L / R;
    constexpr int TWS = L / (NS * R);
    std::vector<double2> y(L);
    for (int jj = 0; jj < NBF; ++jj) {
      double2 v[R];
      const int k = jj % NS;
      for (int r = 0; r < R; ++r) v[r] = x[jj + r * NBF];
      if (NS > 1) {
        for (int r = 1; r < R; ++r) {
          double2 w = tw[k * r * TWS];
          if (SIGN > 0) w.y = -w.y;
          v[r] = psx::cmul(v[r], w);
        }
      }
      psx::Dft<R, SIGN>::run(v);
      const int base = (jj / NS) * NS * R + k;
      for (int r = 0; r < R; ++r) y[base + r * NS] = v[r];
    }
    x = y;
    stages<L, SIGN, S + 1, NS * R>(x, tw);
  }
}

template <int L>
int run(int sign, double* data) {
  std::vector<double2> tw(L), x(L);
  for (int m = 0; m < L; ++m) {
    const double a = 2.0 * 3.14159265358979323846264338327950288 * m / L;
    tw[m] = make_double2(std::cos(a), -std::sin(a));
  }
  for (int i = 0; i < L; ++i) x[i] = make_double2(data[2 * i], data[2 * i + 1]);
  if (sign < 0)
    stages<L, -1, 0, 1>(x, tw);
  else
    stages<L, +1, 0, 1>(x, tw);
  for (int i = 0; i < L; ++i) {
    data[2 * i] = x[i].x;
    data[2 * i + 1] = x[i].y;
  }
  return 0;
}

template <int R>
int dft(int sign, double* data) {
  double2 v[R];
  for (int i = 0; i < R; ++i) v[i] = make_double2(data[2 * i], data[2 * i + 1]);
  if (sign < 0)
    psx::Dft<R, -1>::run(v);
  else
    psx::Dft<R, +1>::run(v);
  for (int i = 0; i < R; ++i) {
    data[2 * i] = v[i].x;
    data[2 * i + 1] = v[i].y;
  }
  return 0;
}
}  // namespace

// One radix-R butterfly, as the one-wavefront z pass's register step runs it.
extern "C" int psx_host_dft(int R, int sign, double* data) {
  switch (R) {
    case 2: return dft<2>(sign, data);
    case 3: return dft<3>(sign, data);
    case 4: return dft<4>(sign, data);
    case 6: return dft<6>(sign, data);
    case 8: return dft<8>(sign, data);
    case 12: return dft<12>(sign, data);
    default: return -1;
  }
}

extern "C" int psx_host_fft(int L, int sign, double* data) {
#define PSX_CASE(LL)                                     \
  case LL:                                               \
    return run<LL>(sign, data);
  switch (L) {
    PSX_CASE(8) PSX_CASE(12) PSX_CASE(16) PSX_CASE(24) PSX_CASE(32) PSX_CASE(36) PSX_CASE(48) PSX_CASE(64)
    PSX_CASE(72) PSX_CASE(96) PSX_CASE(128) PSX_CASE(144) PSX_CASE(192) PSX_CASE(256) PSX_CASE(288)
    PSX_CASE(384) PSX_CASE(512) PSX_CASE(576) PSX_CASE(768)
    default:
      return -1;
  }
#undef PSX_CASE
}
