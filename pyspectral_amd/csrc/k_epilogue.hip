// x-forward + spectral-epilogue instantiations: (transform length, spectral length).
#include "launch.hpp"

namespace psx {

// 3/2 rule pairs (M, N) and 2x pairs (2N, N)
#define PSX_PAIRS32(X) X(24, 16) X(48, 32) X(72, 48) X(96, 64) X(144, 96) X(192, 128) X(288, 192) X(384, 256) X(576, 384) X(768, 512)
#define PSX_PAIRS2X(X) X(32, 16) X(64, 32) X(96, 48) X(128, 64) X(192, 96) X(256, 128) X(384, 192) X(512, 256) X(768, 384)

unsigned epilogue_blocks(int L, const EpiArgs& a) {
  const int T = epi_T(L);
  unsigned main_blocks = (unsigned)(a.N2 * ((a.H + T - 1) / T));
  long long rest = (long long)a.N1 * a.N2 * (a.Nc - a.H);
  unsigned extra = 0;
  if (rest > 0) {
    long long b = (rest + NT - 1) / NT;
    extra = (unsigned)(b < 1024 ? b : 1024);
  }
  return main_blocks + extra;
}

hipError_t launch_epilogue(int L, int ns1, int epi, bool w0, const EpiArgs& a, const double2* tw, hipStream_t s) {
  const dim3 grid(epilogue_blocks(L, a), (unsigned)(a.members > 1 ? a.members : 1));
  const int key = L * 4096 + ns1;
  if (epi == E_DNS) {
    switch (key) {
#define PSX_CASE(LL, NN)                                                                                   \
  case LL * 4096 + NN:                                                                                     \
    hipLaunchKernelGGL((xforward_epilogue<LL, NN, epi_T(LL), E_DNS, false>), grid, dim3(NT), 0, s, a, tw); \
    return hipGetLastError();
      PSX_PAIRS32(PSX_CASE)
#undef PSX_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  if (epi == E_SMAG) {
    if (w0) {
      switch (key) {
#define PSX_CASE(LL, NN)                                                                                   \
  case LL * 4096 + NN:                                                                                     \
    hipLaunchKernelGGL((xforward_epilogue<LL, NN, epi_T(LL), E_SMAG, true>), grid, dim3(NT), 0, s, a, tw); \
    return hipGetLastError();
        PSX_PAIRS32(PSX_CASE)
#undef PSX_CASE
        default:
          return hipErrorInvalidValue;
      }
    }
    switch (key) {
#define PSX_CASE(LL, NN)                                                                                    \
  case LL * 4096 + NN:                                                                                      \
    hipLaunchKernelGGL((xforward_epilogue<LL, NN, epi_T(LL), E_SMAG, false>), grid, dim3(NT), 0, s, a, tw); \
    return hipGetLastError();
      PSX_PAIRS32(PSX_CASE)
#undef PSX_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  if (epi == E_SMAGF) {
    if (w0) return hipErrorInvalidValue;  // the folded fields hold no separate stress
    switch (key) {
#define PSX_CASE(LL, NN)                                                                                     \
  case LL * 4096 + NN:                                                                                       \
    hipLaunchKernelGGL((xforward_epilogue<LL, NN, epi_T(LL), E_SMAGF, false>), grid, dim3(NT), 0, s, a, tw); \
    return hipGetLastError();
      PSX_PAIRS32(PSX_CASE)
#undef PSX_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  if (epi == E_SMAG9) {
    if (w0) return hipErrorInvalidValue;  // the folded fields hold no separate stress
    switch (key) {
#define PSX_CASE(LL, NN)                                                                                     \
  case LL * 4096 + NN:                                                                                       \
    hipLaunchKernelGGL((xforward_epilogue<LL, NN, epi_T(LL), E_SMAG9, false>), grid, dim3(NT), 0, s, a, tw); \
    return hipGetLastError();
      PSX_PAIRS32(PSX_CASE)
#undef PSX_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  if (epi == E_TAU2) {
    switch (key) {
#define PSX_CASE(LL, NN)                                                                                    \
  case LL * 4096 + NN:                                                                                      \
    hipLaunchKernelGGL((xforward_epilogue<LL, NN, epi_T(LL), E_TAU2, false>), grid, dim3(NT), 0, s, a, tw); \
    return hipGetLastError();
      PSX_PAIRS2X(PSX_CASE)
#undef PSX_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  return hipErrorInvalidValue;
}

unsigned sgs_blocks(int L, const SgsArgs& a) {
  const int T = epi_T(L);
  unsigned main_blocks = (unsigned)(a.N2 * ((a.H + T - 1) / T));
  long long rest = (long long)a.N1 * a.N2 * (a.Nc - a.H);
  long long b = (rest + NT - 1) / NT;
  return main_blocks + (unsigned)(b < 1 ? 0 : (b < 1024 ? b : 1024));
}

hipError_t launch_xforward_sgs(int L, int ns1, const SgsArgs& a, const double2* tw, hipStream_t s) {
  const dim3 grid(sgs_blocks(L, a));
  switch (L * 4096 + ns1) {
#define PSX_CASE(LL, NN)                                                                 \
  case LL * 4096 + NN:                                                                   \
    hipLaunchKernelGGL((xforward_sgs<LL, NN, epi_T(LL)>), grid, dim3(NT), 0, s, a, tw); \
    return hipGetLastError();
    PSX_PAIRS32(PSX_CASE)
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace psx
