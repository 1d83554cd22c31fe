// 2x-grid PL pass instantiations (t-model family, first round).
#include "launch.hpp"

namespace psx {

#define PSX_LENGTHS2X(X) X(32) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

unsigned pl_pass_blocks(int L, const PLArgs& a) {
  const int T = pl_T(L);
  return (unsigned)(a.P2 * ((a.C + T - 1) / T));
}

hipError_t launch_pl_pass(int L, bool dyn, const PLArgs& a, const double2* tw, hipStream_t s) {
  const dim3 grid(pl_pass_blocks(L, a));
  if (dyn) {
    switch (L) {
#define PSX_CASE(LL)                                                                           \
  case LL:                                                                                     \
    hipLaunchKernelGGL((pl_pass<LL, pl_T(LL), true, pl_NT(LL)>), grid, dim3(pl_NT(LL)), 0, s, a, tw);           \
    return hipGetLastError();
      PSX_LENGTHS2X(PSX_CASE)
#undef PSX_CASE
      default:
        return hipErrorInvalidValue;
    }
  }
  switch (L) {
#define PSX_CASE(LL)                                                                           \
  case LL:                                                                                     \
    hipLaunchKernelGGL((pl_pass<LL, pl_T(LL), false, pl_NT(LL)>), grid, dim3(pl_NT(LL)), 0, s, a, tw);          \
    return hipGetLastError();
    PSX_LENGTHS2X(PSX_CASE)
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace psx
