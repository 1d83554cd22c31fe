// Pencil-pass instantiations (x and y axes) for every supported length.
#include "launch.hpp"

namespace psx {

#define PSX_LENGTHS(X) \
  X(16) X(24) X(32) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(576) X(768)

bool length_supported(int L) {
  switch (L) {
#define PSX_CASE(LL) case LL:
    PSX_LENGTHS(PSX_CASE)
#undef PSX_CASE
    return true;
    default:
      return false;
  }
}

static inline unsigned pencil_blocks(int L, const PencilArgs& a) {
  const int T = pencil_T(L);
  return (unsigned)(a.A * ((a.C + T - 1) / T));
}

hipError_t launch_pencil_inverse(int L, const PencilArgs& a, const double2* tw, int njobs, hipStream_t s) {
  const dim3 grid(pencil_blocks(L, a), (unsigned)njobs);
  switch (L) {
#define PSX_CASE(LL)                                                                  \
  case LL:                                                                            \
    hipLaunchKernelGGL((pencil_inverse<LL, pencil_T(LL)>), grid, dim3(NT), 0, s, a, tw); \
    break;
    PSX_LENGTHS(PSX_CASE)
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_pencil_forward(int L, const PencilArgs& a, const double2* tw, int njobs, hipStream_t s) {
  const dim3 grid(pencil_blocks(L, a), (unsigned)njobs);
  switch (L) {
#define PSX_CASE(LL)                                                                  \
  case LL:                                                                            \
    hipLaunchKernelGGL((pencil_forward<LL, pencil_T(LL)>), grid, dim3(NT), 0, s, a, tw); \
    break;
    PSX_LENGTHS(PSX_CASE)
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace psx
