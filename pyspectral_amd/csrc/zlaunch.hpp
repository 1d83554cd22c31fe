// Launch one z-pass instantiation: the one-wavefront kernel (zwave.hpp) for
// lengths 64 * {1,2,3,4,6,8,12}, the block-wide Stockham kernel otherwise (the fused
// learned closure has only the one-wavefront form).
#pragma once
#include "launch.hpp"
#include "zwave.hpp"

#ifndef PSX_ZWAVE
#define PSX_ZWAVE 1
#endif

namespace psx {

template <int L, int MODE>
hipError_t zpass_launch(const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  constexpr int P = z_P<MODE>(L);
  const dim3 grid((unsigned)(a.npos / P));
  if constexpr (MODE == Z_LRNF) {
    // the fused closure exists as the one-wavefront kernel only
    if constexpr (ZWave<L>::ok) {
      hipLaunchKernelGGL((zwave<L, P, MODE>), grid, dim3(64), 0, s, a, tw);
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;
    }
  } else {
    if constexpr (PSX_ZWAVE != 0 && ZWave<L>::ok) {
      if (variant == 0) {
        hipLaunchKernelGGL((zwave<L, P, MODE>), grid, dim3(64), 0, s, a, tw);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL((zpass<L, P, MODE, z_NT<MODE>(L)>), grid, dim3(z_NT<MODE>(L)), 0, s, a, tw);
    return hipGetLastError();
  }
}

}  // namespace psx
