// Launch one z-pass instantiation: the one-wavefront kernel (zwave.hpp) for
// lengths 64 * {1,2,3,4,6,8,12}, the lane-group kernel (zsub.hpp) for the 3/2-rule
// products at 48, 72 and 96, the block-wide Stockham kernel otherwise (the fused
// learned closure has only the one-wavefront form).
#pragma once
#include "launch.hpp"
#include "zwave.hpp"
#include "zsub.hpp"

namespace psx {

// The 3/2-rule modes with L/3 half-spectrum entries in and out (the padded length a
// multiple of 192) run an instantiation with that length fixed at compile time (not
// the fused closure at 768, which spills more that way).
template <int L, int MODE> constexpr bool zw_third_mode() {
  return ZWave<L>::ok && L % 192 == 0 &&
         (MODE == Z_DNS || MODE == Z_SMAG || (MODE == Z_LRNF && L < 768));
}
template <int L, int MODE> bool zw_third(const ZArgs& a) {
  return a.Hin0 == L / 3 && (a.nin0 >= ZFields<MODE>::n || a.Hin1 == L / 3) && a.Hout == L / 3;
}

template <int L, int MODE>
hipError_t zpass_launch(const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  constexpr int P = z_P<MODE>(L);
  const dim3 grid((unsigned)(a.npos / P));
  if constexpr (MODE == Z_LRNF) {
    // the fused closure exists as the one-wavefront and the lane-group kernels only
    if constexpr (ZSub<L>::ok) {
      constexpr int NW = zs_waves<L, MODE>(), S = ZSubGroups<L, NW>::S;
      if (zw_third<L, MODE>(a) && a.npos % S == 0) {
        hipLaunchKernelGGL((zsub<L, MODE, L / 3>), dim3((unsigned)(a.npos / S)), dim3(64 * NW), 0, s, a, tw);
        return hipGetLastError();
      }
      return hipErrorInvalidValue;
    } else if constexpr (ZWave<L>::ok) {
      if constexpr (zw_third_mode<L, MODE>()) {
        if (zw_third<L, MODE>(a)) {
          hipLaunchKernelGGL((zwave<L, P, MODE, L / 3>), grid, dim3(64), 0, s, a, tw);
          return hipGetLastError();
        }
      }
      hipLaunchKernelGGL((zwave<L, P, MODE>), grid, dim3(64), 0, s, a, tw);
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;
    }
  } else {
    if constexpr (ZSub<L>::ok && (MODE == Z_DNS || MODE == Z_SMAG)) {
      // the short lengths of small grids: a lane group per position
      constexpr int NW = zs_waves<L, MODE>(), S = ZSubGroups<L, NW>::S;
      if (variant == 0 && zw_third<L, MODE>(a) && a.npos % S == 0) {
        hipLaunchKernelGGL((zsub<L, MODE, L / 3>), dim3((unsigned)(a.npos / S)), dim3(64 * NW), 0, s, a, tw);
        return hipGetLastError();
      }
    }
    if constexpr (ZWave<L>::ok) {
      if (variant == 0) {
        if constexpr (zw_third_mode<L, MODE>()) {
          if (zw_third<L, MODE>(a)) {
            hipLaunchKernelGGL((zwave<L, P, MODE, L / 3>), grid, dim3(64), 0, s, a, tw);
            return hipGetLastError();
          }
        }
        hipLaunchKernelGGL((zwave<L, P, MODE>), grid, dim3(64), 0, s, a, tw);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL((zpass<L, P, MODE, z_NT<MODE>(L)>), grid, dim3(z_NT<MODE>(L)), 0, s, a, tw);
    return hipGetLastError();
  }
}

}  // namespace psx
