// Fused z-pass instantiations, mode Z_SMAG.
#include "launch.hpp"

namespace psx {

constexpr int zp_smag(int L) { return z_P<Z_SMAG>(L); }

int zpass_positions_smag(int L) {
  switch (L) {
    case 24: return zp_smag(24);
    case 48: return zp_smag(48);
    case 72: return zp_smag(72);
    case 96: return zp_smag(96);
    case 144: return zp_smag(144);
    case 192: return zp_smag(192);
    case 288: return zp_smag(288);
    case 384: return zp_smag(384);
    default: return 0;
  }
}

hipError_t launch_zpass_smag(int L, const ZArgs& a, const double2* tw, hipStream_t s) {
  switch (L) {
    case 24: {
      constexpr int P = zp_smag(24);
      hipLaunchKernelGGL((zpass<24, P, Z_SMAG, z_NT<Z_SMAG>(24)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_SMAG>(24)), 0, s, a, tw);
      break;
    }
    case 48: {
      constexpr int P = zp_smag(48);
      hipLaunchKernelGGL((zpass<48, P, Z_SMAG, z_NT<Z_SMAG>(48)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_SMAG>(48)), 0, s, a, tw);
      break;
    }
    case 72: {
      constexpr int P = zp_smag(72);
      hipLaunchKernelGGL((zpass<72, P, Z_SMAG, z_NT<Z_SMAG>(72)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_SMAG>(72)), 0, s, a, tw);
      break;
    }
    case 96: {
      constexpr int P = zp_smag(96);
      hipLaunchKernelGGL((zpass<96, P, Z_SMAG, z_NT<Z_SMAG>(96)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_SMAG>(96)), 0, s, a, tw);
      break;
    }
    case 144: {
      constexpr int P = zp_smag(144);
      hipLaunchKernelGGL((zpass<144, P, Z_SMAG, z_NT<Z_SMAG>(144)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_SMAG>(144)), 0, s, a, tw);
      break;
    }
    case 192: {
      constexpr int P = zp_smag(192);
      hipLaunchKernelGGL((zpass<192, P, Z_SMAG, z_NT<Z_SMAG>(192)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_SMAG>(192)), 0, s, a, tw);
      break;
    }
    case 288: {
      constexpr int P = zp_smag(288);
      hipLaunchKernelGGL((zpass<288, P, Z_SMAG, z_NT<Z_SMAG>(288)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_SMAG>(288)), 0, s, a, tw);
      break;
    }
    case 384: {
      constexpr int P = zp_smag(384);
      hipLaunchKernelGGL((zpass<384, P, Z_SMAG, z_NT<Z_SMAG>(384)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_SMAG>(384)), 0, s, a, tw);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace psx
