// Fused z-pass instantiations, mode Z_TM2.
#include "launch.hpp"

namespace psx {

constexpr int zp_tm2(int L) { return z_P<Z_TM2>(L); }

int zpass_positions_tm2(int L) {
  switch (L) {
    case 32: return zp_tm2(32);
    case 64: return zp_tm2(64);
    case 96: return zp_tm2(96);
    case 128: return zp_tm2(128);
    case 192: return zp_tm2(192);
    case 256: return zp_tm2(256);
    case 384: return zp_tm2(384);
    case 512: return zp_tm2(512);
    case 768: return zp_tm2(768);
    default: return 0;
  }
}

hipError_t launch_zpass_tm2(int L, const ZArgs& a, const double2* tw, hipStream_t s) {
  switch (L) {
    case 32: {
      constexpr int P = zp_tm2(32);
      hipLaunchKernelGGL((zpass<32, P, Z_TM2, z_NT<Z_TM2>(32)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_TM2>(32)), 0, s, a, tw);
      break;
    }
    case 64: {
      constexpr int P = zp_tm2(64);
      hipLaunchKernelGGL((zpass<64, P, Z_TM2, z_NT<Z_TM2>(64)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_TM2>(64)), 0, s, a, tw);
      break;
    }
    case 96: {
      constexpr int P = zp_tm2(96);
      hipLaunchKernelGGL((zpass<96, P, Z_TM2, z_NT<Z_TM2>(96)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_TM2>(96)), 0, s, a, tw);
      break;
    }
    case 128: {
      constexpr int P = zp_tm2(128);
      hipLaunchKernelGGL((zpass<128, P, Z_TM2, z_NT<Z_TM2>(128)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_TM2>(128)), 0, s, a, tw);
      break;
    }
    case 192: {
      constexpr int P = zp_tm2(192);
      hipLaunchKernelGGL((zpass<192, P, Z_TM2, z_NT<Z_TM2>(192)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_TM2>(192)), 0, s, a, tw);
      break;
    }
    case 256: {
      constexpr int P = zp_tm2(256);
      hipLaunchKernelGGL((zpass<256, P, Z_TM2, z_NT<Z_TM2>(256)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_TM2>(256)), 0, s, a, tw);
      break;
    }
    case 384: {
      constexpr int P = zp_tm2(384);
      hipLaunchKernelGGL((zpass<384, P, Z_TM2, z_NT<Z_TM2>(384)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_TM2>(384)), 0, s, a, tw);
      break;
    }
    case 512: {
      constexpr int P = zp_tm2(512);
      hipLaunchKernelGGL((zpass<512, P, Z_TM2, z_NT<Z_TM2>(512)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_TM2>(512)), 0, s, a, tw);
      break;
    }
    case 768: {
      constexpr int P = zp_tm2(768);
      hipLaunchKernelGGL((zpass<768, P, Z_TM2, z_NT<Z_TM2>(768)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_TM2>(768)), 0, s, a, tw);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace psx
