// Fused z-pass instantiations, mode Z_DT.
#include "launch.hpp"

namespace psx {

constexpr int zp_dt(int L) { return z_P<Z_DT>(L); }

int zpass_positions_dt(int L) {
  switch (L) {
    case 16: return zp_dt(16);
    case 32: return zp_dt(32);
    case 48: return zp_dt(48);
    case 64: return zp_dt(64);
    case 96: return zp_dt(96);
    case 128: return zp_dt(128);
    case 192: return zp_dt(192);
    case 256: return zp_dt(256);
    case 384: return zp_dt(384);
    case 512: return zp_dt(512);
    default: return 0;
  }
}

hipError_t launch_zpass_dt(int L, const ZArgs& a, const double2* tw, hipStream_t s) {
  switch (L) {
    case 16: {
      constexpr int P = zp_dt(16);
      hipLaunchKernelGGL((zpass<16, P, Z_DT, z_NT<Z_DT>(16)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DT>(16)), 0, s, a, tw);
      break;
    }
    case 32: {
      constexpr int P = zp_dt(32);
      hipLaunchKernelGGL((zpass<32, P, Z_DT, z_NT<Z_DT>(32)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DT>(32)), 0, s, a, tw);
      break;
    }
    case 48: {
      constexpr int P = zp_dt(48);
      hipLaunchKernelGGL((zpass<48, P, Z_DT, z_NT<Z_DT>(48)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DT>(48)), 0, s, a, tw);
      break;
    }
    case 64: {
      constexpr int P = zp_dt(64);
      hipLaunchKernelGGL((zpass<64, P, Z_DT, z_NT<Z_DT>(64)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DT>(64)), 0, s, a, tw);
      break;
    }
    case 96: {
      constexpr int P = zp_dt(96);
      hipLaunchKernelGGL((zpass<96, P, Z_DT, z_NT<Z_DT>(96)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DT>(96)), 0, s, a, tw);
      break;
    }
    case 128: {
      constexpr int P = zp_dt(128);
      hipLaunchKernelGGL((zpass<128, P, Z_DT, z_NT<Z_DT>(128)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DT>(128)), 0, s, a, tw);
      break;
    }
    case 192: {
      constexpr int P = zp_dt(192);
      hipLaunchKernelGGL((zpass<192, P, Z_DT, z_NT<Z_DT>(192)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DT>(192)), 0, s, a, tw);
      break;
    }
    case 256: {
      constexpr int P = zp_dt(256);
      hipLaunchKernelGGL((zpass<256, P, Z_DT, z_NT<Z_DT>(256)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DT>(256)), 0, s, a, tw);
      break;
    }
    case 384: {
      constexpr int P = zp_dt(384);
      hipLaunchKernelGGL((zpass<384, P, Z_DT, z_NT<Z_DT>(384)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DT>(384)), 0, s, a, tw);
      break;
    }
    case 512: {
      constexpr int P = zp_dt(512);
      hipLaunchKernelGGL((zpass<512, P, Z_DT, z_NT<Z_DT>(512)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DT>(512)), 0, s, a, tw);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace psx
