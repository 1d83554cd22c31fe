// Fused z-pass instantiations, mode Z_DNS.
#include "launch.hpp"

namespace psx {

constexpr int zp_dns(int L) { return z_P<Z_DNS>(L); }

int zpass_positions_dns(int L) {
  switch (L) {
    case 24: return zp_dns(24);
    case 32: return zp_dns(32);
    case 48: return zp_dns(48);
    case 64: return zp_dns(64);
    case 72: return zp_dns(72);
    case 96: return zp_dns(96);
    case 128: return zp_dns(128);
    case 144: return zp_dns(144);
    case 192: return zp_dns(192);
    case 256: return zp_dns(256);
    case 288: return zp_dns(288);
    case 384: return zp_dns(384);
    case 512: return zp_dns(512);
    case 576: return zp_dns(576);
    case 768: return zp_dns(768);
    default: return 0;
  }
}

hipError_t launch_zpass_dns(int L, const ZArgs& a, const double2* tw, hipStream_t s) {
  switch (L) {
    case 24: {
      constexpr int P = zp_dns(24);
      hipLaunchKernelGGL((zpass<24, P, Z_DNS, z_NT<Z_DNS>(24)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(24)), 0, s, a, tw);
      break;
    }
    case 32: {
      constexpr int P = zp_dns(32);
      hipLaunchKernelGGL((zpass<32, P, Z_DNS, z_NT<Z_DNS>(32)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(32)), 0, s, a, tw);
      break;
    }
    case 48: {
      constexpr int P = zp_dns(48);
      hipLaunchKernelGGL((zpass<48, P, Z_DNS, z_NT<Z_DNS>(48)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(48)), 0, s, a, tw);
      break;
    }
    case 64: {
      constexpr int P = zp_dns(64);
      hipLaunchKernelGGL((zpass<64, P, Z_DNS, z_NT<Z_DNS>(64)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(64)), 0, s, a, tw);
      break;
    }
    case 72: {
      constexpr int P = zp_dns(72);
      hipLaunchKernelGGL((zpass<72, P, Z_DNS, z_NT<Z_DNS>(72)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(72)), 0, s, a, tw);
      break;
    }
    case 96: {
      constexpr int P = zp_dns(96);
      hipLaunchKernelGGL((zpass<96, P, Z_DNS, z_NT<Z_DNS>(96)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(96)), 0, s, a, tw);
      break;
    }
    case 128: {
      constexpr int P = zp_dns(128);
      hipLaunchKernelGGL((zpass<128, P, Z_DNS, z_NT<Z_DNS>(128)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(128)), 0, s, a, tw);
      break;
    }
    case 144: {
      constexpr int P = zp_dns(144);
      hipLaunchKernelGGL((zpass<144, P, Z_DNS, z_NT<Z_DNS>(144)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(144)), 0, s, a, tw);
      break;
    }
    case 192: {
      constexpr int P = zp_dns(192);
      hipLaunchKernelGGL((zpass<192, P, Z_DNS, z_NT<Z_DNS>(192)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(192)), 0, s, a, tw);
      break;
    }
    case 256: {
      constexpr int P = zp_dns(256);
      hipLaunchKernelGGL((zpass<256, P, Z_DNS, z_NT<Z_DNS>(256)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(256)), 0, s, a, tw);
      break;
    }
    case 288: {
      constexpr int P = zp_dns(288);
      hipLaunchKernelGGL((zpass<288, P, Z_DNS, z_NT<Z_DNS>(288)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(288)), 0, s, a, tw);
      break;
    }
    case 384: {
      constexpr int P = zp_dns(384);
      hipLaunchKernelGGL((zpass<384, P, Z_DNS, z_NT<Z_DNS>(384)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(384)), 0, s, a, tw);
      break;
    }
    case 512: {
      constexpr int P = zp_dns(512);
      hipLaunchKernelGGL((zpass<512, P, Z_DNS, z_NT<Z_DNS>(512)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(512)), 0, s, a, tw);
      break;
    }
    case 576: {
      constexpr int P = zp_dns(576);
      hipLaunchKernelGGL((zpass<576, P, Z_DNS, z_NT<Z_DNS>(576)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(576)), 0, s, a, tw);
      break;
    }
    case 768: {
      constexpr int P = zp_dns(768);
      hipLaunchKernelGGL((zpass<768, P, Z_DNS, z_NT<Z_DNS>(768)>), dim3((unsigned)(a.npos / P)), dim3(z_NT<Z_DNS>(768)), 0, s, a, tw);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace psx
