// Fused z-pass instantiations, mode Z_DNS.
#include "zlaunch.hpp"

namespace psx {

int zpass_positions_dns(int L) {
  switch (L) {
    case 24: return z_P<Z_DNS>(24);
    case 32: return z_P<Z_DNS>(32);
    case 48: return z_P<Z_DNS>(48);
    case 64: return z_P<Z_DNS>(64);
    case 72: return z_P<Z_DNS>(72);
    case 96: return z_P<Z_DNS>(96);
    case 128: return z_P<Z_DNS>(128);
    case 144: return z_P<Z_DNS>(144);
    case 192: return z_P<Z_DNS>(192);
    case 256: return z_P<Z_DNS>(256);
    case 288: return z_P<Z_DNS>(288);
    case 384: return z_P<Z_DNS>(384);
    case 512: return z_P<Z_DNS>(512);
    case 576: return z_P<Z_DNS>(576);
    case 768: return z_P<Z_DNS>(768);
    default: return 0;
  }
}

hipError_t launch_zpass_dns(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  switch (L) {
    case 24: return zpass_launch<24, Z_DNS>(a, tw, s, variant);
    case 32: return zpass_launch<32, Z_DNS>(a, tw, s, variant);
    case 48: return zpass_launch<48, Z_DNS>(a, tw, s, variant);
    case 64: return zpass_launch<64, Z_DNS>(a, tw, s, variant);
    case 72: return zpass_launch<72, Z_DNS>(a, tw, s, variant);
    case 96: return zpass_launch<96, Z_DNS>(a, tw, s, variant);
    case 128: return zpass_launch<128, Z_DNS>(a, tw, s, variant);
    case 144: return zpass_launch<144, Z_DNS>(a, tw, s, variant);
    case 192: return zpass_launch<192, Z_DNS>(a, tw, s, variant);
    case 256: return zpass_launch<256, Z_DNS>(a, tw, s, variant);
    case 288: return zpass_launch<288, Z_DNS>(a, tw, s, variant);
    case 384: return zpass_launch<384, Z_DNS>(a, tw, s, variant);
    case 512: return zpass_launch<512, Z_DNS>(a, tw, s, variant);
    case 576: return zpass_launch<576, Z_DNS>(a, tw, s, variant);
    case 768: return zpass_launch<768, Z_DNS>(a, tw, s, variant);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace psx
