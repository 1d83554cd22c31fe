// Fused z-pass instantiations, mode Z_TM2.
#include "zlaunch.hpp"

namespace psx {

int zpass_positions_tm2(int L) {
  switch (L) {
    case 32: return z_P<Z_TM2>(32);
    case 64: return z_P<Z_TM2>(64);
    case 96: return z_P<Z_TM2>(96);
    case 128: return z_P<Z_TM2>(128);
    case 192: return z_P<Z_TM2>(192);
    case 256: return z_P<Z_TM2>(256);
    case 384: return z_P<Z_TM2>(384);
    case 512: return z_P<Z_TM2>(512);
    case 768: return z_P<Z_TM2>(768);
    default: return 0;
  }
}

hipError_t launch_zpass_tm2(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  switch (L) {
    case 32: return zpass_launch<32, Z_TM2>(a, tw, s, variant);
    case 64: return zpass_launch<64, Z_TM2>(a, tw, s, variant);
    case 96: return zpass_launch<96, Z_TM2>(a, tw, s, variant);
    case 128: return zpass_launch<128, Z_TM2>(a, tw, s, variant);
    case 192: return zpass_launch<192, Z_TM2>(a, tw, s, variant);
    case 256: return zpass_launch<256, Z_TM2>(a, tw, s, variant);
    case 384: return zpass_launch<384, Z_TM2>(a, tw, s, variant);
    case 512: return zpass_launch<512, Z_TM2>(a, tw, s, variant);
    case 768: return zpass_launch<768, Z_TM2>(a, tw, s, variant);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace psx
