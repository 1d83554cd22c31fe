// Fused z-pass instantiations, mode Z_SMAG.
#include "zlaunch.hpp"

namespace psx {

int zpass_positions_smag(int L) {
  switch (L) {
    case 24: return z_P<Z_SMAG>(24);
    case 48: return z_P<Z_SMAG>(48);
    case 72: return z_P<Z_SMAG>(72);
    case 96: return z_P<Z_SMAG>(96);
    case 144: return z_P<Z_SMAG>(144);
    case 192: return z_P<Z_SMAG>(192);
    case 288: return z_P<Z_SMAG>(288);
    case 384: return z_P<Z_SMAG>(384);
    default: return 0;
  }
}

hipError_t launch_zpass_smag(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  switch (L) {
    case 24: return zpass_launch<24, Z_SMAG>(a, tw, s, variant);
    case 48: return zpass_launch<48, Z_SMAG>(a, tw, s, variant);
    case 72: return zpass_launch<72, Z_SMAG>(a, tw, s, variant);
    case 96: return zpass_launch<96, Z_SMAG>(a, tw, s, variant);
    case 144: return zpass_launch<144, Z_SMAG>(a, tw, s, variant);
    case 192: return zpass_launch<192, Z_SMAG>(a, tw, s, variant);
    case 288: return zpass_launch<288, Z_SMAG>(a, tw, s, variant);
    case 384: return zpass_launch<384, Z_SMAG>(a, tw, s, variant);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace psx
