// Fused z-pass instantiations, mode Z_SGS (filtered-DNS stress products).
#include "zlaunch.hpp"

namespace psx {

int zpass_positions_sgs(int L) {
  switch (L) {
    case 24: return z_P<Z_SGS>(24);
    case 48: return z_P<Z_SGS>(48);
    case 72: return z_P<Z_SGS>(72);
    case 96: return z_P<Z_SGS>(96);
    case 144: return z_P<Z_SGS>(144);
    case 192: return z_P<Z_SGS>(192);
    case 288: return z_P<Z_SGS>(288);
    case 384: return z_P<Z_SGS>(384);
    case 576: return z_P<Z_SGS>(576);
    case 768: return z_P<Z_SGS>(768);
    default: return 0;
  }
}

hipError_t launch_zpass_sgs(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  switch (L) {
    case 24: return zpass_launch<24, Z_SGS>(a, tw, s, variant);
    case 48: return zpass_launch<48, Z_SGS>(a, tw, s, variant);
    case 72: return zpass_launch<72, Z_SGS>(a, tw, s, variant);
    case 96: return zpass_launch<96, Z_SGS>(a, tw, s, variant);
    case 144: return zpass_launch<144, Z_SGS>(a, tw, s, variant);
    case 192: return zpass_launch<192, Z_SGS>(a, tw, s, variant);
    case 288: return zpass_launch<288, Z_SGS>(a, tw, s, variant);
    case 384: return zpass_launch<384, Z_SGS>(a, tw, s, variant);
    case 576: return zpass_launch<576, Z_SGS>(a, tw, s, variant);
    case 768: return zpass_launch<768, Z_SGS>(a, tw, s, variant);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace psx
