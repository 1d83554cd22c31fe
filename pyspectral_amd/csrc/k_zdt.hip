// Fused z-pass instantiations, mode Z_DT.
#include "zlaunch.hpp"

namespace psx {

int zpass_positions_dt(int L) {
  switch (L) {
    case 16: return z_P<Z_DT>(16);
    case 32: return z_P<Z_DT>(32);
    case 48: return z_P<Z_DT>(48);
    case 64: return z_P<Z_DT>(64);
    case 96: return z_P<Z_DT>(96);
    case 128: return z_P<Z_DT>(128);
    case 192: return z_P<Z_DT>(192);
    case 256: return z_P<Z_DT>(256);
    case 384: return z_P<Z_DT>(384);
    case 512: return z_P<Z_DT>(512);
    default: return 0;
  }
}

hipError_t launch_zpass_dt(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  switch (L) {
    case 16: return zpass_launch<16, Z_DT>(a, tw, s, variant);
    case 32: return zpass_launch<32, Z_DT>(a, tw, s, variant);
    case 48: return zpass_launch<48, Z_DT>(a, tw, s, variant);
    case 64: return zpass_launch<64, Z_DT>(a, tw, s, variant);
    case 96: return zpass_launch<96, Z_DT>(a, tw, s, variant);
    case 128: return zpass_launch<128, Z_DT>(a, tw, s, variant);
    case 192: return zpass_launch<192, Z_DT>(a, tw, s, variant);
    case 256: return zpass_launch<256, Z_DT>(a, tw, s, variant);
    case 384: return zpass_launch<384, Z_DT>(a, tw, s, variant);
    case 512: return zpass_launch<512, Z_DT>(a, tw, s, variant);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace psx
