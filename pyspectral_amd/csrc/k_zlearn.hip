// Fused z-pass instantiations for the learned closure and physical-space output:
// Z_PHYS (c2r to the physical grid), Z_GRAD (the 9 velocity gradients), Z_LRN
// (products plus a given stress) and Z_LRNF (products plus the network's stress,
// evaluated inside the pass).
#include "zlaunch.hpp"

namespace psx {

int zpass_positions_phys(int L) {
  switch (L) {
    case 16: return z_P<Z_PHYS>(16);
    case 24: return z_P<Z_PHYS>(24);
    case 32: return z_P<Z_PHYS>(32);
    case 48: return z_P<Z_PHYS>(48);
    case 64: return z_P<Z_PHYS>(64);
    case 72: return z_P<Z_PHYS>(72);
    case 96: return z_P<Z_PHYS>(96);
    case 128: return z_P<Z_PHYS>(128);
    case 144: return z_P<Z_PHYS>(144);
    case 192: return z_P<Z_PHYS>(192);
    case 256: return z_P<Z_PHYS>(256);
    case 288: return z_P<Z_PHYS>(288);
    case 384: return z_P<Z_PHYS>(384);
    case 512: return z_P<Z_PHYS>(512);
    default: return 0;
  }
}

hipError_t launch_zpass_phys(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  switch (L) {
    case 16: return zpass_launch<16, Z_PHYS>(a, tw, s, variant);
    case 24: return zpass_launch<24, Z_PHYS>(a, tw, s, variant);
    case 32: return zpass_launch<32, Z_PHYS>(a, tw, s, variant);
    case 48: return zpass_launch<48, Z_PHYS>(a, tw, s, variant);
    case 64: return zpass_launch<64, Z_PHYS>(a, tw, s, variant);
    case 72: return zpass_launch<72, Z_PHYS>(a, tw, s, variant);
    case 96: return zpass_launch<96, Z_PHYS>(a, tw, s, variant);
    case 128: return zpass_launch<128, Z_PHYS>(a, tw, s, variant);
    case 144: return zpass_launch<144, Z_PHYS>(a, tw, s, variant);
    case 192: return zpass_launch<192, Z_PHYS>(a, tw, s, variant);
    case 256: return zpass_launch<256, Z_PHYS>(a, tw, s, variant);
    case 288: return zpass_launch<288, Z_PHYS>(a, tw, s, variant);
    case 384: return zpass_launch<384, Z_PHYS>(a, tw, s, variant);
    case 512: return zpass_launch<512, Z_PHYS>(a, tw, s, variant);
    default: return hipErrorInvalidValue;
  }
}

int zpass_positions_grad(int L) {
  switch (L) {
    case 16: return z_P<Z_GRAD>(16);
    case 24: return z_P<Z_GRAD>(24);
    case 32: return z_P<Z_GRAD>(32);
    case 48: return z_P<Z_GRAD>(48);
    case 64: return z_P<Z_GRAD>(64);
    case 72: return z_P<Z_GRAD>(72);
    case 96: return z_P<Z_GRAD>(96);
    case 128: return z_P<Z_GRAD>(128);
    case 144: return z_P<Z_GRAD>(144);
    case 192: return z_P<Z_GRAD>(192);
    case 256: return z_P<Z_GRAD>(256);
    case 288: return z_P<Z_GRAD>(288);
    case 384: return z_P<Z_GRAD>(384);
    default: return 0;
  }
}

hipError_t launch_zpass_grad(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  switch (L) {
    case 16: return zpass_launch<16, Z_GRAD>(a, tw, s, variant);
    case 24: return zpass_launch<24, Z_GRAD>(a, tw, s, variant);
    case 32: return zpass_launch<32, Z_GRAD>(a, tw, s, variant);
    case 48: return zpass_launch<48, Z_GRAD>(a, tw, s, variant);
    case 64: return zpass_launch<64, Z_GRAD>(a, tw, s, variant);
    case 72: return zpass_launch<72, Z_GRAD>(a, tw, s, variant);
    case 96: return zpass_launch<96, Z_GRAD>(a, tw, s, variant);
    case 128: return zpass_launch<128, Z_GRAD>(a, tw, s, variant);
    case 144: return zpass_launch<144, Z_GRAD>(a, tw, s, variant);
    case 192: return zpass_launch<192, Z_GRAD>(a, tw, s, variant);
    case 256: return zpass_launch<256, Z_GRAD>(a, tw, s, variant);
    case 288: return zpass_launch<288, Z_GRAD>(a, tw, s, variant);
    case 384: return zpass_launch<384, Z_GRAD>(a, tw, s, variant);
    default: return hipErrorInvalidValue;
  }
}

int zpass_positions_lrn(int L) {
  switch (L) {
    case 24: return z_P<Z_LRN>(24);
    case 48: return z_P<Z_LRN>(48);
    case 72: return z_P<Z_LRN>(72);
    case 96: return z_P<Z_LRN>(96);
    case 144: return z_P<Z_LRN>(144);
    case 192: return z_P<Z_LRN>(192);
    case 288: return z_P<Z_LRN>(288);
    case 384: return z_P<Z_LRN>(384);
    case 576: return z_P<Z_LRN>(576);
    default: return 0;
  }
}

hipError_t launch_zpass_lrn(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  switch (L) {
    case 24: return zpass_launch<24, Z_LRN>(a, tw, s, variant);
    case 48: return zpass_launch<48, Z_LRN>(a, tw, s, variant);
    case 72: return zpass_launch<72, Z_LRN>(a, tw, s, variant);
    case 96: return zpass_launch<96, Z_LRN>(a, tw, s, variant);
    case 144: return zpass_launch<144, Z_LRN>(a, tw, s, variant);
    case 192: return zpass_launch<192, Z_LRN>(a, tw, s, variant);
    case 288: return zpass_launch<288, Z_LRN>(a, tw, s, variant);
    case 384: return zpass_launch<384, Z_LRN>(a, tw, s, variant);
    case 576: return zpass_launch<576, Z_LRN>(a, tw, s, variant);
    default: return hipErrorInvalidValue;
  }
}

// the fused closure (network inside the pass): the one-wavefront lengths, and the
// lane-group lengths 48, 72, 96 (positions per wavefront)
int zpass_positions_lrnf(int L) {
  switch (L) {
    case 48: return ZSubGroups<48, zs_waves<48, Z_LRNF>()>::S;
    case 72: return ZSubGroups<72, zs_waves<72, Z_LRNF>()>::S;
    case 96: return ZSubGroups<96, zs_waves<96, Z_LRNF>()>::S;
    case 64: return z_P<Z_LRNF>(64);
    case 128: return z_P<Z_LRNF>(128);
    case 192: return z_P<Z_LRNF>(192);
    case 256: return z_P<Z_LRNF>(256);
    case 384: return z_P<Z_LRNF>(384);
    case 512: return z_P<Z_LRNF>(512);
    case 768: return z_P<Z_LRNF>(768);
    default: return 0;
  }
}

hipError_t launch_zpass_lrnf(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  switch (L) {
    case 48: return zpass_launch<48, Z_LRNF>(a, tw, s, variant);
    case 72: return zpass_launch<72, Z_LRNF>(a, tw, s, variant);
    case 96: return zpass_launch<96, Z_LRNF>(a, tw, s, variant);
    case 64: return zpass_launch<64, Z_LRNF>(a, tw, s, variant);
    case 128: return zpass_launch<128, Z_LRNF>(a, tw, s, variant);
    case 192: return zpass_launch<192, Z_LRNF>(a, tw, s, variant);
    case 256: return zpass_launch<256, Z_LRNF>(a, tw, s, variant);
    case 384: return zpass_launch<384, Z_LRNF>(a, tw, s, variant);
    case 512: return zpass_launch<512, Z_LRNF>(a, tw, s, variant);
    case 768: return zpass_launch<768, Z_LRNF>(a, tw, s, variant);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace psx
