// Host-side launchers: each kernel family is instantiated for the supported
// transform lengths in its own translation unit; these entry points pick the
// instantiation for a runtime length.  All return a hipError_t.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace psx {

// Tile widths (contiguous k3 columns per block) and z-pass position counts,
// sized so a block's LDS buffer stays near 48 KiB.
constexpr int pencil_T(int L) { return L <= 96 ? 16 : (L <= 384 ? 8 : 4); }
constexpr int epi_T(int L) { return L <= 192 ? 8 : (L <= 384 ? 4 : 2); }
// PL pass (t-model family): columns and threads per block -- at length 256, 8 columns
// and 512 threads (four entries per thread, 128-byte row segments)
constexpr int pl_T(int L) { return L == 256 ? 8 : epi_T(L); }
constexpr int pl_NT(int L) { return L == 256 ? 512 : NT; }
// z pass: positions per block (pairing two positions, or two fields when P = 1)
// and threads per block, sized to the first (radix-8 or -4) stage's butterflies.
template <int MODE> constexpr int z_P(int L) {
  return (MODE == Z_SMAG || MODE == Z_TM2 || MODE == Z_SGS || MODE == Z_GRAD || MODE == Z_LRN || MODE == Z_LRNF)
             ? 1
             : 2;
}
template <int MODE> constexpr int z_NT(int L) {
  constexpr int nin = ZShape<MODE>::nin, nout = ZShape<MODE>::nout;
  const int P = z_P<MODE>(L);
  const int nc = ((nin > nout ? nin : nout) * P + 1) / 2;
  const int bf = nc * (L / 8);
  const int nt = ((bf + 63) / 64) * 64;
  return nt < 64 ? 64 : (nt > 256 ? 256 : nt);
}

// Every transform length the kernels are instantiated for.
#define PSX_LENGTHS(X) \
  X(16) X(24) X(32) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(576) X(768)

bool length_supported(int L);

// r2c along z (the contiguous axis) of nfields real fields phys[f*in_sf + pos*L + z]
// into out[f*out_sf + pos*(L/2+1) + k], k = 0..L/2, unnormalised (k_spectral.hip)
hipError_t launch_zreal_forward(int L, const double* phys, long long in_sf, double2* out, long long out_sf, int npos,
                                int nfields, const double2* tw, hipStream_t s);

// T: tile width (0 = the length's default pencil_T, or 24 or 8 where 16 does not
// divide the column count but 24 or 8 does); pencil_tile(L, T) gives the width used
// for T > 0, or -1 when the width is not available for that length
int pencil_tile(int L, int T);
hipError_t launch_pencil_inverse(int L, const PencilArgs& a, const double2* tw, int njobs, hipStream_t s, int T = 0);
hipError_t launch_pencil_forward(int L, const PencilArgs& a, const double2* tw, int njobs, hipStream_t s, int T = 0);
hipError_t launch_pencil_forward_sum(int L, const SumArgs& a, const double2* tw, int njobs, hipStream_t s, int T = 0);
// mode: ZMode; returns hipErrorInvalidValue for an unsupported (L, mode)
// variant: 0 = the one-wavefront kernel where the length allows it, 1 = the block kernel
hipError_t launch_zpass(int L, int mode, const ZArgs& a, const double2* tw, hipStream_t s, int variant = 0);
int zpass_positions(int L, int mode);
// epi: EpiMode; ns1 = spectral length along x
hipError_t launch_epilogue(int L, int ns1, int epi, bool w0, const EpiArgs& a, const double2* tw, hipStream_t s);
unsigned epilogue_blocks(int L, const EpiArgs& a);
hipError_t launch_pl_pass(int L, bool dyn, const PLArgs& a, const double2* tw, hipStream_t s);
// filtered-DNS stress x-forward (3/2 pairs only)
hipError_t launch_xforward_sgs(int L, int ns1, const SgsArgs& a, const double2* tw, hipStream_t s);
unsigned sgs_blocks(int L, const SgsArgs& a);
unsigned pl_pass_blocks(int L, const PLArgs& a);

}  // namespace psx
