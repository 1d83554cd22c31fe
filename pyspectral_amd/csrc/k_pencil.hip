// Pencil-pass instantiations (x and y axes) for every supported length.
#include "launch.hpp"

namespace psx {

bool length_supported(int L) {
  switch (L) {
#define PSX_CASE(LL) case LL:
    PSX_LENGTHS(PSX_CASE)
#undef PSX_CASE
    return true;
    default:
      return false;
  }
}

// Tile widths (k3 columns per block) a pencil pass may use: 4, 8 and 16 at every
// length, 24 and 32 (a small grid's whole column range in one tile) up to length 96,
// as long as the block's LDS rows fit in 64 KiB.
template <int L, int T> constexpr bool tile_ok() {
  return (T <= 16 || L <= 96) && T * Row<L>::LS * 16 <= 65536;
}

int pencil_tile(int L, int T) {
  if (T <= 0) return pencil_T(L);
  switch (L) {
#define PSX_CASE(LL)                                                                   \
  case LL:                                                                             \
    switch (T) {                                                                       \
      case 4: return tile_ok<LL, 4>() ? T : -1;                                        \
      case 8: return tile_ok<LL, 8>() ? T : -1;                                        \
      case 16: return tile_ok<LL, 16>() ? T : -1;                                      \
      case 24: return tile_ok<LL, 24>() ? T : -1;                                      \
      case 32: return tile_ok<LL, 32>() ? T : -1;                                      \
      default: return -1;                                                              \
    }
    PSX_LENGTHS(PSX_CASE)
#undef PSX_CASE
    default:
      return -1;
  }
}

// The width a pass launches with: T when given, else the length's default; where that
// is 16 and 16 does not divide the column count C (C = 24: two 16-wide tiles would
// leave a quarter of the block's columns empty), 24 if it divides C, else 8 if that
// does.  A 24-column tile reads 384 contiguous bytes per row, an 8-column one 128.
static int launch_width(int L, int T, int C) {
  if (T > 0) return pencil_tile(L, T);
  const int t = pencil_T(L);
  if (t != 16 || C % 16 == 0) return t;
  if (C % 24 == 0 && pencil_tile(L, 24) == 24) return 24;
  return C % 8 == 0 ? 8 : t;
}

// kind: 0 forward, 1 inverse
template <int L, int T, int KIND>
static hipError_t launch_tile(const PencilArgs& a, const double2* tw, int njobs, hipStream_t s) {
  if constexpr (tile_ok<L, T>()) {
    const dim3 grid((unsigned)(a.A * ((a.C + T - 1) / T)), (unsigned)njobs);
    if constexpr (KIND == 1)
      hipLaunchKernelGGL((pencil_inverse<L, T>), grid, dim3(NT), 0, s, a, tw);
    else
      hipLaunchKernelGGL((pencil_forward<L, T>), grid, dim3(NT), 0, s, a, tw);
    return hipGetLastError();
  } else {
    return hipErrorInvalidValue;
  }
}

template <int L, int KIND>
static hipError_t launch_len(int T, const PencilArgs& a, const double2* tw, int njobs, hipStream_t s) {
  switch (T) {
    case 4: return launch_tile<L, 4, KIND>(a, tw, njobs, s);
    case 8: return launch_tile<L, 8, KIND>(a, tw, njobs, s);
    case 16: return launch_tile<L, 16, KIND>(a, tw, njobs, s);
    case 24: return launch_tile<L, 24, KIND>(a, tw, njobs, s);
    case 32: return launch_tile<L, 32, KIND>(a, tw, njobs, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_pencil_inverse(int L, const PencilArgs& a, const double2* tw, int njobs, hipStream_t s, int T) {
  const int t = launch_width(L, T, a.C);
  switch (L) {
#define PSX_CASE(LL) \
  case LL:           \
    return launch_len<LL, 1>(t, a, tw, njobs, s);
    PSX_LENGTHS(PSX_CASE)
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_pencil_forward(int L, const PencilArgs& a, const double2* tw, int njobs, hipStream_t s, int T) {
  const int t = launch_width(L, T, a.C);
  switch (L) {
#define PSX_CASE(LL) \
  case LL:           \
    return launch_len<LL, 0>(t, a, tw, njobs, s);
    PSX_LENGTHS(PSX_CASE)
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
}

template <int L, int T>
static hipError_t launch_sum_tile(const SumArgs& a, const double2* tw, int njobs, hipStream_t s) {
  if constexpr (tile_ok<L, T>()) {
    const dim3 grid((unsigned)(a.p.A * ((a.p.C + T - 1) / T)), (unsigned)njobs);
    hipLaunchKernelGGL((pencil_forward_sum<L, T>), grid, dim3(NT), 0, s, a, tw);
    return hipGetLastError();
  } else {
    return hipErrorInvalidValue;
  }
}

hipError_t launch_pencil_forward_sum(int L, const SumArgs& a, const double2* tw, int njobs, hipStream_t s, int T) {
  if (3 * a.p.Nspec != 2 * L) return hipErrorInvalidValue;  // the kernel's compile-time truncation
  const int t = launch_width(L, T, a.p.C);
  switch (L) {
#define PSX_CASE(LL)                                                     \
  case LL:                                                               \
    switch (t) {                                                         \
      case 4: return launch_sum_tile<LL, 4>(a, tw, njobs, s);            \
      case 8: return launch_sum_tile<LL, 8>(a, tw, njobs, s);            \
      case 16: return launch_sum_tile<LL, 16>(a, tw, njobs, s);          \
      case 24: return launch_sum_tile<LL, 24>(a, tw, njobs, s);          \
      case 32: return launch_sum_tile<LL, 32>(a, tw, njobs, s);          \
      default: return hipErrorInvalidValue;                              \
    }
    PSX_LENGTHS(PSX_CASE)
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace psx
