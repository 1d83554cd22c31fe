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

// Tile widths (k3 columns per block) a pencil pass may use: the length's default and
// its neighbours, as long as the block's LDS rows fit in 64 KiB.
template <int L, int T> constexpr bool tile_ok() { return T * Row<L>::LS * 16 <= 65536; }

int pencil_tile(int L, int T) {
  if (T <= 0) return pencil_T(L);
  if (T != 4 && T != 8 && T != 16) return -1;
  switch (L) {
#define PSX_CASE(LL)                                                                    \
  case LL:                                                                              \
    return (T == 4 ? tile_ok<LL, 4>() : (T == 8 ? tile_ok<LL, 8>() : tile_ok<LL, 16>())) ? T : -1;
    PSX_LENGTHS(PSX_CASE)
#undef PSX_CASE
    default:
      return -1;
  }
}

// The width a pass launches with: T when given, else the length's default, narrowed
// to 8 where 16 does not divide the column count C but 8 does (C = 24: two 16-wide
// tiles would leave a quarter of the block's columns empty).
static int launch_width(int L, int T, int C) {
  if (T > 0) return pencil_tile(L, T);
  const int t = pencil_T(L);
  return (t == 16 && C % 16 != 0 && C % 8 == 0) ? 8 : t;
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
      default: return hipErrorInvalidValue;                              \
    }
    PSX_LENGTHS(PSX_CASE)
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace psx
