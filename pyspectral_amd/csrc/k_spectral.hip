// Forward transform of real grid fields along z, the contiguous axis: the first pass
// of psx_spectral (the r2c of the reference's fft_obj plans, src/Classes.py:608-613),
// whose y and x passes are the pencil kernels.  Post-processing, not the step: a
// block transforms T z-lines as complex lines with zero imaginary parts.
#include "launch.hpp"

namespace psx {

template <int L, int T>
__global__ __launch_bounds__(NT) void zreal_forward(const double* __restrict__ phys, long long in_sf,
                                                    double2* __restrict__ out, long long out_sf, int npos,
                                                    const double2* __restrict__ tw) {
  constexpr int LS = Row<L>::LS;
  constexpr int NC = L / 2 + 1;
  __shared__ double2 buf[T * LS];
  const int tid = threadIdx.x;
  PSX_STAGE_TW(NT)
  const long long pos0 = (long long)blockIdx.x * T;
  const double* src = phys + (long long)blockIdx.y * in_sf + pos0 * L;
  // consecutive threads read consecutive z of one line
  for (int e = tid; e < L * T; e += NT) {
    const int c = e / L, l = e - c * L;
    buf[c * LS + RW(l)] = make_double2(pos0 + c < npos ? src[(long long)c * L + l] : 0.0, 0.0);
  }
  __syncthreads();
  fft_batch<L, T, NT, LS, -1>(buf, twp, tid);
  double2* dst = out + (long long)blockIdx.y * out_sf + pos0 * NC;
  for (int e = tid; e < NC * T; e += NT) {
    const int c = e / NC, k = e - c * NC;
    if (pos0 + c < npos) dst[(long long)c * NC + k] = buf[c * LS + RW(k)];
  }
}

template <int L>
static hipError_t zreal_launch(const double* phys, long long in_sf, double2* out, long long out_sf, int npos,
                               int nfields, const double2* tw, hipStream_t s) {
  constexpr int T = pencil_T(L);
  const dim3 grid((unsigned)((npos + T - 1) / T), (unsigned)nfields);
  hipLaunchKernelGGL((zreal_forward<L, T>), grid, dim3(NT), 0, s, phys, in_sf, out, out_sf, npos, tw);
  return hipGetLastError();
}

hipError_t launch_zreal_forward(int L, const double* phys, long long in_sf, double2* out, long long out_sf, int npos,
                                int nfields, const double2* tw, hipStream_t s) {
  switch (L) {
#define PSX_CASE(LL) \
  case LL:           \
    return zreal_launch<LL>(phys, in_sf, out, out_sf, npos, nfields, tw, s);
    PSX_LENGTHS(PSX_CASE)
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace psx
