// libpsx: context, plans and the C ABI declared in include/psx.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/psx.h"
#include "launch.hpp"
#include "small_kernels.hpp"

namespace psx {
hipError_t launch_zpass_dns(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant);
hipError_t launch_zpass_smag(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant);
hipError_t launch_zpass_tm2(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant);
hipError_t launch_zpass_dt(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant);
hipError_t launch_zpass_sgs(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant);
hipError_t launch_zpass_phys(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant);
hipError_t launch_zpass_grad(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant);
hipError_t launch_zpass_lrn(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant);
hipError_t mlp_closure(const psx_mlp* m, const float* g, float* tau, long long n, float c2, float bound, int clip,
                       hipStream_t s);
void mlp_closure_args(const psx_mlp* m, const short** pk, const float** w, psx_mlp_impl::Norm* nm);
hipError_t launch_zpass_lrnf(int L, const ZArgs& a, const double2* tw, hipStream_t s, int variant);
int zpass_positions_lrnf(int L);
int zpass_positions_dns(int L);
int zpass_positions_smag(int L);
int zpass_positions_tm2(int L);
int zpass_positions_dt(int L);
int zpass_positions_sgs(int L);
int zpass_positions_phys(int L);
int zpass_positions_grad(int L);
int zpass_positions_lrn(int L);

int zpass_positions(int L, int mode) {
  switch (mode) {
    case Z_DNS: return zpass_positions_dns(L);
    case Z_SMAG: return zpass_positions_smag(L);
    case Z_TM2: return zpass_positions_tm2(L);
    case Z_DT: return zpass_positions_dt(L);
    case Z_SGS: return zpass_positions_sgs(L);
    case Z_PHYS: return zpass_positions_phys(L);
    case Z_GRAD: return zpass_positions_grad(L);
    case Z_LRN: return zpass_positions_lrn(L);
    case Z_LRNF: return zpass_positions_lrnf(L);
  }
  return 0;
}

hipError_t launch_zpass(int L, int mode, const ZArgs& a, const double2* tw, hipStream_t s, int variant) {
  switch (mode) {
    case Z_DNS: return launch_zpass_dns(L, a, tw, s, variant);
    case Z_SMAG: return launch_zpass_smag(L, a, tw, s, variant);
    case Z_TM2: return launch_zpass_tm2(L, a, tw, s, variant);
    case Z_DT: return launch_zpass_dt(L, a, tw, s, variant);
    case Z_SGS: return launch_zpass_sgs(L, a, tw, s, variant);
    case Z_PHYS: return launch_zpass_phys(L, a, tw, s, variant);
    case Z_GRAD: return launch_zpass_grad(L, a, tw, s, variant);
    case Z_LRN: return launch_zpass_lrn(L, a, tw, s, variant);
    case Z_LRNF: return launch_zpass_lrnf(L, a, tw, s, variant);
  }
  return hipErrorInvalidValue;
}
}  // namespace psx

using namespace psx;

namespace {
constexpr double kPi = 3.14159265358979323846264338327950288;

struct Err {
  int code;
};
}  // namespace

struct psx_context {
  int N1 = 0, N2 = 0, N3 = 0, Nc = 0, H = 0;
  int M1 = 0, M2 = 0, M3 = 0;  // 3/2-padded lengths
  int model = 0;
  psx_params p{};
  int device = 0;
  hipStream_t stream = nullptr;
  // pipeline scratch
  double2* bufA = nullptr;
  double2* bufB = nullptr;
  size_t nA = 0, nB = 0;  // per member
  // batch contexts: `members` instances of the same grid, one after another in every
  // state and scratch array, with per-member viscosity nu_m (device) and dt
  int members = 1;
  double* nu_m = nullptr;
  // kernel shapes (psx_autotune / psx_set_tuning): tile widths of the x-inverse,
  // y-inverse and y-forward passes of the 3/2 pipeline (0 = default), z-pass variant
  int tune[PSX_TUNE_N] = {};
  // x-chunk pipelining of the y-inverse / z / y-forward passes over two streams
  // (tune[4] chunks): separate X and W buffers so that chunks never alias, the
  // second stream and its fork/stagger/join events
  double2* bufX = nullptr;
  double2* bufW = nullptr;
  size_t nX = 0, nW = 0;  // per member
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_stagger = nullptr, ev_join = nullptr;
  // t-model family: Y1 (velocity after the 2x y-inverse), PL and PLQL on the base
  // grid, the filtered velocity and the filtered Leonard block (dynamic tau)
  double2* bufC = nullptr;
  double2* pln = nullptr;
  double2* plql = nullptr;
  double2* uf = nullptr;
  double2* glf = nullptr;
  // dynamic Smagorinsky: spectral store of 24 product fields and its reduction partials
  double2* dsD = nullptr;
  double* dsPart = nullptr;
  int ndsPart = 0;
  // filtered-DNS stress work buffers, allocated on first use
  double2* sgsA = nullptr;
  double2* sgsB = nullptr;
  double2* sgsU = nullptr;
  double2* sgsT = nullptr;
  size_t nC = 0;
  double* partLE = nullptr;
  double* partA = nullptr;
  double* partB = nullptr;
  int npartT = 0;
  // learned closure: the attached network (not owned), its input range bound, and
  // the FP32 gradient / stress arrays of the padded grid (9 and 6 values per point)
  psx_mlp* net = nullptr;
  double net_bound = 0.0;
  int net_flags = 0;
  float* lgrad = nullptr;
  float* ltau = nullptr;
  double* cfl_part = nullptr;  // per-block CFL maxima (N1*N2/P entries)
  // the step's CFL passes on their own stream and buffers (3/2-pipeline models with
  // cfl > 0), overlapping the first stage's x/y/z passes; the epilogue that first
  // needs dt waits for ev_cfl_join (cfl_pending)
  double2* cflA = nullptr;
  double2* cflB = nullptr;
  hipStream_t stream_cfl = nullptr;
  hipEvent_t ev_cfl_fork = nullptr, ev_cfl_join = nullptr;
  bool cfl_pending = false;
  double* scal = nullptr;  // device scalars: [0] dt (internal), [1..4] tau family
  double* part = nullptr;
  int npart = 0;
  std::map<int, double2*> tw;
  long long scratch_bytes = 0;
  std::string err;
  // captured step (psx_step_graph), the pointers it was captured for and the
  // settings generation it holds by value (gen: bumped by every call that changes
  // parameters, tuning or the closure)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  const void* graph_key[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  unsigned long long gen = 1, graph_gen = 0;
  // non-finite flag: set by the CFL and diagnostics reductions when they meet a NaN or
  // an infinite value; host-mapped, so the host reads what completed kernels wrote
  // without a copy (nf_dev is the device view of nf_host)
  int* nf_host = nullptr;
  int* nf_dev = nullptr;
  // optional per-kernel timing with HIP events on the context stream
  bool timing = false;
  std::vector<hipEvent_t> events;
  size_t events_used = 0;
  std::vector<int> event_kinds;
};

namespace {

void check(psx_context* c, hipError_t e, const char* what) {
  if (e != hipSuccess) {
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    throw Err{e == hipErrorOutOfMemory ? PSX_E_OOM : PSX_E_HIP};
  }
}

void fail(psx_context* c, int code, const std::string& msg) {
  c->err = msg;
  throw Err{code};
}

template <typename F>
int guarded(psx_context* c, F&& f) {
  if (!c) return PSX_E_ARG;
  try {
    c->err.clear();
    f();
    return PSX_OK;
  } catch (const Err& e) {
    return e.code;
  } catch (const std::exception& e) {
    c->err = e.what();
    return PSX_E_HIP;
  }
}

// A captured step holds parameters, tuning and the closure by value: drop it.
void drop_graph(psx_context* c) {
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  for (auto& k : c->graph_key) k = nullptr;
  ++c->gen;
}

bool nonfinite_seen(const psx_context* c) { return c->nf_host && *(volatile const int*)c->nf_host != 0; }

// PSX_OK, or PSX_E_NONFINITE when a reduction that has completed met a NaN or an
// infinite value (the call's work is enqueued either way)
int with_status(const psx_context* c, int rc) { return (rc == PSX_OK && nonfinite_seen(c)) ? PSX_E_NONFINITE : rc; }

bool truncating(int model) {
  return model != PSX_MODEL_SMAGORINSKY && model != PSX_MODEL_DSMAG && model != PSX_MODEL_LEARNED;
}

// Kernel classes reported by psx_timing_end (order of PSX_KERNEL_* in psx.h).
enum Kind { K_XINV = 0, K_YINV, K_ZPASS, K_YFWD, K_XEPI, K_CFL_X, K_CFL_Y, K_CFL_Z, K_OTHER, K_CLOSURE, K_NKIND };

hipEvent_t next_event(psx_context* c) {
  if (c->events_used == c->events.size()) {
    hipEvent_t e;
    check(c, hipEventCreate(&e), "hipEventCreate");
    c->events.push_back(e);
  }
  return c->events[c->events_used++];
}

// Brackets one launch with start/stop events when timing is on.
struct TimeScope {
  psx_context* c;
  TimeScope(psx_context* ctx, int kind) : c(ctx) {
    if (c->timing) {
      check(c, hipEventRecord(next_event(c), c->stream), "hipEventRecord");
      c->event_kinds.push_back(kind);
    }
  }
  ~TimeScope() {
    if (c->timing) hipEventRecord(next_event(c), c->stream);
  }
};

const double2* twiddles(psx_context* c, int L) {
  auto it = c->tw.find(L);
  if (it != c->tw.end()) return it->second;
  if (!length_supported(L)) fail(c, PSX_E_UNSUPPORTED, "transform length " + std::to_string(L) + " not supported");
  std::vector<double2> h(L);
  for (int m = 0; m < L; ++m) {
    // exp(-2 pi i m / L), evaluated on the reduced angle for accuracy
    const double ang = 2.0 * kPi * (double)m / (double)L;
    h[m] = make_double2(std::cos(ang), -std::sin(ang));
  }
  double2* d = nullptr;
  check(c, hipMalloc(&d, sizeof(double2) * L), "hipMalloc(twiddles)");
  c->scratch_bytes += sizeof(double2) * L;
  check(c, hipMemcpy(d, h.data(), sizeof(double2) * L, hipMemcpyHostToDevice), "hipMemcpy(twiddles)");
  c->tw[L] = d;
  return d;
}

PencilArgs pencil_base() {
  PencilArgs a;
  std::memset(&a, 0, sizeof(a));
  for (int j = 0; j < MAXJOBS; ++j) {
    a.job_in[j] = a.job_plain[j] = a.job_deriv[j] = a.job_in2[j] = -1;
    a.job_dcoef[j] = 1.0;
  }
  a.scale = 1.0;
  return a;
}

// ---- pipeline building blocks --------------------------------------------------
// Per-member stride (complex elements) of an array the pipeline reads or writes:
// the scratch buffers, or the state layout for everything else.
long long member_stride(const psx_context* c, const void* p) {
  if (p == c->bufA) return (long long)c->nA;
  if (p == c->bufB) return (long long)c->nB;
  if (p == c->bufX) return (long long)c->nX;
  if (p == c->bufW) return (long long)c->nW;
  return 3LL * c->N1 * c->N2 * c->Nc;
}
void set_members(const psx_context* c, PencilArgs& a, int njobs) {
  a.jpm = njobs;
  a.in_ms = member_stride(c, a.in);
  a.out_ms = member_stride(c, a.out);
}

// x-inverse from the state layout: (nf fields of uin) -> X[f][L1][N2][C]
void x_inverse(psx_context* c, const double2* uin, double2* X, int L1, int C, int keep_nyq, int njobs,
               const int* jin, const int* jplain, const int* jderiv, int kind = K_XINV) {
  PencilArgs a = pencil_base();
  a.in = uin;
  a.out = X;
  a.in_sf = (long long)c->N1 * c->N2 * c->Nc;
  a.in_sa = c->Nc;
  a.in_sl = (long long)c->N2 * c->Nc;
  a.out_sf = (long long)L1 * c->N2 * C;
  a.out_sa = C;
  a.out_sl = (long long)c->N2 * C;
  a.A = c->N2;
  a.C = C;
  a.Nspec = c->N1;
  a.keep_nyq = keep_nyq;
  a.kscale = 2.0 * kPi / c->p.L1;
  for (int j = 0; j < njobs; ++j) {
    a.job_in[j] = jin[j];
    a.job_plain[j] = jplain[j];
    a.job_deriv[j] = jderiv ? jderiv[j] : -1;
  }
  set_members(c, a, njobs);
  const double2* tw = twiddles(c, L1);
  if (kind == K_XINV) a.order = c->tune[7];
  TimeScope ts(c, kind);
  check(c, launch_pencil_inverse(L1, a, tw, njobs * c->members, c->stream, kind == K_XINV ? c->tune[0] : 0),
        "x_inverse");
}

// y-inverse: X[f][L1][N2][C] -> Y[f][L1][L2][C]
void y_inverse(psx_context* c, const double2* X, double2* Y, int L1, int L2, int C, int keep_nyq, int njobs,
               const int* jin, const int* jplain, const int* jderiv, int kind = K_YINV, int ns2 = 0, int x0 = 0,
               int x1 = -1, const int* jin2 = nullptr, const double* jdcoef = nullptr) {
  if (ns2 == 0) ns2 = c->N2;
  if (x1 < 0) x1 = L1;
  PencilArgs a = pencil_base();
  a.in = X;
  a.out = Y;
  a.in_sf = (long long)L1 * ns2 * C;
  a.in_sa = (long long)ns2 * C;
  a.in_sl = C;
  a.out_sf = (long long)L1 * L2 * C;
  a.out_sa = (long long)L2 * C;
  a.out_sl = C;
  a.A = x1 - x0;
  a.C = C;
  a.Nspec = ns2;
  a.keep_nyq = keep_nyq;
  a.kscale = 2.0 * kPi / c->p.L2;
  for (int j = 0; j < njobs; ++j) {
    a.job_in[j] = jin[j];
    a.job_plain[j] = jplain[j];
    a.job_deriv[j] = jderiv ? jderiv[j] : -1;
    if (jin2) a.job_in2[j] = jin2[j];
    if (jdcoef) a.job_dcoef[j] = jdcoef[j];
  }
  set_members(c, a, njobs);
  a.in += x0 * a.in_sa;
  a.out += x0 * a.out_sa;
  const double2* tw = twiddles(c, L2);
  if (kind == K_YINV) a.order = c->tune[8];
  TimeScope ts(c, kind);
  check(c, launch_pencil_inverse(L2, a, tw, njobs * c->members, c->stream, kind == K_YINV ? c->tune[1] : 0),
        "y_inverse");
}

// y-forward with truncation: Z[f][L1][L2][C] -> W[f][L1][N2][C]
void y_forward(psx_context* c, const double2* Z, double2* W, int L1, int L2, int C, int keep_nyq, int nf,
               int ns2 = 0, int x0 = 0, int x1 = -1) {
  if (ns2 == 0) ns2 = c->N2;
  if (x1 < 0) x1 = L1;
  PencilArgs a = pencil_base();
  a.in = Z;
  a.out = W;
  a.in_sf = (long long)L1 * L2 * C;
  a.in_sa = (long long)L2 * C;
  a.in_sl = C;
  a.out_sf = (long long)L1 * ns2 * C;
  a.out_sa = (long long)ns2 * C;
  a.out_sl = C;
  a.A = x1 - x0;
  a.C = C;
  a.Nspec = ns2;
  a.keep_nyq = keep_nyq;
  for (int j = 0; j < nf; ++j) {
    a.job_in[j] = j;
    a.job_plain[j] = j;
  }
  set_members(c, a, nf);
  a.in += x0 * a.in_sa;
  a.out += x0 * a.out_sa;
  const double2* tw = twiddles(c, L2);
  a.order = c->tune[9];
  TimeScope ts(c, K_YFWD);
  check(c, launch_pencil_forward(L2, a, tw, nf * c->members, c->stream, c->tune[2]), "y_forward");
}

// y-forward of the folded Smagorinsky-form fields (D_mn at 0..5, tau_mn at 6..11, mn =
// 11 22 33 12 13 23) into the 9 fields the epilogue transforms: D_mn, A = tau_11,
// B = 2 k2 tau_12 + 2 k3 tau_13, C = k2^2 tau_22 + 2 k2 k3 tau_23 + k3^2 tau_33.  The
// three-term job runs first, so the longest blocks start earliest.
void y_forward_sum(psx_context* c, const double2* Z, double2* W, int L1, int L2, int C, int x0 = 0, int x1 = -1) {
  if (x1 < 0) x1 = L1;
  SumArgs sa;
  std::memset(&sa, 0, sizeof(sa));
  sa.p = pencil_base();
  PencilArgs& a = sa.p;
  a.in = Z;
  a.out = W;
  a.in_sf = (long long)L1 * L2 * C;
  a.in_sa = (long long)L2 * C;
  a.in_sl = C;
  a.out_sf = (long long)L1 * c->N2 * C;
  a.out_sa = (long long)c->N2 * C;
  a.out_sl = C;
  a.A = x1 - x0;
  a.C = C;
  a.Nspec = c->N2;
  a.keep_nyq = 0;
  sa.ks2 = 2.0 * kPi / c->p.L2;
  sa.ks3 = 2.0 * kPi / c->p.L3;
  struct Term { int f; double coef; int p2, p3; };
  const std::vector<std::pair<int, std::vector<Term>>> jobs = {
      {8, {{7, 1.0, 2, 0}, {11, 2.0, 1, 1}, {8, 1.0, 0, 2}}},  // C
      {7, {{9, 2.0, 1, 0}, {10, 2.0, 0, 1}}},                  // B
      {0, {{0, 1.0, 0, 0}}}, {1, {{1, 1.0, 0, 0}}}, {2, {{2, 1.0, 0, 0}}},
      {3, {{3, 1.0, 0, 0}}}, {4, {{4, 1.0, 0, 0}}}, {5, {{5, 1.0, 0, 0}}},
      {6, {{6, 1.0, 0, 0}}}};                                  // D_mn, A
  for (size_t j = 0; j < jobs.size(); ++j) {
    a.job_in[j] = jobs[j].second[0].f;
    a.job_plain[j] = jobs[j].first;
    sa.nterm[j] = (int)jobs[j].second.size();
    for (size_t t = 0; t < jobs[j].second.size(); ++t) {
      sa.tf[j][t] = jobs[j].second[t].f;
      sa.tc[j][t] = jobs[j].second[t].coef;
      sa.tp2[j][t] = jobs[j].second[t].p2;
      sa.tp3[j][t] = jobs[j].second[t].p3;
    }
  }
  const int nj = (int)jobs.size();
  set_members(c, a, nj);
  a.in += x0 * a.in_sa;
  a.out += x0 * a.out_sa;
  const double2* tw = twiddles(c, L2);
  a.order = c->tune[9];
  TimeScope ts(c, K_YFWD);
  check(c, launch_pencil_forward_sum(L2, sa, tw, nj * c->members, c->stream, c->tune[2]), "y_forward(stress sums)");
}

// Whether the folded Smagorinsky-form fields leave the y-forward as the 9 fields of
// y_forward_sum (tune[10] = 0) or as all 12, combined by the epilogue (tune[10] = 1).  At
// 128^3 the 9-field form saves 150 MB per stage: y-forward 161 -> 166 us, epilogue
// 114 -> 92 us.
bool sums_in_y(const psx_context* c) { return c->tune[10] == 0; }

// K1-K4 of the 3/2-rule pipeline; returns the buffer holding W (the y-forward of the
// z-pass products of uin).  smag adds the strain-rate pencils and the stress
// -2 c2 |S| S_ij (c2 = (Cs Delta)^2 for the static closure); fold (with smag) has the
// z pass write D_mn = N_mn + tau_mn in place of the products (for E_SMAGF).  With tune[4] = n > 1
// the y-inverse, z and y-forward passes run in n chunks of x rows, alternating between
// the context stream and a second stream (the second starting one y-inverse later), so
// an HBM-bound y pass of one chunk overlaps the z pass of another; X and W then live
// in their own buffers so that no chunk reads what another writes.
double2* front_32(psx_context* c, const double2* uin, bool smag, double c2, bool fold = false) {
  const int M1 = c->M1, M2 = c->M2, M3 = c->M3, H = c->H;
  const int nch = c->tune[4] > 1 ? c->tune[4] : 1;
  double2* X = nch > 1 ? c->bufX : c->bufA;
  double2* Y = c->bufB;
  double2* Z = c->bufA;
  double2* W = nch > 1 ? c->bufW : c->bufB;
  // K1: x-inverse (+ d/dx for the closure)
  {
    const int jin[3] = {0, 1, 2}, jp[3] = {0, 1, 2}, jd[3] = {3, 4, 5};
    x_inverse(c, uin, X, M1, H, 0, 3, jin, jp, smag ? jd : nullptr);
  }
  const int nY = smag ? 8 : 3;
  const int nF = smag ? 12 : 6;
  ZArgs z;
  std::memset(&z, 0, sizeof(z));
  z.in_sf0 = (long long)M1 * M2 * H;
  z.Hin0 = H;
  z.nin0 = nY;
  z.in1 = nullptr;
  z.out_sf = (long long)M1 * M2 * H;
  z.Hout = H;
  z.ms_in0 = member_stride(c, Y);
  z.ms_out = member_stride(c, Z);
  z.kzscale = 2.0 * kPi / c->p.L3;
  const double Nt = (double)c->N1 * c->N2 * c->N3;
  z.out_scale = 1.0 / (std::sqrt(Nt) * (double)M1 * M2 * M3);
  z.c2 = c2;
  z.fold = smag && fold ? 1 : 0;
  auto one = [](Recipe& r, int f) { r.nterm = 1; r.f[0] = f; r.coef[0] = 1.0; r.dz[0] = 0; };
  one(z.rec[0], 0);
  one(z.rec[1], 1);
  one(z.rec[2], 2);
  if (smag) {
    // Y fields: u v w ux vy S12 wx wy; pencils u v w S11 S22 S33 S12 S13 S23
    one(z.rec[3], 3);
    one(z.rec[4], 4);
    z.rec[5] = Recipe{1, {2, 0}, {1.0, 0.0}, {1, 0}};
    one(z.rec[6], 5);
    z.rec[7] = Recipe{2, {0, 6}, {0.5, 0.5}, {1, 0}};
    z.rec[8] = Recipe{2, {1, 7}, {0.5, 0.5}, {1, 0}};
  }
  const int mode = smag ? Z_SMAG : Z_DNS;
  const int P = zpass_positions(M3, mode);
  if (P == 0 || (M2 % P) != 0) fail(c, PSX_E_UNSUPPORTED, "z pass: unsupported padded length");
  const double2* tw3 = twiddles(c, M3);
  const hipStream_t main = c->stream;
  struct Restore {  // the context stream is switched per chunk; put it back on any exit
    psx_context* c;
    hipStream_t s;
    ~Restore() { c->stream = s; }
  } restore{c, main};
  // tune[5] = 1: the chunks run in order on the context stream (each chunk's Y and Z
  // stay small enough to be re-read from the last-level cache); else two streams
  const bool two = nch > 1 && c->tune[5] != 1;
  if (two) {
    check(c, hipEventRecord(c->ev_fork, main), "hipEventRecord");
    check(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0), "hipStreamWaitEvent");
  }
  for (int ch = 0; ch < nch; ++ch) {
    const int x0 = (int)((long long)M1 * ch / nch), x1 = (int)((long long)M1 * (ch + 1) / nch);
    c->stream = (two && (ch & 1)) ? c->stream2 : main;
    if (two && ch == 1) check(c, hipStreamWaitEvent(c->stream, c->ev_stagger, 0), "hipStreamWaitEvent");
    // K2: y-inverse.  With the closure, X holds u v w ux vx wx and Y gets u v w ux vy
    // S12 = (uy + vx)/2 wx wy: S12 is one transform of (i k2 u + vx)/2, and vx needs
    // no field of its own
    if (smag) {
      const int jin[5] = {0, 1, 2, 3, 5}, jp[5] = {0, 1, 2, 3, 6}, jd[5] = {5, 4, 7, -1, -1};
      const int jin2[5] = {4, -1, -1, -1, -1};
      const double dc[5] = {0.5, 1.0, 1.0, 1.0, 1.0};
      y_inverse(c, X, Y, M1, M2, H, 0, 5, jin, jp, jd, K_YINV, 0, x0, x1, jin2, dc);
    } else {
      const int jin[3] = {0, 1, 2}, jp[3] = {0, 1, 2};
      y_inverse(c, X, Y, M1, M2, H, 0, 3, jin, jp, nullptr, K_YINV, 0, x0, x1);
    }
    if (ch == 0 && two) check(c, hipEventRecord(c->ev_stagger, c->stream), "hipEventRecord");
    // K3: z c2r -> products (+ Smagorinsky stress) -> z r2c, on positions [x0 M2, x1 M2)
    {
      ZArgs zc = z;
      zc.in0 = Y + (long long)x0 * M2 * H;
      zc.out = Z + (long long)x0 * M2 * H;
      zc.npos = (x1 - x0) * M2 * c->members;
      zc.npos_m = c->members > 1 ? (x1 - x0) * M2 : 0;
      TimeScope ts(c, K_ZPASS);
      check(c, launch_zpass(M3, mode, zc, tw3, c->stream, c->tune[3]), "z_pass");
    }
    // K4: y-forward + truncation
    if (smag && fold && sums_in_y(c)) {
      y_forward_sum(c, Z, W, M1, M2, H, x0, x1);
    } else {
      y_forward(c, Z, W, M1, M2, H, 0, nF, 0, x0, x1);
    }
  }
  c->stream = main;
  if (two) {
    check(c, hipEventRecord(c->ev_join, c->stream2), "hipEventRecord");
    check(c, hipStreamWaitEvent(main, c->ev_join, 0), "hipStreamWaitEvent");
  }
  return W;
}

double delta_grid(const psx_context* c) { return std::cbrt(c->p.dx * c->p.dy * c->p.dz); }

void join_cfl(psx_context* c);

// K5 of the 3/2-rule models: x-forward + truncation + projection/closure/viscous +
// stage update, from the y-forwarded products in B (with the stress fields when smag;
// folded: D_mn = N_mn + tau_mn in place of the products, no w0)
void epilogue_32(psx_context* c, const double2* B, const double2* uin, const double2* q0, double2* out,
                 const double* dt, double coef, double2* w0, int zero_nyq_out, bool smag, bool folded = false) {
  if (folded && (!smag || w0)) fail(c, PSX_E_ARG, "epilogue: the folded fields carry no separate stress");
  join_cfl(c);  // dt of an overlapped CFL reduction
  const int M1 = c->M1, H = c->H;
  EpiArgs e;
  std::memset(&e, 0, sizeof(e));
  e.in = B;
  e.in_sf = (long long)M1 * c->N2 * H;
  e.N1 = c->N1;
  e.N2 = c->N2;
  e.N3 = c->N3;
  e.Nc = c->Nc;
  e.H = H;
  e.keep_nyq = 0;
  e.uprev = uin;
  e.q0 = q0;
  e.out = out;
  e.w0 = w0;
  e.dt = dt;
  e.c = coef;
  e.nu = c->p.nu;
  e.ks[0] = 2.0 * kPi / c->p.L1;
  e.ks[1] = 2.0 * kPi / c->p.L2;
  e.ks[2] = 2.0 * kPi / c->p.L3;
  e.trunc_u = truncating(c->model) ? 1 : 0;
  e.zero_nyq_out = zero_nyq_out;
  e.rotate = c->p.rotate;
  e.Om[0] = c->p.Om1;
  e.Om[1] = c->p.Om2;
  e.Om[2] = c->p.Om3;
  e.members = c->members;
  e.force = c->p.force;
  e.ms_in = member_stride(c, e.in);
  e.ms_state = 3LL * c->N1 * c->N2 * c->Nc;
  e.nu_m = c->nu_m;
  const double2* tw = twiddles(c, M1);
  TimeScope ts(c, K_XEPI);
  const int epi = folded ? (sums_in_y(c) ? E_SMAG9 : E_SMAGF) : (smag ? E_SMAG : E_DNS);
  check(c, launch_epilogue(M1, c->N1, epi, !folded && smag && w0 != nullptr,
                           e, tw, c->stream),
        "x_forward_epilogue");
}

// RHS / stage for the 3/2-rule models (DNS, Smagorinsky)
void stage_32(psx_context* c, const double2* uin, const double2* q0, double2* out, const double* dt, double coef,
              double2* w0, int zero_nyq_out) {
  const bool smag = c->model == PSX_MODEL_SMAGORINSKY;
  const double cd = c->p.Cs * delta_grid(c);
  // without a w0 output the products and stresses fold into 9 transformed fields
  const bool fold = smag && w0 == nullptr;
  double2* B = front_32(c, uin, smag, cd * cd, fold);
  epilogue_32(c, B, uin, q0, out, dt, coef, w0, zero_nyq_out, smag, fold);
}

// Learned closure: the 3/2-rule pipeline with the stress of an attached network.
// The z pass first writes the nine velocity gradients du_i/dx_j of every padded grid
// point (FP32, point-major); the network maps them to the six stresses (order 11 22 33
// 12 13 23, Smagorinsky outside its input range); a second z pass forms the products
// and joins the stresses; then the y-forward and the Smagorinsky-form epilogue
// (w0_i = -i k_j tau_ij - k_i k_m k_j tau_mj / k^2, as src/RHSfunctions.py:215-222).
void stage_learned(psx_context* c, const double2* uin, const double2* q0, double2* out, const double* dt,
                   double coef, double2* w0, int zero_nyq_out) {
  if (!c->net) fail(c, PSX_E_ARG, "learned closure: no network attached (psx_set_closure)");
  const int M1 = c->M1, M2 = c->M2, M3 = c->M3, H = c->H;
  double2* X = c->bufA;
  double2* Y = c->bufB;
  double2* Z = c->bufA;
  double2* W = c->bufB;
  {
    const int jin[3] = {0, 1, 2}, jp[3] = {0, 1, 2}, jd[3] = {3, 4, 5};
    x_inverse(c, uin, X, M1, H, 0, 3, jin, jp, jd);
  }
  {
    const int jin[6] = {0, 1, 2, 3, 4, 5}, jp[6] = {0, 1, 2, 3, 4, 5}, jd[6] = {6, 7, 8, -1, -1, -1};
    y_inverse(c, X, Y, M1, M2, H, 0, 6, jin, jp, jd);
  }
  const double Nt = (double)c->N1 * c->N2 * c->N3;
  const double2* tw3 = twiddles(c, M3);
  ZArgs z;
  std::memset(&z, 0, sizeof(z));
  z.in0 = Y;
  z.in_sf0 = (long long)M1 * M2 * H;
  z.Hin0 = H;
  z.nin0 = 9;
  z.ms_in0 = member_stride(c, Y);
  z.npos = M1 * M2 * c->members;
  z.npos_m = c->members > 1 ? M1 * M2 : 0;
  z.kzscale = 2.0 * kPi / c->p.L3;
  const double cd = c->p.Cs * delta_grid(c);
  const int clip = (c->net_flags & PSX_CLOSURE_CLIP_BACKSCATTER) ? 1 : 0;
  const int src[9] = {3, 6, 0, 4, 7, 1, 5, 8, 2};
  const int PF = zpass_positions(M3, Z_LRNF);
  if (c->tune[6] == 0 && PF > 0 && M2 % PF == 0) {
    // the network inside the z pass: velocity and gradient pencils in, products and
    // stresses out, no gradient or stress arrays in HBM
    ZArgs q = z;
    q.out = Z;
    q.out_sf = (long long)M1 * M2 * H;
    q.Hout = H;
    q.ms_out = member_stride(c, Z);
    for (int r = 0; r < 3; ++r) q.rec[r] = Recipe{1, {r, 0}, {1.0, 0.0}, {0, 0}};
    for (int r = 0; r < 9; ++r) q.rec[3 + r] = Recipe{1, {src[r], 0}, {1.0, 0.0}, {r % 3 == 2 ? 1 : 0, 0}};
    q.out_scale = 1.0 / (std::sqrt(Nt) * (double)M1 * M2 * M3);
    q.c2 = q.out_scale * Nt;
    q.gscale = 1.0 / std::sqrt(Nt);
    mlp_closure_args(c->net, &q.npk, &q.nw, &q.nnorm);
    q.nc2 = (float)(cd * cd);
    q.nbound = (float)c->net_bound;
    q.nclip = clip;
    q.fold = w0 == nullptr ? 1 : 0;
    {
      TimeScope ts(c, K_ZPASS);
      check(c, launch_zpass(M3, Z_LRNF, q, tw3, c->stream, 0), "z_pass(fused closure)");
    }
    if (w0 == nullptr && sums_in_y(c)) {
      y_forward_sum(c, Z, W, M1, M2, H);
    } else {
      y_forward(c, Z, W, M1, M2, H, 0, 12);
    }
    epilogue_32(c, W, uin, q0, out, dt, coef, w0, zero_nyq_out, true, w0 == nullptr);
    return;
  }
  {
    // Y fields u v w ux vx wx uy vy wy -> du_i/dx_j, (i, j) row-major
    ZArgs g = z;
    for (int r = 0; r < 9; ++r) g.rec[r] = Recipe{1, {src[r], 0}, {1.0, 0.0}, {r % 3 == 2 ? 1 : 0, 0}};
    g.out_scale = 1.0 / std::sqrt(Nt);  // pencil values are sqrt(Nt) times physical
    g.gout = c->lgrad;
    const int P = zpass_positions(M3, Z_GRAD);
    if (P == 0 || (M2 % P) != 0) fail(c, PSX_E_UNSUPPORTED, "z pass: unsupported padded length");
    TimeScope ts(c, K_ZPASS);
    check(c, launch_zpass(M3, Z_GRAD, g, tw3, c->stream, c->tune[3]), "z_pass(gradients)");
  }
  {
    const long long npts = (long long)M1 * M2 * M3 * c->members;
    TimeScope ts(c, K_CLOSURE);
    check(c, mlp_closure(c->net, c->lgrad, c->ltau, npts, (float)(cd * cd), (float)c->net_bound, clip, c->stream),
          "closure network");
  }
  {
    ZArgs q = z;
    q.out = Z;
    q.out_sf = (long long)M1 * M2 * H;
    q.Hout = H;
    q.ms_out = member_stride(c, Z);
    for (int r = 0; r < 3; ++r) q.rec[r] = Recipe{1, {r, 0}, {1.0, 0.0}, {0, 0}};
    q.out_scale = 1.0 / (std::sqrt(Nt) * (double)M1 * M2 * M3);
    // physical stress to the products' units: pencil products are Nt times physical
    q.c2 = q.out_scale * Nt;
    q.tin = c->ltau;
    q.fold = w0 == nullptr ? 1 : 0;
    const int P = zpass_positions(M3, Z_LRN);
    if (P == 0 || (M2 % P) != 0) fail(c, PSX_E_UNSUPPORTED, "z pass: unsupported padded length");
    TimeScope ts(c, K_ZPASS);
    check(c, launch_zpass(M3, Z_LRN, q, tw3, c->stream, c->tune[3]), "z_pass(learned)");
  }
  if (w0 == nullptr && sums_in_y(c)) {
    y_forward_sum(c, Z, W, M1, M2, H);
  } else {
    y_forward(c, Z, W, M1, M2, H, 0, 12);
  }
  epilogue_32(c, W, uin, q0, out, dt, coef, w0, zero_nyq_out, true, w0 == nullptr);
}

// computeRHS_DSMAG (src/RHSfunctions.py:1663-1862): the Smagorinsky pipeline on the
// velocity (c2 = Delta^2) and on its test-filtered copy (c2 = alpha^2 Delta^2), each
// x-forwarded into the spectral store; Cs^2 from the store by a deterministic
// reduction; then the elementwise RHS with tau = Cs^2 T and the stage update.  The
// whole stage stays on the device (graph-capturable).
constexpr double kDsmagAlpha = 2.0;

void store_x_forward(psx_context* c, const double2* W, int first) {
  const int M1 = c->M1, H = c->H;
  PencilArgs a = pencil_base();
  a.in = W;
  a.out = c->dsD;
  a.in_sf = (long long)M1 * c->N2 * H;
  a.in_sa = H;
  a.in_sl = (long long)c->N2 * H;
  a.out_sf = (long long)c->N1 * c->N2 * H;
  a.out_sa = H;
  a.out_sl = (long long)c->N2 * H;
  a.A = c->N2;
  a.C = H;
  a.Nspec = c->N1;
  a.keep_nyq = 0;
  for (int j = 0; j < 12; ++j) {
    a.job_in[j] = j;
    a.job_plain[j] = first + j;
  }
  const double2* tw = twiddles(c, M1);
  TimeScope ts(c, K_XEPI);
  check(c, launch_pencil_forward(M1, a, tw, 12, c->stream), "x_forward(store)");
}

void stage_dsmag(psx_context* c, const double2* uin, const double2* q0, double2* out, const double* dt, double coef,
                 double2* w0) {
  const double d = delta_grid(c);
  const double kL = (double)(int)((double)c->p.kc / kDsmagAlpha);  // myFilter(., kc/alpha): int slice bound
  const long long total = 3LL * c->N1 * c->N2 * c->Nc;
  const unsigned eb = (unsigned)std::min<long long>((total + 255) / 256, 8192);
  hipLaunchKernelGGL(apply_filter, dim3(eb), dim3(256), 0, c->stream, uin, c->uf, 3, c->N1, c->N2, c->Nc, F_SHARP,
                     kL, 0.0, 0.0, 0.0, 0.0);
  check(c, hipGetLastError(), "apply_filter");
  store_x_forward(c, front_32(c, uin, true, d * d), 0);
  store_x_forward(c, front_32(c, c->uf, true, kDsmagAlpha * kDsmagAlpha * d * d), 12);
  {
    TimeScope ts(c, K_OTHER);
    hipLaunchKernelGGL(dsmag_partials, dim3(c->ndsPart), dim3(NT), 0, c->stream, (const double2*)c->dsD, c->dsPart,
                       c->N1, c->N2, c->H, kL);
    check(c, hipGetLastError(), "dsmag_partials");
    hipLaunchKernelGGL(dsmag_finalize, dim3(1), dim3(NT), 0, c->stream, (const double*)c->dsPart, c->ndsPart,
                       c->scal + 1);
    check(c, hipGetLastError(), "dsmag_finalize");
  }
  DsmagArgs a;
  std::memset(&a, 0, sizeof(a));
  a.D = c->dsD;
  a.uprev = uin;
  a.q0 = q0;
  a.out = out;
  a.w0 = w0;
  a.cs2 = c->scal + 1;
  a.dt = dt;
  a.c = coef;
  a.nu = c->p.nu;
  a.ks[0] = 2.0 * kPi / c->p.L1;
  a.ks[1] = 2.0 * kPi / c->p.L2;
  a.ks[2] = 2.0 * kPi / c->p.L3;
  a.N1 = c->N1;
  a.N2 = c->N2;
  a.N3 = c->N3;
  a.Nc = c->Nc;
  a.H = c->H;
  a.rotate = c->p.rotate;
  a.Om[0] = c->p.Om1;
  a.Om[1] = c->p.Om2;
  a.Om[2] = c->p.Om3;
  a.force = c->p.force;
  TimeScope ts(c, K_XEPI);
  hipLaunchKernelGGL(dsmag_final, dim3((unsigned)std::min<long long>((total / 3 + 255) / 256, 8192)), dim3(256), 0,
                     c->stream, a);
  check(c, hipGetLastError(), "dsmag_final");
}

// One pass of the 2x-grid t-model pipeline on velocity u (already filtered for the
// test-filtered pass of the dynamic model): a first 2x-grid sweep (products, PL,
// PL_q), a second (u_j (PL_q)_i on the 2x grid) and the final x-forward epilogue.
// src/RHSfunctions.py:240-363 (t-model), 366-497 (static tau), 592-878 (dynamic).
void tau_pass(psx_context* c, const double2* u, bool filtered, const double2* uin, const double2* q0, double2* out,
              const double* dt, double stage_c, double2* w0, int zero_nyq_out, double tcoef,
              const double* tcoef_dev = nullptr) {
  const bool dyn = c->model == PSX_MODEL_DYNAMIC_TAU;
  const int N1 = c->N1, N2 = c->N2, N3 = c->N3, H = c->H;
  const int P1 = 2 * N1, P2 = 2 * N2, P3 = 2 * N3;
  double2* A = c->bufA;
  double2* B = c->bufB;
  double2* Y1 = c->bufC;
  const double Nt = (double)N1 * N2 * N3;
  const double kfd = (double)c->p.kc * 2.0 / 3.0;
  const int kf = (int)kfd;  // int(kc*2/3), src/Classes.py:131
  {
    const int jin[3] = {0, 1, 2}, jp[3] = {0, 1, 2};
    x_inverse(c, u, A, P1, H, 0, 3, jin, jp, nullptr);
    y_inverse(c, A, Y1, P1, P2, H, 0, 3, jin, jp, nullptr);
  }
  {
    ZArgs z;
    std::memset(&z, 0, sizeof(z));
    z.in0 = Y1;
    z.in_sf0 = (long long)P1 * P2 * H;
    z.Hin0 = H;
    z.nin0 = 3;
    z.out = A;
    z.out_sf = (long long)P1 * P2 * N3;
    z.Hout = N3;
    z.npos = P1 * P2;
    z.kzscale = 2.0 * kPi / c->p.L3;
    z.out_scale = 1.0 / (std::sqrt(Nt) * (double)P1 * P2 * P3);
    for (int r = 0; r < 3; ++r) z.rec[r] = Recipe{1, {r, 0}, {1.0, 0.0}, {0, 0}};
    const int P = zpass_positions(P3, Z_DNS);
    if (P == 0 || (z.npos % P) != 0) fail(c, PSX_E_UNSUPPORTED, "z pass: unsupported 2x length");
    const double2* tw = twiddles(c, P3);
    TimeScope ts(c, K_ZPASS);
    check(c, launch_zpass(P3, Z_DNS, z, tw, c->stream), "z_pass(2x)");
  }
  y_forward(c, A, B, P1, P2, N3, 1, 6, P2);
  {
    PLArgs a;
    std::memset(&a, 0, sizeof(a));
    a.in = B;
    a.in_sf = (long long)P1 * P2 * N3;
    a.P1 = P1;
    a.P2 = P2;
    a.C = N3;
    a.N1 = N1;
    a.N2 = N2;
    a.N3 = N3;
    a.Nc = c->Nc;
    a.u = u;
    a.nu = c->p.nu;
    a.ks[0] = 2.0 * kPi / c->p.L1;
    a.ks[1] = 2.0 * kPi / c->p.L2;
    a.ks[2] = 2.0 * kPi / c->p.L3;
    if (filtered) {
      a.kpx = a.kpy = a.kpz = kf;  // seperateModes_testFilter (src/padding.py:52-64)
    } else {
      a.kpx = N1 / 2;              // seperateModes (src/padding.py:67-79)
      a.kpy = N2 / 2;
      a.kpz = N3 / 2;
    }
    a.pln = filtered ? nullptr : c->pln;
    a.xq = A;
    if (dyn) {
      a.kf = kfd;
      a.uf = c->uf;
      if (filtered) {
        a.g_out = c->glf;
      } else {
        a.g_ref = c->glf;
        a.part = c->partLE;
      }
    }
    const double2* tw = twiddles(c, P1);
    TimeScope ts(c, K_OTHER);
    check(c, launch_pl_pass(P1, dyn, a, tw, c->stream), "pl_pass");
  }
  {
    const int jin[3] = {0, 1, 2}, jp[3] = {0, 1, 2};
    y_inverse(c, A, B, P1, P2, N3, 1, 3, jin, jp, nullptr, K_YINV, P2);
  }
  {
    ZArgs z;
    std::memset(&z, 0, sizeof(z));
    z.in0 = Y1;
    z.in_sf0 = (long long)P1 * P2 * H;
    z.Hin0 = H;
    z.nin0 = 3;
    z.in1 = B;
    z.in_sf1 = (long long)P1 * P2 * N3;
    z.Hin1 = N3;
    z.out = A;
    z.out_sf = (long long)P1 * P2 * H;
    z.Hout = H;
    z.npos = P1 * P2;
    z.kzscale = 2.0 * kPi / c->p.L3;
    z.out_scale = 1.0 / (std::sqrt(Nt) * (double)P1 * P2 * P3);
    for (int r = 0; r < 6; ++r) z.rec[r] = Recipe{1, {r, 0}, {1.0, 0.0}, {0, 0}};
    const int P = zpass_positions(P3, Z_TM2);
    if (P == 0 || (z.npos % P) != 0) fail(c, PSX_E_UNSUPPORTED, "z pass: unsupported 2x length");
    const double2* tw = twiddles(c, P3);
    TimeScope ts(c, K_ZPASS);
    check(c, launch_zpass(P3, Z_TM2, z, tw, c->stream), "z_pass(2x, second sweep)");
  }
  y_forward(c, A, B, P1, P2, H, 1, 9, N2);
  {
    EpiArgs e;
    std::memset(&e, 0, sizeof(e));
    e.in = B;
    e.in_sf = (long long)P1 * N2 * H;
    e.N1 = N1;
    e.N2 = N2;
    e.N3 = N3;
    e.Nc = c->Nc;
    e.H = H;
    e.keep_nyq = 1;
    e.uprev = uin;
    e.pl = c->pln;
    e.dt = dt;
    e.c = stage_c;
    e.nu = c->p.nu;
    e.ks[0] = 2.0 * kPi / c->p.L1;
    e.ks[1] = 2.0 * kPi / c->p.L2;
    e.ks[2] = 2.0 * kPi / c->p.L3;
    e.trunc_u = 1;
    e.zero_nyq_out = zero_nyq_out;
    e.force = c->p.force;
    e.rotate = c->p.rotate;
    e.Om[0] = c->p.Om1;
    e.Om[1] = c->p.Om2;
    e.Om[2] = c->p.Om3;
    if (dyn) {
      e.store_only = 1;
      e.w0 = filtered ? nullptr : c->plql;
      e.red_u = c->uf;
      e.red_part = filtered ? c->partA : c->partB;
    } else {
      e.q0 = q0;
      e.out = out;
      e.w0 = w0;
      e.coef = tcoef;
      e.coef_dev = tcoef_dev;
    }
    const double2* tw = twiddles(c, P1);
    TimeScope ts(c, K_XEPI);
    check(c, launch_epilogue(P1, N1, E_TAU2, false, e, tw, c->stream), "x_forward_epilogue(2x)");
  }
}

void stage_tau(psx_context* c, const double2* uin, const double2* q0, double2* out, const double* dt, double stage_c,
               double2* w0, int zero_nyq_out, double t, const double* t_dev) {
  const int N1 = c->N1, N2 = c->N2, N3 = c->N3, Nc = c->Nc;
  const long long total = 3LL * N1 * N2 * Nc;
  const unsigned eb = (unsigned)std::min<long long>((total + 255) / 256, 4096);
  const double ks0 = 2.0 * kPi / c->p.L1, ks1 = 2.0 * kPi / c->p.L2, ks2 = 2.0 * kPi / c->p.L3;
  if (c->model == PSX_MODEL_TMODEL || c->model == PSX_MODEL_STATIC_TAU) {
    const bool tm = c->model == PSX_MODEL_TMODEL;
    const double* coef_dev = nullptr;
    if (t_dev) {
      // the step-start time lives on the device (graph replay): coefficient from it
      hipLaunchKernelGGL(time_coef, dim3(1), dim3(1), 0, c->stream, t_dev, tm ? 1 : 0, c->p.Ct, c->p.tau_a,
                         c->p.tau_b, c->scal + 9);
      check(c, hipGetLastError(), "time_coef");
      coef_dev = c->scal + 9;
    }
    const double coef = tm ? c->p.Ct * t : std::max(0.0, c->p.tau_a + c->p.tau_b * t);
    tau_pass(c, uin, false, uin, q0, out, dt, stage_c, w0, zero_nyq_out, coef, coef_dev);
    return;
  }
  // dynamic tau-model: filtered pass, unfiltered pass, tau, update
  const double kf = (double)(int)((double)c->p.kc * 2.0 / 3.0);
  hipLaunchKernelGGL(filter_copy, dim3(eb), dim3(256), 0, c->stream, uin, c->uf, N1, N2, N3, Nc, ks0, ks1, ks2, kf);
  check(c, hipGetLastError(), "filter_copy");
  tau_pass(c, c->uf, true, uin, nullptr, nullptr, dt, stage_c, nullptr, 0, 0.0);
  tau_pass(c, uin, false, uin, nullptr, nullptr, dt, stage_c, nullptr, 0, 0.0);
  const int P1 = 2 * N1, P2 = 2 * N2;
  PLArgs pa;
  std::memset(&pa, 0, sizeof(pa));
  pa.P2 = P2;
  pa.C = N3;
  EpiArgs ea;
  std::memset(&ea, 0, sizeof(ea));
  ea.N1 = N1;
  ea.N2 = N2;
  ea.Nc = Nc;
  ea.H = c->H;
  const int nLE = (int)pl_pass_blocks(P1, pa);
  const int nE = (int)epilogue_blocks(P1, ea);
  hipLaunchKernelGGL(tau_finalize, dim3(1), dim3(NT), 0, c->stream, c->partLE, nLE, c->partA, nE, c->partB, nE,
                     1.0 / ((double)N1 * N2 * N3), c->scal + 1);
  check(c, hipGetLastError(), "tau_finalize");
  hipLaunchKernelGGL(tau_update, dim3(eb), dim3(256), 0, c->stream, c->pln, c->plql, c->scal + 2, uin, q0, out, w0, dt,
                     stage_c, N1, N2, N3, Nc, ks0, ks1, ks2, zero_nyq_out, c->p.rotate, c->p.Om1, c->p.Om2, c->p.Om3, c->p.force);
  check(c, hipGetLastError(), "tau_update");
}

void run_stage(psx_context* c, const double2* uin, const double2* q0, double2* out, const double* dt, double coef,
               double2* w0, int zero_nyq_out, double t, const double* t_dev = nullptr) {
  switch (c->model) {
    case PSX_MODEL_DNS:
    case PSX_MODEL_SMAGORINSKY:
      stage_32(c, uin, q0, out, dt, coef, w0, zero_nyq_out);
      break;
    case PSX_MODEL_TMODEL:
    case PSX_MODEL_STATIC_TAU:
    case PSX_MODEL_DYNAMIC_TAU:
      stage_tau(c, uin, q0, out, dt, coef, w0, zero_nyq_out, t, t_dev);
      break;
    case PSX_MODEL_DSMAG:
      stage_dsmag(c, uin, q0, out, dt, coef, w0);
      break;
    case PSX_MODEL_LEARNED:
      stage_learned(c, uin, q0, out, dt, coef, w0, zero_nyq_out);
      break;
    default:
      fail(c, PSX_E_UNSUPPORTED, "unknown model");
  }
}

// overlap: run the passes on stream_cfl (forked from the context stream) into the CFL
// buffers and leave the join to the next epilogue_32 (or join_cfl)
void compute_dt(psx_context* c, const double2* u, double* dt, bool overlap = false) {
  if (c->p.cfl <= 0) {
    hipLaunchKernelGGL(set_f64, dim3(1), dim3(1), 0, c->stream, dt, -c->p.cfl);
    check(c, hipGetLastError(), "set_f64");
    return;
  }
  overlap = overlap && c->cflA != nullptr;
  struct Restore {  // the context stream is switched to stream_cfl; put it back on any exit
    psx_context* c;
    hipStream_t s;
    ~Restore() { c->stream = s; }
  } restore{c, c->stream};
  if (overlap) {
    check(c, hipEventRecord(c->ev_cfl_fork, c->stream), "hipEventRecord");
    check(c, hipStreamWaitEvent(c->stream_cfl, c->ev_cfl_fork, 0), "hipStreamWaitEvent");
    c->stream = c->stream_cfl;
  }
  double2* XA = overlap ? c->cflA : c->bufA;
  double2* XB = overlap ? c->cflB : c->bufB;
  const int N1 = c->N1, N2 = c->N2, N3 = c->N3, Nc = c->Nc;
  const int jin[3] = {0, 1, 2}, jp[3] = {0, 1, 2};
  x_inverse(c, u, XA, N1, Nc, 1, 3, jin, jp, nullptr, K_CFL_X);
  y_inverse(c, XA, XB, N1, N2, Nc, 1, 3, jin, jp, nullptr, K_CFL_Y);
  ZArgs z;
  std::memset(&z, 0, sizeof(z));
  z.in0 = XB;
  z.in_sf0 = (long long)N1 * N2 * Nc;
  z.Hin0 = Nc;
  z.nin0 = 3;
  z.npos = N1 * N2 * c->members;
  z.npos_m = c->members > 1 ? N1 * N2 : 0;
  z.ms_in0 = member_stride(c, z.in0);
  z.kzscale = 2.0 * kPi / c->p.L3;
  for (int r = 0; r < 3; ++r) z.rec[r] = Recipe{1, {r, 0}, {1.0, 0.0}, {0, 0}};
  z.dinv[0] = 1.0 / c->p.dx;
  z.dinv[1] = 1.0 / c->p.dy;
  z.dinv[2] = 1.0 / c->p.dz;
  z.red = c->cfl_part;
  const int P = zpass_positions(N3, Z_DT);
  if (P == 0 || ((N1 * N2) % P) != 0) fail(c, PSX_E_UNSUPPORTED, "CFL pass: unsupported length");
  const int nparts = N1 * N2 / P;  // per member
  {
    const double2* tw = twiddles(c, N3);
    TimeScope ts(c, K_CFL_Z);
    check(c, launch_zpass(N3, Z_DT, z, tw, c->stream), "cfl_pass");
  }
  const double Nt = (double)N1 * N2 * N3;
  const double dx = c->p.dx, dy = c->p.dy, dz = c->p.dz;
  const double ksum = 1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz);
  hipLaunchKernelGGL(cfl_finalize, dim3(c->members), dim3(NT), 0, c->stream, c->cfl_part, nparts, dt,
                     1.0 / std::sqrt(Nt), c->p.cfl, c->p.nu, (const double*)c->nu_m, ksum, c->nf_dev);
  check(c, hipGetLastError(), "cfl_finalize");
  if (overlap) {
    check(c, hipEventRecord(c->ev_cfl_join, c->stream_cfl), "hipEventRecord");
    c->cfl_pending = true;
  }
}

// the context stream waits for the overlapped CFL passes (dt is ready after this)
void join_cfl(psx_context* c) {
  if (!c->cfl_pending) return;
  c->cfl_pending = false;
  check(c, hipStreamWaitEvent(c->stream, c->ev_cfl_join, 0), "hipStreamWaitEvent");
}

void validate_params(psx_context* c, const psx_params* p, int model) {
  if (!p) fail(c, PSX_E_ARG, "params is NULL");
  if (!(p->L1 > 0 && p->L2 > 0 && p->L3 > 0)) fail(c, PSX_E_ARG, "box lengths must be positive");
  if (!(p->dx > 0 && p->dy > 0 && p->dz > 0)) fail(c, PSX_E_ARG, "grid spacings must be positive");
  if (p->rotate != 0 && p->rotate != 1) fail(c, PSX_E_ARG, "rotate must be 0 or 1");
}

}  // namespace

extern "C" {

int psx_create(psx_context** out, int n1, int n2, int n3, int model, const psx_params* params, int device,
               void* stream) {
  return psx_create_batch(out, n1, n2, n3, model, params, 1, nullptr, device, stream);
}

int psx_create_batch(psx_context** out, int n1, int n2, int n3, int model, const psx_params* params, int members,
                     const double* nu_host, int device, void* stream) {
  if (!out) return PSX_E_ARG;
  *out = nullptr;
  psx_context* c = new psx_context();
  int rc = guarded(c, [&] {
    if (n1 < 4 || n2 < 4 || n3 < 4 || (n1 % 2) || (n2 % 2) || (n3 % 4))
      fail(c, PSX_E_ARG, "grid sizes must be even (N3 a multiple of 4)");
    if (model < PSX_MODEL_DNS || model > PSX_MODEL_LEARNED) fail(c, PSX_E_ARG, "unknown model");
    if (members < 1) fail(c, PSX_E_ARG, "members must be >= 1");
    if (members > 1 && model != PSX_MODEL_DNS && model != PSX_MODEL_SMAGORINSKY && model != PSX_MODEL_LEARNED)
      fail(c, PSX_E_UNSUPPORTED, "batched instances are available for the DNS, Smagorinsky and learned models");
    validate_params(c, params, model);
    c->members = members;
    c->N1 = n1;
    c->N2 = n2;
    c->N3 = n3;
    c->Nc = n3 / 2 + 1;
    c->H = n3 / 2;
    c->M1 = 3 * n1 / 2;
    c->M2 = 3 * n2 / 2;
    c->M3 = 3 * n3 / 2;
    c->model = model;
    c->p = *params;
    c->device = device;
    c->stream = (hipStream_t)stream;
    check(c, hipSetDevice(device), "hipSetDevice");
    const bool tau = model >= PSX_MODEL_TMODEL && model <= PSX_MODEL_DYNAMIC_TAU;
    std::vector<int> lens = {n1, n2, n3};
    if (tau) {
      for (int L : {2 * n1, 2 * n2, 2 * n3}) lens.push_back(L);
    } else {
      for (int L : {c->M1, c->M2, c->M3}) lens.push_back(L);
    }
    for (int L : lens)
      if (!length_supported(L)) fail(c, PSX_E_UNSUPPORTED, "grid size not supported: length " + std::to_string(L));
    const bool smag = model == PSX_MODEL_SMAGORINSKY || model == PSX_MODEL_DSMAG || model == PSX_MODEL_LEARNED;
    const size_t state = (size_t)3 * n1 * n2 * c->Nc;
    if (!tau) {
      const size_t MMH = (size_t)c->M1 * c->M2 * c->H;
      const size_t MNH = (size_t)c->M1 * c->N2 * c->H;
      const size_t nZ = (smag ? 12 : 6) * MMH;
      const size_t nX = (smag ? 6 : 3) * MNH;
      const size_t nY = (smag ? 9 : 3) * MMH;
      const size_t nW = (smag ? 12 : 6) * MNH;
      c->nA = std::max(std::max(nX, nZ), state);
      c->nB = std::max(std::max(nY, nW), state);
      if (model == PSX_MODEL_DSMAG) {
        const size_t nD = (size_t)24 * n1 * n2 * c->H;
        c->ndsPart = 1024;
        check(c, hipMalloc(&c->dsD, nD * sizeof(double2)), "hipMalloc(dsD)");
        check(c, hipMalloc(&c->uf, state * sizeof(double2)), "hipMalloc(uf)");
        check(c, hipMalloc(&c->dsPart, 2 * sizeof(double) * c->ndsPart), "hipMalloc(dsPart)");
        c->scratch_bytes += (long long)(nD + state) * sizeof(double2) + 16LL * c->ndsPart;
      }
      if (model == PSX_MODEL_LEARNED) {
        const size_t npts = (size_t)c->M1 * c->M2 * c->M3 * members;
        check(c, hipMalloc(&c->lgrad, 9 * npts * sizeof(float)), "hipMalloc(lgrad)");
        check(c, hipMalloc(&c->ltau, 6 * npts * sizeof(float)), "hipMalloc(ltau)");
        c->scratch_bytes += (long long)(15 * npts * sizeof(float));
      }
    } else {
      const size_t P1 = 2 * (size_t)n1, P2 = 2 * (size_t)n2;
      const size_t full = P1 * P2 * (size_t)n3;       // one 2x-grid field, k3 < N3
      const size_t half = P1 * P2 * (size_t)c->H;     // k3 < N3/2
      const size_t X1 = 3 * P1 * (size_t)n2 * c->H, W2 = 9 * P1 * (size_t)n2 * c->H;
      c->nA = std::max({X1, 6 * full, 3 * full, 9 * half, state});
      c->nB = std::max({6 * full, 3 * full, W2, state});
      c->nC = 3 * half;
      check(c, hipMalloc(&c->bufC, c->nC * sizeof(double2)), "hipMalloc(bufC)");
      check(c, hipMalloc(&c->pln, state * sizeof(double2)), "hipMalloc(pl)");
      c->scratch_bytes += (long long)(c->nC + state) * sizeof(double2);
      if (model == PSX_MODEL_DYNAMIC_TAU) {
        check(c, hipMalloc(&c->plql, state * sizeof(double2)), "hipMalloc(plql)");
        check(c, hipMalloc(&c->uf, state * sizeof(double2)), "hipMalloc(uf)");
        check(c, hipMalloc(&c->glf, state * sizeof(double2)), "hipMalloc(glf)");
        c->npartT = 1 << 20;
        check(c, hipMalloc(&c->partLE, 2 * sizeof(double) * c->npartT), "hipMalloc(partLE)");
        check(c, hipMalloc(&c->partA, 2 * sizeof(double) * c->npartT), "hipMalloc(partA)");
        check(c, hipMalloc(&c->partB, 2 * sizeof(double) * c->npartT), "hipMalloc(partB)");
        c->scratch_bytes += (long long)3 * state * sizeof(double2) + 6LL * sizeof(double) * c->npartT;
      }
    }
    const size_t mb = (size_t)members;
    check(c, hipMalloc(&c->bufA, mb * c->nA * sizeof(double2)), "hipMalloc(bufA)");
    check(c, hipMalloc(&c->bufB, mb * c->nB * sizeof(double2)), "hipMalloc(bufB)");
    check(c, hipMalloc(&c->cfl_part, mb * sizeof(double) * ((size_t)n1 * n2 / 2 + 1)), "hipMalloc(cfl_part)");
    if (params->cfl > 0 &&
        (model == PSX_MODEL_DNS || model == PSX_MODEL_SMAGORINSKY || model == PSX_MODEL_LEARNED)) {
      const size_t nst = 3 * (size_t)n1 * n2 * (n3 / 2 + 1);
      check(c, hipMalloc(&c->cflA, mb * nst * sizeof(double2)), "hipMalloc(cflA)");
      check(c, hipMalloc(&c->cflB, mb * nst * sizeof(double2)), "hipMalloc(cflB)");
      c->scratch_bytes += (long long)(2 * mb * nst * sizeof(double2));
      check(c, hipStreamCreateWithFlags(&c->stream_cfl, hipStreamNonBlocking), "hipStreamCreate");
      check(c, hipEventCreateWithFlags(&c->ev_cfl_fork, hipEventDisableTiming), "hipEventCreate");
      check(c, hipEventCreateWithFlags(&c->ev_cfl_join, hipEventDisableTiming), "hipEventCreate");
    }
    if (members > 1) {
      check(c, hipMalloc(&c->nu_m, mb * sizeof(double)), "hipMalloc(nu_m)");
      std::vector<double> nu(mb, params->nu);
      if (nu_host) nu.assign(nu_host, nu_host + mb);
      for (double v : nu)
        if (!(v >= 0.0)) fail(c, PSX_E_ARG, "member viscosities must be >= 0");
      check(c, hipMemcpy(c->nu_m, nu.data(), mb * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy(nu_m)");
    }
    check(c, hipHostMalloc((void**)&c->nf_host, 64, hipHostMallocMapped | hipHostMallocCoherent),
          "hipHostMalloc(flag)");
    *(volatile int*)c->nf_host = 0;
    check(c, hipHostGetDevicePointer((void**)&c->nf_dev, c->nf_host, 0), "hipHostGetDevicePointer(flag)");
    check(c, hipMalloc(&c->scal, 16 * sizeof(double)), "hipMalloc(scal)");
    check(c, hipMemset(c->scal, 0, 16 * sizeof(double)), "hipMemset(scal)");
    c->npart = 1024;
    check(c, hipMalloc(&c->part, 3 * c->npart * sizeof(double)), "hipMalloc(part)");
    c->scratch_bytes += (long long)members * ((c->nA + c->nB) * sizeof(double2) + 8LL * (n1 * n2 / 2 + 1)) +
                        16 * 8 + 3 * c->npart * 8 + (members > 1 ? 8LL * members : 0);
    for (int L : lens) twiddles(c, L);
  });
  if (rc != PSX_OK) {
    // keep the message reachable through a static buffer
    static thread_local std::string last;
    last = c->err;
    psx_destroy(c);
    std::fprintf(stderr, "psx_create: %s\n", last.c_str());
    return rc;
  }
  *out = c;
  return PSX_OK;
}

int psx_destroy(psx_context* c) {
  if (!c) return PSX_E_ARG;
  hipFree(c->bufA);
  hipFree(c->bufB);
  hipFree(c->bufX);
  hipFree(c->bufW);
  if (c->stream2) hipStreamDestroy(c->stream2);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_stagger) hipEventDestroy(c->ev_stagger);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  hipFree(c->nu_m);
  hipFree(c->cfl_part);
  hipFree(c->cflA);
  hipFree(c->cflB);
  if (c->stream_cfl) hipStreamDestroy(c->stream_cfl);
  if (c->ev_cfl_fork) hipEventDestroy(c->ev_cfl_fork);
  if (c->ev_cfl_join) hipEventDestroy(c->ev_cfl_join);
  hipFree(c->scal);
  hipFree(c->part);
  hipFree(c->bufC);
  hipFree(c->pln);
  hipFree(c->plql);
  hipFree(c->uf);
  hipFree(c->glf);
  hipFree(c->dsD);
  hipFree(c->dsPart);
  hipFree(c->sgsA);
  hipFree(c->sgsB);
  hipFree(c->sgsU);
  hipFree(c->sgsT);
  hipFree(c->partLE);
  hipFree(c->partA);
  hipFree(c->partB);
  hipFree(c->lgrad);
  hipFree(c->ltau);
  for (auto& kv : c->tw) hipFree(kv.second);
  for (auto e : c->events) hipEventDestroy(e);
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  if (c->graph) hipGraphDestroy(c->graph);
  // hipHostFree waits for the device, as hipFree does for the buffers above
  if (c->nf_host) hipHostFree(c->nf_host);
  delete c;
  return PSX_OK;
}

int psx_set_stream(psx_context* c, void* stream) {
  return guarded(c, [&] { c->stream = (hipStream_t)stream; });
}

int psx_set_params(psx_context* c, const psx_params* params) {
  return guarded(c, [&] {
    validate_params(c, params, c->model);
    c->p = *params;
    drop_graph(c);  // the captured step holds the old values
  });
}

int psx_status(psx_context* c) {
  const int rc = guarded(c, [&] { check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize"); });
  return with_status(c, rc);
}

int psx_clear_status(psx_context* c) {
  return guarded(c, [&] {
    check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    *(volatile int*)c->nf_host = 0;
  });
}

const char* psx_last_error(const psx_context* c) { return c ? c->err.c_str() : "null context"; }

long long psx_scratch_bytes(const psx_context* c) { return c ? c->scratch_bytes : 0; }

int psx_rhs(psx_context* c, const double* uhat, double* rhs, double* uhat_model, double* w0, double t) {
  return with_status(c, guarded(c, [&] {
    if (!uhat || !rhs) fail(c, PSX_E_ARG, "null pointer");
    if ((const double*)rhs == uhat) fail(c, PSX_E_ARG, "rhs must not alias uhat");
    run_stage(c, (const double2*)uhat, nullptr, (double2*)rhs, nullptr, 0.0, (double2*)w0, 0, t);
    if (uhat_model) {
      const long long total = 3LL * c->N1 * c->N2 * c->Nc;
      const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 4096);
      for (int m = 0; m < c->members; ++m) {
        hipLaunchKernelGGL(copy_truncate, dim3(blocks), dim3(256), 0, c->stream, (const double2*)uhat + m * total,
                           (double2*)uhat_model + m * total, c->N1, c->N2, c->Nc, c->N3,
                           truncating(c->model) ? 1 : 0);
        check(c, hipGetLastError(), "copy_truncate");
      }
    }
  }));
}

int psx_rk_stage(psx_context* c, const double* q0, const double* uin, double* out, const double* dt_dev, double coef,
                 double t, int zero_nyquist_out) {
  return with_status(c, guarded(c, [&] {
    if (!q0 || !uin || !out || !dt_dev) fail(c, PSX_E_ARG, "null pointer");
    run_stage(c, (const double2*)uin, (const double2*)q0, (double2*)out, dt_dev, coef, nullptr,
              zero_nyquist_out ? 1 : 0, t);
  }));
}

int psx_compute_dt(psx_context* c, const double* uhat, double* dt_dev) {
  return with_status(c, guarded(c, [&] {
    if (!uhat || !dt_dev) fail(c, PSX_E_ARG, "null pointer");
    compute_dt(c, (const double2*)uhat, dt_dev);
  }));
}

namespace {
// The body of advanceQ_RK4 (src/PySpec_3d.py:96-105).  With t_dev the step-start
// time is read from, and advanced by dt on, the device.
void step_body(psx_context* c, double2* Q, double2* S, double* dt_dev, double t, double* t_dev) {
  // the 3/2-pipeline stages read dt only in their epilogue and write S only there, so
  // the CFL reduction of S overlaps the first stage's x/y/z passes (in order while
  // per-kernel timing is on, so that each kernel's time is its own)
  const bool overlap =
      !c->timing &&
      (c->model == PSX_MODEL_DNS || c->model == PSX_MODEL_SMAGORINSKY || c->model == PSX_MODEL_LEARNED);
  compute_dt(c, S, dt_dev, overlap);
  const int z = truncating(c->model) ? 1 : 0;
  // src/PySpec_3d.py:99-104: Q = Q0 + dt*c_i*RHS(Q), c = (1/4, 1/3, 1/2, 1)
  run_stage(c, Q, Q, S, dt_dev, 0.25, nullptr, z, t, t_dev);
  join_cfl(c);
  run_stage(c, S, Q, S, dt_dev, 1.0 / 3.0, nullptr, z, t, t_dev);
  run_stage(c, S, Q, S, dt_dev, 0.5, nullptr, z, t, t_dev);
  run_stage(c, S, Q, Q, dt_dev, 1.0, nullptr, 0, t, t_dev);
  if (t_dev) {
    hipLaunchKernelGGL(advance_time, dim3(1), dim3(1), 0, c->stream, t_dev, dt_dev);
    check(c, hipGetLastError(), "advance_time");
  }
}
}  // namespace

int psx_step(psx_context* c, double* q, double* qs, double* dt_dev, double t) {
  return with_status(c, guarded(c, [&] {
    if (!q || !qs || !dt_dev) fail(c, PSX_E_ARG, "null pointer");
    if (q == qs) fail(c, PSX_E_ARG, "q and qs must be distinct buffers");
    step_body(c, (double2*)q, (double2*)qs, dt_dev, t, nullptr);
  }));
}

int psx_step_graph(psx_context* c, double* q, double* qs, double* dt_dev, double* t_dev) {
  return with_status(c, guarded(c, [&] {
    if (!q || !qs || !dt_dev || !t_dev) fail(c, PSX_E_ARG, "null pointer");
    if (q == qs) fail(c, PSX_E_ARG, "q and qs must be distinct buffers");
    if (c->stream == nullptr) fail(c, PSX_E_ARG, "graph capture needs a non-default stream (psx_set_stream)");
    const void* key[5] = {q, qs, dt_dev, t_dev, c->stream};
    bool same = c->graph_exec != nullptr && c->graph_gen == c->gen;
    for (int i = 0; i < 5 && same; ++i) same = c->graph_key[i] == key[i];
    if (!same) {
      if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
      if (c->graph) hipGraphDestroy(c->graph);
      c->graph_exec = nullptr;
      c->graph = nullptr;
      const bool timing = c->timing;
      c->timing = false;
      check(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        step_body(c, (double2*)q, (double2*)qs, dt_dev, 0.0, t_dev);
      } catch (...) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(c->stream, &g);
        if (g) hipGraphDestroy(g);
        c->timing = timing;
        throw;
      }
      c->timing = timing;
      check(c, hipStreamEndCapture(c->stream, &c->graph), "hipStreamEndCapture");
      check(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      for (int i = 0; i < 5; ++i) c->graph_key[i] = key[i];
      c->graph_gen = c->gen;
    }
    check(c, hipGraphLaunch(c->graph_exec, c->stream), "hipGraphLaunch");
  }));
}

int psx_stats(psx_context* c, const double* uhat, double* out_dev) {
  return with_status(c, guarded(c, [&] {
    if (!uhat || !out_dev) fail(c, PSX_E_ARG, "null pointer");
    const long long total = (long long)c->N1 * c->N2 * c->Nc;
    int nblk = (int)std::min<long long>((total + NT - 1) / NT, c->npart);
    const double Nt = (double)c->N1 * c->N2 * c->N3;
    for (int m = 0; m < c->members; ++m) {
      const double2* u = (const double2*)uhat + (long long)m * 3 * total;
      hipLaunchKernelGGL(stats_partial, dim3(nblk), dim3(NT), 0, c->stream, u, c->part, c->N1, c->N2, c->N3, c->Nc,
                         2.0 * kPi / c->p.L1, 2.0 * kPi / c->p.L2, 2.0 * kPi / c->p.L3, c->p.kc);
      check(c, hipGetLastError(), "stats_partial");
      hipLaunchKernelGGL(stats_final, dim3(1), dim3(NT), 0, c->stream, c->part, nblk, out_dev + 3 * m, 1.0 / Nt,
                         c->nf_dev);
      check(c, hipGetLastError(), "stats_final");
    }
  }));
}

// ---- diagnostics without allocation ------------------------------------------------
// The reduction partials of psx_stats_noalloc live in this file-scope device array, one
// copy per GPU for the life of the process.  Host code never passes `scratch` itself (in
// host code the name is a stand-in, not a device address): the address comes from
// hipGetSymbolAddress, once per device, with that device current.
namespace {
constexpr int kNoallocParts = 1024;  // partial sums per reduction (3 doubles each)
constexpr int kNoallocMaxDevices = 64;
}  // namespace
static __device__ __align__(256) double scratch[3 * kNoallocParts];

namespace {
double* noalloc_ptr[kNoallocMaxDevices] = {};
std::mutex noalloc_mu;

// The device address of this GPU's scratch array (nullptr on failure).
double* noalloc_scratch(int device) {
  if (device < 0 || device >= kNoallocMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (!noalloc_ptr[device]) {
    double* p = nullptr;
    if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    noalloc_ptr[device] = p;
  }
  return noalloc_ptr[device];
}
}  // namespace

extern "C" int psx_stats_noalloc(const double* uhat, int members, int n1, int n2, int n3, double L1, double L2,
                                 double L3, int kc, double* out_dev, void* stream) {
  if (!uhat || !out_dev || members < 1) return PSX_E_ARG;
  if (n1 < 2 || n2 < 2 || n3 < 2 || (n1 % 2) || (n2 % 2) || (n3 % 2)) return PSX_E_ARG;
  if (!(L1 > 0 && L2 > 0 && L3 > 0)) return PSX_E_ARG;
  const int nc = n3 / 2 + 1;
  const long long total = (long long)n1 * n2 * nc;
  // the grid-stride reduction covers any grid with at most kNoallocParts blocks, so the
  // partials never exceed the array
  const int nblk = (int)std::min<long long>((total + NT - 1) / NT, kNoallocParts);
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return PSX_E_HIP;
  double* part = noalloc_scratch(device);
  if (!part) return PSX_E_HIP;
  const hipStream_t s = (hipStream_t)stream;
  const double Nt = (double)n1 * n2 * n3;
  for (int m = 0; m < members; ++m) {
    const double2* u = (const double2*)uhat + (long long)m * 3 * total;
    hipLaunchKernelGGL(stats_partial, dim3(nblk), dim3(NT), 0, s, u, part, n1, n2, n3, nc, 2.0 * kPi / L1,
                       2.0 * kPi / L2, 2.0 * kPi / L3, kc);
    if (hipGetLastError() != hipSuccess) return PSX_E_HIP;
    hipLaunchKernelGGL(stats_final, dim3(1), dim3(NT), 0, s, (const double*)part, nblk, out_dev + 3 * m, 1.0 / Nt,
                       (int*)nullptr);
    if (hipGetLastError() != hipSuccess) return PSX_E_HIP;
  }
  return PSX_OK;
}

namespace {
void check_filter(psx_context* c, int filter, double delta) {
  if (filter != PSX_FILTER_SHARP && filter != PSX_FILTER_BOX) fail(c, PSX_E_ARG, "unknown filter");
  if (filter == PSX_FILTER_BOX && !(delta > 0.0)) fail(c, PSX_E_ARG, "box filter width must be positive");
}

void launch_filter(psx_context* c, const double2* in, double2* out, int filter, double delta) {
  const long long total = 3LL * c->N1 * c->N2 * c->Nc;
  const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 8192);
  hipLaunchKernelGGL(apply_filter, dim3(blocks), dim3(256), 0, c->stream, in, out, 3, c->N1, c->N2, c->Nc, filter,
                     (double)c->p.kc, delta, 2.0 * kPi / c->p.L1, 2.0 * kPi / c->p.L2, 2.0 * kPi / c->p.L3);
  check(c, hipGetLastError(), "apply_filter");
}
}  // namespace

int psx_filter(psx_context* c, const double* uhat, double* out, int filter, double delta) {
  return guarded(c, [&] {
    if (c->members > 1) fail(c, PSX_E_UNSUPPORTED, "one instance per call: use a single-member context");
    if (!uhat || !out) fail(c, PSX_E_ARG, "null pointer");
    check_filter(c, filter, delta);
    launch_filter(c, (const double2*)uhat, (double2*)out, filter, delta);
  });
}

int psx_sgs_dns(psx_context* c, const double* uhat, double* w0_out, double* tau_out, int filter, double delta) {
  return guarded(c, [&] {
    if (c->members > 1) fail(c, PSX_E_UNSUPPORTED, "one instance per call: use a single-member context");
    if (!uhat) fail(c, PSX_E_ARG, "null pointer");
    check_filter(c, filter, delta);
    const int M1 = c->M1, M2 = c->M2, M3 = c->M3, H = c->H;
    for (int L : {M1, M2, M3})
      if (!length_supported(L)) fail(c, PSX_E_UNSUPPORTED, "3/2-rule length " + std::to_string(L) + " not supported");
    const size_t state = (size_t)3 * c->N1 * c->N2 * c->Nc;
    if (!c->sgsA) {
      const size_t MMH = (size_t)M1 * M2 * H, MNH = (size_t)M1 * c->N2 * H;
      const size_t nA = std::max(6 * MNH, 12 * MMH), nB = std::max(6 * MMH, 12 * MNH);
      check(c, hipMalloc(&c->sgsA, nA * sizeof(double2)), "hipMalloc(sgsA)");
      check(c, hipMalloc(&c->sgsB, nB * sizeof(double2)), "hipMalloc(sgsB)");
      check(c, hipMalloc(&c->sgsU, state * sizeof(double2)), "hipMalloc(sgsU)");
      check(c, hipMalloc(&c->sgsT, 2 * state * sizeof(double2)), "hipMalloc(sgsT)");
      c->scratch_bytes += (long long)(nA + nB + 3 * state) * sizeof(double2);
    }
    double2* A = c->sgsA;
    double2* B = c->sgsB;
    const double2* u = (const double2*)uhat;
    launch_filter(c, u, c->sgsU, filter, delta);
    {
      const int jin[3] = {0, 1, 2}, jp0[3] = {0, 1, 2}, jp1[3] = {3, 4, 5};
      x_inverse(c, u, A, M1, H, 0, 3, jin, jp0, nullptr, K_OTHER);
      x_inverse(c, c->sgsU, A, M1, H, 0, 3, jin, jp1, nullptr, K_OTHER);
      const int jin6[6] = {0, 1, 2, 3, 4, 5};
      y_inverse(c, A, B, M1, M2, H, 0, 6, jin6, jin6, nullptr, K_OTHER);
    }
    {
      ZArgs z;
      std::memset(&z, 0, sizeof(z));
      z.in0 = B;
      z.in_sf0 = (long long)M1 * M2 * H;
      z.Hin0 = H;
      z.nin0 = 6;
      z.out = A;
      z.out_sf = (long long)M1 * M2 * H;
      z.Hout = H;
      z.npos = M1 * M2;
      z.kzscale = 2.0 * kPi / c->p.L3;
      const double Nt = (double)c->N1 * c->N2 * c->N3;
      z.out_scale = 1.0 / (std::sqrt(Nt) * (double)M1 * M2 * M3);
      for (int r = 0; r < 6; ++r) z.rec[r] = Recipe{1, {r, 0}, {1.0, 0.0}, {0, 0}};
      const int P = zpass_positions(M3, Z_SGS);
      if (P == 0 || (z.npos % P) != 0) fail(c, PSX_E_UNSUPPORTED, "z pass: unsupported padded length");
      const double2* tw = twiddles(c, M3);
      TimeScope ts(c, K_OTHER);
      check(c, launch_zpass(M3, Z_SGS, z, tw, c->stream), "z_pass(sgs)");
    }
    y_forward(c, A, B, M1, M2, H, 0, 12);
    double2* tau = tau_out ? (double2*)tau_out : c->sgsT;
    {
      SgsArgs s;
      std::memset(&s, 0, sizeof(s));
      s.in = B;
      s.in_sf = (long long)M1 * c->N2 * H;
      s.N1 = c->N1;
      s.N2 = c->N2;
      s.N3 = c->N3;
      s.Nc = c->Nc;
      s.H = H;
      s.tau = tau;
      s.filter = filter;
      s.kc = (double)c->p.kc;
      s.delta = delta;
      s.ks[0] = 2.0 * kPi / c->p.L1;
      s.ks[1] = 2.0 * kPi / c->p.L2;
      s.ks[2] = 2.0 * kPi / c->p.L3;
      const double2* tw = twiddles(c, M1);
      TimeScope ts(c, K_OTHER);
      check(c, launch_xforward_sgs(M1, c->N1, s, tw, c->stream), "x_forward(sgs)");
    }
    if (w0_out) {
      const long long total = (long long)c->N1 * c->N2 * c->Nc;
      const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 8192);
      hipLaunchKernelGGL(sgs_w0, dim3(blocks), dim3(256), 0, c->stream, (const double2*)tau, (double2*)w0_out,
                         c->N1, c->N2, c->Nc, 2.0 * kPi / c->p.L1, 2.0 * kPi / c->p.L2, 2.0 * kPi / c->p.L3);
      check(c, hipGetLastError(), "sgs_w0");
    }
  });
}

namespace {
// x- and y-inverse of nj spectral fields onto the N-grid (Nyquist rows kept, so the
// grid values are those of numpy's irfftn), optionally with d/dx and d/dy outputs
// (field order of the y-inverse: the fields, their d/dx, their d/dy).
void ngrid_xy(psx_context* c, const double2* in, int nj, bool deriv) {
  const int N1 = c->N1, N2 = c->N2, Nc = c->Nc;
  const size_t field = (size_t)N1 * N2 * Nc;
  const size_t nx = (deriv ? 2 : 1) * nj * field, ny = (deriv ? 3 : 1) * nj * field;
  if (nx > c->nA || ny > c->nB) fail(c, PSX_E_UNSUPPORTED, "grid-space output: work buffers too small for this model");
  int jin[MAXJOBS], jp[MAXJOBS], jd[MAXJOBS];
  for (int j = 0; j < nj; ++j) {
    jin[j] = j;
    jp[j] = j;
    jd[j] = nj + j;
  }
  x_inverse(c, in, c->bufA, N1, Nc, 1, nj, jin, jp, deriv ? jd : nullptr, K_OTHER);
  const int ny_jobs = deriv ? 2 * nj : nj;
  for (int j = 0; j < ny_jobs; ++j) {
    jin[j] = j;
    jp[j] = j;
    jd[j] = j < nj ? 2 * nj + j : -1;
  }
  y_inverse(c, c->bufA, c->bufB, N1, N2, Nc, 1, ny_jobs, jin, jp, deriv ? jd : nullptr, K_OTHER);
}

ZArgs ngrid_z(psx_context* c, int nin) {
  ZArgs z;
  std::memset(&z, 0, sizeof(z));
  z.in0 = c->bufB;
  z.in_sf0 = (long long)c->N1 * c->N2 * c->Nc;
  z.Hin0 = c->Nc;
  z.nin0 = nin;
  z.npos = c->N1 * c->N2;
  z.kzscale = 2.0 * kPi / c->p.L3;
  z.out_scale = 1.0 / std::sqrt((double)c->N1 * c->N2 * c->N3);
  return z;
}
}  // namespace

int psx_physical(psx_context* c, const double* fhat, int nfields, double* out) {
  return guarded(c, [&] {
    if (c->members > 1) fail(c, PSX_E_UNSUPPORTED, "one instance per call: use a single-member context");
    if (!fhat || !out || nfields < 1) fail(c, PSX_E_ARG, "null pointer or no fields");
    const long long field = (long long)c->N1 * c->N2 * c->Nc, grid = (long long)c->N1 * c->N2 * c->N3;
    const int P = zpass_positions(c->N3, Z_PHYS);
    if (P == 0 || (c->N2 % P) != 0) fail(c, PSX_E_UNSUPPORTED, "grid-space output: unsupported length");
    const double2* tw = twiddles(c, c->N3);
    for (int f0 = 0; f0 < nfields; f0 += 3) {
      const int nj = std::min(3, nfields - f0);
      ngrid_xy(c, (const double2*)fhat + f0 * field, nj, false);
      ZArgs z = ngrid_z(c, 3);
      for (int r = 0; r < 3; ++r) z.rec[r] = Recipe{1, {std::min(r, nj - 1), 0}, {1.0, 0.0}, {0, 0}};
      z.phys = out + f0 * grid;
      z.out_sf = grid;
      z.nphys = nj;
      TimeScope ts(c, K_OTHER);
      check(c, launch_zpass(c->N3, Z_PHYS, z, tw, c->stream, 0), "z_pass(physical)");
    }
  });
}

int psx_spectral(psx_context* c, const double* phys, int nfields, double* fhat) {
  return guarded(c, [&] {
    if (c->members > 1) fail(c, PSX_E_UNSUPPORTED, "one instance per call: use a single-member context");
    if (!phys || !fhat || nfields < 1) fail(c, PSX_E_ARG, "null pointer or no fields");
    const int N1 = c->N1, N2 = c->N2, N3 = c->N3, Nc = c->Nc;
    const long long field = (long long)N1 * N2 * Nc, grid = (long long)N1 * N2 * N3;
    const double2* tw3 = twiddles(c, N3);
    const double2* tw2 = twiddles(c, N2);
    const double2* tw1 = twiddles(c, N1);
    for (int f0 = 0; f0 < nfields; f0 += 3) {
      const int nj = std::min(3, nfields - f0);
      {
        TimeScope ts(c, K_OTHER);
        check(c, launch_zreal_forward(N3, phys + f0 * grid, grid, c->bufA, field, N1 * N2, nj, tw3, c->stream),
              "z_forward(real)");
      }
      // y then x, c2c forward on the N grid (every index kept)
      PencilArgs a = pencil_base();
      a.in = c->bufA;
      a.out = c->bufB;
      a.in_sf = a.out_sf = field;
      a.in_sa = a.out_sa = (long long)N2 * Nc;
      a.in_sl = a.out_sl = Nc;
      a.A = N1;
      a.C = Nc;
      a.Nspec = N2;
      a.keep_nyq = 1;
      for (int j = 0; j < nj; ++j) a.job_in[j] = a.job_plain[j] = j;
      set_members(c, a, nj);
      {
        TimeScope ts(c, K_OTHER);
        check(c, launch_pencil_forward(N2, a, tw2, nj, c->stream), "y_forward(spectral)");
      }
      PencilArgs b = pencil_base();
      b.in = c->bufB;
      b.out = (double2*)fhat + f0 * field;
      b.in_sf = b.out_sf = field;
      b.in_sa = b.out_sa = Nc;
      b.in_sl = b.out_sl = (long long)N2 * Nc;
      b.A = N2;
      b.C = Nc;
      b.Nspec = N1;
      b.keep_nyq = 1;
      b.scale = 1.0 / std::sqrt((double)N1 * N2 * N3);  // uhat = rfftn(u) / sqrt(Nt)
      for (int j = 0; j < nj; ++j) b.job_in[j] = b.job_plain[j] = j;
      b.jpm = nj;
      b.in_ms = b.out_ms = 0;
      {
        TimeScope ts(c, K_OTHER);
        check(c, launch_pencil_forward(N1, b, tw1, nj, c->stream), "x_forward(spectral)");
      }
    }
  });
}

int psx_gradients(psx_context* c, const double* uhat, float* out) {
  return guarded(c, [&] {
    if (c->members > 1) fail(c, PSX_E_UNSUPPORTED, "one instance per call: use a single-member context");
    if (!uhat || !out) fail(c, PSX_E_ARG, "null pointer");
    const int P = zpass_positions(c->N3, Z_GRAD);
    if (P == 0 || (c->N2 % P) != 0) fail(c, PSX_E_UNSUPPORTED, "gradients: unsupported length");
    ngrid_xy(c, (const double2*)uhat, 3, true);
    ZArgs z = ngrid_z(c, 9);
    // y-inverse fields u v w ux vx wx uy vy wy -> du_i/dx_j, (i, j) row-major
    const int src[9] = {3, 6, 0, 4, 7, 1, 5, 8, 2};
    for (int r = 0; r < 9; ++r) z.rec[r] = Recipe{1, {src[r], 0}, {1.0, 0.0}, {r % 3 == 2 ? 1 : 0, 0}};
    z.gout = out;
    const double2* tw = twiddles(c, c->N3);
    TimeScope ts(c, K_OTHER);
    check(c, launch_zpass(c->N3, Z_GRAD, z, tw, c->stream, 0), "z_pass(gradients)");
  });
}

int psx_restrict(psx_context* c, const double* src, int s1, int s2, int s3, int nfields, double* out) {
  return guarded(c, [&] {
    if (!src || !out || nfields < 1) fail(c, PSX_E_ARG, "null pointer or no fields");
    if (s1 < c->N1 || s2 < c->N2 || s3 < c->N3 || (s1 % 2) || (s2 % 2) || (s3 % 2))
      fail(c, PSX_E_ARG, "the source grid must be even and at least as fine as the context grid");
    const long long total = (long long)nfields * c->N1 * c->N2 * c->Nc;
    const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 8192);
    const double scale = std::sqrt(((double)c->N1 * c->N2 * c->N3) / ((double)s1 * s2 * s3));
    hipLaunchKernelGGL(restrict_modes, dim3(blocks), dim3(256), 0, c->stream, (const double2*)src, (double2*)out,
                       nfields, c->N1, c->N2, c->Nc, s1, s2, s3 / 2 + 1, scale);
    check(c, hipGetLastError(), "restrict_modes");
  });
}

int psx_set_closure(psx_context* c, psx_mlp* net, double bound, int flags) {
  return guarded(c, [&] {
    if (c->model != PSX_MODEL_LEARNED) fail(c, PSX_E_ARG, "psx_set_closure needs a PSX_MODEL_LEARNED context");
    if (std::isnan(bound)) fail(c, PSX_E_ARG, "bound is NaN");
    if (flags & ~PSX_CLOSURE_CLIP_BACKSCATTER) fail(c, PSX_E_ARG, "unknown closure flags");
    c->net = net;
    c->net_bound = bound;
    c->net_flags = flags;
    drop_graph(c);  // a captured step holds the network's normalisation and the bound by value
  });
}

int psx_spectrum_bins(const psx_context* c) {
  if (!c) return PSX_E_ARG;
  const double a = c->N1 / 2, b = c->N2 / 2, d = c->N3 / 2 - 1;
  return (int)std::lrint(std::sqrt(a * a + b * b + d * d)) + 1;
}

int psx_spectrum(psx_context* c, const double* uhat, double* spec_dev, int resolved) {
  return guarded(c, [&] {
    if (c->members > 1) fail(c, PSX_E_UNSUPPORTED, "one instance per call: use a single-member context");
    if (!uhat || !spec_dev) fail(c, PSX_E_ARG, "null pointer");
    const int nb = psx_spectrum_bins(c);
    hipLaunchKernelGGL(fill_f64, dim3((3 * nb + 255) / 256), dim3(256), 0, c->stream, spec_dev, 3 * nb, 0.0);
    check(c, hipGetLastError(), "fill_f64");
    const long long total = (long long)c->N1 * c->N2 * c->Nc;
    const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 4096);
    const double Nt = (double)c->N1 * c->N2 * c->N3;
    hipLaunchKernelGGL(spectrum_kernel, dim3(blocks), dim3(256), 0, c->stream, (const double2*)uhat, spec_dev, nb,
                       c->N1, c->N2, c->N3, c->Nc, c->p.kc, resolved ? 1 : 0, 1.0 / Nt);
    check(c, hipGetLastError(), "spectrum");
  });
}

int psx_tau(psx_context* c, double* out) {
  return guarded(c, [&] {
    if (!out) fail(c, PSX_E_ARG, "null pointer");
    check(c, hipMemcpyAsync(out, c->scal + 1, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
    check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  });
}

int psx_timing_begin(psx_context* c) {
  return guarded(c, [&] {
    c->timing = true;
    c->events_used = 0;
    c->event_kinds.clear();
  });
}

int psx_timing_end(psx_context* c, double* ms, int* count) {
  return guarded(c, [&] {
    if (!ms || !count) fail(c, PSX_E_ARG, "null pointer");
    c->timing = false;
    check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (int k = 0; k < K_NKIND; ++k) {
      ms[k] = 0.0;
      count[k] = 0;
    }
    for (size_t i = 0; i < c->event_kinds.size(); ++i) {
      float t = 0.f;
      check(c, hipEventElapsedTime(&t, c->events[2 * i], c->events[2 * i + 1]), "hipEventElapsedTime");
      ms[c->event_kinds[i]] += t;
      count[c->event_kinds[i]] += 1;
    }
    c->events_used = 0;
    c->event_kinds.clear();
  });
}

}  // extern "C"

// Buffers, stream and events of the chunked y/z pipeline (allocated outside any capture).
void ensure_pipeline(psx_context* c) {
  if (c->bufX) return;
  const bool smag = c->model == PSX_MODEL_SMAGORINSKY || c->model == PSX_MODEL_DSMAG;
  const size_t MNH = (size_t)c->M1 * c->N2 * c->H;
  c->nX = (smag ? 6 : 3) * MNH;
  c->nW = (smag ? 12 : 6) * MNH;
  const size_t mb = (size_t)c->members;
  check(c, hipMalloc(&c->bufX, mb * c->nX * sizeof(double2)), "hipMalloc(bufX)");
  check(c, hipMalloc(&c->bufW, mb * c->nW * sizeof(double2)), "hipMalloc(bufW)");
  c->scratch_bytes += (long long)(mb * (c->nX + c->nW) * sizeof(double2));
  check(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking), "hipStreamCreate");
  check(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), "hipEventCreate");
  check(c, hipEventCreateWithFlags(&c->ev_stagger, hipEventDisableTiming), "hipEventCreate");
  check(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming), "hipEventCreate");
}

// ---- kernel-shape autotuning -----------------------------------------------------
extern "C" {

namespace {
bool tau_family(int model) { return model >= PSX_MODEL_TMODEL && model <= PSX_MODEL_DYNAMIC_TAU; }
// Transform lengths of the tuned x-inverse, y-inverse and y-forward passes: the 3/2-rule
// lengths, or the 2x lengths of the t-model family.
void tuned_lengths(const psx_context* c, int* L) {
  const bool tau = tau_family(c->model);
  L[0] = tau ? 2 * c->N1 : c->M1;
  L[1] = L[2] = tau ? 2 * c->N2 : c->M2;
}
}  // namespace

int psx_set_tuning(psx_context* c, const int* tune) {
  return guarded(c, [&] {
    if (!tune) fail(c, PSX_E_ARG, "null pointer");
    int L[3];
    tuned_lengths(c, L);
    for (int k = 0; k < 3; ++k)
      if (tune[k] != 0 && pencil_tile(L[k], tune[k]) < 0) fail(c, PSX_E_ARG, "tile width not available");
    if (tune[3] != 0 && tune[3] != 1) fail(c, PSX_E_ARG, "z variant must be 0 or 1");
    if (tune[4] < 0 || tune[4] > 8 || (tune[4] > 1 && c->M1 < 2 * tune[4])) fail(c, PSX_E_ARG, "bad chunk count");
    if (tune[5] < 0 || tune[5] > 2) fail(c, PSX_E_ARG, "chunk streams must be 0, 1 or 2");
    if (tune[6] != 0 && tune[6] != 1) fail(c, PSX_E_ARG, "closure placement must be 0 or 1");
    for (int k = 7; k < 10; ++k)
      if (tune[k] < 0 || tune[k] > 2) fail(c, PSX_E_ARG, "block order must be 0, 1 or 2");
    if (tune[10] != 0 && tune[10] != 1) fail(c, PSX_E_ARG, "stress sums must be 0 (y-forward) or 1 (epilogue)");
    if (tune[4] > 1) {
      if (c->model != PSX_MODEL_DNS && c->model != PSX_MODEL_SMAGORINSKY && c->model != PSX_MODEL_DSMAG)
        fail(c, PSX_E_UNSUPPORTED, "chunked pipeline: DNS, Smagorinsky and DSMAG");
      ensure_pipeline(c);
    }
    for (int k = 0; k < PSX_TUNE_N; ++k) c->tune[k] = tune[k];
    drop_graph(c);  // the captured step was built with the old shapes
  });
}

int psx_get_tuning(const psx_context* c, int* tune) {
  if (!c || !tune) return PSX_E_ARG;
  for (int k = 0; k < PSX_TUNE_N; ++k) tune[k] = c->tune[k];
  return PSX_OK;
}

int psx_autotune(psx_context* c, const double* uhat, int reps, int* tune_out) {
  return guarded(c, [&] {
    if (!uhat) fail(c, PSX_E_ARG, "null pointer");
    const bool learned = c->model == PSX_MODEL_LEARNED;
    const bool tau = tau_family(c->model);
    if (reps < 1) reps = 1;
    const size_t state = (size_t)c->members * 3 * c->N1 * c->N2 * c->Nc;
    double2* tmp = nullptr;
    check(c, hipMalloc(&tmp, state * sizeof(double2)), "hipMalloc(autotune)");
    const int kinds[PSX_TUNE_N] = {K_XINV, K_YINV, K_YFWD, K_ZPASS, -1, -1, -1,  // -1: the whole stage
                                   K_XINV, K_YINV, K_YFWD, -1};
    const bool smag_form = c->model != PSX_MODEL_DNS && !tau;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    check(c, hipEventCreate(&t0), "hipEventCreate");
    check(c, hipEventCreate(&t1), "hipEventCreate");
    int L[3];
    tuned_lengths(c, L);
    const bool was_timing = c->timing;
    try {
      for (int knob = 0; knob < PSX_TUNE_N; ++knob) {
        if (learned && (knob == 4 || knob == 5)) continue;  // the learned stage is not chunked
        if (!learned && knob == 6) continue;                 // closure placement: learned only
        if (!smag_form && knob == 10) continue;              // stress sums: Smagorinsky-form stages
        // the t-model family tunes its pencil passes' tiles and block orders (its z passes
        // have one variant, and it is not chunked)
        if (tau && (knob >= 3 && knob <= 6)) continue;
        // knob 4: chunk count and streams together (v = chunks + 16 * streams)
        if (knob == 5) continue;
        std::vector<int> cand;
        if (knob < 3) {
          for (int T : {4, 8, 16, 24, 32})
            if (pencil_tile(L[knob], T) == T) cand.push_back(T);
        } else if (knob >= 7 && knob <= 9) {
          cand = {0, 1, 2};  // block orders
        } else if (knob == 3 || knob >= 6) {
          cand = {0, 1};
          // the stress sums change the folded stage only (the autotuner's RHS calls pass
          // no w0, so they time exactly that stage)
        } else {
          cand = {1 + 16 * 2, 2 + 16 * 2, 4 + 16 * 2, 2 + 16 * 1, 4 + 16 * 1, 8 + 16 * 1};
          ensure_pipeline(c);
        }
        double best = 1e300;
        int pick = knob == 4 ? c->tune[4] + 16 * c->tune[5] : c->tune[knob];
        for (int v : cand) {
          if (knob == 4) {
            if ((v & 15) > 1 && c->M1 < 2 * (v & 15)) continue;
            c->tune[4] = v & 15;
            c->tune[5] = v >> 4;
          } else {
            c->tune[knob] = v;
          }
          run_stage(c, (const double2*)uhat, nullptr, tmp, nullptr, 0.0, nullptr, 0, 0.0);  // warm-up
          double ms = 0.0;
          int cnt = 0;
          if (kinds[knob] < 0) {
            check(c, hipEventRecord(t0, c->stream), "hipEventRecord");
            for (int r = 0; r < reps; ++r)
              run_stage(c, (const double2*)uhat, nullptr, tmp, nullptr, 0.0, nullptr, 0, 0.0);
            check(c, hipEventRecord(t1, c->stream), "hipEventRecord");
            check(c, hipEventSynchronize(t1), "hipEventSynchronize");
            float t = 0.f;
            check(c, hipEventElapsedTime(&t, t0, t1), "hipEventElapsedTime");
            ms = t;
            cnt = reps;
          } else {
            c->timing = true;
            c->events_used = 0;
            c->event_kinds.clear();
            for (int r = 0; r < reps; ++r)
              run_stage(c, (const double2*)uhat, nullptr, tmp, nullptr, 0.0, nullptr, 0, 0.0);
            check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            for (size_t i = 0; i < c->event_kinds.size(); ++i) {
              if (c->event_kinds[i] != kinds[knob]) continue;
              float t = 0.f;
              check(c, hipEventElapsedTime(&t, c->events[2 * i], c->events[2 * i + 1]), "hipEventElapsedTime");
              ms += t;
              ++cnt;
            }
          }
          c->timing = false;
          c->events_used = 0;
          c->event_kinds.clear();
          const double avg = cnt ? ms / cnt : 1e300;
          // whole-stage timings (chunking) vary by a few percent from run to run: a
          // chunked pipeline must beat the best stage so far by 5% to be picked
          const double margin = (knob == 4 && best < 1e300 && (v & 15) > 1) ? 0.95 : 1.0;
          if (avg < best * margin) {
            best = avg;
            pick = v;
          }
        }
        if (knob == 4) {
          c->tune[4] = pick & 15;
          c->tune[5] = pick >> 4;
        } else {
          c->tune[knob] = pick;
        }
      }
    } catch (...) {
      c->timing = was_timing;
      hipFree(tmp);
      hipEventDestroy(t0);
      hipEventDestroy(t1);
      throw;
    }
    c->timing = was_timing;
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    check(c, hipFree(tmp), "hipFree(autotune)");
    if (tune_out)
      for (int k = 0; k < PSX_TUNE_N; ++k) tune_out[k] = c->tune[k];
    drop_graph(c);
  });
}

}  // extern "C"

// ---- ensembles of uneven sizes ------------------------------------------------------
// Members are grouped by grid size into batch contexts (one per size class, in order of
// first appearance); the ensemble state is class-major, each class its members in
// order.  A step runs every class's batched step; the graph variant captures all of
// them as one hipGraph on the ensemble's stream.
struct psx_ensemble {
  std::vector<psx_context*> cls;
  std::vector<long long> cls_offset;  // complex offset of a class's first member
  std::vector<int> cls_first_slot;
  std::vector<int> slot;              // member -> slot (class-major order)
  std::vector<long long> offset;      // member -> complex offset
  long long total = 0;
  int count = 0;
  hipStream_t stream = nullptr;
  const void* key[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<unsigned long long> gens;  // each class context's settings generation at capture
  // size classes run concurrently: class k on cls_stream[k], forked from and joined to the
  // ensemble stream (each class's launches fill only part of the GPU)
  std::vector<hipStream_t> cls_stream;
  std::vector<hipEvent_t> cls_join;
  hipEvent_t fork = nullptr;
  // psx_ensemble_step_graph: one graph per class, captured on the class's stream (a
  // single-stream capture, as psx_step_graph's), replayed concurrently
  std::vector<hipGraph_t> cls_graph;
  std::vector<hipGraphExec_t> cls_exec;
  std::string err;
};

namespace {
// One ensemble step of every class, each class on its own stream forked from the
// ensemble stream and joined back (t: step-start time for the eager step).
int ensemble_step_classes(psx_ensemble* e, double* q, double* qs, double* dt_dev, double t) {
  if (hipEventRecord(e->fork, e->stream) != hipSuccess) {
    e->err = "hipEventRecord failed";
    return PSX_E_HIP;
  }
  int rc = PSX_OK;
  for (size_t k = 0; k < e->cls.size() && rc == PSX_OK; ++k) {
    psx_context* c = e->cls[k];
    if (hipStreamWaitEvent(e->cls_stream[k], e->fork, 0) != hipSuccess) {
      e->err = "hipStreamWaitEvent failed";
      return PSX_E_HIP;
    }
    c->stream = e->cls_stream[k];
    rc = guarded(c, [&] {
      step_body(c, (double2*)q + e->cls_offset[k], (double2*)qs + e->cls_offset[k], dt_dev + e->cls_first_slot[k], t,
                nullptr);
      check(c, hipEventRecord(e->cls_join[k], c->stream), "hipEventRecord");
      check(c, hipStreamWaitEvent(e->stream, e->cls_join[k], 0), "hipStreamWaitEvent");
    });
    c->stream = e->stream;
    if (rc != PSX_OK) e->err = c->err;
  }
  return rc;
}

// PSX_E_NONFINITE when any class context's non-finite flag is set
int ensemble_status(const psx_ensemble* e, int rc) {
  for (const psx_context* c : e->cls) rc = with_status(c, rc);
  return rc;
}
}  // namespace


extern "C" {

int psx_ensemble_create(psx_ensemble** out, int count, const int* dims, const double* nu, int model,
                        const psx_params* common, int device, void* stream) {
  if (!out || !dims || count < 1 || !common) return PSX_E_ARG;
  *out = nullptr;
  psx_ensemble* e = new psx_ensemble();
  e->count = count;
  e->stream = (hipStream_t)stream;
  std::vector<std::array<int, 3>> sizes;
  std::vector<std::vector<int>> members;
  for (int i = 0; i < count; ++i) {
    const std::array<int, 3> d = {dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]};
    size_t k = 0;
    while (k < sizes.size() && sizes[k] != d) ++k;
    if (k == sizes.size()) {
      sizes.push_back(d);
      members.emplace_back();
    }
    members[k].push_back(i);
  }
  e->slot.assign(count, 0);
  e->offset.assign(count, 0);
  int next_slot = 0;
  for (size_t k = 0; k < sizes.size(); ++k) {
    const int n1 = sizes[k][0], n2 = sizes[k][1], n3 = sizes[k][2];
    psx_params p = *common;
    p.dx = p.L1 / n1;
    p.dy = p.L2 / n2;
    p.dz = p.L3 / n3;
    p.kc = n1 / 2;
    std::vector<double> nuk;
    for (int i : members[k]) nuk.push_back(nu ? nu[i] : common->nu);
    p.nu = nuk[0];
    psx_context* c = nullptr;
    const int rc = psx_create_batch(&c, n1, n2, n3, model, &p, (int)nuk.size(), nuk.data(), device, stream);
    if (rc != PSX_OK) {
      for (psx_context* x : e->cls) psx_destroy(x);
      delete e;
      return rc;
    }
    e->cls.push_back(c);
    e->cls_offset.push_back(e->total);
    e->cls_first_slot.push_back(next_slot);
    const long long per = 3LL * n1 * n2 * (n3 / 2 + 1);
    for (size_t j = 0; j < members[k].size(); ++j) {
      e->slot[members[k][j]] = next_slot + (int)j;
      e->offset[members[k][j]] = e->total + (long long)j * per;
    }
    next_slot += (int)members[k].size();
    e->total += per * (long long)members[k].size();
  }
  bool ok = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess;
  for (size_t k = 0; k < e->cls.size() && ok; ++k) {
    hipStream_t s = nullptr;
    hipEvent_t j = nullptr;
    ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
    if (ok) e->cls_stream.push_back(s);
    ok = ok && hipEventCreateWithFlags(&j, hipEventDisableTiming) == hipSuccess;
    if (ok) e->cls_join.push_back(j);
  }
  if (!ok) {
    psx_ensemble_destroy(e);
    return PSX_E_HIP;
  }
  *out = e;
  return PSX_OK;
}

int psx_ensemble_destroy(psx_ensemble* e) {
  if (!e) return PSX_E_ARG;
  for (hipGraphExec_t x : e->cls_exec)
    if (x) hipGraphExecDestroy(x);
  for (hipGraph_t g : e->cls_graph)
    if (g) hipGraphDestroy(g);
  for (psx_context* c : e->cls) psx_destroy(c);
  for (hipStream_t s : e->cls_stream) hipStreamDestroy(s);
  for (hipEvent_t j : e->cls_join) hipEventDestroy(j);
  if (e->fork) hipEventDestroy(e->fork);
  delete e;
  return PSX_OK;
}

const char* psx_ensemble_last_error(const psx_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

int psx_ensemble_set_closure(psx_ensemble* e, psx_mlp* net, double bound, int flags) {
  if (!e) return PSX_E_ARG;
  for (psx_context* c : e->cls) {
    const int rc = psx_set_closure(c, net, bound, flags);
    if (rc != PSX_OK) {
      e->err = c->err;
      return rc;
    }
  }
  // the captured ensemble step is rebuilt: psx_set_closure advanced every class's generation
  return PSX_OK;
}
int psx_ensemble_classes(const psx_ensemble* e) { return e ? (int)e->cls.size() : PSX_E_ARG; }

int psx_ensemble_autotune(psx_ensemble* e, const double* q, int reps, int* tune_out) {
  if (!e || !q) return PSX_E_ARG;
  if (reps < 1) reps = 1;
  // Coordinate descent on the time of the whole ensemble step, the classes running
  // concurrently as psx_ensemble_step runs them (a class timed alone favours shapes
  // that fill the GPU by themselves): for each class and knob, every candidate is timed
  // on reps steps from q (copied, so q is not changed) and the fastest kept.
  const size_t bytes = (size_t)e->total * sizeof(double2);
  double2* tq = nullptr;
  double2* tqs = nullptr;
  double* tdt = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  int rc = PSX_OK;
  if (hipMalloc(&tq, bytes) != hipSuccess || hipMalloc(&tqs, bytes) != hipSuccess ||
      hipMalloc(&tdt, sizeof(double) * e->count) != hipSuccess || hipEventCreate(&t0) != hipSuccess ||
      hipEventCreate(&t1) != hipSuccess) {
    e->err = "psx_ensemble_autotune: allocation failed";
    rc = PSX_E_HIP;
  }
  auto trial = [&](double& ms) -> int {
    if (hipMemcpyAsync(tq, q, bytes, hipMemcpyDeviceToDevice, e->stream) != hipSuccess ||
        hipMemcpyAsync(tqs, q, bytes, hipMemcpyDeviceToDevice, e->stream) != hipSuccess) {
      e->err = "psx_ensemble_autotune: hipMemcpyAsync failed";
      return PSX_E_HIP;
    }
    int r = ensemble_step_classes(e, (double*)tq, (double*)tqs, tdt, 0.0);  // warm-up
    if (r != PSX_OK) return r;
    if (hipEventRecord(t0, e->stream) != hipSuccess) return PSX_E_HIP;
    for (int i = 0; i < reps && r == PSX_OK; ++i) r = ensemble_step_classes(e, (double*)tq, (double*)tqs, tdt, 0.0);
    if (r != PSX_OK) return r;
    float f = 0.f;
    if (hipEventRecord(t1, e->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
        hipEventElapsedTime(&f, t0, t1) != hipSuccess) {
      e->err = "psx_ensemble_autotune: event timing failed";
      return PSX_E_HIP;
    }
    ms = f;
    return PSX_OK;
  };
  // tiles, z-pass variant, block orders, stress sums (chunking and the closure's
  // placement stay as they are)
  const int knobs[] = {0, 1, 2, 3, 7, 8, 9, 10};
  for (size_t k = 0; k < e->cls.size() && rc == PSX_OK; ++k) {
    psx_context* c = e->cls[k];
    int L[3];
    tuned_lengths(c, L);
    const bool smag_form = c->model != PSX_MODEL_DNS && !tau_family(c->model);
    for (int knob : knobs) {
      if (rc != PSX_OK) break;
      if (knob == 10 && !smag_form) continue;
      std::vector<int> cand{c->tune[knob]};  // the current choice first: others must beat it
      if (knob < 3) {
        for (int T : {4, 8, 16, 24, 32})
          if (pencil_tile(L[knob], T) == T && T != c->tune[knob]) cand.push_back(T);
      } else if (knob >= 7 && knob <= 9) {
        for (int v : {0, 1, 2})
          if (v != c->tune[knob]) cand.push_back(v);
      } else {
        cand.push_back(c->tune[knob] ? 0 : 1);
      }
      double best = 1e300;
      int pick = c->tune[knob];
      for (int v : cand) {
        c->tune[knob] = v;
        double ms = 0.0;
        rc = trial(ms);
        if (rc != PSX_OK) break;
        if (ms < (best < 1e300 ? 0.99 * best : best)) {
          best = ms;
          pick = v;
        }
      }
      c->tune[knob] = pick;
    }
  }
  for (size_t k = 0; k < e->cls.size(); ++k) {
    psx_context* c = e->cls[k];
    drop_graph(c);  // the captured ensemble step is rebuilt: every class's generation advanced
    if (tune_out)
      for (int i = 0; i < PSX_TUNE_N; ++i) tune_out[k * PSX_TUNE_N + i] = c->tune[i];
  }
  if (hipStreamSynchronize(e->stream) != hipSuccess && rc == PSX_OK) rc = PSX_E_HIP;
  hipFree(tq);
  hipFree(tqs);
  hipFree(tdt);
  if (t0) hipEventDestroy(t0);
  if (t1) hipEventDestroy(t1);
  return rc;
}
long long psx_ensemble_size(const psx_ensemble* e) { return e ? e->total : PSX_E_ARG; }
long long psx_ensemble_offset(const psx_ensemble* e, int member) {
  return (e && member >= 0 && member < e->count) ? e->offset[member] : PSX_E_ARG;
}
int psx_ensemble_slot(const psx_ensemble* e, int member) {
  return (e && member >= 0 && member < e->count) ? e->slot[member] : PSX_E_ARG;
}

int psx_ensemble_set_stream(psx_ensemble* e, void* stream) {
  if (!e) return PSX_E_ARG;
  e->stream = (hipStream_t)stream;
  for (psx_context* c : e->cls) c->stream = (hipStream_t)stream;
  return PSX_OK;
}

int psx_ensemble_step(psx_ensemble* e, double* q, double* qs, double* dt_dev, double t) {
  if (!e) return PSX_E_ARG;
  if (!q || !qs || !dt_dev || q == qs) return PSX_E_ARG;
  const int rc = ensemble_step_classes(e, q, qs, dt_dev, t);
  return rc != PSX_OK ? rc : ensemble_status(e, PSX_OK);
}

int psx_ensemble_status(psx_ensemble* e) {
  if (!e) return PSX_E_ARG;
  if (hipStreamSynchronize(e->stream) != hipSuccess) {
    e->err = "hipStreamSynchronize failed";
    return PSX_E_HIP;
  }
  return ensemble_status(e, PSX_OK);
}

int psx_ensemble_step_graph(psx_ensemble* e, double* q, double* qs, double* dt_dev) {
  if (!e) return PSX_E_ARG;
  if (!q || !qs || !dt_dev || q == qs) return PSX_E_ARG;
  if (e->stream == nullptr) {
    e->err = "graph capture needs a non-default stream (psx_ensemble_set_stream)";
    return PSX_E_ARG;
  }
  const void* key[4] = {q, qs, dt_dev, e->stream};
  bool same = e->cls_exec.size() == e->cls.size() && e->gens.size() == e->cls.size();
  for (int i = 0; i < 4 && same; ++i) same = e->key[i] == key[i];
  for (size_t k = 0; k < e->cls.size() && same; ++k) same = e->gens[k] == e->cls[k]->gen;
  if (!same) {
    for (hipGraphExec_t x : e->cls_exec)
      if (x) hipGraphExecDestroy(x);
    for (hipGraph_t g : e->cls_graph)
      if (g) hipGraphDestroy(g);
    e->cls_exec.assign(e->cls.size(), nullptr);
    e->cls_graph.assign(e->cls.size(), nullptr);
    e->gens.clear();
    // each class's step captured on its own stream (the ensemble stream is not in the capture)
    for (size_t k = 0; k < e->cls.size(); ++k) {
      psx_context* c = e->cls[k];
      const hipStream_t cs = e->cls_stream[k];
      const bool timing = c->timing;
      c->timing = false;
      c->stream = cs;
      int rc = PSX_OK;
      if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        e->err = "hipStreamBeginCapture failed";
        rc = PSX_E_HIP;
      } else {
        rc = guarded(c, [&] {
          step_body(c, (double2*)q + e->cls_offset[k], (double2*)qs + e->cls_offset[k], dt_dev + e->cls_first_slot[k],
                    0.0, nullptr);
        });
        if (rc != PSX_OK) e->err = c->err;
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(cs, &g);
        if (rc == PSX_OK && ec != hipSuccess) {
          e->err = "hipStreamEndCapture failed";
          rc = PSX_E_HIP;
        }
        e->cls_graph[k] = g;
        if (rc == PSX_OK && hipGraphInstantiate(&e->cls_exec[k], g, nullptr, nullptr, 0) != hipSuccess) {
          e->err = "hipGraphInstantiate failed";
          rc = PSX_E_HIP;
        }
      }
      c->timing = timing;
      c->stream = e->stream;
      if (rc != PSX_OK) return rc;  // gens stays short: the next call captures again
    }
    for (int i = 0; i < 4; ++i) e->key[i] = key[i];
    for (const psx_context* c : e->cls) e->gens.push_back(c->gen);
  }
  // the class graphs run concurrently, forked from and joined to the ensemble stream
  if (hipEventRecord(e->fork, e->stream) != hipSuccess) {
    e->err = "hipEventRecord failed";
    return PSX_E_HIP;
  }
  for (size_t k = 0; k < e->cls.size(); ++k) {
    if (hipStreamWaitEvent(e->cls_stream[k], e->fork, 0) != hipSuccess ||
        hipGraphLaunch(e->cls_exec[k], e->cls_stream[k]) != hipSuccess ||
        hipEventRecord(e->cls_join[k], e->cls_stream[k]) != hipSuccess ||
        hipStreamWaitEvent(e->stream, e->cls_join[k], 0) != hipSuccess) {
      e->err = "hipGraphLaunch failed";
      return PSX_E_HIP;
    }
  }
  return ensemble_status(e, PSX_OK);
}

int psx_ensemble_stats(psx_ensemble* e, const double* q, double* out_dev) {
  if (!e || !q || !out_dev) return PSX_E_ARG;
  for (size_t k = 0; k < e->cls.size(); ++k) {
    const int rc = psx_stats(e->cls[k], (const double*)((const double2*)q + e->cls_offset[k]),
                             out_dev + 3 * e->cls_first_slot[k]);
    if (rc != PSX_OK && rc != PSX_E_NONFINITE) {
      e->err = e->cls[k]->err;
      return rc;
    }
  }
  return ensemble_status(e, PSX_OK);
}

}  // extern "C"
