// extern "C" boundary of libnusht_amd.so (declared in include/nusht_amd.h).
#include "../../include/nusht_amd.h"

#include <algorithm>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "fft2d.h"
#include "legendre.h"
#include "nufft.h"
#include "points.h"
#include "ring.h"
#include "stages.h"
#include "timing.h"

using namespace nusht;

struct nusht_plan_s {
  NufftParams np;
  int lmax = 0, spin = 0, ncomp = 1;
  long npts_max = 0;
  LegendreTables tab;
  std::unique_ptr<NufftPlan> nu;
  DevBuf coef, G, ptg;
  int nmrows = 0;
  int max_maps = kMaxMaps;  // maps per batched synthesis the G buffer holds
  NufftTune tune;           // plan-time nuFFT measurement (candidates = 0: not tuned)
  static constexpr int kMaxMaps = 8;
  size_t bytes() const { return tab.bytes() + nu->bytes() + coef.bytes() + G.bytes() + ptg.bytes(); }
};

struct nusht_ringplan_s {
  std::unique_ptr<RingPlan> rp;
};

struct nusht_pointplan_s {
  std::unique_ptr<PointPlan> pp;
};

namespace {

template <typename F>
int guard(F&& f) {
  try {
    f();
    return NUSHT_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return NUSHT_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return NUSHT_ERR_HIP;
  }
}

inline hipStream_t S(nusht_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

void check_plan(nusht_plan_t p) { NUSHT_REQUIRE(p != nullptr, kErrArg, "null plan"); }

}  // namespace

extern "C" {

const char* nusht_last_error(void) { return last_error(); }
int nusht_version(void) { return 1; }

// ------------------------------------------------------------------ stages
int nusht_doubling_f64(const double* map, int ntheta, int nphi, int spin, double* dmap, nusht_stream_t s) {
  return guard([&] { doubling<double>(map, ntheta, nphi, spin, dmap, S(s)); });
}
int nusht_doubling_f32(const float* map, int ntheta, int nphi, int spin, float* dmap, nusht_stream_t s) {
  return guard([&] { doubling<float>(map, ntheta, nphi, spin, dmap, S(s)); });
}
int nusht_doubling_c128(const double* map, int ntheta, int nphi, int spin, double* dmap, nusht_stream_t s) {
  return guard([&] {
    doubling<double2>(reinterpret_cast<const double2*>(map), ntheta, nphi, spin, reinterpret_cast<double2*>(dmap), S(s));
  });
}
int nusht_adjoint_doubling_f64(const double* dmap, int ntheta, int nphi, int spin, double* map, nusht_stream_t s) {
  return guard([&] { adjoint_doubling<double>(dmap, ntheta, nphi, spin, map, S(s)); });
}
int nusht_adjoint_doubling_f32(const float* dmap, int ntheta, int nphi, int spin, float* map, nusht_stream_t s) {
  return guard([&] { adjoint_doubling<float>(dmap, ntheta, nphi, spin, map, S(s)); });
}
int nusht_adjoint_doubling_c128(const double* dmap, int ntheta, int nphi, int spin, double* map, nusht_stream_t s) {
  return guard([&] {
    adjoint_doubling<double2>(reinterpret_cast<const double2*>(dmap), ntheta, nphi, spin,
                              reinterpret_cast<double2*>(map), S(s));
  });
}
int nusht_pointing_f64(const double* theta, const double* phi0, const int64_t* nph, const int64_t* ofs, int nring,
                       int64_t max_nph, int64_t npix, const double* d_theta, const double* d_phi, double* out_theta,
                       double* out_phi, double* out_gamma, nusht_stream_t s) {
  return guard([&] {
    NUSHT_REQUIRE(theta && phi0 && nph && ofs && d_theta && d_phi && out_theta && out_phi, kErrArg, "null array");
    pointing(theta, phi0, reinterpret_cast<const long*>(nph), reinterpret_cast<const long*>(ofs), nring, max_nph,
             npix, d_theta, d_phi, out_theta, out_phi, out_gamma, S(s));
  });
}

int nusht_magnification_f64(const double* kappa, const double* g1, const double* g2, const double* d1,
                            const double* d2, const double* omega, int64_t npix, int exact, double* A,
                            nusht_stream_t s) {
  return guard([&] {
    NUSHT_REQUIRE(kappa && g1 && g2 && d1 && d2 && A, kErrArg, "null array");
    magnification(kappa, g1, g2, d1, d2, omega, npix, exact != 0, A, S(s));
  });
}

int nusht_almxfl_c128(const double* alm, int lmax, int ncomp, const double* fl, double* out, nusht_stream_t s) {
  return guard([&] {
    NUSHT_REQUIRE(alm && fl && out, kErrArg, "null array");
    almxfl(reinterpret_cast<const double2*>(alm), lmax, ncomp, fl, reinterpret_cast<double2*>(out), S(s));
  });
}

// ------------------------------------------------------------------ point plan
int nusht_pointplan_create(int lmax, int spin, int64_t npts, const double* theta, nusht_stream_t s,
                           nusht_pointplan_t* out) {
  return guard([&] {
    NUSHT_REQUIRE(theta && out, kErrArg, "null argument");
    NUSHT_REQUIRE(lmax >= 0 && spin >= 0 && spin <= 3, kErrArg, "bad lmax or spin");
    NUSHT_REQUIRE(lmax >= spin, kErrArg, "lmax must be >= spin (a spin-s field has no modes l < s)");
    std::vector<double> th(theta, theta + npts);
    auto p = std::make_unique<nusht_pointplan_s>();
    p->pp = std::make_unique<PointPlan>(lmax, spin, th, S(s));
    NUSHT_HIP(hipStreamSynchronize(S(s)));
    *out = p.release();
  });
}

int nusht_pointplan_destroy(nusht_pointplan_t p) {
  delete p;
  return NUSHT_OK;
}

int nusht_pointplan_synthesis(nusht_pointplan_t p, const double* alm, int ncomp, const double* phi,
                              const double* gamma, double* out0, double* out1, nusht_stream_t s) {
  return guard([&] {
    NUSHT_REQUIRE(p, kErrArg, "null point plan");
    p->pp->synthesis(reinterpret_cast<const double2*>(alm), ncomp, phi, gamma, out0, out1, S(s));
  });
}

int nusht_pointplan_synthesis_noalloc(int lmax, int spin, int64_t npts, const double* theta, const double* alm,
                                      int ncomp, const double* phi, const double* gamma, double* out0, double* out1,
                                      nusht_stream_t s) {
  return guard([&] {
    point_synthesis_noalloc(lmax, spin, npts, theta, reinterpret_cast<const double2*>(alm), ncomp, phi, gamma, out0,
                            out1, S(s));
  });
}

// ------------------------------------------------------------------ ring plan
int nusht_ringplan_create(const nusht_ring_desc* d, nusht_stream_t s, nusht_ringplan_t* out) {
  return guard([&] {
    NUSHT_REQUIRE(d && out, kErrArg, "null argument");
    NUSHT_REQUIRE(d->nring >= 1 && d->theta && d->phi0 && d->nph && d->ofs, kErrArg, "bad ring description");
    NUSHT_REQUIRE(d->spin >= 0 && d->spin <= 2, kErrArg, "spin must be 0, 1 or 2");
    NUSHT_REQUIRE(d->lmax >= d->spin, kErrArg, "lmax must be >= spin (a spin-s field has no modes l < s)");
    std::vector<double> th(d->theta, d->theta + d->nring), p0(d->phi0, d->phi0 + d->nring);
    std::vector<long> nph(d->nph, d->nph + d->nring), ofs(d->ofs, d->ofs + d->nring);
    std::vector<double> wt;
    if (d->weight) wt.assign(d->weight, d->weight + d->nring);
    auto rp = std::make_unique<nusht_ringplan_s>();
    rp->rp = std::make_unique<RingPlan>(d->lmax, d->spin, th, p0, nph, ofs, wt, S(s));
    NUSHT_HIP(hipStreamSynchronize(S(s)));
    *out = rp.release();
  });
}

int nusht_ringplan_destroy(nusht_ringplan_t p) {
  return guard([&] { delete p; });
}

int64_t nusht_ringplan_npix(nusht_ringplan_t p) { return p ? p->rp->npix() : -1; }

int nusht_legendre_variants(void) { return legendre_variant_count(); }

int nusht_ringplan_get_info(nusht_ringplan_t p, nusht_ringplan_info* info) {
  return guard([&] {
    NUSHT_REQUIRE(p, kErrArg, "null ring plan");
    NUSHT_REQUIRE(info, kErrArg, "null info");
    const LegendreTables& t = p->rp->tables();
    info->npix = p->rp->npix();
    info->nring = p->rp->nring();
    info->ring_pairs = t.nslot;
    info->ring_dft = (int)p->rp->method();
    info->ring_lengths = p->rp->nlengths();
    info->workspace_bytes = (int64_t)p->rp->bytes();
    info->legendre_pair_steps = t.pair_steps;
    info->legendre_variant = t.variant;
    info->legendre_ms = t.tuned_ms;
  });
}

static void check_ring_call(nusht_ringplan_t p, const void* map0, const void* map1, const void* alm, int ncomp) {
  NUSHT_REQUIRE(p, kErrArg, "null ring plan");
  NUSHT_REQUIRE(map0 && alm, kErrArg, "null map or alm pointer");
  NUSHT_REQUIRE(p->rp->spin() == 0 || map1, kErrArg, "spin > 0 needs two maps (Q, U)");
  NUSHT_REQUIRE(ncomp == 1 || (p->rp->spin() > 0 && ncomp == 2), kErrArg, "ncomp must be 1, or 2 for spin > 0");
}

int nusht_ringplan_synthesis(nusht_ringplan_t p, const double* alm, int ncomp, double* map0, double* map1,
                             nusht_stream_t s) {
  return guard([&] {
    check_ring_call(p, map0, map1, alm, ncomp);
    p->rp->synthesis(reinterpret_cast<const double2*>(alm), ncomp, map0, map1, S(s));
  });
}

int nusht_ringplan_analysis(nusht_ringplan_t p, const double* map0, const double* map1, double* alm, int ncomp,
                            nusht_stream_t s) {
  return guard([&] {
    check_ring_call(p, map0, map1, alm, ncomp);
    p->rp->analysis(map0, map1, reinterpret_cast<double2*>(alm), ncomp, S(s));
  });
}

int nusht_ringplan_adjoint(nusht_ringplan_t p, const double* map0, const double* map1, double* alm, int ncomp,
                           nusht_stream_t s) {
  return guard([&] {
    check_ring_call(p, map0, map1, alm, ncomp);
    p->rp->adjoint(map0, map1, reinterpret_cast<double2*>(alm), ncomp, S(s));
  });
}

int nusht_sht_synth_grad_gl(nusht_ringplan_t p, const double* alm, double* dtheta, double* dphi, nusht_stream_t s) {
  return guard([&] {
    NUSHT_REQUIRE(p, kErrArg, "null ring plan");
    NUSHT_REQUIRE(alm && dtheta && dphi, kErrArg, "null array");
    p->rp->synthesis_grad(reinterpret_cast<const double2*>(alm), dtheta, dphi, S(s));
  });
}

int nusht_ringplan_dlm2pointing(nusht_ringplan_t p, const double* dlm_scaled, double* ptg_theta, double* ptg_phi,
                                double* gamma, nusht_stream_t s) {
  return guard([&] {
    NUSHT_REQUIRE(p, kErrArg, "null ring plan");
    p->rp->dlm2pointing(reinterpret_cast<const double2*>(dlm_scaled), ptg_theta, ptg_phi, gamma, S(s));
  });
}

// ------------------------------------------------------------------ nuSHT plan
int nusht_plan_create(const nusht_plan_desc* d, nusht_stream_t s, nusht_plan_t* out) {
  return guard([&] {
    NUSHT_REQUIRE(d && out, kErrArg, "null argument");
    NUSHT_REQUIRE(d->mmax == d->lmax, kErrArg, "mmax must equal lmax");
    NUSHT_REQUIRE(d->spin >= 0 && d->spin <= 2, kErrArg, "spin must be 0, 1 or 2");
    NUSHT_REQUIRE(d->lmax >= d->spin, kErrArg, "lmax must be >= spin (a spin-s field has no modes l < s)");
    NUSHT_REQUIRE(d->ncomp == 1 || (d->ncomp == 2 && d->spin > 0), kErrArg, "ncomp must be 1, or 2 for spin > 0");
    auto p = std::make_unique<nusht_plan_s>();
    p->lmax = d->lmax;
    p->spin = d->spin;
    p->ncomp = d->ncomp;
    p->npts_max = std::max<long>(d->npts_max, 1);
    const char* at = getenv("NUSHT_NUFFT_AUTOTUNE");
    const bool tune = d->sigma < 0 || (d->sigma == 0 && at && at[0] == '1');
    p->np = NufftParams::choose(d->lmax, d->spin, d->epsilon, d->sigma > 0 ? d->sigma : 0.0);
    std::vector<double> th(p->np.ntheta);
    for (int j = 0; j < p->np.ntheta; ++j) th[j] = M_PI * j / (p->np.ntheta - 1);
    p->tab.tag = "cc";
    p->tab.build(d->lmax, d->spin, th, S(s));
    NUSHT_REQUIRE(d->max_maps >= 0 && d->max_maps <= nusht_plan_s::kMaxMaps, kErrArg, "max_maps must be in [0, 8]");
    p->max_maps = d->max_maps ? d->max_maps : nusht_plan_s::kMaxMaps;
    p->nmrows = d->spin ? 2 * d->lmax + 1 : d->lmax + 1;
    p->coef.alloc(legendre_coef_bytes(p->tab, p->max_maps));
    p->G.alloc(sizeof(double2) * (size_t)p->nmrows * p->np.ntheta * p->max_maps);
    p->ptg.alloc(sizeof(double) * 3 * (size_t)p->npts_max);
    if (tune) {  // the CC grid (ntheta) does not depend on sigma: the Legendre tables stay
      p->nu = nufft_autotune(d->lmax, d->spin, d->epsilon, p->npts_max, p->G.as<double2>(), p->np.ntheta, S(s),
                             p->tune);
      p->np = p->nu->params();
    } else {
      p->nu = std::make_unique<NufftPlan>(p->np, p->npts_max, S(s));
    }
    legendre_autotune(p->tab, p->ncomp, p->coef.as<double2>(), p->G.as<double2>(), p->np.ntheta, 1, S(s));
    NUSHT_HIP(hipStreamSynchronize(S(s)));
    *out = p.release();
  });
}

int nusht_plan_destroy(nusht_plan_t p) {
  return guard([&] { delete p; });
}

int nusht_plan_get_info(nusht_plan_t p, nusht_plan_info* info) {
  return guard([&] {
    check_plan(p);
    NUSHT_REQUIRE(info, kErrArg, "null info");
    info->ntheta = p->np.ntheta;
    info->ntheta_dfs = p->np.ntd;
    info->nfine = p->np.nf;
    info->nrow_fine = p->np.nrow;
    info->kernel_width = p->np.w;
    info->single_precision = p->np.single ? 1 : 0;
    info->sigma = p->np.sigma;
    info->beta = p->np.beta;
    info->workspace_bytes = (int64_t)p->bytes();
    info->legendre_pair_steps = p->tab.pair_steps;
    info->legendre_variant = p->tab.variant;
    info->legendre_ms = p->tab.tuned_ms;
    info->nufft_tile_rows = p->nu->tile_rows();
    info->nufft_tile_cols = p->nu->tile_cols();
    info->nufft_candidates = p->tune.candidates;
    info->nufft_type2_ms = p->tune.type2_ms;
    info->nufft_type1_ms = p->tune.type1_ms;
    info->nufft_spread_rows = p->nu->rows_spread() ? 1 : 0;
  });
}

int64_t nusht_plan_workspace_bytes(nusht_plan_t p) { return p ? (int64_t)p->bytes() : -1; }

int nusht_plan_set_pts(nusht_plan_t p, const double* theta, const double* phi, int64_t npts, nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    p->nu->set_pts(theta, phi, npts, S(s));
  });
}

int nusht_plan_set_pts_async(nusht_plan_t p, const double* theta, const double* phi, int64_t npts,
                             nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    p->nu->set_pts(theta, phi, npts, S(s), /*check=*/false);
  });
}

}  // extern "C"

static void synth_batch(nusht_plan_t p, int nmaps, const double* alm, long alm_stride, double* out0, double* out1,
                        long out_stride, const double* gamma, hipStream_t st) {
  const long gmap = (long)p->nmrows * p->np.ntheta;
  legendre_synth(p->tab, nmaps, p->ncomp, reinterpret_cast<const double2*>(alm), alm_stride / 2, nullptr,
                 p->coef.as<double2>(), p->G.as<double2>(), p->np.ntheta, 1, gmap, st);
  for (int b = 0; b < nmaps; ++b)
    p->nu->type2(p->G.as<double2>() + b * gmap, p->np.ntheta, gamma, out0 + b * out_stride,
                 out1 ? out1 + b * out_stride : nullptr, st);
}

// ------------------------------------------------------------------ composed operations (no guard; shared by
// the single-call entry points and the batch executor)
namespace {

void op_nusht2d2(nusht_plan_t p, const double* alm, double* out0, double* out1, const double* gamma, hipStream_t st) {
  synth_batch(p, 1, alm, 0, out0, out1, 0, gamma, st);
}

void op_nusht2d1(nusht_plan_t p, const double* in0, const double* in1, const double* gamma, double* alm,
                 hipStream_t st) {
  p->nu->type1(in0, in1, gamma, p->G.as<double2>(), p->np.ntheta, st);
  legendre_adjoint(p->tab, p->ncomp, p->G.as<double2>(), p->np.ntheta, 1, reinterpret_cast<double2*>(alm), st);
}

void check_lensing(nusht_plan_t p, nusht_ringplan_t ptg) {
  NUSHT_REQUIRE(ptg && ptg->rp->spin() == 1, kErrArg, "need a spin-1 ring plan for the pointing");
  NUSHT_REQUIRE(ptg->rp->lmax() == p->lmax, kErrArg, "ring plan lmax differs from the nuSHT plan");
  NUSHT_REQUIRE(ptg->rp->npix() <= p->npts_max, kErrArg, "pointing grid larger than the plan's npts_max");
}

// pointing of the ring plan's grid into the plan's point buffers; returns gamma (or null)
const double* lensing_points(nusht_plan_t p, nusht_ringplan_t ptg, const double* dlm_scaled, int polrot,
                             hipStream_t st) {
  double* th = p->ptg.as<double>();
  double* ph = th + p->npts_max;
  double* ga = ph + p->npts_max;
  const bool rot = polrot && p->spin > 0;
  ptg->rp->dlm2pointing(reinterpret_cast<const double2*>(dlm_scaled), th, ph, rot ? ga : nullptr, st);
  p->nu->set_pts(th, ph, ptg->rp->npix(), st, /*check=*/false);  // pointing from the ring plan
  return rot ? ga : nullptr;
}

void op_gclm2lenmap(nusht_plan_t p, nusht_ringplan_t ptg, const double* gclm, const double* dlm_scaled,
                    double* lenmap0, double* lenmap1, int polrot, hipStream_t st) {
  const double* ga = lensing_points(p, ptg, dlm_scaled, polrot, st);
  synth_batch(p, 1, gclm, 0, lenmap0, lenmap1, 0, ga, st);
}

// nmaps (>= 1) lensed maps sharing one deflection: the pointing and binning once, then the synthesis of all maps
// (the matrix-core launch for 3 or more) and their interpolations, in order on `st`.  Running the Legendre
// synthesis on a second stream beside the pointing gained about 1 % of a batched step (both kernels fill the GPU
// and neither leaves room for the other on a CU) and made the Legendre kernels' timings include the time they
// shared the GPU; a fork from a stream that had joined a graph capture through an event also made
// hipStreamEndCapture crash.
void op_gclm2lenmap_n(nusht_plan_t p, nusht_ringplan_t ptg, int nmaps, const double* gclm, long alm_stride,
                      const double* dlm_scaled, double* out0, double* out1, long out_stride, int polrot,
                      hipStream_t st) {
  const double* ga = lensing_points(p, ptg, dlm_scaled, polrot, st);
  synth_batch(p, nmaps, gclm, alm_stride, out0, out1, out_stride, ga, st);
}

void op_lenmap2gclm(nusht_plan_t p, nusht_ringplan_t ptg, const double* lenmap0, const double* lenmap1,
                    const double* dlm_scaled, double* gclm, int polrot, hipStream_t st) {
  const double* ga = lensing_points(p, ptg, dlm_scaled, polrot, st);
  op_nusht2d1(p, lenmap0, lenmap1, ga, gclm, st);
}

}  // namespace

extern "C" {

int nusht_nusht2d2(nusht_plan_t p, const double* alm, double* out0, double* out1, const double* gamma,
                   nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    NUSHT_REQUIRE(alm && out0, kErrArg, "null array");
    NUSHT_REQUIRE(p->spin == 0 || out1, kErrArg, "spin > 0 needs two output arrays");
    op_nusht2d2(p, alm, out0, out1, gamma, S(s));
  });
}

int nusht_nusht2d2_batch(nusht_plan_t p, int nmaps, const double* alm, int64_t alm_stride, double* out0,
                         double* out1, int64_t out_stride, const double* gamma, nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    NUSHT_REQUIRE(nmaps >= 1 && nmaps <= p->max_maps, kErrArg,
                  "nmaps must be in [1, " + std::to_string(p->max_maps) + "] (the plan's max_maps)");
    NUSHT_REQUIRE(alm && out0, kErrArg, "null array");
    NUSHT_REQUIRE(p->spin == 0 || out1, kErrArg, "spin > 0 needs two output arrays");
    synth_batch(p, nmaps, alm, alm_stride, out0, out1, out_stride, gamma, S(s));
  });
}

int nusht_nusht2d1(nusht_plan_t p, const double* in0, const double* in1, const double* gamma, double* alm,
                   nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    NUSHT_REQUIRE(in0 && alm, kErrArg, "null array");
    NUSHT_REQUIRE(p->spin == 0 || in1, kErrArg, "spin > 0 needs two input arrays");
    op_nusht2d1(p, in0, in1, gamma, alm, S(s));
  });
}

int nusht_gclm2lenmap(nusht_plan_t p, nusht_ringplan_t ptg, const double* gclm, const double* dlm_scaled,
                      double* lenmap0, double* lenmap1, int polrot, nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    check_lensing(p, ptg);
    NUSHT_REQUIRE(gclm && dlm_scaled && lenmap0, kErrArg, "null array");
    NUSHT_REQUIRE(p->spin == 0 || lenmap1, kErrArg, "spin > 0 needs two output arrays");
    op_gclm2lenmap(p, ptg, gclm, dlm_scaled, lenmap0, lenmap1, polrot, S(s));
  });
}

int nusht_gclm2lenmap_batch(nusht_plan_t p, nusht_ringplan_t ptg, int nmaps, const double* gclm,
                            int64_t alm_stride, const double* dlm_scaled, double* lenmap0, double* lenmap1,
                            int64_t out_stride, int polrot, nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    check_lensing(p, ptg);
    NUSHT_REQUIRE(nmaps >= 1 && nmaps <= p->max_maps, kErrArg,
                  "nmaps must be in [1, " + std::to_string(p->max_maps) + "] (the plan's max_maps)");
    NUSHT_REQUIRE(gclm && dlm_scaled && lenmap0, kErrArg, "null array");
    NUSHT_REQUIRE(p->spin == 0 || lenmap1, kErrArg, "spin > 0 needs two output arrays");
    op_gclm2lenmap_n(p, ptg, nmaps, gclm, alm_stride, dlm_scaled, lenmap0, lenmap1, out_stride, polrot, S(s));
  });
}

int nusht_lenmap2gclm(nusht_plan_t p, nusht_ringplan_t ptg, const double* lenmap0, const double* lenmap1,
                      const double* dlm_scaled, double* gclm, int polrot, nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    check_lensing(p, ptg);
    NUSHT_REQUIRE(lenmap0 && dlm_scaled && gclm, kErrArg, "null array");
    NUSHT_REQUIRE(p->spin == 0 || lenmap1, kErrArg, "spin > 0 needs two input arrays");
    op_lenmap2gclm(p, ptg, lenmap0, lenmap1, dlm_scaled, gclm, polrot, S(s));
  });
}

// ------------------------------------------------------------------ 2D FFT
int nusht_fft2d_c2c_f64(double* data, int n0, int n1, int batch, int direction, nusht_stream_t s) {
  return guard([&] {
    NUSHT_REQUIRE(direction == -1 || direction == 1, kErrArg, "direction must be -1 or +1");
    fft2d_c2c(reinterpret_cast<double2*>(data), n0, n1, batch, direction > 0, S(s));
  });
}

int nusht_fft2d_c2c_f32(float* data, int n0, int n1, int batch, int direction, nusht_stream_t s) {
  return guard([&] {
    NUSHT_REQUIRE(direction == -1 || direction == 1, kErrArg, "direction must be -1 or +1");
    fft2d_c2c(reinterpret_cast<float2*>(data), n0, n1, batch, direction > 0, S(s));
  });
}

// ------------------------------------------------------------------ plan stages
int nusht_sht_synth_cc(nusht_plan_t p, const double* alm, double* G, nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    NUSHT_REQUIRE(alm && G, kErrArg, "null array");
    legendre_synth(p->tab, 1, p->ncomp, reinterpret_cast<const double2*>(alm), 0, nullptr, p->coef.as<double2>(),
                   reinterpret_cast<double2*>(G), p->np.ntheta, 1, 0, S(s));
  });
}

int nusht_sht_adjoint_cc(nusht_plan_t p, const double* G, double* alm, nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    NUSHT_REQUIRE(alm && G, kErrArg, "null array");
    legendre_adjoint(p->tab, p->ncomp, reinterpret_cast<const double2*>(G), p->np.ntheta, 1,
                     reinterpret_cast<double2*>(alm), S(s));
  });
}

int nusht_nufft_setpts(nusht_plan_t p, const double* theta, const double* phi, int64_t npts, nusht_stream_t s) {
  return nusht_plan_set_pts(p, theta, phi, npts, s);
}

int nusht_nufft_interp(nusht_plan_t p, const double* G, const double* gamma, double* out0, double* out1,
                       nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    NUSHT_REQUIRE(G && out0, kErrArg, "null array");
    NUSHT_REQUIRE(p->spin == 0 || out1, kErrArg, "spin > 0 needs two output arrays");
    p->nu->type2(reinterpret_cast<const double2*>(G), p->np.ntheta, gamma, out0, out1, S(s));
  });
}

int nusht_nufft_spread(nusht_plan_t p, const double* in0, const double* in1, const double* gamma, double* G,
                       nusht_stream_t s) {
  return guard([&] {
    check_plan(p);
    NUSHT_REQUIRE(in0 && G, kErrArg, "null array");
    NUSHT_REQUIRE(p->spin == 0 || in1, kErrArg, "spin > 0 needs two input arrays");
    p->nu->type1(in0, in1, gamma, reinterpret_cast<double2*>(G), p->np.ntheta, S(s));
  });
}

}  // extern "C"

// ------------------------------------------------------------------ batches of independent transforms
namespace {

// Internal streams of one device for the batch executor, created on first use and kept for the life
// of the process (default priority, non-blocking with respect to the legacy stream).
struct BatchStreams {
  static constexpr int K = 4;
  hipStream_t s[K] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[K] = {};
  bool ready = false;
  std::mutex mu;
};

BatchStreams& batch_streams() {
  static BatchStreams pools[64];
  int dev = 0;
  NUSHT_HIP(hipGetDevice(&dev));
  NUSHT_REQUIRE(dev >= 0 && dev < 64, kErrUnsupported, "batch: device index out of range");
  return pools[dev];
}

void check_job(const nusht_job& j, int i) {
  const std::string at = "batch job " + std::to_string(i) + ": ";
  NUSHT_REQUIRE(j.plan, kErrArg, at + "null plan");
  nusht_plan_t p = j.plan;
  switch (j.kind) {
    case NUSHT_JOB_GCLM2LENMAP:
    case NUSHT_JOB_LENMAP2GCLM:
      check_lensing(p, j.ring);
      NUSHT_REQUIRE(j.dlm_scaled, kErrArg, at + "null dlm_scaled");
      break;
    case NUSHT_JOB_NUSHT2D2:
    case NUSHT_JOB_NUSHT2D1:
      NUSHT_REQUIRE(!j.ring, kErrArg, at + "nusht2d2 / nusht2d1 jobs take no ring plan");
      break;
    default:
      NUSHT_REQUIRE(false, kErrArg, at + "unknown job kind " + std::to_string(j.kind));
  }
  const bool fwd = j.kind == NUSHT_JOB_GCLM2LENMAP || j.kind == NUSHT_JOB_NUSHT2D2;
  if (fwd) {
    NUSHT_REQUIRE(j.alm_in && j.map_out0, kErrArg, at + "null alm_in or map_out0");
    NUSHT_REQUIRE(p->spin == 0 || j.map_out1, kErrArg, at + "spin > 0 needs map_out1");
  } else {
    NUSHT_REQUIRE(j.map_in0 && j.alm_out, kErrArg, at + "null map_in0 or alm_out");
    NUSHT_REQUIRE(p->spin == 0 || j.map_in1, kErrArg, at + "spin > 0 needs map_in1");
  }
}

void run_job(const nusht_job& j, hipStream_t st) {
  switch (j.kind) {
    case NUSHT_JOB_GCLM2LENMAP:
      op_gclm2lenmap(j.plan, j.ring, j.alm_in, j.dlm_scaled, j.map_out0, j.map_out1, j.polrot, st);
      break;
    case NUSHT_JOB_LENMAP2GCLM:
      op_lenmap2gclm(j.plan, j.ring, j.map_in0, j.map_in1, j.dlm_scaled, j.alm_out, j.polrot, st);
      break;
    case NUSHT_JOB_NUSHT2D2:
      op_nusht2d2(j.plan, j.alm_in, j.map_out0, j.map_out1, j.gamma, st);
      break;
    default:
      op_nusht2d1(j.plan, j.map_in0, j.map_in1, j.gamma, j.alm_out, st);
      break;
  }
}

int find_root(std::vector<int>& par, int a) {
  while (par[a] != a) a = par[a] = par[par[a]];
  return a;
}

// Forward jobs `idx` of one (lmax, spin, ncomp), in order on `st`: their Legendre syntheses share launches (the
// matrix-core kernel takes up to four maps, whatever their deflections) into the G buffer of the first job's plan,
// max_maps maps at a time; each job then computes its own pointing and interpolates from its slice of that G.
void run_fused_forward(const nusht_job* jobs, const std::vector<int>& idx, hipStream_t st) {
  const nusht_plan_t lead = jobs[idx[0]].plan;
  const long gmap = (long)lead->nmrows * lead->np.ntheta;
  for (size_t c0 = 0; c0 < idx.size(); c0 += lead->max_maps) {
    const int nb = (int)std::min<size_t>(lead->max_maps, idx.size() - c0);
    std::vector<const double2*> alm(nb);
    for (int b = 0; b < nb; ++b) alm[b] = reinterpret_cast<const double2*>(jobs[idx[c0 + b]].alm_in);
    legendre_synth_maps(lead->tab, nb, lead->ncomp, alm.data(), lead->coef.as<double2>(), lead->G.as<double2>(),
                        lead->np.ntheta, 1, gmap, st);
    for (int b = 0; b < nb; ++b) {
      const nusht_job& j = jobs[idx[c0 + b]];
      const double* ga = lensing_points(j.plan, j.ring, j.dlm_scaled, j.polrot, st);
      j.plan->nu->type2(lead->G.as<double2>() + b * gmap, lead->np.ntheta, ga, j.map_out0, j.map_out1, st);
    }
  }
}

}  // namespace

extern "C" {

int nusht_batch_execute(const nusht_job* jobs, int njobs, nusht_stream_t s) {
  return guard([&] {
    NUSHT_REQUIRE(njobs >= 0 && (jobs || njobs == 0), kErrArg, "batch: bad job array");
    for (int i = 0; i < njobs; ++i) check_job(jobs[i], i);
    if (njobs == 0) return;
    // jobs that share a plan or a ring plan form one ordered group (union-find over the handles)
    std::vector<int> par(njobs);
    for (int i = 0; i < njobs; ++i) par[i] = i;
    // forward jobs of one (lmax, spin, ncomp), three or more, whose first plan holds G for four or more maps:
    // one fused, ordered group (shared Legendre launches)
    std::map<std::tuple<int, int, int>, std::vector<int>> fwd;
    for (int i = 0; i < njobs; ++i)
      if (jobs[i].kind == NUSHT_JOB_GCLM2LENMAP)
        fwd[std::make_tuple(jobs[i].plan->lmax, jobs[i].plan->spin, jobs[i].plan->ncomp)].push_back(i);
    std::vector<int> fused_head(njobs, -1);  // fused group: index of its first job; -1: runs alone
    std::vector<std::vector<int>> fused;
    for (auto& kv : fwd) {
      const std::vector<int>& v = kv.second;
      if (v.size() < 3 || jobs[v[0]].plan->max_maps < kLegendreMfmaMaps || !jobs[v[0]].plan->tab.mfma) continue;
      for (int i : v) {
        fused_head[i] = v[0];
        par[find_root(par, i)] = find_root(par, v[0]);
      }
      fused.push_back(v);
    }
    std::unordered_map<const void*, int> owner;
    for (int i = 0; i < njobs; ++i)
      for (const void* h : {(const void*)jobs[i].plan, (const void*)jobs[i].ring}) {
        if (!h) continue;
        auto it = owner.find(h);
        if (it == owner.end()) owner.emplace(h, i);
        else par[find_root(par, i)] = find_root(par, it->second);
      }
    std::unordered_map<int, int> group_of_root;
    std::vector<int> group(njobs);
    std::vector<double> cost;
    for (int i = 0; i < njobs; ++i) {
      const int r = find_root(par, i);
      auto it = group_of_root.find(r);
      if (it == group_of_root.end()) it = group_of_root.emplace(r, (int)cost.size()).first, cost.push_back(0.0);
      group[i] = it->second;
      const nusht_plan_t p = jobs[i].plan;
      cost[group[i]] += (double)(p->lmax + 1) * (p->lmax + 1) * (p->spin ? 2.0 : 1.0) * p->np.ntheta;
    }
    const int ngroups = (int)cost.size();
    BatchStreams& bs = batch_streams();
    std::lock_guard<std::mutex> lock(bs.mu);
    if (!bs.ready) {
      for (int k = 0; k < BatchStreams::K; ++k) {
        NUSHT_HIP(hipStreamCreateWithFlags(&bs.s[k], hipStreamNonBlocking));
        NUSHT_HIP(hipEventCreateWithFlags(&bs.join[k], hipEventDisableTiming));
      }
      NUSHT_HIP(hipEventCreateWithFlags(&bs.fork, hipEventDisableTiming));
      bs.ready = true;
    }
    // longest group first onto the least loaded stream
    const int K = std::min(BatchStreams::K, ngroups);
    std::vector<int> order(ngroups), stream_of(ngroups);
    for (int g = 0; g < ngroups; ++g) order[g] = g;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<double> load(K, 0.0);
    for (int g : order) {
      const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      stream_of[g] = k;
      load[k] += cost[g];
    }
    NUSHT_HIP(hipEventRecord(bs.fork, S(s)));
    for (int k = 0; k < K; ++k) NUSHT_HIP(hipStreamWaitEvent(bs.s[k], bs.fork, 0));
    for (int i = 0; i < njobs; ++i) {
      hipStream_t si = bs.s[stream_of[group[i]]];
      if (fused_head[i] < 0) {
        run_job(jobs[i], si);
      } else if (fused_head[i] == i) {
        for (const std::vector<int>& v : fused)
          if (v[0] == i) run_fused_forward(jobs, v, si);
      }
    }
    for (int k = 0; k < K; ++k) {
      NUSHT_HIP(hipEventRecord(bs.join[k], bs.s[k]));
      NUSHT_HIP(hipStreamWaitEvent(S(s), bs.join[k], 0));
    }
  });
}

// ------------------------------------------------------------------ stage timing
int nusht_timing_enable(int on) {
  return guard([&] { StageTimer::get().enable(on != 0); });
}
int nusht_timing_reset(void) {
  return guard([&] { StageTimer::get().reset(); });
}
int nusht_timing_count(void) {
  int n = -1;
  int rc = guard([&] { n = StageTimer::get().collect(); });
  return rc == NUSHT_OK ? n : -rc;
}
int nusht_timing_get(int i, const char** name, double* total_ms, int64_t* calls) {
  return guard([&] {
    auto& t = StageTimer::get();
    NUSHT_REQUIRE(i >= 0 && i < t.count(), kErrArg, "timing index out of range");
    if (name) *name = t.name(i);
    if (total_ms) *total_ms = t.total_ms(i);
    if (calls) *calls = t.calls(i);
  });
}

}  // extern "C"
