// C ABI of the nedyson_amd library (declarations: include/nedyson.h).
#include "../../include/nedyson.h"
#include "ctx.h"
#include <memory>
#include <thread>
#include <vector>
#include "dyson.h"
#include "predictor.h"
#include "quadrature.h"
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <new>
#include <string>
#include <vector>

struct ned_ctx {
  ned::Ctx c;
};

namespace {
thread_local std::string g_err;

template <class F>
ned_status guard(ned_ctx *ctx, F &&f) {
  try {
    if (ctx) NED_HIP(hipSetDevice(ctx->c.device));
    f();
    return NED_OK;
  } catch (const std::invalid_argument &e) {
    g_err = e.what();
    return NED_ERR_ARG;
  } catch (const std::bad_alloc &e) {
    g_err = e.what();
    return NED_ERR_NOMEM;
  } catch (const std::runtime_error &e) {
    g_err = e.what();
    if (g_err.find("out of memory") != std::string::npos) return NED_ERR_NOMEM;
    if (g_err.find("non-finite") != std::string::npos) return NED_ERR_NONFINITE;
    if (g_err.find("not converge") != std::string::npos) return NED_ERR_NOCONV;
    if (g_err.find("state:") != std::string::npos) return NED_ERR_STATE;
    return NED_ERR_DEVICE;
  } catch (...) {
    g_err = "unknown error";
    return NED_ERR_DEVICE;
  }
}

ned::DevBuf &component(ned::Ctx &c, int which, int comp) {
  ned::Green &g = which == NED_SIGMA ? c.S : c.G;
  switch (comp) {
    case NED_RET: return g.ret;
    case NED_LES: return g.les;
    case NED_TV: return g.tv;
    case NED_MAT: return g.mat;
  }
  throw std::invalid_argument("component must be NED_RET, NED_LES, NED_TV or NED_MAT");
}

long long nblocks_of(const ned::Ctx &c, int comp) {
  if (comp == NED_RET || comp == NED_LES) return (long long)(c.nt + 1) * (c.nt + 2) / 2;
  if (comp == NED_TV) return (long long)(c.nt + 1) * c.np;
  return c.np;
}

// host <-> device copy of the packed blocks [first, first + nb) of one
// component (complex FP64 on the host): Sigma^R / Sigma^< of a context with
// FP32 Sigma storage are rounded on the way in and widened on the way out;
// G^R / G^< of a context with ring storage are copied row by row to and from
// their ring slots (a read needs the rows held; a write makes its last row
// the newest, and every row it writes must stay held)
void triangle_copy(ned::Ctx &c, int which, int comp, ned::DevBuf &b, long long first, long long nb, void *host,
                   bool to_device) {
  const size_t es = c.es, bs = es * sizeof(ned::z2);
  const bool tri = comp == NED_RET || comp == NED_LES;
  if (nb == 0) return;
  if (tri && which == NED_SIGMA && c.s32) {
    std::vector<float> f((size_t)nb * es * 2);
    double *h = static_cast<double *>(host);
    char *d = (char *)b.p + (size_t)first * es * ned::s_elem(c);
    if (to_device) {
      for (size_t i = 0; i < f.size(); i++) f[i] = (float)h[i];
      NED_HIP(hipMemcpyAsync(d, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, c.stream));
      NED_HIP(hipStreamSynchronize(c.stream));
    } else {
      NED_HIP(hipMemcpyAsync(f.data(), d, f.size() * sizeof(float), hipMemcpyDeviceToHost, c.stream));
      NED_HIP(hipStreamSynchronize(c.stream));
      for (size_t i = 0; i < f.size(); i++) h[i] = (double)f[i];
    }
    return;
  }
  if (tri && which == NED_G && c.gring) {
    auto row_of = [](long long blk) {
      long long t = (long long)((std::sqrt(8.0 * (double)blk + 1.0) - 1.0) / 2.0);
      while ((t + 1) * (t + 2) / 2 <= blk) t++;
      while (t * (t + 1) / 2 > blk) t--;
      return (int)t;
    };
    const int t0 = row_of(first), t1 = row_of(first + nb - 1);
    if (to_device) {
      ned::g_wrote(c, t1);
      ned::g_need(c, t0, t1, "green_write");
    } else {
      ned::g_need(c, t0, t1, "green_read");
    }
    for (long long blk = first; blk < first + nb;) {
      const int t = row_of(blk);
      const long long s0 = blk - (long long)t * (t + 1) / 2;
      const long long cnt = std::min<long long>(first + nb - blk, t + 1 - s0);
      char *d = (char *)b.p + (ned::g_row(c, t) + (size_t)s0) * bs;
      char *h = (char *)host + (size_t)(blk - first) * bs;
      if (to_device) NED_HIP(hipMemcpyAsync(d, h, cnt * bs, hipMemcpyHostToDevice, c.stream));
      else NED_HIP(hipMemcpyAsync(h, d, cnt * bs, hipMemcpyDeviceToHost, c.stream));
      blk += cnt;
    }
    NED_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  char *d = (char *)b.p + (size_t)first * bs;
  if (to_device) NED_HIP(hipMemcpyAsync(d, host, nb * bs, hipMemcpyHostToDevice, c.stream));
  else NED_HIP(hipMemcpyAsync(host, d, nb * bs, hipMemcpyDeviceToHost, c.stream));
  NED_HIP(hipStreamSynchronize(c.stream));
}

void state(bool ok, const char *msg) {
  if (!ok) throw std::runtime_error(std::string("state: ") + msg);
}

void release_streams(ned::Ctx &c) {
  for (hipEvent_t &e : c.ev_side)
    if (e) (void)hipEventDestroy(e);
  if (c.ev_main) (void)hipEventDestroy(c.ev_main);
  if (c.ev_fork) (void)hipEventDestroy(c.ev_fork);
  if (c.ev_tv) (void)hipEventDestroy(c.ev_tv);
  if (c.side) (void)hipStreamDestroy(c.side);
  if (c.side2) (void)hipStreamDestroy(c.side2);
  if (c.stream) (void)hipStreamDestroy(c.stream);
}
}  // namespace

struct ned_mlp {
  std::unique_ptr<ned::Mlp> m;
};

// independent systems stepped together: grouped (one launch per link of the
// Dyson recurrences for all members, ned::DysonGroup) when the members share a
// device, else one host thread and stream per member
struct ned_batch {
  std::vector<ned_ctx *> ctx;
  bool grouped = true;
  std::unique_ptr<ned::DysonGroup> group;
};

template <class F>
ned_status mlp_guard(ned_mlp *m, F &&f) {
  if (!m) { g_err = "null network"; return NED_ERR_ARG; }
  return guard(nullptr, [&] {
    NED_HIP(hipSetDevice(m->m->device));
    f(*m->m);
  });
}

extern "C" {

const char *ned_last_error(void) { return g_err.c_str(); }

ned_status ned_ctx_create(ned_ctx **out, const ned_dims *d, int device) { return ned_ctx_create_ex(out, d, device, 0); }

ned_status ned_ctx_create_ex(ned_ctx **out, const ned_dims *d, int device, unsigned flags) {
  if (!out || !d) { g_err = "null argument"; return NED_ERR_ARG; }
  if (flags & ~(unsigned)(NED_STORE_SIGMA_F32 | NED_STORE_G_RING)) { g_err = "unknown storage flags"; return NED_ERR_ARG; }
  *out = nullptr;
  ned_ctx *ctx = nullptr;
  ned_status st = guard(nullptr, [&] {
    ned::require(d->nao >= 1 && d->nao <= 64, "nao must be in 1..64");
    ned::require(d->naux >= 1, "naux must be >= 1");
    ned::require(d->k >= 1 && d->k <= ned::KMAX, "k must be in 1..8");
    ned::require(d->nt >= d->k, "nt must be >= k");
    ned::require(d->ntau >= d->k, "ntau must be >= k (src/params.h:56-102)");
    ned::require(d->sig == -1, "only fermions (sig = -1) are supported");
    int ndev = 0;
    NED_HIP(hipGetDeviceCount(&ndev));
    ned::require(device >= 0 && device < ndev, "no such HIP device");
    NED_HIP(hipSetDevice(device));
    // address of this device's copy of the no-allocation Sigma scratch
    ned::sigma_scratch_init(device);
    ctx = new ned_ctx();
    ned::Ctx &c = ctx->c;
    c.nao = d->nao; c.naux = d->naux; c.nt = d->nt; c.ntau = d->ntau; c.k = d->k; c.sig = d->sig;
    c.device = device;
    c.es = c.nao * c.nao;
    c.np = c.ntau + 1;
    NED_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    NED_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
    NED_HIP(hipStreamCreateWithFlags(&c.side2, hipStreamNonBlocking));
    NED_HIP(hipEventCreateWithFlags(&c.ev_main, hipEventDisableTiming));
    NED_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
    NED_HIP(hipEventCreateWithFlags(&c.ev_tv, hipEventDisableTiming));
    for (hipEvent_t &e : c.ev_side) NED_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c.s32 = (flags & NED_STORE_SIGMA_F32) != 0;
    // the ring holds rows T - k - 1 .. T of a timestep, and one more
    c.gring = (flags & NED_STORE_G_RING) ? std::min(c.nt + 1, c.k + 3) : 0;
    const size_t es = c.es, ntri = (size_t)(c.nt + 1) * (c.nt + 2) / 2;
    const size_t gtri = c.gring ? (size_t)c.gring * (c.nt + 1) : ntri;
    for (ned::Green *g : {&c.G, &c.S}) {
      const size_t nblk = g == &c.G ? gtri : ntri, el = g == &c.S ? ned::s_elem(c) : sizeof(ned::z2);
      g->ret.ensure(nblk * es * el);
      g->les.ensure(nblk * es * el);
      g->tv.ensure((size_t)(c.nt + 1) * c.np * es * sizeof(ned::z2));
      g->mat.ensure((size_t)c.np * es * sizeof(ned::z2));
      NED_HIP(hipMemsetAsync(g->ret.p, 0, g->ret.bytes, c.stream));
      NED_HIP(hipMemsetAsync(g->les.p, 0, g->les.bytes, c.stream));
      NED_HIP(hipMemsetAsync(g->tv.p, 0, g->tv.bytes, c.stream));
      NED_HIP(hipMemsetAsync(g->mat.p, 0, g->mat.bytes, c.stream));
    }
    c.hmf.ensure((size_t)(c.nt + 1) * es * sizeof(ned::z2));
    NED_HIP(hipMemsetAsync(c.hmf.p, 0, c.hmf.bytes, c.stream));
    c.Minv.ensure((size_t)(c.nt + 1) * es * sizeof(ned::z2));
    c.minv_valid.assign(c.nt + 1, 0);
    c.dy_info.ensure(sizeof(int));
    // items per Sigma chunk (tests and tuning; 0 = sized to ~3 GiB of scratch)
    if (const char *ch = std::getenv("NED_SIGMA_CHUNK")) c.sigma_chunk = std::atoi(ch);
    // schedule of the Dyson recurrences (tests and tuning)
    if (const char *e = std::getenv("NED_RPANEL")) c.rpanel = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("NED_PANEL_BAND_MAX")) c.panel_band_max = std::atoi(e);
    if (const char *e = std::getenv("NED_CHAIN")) c.chain_mode = std::atoi(e);
    ned::require(c.chain_mode >= 0 && c.chain_mode <= 2, "NED_CHAIN must be 0, 1 or 2");
    if (const char *e = std::getenv("NED_GRAPH")) c.graph_mode = std::atoi(e) != 0;
    NED_HIP(hipStreamSynchronize(c.stream));
  });
  if (st != NED_OK) {
    if (ctx) {
      release_streams(ctx->c);
      delete ctx;
    }
    if (g_err.find("out of memory") != std::string::npos) st = NED_ERR_NOMEM;
    return st;
  }
  *out = ctx;
  return NED_OK;
}

ned_status ned_ctx_destroy(ned_ctx *ctx) {
  if (!ctx) return NED_OK;
  ned_status st = guard(ctx, [&] {
    NED_HIP(hipStreamSynchronize(ctx->c.side));
    NED_HIP(hipStreamSynchronize(ctx->c.side2));
    NED_HIP(hipStreamSynchronize(ctx->c.stream));
  });
  release_streams(ctx->c);
  delete ctx;
  return st;
}

ned_status ned_ctx_get_dims(const ned_ctx *ctx, ned_dims *out) {
  if (!ctx || !out) { g_err = "null argument"; return NED_ERR_ARG; }
  const ned::Ctx &c = ctx->c;
  out->nao = c.nao; out->naux = c.naux; out->nt = c.nt; out->ntau = c.ntau; out->k = c.k; out->sig = c.sig;
  return NED_OK;
}

ned_status ned_ctx_get_storage(const ned_ctx *ctx, unsigned *flags, int *g_first, int *g_last) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  const ned::Ctx &c = ctx->c;
  if (flags) *flags = (c.s32 ? NED_STORE_SIGMA_F32 : 0u) | (c.gring ? NED_STORE_G_RING : 0u);
  if (g_first) *g_first = c.gring ? std::max(0, c.g_top - c.gring + 1) : 0;
  if (g_last) *g_last = c.gring ? c.g_top : c.nt;
  return NED_OK;
}

ned_status ned_sync(ned_ctx *ctx) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  return guard(ctx, [&] { NED_HIP(hipStreamSynchronize(ctx->c.stream)); });
}

ned_status ned_set_coulomb(ned_ctx *ctx, const double *B) {
  if (!ctx || !B) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(ctx, [&] { ned::build_coulomb(ctx->c, B); });
}

ned_status ned_green_write(ned_ctx *ctx, int which, int comp, long long first, long long nb, const void *host) {
  if (!ctx || (!host && nb > 0)) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::Ctx &c = ctx->c;
    ned::require(which == NED_G || which == NED_SIGMA, "which must be NED_G or NED_SIGMA");
    ned::DevBuf &b = component(c, which, comp);
    ned::require(first >= 0 && nb >= 0 && first + nb <= nblocks_of(c, comp), "block range out of bounds");
    triangle_copy(c, which, comp, b, first, nb, const_cast<void *>(host), true);
    if (which == NED_SIGMA || (which == NED_G && comp == NED_MAT)) {
      ned::invalidate_minv(c);
      if (comp == NED_MAT) c.have_mix = false;
    }
  });
}

ned_status ned_green_read(ned_ctx *ctx, int which, int comp, long long first, long long nb, void *host) {
  if (!ctx || (!host && nb > 0)) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::Ctx &c = ctx->c;
    ned::require(which == NED_G || which == NED_SIGMA, "which must be NED_G or NED_SIGMA");
    ned::DevBuf &b = component(c, which, comp);
    ned::require(first >= 0 && nb >= 0 && first + nb <= nblocks_of(c, comp), "block range out of bounds");
    triangle_copy(c, which, comp, b, first, nb, host, false);
  });
}

ned_status ned_hmf_write(ned_ctx *ctx, int first, int count, const void *host) {
  if (!ctx || (!host && count > 0)) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::Ctx &c = ctx->c;
    ned::require(first >= 0 && count >= 0 && first + count <= c.nt + 1, "hmf range out of bounds");
    const size_t bs = c.es * sizeof(ned::z2);
    NED_HIP(hipMemcpyAsync((char *)c.hmf.p + first * bs, host, count * bs, hipMemcpyHostToDevice, c.stream));
    NED_HIP(hipStreamSynchronize(c.stream));
    for (int j = first; j < first + count; j++) ned::touch_time(c, j);
  });
}

ned_status ned_hmf_read(ned_ctx *ctx, int first, int count, void *host) {
  if (!ctx || (!host && count > 0)) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::Ctx &c = ctx->c;
    ned::require(first >= 0 && count >= 0 && first + count <= c.nt + 1, "hmf range out of bounds");
    const size_t bs = c.es * sizeof(ned::z2);
    NED_HIP(hipMemcpyAsync(host, (char *)c.hmf.p + first * bs, count * bs, hipMemcpyDeviceToHost, c.stream));
    NED_HIP(hipStreamSynchronize(c.stream));
  });
}

ned_status ned_sigma_gf2_decomp(ned_ctx *ctx, int tstp, int prec) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::require(prec >= 0 && prec <= 2, "prec must be NED_PREC_F64, _F32 or _BF16X3");
    state(ctx->c.have_coulomb, "interaction not set (ned_set_coulomb)");
    ned::sigma_gf2(ctx->c, tstp, (ned::Prec)prec);
    ned::touch_time(ctx->c, tstp);
  });
}

ned_status ned_sigma_gf2_decomp_noalloc(ned_ctx *ctx, int tstp, int prec) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::require(prec >= 0 && prec <= 2, "prec must be NED_PREC_F64, _F32 or _BF16X3");
    state(ctx->c.have_coulomb, "interaction not set (ned_set_coulomb)");
    ned::sigma_gf2_noalloc(ctx->c, tstp, (ned::Prec)prec);
    ned::touch_time(ctx->c, tstp);
  });
}

size_t ned_sigma_noalloc_scratch_bytes(void) { return ned::sigma_scratch_bytes(); }

ned_status ned_sigma_gf2_full(ned_ctx *ctx, int tstp, int prec) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::require(prec >= 0 && prec <= 2, "prec must be NED_PREC_F64, _F32 or _BF16X3");
    state(ctx->c.have_coulomb, "interaction not set (ned_set_coulomb)");
    ned::sigma_gf2_full(ctx->c, tstp, (ned::Prec)prec);
    ned::touch_time(ctx->c, tstp);
  });
}

ned_status ned_hf_decomp(ned_ctx *ctx, int tstp, const double *h0) {
  if (!ctx || !h0) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    state(ctx->c.have_coulomb, "interaction not set (ned_set_coulomb)");
    ned::hf_decomp(ctx->c, tstp, h0);
    ned::touch_time(ctx->c, tstp);  // hmf(tstp) enters the cached M_tstp
  });
}

ned_status ned_set_matsubara(ned_ctx *ctx, double beta) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::require(beta > 0 && std::isfinite(beta), "beta must be positive");
    ned::set_matsubara(ctx->c, beta);
  });
}

ned_status ned_matsubara_gf2(ned_ctx *ctx, const double *h1, double nelec, double beta, int mode, int fix_mu,
                             double *mu_io, int maxiter, double tol, double *fock_out, int *iters_out,
                             double *err_out) {
  if (!ctx || !h1 || !mu_io || !fock_out) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::require(beta > 0 && std::isfinite(beta), "beta must be positive");
    ned::require(mode == NED_MAT_HF || mode == NED_MAT_GF2, "mode must be NED_MAT_HF or NED_MAT_GF2");
    ned::require(fix_mu || (nelec > 0 && nelec < 2.0 * ctx->c.nao), "nelec must lie in (0, 2 nao)");
    ned::require(maxiter > 0 && tol > 0, "maxiter and tol must be positive");
    int it = 0;
    double err = 0;
    ned::matsubara_gf2(ctx->c, h1, nelec, beta, mode == NED_MAT_GF2, fix_mu != 0, mu_io, maxiter, tol, fock_out, &it,
                       &err);
    if (iters_out) *iters_out = it;
    if (err_out) *err_out = err;
    if (!(err < tol)) throw std::runtime_error("matsubara: self-consistency did not converge");
  });
}

ned_status ned_dyson_step(ned_ctx *ctx, int tstp, double mu, double beta, double dt, double *err_out) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::require(dt > 0, "dt must be positive");
    state(ctx->c.have_mix && ctx->c.mix_beta == beta, "ned_set_matsubara(beta) not called for this beta");
    double e[3];
    ned::dyson_step(ctx->c, tstp, mu, beta, dt, e);
    if (err_out) { err_out[0] = e[0]; err_out[1] = e[1]; err_out[2] = e[2]; }
    if (!std::isfinite(e[0] + e[1] + e[2])) throw std::runtime_error("non-finite update in dyson_step");
  });
}

ned_status ned_extrapolate(ned_ctx *ctx, int tstp) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  return guard(ctx, [&] { ned::extrapolate(ctx->c, tstp); });
}

ned_status ned_dyson_start(ned_ctx *ctx, double mu, double beta, double dt, double *err_out) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::require(dt > 0, "dt must be positive");
    state(ctx->c.have_mix && ctx->c.mix_beta == beta, "ned_set_matsubara(beta) not called for this beta");
    double e = ned::dyson_start(ctx->c, mu, beta, dt);
    if (err_out) *err_out = e;
    if (!std::isfinite(e)) throw std::runtime_error("non-finite update in dyson_start");
  });
}

ned_status ned_do_boot(ned_ctx *ctx, const double *fock, double mu, double beta, double dt, int prec, int maxiter,
                       double tol, int *iters_out, double *err_out) {
  if (!ctx || !fock) { g_err = "null argument"; return NED_ERR_ARG; }
  ned_status st = guard(ctx, [&] {
    ned::Ctx &c = ctx->c;
    ned::require(prec >= 0 && prec <= 2, "bad prec");
    state(c.have_coulomb, "interaction not set (ned_set_coulomb)");
    state(c.have_mix && c.mix_beta == beta, "ned_set_matsubara(beta) not called for this beta");
    std::vector<ned::z2> h((size_t)(c.k + 1) * c.es);
    for (int t = 0; t <= c.k; t++)
      for (int e = 0; e < c.es; e++) h[(size_t)t * c.es + e] = {fock[e], 0.0};
    NED_HIP(hipMemcpy(c.hmf.p, h.data(), h.size() * sizeof(ned::z2), hipMemcpyHostToDevice));
    double err = 0;
    int it = 0;
    bool conv = false;
    for (; it <= maxiter; it++) {
      for (int t = 0; t <= c.k; t++) ned::sigma_gf2(c, t, (ned::Prec)prec);
      err = ned::dyson_start(c, mu, beta, dt);
      if (!std::isfinite(err)) throw std::runtime_error("non-finite update in bootstrap");
      if (err < tol) { conv = true; it++; break; }
    }
    if (iters_out) *iters_out = it;
    if (err_out) *err_out = err;
    ned::invalidate_minv(c);
    if (!conv) throw std::runtime_error("bootstrap did not converge");
  });
  return st;
}

ned_status ned_do_tstp(ned_ctx *ctx, int tstp, const double *h0, double mu, double beta, double dt, int prec,
                       int corr_max, double corr_tol, int *iters_out, double *err_out) {
  if (!ctx || !h0) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::Ctx &c = ctx->c;
    ned::require(prec >= 0 && prec <= 2, "bad prec");
    ned::require(corr_max >= 1, "corr_max must be >= 1");
    state(c.have_coulomb, "interaction not set (ned_set_coulomb)");
    state(c.have_mix && c.mix_beta == beta, "ned_set_matsubara(beta) not called for this beta");
    ned::do_tstp(c, tstp, h0, mu, beta, dt, (ned::Prec)prec, corr_max, corr_tol, iters_out, err_out);
  });
}

ned_status ned_energy(ned_ctx *ctx, int tstp, const double *h0, double beta, double dt, double *out3) {
  if (!ctx || !h0 || !out3) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::require(tstp >= 0 && tstp <= ctx->c.nt, "tstp out of range");
    ned::energy(ctx->c, tstp, h0, beta, dt, out3);
  });
}

ned_status ned_g0_from_h0(ned_ctx *ctx, int tmax, const double *fock, double mu, double beta, double dt) {
  if (!ctx || !fock) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(ctx, [&] { ned::g0_from_h0(ctx->c, tmax, fock, mu, beta, dt); });
}

ned_status ned_tti_extend(ned_ctx *ctx, int which, int tsrc) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    ned::require(which == NED_G || which == NED_SIGMA, "which must be NED_G or NED_SIGMA");
    ned::tti_extend(ctx->c, which, tsrc);
    if (which == NED_SIGMA) ned::invalidate_minv(ctx->c);  // Sigma^R(t,t) enters the cached M_t
  });
}

ned_status ned_set_graph(ned_ctx *ctx, int enable) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  ctx->c.graph_mode = enable != 0;
  return NED_OK;
}

int ned_gemm_variant_count(void) { return ned::nt_variant_count(); }
const char *ned_gemm_variant_name(int v) { return ned::nt_variant_name(v); }

ned_status ned_profile(ned_ctx *ctx, int enable) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  return guard(ctx, [&] { ctx->c.prof = enable != 0; });
}

ned_status ned_profile_read(ned_ctx *ctx, double *out8) {
  if (!ctx || !out8) { g_err = "null argument"; return NED_ERR_ARG; }
  double *out6 = out8;
  return guard(ctx, [&] {
    ned::Ctx &c = ctx->c;
    NED_HIP(hipStreamSynchronize(c.stream));
    for (int ph = 0; ph < 3; ph++) {
      auto &a = c.ev[2 * ph], &b = c.ev[2 * ph + 1];
      double tot = 0;
      size_t cnt = std::min(a.size(), b.size());
      for (size_t i = 0; i < cnt; i++) {
        float ms = 0;
        NED_HIP(hipEventElapsedTime(&ms, a[i], b[i]));
        tot += ms;
      }
      for (auto e : a) (void)hipEventDestroy(e);
      for (auto e : b) (void)hipEventDestroy(e);
      a.clear();
      b.clear();
      out6[2 * ph] = tot;
      out6[2 * ph + 1] = (double)cnt;
    }
    out8[6] = c.prof_flops;
    out8[7] = 0;
    c.prof_flops = 0;
  });
}

ned_status ned_mlp_create(ned_mlp **out, int n_in, int width, int n_out, int device, unsigned long long seed) {
  if (!out) { g_err = "null argument"; return NED_ERR_ARG; }
  *out = nullptr;
  ned_mlp *h = nullptr;
  ned_status st = guard(nullptr, [&] {
    ned::require(n_in > 0 && width > 0 && n_out > 0, "network sizes must be positive");
    int ndev = 0;
    NED_HIP(hipGetDeviceCount(&ndev));
    ned::require(device >= 0 && device < ndev, "no such HIP device");
    NED_HIP(hipSetDevice(device));
    h = new ned_mlp();
    h->m.reset(new ned::Mlp(n_in, width, n_out, device, seed));
  });
  if (st != NED_OK) {
    delete h;
    return st;
  }
  *out = h;
  return NED_OK;
}

ned_status ned_mlp_destroy(ned_mlp *m) {
  if (!m) return NED_OK;
  (void)hipSetDevice(m->m->device);
  delete m;
  return NED_OK;
}

ned_status ned_mlp_nparams(ned_mlp *m, long long *n) {
  if (!n) { g_err = "null argument"; return NED_ERR_ARG; }
  return mlp_guard(m, [&](ned::Mlp &x) { *n = x.nparam; });
}

ned_status ned_mlp_get_params(ned_mlp *m, float *out) {
  if (!out) { g_err = "null argument"; return NED_ERR_ARG; }
  return mlp_guard(m, [&](ned::Mlp &x) {
    NED_HIP(hipMemcpy(out, x.params.p, x.nparam * sizeof(float), hipMemcpyDeviceToHost));
  });
}

ned_status ned_mlp_set_params(ned_mlp *m, const float *in) {
  if (!in) { g_err = "null argument"; return NED_ERR_ARG; }
  return mlp_guard(m, [&](ned::Mlp &x) {
    NED_HIP(hipMemcpy(x.params.p, in, x.nparam * sizeof(float), hipMemcpyHostToDevice));
    x.version++;
    x.reset_adam();
  });
}

ned_status ned_mlp_loss_grad(ned_mlp *m, const float *X, const float *T, int rows, double *loss, float *grads) {
  if (!X || !T || !loss) { g_err = "null argument"; return NED_ERR_ARG; }
  return mlp_guard(m, [&](ned::Mlp &x) {
    ned::require(rows > 0, "rows must be positive");
    ned::DevBuf dX, dT;
    dX.ensure((size_t)rows * x.n_in * sizeof(float));
    dT.ensure((size_t)rows * x.n_out * sizeof(float));
    NED_HIP(hipMemcpy(dX.p, X, (size_t)rows * x.n_in * sizeof(float), hipMemcpyHostToDevice));
    NED_HIP(hipMemcpy(dT.p, T, (size_t)rows * x.n_out * sizeof(float), hipMemcpyHostToDevice));
    *loss = x.loss_grad(dX.as<float>(), dT.as<float>(), rows, x.stream);
    if (grads) {
      NED_HIP(hipMemcpyAsync(grads, x.grads.p, x.nparam * sizeof(float), hipMemcpyDeviceToHost, x.stream));
      NED_HIP(hipStreamSynchronize(x.stream));
    }
  });
}

ned_status ned_mlp_adam_step(ned_mlp *m, float lr) {
  return mlp_guard(m, [&](ned::Mlp &x) {
    x.adam(lr, x.stream);
    NED_HIP(hipStreamSynchronize(x.stream));
  });
}

ned_status ned_mlp_train(ned_mlp *m, const float *X, const float *T, int rows, int epochs, int batch, float lr,
                         unsigned long long seed, float *loss_per_epoch) {
  if (!X || !T) { g_err = "null argument"; return NED_ERR_ARG; }
  return mlp_guard(m, [&](ned::Mlp &x) { x.train(X, T, rows, epochs, batch, lr, seed, loss_per_epoch); });
}

ned_status ned_mlp_predict(ned_mlp *m, const float *X, int rows, float *Y) {
  if (!X || !Y) { g_err = "null argument"; return NED_ERR_ARG; }
  return mlp_guard(m, [&](ned::Mlp &x) {
    ned::require(rows > 0, "rows must be positive");
    ned::DevBuf dX, dY;
    dX.ensure((size_t)rows * x.n_in * sizeof(float));
    dY.ensure((size_t)rows * x.n_out * sizeof(float));
    NED_HIP(hipMemcpy(dX.p, X, (size_t)rows * x.n_in * sizeof(float), hipMemcpyHostToDevice));
    x.infer(dX.as<float>(), rows, dY.as<float>(), x.stream);
    NED_HIP(hipMemcpyAsync(Y, dY.p, (size_t)rows * x.n_out * sizeof(float), hipMemcpyDeviceToHost, x.stream));
    NED_HIP(hipStreamSynchronize(x.stream));
  });
}

ned_status ned_mlp_set_bf16(ned_mlp *m, int on) {
  return mlp_guard(m, [&](ned::Mlp &x) {
    if (on)
      ned::require(x.n_in % 64 == 0 && x.w % 128 == 0 && x.n_out % 128 == 0,
                   "BF16 inference needs n_in a multiple of 64, width and n_out of 128");
    x.bf16_infer = on != 0;
  });
}

ned_status ned_set_predictor(ned_ctx *ctx, ned_mlp *m) {
  if (!ctx) { g_err = "null context"; return NED_ERR_ARG; }
  return guard(ctx, [&] { ned::set_predictor(ctx->c, m ? m->m.get() : nullptr); });
}

ned_status ned_predictor_residual(ned_ctx *ctx, int tstp, float *out) {
  if (!ctx || !out) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(ctx, [&] {
    const float *r = ned::predictor_residual(ctx->c, tstp);
    if (!r) throw std::invalid_argument("no extrapolation error recorded for this step (the last 7 steps are kept)");
    const size_t bytes = (size_t)ctx->c.es * ned::predictor_outputs() * sizeof(float);
    NED_HIP(hipMemcpyAsync(out, r, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
    NED_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

ned_status ned_predictor_stats(ned_ctx *ctx, long long *used, long long *fallbacks) {
  if (!ctx || !used || !fallbacks) { g_err = "null argument"; return NED_ERR_ARG; }
  *used = ctx->c.pred_used;
  *fallbacks = ctx->c.pred_fallback;
  return NED_OK;
}

ned_status ned_batch_create(ned_batch **out, ned_ctx *const *ctxs, int n) {
  if (!out || !ctxs || n <= 0) { g_err = "null argument or empty batch"; return NED_ERR_ARG; }
  *out = nullptr;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++)
      if (!ctxs[i] || ctxs[i] == ctxs[j]) { g_err = "batch members must be distinct contexts"; return NED_ERR_ARG; }
  ned_batch *b = new ned_batch();
  b->ctx.assign(ctxs, ctxs + n);
  *out = b;
  return NED_OK;
}

ned_status ned_batch_destroy(ned_batch *b) {
  delete b;
  return NED_OK;
}

ned_status ned_batch_set_grouped(ned_batch *b, int grouped) {
  if (!b) { g_err = "null batch"; return NED_ERR_ARG; }
  b->grouped = grouped != 0;
  return NED_OK;
}

ned_status ned_batch_do_tstp(ned_batch *b, int tstp, const double *const *h0, const double *mu, const double *beta,
                             const double *dt, int prec, int corr_max, double corr_tol, int *iters, double *err,
                             int *status) {
  if (!b || !h0 || !mu || !beta || !dt) { g_err = "null argument"; return NED_ERR_ARG; }
  const int n = (int)b->ctx.size();
  bool same_device = true;
  for (ned_ctx *c : b->ctx) same_device = same_device && c->c.device == b->ctx[0]->c.device;
  if (b->grouped && same_device) {
    std::vector<int> stv(n, 0);
    ned_status r = guard(b->ctx[0], [&] {
      ned::require(prec >= 0 && prec <= 2, "bad prec");
      ned::require(corr_max >= 1, "corr_max must be >= 1");
      std::vector<ned::Ctx *> cs;
      for (ned_ctx *c : b->ctx) {
        state(c->c.have_coulomb, "interaction not set (ned_set_coulomb)");
        state(c->c.have_mix, "ned_set_matsubara not called");
        cs.push_back(&c->c);
      }
      if (!b->group) b->group.reset(new ned::DysonGroup(b->ctx[0]->c.device));
      ned::batch_do_tstp(*b->group, cs, tstp, h0, mu, beta, dt, (ned::Prec)prec, corr_max, corr_tol, iters, err,
                         stv.data());
    });
    if (status)
      for (int i = 0; i < n; i++) status[i] = r != NED_OK ? r : stv[i];
    if (r != NED_OK) return r;
    for (int i = 0; i < n; i++)
      if (stv[i] != NED_OK) {
        g_err = "system " + std::to_string(i) +
                (stv[i] == NED_ERR_ARG ? ": tstp must be in (k, nt]" : ": non-finite update in corrector");
        return (ned_status)stv[i];
      }
    return NED_OK;
  }
  std::vector<ned_status> st(n, NED_OK);
  std::vector<std::string> msg(n);
  std::vector<std::thread> th;
  th.reserve(n);
  for (int i = 0; i < n; i++)
    th.emplace_back([&, i] {
      st[i] = ned_do_tstp(b->ctx[i], tstp, h0[i], mu[i], beta[i], dt[i], prec, corr_max, corr_tol,
                          iters ? iters + i : nullptr, err ? err + i : nullptr);
      if (st[i] != NED_OK) msg[i] = g_err;
    });
  for (auto &t : th) t.join();
  ned_status first = NED_OK;
  for (int i = 0; i < n; i++) {
    if (status) status[i] = st[i];
    if (st[i] != NED_OK && first == NED_OK) {
      first = st[i];
      g_err = "system " + std::to_string(i) + ": " + msg[i];
    }
  }
  return first;
}

ned_status ned_gll_nodes(int ntau, double *x) {
  if (!x) { g_err = "null argument"; return NED_ERR_ARG; }
  return guard(nullptr, [&] {
    ned::GLL g(ntau);
    std::memcpy(x, g.x.data(), sizeof(double) * g.np);
  });
}

ned_status ned_integrator(int k, double *interp, double *diff, double *integ, double *bd, double *start,
                          double *Omega, double *ex) {
  return guard(nullptr, [&] {
    ned::Weights W(k);
    const int k1 = k + 1;
    if (interp) std::memcpy(interp, W.interp.data(), sizeof(double) * k1 * k1);
    if (diff) std::memcpy(diff, W.diff.data(), sizeof(double) * k1 * k1);
    if (integ) std::memcpy(integ, W.integ.data(), sizeof(double) * k1 * k1 * k1);
    if (bd) std::memcpy(bd, W.bd.data(), sizeof(double) * (k + 2));
    if (start) std::memcpy(start, W.start.data(), sizeof(double) * k1 * k1);
    if (Omega) std::memcpy(Omega, W.Omega.data(), sizeof(double) * k1 * k1);
    if (ex) std::memcpy(ex, W.ex.data(), sizeof(double) * k1);
  });
}

}  // extern "C"
