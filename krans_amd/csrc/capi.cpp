// C ABI of libkrans_hip.so (declared in include/krans_hip.h).
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "../../include/krans_hip.h"
#include "capi_internal.hpp"
#include "dense.hpp"
#include "ingest.hpp"
#include "server.hpp"
#include "solver.hpp"

struct kh_server_s {
  std::unique_ptr<kh::Server> srv;
};
struct kh_host_csr_s {
  kh::HostCsr H;
};

namespace {
thread_local std::string g_err;

template <class F> int guard(F &&fn) {
  try {
    fn();
    return KH_OK;
  } catch (const std::invalid_argument &e) {
    g_err = e.what();
    return KH_EINVAL;
  } catch (const kh::DeviceOOM &e) {
    g_err = e.what();
    return KH_ENOMEM;
  } catch (const std::bad_alloc &e) {
    g_err = e.what();
    return KH_ENOMEM;
  } catch (const kh::DeviceError &e) {
    g_err = e.what();
    return KH_EDEVICE;
  } catch (const std::exception &e) {
    g_err = e.what();
    return KH_ERUNTIME;
  } catch (...) {
    g_err = "unknown error";
    return KH_ERUNTIME;
  }
}

void need(bool cond, const char *msg) {
  if (!cond)
    throw std::invalid_argument(msg);
}

kh::Prec prec_of(int p) {
  if (p == KH_PREC_FP64)
    return kh::Prec::F64;
  if (p == KH_PREC_MIXED)
    return kh::Prec::F32;
  throw std::invalid_argument("precision must be KH_PREC_FP64 or KH_PREC_MIXED");
}

kh::CyclicParams to_params(const kh_cyclic_options &o, int precision) {
  kh::CyclicParams P;
  P.k = o.k;
  P.ncv = o.ncv;
  P.max_restarts = o.max_restarts;
  P.tol_abs = o.tol_abs;
  P.tol_rel = o.tol_rel;
  need(o.reorth >= 0 && o.reorth <= 2, "reorth must be 0, 1 or 2");
  P.reorth = o.reorth;
  P.use_harmonic = o.use_harmonic != 0;
  P.target_tau = o.target_tau;
  P.nullity_known = o.nullity_known;
  P.rank_known = o.rank_known;
  P.cg_maxit = o.cg_maxit;
  P.cg_rel_tol = o.cg_rel_tol;
  P.cg_breakdown_tol = o.cg_breakdown_tol;
  P.projector_passes = o.projector_passes;
  P.grad_max_iters = o.grad_max_iters;
  P.frob_samples = o.frob_samples;
  P.prescreen_factor = o.prescreen_factor;
  P.stagnation_restarts = o.stagnation_restarts;
  P.return_left_space = o.return_left_space != 0;
  P.verbose = o.verbose != 0;
  P.jacobi_AAt = o.precond_jacobi != 0;
  P.graphs = o.graphs;
  need(o.screen >= -1 && o.screen <= 1, "screen must be -1, 0 or 1");
  P.screen = o.screen;
  need(o.second_pass >= 0 && o.second_pass <= 2, "second_pass must be 0, 1 or 2");
  P.second_pass = o.second_pass;
  need(o.second_pass_eta >= 0.0 && o.second_pass_eta <= 1.0, "second_pass_eta must be in [0, 1]");
  if (o.second_pass_eta > 0.0)
    P.second_pass_eta = o.second_pass_eta;
  P.krylov = prec_of(precision);
  P.polish = P.krylov;
  need(P.k >= 0 && P.ncv >= 0 && P.max_restarts >= 0 && P.cg_maxit >= 0,
       "negative size option");
  return P;
}

// out: host memory, or device memory of the context's device (dev)
void put_basis(kh::Context &ctx, const kh::DeviceBasis &B, double *out, int64_t cap,
               bool dev = false) {
  if (!out)
    return;
  const int64_t k = std::min(B.count, cap);
  if (k > 0)
    KH_CHECK(hipMemcpyAsync(out, B.buf.p, (std::size_t)(k * B.len) * 8,
                            dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
}
}  // namespace

extern "C" {

const char *kh_last_error(void) { return g_err.c_str(); }
int kh_version(void) { return 1; }

int kh_device_count(int *n) {
  return guard([&] {
    need(n, "null out");
    KH_CHECK(hipGetDeviceCount(n));
  });
}

int kh_context_create(int device, kh_context **out) {
  return guard([&] {
    need(out, "null out");
    auto c = std::make_unique<kh_context_s>();
    c->ctx = std::make_unique<kh::Context>(device);
    kh::noalloc_prepare(device);  // the scratch of kh_cgs2_dev_noalloc on this device
    *out = c.release();
  });
}

int kh_context_destroy(kh_context *ctx) {
  return guard([&] { delete ctx; });
}

int kh_context_sync(kh_context *ctx) {
  return guard([&] {
    need(ctx, "null context");
    ctx->ctx->sync();
  });
}

int kh_matrix_create_csr(kh_context *ctx, int64_t m, int64_t n, const int32_t *rowptr,
                         const int32_t *col, const double *val, kh_matrix **out) {
  return guard([&] {
    need(ctx && out, "null argument");
    auto M = std::make_unique<kh_matrix_s>();
    M->owner = ctx;
    M->A.reset(kh::csr_create(*ctx->ctx, m, n, rowptr, col, val));
    *out = M.release();
  });
}

int kh_load_ija(const char *path, int64_t rows, int64_t cols, kh_host_csr **out) {
  return guard([&] {
    need(path && out, "null argument");
    auto h = std::make_unique<kh_host_csr_s>();
    h->H = kh::load_ija(path, rows, cols);
    *out = h.release();
  });
}

int kh_load_matrix_market(const char *path, kh_host_csr **out) {
  return guard([&] {
    need(path && out, "null argument");
    auto h = std::make_unique<kh_host_csr_s>();
    h->H = kh::load_matrix_market(path);
    *out = h.release();
  });
}

int kh_csr_from_triplets(int64_t m, int64_t n, int64_t count, const int32_t *I,
                         const int32_t *J, const double *V, kh_host_csr **out) {
  return guard([&] {
    need(out, "null argument");
    auto h = std::make_unique<kh_host_csr_s>();
    h->H = kh::csr_from_triplets(m, n, I, J, V, count);
    *out = h.release();
  });
}

int kh_host_csr_info(const kh_host_csr *h, int64_t *m, int64_t *n, int64_t *nnz,
                     int64_t *declared_nnz, int64_t *entries_read) {
  return guard([&] {
    need(h, "null handle");
    if (m)
      *m = h->H.m;
    if (n)
      *n = h->H.n;
    if (nnz)
      *nnz = (int64_t)h->H.col.size();
    if (declared_nnz)
      *declared_nnz = h->H.declared_nnz;
    if (entries_read)
      *entries_read = h->H.entries_read;
  });
}

int kh_host_csr_data(const kh_host_csr *h, const int32_t **rowptr, const int32_t **col,
                     const double **val) {
  return guard([&] {
    need(h && rowptr && col && val, "null argument");
    *rowptr = h->H.rowptr.data();
    *col = h->H.col.data();
    *val = h->H.val.data();
  });
}

int kh_host_csr_free(kh_host_csr *h) {
  return guard([&] { delete h; });
}

int kh_matrix_create_shifted_difference(kh_context *ctx, int64_t n, int64_t N, double c,
                                        kh_matrix **out) {
  return guard([&] {
    need(ctx && out, "null argument");
    auto M = std::make_unique<kh_matrix_s>();
    M->owner = ctx;
    M->op = std::make_unique<kh::StencilOp>(*ctx->ctx, n, N, c);
    M->nnz_equiv = 2 * (n - N);
    *out = M.release();
  });
}

int kh_matrix_create_operator(kh_context *ctx, int64_t m, int64_t n, kh_apply_fn apply,
                              void *user, kh_matrix **out) {
  return guard([&] {
    need(ctx && out && apply, "null argument");
    need(m >= 0 && n >= 0, "negative dimensions");
    auto M = std::make_unique<kh_matrix_s>();
    M->owner = ctx;
    M->op = std::make_unique<kh::FunctorOp>(m, n, apply, user);
    M->nnz_equiv = -1;
    *out = M.release();
  });
}

int kh_matrix_destroy(kh_matrix *A) {
  return guard([&] { delete A; });
}

int kh_matrix_info(const kh_matrix *A, int64_t *m, int64_t *n, int64_t *nnz, double *fro) {
  return guard([&] {
    need(A, "null matrix");
    if (m)
      *m = A->rows();
    if (n)
      *n = A->cols();
    if (nnz)
      *nnz = A->A ? A->A->nnz : A->nnz_equiv;
    if (fro)
      *fro = A->fro();
  });
}

int kh_matrix_download_csr(const kh_matrix *A, int32_t *rowptr, int32_t *col, double *val) {
  return guard([&] {
    need(A, "null matrix");
    auto &c = *A->owner->ctx;
    const auto &D = A->csr("kh_matrix_download_csr");
    if (rowptr)
      KH_CHECK(hipMemcpyAsync(rowptr, D.rp.p, (std::size_t)(D.m + 1) * 4,
                              hipMemcpyDeviceToHost, c.stream));
    if (col && D.nnz)
      KH_CHECK(hipMemcpyAsync(col, D.ci.p, (std::size_t)D.nnz * 4, hipMemcpyDeviceToHost,
                              c.stream));
    if (val && D.nnz)
      KH_CHECK(hipMemcpyAsync(val, D.va.p, (std::size_t)D.nnz * 8, hipMemcpyDeviceToHost,
                              c.stream));
    c.sync();
  });
}

int kh_matrix_download_csc(const kh_matrix *A, int32_t *colptr, int32_t *rowidx, double *val,
                           int32_t *perm) {
  return guard([&] {
    need(A, "null matrix");
    auto &c = *A->owner->ctx;
    const auto &D = A->csr("kh_matrix_download_csc");
    if (colptr)
      KH_CHECK(hipMemcpyAsync(colptr, D.cp.p, (std::size_t)(D.n + 1) * 4,
                              hipMemcpyDeviceToHost, c.stream));
    if (rowidx && D.nnz)
      KH_CHECK(hipMemcpyAsync(rowidx, D.ri.p, (std::size_t)D.nnz * 4, hipMemcpyDeviceToHost,
                              c.stream));
    if (val && D.nnz)
      KH_CHECK(hipMemcpyAsync(val, D.vt.p, (std::size_t)D.nnz * 8, hipMemcpyDeviceToHost,
                              c.stream));
    if (perm && D.nnz)
      KH_CHECK(hipMemcpyAsync(perm, D.perm.p, (std::size_t)D.nnz * 4, hipMemcpyDeviceToHost,
                              c.stream));
    c.sync();
  });
}

int kh_matrix_multiply(kh_matrix *A, int transpose, int precision, const double *x,
                       double *y) {
  return guard([&] {
    need(A && x && y, "null argument");
    auto &c = *A->owner->ctx;
    const auto V = A->view(transpose != 0);
    const kh::Prec p = prec_of(precision);
    const int64_t nin = (*V).cols(), nout = (*V).rows();
    kh::DBuf dx((std::size_t)std::max<int64_t>(1, nin) * 8),
        dy((std::size_t)std::max<int64_t>(1, nout) * 8);
    KH_CHECK(hipMemcpyAsync(dx.p, x, (std::size_t)nin * 8, hipMemcpyHostToDevice, c.stream));
    if (p == kh::Prec::F64) {
      (*V).mv(c, p, dx.p, dy.p);
    } else {
      kh::DBuf x32((std::size_t)std::max<int64_t>(1, nin) * 4),
          y32((std::size_t)std::max<int64_t>(1, nout) * 4);
      kh::convert(c, kh::Prec::F64, dx.p, kh::Prec::F32, x32.p, nin);
      (*V).mv(c, p, x32.p, y32.p);
      kh::convert(c, kh::Prec::F32, y32.p, kh::Prec::F64, dy.p, nout);
      c.sync();
    }
    KH_CHECK(hipMemcpyAsync(y, dy.p, (std::size_t)nout * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int kh_rng_create(uint32_t seed, kh_rng **out) {
  return guard([&] {
    need(out, "null out");
    *out = new kh_rng_s(seed);
  });
}

int kh_rng_destroy(kh_rng *r) {
  return guard([&] { delete r; });
}

int kh_rng_normal(kh_rng *r, int64_t n, double *out) {
  return guard([&] {
    need(r && (out || n == 0) && n >= 0, "bad argument");
    r->s.fill(out, (std::size_t)n);
  });
}

int kh_tiny_svd_right(const double *a, int32_t rows, int32_t cols, int32_t mode, double *sv,
                      double *v) {
  return guard([&] {
    need(a && sv && v && cols >= 1 && rows >= cols && mode >= 0 && mode <= 2, "bad argument");
    kh::dense::Mat A(rows, cols);
    std::copy(a, a + (std::size_t)rows * cols, A.a.begin());
    std::vector<double> s;
    kh::dense::Mat V;
    kh::dense::svd_right(A, s, V, mode >= 1, mode == 2);
    std::copy(s.begin(), s.end(), sv);
    std::copy(V.a.begin(), V.a.end(), v);
  });
}

namespace {
int cyclic_entry(kh_matrix *A, kh_rng *rng, const kh_cyclic_options *opt, int precision,
                 const double *y0, double *right, int64_t right_cap, double *left,
                 int64_t left_cap, kh_cyclic_result *res, bool dev) {
  return guard([&] {
    need(A && rng && opt && res, "null argument");
    auto &c = *A->owner->ctx;
    kh::CyclicParams P = to_params(*opt, precision);
    P.y0_host = y0;
    const auto op = A->view(false);
    kh::CyclicStats S;
    kh::DeviceBasis R = kh::cyclic_right(c, *op, rng->s, P, S);
    std::memset(res, 0, sizeof(*res));
    res->n_right = R.count;
    put_basis(c, R, right, right_cap, dev);
    kh::CyclicStats SL;
    if (P.return_left_space && A->rows() > 0 && A->cols() > 0) {
      // CyclicLanczosSolver.cpp:626-659: a second run on A^T with the listed
      // options copied and no preconditioner.
      kh::CyclicParams L = P;
      L.y0_host = nullptr;
      L.jacobi_AAt = false;
      L.nullity_known = -1;
      if (P.rank_known >= 0) {
        const int left_nullity = (int)A->rows() - P.rank_known;
        L.nullity_known = std::max(0, std::min(left_nullity, (int)P.k));
      }
      const auto opt_t = A->view(true);
      kh::DeviceBasis Lb = kh::cyclic_right(c, *opt_t, rng->s, L, SL);
      res->n_left = Lb.count;
      put_basis(c, Lb, left, left_cap, dev);
    }
    res->restarts = S.restarts;
    res->krylov_steps_total = S.krylov_steps_total;
    res->frob_norm_est = S.frob_norm_est;
    res->converged_to_target = S.converged_to_target;
    res->cg_solves = S.cg_solves + SL.cg_solves;
    res->cg_maxit_hits = S.cg_maxit_hits + SL.cg_maxit_hits;
    res->spmv = S.spmv + SL.spmtv;
    res->spmtv = S.spmtv + SL.spmv;
    res->cg_iters = S.cg_iters + SL.cg_iters;
    res->t_gk = S.t.gk + SL.t.gk;
    res->t_tiny = S.t.tiny + SL.t.tiny;
    res->t_cand = S.t.cand + SL.t.cand;
    res->t_polish = S.t.polish + SL.t.polish;
    res->t_fallback = S.t.fallback + SL.t.fallback;
    res->t_rng = S.t.rng + SL.t.rng;
    res->screened = S.screened + SL.screened;
    res->refined = S.refined + SL.refined;
    res->escalated = S.escalated + SL.escalated;
    res->graph_captures = S.graph_captures + SL.graph_captures;
    res->graph_replays = S.graph_replays + SL.graph_replays;
    res->cgs_pairs = S.cgs_pairs + SL.cgs_pairs;
    res->second_passes = S.second_passes + SL.second_passes;
    res->gk_requeued = S.gk_requeued + SL.gk_requeued;
  });
}
}  // namespace

int kh_cyclic_lanczos(kh_matrix *A, kh_rng *rng, const kh_cyclic_options *opt,
                      int precision, const double *y0, double *right, int64_t right_cap,
                      double *left, int64_t left_cap, kh_cyclic_result *res) {
  return cyclic_entry(A, rng, opt, precision, y0, right, right_cap, left, left_cap, res, false);
}

int kh_cyclic_lanczos_dev(kh_matrix *A, kh_rng *rng, const kh_cyclic_options *opt,
                          int precision, const double *y0, double *right_dev,
                          int64_t right_cap, double *left_dev, int64_t left_cap,
                          kh_cyclic_result *res) {
  return cyclic_entry(A, rng, opt, precision, y0, right_dev, right_cap, left_dev, left_cap, res,
                      true);
}

int kh_lobpcg(kh_matrix *A, kh_rng *rng, const kh_lobpcg_options *opt, int precision,
              const double *initial_block, int64_t n_initial, double *right, int64_t right_cap,
              double *theta, double *residuals, kh_lobpcg_result *res) {
  return guard([&] {
    need(A && rng && opt && res, "null argument");
    auto &c = *A->owner->ctx;
    kh::LobpcgParams P;
    P.nev = opt->nev;
    P.block_size = opt->block_size;
    P.max_iters = opt->max_iters;
    P.tol_abs = opt->tol_abs;
    P.tol_rel = opt->tol_rel;
    P.soft_locking = opt->soft_locking != 0;
    P.gs_reorth = opt->gs_reorth != 0;
    P.restart = opt->restart;
    P.gs_tol = opt->gs_tol;
    P.frob_samples = opt->frob_samples;
    P.nullity_known = opt->nullity_known;
    P.rank_known = opt->rank_known;
    P.verbose = opt->verbose != 0;
    P.max_tiny_dim = opt->max_tiny_dim;
    P.progress_every = opt->progress_every;
    P.block = prec_of(precision);
    P.initial_block_host = initial_block;
    P.n_initial = initial_block ? n_initial : 0;
    const auto op = A->view(false);
    kh::LobpcgStats S;
    kh::DeviceBasis R = kh::lobpcg(c, *op, rng->s, P, S);
    std::memset(res, 0, sizeof(*res));
    res->n_right = R.count;
    put_basis(c, R, right, right_cap);
    for (int64_t i = 0; i < R.count && i < right_cap; ++i) {
      if (theta)
        theta[i] = S.theta[(std::size_t)i];
      if (residuals)
        residuals[i] = S.residuals_Av[(std::size_t)i];
    }
    res->iters = S.iters;
    res->locked = S.locked;
    res->frob_norm_est = S.frob_norm_est;
    res->accept_threshold = S.accept_threshold;
    res->best_unaccepted_residual = S.best_unaccepted_residual;
    res->block_size = S.block_size;
    res->converged_to_target = S.converged_to_target;
    res->spmv = S.spmv;
    res->spmtv = S.spmtv;
    res->t_gram = S.t_gram;
    res->t_sparse = S.t_sparse;
    res->t_host = S.t_host;
    res->cholqr_blocks = S.cholqr_blocks;
    res->cholqr_fallbacks = S.cholqr_fallbacks;
  });
}

int kh_server_create(int device, int slots, double hbm_budget_bytes, kh_server **out) {
  return guard([&] {
    need(out, "null out");
    auto s = std::make_unique<kh_server_s>();
    s->srv = std::make_unique<kh::Server>(device, slots, hbm_budget_bytes);
    *out = s.release();
  });
}

int kh_server_destroy(kh_server *s) {
  return guard([&] { delete s; });
}

int kh_server_submit_cyclic(kh_server *s, int64_t m, int64_t n, const int32_t *rowptr,
                            const int32_t *col, const double *val, const kh_cyclic_options *opt,
                            int precision, uint32_t seed, int64_t *id) {
  return guard([&] {
    need(s && opt && id, "null argument");
    *id = s->srv->submit(m, n, rowptr, col, val, to_params(*opt, precision), seed);
  });
}

int kh_server_submit_resident(kh_server *s, kh_matrix *A, const kh_cyclic_options *opt,
                              int precision, uint32_t seed, int64_t *id) {
  return guard([&] {
    need(s && A && opt && id, "null argument");
    *id = s->srv->submit_resident(&A->csr("kh_server_submit_resident"),
                                  to_params(*opt, precision), seed);
  });
}

int kh_server_wait(kh_server *s, int64_t id, double *right, int64_t cap, kh_cyclic_result *res,
                   double *latency_s, double *queued_s) {
  return guard([&] {
    need(s && res, "null argument");
    need(cap >= 0, "negative basis capacity");
    kh::Server::Result r = s->srv->wait(id);
    std::memset(res, 0, sizeof(*res));
    if (r.status == 1)
      throw std::invalid_argument(r.error);
    if (r.status == 3)
      throw kh::DeviceOOM(r.error);
    if (r.status != 0)
      throw std::runtime_error(r.error);
    res->n_right = r.count;
    if (right)
      std::memcpy(right, r.basis.data(),
                  (std::size_t)(std::min(r.count, cap) * r.n) * sizeof(double));
    res->restarts = r.stats.restarts;
    res->krylov_steps_total = r.stats.krylov_steps_total;
    res->frob_norm_est = r.stats.frob_norm_est;
    res->converged_to_target = r.stats.converged_to_target;
    res->cg_solves = r.stats.cg_solves;
    res->cg_maxit_hits = r.stats.cg_maxit_hits;
    res->spmv = r.stats.spmv;
    res->spmtv = r.stats.spmtv;
    res->cg_iters = r.stats.cg_iters;
    res->t_gk = r.stats.t.gk;
    res->t_tiny = r.stats.t.tiny;
    res->t_cand = r.stats.t.cand;
    res->t_polish = r.stats.t.polish;
    res->t_fallback = r.stats.t.fallback;
    res->t_rng = r.stats.t.rng;
    res->screened = r.stats.screened;
    res->refined = r.stats.refined;
    res->escalated = r.stats.escalated;
    res->packed = r.packed;
    res->graph_captures = r.stats.graph_captures;
    res->graph_replays = r.stats.graph_replays;
    res->cgs_pairs = r.stats.cgs_pairs;
    res->second_passes = r.stats.second_passes;
    res->gk_requeued = r.stats.gk_requeued;
    if (latency_s)
      *latency_s = r.latency;
    if (queued_s)
      *queued_s = r.queued;
  });
}

int kh_server_set_packing(kh_server *s, int pack_max, int64_t pack_len) {
  return guard([&] {
    need(s, "null server");
    s->srv->set_packing(pack_max, pack_len);
  });
}

int kh_server_stats(kh_server *s, int64_t *done, int64_t *vectors, double *p50, double *p99,
                    double *wall_s) {
  return guard([&] {
    need(s, "null server");
    kh::Server::Stats st = s->srv->stats();
    if (done)
      *done = st.done;
    if (vectors)
      *vectors = st.vectors;
    if (p50)
      *p50 = st.p50;
    if (p99)
      *p99 = st.p99;
    if (wall_s)
      *wall_s = st.wall;
  });
}

int kh_server_reset_clock(kh_server *s) {
  return guard([&] {
    need(s, "null server");
    s->srv->reset_clock();
  });
}

int kh_dev_alloc(kh_context *ctx, size_t bytes, void **out) {
  return guard([&] {
    need(ctx && out, "null argument");
    KH_CHECK(hipSetDevice(ctx->ctx->device));
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      throw kh::DeviceOOM("hipMalloc: out of device memory");
    }
    KH_CHECK(e);
    *out = p;
  });
}

int kh_dev_free(kh_context *ctx, void *p) {
  return guard([&] {
    need(ctx, "null context");
    ctx->ctx->sync();
    KH_CHECK(hipFree(p));
  });
}

int kh_h2d(kh_context *ctx, void *dst, const void *src, size_t bytes) {
  return guard([&] {
    need(ctx, "null context");
    KH_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->ctx->stream));
    ctx->ctx->sync();
  });
}

int kh_d2h(kh_context *ctx, void *dst, const void *src, size_t bytes) {
  return guard([&] {
    need(ctx, "null context");
    KH_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->ctx->stream));
    ctx->ctx->sync();
  });
}

int kh_vec_axpby(kh_context *ctx, int64_t len, double alpha, const double *x, double beta,
                 double *y) {
  return guard([&] {
    need(ctx && y && len >= 0, "bad argument");
    auto &c = *ctx->ctx;
    kh::axpby(c, kh::Prec::F64, len, alpha, x, beta, y);
    c.sync();
  });
}

int kh_matrix_jacobi_AAt(kh_matrix *A, const double *r, double *y) {
  return guard([&] {
    need(A && r && y, "null argument");
    auto &c = *A->owner->ctx;
    const double *d = A->A ? A->A->diag_AAt.as<double>() : A->op->diag_AAt();
    need(d != nullptr, "kh_matrix_jacobi_AAt: this operator has no diag(A A^T)");
    kh::diag_solve(c, kh::Prec::F64, A->rows(), d, r, y);
    c.sync();
  });
}

int kh_spmm_dev(kh_matrix *A, int transpose, int precision, int s, const void *x, void *y) {
  return guard([&] {
    need(A && x && y && s >= 1, "bad argument");
    auto &c = *A->owner->ctx;
    const auto V = A->view(transpose != 0);
    const int64_t nin = (*V).cols(), nout = (*V).rows();
    if (s == 1)
      (*V).mv(c, prec_of(precision), x, y);
    else
      (*V).mm(c, prec_of(precision), s, x, nin, y, nout, nullptr);
    c.sync();
  });
}

int kh_gemm_tn_dev(kh_context *ctx, int precision, const void *V, int s, const void *W, int q,
                   int64_t len, double *H) {
  return guard([&] {
    need(ctx && V && W && H && s >= 1 && q >= 1 && len >= 1, "bad argument");
    auto &c = *ctx->ctx;
    const kh::Prec p = prec_of(precision);
    kh::DBuf d((std::size_t)s * q * 8);
    kh::gemm_tn(c, p, V, len, s, p, W, len, q, len, d.as<double>());
    KH_CHECK(hipMemcpyAsync(H, d.p, (std::size_t)s * q * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int kh_gemm_nn_dev(kh_context *ctx, int precision, const void *V, int s, const double *H, int q,
                   void *W, int64_t len, int subtract) {
  return guard([&] {
    need(ctx && V && H && W && s >= 1 && q >= 1 && len >= 1, "bad argument");
    auto &c = *ctx->ctx;
    const kh::Prec p = prec_of(precision);
    kh::DBuf d((std::size_t)s * q * 8);
    KH_CHECK(hipMemcpyAsync(d.p, H, (std::size_t)s * q * 8, hipMemcpyHostToDevice, c.stream));
    if (subtract)
      kh::gemm_nn_sub(c, p, V, len, s, d.as<double>(), q, p, W, len, len, nullptr);
    else
      kh::gemm_nn(c, p, V, len, s, d.as<double>(), q, p, W, len, len);
    c.sync();
  });
}

int kh_spmm_rows_dev(kh_matrix *A, int transpose, int precision, int s, const void *x,
                     void *y) {
  return guard([&] {
    need(A && x && y && s >= 1, "bad argument");
    auto &c = *A->owner->ctx;
    const auto V = A->view(transpose != 0);
    need((*V).has_rows(), "kh_spmm_rows_dev: this operator has no row-major block product");
    (*V).mm_rows(c, prec_of(precision), s, x, y, false);
    c.sync();
  });
}

int kh_dot_dev(kh_context *ctx, int precision, const void *a, const void *b, int64_t len,
               double *out) {
  return guard([&] {
    need(ctx && a && b && out, "null argument");
    auto &c = *ctx->ctx;
    const kh::Prec p = prec_of(precision);
    kh::dot(c, p, a, p, b, len, c.dscal, false);
    KH_CHECK(hipMemcpyAsync(out, c.dscal, 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int kh_cgs2_dev_ex(kh_context *ctx, int precision, const void *V, int s, void *w, int64_t len,
                   int fused, double *h_out, double *norm_out) {
  return guard([&] {
    need(ctx && V && w && s >= 0 && s <= 4096, "bad argument");
    need(!fused || s <= 1024, "fused CGS2 takes at most 1024 columns");
    auto &c = *ctx->ctx;
    const kh::Prec p = prec_of(precision);
    double *h1 = c.dscal + 16, *h2 = h1 + 4096, *nrm = c.dscal + 8;
    kh::gemv_t(c, p, V, len, s, p, w, len, h1);
    if (fused && s > 0) {
      kh::gemv_n_sub_t(c, p, V, len, s, h1, p, w, len, h2);
    } else {
      kh::gemv_n_sub(c, p, V, len, s, h1, p, w, len, nullptr);
      kh::gemv_t(c, p, V, len, s, p, w, len, h2);
    }
    kh::gemv_n_sub(c, p, V, len, s, h2, p, w, len, nrm);
    std::vector<double> a((std::size_t)s), b((std::size_t)s);
    KH_CHECK(hipMemcpyAsync(a.data(), h1, (std::size_t)s * 8, hipMemcpyDeviceToHost, c.stream));
    KH_CHECK(hipMemcpyAsync(b.data(), h2, (std::size_t)s * 8, hipMemcpyDeviceToHost, c.stream));
    double nn = 0;
    KH_CHECK(hipMemcpyAsync(&nn, nrm, 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (h_out)
      for (int j = 0; j < s; ++j)
        h_out[j] = a[(std::size_t)j] + b[(std::size_t)j];
    if (norm_out)
      *norm_out = nn;
  });
}

int kh_cgs2_dev_select(kh_context *ctx, int precision, const void *V, int s, void *w,
                       int64_t len, double *h_out, double *norm_out, int32_t *second_out) {
  return guard([&] {
    need(ctx && V && w && s >= 1 && s < 512 && len >= 1, "bad argument (1 <= s < 512)");
    auto &c = *ctx->ctx;
    const kh::Prec p = prec_of(precision);
    double *g = c.dscal + 16, *h2 = g + 4096, *nrm = c.dscal + 8, *ran = c.dscal + 9;
    kh::gemv_t(c, p, V, len, s, p, w, len, g, /*wsq=*/true);  // h1 and ||w||^2
    kh::gemv_n_sub_t(c, p, V, len, s, g, p, w, len, h2, nullptr, g);
    kh::gemv_n_sub(c, p, V, len, s, h2, p, w, len, nrm, g, ran);
    std::vector<double> a((std::size_t)s), b((std::size_t)s);
    double tail[2] = {0, 0};
    KH_CHECK(hipMemcpyAsync(a.data(), g, (std::size_t)s * 8, hipMemcpyDeviceToHost, c.stream));
    KH_CHECK(hipMemcpyAsync(b.data(), h2, (std::size_t)s * 8, hipMemcpyDeviceToHost, c.stream));
    KH_CHECK(hipMemcpyAsync(tail, nrm, 16, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (h_out)
      for (int j = 0; j < s; ++j)
        h_out[j] = a[(std::size_t)j] + b[(std::size_t)j];
    if (norm_out)
      *norm_out = tail[0];
    if (second_out)
      *second_out = tail[1] > 0.5 ? 1 : 0;
  });
}

int kh_cgs2_dev(kh_context *ctx, int precision, const void *V, int s, void *w, int64_t len,
                double *h_out, double *norm_out) {
  return kh_cgs2_dev_ex(ctx, precision, V, s, w, len, s <= 1024 ? 1 : 0, h_out, norm_out);
}

// kh_cgs2_dev's kernels (fused form) with their reduction partials and the
// second-pass coefficients in the per-device static scratch of noalloc.hip
// instead of the context's buffers, h and ||w|| written to device memory and
// no wait: no allocation, no host copy, no synchronisation.
int kh_cgs2_dev_noalloc(kh_context *ctx, int precision, const void *V, int s, void *w,
                        int64_t len, double *h_out, double *norm_out) {
  return guard([&] {
    need(ctx && V && w && h_out && norm_out, "null argument");
    need(s >= 0 && s <= 1024 && len >= 0, "bad argument (0 <= s <= 1024, len >= 0)");
    auto &c = *ctx->ctx;
    need(c.trace_mask == 0, "kernel tracing is on for this context (it records events)");
    need(c.part_ext == nullptr, "a no-alloc call is already running on this context");
    const kh::Prec p = prec_of(precision);
    std::size_t cap = 0;
    double *scr = kh::noalloc_scratch(c.device, &cap);
    // layout: h2 (s doubles, padded to 256 bytes) | reduction partials
    const std::size_t hpad = ((std::size_t)s + 31) / 32 * 32;
    const std::size_t part = kh::cgs2_partials_bound(len, s);
    if (hpad + part > cap)
      throw std::invalid_argument(
          "kh_cgs2_dev_noalloc: len x s too large for the 32 MiB device scratch (" +
          std::to_string(hpad + part) + " doubles needed, " + std::to_string(cap) +
          " available)");
    struct Redirect {  // partials go to the scratch for this call only
      kh::Context &c;
      Redirect(kh::Context &cc, double *pp, std::size_t n) : c(cc) {
        c.part_ext = pp;
        c.part_ext_cap = n;
      }
      ~Redirect() {
        c.part_ext = nullptr;
        c.part_ext_cap = 0;
      }
    } redirect(c, scr + hpad, cap - hpad);
    double *h2 = scr;
    kh::gemv_t(c, p, V, len, s, p, w, len, h_out);  // h1
    if (s > 0)
      kh::gemv_n_sub_t(c, p, V, len, s, h_out, p, w, len, h2);
    kh::gemv_n_sub(c, p, V, len, s, h2, p, w, len, norm_out);
    if (s > 0)
      kh::axpby(c, kh::Prec::F64, s, 1.0, h2, 1.0, h_out);  // h1 + h2
  });
}

int kh_gram_dev(kh_context *ctx, int precision, const void *Z, const void *AZ, int k,
                int64_t len, double *S, double *G) {
  return guard([&] {
    need(ctx && Z && AZ && S && G && k >= 1, "bad argument");
    auto &c = *ctx->ctx;
    const kh::Prec p = prec_of(precision);
    kh::DBuf d((std::size_t)k * k * 2 * 8);
    if (p == kh::Prec::F32 && k <= 128) {  // fused FP32 MFMA pass
      kh::gram2_mfma(c, (const float *)Z, len, (const float *)AZ, len, k, len, d.as<double>());
    } else if (k <= 128) {  // fused FP64 MFMA pass
      kh::gram2_mfma(c, (const double *)Z, len, (const double *)AZ, len, k, len, d.as<double>());
    } else if (p == kh::Prec::F32) {  // FP32 MFMA tiles
      kh::gram_mfma(c, (const float *)Z, len, k, (const float *)Z, len, k, len, d.as<double>());
      kh::gram_mfma(c, (const float *)Z, len, k, (const float *)AZ, len, k, len,
                    d.as<double>() + (std::size_t)k * k);
    } else {
      kh::gemm_tn(c, p, Z, len, k, p, Z, len, k, len, d.as<double>());
      kh::gemm_tn(c, p, Z, len, k, p, AZ, len, k, len, d.as<double>() + (std::size_t)k * k);
    }
    KH_CHECK(hipMemcpyAsync(S, d.p, (std::size_t)k * k * 8, hipMemcpyDeviceToHost, c.stream));
    KH_CHECK(hipMemcpyAsync(G, d.as<double>() + (std::size_t)k * k, (std::size_t)k * k * 8,
                            hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int kh_trace_begin(kh_context *ctx, int kernel_class) {
  return guard([&] {
    need(ctx && kernel_class >= 0 && kernel_class <= 5, "bad argument");
    auto &c = *ctx->ctx;
    c.sync();
    c.trace_reset();
    c.trace_mask = kernel_class ? (1u << kernel_class) : 0u;
  });
}

int kh_trace_begin_mask(kh_context *ctx, unsigned mask) {
  return guard([&] {
    need(ctx && (mask & ~0x3eu) == 0, "bad class mask (bits 1..5)");
    auto &c = *ctx->ctx;
    c.sync();
    c.trace_reset();
    c.trace_mask = mask;
  });
}

int kh_trace_class(kh_context *ctx, int kernel_class, int64_t *launches, double *bytes,
                   double *ms) {
  return guard([&] {
    need(ctx && launches && bytes && ms, "null argument");
    auto &c = *ctx->ctx;
    c.sync();
    double tb = 0.0, tm = 0.0;
    int64_t nl = 0;
    for (auto &r : c.trace) {
      if (r.cls != kernel_class)
        continue;
      float e = 0.0f;
      KH_CHECK(hipEventElapsedTime(&e, r.a, r.b));
      tb += r.bytes;
      tm += (double)e;
      ++nl;
    }
    *launches = nl;
    *bytes = tb;
    *ms = tm;
  });
}

int kh_trace_end(kh_context *ctx, int64_t *launches, double *bytes, double *ms) {
  return guard([&] {
    need(ctx && launches && bytes && ms, "null argument");
    auto &c = *ctx->ctx;
    c.sync();
    double tb = 0.0, tm = 0.0;
    int64_t nl = 0;
    for (auto &r : c.trace) {
      if (r.cls == kh::kTraceNone)  // a launch the device skipped (no bytes moved)
        continue;
      float e = 0.0f;
      KH_CHECK(hipEventElapsedTime(&e, r.a, r.b));
      tb += r.bytes;
      tm += (double)e;
      ++nl;
    }
    *launches = nl;
    *bytes = tb;
    *ms = tm;
    c.trace_mask = 0;
    c.trace_reset();
  });
}

int kh_trace_shapes(kh_context *ctx, int kernel_class, int max, int64_t *len, int32_t *s,
                    int64_t *count, int32_t *n_shapes) {
  return guard([&] {
    need(ctx && n_shapes && max >= 0, "bad argument");
    need(max == 0 || (len && s && count), "null argument");
    auto &c = *ctx->ctx;
    int n = 0;
    for (auto &r : c.trace) {
      if (r.cls != kernel_class || r.s <= 0)
        continue;
      int i = 0;
      while (i < n && !(len[i] == r.len && s[i] == r.s))
        ++i;
      if (i < n) {
        ++count[i];
      } else if (n < max) {
        len[n] = r.len;
        s[n] = r.s;
        count[n] = 1;
        ++n;
      } else {
        throw std::length_error("kh_trace_shapes: more distinct shapes than max");
      }
    }
    *n_shapes = n;
  });
}

int kh_time_vtw(kh_context *ctx, int precision, int64_t len, int s, int reps, double *avg_ms) {
  return guard([&] {
    need(ctx && avg_ms, "null argument");
    *avg_ms = kh::time_vtw(*ctx->ctx, prec_of(precision), len, s, reps);
  });
}

int kh_spmv_kernel(kh_matrix *A, int transpose, int precision, int set, int *path) {
  return guard([&] {
    need(A && path, "null argument");
    const kh::DeviceCsr &M = A->csr("kh_spmv_kernel");
    const kh::Prec p = prec_of(precision);
    if (set) {
      const int v = *path;
      need(v >= 0 && v <= 3, "path must be 0..3");
      const int d = transpose ? 1 : 0;
      need(v < 2 || M.span_rb[d][v - 2] > 0, "no window table for that block size");
      if (v >= 2 && (int64_t)M.span_len[d][v - 2] * (p == kh::Prec::F64 ? 8 : 4) >
                        A->owner->ctx->lds_max)
        throw std::invalid_argument("that window exceeds the LDS of one workgroup (" +
                                    std::to_string(A->owner->ctx->lds_max) + " bytes)");
      kh::set_spmv_path(M, transpose != 0, p, v);
    }
    *path = kh::spmv_path(M, transpose != 0, p);
  });
}

int kh_time_spmv(kh_matrix *A, int transpose, int precision, int reps, double *avg_ms) {
  return guard([&] {
    need(A && avg_ms && reps >= 1, "bad argument");
    auto &c = *A->owner->ctx;
    const auto V = A->view(transpose != 0);
    const kh::Prec p = prec_of(precision);
    const std::size_t e = p == kh::Prec::F64 ? 8 : 4;
    const int64_t nin = (*V).cols(), nout = (*V).rows();
    kh::DBuf x((std::size_t)std::max<int64_t>(1, nin) * e),
        y((std::size_t)std::max<int64_t>(1, nout) * e);
    kh::fill_zero(c, p, nin, x.p);
    hipEvent_t e0, e1;
    KH_CHECK(hipEventCreate(&e0));
    KH_CHECK(hipEventCreate(&e1));
    (*V).mv(c, p, x.p, y.p);
    KH_CHECK(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; ++r)
      (*V).mv(c, p, x.p, y.p);
    KH_CHECK(hipEventRecord(e1, c.stream));
    KH_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    KH_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / reps;
  });
}

}  // extern "C"
