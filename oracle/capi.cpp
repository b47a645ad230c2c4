// C ABI of the oracle (liboracle.so), loaded by tests/ and bench.py's CPU
// baseline through ctypes. Test infrastructure only.
//
// Parity status (how each restated function is pinned; see tests/):
//  - CsrOp products: exact closed forms (shifted difference stencil vs CSR,
//    bit-identical), known-answer values of tests/test_sparse_adapter.cpp:33-45,
//    adjoint identity <Ax,y> = <x,A^T y>; CSC twin == numpy transpose.
//  - Factory: libstdc++ std::mt19937 + normal_distribution (the same library
//    the reference links); pinned by the first mt19937 output 3499211612 for
//    the default seed 5489 (C++ standard [rand.predef]) and by moments.
//  - svd_jacobi / sym_eig / cholesky: LAPACK (numpy) fixtures.
//  - cyclic_lanczos: README.md:76-80 quick-start results (3 vectors,
//    max||Au|| ~1e-15, dist to the exact kernel ~1e-15); the scenarios of
//    tests/test_cyclic_lanczos.cpp (2x3, identity, 2x2 rank-1 with left space,
//    1e6 scaling); dense-SVD null spaces of data/mat1-style matrices; exact
//    kernels of the generators.
//  - lobpcg: README.md:79 (3 vectors, max||Au|| = 0.01*thr = 3.5e-11 target);
//    scenarios of tests/test_lobpcg.cpp; same exact kernels.
//  - subspace_diff: tests/test_subspace_compare.cpp:14-96 values.
// Trajectory-level counters (restarts, cg_solves, iters) are parity unpinned:
// only a build of the reference with Eigen could pin them.
#include "oracle.hpp"

#include <algorithm>
#include <cstring>
#include <string>

namespace orc {

CsrOp::CsrOp(std::size_t m_, std::size_t n_, const int32_t *rowptr,
             const int32_t *col, const double *val)
    : m(m_), n(n_) {
  const std::size_t nnz = (std::size_t)rowptr[m];
  rp.assign(rowptr, rowptr + m + 1);
  ci.assign(col, col + nnz);
  va.assign(val, val + nnz);
  // CSC twin by a stable counting sort over rows (rows ascend within a column)
  cp.assign(n + 1, 0);
  for (std::size_t p = 0; p < nnz; ++p)
    cp[(std::size_t)ci[p] + 1]++;
  for (std::size_t j = 0; j < n; ++j)
    cp[j + 1] += cp[j];
  std::vector<int32_t> next(cp.begin(), cp.end() - 1);
  ri.resize(nnz);
  vt.resize(nnz);
  perm.resize(nnz);
  for (std::size_t i = 0; i < m; ++i)
    for (int32_t p = rp[i]; p < rp[i + 1]; ++p) {
      const int32_t q = next[(std::size_t)ci[(std::size_t)p]]++;
      ri[(std::size_t)q] = (int32_t)i;
      vt[(std::size_t)q] = va[(std::size_t)p];
      perm[(std::size_t)q] = p;
    }
}

void CsrOp::mv(const Vec &x, Vec &y) const {
  for (std::size_t i = 0; i < m; ++i) {
    double s = 0.0;
    for (int32_t p = rp[i]; p < rp[i + 1]; ++p)
      s += va[(std::size_t)p] * x[(std::size_t)ci[(std::size_t)p]];
    y[i] = s;
  }
}

void CsrOp::mtv(const Vec &x, Vec &y) const {
  for (std::size_t j = 0; j < n; ++j) {
    double s = 0.0;
    for (int32_t p = cp[j]; p < cp[j + 1]; ++p)
      s += vt[(std::size_t)p] * x[(std::size_t)ri[(std::size_t)p]];
    y[j] = s;
  }
}

// SubspaceCompare.hpp:17-78
void compute_rref(std::vector<Vec> &rows, double tol) {
  if (rows.empty())
    return;
  const std::size_t num_rows = rows.size(), num_cols = rows[0].size();
  std::size_t pivot_row = 0;
  for (std::size_t pc = 0; pc < num_cols && pivot_row < num_rows; ++pc) {
    std::size_t best_row = pivot_row;
    double max_val = 0.0;
    for (std::size_t r = pivot_row; r < num_rows; ++r) {
      const double val = std::abs(rows[r][pc]);
      if (val > max_val) {
        max_val = val;
        best_row = r;
      }
    }
    if (max_val < tol) {
      for (std::size_t r = pivot_row; r < num_rows; ++r)
        rows[r][pc] = 0.0;
      continue;
    }
    if (best_row != pivot_row)
      std::swap(rows[pivot_row], rows[best_row]);
    const double pv = rows[pivot_row][pc];
    scal(rows[pivot_row], 1.0 / pv);
    rows[pivot_row][pc] = 1.0;
    for (std::size_t r = 0; r < num_rows; ++r)
      if (r != pivot_row) {
        const double f = rows[r][pc];
        if (std::abs(f) > tol)
          axpy(rows[r], rows[pivot_row], -f);
        rows[r][pc] = 0.0;
      }
    pivot_row++;
  }
  for (auto &row : rows)
    for (double &v : row)
      if (std::abs(v) < tol)
        v = 0.0;
}

// SubspaceCompare.hpp:86-110
double subspace_diff(const std::vector<Vec> &basis1,
                     const std::vector<Vec> &basis2, double tol) {
  if (basis1.empty() && basis2.empty())
    return 0.0;
  if (basis1.size() != basis2.size())
    return 1e9;
  if (basis1[0].size() != basis2[0].size())
    return 1e9;
  std::vector<Vec> b1 = basis1, b2 = basis2;
  compute_rref(b1, tol);
  compute_rref(b2, tol);
  double sum_sq = 0.0;
  for (std::size_t i = 0; i < b1.size(); ++i) {
    axpy(b1[i], b2[i], -1.0);
    const double nn = nrm2(b1[i]);
    sum_sq += nn * nn;
  }
  return std::sqrt(sum_sq);
}

} // namespace orc

// ============================== C ABI ========================================
using namespace orc;

namespace {
thread_local std::string g_err;

// Plain-old-data option blocks, field for field the reference structs.
struct OrcCyclicOpts {
  int64_t k, ncv, max_restarts;
  double tol_abs, tol_rel;
  int32_t reorth, use_harmonic;
  double target_tau;
  int32_t nullity_known, rank_known;
  int64_t cg_maxit;
  double cg_rel_tol, cg_breakdown_tol;
  int32_t projector_passes, grad_max_iters, frob_samples;
  double prescreen_factor;
  int64_t stagnation_restarts;
  int32_t return_left_space, verbose;
  int32_t precond_jacobi; // built-in M^{-1} = diag(A A^T)^{-1}
  int32_t pad_;
};
struct OrcCyclicResult {
  int64_t n_right, n_left, restarts, krylov_steps_total;
  double frob_norm_est;
  int64_t converged_to_target, cg_solves, cg_maxit_hits, spmv, spmtv;
};
struct OrcLobpcgOpts {
  int64_t nev, block_size, max_iters;
  double tol_abs, tol_rel;
  int32_t soft_locking, gs_reorth;
  double restart, gs_tol;
  int32_t frob_samples, nullity_known, rank_known, verbose;
  int64_t max_tiny_dim, progress_every;
};
struct OrcLobpcgResult {
  int64_t n_right, iters, locked;
  double frob_norm_est, accept_threshold, best_unaccepted_residual;
  int64_t block_size, converged_to_target, spmv, spmtv;
};

struct OpHolder {
  std::unique_ptr<Op> op;
};

CyclicOpts to_opts(const OrcCyclicOpts &o) {
  CyclicOpts c;
  c.k = (std::size_t)o.k;
  c.ncv = (std::size_t)o.ncv;
  c.max_restarts = (std::size_t)o.max_restarts;
  c.tol_abs = o.tol_abs;
  c.tol_rel = o.tol_rel;
  c.reorth = (Reorth)o.reorth;
  c.use_harmonic = o.use_harmonic != 0;
  c.target_tau = o.target_tau;
  c.nullity_known = o.nullity_known;
  c.rank_known = o.rank_known;
  c.cg_maxit = (std::size_t)o.cg_maxit;
  c.cg_rel_tol = o.cg_rel_tol;
  c.cg_breakdown_tol = o.cg_breakdown_tol;
  c.projector_passes = o.projector_passes;
  c.grad_max_iters = o.grad_max_iters;
  c.frob_samples = o.frob_samples;
  c.prescreen_factor = o.prescreen_factor;
  c.stagnation_restarts = (std::size_t)o.stagnation_restarts;
  c.return_left_space = o.return_left_space != 0;
  c.verbose = o.verbose != 0;
  return c;
}

LobpcgOpts to_opts(const OrcLobpcgOpts &o) {
  LobpcgOpts c;
  c.nev = (std::size_t)o.nev;
  c.block_size = (std::size_t)o.block_size;
  c.max_iters = (std::size_t)o.max_iters;
  c.tol_abs = o.tol_abs;
  c.tol_rel = o.tol_rel;
  c.soft_locking = o.soft_locking != 0;
  c.gs_reorth = o.gs_reorth != 0;
  c.restart = o.restart;
  c.gs_tol = o.gs_tol;
  c.frob_samples = o.frob_samples;
  c.nullity_known = o.nullity_known;
  c.rank_known = o.rank_known;
  c.verbose = o.verbose != 0;
  c.max_tiny_dim = (std::size_t)o.max_tiny_dim;
  c.progress_every = (std::size_t)o.progress_every;
  return c;
}

// Factories persist across solver calls so that one stream can be shared by
// consecutive solves (examples/quickstart.cpp:39,77,84).
struct FactoryHolder {
  Factory f;
  explicit FactoryHolder(unsigned s) : f(s) {}
};

template <class F> int guard(F &&fn) {
  try {
    fn();
    return 0;
  } catch (const std::invalid_argument &e) {
    g_err = e.what();
    return 1;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 2;
  }
}

void put_basis(const std::vector<Vec> &B, double *out, int64_t cap,
               std::size_t len) {
  if (!out)
    return;
  for (std::size_t j = 0; j < B.size() && (int64_t)j < cap; ++j)
    std::memcpy(out + j * len, B[j].data(), len * sizeof(double));
}
} // namespace

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

// Loaders: a handle to an Eigen-style CSC, or null with orc_last_error set.
void *orc_load_ija(const char *path, int rows, int cols) {
  Csc *out = nullptr;
  guard([&] { out = new Csc(load_ija(path, rows, cols)); });
  return out;
}

void *orc_load_mtx(const char *path) {
  Csc *out = nullptr;
  guard([&] { out = new Csc(load_mtx(path)); });
  return out;
}

void *orc_triplets(int64_t rows, int64_t cols, int64_t count, const int32_t *I,
                   const int32_t *J, const double *V) {
  Csc *out = nullptr;
  guard([&] {
    std::vector<int> i(I, I + count), j(J, J + count);
    std::vector<double> v(V, V + count);
    out = new Csc(set_from_triplets(rows, cols, i, j, v));
    out->entries_read = count;
  });
  return out;
}

void orc_csc_info(const void *h, int64_t *rows, int64_t *cols, int64_t *nnz,
                  int64_t *entries_read) {
  const Csc &C = *static_cast<const Csc *>(h);
  *rows = C.rows;
  *cols = C.cols;
  *nnz = (int64_t)C.rowidx.size();
  *entries_read = C.entries_read;
}

void orc_csc_copy(const void *h, int64_t *colptr, int32_t *rowidx, double *val) {
  const Csc &C = *static_cast<const Csc *>(h);
  std::copy(C.colptr.begin(), C.colptr.end(), colptr);
  std::copy(C.rowidx.begin(), C.rowidx.end(), rowidx);
  std::copy(C.val.begin(), C.val.end(), val);
}

void orc_csc_free(void *h) { delete static_cast<Csc *>(h); }

void *orc_op_csr(int64_t m, int64_t n, const int32_t *rowptr,
                 const int32_t *col, const double *val) {
  auto *h = new OpHolder;
  h->op = std::make_unique<CsrOp>((std::size_t)m, (std::size_t)n, rowptr, col, val);
  return h;
}
void *orc_op_shiftdiff(int64_t n, int64_t N, double c) {
  auto *h = new OpHolder;
  h->op = std::make_unique<ShiftDiffOp>((std::size_t)n, (std::size_t)N, c);
  return h;
}
void orc_op_free(void *h) { delete static_cast<OpHolder *>(h); }
int64_t orc_op_rows(void *h) { return (int64_t) static_cast<OpHolder *>(h)->op->rows(); }
int64_t orc_op_cols(void *h) { return (int64_t) static_cast<OpHolder *>(h)->op->cols(); }

// CSC twin export: colptr (n+1), rowidx (nnz), values (nnz), perm (nnz).
int orc_op_csc(void *h, int32_t *cp, int32_t *ri, double *vt, int32_t *perm) {
  auto *op = dynamic_cast<CsrOp *>(static_cast<OpHolder *>(h)->op.get());
  if (!op) {
    g_err = "not a CSR operator";
    return 1;
  }
  std::copy(op->cp.begin(), op->cp.end(), cp);
  std::copy(op->ri.begin(), op->ri.end(), ri);
  std::copy(op->vt.begin(), op->vt.end(), vt);
  std::copy(op->perm.begin(), op->perm.end(), perm);
  return 0;
}

void orc_mv(void *h, const double *x, double *y) {
  const Op &A = *static_cast<OpHolder *>(h)->op;
  Vec xv(x, x + A.cols()), yv(A.rows(), 0.0);
  A.mv(xv, yv);
  std::memcpy(y, yv.data(), yv.size() * sizeof(double));
}
void orc_mtv(void *h, const double *x, double *y) {
  const Op &A = *static_cast<OpHolder *>(h)->op;
  Vec xv(x, x + A.rows()), yv(A.cols(), 0.0);
  A.mtv(xv, yv);
  std::memcpy(y, yv.data(), yv.size() * sizeof(double));
}

void *orc_factory(uint32_t seed) { return new FactoryHolder(seed); }
void orc_factory_free(void *f) { delete static_cast<FactoryHolder *>(f); }
void orc_randn(void *f, int64_t n, double *out) {
  Vec v = static_cast<FactoryHolder *>(f)->f.randn((std::size_t)n);
  std::memcpy(out, v.data(), (std::size_t)n * sizeof(double));
}
uint32_t orc_mt19937_first(uint32_t seed) {
  std::mt19937 g(seed);
  return (uint32_t)g();
}

int orc_cyclic(void *h, void *fh, const OrcCyclicOpts *o, const double *y0,
               double *right_out, int64_t right_cap, double *left_out,
               int64_t left_cap, OrcCyclicResult *res) {
  return guard([&] {
    const Op &A = *static_cast<OpHolder *>(h)->op;
    Factory &vf = static_cast<FactoryHolder *>(fh)->f;
    CyclicOpts c = to_opts(*o);
    if (y0)
      c.initial_right_space.emplace_back(y0, y0 + A.cols());
    Vec diag;
    if (o->precond_jacobi) {
      auto *csr = dynamic_cast<const CsrOp *>(&A);
      if (!csr)
        throw std::invalid_argument("jacobi preconditioner needs CSR");
      diag.assign(csr->m, 0.0);
      for (std::size_t i = 0; i < csr->m; ++i) {
        double s = 0.0;
        for (int32_t p = csr->rp[i]; p < csr->rp[i + 1]; ++p)
          s += csr->va[(std::size_t)p] * csr->va[(std::size_t)p];
        diag[i] = s;
      }
      c.precond_AAt = [&diag](const Vec &r, Vec &y) {
        for (std::size_t i = 0; i < r.size(); ++i)
          y[i] = diag[i] > 0.0 ? r[i] / diag[i] : r[i];
      };
    }
    CyclicResult R = cyclic_lanczos(A, vf, c);
    put_basis(R.right_null, right_out, right_cap, A.cols());
    put_basis(R.left_null, left_out, left_cap, A.rows());
    res->n_right = (int64_t)R.right_null.size();
    res->n_left = (int64_t)R.left_null.size();
    res->restarts = (int64_t)R.restarts;
    res->krylov_steps_total = (int64_t)R.krylov_steps_total;
    res->frob_norm_est = R.frob_norm_est;
    res->converged_to_target = R.converged_to_target;
    res->cg_solves = (int64_t)R.cg_solves;
    res->cg_maxit_hits = (int64_t)R.cg_maxit_hits;
    res->spmv = (int64_t)R.spmv;
    res->spmtv = (int64_t)R.spmtv;
  });
}

int orc_lobpcg(void *h, void *fh, const OrcLobpcgOpts *o, double *right_out,
               int64_t right_cap, double *theta_out, double *resid_out,
               OrcLobpcgResult *res) {
  return guard([&] {
    const Op &A = *static_cast<OpHolder *>(h)->op;
    Factory &vf = static_cast<FactoryHolder *>(fh)->f;
    LobpcgResult R = lobpcg(A, vf, to_opts(*o));
    put_basis(R.right_null, right_out, right_cap, A.cols());
    for (std::size_t i = 0; i < R.right_null.size() && (int64_t)i < right_cap; ++i) {
      if (theta_out)
        theta_out[i] = R.theta[i];
      if (resid_out)
        resid_out[i] = R.residuals_Av[i];
    }
    res->n_right = (int64_t)R.right_null.size();
    res->iters = (int64_t)R.iters;
    res->locked = (int64_t)R.locked;
    res->frob_norm_est = R.frob_norm_est;
    res->accept_threshold = R.accept_threshold;
    res->best_unaccepted_residual = R.best_unaccepted_residual;
    res->block_size = (int64_t)R.block_size;
    res->converged_to_target = R.converged_to_target;
    res->spmv = (int64_t)R.spmv;
    res->spmtv = (int64_t)R.spmtv;
  });
}

// Dense kernels, column-major.
void orc_svd(int32_t r, int32_t c, const double *A, double *sv, double *V) {
  Mat M(r, c);
  std::copy(A, A + (std::size_t)r * c, M.a.begin());
  std::vector<double> s;
  Mat Vm;
  svd_jacobi(M, s, Vm);
  std::copy(s.begin(), s.end(), sv);
  std::copy(Vm.a.begin(), Vm.a.end(), V);
}
void orc_sym_eig(int32_t n, const double *A, double *w, double *Q) {
  Mat M(n, n);
  std::copy(A, A + (std::size_t)n * n, M.a.begin());
  std::vector<double> ww;
  Mat Qm;
  sym_eig(M, ww, Qm);
  std::copy(ww.begin(), ww.end(), w);
  std::copy(Qm.a.begin(), Qm.a.end(), Q);
}
int orc_cholesky(int32_t n, const double *A, double *L) {
  Mat M(n, n);
  std::copy(A, A + (std::size_t)n * n, M.a.begin());
  Mat Lm;
  if (!cholesky(M, Lm))
    return 1;
  std::copy(Lm.a.begin(), Lm.a.end(), L);
  return 0;
}

// Gram-Schmidt utilities on k column vectors of length len (column-major).
int orc_mgs(int64_t len, int64_t k, double *V, double tol, int32_t reorth) {
  std::vector<Vec> vs;
  for (int64_t j = 0; j < k; ++j)
    vs.emplace_back(V + j * len, V + (j + 1) * len);
  const bool ok = modified_gram_schmidt(vs, tol, reorth != 0);
  for (int64_t j = 0; j < k; ++j)
    std::copy(vs[(std::size_t)j].begin(), vs[(std::size_t)j].end(), V + j * len);
  return ok ? 1 : 0;
}
double orc_orth_against(int64_t len, int64_t k, const double *B, double *v,
                        double tol, int32_t reorth) {
  std::vector<Vec> bs;
  for (int64_t j = 0; j < k; ++j)
    bs.emplace_back(B + j * len, B + (j + 1) * len);
  Vec vv(v, v + len);
  const double r = orthogonalize_against_basis(vv, bs, tol, reorth != 0);
  std::copy(vv.begin(), vv.end(), v);
  return r;
}

double orc_subspace_diff(int64_t len, int64_t k1, const double *B1, int64_t k2,
                         const double *B2, double tol) {
  std::vector<Vec> b1, b2;
  for (int64_t j = 0; j < k1; ++j)
    b1.emplace_back(B1 + j * len, B1 + (j + 1) * len);
  for (int64_t j = 0; j < k2; ++j)
    b2.emplace_back(B2 + j * len, B2 + (j + 1) * len);
  return subspace_diff(b1, b2, tol);
}

} // extern "C"
