// Device context, buffers and the device CSR/CSC build of the KraNS-MI
// engine. The CSC twin is produced by a stable counting sort over the rows of
// the canonical CSR, so within every column the rows ascend and perm[q] is
// the CSR position of CSC entry q: the same structure as Eigen's
// A_.transpose() of EigenSparseMatrixAdapter.hpp:20-24.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "engine.hpp"

namespace kh {

void hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess)
    return;
  std::string msg = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory)
    throw DeviceOOM(msg);
  throw DeviceError(msg);
}

Context::Context(int dev, std::size_t part_doubles) : device(dev) {
  KH_CHECK(hipSetDevice(dev));
  KH_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  KH_CHECK(hipMalloc((void **)&dscal, sizeof(double) * kScalars));
  KH_CHECK(hipMemset(dscal, 0, sizeof(double) * kScalars));
  KH_CHECK(hipHostMalloc((void **)&hscal, sizeof(double) * kScalars, hipHostMallocDefault));
  KH_CHECK(hipHostMalloc((void **)&hstage, kStage, hipHostMallocDefault));
  part_cap = part_doubles;  // 16 Mi doubles = 128 MiB of reduction partials by default
  KH_CHECK(hipMalloc((void **)&dpart, sizeof(double) * part_cap));
  hipDeviceProp_t prop;
  KH_CHECK(hipGetDeviceProperties(&prop, dev));
  cu_count = prop.multiProcessorCount;
  lds_max = (int)prop.sharedMemPerBlock;
}

hipEvent_t Context::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  KH_CHECK(hipEventCreate(&e));
  return e;
}

void Context::trace_reset() {
  for (auto &r : trace) {
    event_pool.push_back(r.a);
    event_pool.push_back(r.b);
  }
  trace.clear();
}

Context::~Context() {
  if (stream)
    (void)hipStreamSynchronize(stream);
  trace_reset();
  for (hipEvent_t e : event_pool)
    (void)hipEventDestroy(e);
  if (dpart)
    (void)hipFree(dpart);
  if (dscal)
    (void)hipFree(dscal);
  if (hscal)
    (void)hipHostFree(hscal);
  if (hstage)
    (void)hipHostFree(hstage);
  if (stream)
    (void)hipStreamDestroy(stream);
}

void Context::sync() {
  KH_CHECK(hipStreamSynchronize(stream));
  stage_off = 0;
}

void Context::upload(void *dst, const void *src, std::size_t bytes) {
  if (bytes == 0)
    return;
  if (bytes > kStage)
    throw std::length_error("upload larger than the staging area");
  if (stage_off + bytes > kStage)
    sync();  // every staged slice has been consumed
  char *h = hstage + stage_off;
  std::memcpy(h, src, bytes);
  KH_CHECK(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, stream));
  stage_off += (bytes + 255) & ~std::size_t(255);
}

void CapturedWork::reset() {
  if (exec)
    (void)hipGraphExecDestroy(exec);
  exec = nullptr;
  key = -1;
}

void CapturedWork::begin(Context &ctx) {
  KH_CHECK(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
}

void CapturedWork::end(Context &ctx, int64_t k, bool ok) {
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(ctx.stream, &g);
  if (!ok) {  // an exception left the queued code: drop what was recorded
    if (g)
      (void)hipGraphDestroy(g);
    return;
  }
  KH_CHECK(e);
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  KH_CHECK(e);
  reset();
  exec = x;
  key = k;
  ++captures;
}

void CapturedWork::launch(Context &ctx) {
  if (!exec)
    throw std::logic_error("CapturedWork::launch without a captured graph");
  KH_CHECK(hipGraphLaunch(exec, ctx.stream));
  ++replays;
}

void Context::fetch_scalars(int i0, int n) {
  if (i0 < 0 || n < 0 || i0 + n > kScalars)
    throw std::out_of_range("scalar slab index out of range");
  KH_CHECK(hipMemcpyAsync(hscal + i0, dscal + i0, sizeof(double) * (std::size_t)n,
                          hipMemcpyDeviceToHost, stream));
  KH_CHECK(hipStreamSynchronize(stream));
}

double *Context::partials(std::size_t need) {
  if (need > part_capacity())
    throw std::length_error("reduction partial buffer too small for this batch");
  return part_ext ? part_ext : dpart;
}

DBuf::DBuf(std::size_t b) : bytes(b) {
  if (b == 0)
    return;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    p = nullptr;
    bytes = 0;
    (void)hipGetLastError();
    if (e == hipErrorOutOfMemory)
      throw DeviceOOM("hipMalloc(" + std::to_string(b) + " bytes): out of device memory");
    hip_check(e, "hipMalloc");
  }
}

DBuf::~DBuf() {
  if (p)
    (void)hipFree(p);
}

DBuf &DBuf::operator=(DBuf &&o) noexcept {
  if (this != &o) {
    if (p)
      (void)hipFree(p);
    p = o.p;
    bytes = o.bytes;
    o.p = nullptr;
    o.bytes = 0;
  }
  return *this;
}

static int pick_lanes(double avg, int maxlen) {
  // Short rows: one lane per row keeps the reference's summation order.
  if (avg <= 12.0 && maxlen <= 256)
    return 1;
  if (avg <= 48.0)
    return 4;
  if (avg <= 192.0)
    return 16;
  return 64;
}

template <class T> static DBuf upload(Context &ctx, const T *h, std::size_t count) {
  DBuf d(count * sizeof(T) + (count == 0 ? sizeof(T) : 0));
  if (count)
    KH_CHECK(hipMemcpyAsync(d.p, h, count * sizeof(T), hipMemcpyHostToDevice, ctx.stream));
  return d;
}

// Window tables of the LDS-staged gathers for the rows of one direction
// (ptr/idx: CSR rows of A or of the CSC twin) when every row holds R entries
// (R in 1..4 or 8) and there are at least 65,536 rows: the largest block of
// 512 x 2^j rows (at most 32 / R rows per thread, 8 at most) that still
// gives 256 blocks, and half of it; a block size is kept when its largest
// window (from the lowest entry index rounded down to a multiple of 4 to
// the highest, rounded up to a multiple of 4) holds at most 40,960 elements
// (160 KB of FP32).
static void build_spans(Context &ctx, DeviceCsr &A, int d, int64_t rows, int R,
                        const int32_t *ptr, const int32_t *idx) {
  if (!(R >= 1 && R <= 4) && R != 8)
    return;
  if (rows < 65536)
    return;
  int rows_pt = std::min(8, 32 / R);
  int rb = rows_pt * 512;
  while (rb > 512 && (rows + rb - 1) / rb < 256)
    rb /= 2;
  for (int k = 0; k < 2 && rb >= 512; ++k, rb /= 2) {
    const int64_t nb = (rows + rb - 1) / rb;
    std::vector<int32_t> st((std::size_t)nb);
    int64_t len = 0;
    for (int64_t b = 0; b < nb; ++b) {
      const int64_t i0 = b * rb, i1 = std::min<int64_t>(rows, i0 + rb);
      int32_t lo = INT32_MAX, hi = -1;
      for (int64_t p = ptr[i0]; p < ptr[i1]; ++p) {
        lo = std::min(lo, idx[p]);
        hi = std::max(hi, idx[p]);
      }
      if (hi < 0)
        lo = hi = 0;
      const int32_t s0 = lo & ~3;
      st[(std::size_t)b] = s0;
      len = std::max<int64_t>(len, ((int64_t)hi + 1 - s0 + 3) & ~(int64_t)3);
    }
    if (len > 40960)
      continue;
    A.span_rb[d][k] = rb;
    A.span_len[d][k] = (int)len;
    A.span[d][k] = upload(ctx, st.data(), st.size());
    ctx.sync();  // st dies with this iteration
  }
}

int validate_csr(int64_t m, int64_t n, const int32_t *rowptr, const int32_t *col,
                 const double *val) {
  if (m < 0 || n < 0)
    throw std::invalid_argument("csr_create: negative dimension");
  if (m > INT32_MAX || n > INT32_MAX)
    throw std::invalid_argument("csr_create: dimension exceeds int32 indexing");
  if (!rowptr)
    throw std::invalid_argument("csr_create: null rowptr");
  if (rowptr[0] != 0)
    throw std::invalid_argument("csr_create: rowptr[0] != 0");
  const int64_t nnz = rowptr[m];
  if (nnz < 0)
    throw std::invalid_argument("csr_create: negative nnz");
  if (nnz > 0 && (!col || !val))
    throw std::invalid_argument("csr_create: null col/val");
  int max_row = 0;
  for (int64_t i = 0; i < m; ++i) {
    const int32_t a = rowptr[i], b = rowptr[i + 1];
    if (b < a || b > nnz)
      throw std::invalid_argument("csr_create: rowptr not monotone within [0, nnz]");
    max_row = std::max(max_row, b - a);
    for (int32_t p = a; p < b; ++p) {
      if (col[p] < 0 || col[p] >= n)
        throw std::invalid_argument("csr_create: column index out of range");
      if (p > a && col[p] <= col[p - 1])
        throw std::invalid_argument(
            "csr_create: columns must be strictly ascending within a row (canonical CSR)");
    }
  }
  return max_row;
}

DeviceCsr *csr_create(Context &ctx, int64_t m, int64_t n, const int32_t *rowptr,
                      const int32_t *col, const double *val) {
  const int max_row = validate_csr(m, n, rowptr, col, val);
  const int64_t nnz = rowptr[m];
  auto *A = new DeviceCsr;
  try {
    A->m = m;
    A->n = n;
    A->nnz = nnz;
    // CSC twin: stable counting sort over rows
    std::vector<int32_t> cp((std::size_t)n + 1, 0), ri((std::size_t)nnz), perm((std::size_t)nnz);
    std::vector<double> vt((std::size_t)nnz);
    for (int64_t p = 0; p < nnz; ++p)
      cp[(std::size_t)col[p] + 1]++;
    int max_col = 0;
    for (int64_t j = 0; j < n; ++j) {
      max_col = std::max(max_col, cp[(std::size_t)j + 1]);
      cp[(std::size_t)j + 1] += cp[(std::size_t)j];
    }
    std::vector<int32_t> next(cp.begin(), cp.end() - 1);
    double fro2 = 0.0;
    std::vector<double> dAAt((std::size_t)m, 0.0);
    for (int64_t i = 0; i < m; ++i) {
      double rs = 0.0;
      for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
        const int32_t q = next[(std::size_t)col[p]]++;
        ri[(std::size_t)q] = (int32_t)i;
        vt[(std::size_t)q] = val[p];
        perm[(std::size_t)q] = p;
        rs += val[p] * val[p];
      }
      dAAt[(std::size_t)i] = rs;
      fro2 += rs;
    }
    A->fro = std::sqrt(fro2);
    A->max_row = max_row;
    A->max_col = max_col;
    A->lpr_rows = pick_lanes(m ? (double)nnz / (double)m : 0.0, max_row);
    A->lpr_cols = pick_lanes(n ? (double)nnz / (double)n : 0.0, max_col);
    // no row exceeds max_row, so nnz == m max_row means every row has max_row
    A->uniform_rows = (m > 0 && nnz == (int64_t)m * max_row) ? max_row : 0;
    A->uniform_cols = (n > 0 && nnz == (int64_t)n * max_col) ? max_col : 0;
    std::vector<float> va32((std::size_t)nnz), vt32((std::size_t)nnz);
    for (int64_t p = 0; p < nnz; ++p) {
      va32[(std::size_t)p] = (float)val[p];
      vt32[(std::size_t)p] = (float)vt[(std::size_t)p];
    }
    A->rp = upload(ctx, rowptr, (std::size_t)m + 1);
    A->ci = upload(ctx, col, (std::size_t)nnz);
    A->va = upload(ctx, val, (std::size_t)nnz);
    A->va32 = upload(ctx, va32.data(), (std::size_t)nnz);
    A->cp = upload(ctx, cp.data(), (std::size_t)n + 1);
    A->ri = upload(ctx, ri.data(), (std::size_t)nnz);
    A->vt = upload(ctx, vt.data(), (std::size_t)nnz);
    A->vt32 = upload(ctx, vt32.data(), (std::size_t)nnz);
    A->perm = upload(ctx, perm.data(), (std::size_t)nnz);
    A->diag_AAt = upload(ctx, dAAt.data(), (std::size_t)m);
    A->diag_AAt_host = std::move(dAAt);
    build_spans(ctx, *A, 0, m, A->uniform_rows, rowptr, col);
    build_spans(ctx, *A, 1, n, A->uniform_cols, cp.data(), ri.data());
    ctx.sync();  // host staging vectors die at scope exit
  } catch (...) {
    delete A;
    throw;
  }
  return A;
}

}  // namespace kh
