// krans_nullspace on the HIP backend: the command line of the reference's
// apps/krans_nullspace.cpp (same flags, defaults, exit codes and outputs —
// N.csv, optional N_left.csv, report.json), with the matrix loaded into HBM
// (kh_load_ija / kh_load_matrix_market) and solved by the device-resident
// cyclic-Lanczos or LOBPCG driver. Added flags: --precision fp64|mixed and
// --backend hip (the only backend of this build). The direct dense solvers
// and the third-party wrappers are not part of this backend: asking for one
// is "Unknown solver" (exit 2), as for a solver compiled out of the reference.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "krans/Solvers/CyclicLanczosSolver.hpp"
#include "krans/Solvers/LOBPCGSolver.hpp"
#include "krans/adapters/hip/HipAdapters.hpp"

namespace {

using Basis = std::vector<std::unique_ptr<krans::VectorInterface>>;

std::string to_lower(std::string s) {
  for (auto &c : s)
    c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

void usage(const char *argv0) {
  std::cerr
      << "KraNS Nullspace CLI (HIP backend)\n"
      << "---------------------------------\n"
      << "Usage:\n"
      << "  " << argv0 << " --solver [cyclic|lobpcg]\n"
      << "           --input <matrix.(mtx|ija)>\n"
      << "           [--out N.csv] [--report report.json]\n"
      << "           [--tol <rel_tol>]           # accepted for compatibility (direct solvers)\n"
      << "           [--seed <int>]              # RNG seed for iterative solvers (default 1; "
         "0 = random)\n"
      << "           [--precision fp64|mixed]    # device arithmetic (default fp64)\n"
      << "           [--backend hip]\n"
      << "\n"
      << "Cyclic solver options (optional):\n"
      << "  --k <int>              # number of right null vectors (default 1)\n"
      << "  --ncv <int>            # subspace size (default 48)\n"
      << "  --max-restarts <int>   # outer restarts (default 80)\n"
      << "  --tol-abs <float>      # accept if ||A v|| <= max(tol_abs, tol_rel*||A||_F), "
         "default 1e-12\n"
      << "  --tol-rel <float>      # relative part in acceptance, default 1e-10\n"
      << "  --proj-passes <int>    # projector polish passes (1 or 2), default 1\n"
      << "  --cg-maxit <int>       # CG iters for normal eqs, default 80\n"
      << "  --cg-rel <float>       # CG relative tol, default 1e-10\n"
      << "  --prescreen <float>    # skip polishing if ||A v|| > f*||A||_F, default 0.8\n"
      << "  --stagnation <int>     # stop after this many restarts without progress (0 = "
         "off), default 30\n"
      << "  --verbose              # print progress\n"
      << "  --left                 # also export left nullspace to N_left.csv\n"
      << "\n"
      << "LOBPCG options (optional):\n"
      << "  --k <int>                  # number of right null vectors (default 1)\n"
      << "  --lobpcg-block <int>       # block size (default min(16,k))\n"
      << "  --lobpcg-max-it <int>      # max outer iterations (default 1000)\n"
      << "  --lobpcg-restart <float>   # restart fraction in (0.1..1], default 0.9\n"
      << "  --lobpcg-soft-lock <0|1>   # soft locking on/off (default 1)\n"
      << "  --tol-abs <float>          # default 1e-12\n"
      << "  --tol-rel <float>          # default 1e-10\n"
      << "  --verbose                  # print progress\n"
      << std::endl;
}

// write_csv_basis (include/krans/Utils/ExportIO.hpp:17-41): fixed notation,
// 17 digits after the point, one row per vector entry, no header.
bool write_csv(const std::string &path, const Basis &B) {
  std::ofstream out(path);
  if (!out)
    return false;
  out.setf(std::ios::fixed, std::ios::floatfield);
  out << std::setprecision(17);
  if (B.empty())
    return true;
  std::vector<std::vector<double>> cols;
  for (const auto &v : B)
    cols.push_back(static_cast<const krans::hip::HipVector &>(*v).to_host());
  const std::size_t rows = cols.front().size();
  for (std::size_t i = 0; i < rows; ++i) {
    for (std::size_t j = 0; j < cols.size(); ++j) {
      out << cols[j][i];
      if (j + 1 < cols.size())
        out << ',';
    }
    out << '\n';
  }
  return static_cast<bool>(out);
}

// diagnose_basis + write_json_report (ExportIO.hpp:44-105)
bool write_report(const std::string &path, std::size_t m, std::size_t n, std::size_t rank,
                  std::size_t nullity, const std::string &solver, double tol_rel, double load_ms,
                  double solve_ms, const krans::MatrixInterface &A, const Basis &B,
                  kh_context *ctx) {
  std::vector<double> res(B.size(), 0.0);
  krans::hip::HipVector y(ctx, A.rows());
  for (std::size_t j = 0; j < B.size(); ++j) {
    A.multiply(*B[j], y);
    res[j] = y.norm();
  }
  double mx = 0.0;
  for (std::size_t i = 0; i < B.size(); ++i)
    for (std::size_t j = i + 1; j < B.size(); ++j)
      mx = std::max(mx, std::fabs(B[i]->dot(*B[j])));
  std::ofstream out(path);
  if (!out)
    return false;
  out << std::setprecision(17);
  out << "{\n";
  out << "  \"solver\": \"" << solver << "\",\n";
  out << "  \"rows\": " << m << ",\n";
  out << "  \"cols\": " << n << ",\n";
  out << "  \"rank\": " << rank << ",\n";
  out << "  \"nullity\": " << nullity << ",\n";
  out << "  \"tol_rel\": " << tol_rel << ",\n";
  out << "  \"timings_ms\": { \"load\": " << load_ms << ", \"solve\": " << solve_ms << " },\n";
  out << "  \"residuals\": [";
  for (std::size_t i = 0; i < res.size(); ++i) {
    if (i)
      out << ", ";
    out << res[i];
  }
  out << "],\n";
  out << "  \"max_offdiag_abs_dot\": " << mx << "\n";
  out << "}\n";
  return static_cast<bool>(out);
}

struct HostCsr {
  kh_host_csr *h = nullptr;
  ~HostCsr() { kh_host_csr_free(h); }
};

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
      .count();
}

}  // namespace

int main(int argc, char **argv) {
  std::string solver_name = "svd";
  std::string input_path;
  std::string out_csv = "N.csv";
  std::string report_json = "report.json";
  std::string precision = "fp64";
  double tol_rel_direct = 1e-12;
  int cyc_k = 1, cyc_ncv = 48, cyc_max_restarts = 80, cyc_proj_passes = 1, cyc_cg_maxit = 80;
  double cyc_tol_abs = 1e-12, cyc_tol_rel = 1e-10, cyc_cg_rel = 1e-10, cyc_prescreen = 0.8;
  int cyc_stagnation = 30;
  bool cyc_verbose = false, cyc_left = false;
  unsigned seed = 1;
  int lob_block = 0, lob_max_it = 1000, lob_soft_lock = 1;
  double lob_restart = 0.9;

  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto value = [&](const char *&v) {
      if (i + 1 >= argc)
        return false;
      v = argv[++i];
      return true;
    };
    const char *v = nullptr;
    try {
      if (a == "--help" || a == "-h") {
        usage(argv[0]);
        return 0;
      } else if (a == "--verbose") {
        cyc_verbose = true;
      } else if (a == "--left") {
        cyc_left = true;
      } else if (a == "--solver" || a == "--input" || a == "--out" || a == "--report" ||
                 a == "--tol" || a == "--k" || a == "--tol-abs" || a == "--tol-rel" ||
                 a == "--ncv" || a == "--max-restarts" || a == "--proj-passes" ||
                 a == "--cg-maxit" || a == "--cg-rel" || a == "--prescreen" ||
                 a == "--stagnation" || a == "--seed" || a == "--lobpcg-block" ||
                 a == "--lobpcg-max-it" || a == "--lobpcg-restart" ||
                 a == "--lobpcg-soft-lock" || a == "--precision" || a == "--backend") {
        if (!value(v)) {
          usage(argv[0]);
          return 2;
        }
        if (a == "--solver")
          solver_name = to_lower(v);
        else if (a == "--input")
          input_path = v;
        else if (a == "--out")
          out_csv = v;
        else if (a == "--report")
          report_json = v;
        else if (a == "--tol")
          tol_rel_direct = std::stod(v);
        else if (a == "--k")
          cyc_k = std::stoi(v);
        else if (a == "--tol-abs")
          cyc_tol_abs = std::stod(v);
        else if (a == "--tol-rel")
          cyc_tol_rel = std::stod(v);
        else if (a == "--ncv")
          cyc_ncv = std::stoi(v);
        else if (a == "--max-restarts")
          cyc_max_restarts = std::stoi(v);
        else if (a == "--proj-passes")
          cyc_proj_passes = std::stoi(v);
        else if (a == "--cg-maxit")
          cyc_cg_maxit = std::stoi(v);
        else if (a == "--cg-rel")
          cyc_cg_rel = std::stod(v);
        else if (a == "--prescreen")
          cyc_prescreen = std::stod(v);
        else if (a == "--stagnation")
          cyc_stagnation = std::stoi(v);
        else if (a == "--seed")
          seed = static_cast<unsigned>(std::stoul(v));
        else if (a == "--lobpcg-block")
          lob_block = std::stoi(v);
        else if (a == "--lobpcg-max-it")
          lob_max_it = std::stoi(v);
        else if (a == "--lobpcg-restart")
          lob_restart = std::stod(v);
        else if (a == "--lobpcg-soft-lock")
          lob_soft_lock = std::stoi(v);
        else if (a == "--precision")
          precision = to_lower(v);
        else if (a == "--backend" && to_lower(v) != "hip") {
          std::cerr << "Unsupported backend: " << v << " (this build provides hip)\n";
          return 2;
        }
      } else {
        std::cerr << "Unknown argument: " << a << "\n";
        usage(argv[0]);
        return 2;
      }
    } catch (const std::exception &) {  // std::stoi / std::stod on a bad number
      std::cerr << "Invalid value for " << a << "\n";
      usage(argv[0]);
      return 2;
    }
  }
  if (input_path.empty()) {
    std::cerr << "Error: --input is required.\n";
    usage(argv[0]);
    return 2;
  }
  if (solver_name != "cyclic" && solver_name != "lobpcg") {
    std::cerr << "Unknown solver: " << solver_name << "\n";
    usage(argv[0]);
    return 2;
  }
  if (precision != "fp64" && precision != "mixed") {
    std::cerr << "Unknown precision: " << precision << " (fp64 or mixed)\n";
    return 2;
  }
  (void)tol_rel_direct;

  // ---- load ---------------------------------------------------------------
  auto t0 = std::chrono::steady_clock::now();
  const auto dot = input_path.find_last_of('.');
  const std::string ext = dot == std::string::npos ? "" : to_lower(input_path.substr(dot + 1));
  HostCsr H;
  int rc;
  if (ext == "mtx") {
    rc = kh_load_matrix_market(input_path.c_str(), &H.h);
  } else if (ext == "ija") {
    rc = kh_load_ija(input_path.c_str(), -1, -1, &H.h);
  } else {
    std::cerr << "Unsupported input format: expected .mtx or .ija\n";
    return 2;
  }
  if (rc != KH_OK) {
    std::cerr << "Matrix load error: " << kh_last_error() << "\n";
    return 2;
  }
  int64_t m64 = 0, n64 = 0, nnz64 = 0;
  const int32_t *rp = nullptr, *ci = nullptr;
  const double *va = nullptr;
  kh_host_csr_info(H.h, &m64, &n64, &nnz64, nullptr, nullptr);
  kh_host_csr_data(H.h, &rp, &ci, &va);

  try {
    krans::hip::HipContext ctx(0);
    krans::hip::HipCsrMatrixAdapter A(ctx.get(), (std::size_t)m64, (std::size_t)n64, rp, ci, va);
    const double t_load_ms = ms_since(t0);
    const std::size_t m = A.rows(), n = A.cols();
    krans::hip::HipVectorFactory vf(ctx.get(), n, seed,
                                    precision == "mixed" ? krans::hip::Precision::Mixed
                                                         : krans::hip::Precision::FP64);
    // ---- solve ------------------------------------------------------------
    t0 = std::chrono::steady_clock::now();
    Basis right;
    bool wrote_left = false;
    if (solver_name == "cyclic") {
      krans::solvers::CyclicLanczosOptions opt;
      opt.k = cyc_k > 0 ? (std::size_t)cyc_k : 1;
      opt.ncv = cyc_ncv > 0 ? (std::size_t)cyc_ncv : 48;
      opt.max_restarts = cyc_max_restarts >= 0 ? (std::size_t)cyc_max_restarts : 80;
      opt.tol_abs = cyc_tol_abs;
      opt.tol_rel = cyc_tol_rel;
      opt.projector_passes = std::max(1, cyc_proj_passes);
      opt.cg_maxit = cyc_cg_maxit >= 0 ? (std::size_t)cyc_cg_maxit : 80;
      opt.cg_rel_tol = cyc_cg_rel;
      opt.prescreen_factor = cyc_prescreen;
      opt.stagnation_restarts = (std::size_t)std::max(0, cyc_stagnation);
      opt.reorth = krans::solvers::ReorthMode::Double;
      opt.use_harmonic = true;
      opt.target_tau = 0.0;
      opt.return_left_space = cyc_left;
      opt.verbose = cyc_verbose;
      auto R = krans::solvers::find_nullspace_cyclic_lanczos_ex(A, vf, opt);
      right = std::move(R.right_null);
      if (cyc_left && !R.left_null.empty()) {
        if (!write_csv("N_left.csv", R.left_null))
          std::cerr << "Warning: failed to write N_left.csv\n";
        else
          wrote_left = true;
      }
    } else {
      krans::solvers::LOBPCGOptions opt;
      opt.nev = cyc_k > 0 ? (std::size_t)cyc_k : 1;
      opt.block_size =
          lob_block > 0 ? (std::size_t)lob_block : std::min<std::size_t>(16, opt.nev);
      opt.max_iters = lob_max_it > 0 ? (std::size_t)lob_max_it : 1000;
      opt.restart = (lob_restart > 0.1 && lob_restart <= 1.0) ? lob_restart : 0.9;
      opt.soft_locking = lob_soft_lock != 0;
      opt.tol_abs = cyc_tol_abs;
      opt.tol_rel = cyc_tol_rel;
      opt.gs_tol = 1e-14;
      opt.gs_reorth = true;
      opt.frob_samples = 4;
      opt.verbose = cyc_verbose;
      right = krans::solvers::find_nullspace_lobpcg(A, vf, opt);
    }
    const double t_solve_ms = ms_since(t0);
    const std::size_t nullity = right.size();
    const std::size_t rank_est = n >= nullity ? n - nullity : 0;
    if (!write_csv(out_csv, right)) {
      std::cerr << "Failed to write " << out_csv << "\n";
      return 3;
    }
    if (!write_report(report_json, m, n, rank_est, nullity, solver_name, cyc_tol_rel, t_load_ms,
                      t_solve_ms, A, right, ctx.get())) {
      std::cerr << "Failed to write " << report_json << "\n";
      return 4;
    }
    std::cout << "OK [" << solver_name << "]  m=" << m << " n=" << n << " rank≈" << rank_est
              << " nullity=" << nullity << " | load=" << t_load_ms
              << " ms, solve=" << t_solve_ms << " ms\n"
              << "Wrote: " << out_csv << " and " << report_json
              << (wrote_left ? " (plus N_left.csv)" : "") << "\n";
  } catch (const std::exception &e) {
    std::cerr << "Solve failed: " << e.what() << "\n";
    return 5;
  }
  return 0;
}
