// Packed cyclic solves: several small null-space requests whose Golub-Kahan
// cycles run in shared (segmented) launches, the batched form of the serving
// loop (see packed.cpp).
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "solver.hpp"

namespace kh {

// One request of a pack. Inputs: a canonical host CSR (validated by
// csr_create), the options and the seed of its random stream. Outputs are
// written by the request's own thread, which then calls on_done.
struct PackRequest {
  int64_t m = 0, n = 0;
  const int32_t *rowptr = nullptr, *col = nullptr;
  const double *val = nullptr;
  CyclicParams P;
  unsigned seed = 1;
  int status = 0;  // 0 ok, 1 invalid argument, 2 runtime error, 3 device OOM
  std::string error;
  std::vector<double> basis;  // n x count, column-major
  int64_t count = 0;
  CyclicStats stats;
  std::function<void()> on_done;
};

// Whether a request can join a pack: host CSR solves with CGS2
// reorthogonalisation; packs share the Krylov precision and ncv_eff.
bool packable(const CyclicParams &P);
int64_t pack_ncv_eff(const CyclicParams &P);

// Device bytes a pack of these requests holds beyond the members' own
// matrices and solver state: the block-diagonal pack matrix and the packed
// Golub-Kahan storage.
double pack_footprint(const std::vector<const PackRequest *> &reqs);

// Solves every request (same Krylov precision and ncv_eff, packable): the
// Golub-Kahan cycles of the members that are ready run as one packed cycle
// on `leader`'s stream; each member runs its replay, candidates, polish and
// restart logic on its own thread and context (`members[i]`, at least as
// many as requests). A request's result does not depend on the other
// members of its pack: alone in a pack it is bitwise the same.
void solve_packed(Context &leader, const std::vector<Context *> &members,
                  const std::vector<PackRequest *> &reqs);

}  // namespace kh
