// necg_pool_impl.hpp -- the resident-system record that trial kernels read,
// shared by the context-backed and batch-backed pools (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/necg.h"
#include "necg_abi.hpp"
#include "necg_finish.hpp"

struct necg_pool;

namespace necg {

struct PoolSys {
  SysConst sc;
  int n, nst, K;
  const double *bra, *H, *S, *HC, *SC, *scal, *sml, *sms, *scvv;
  // resident factors of the pencil (H, S), or null: V (K x K row-major) with
  // V^T S V = I and V^T H V = diag(lam), lam ascending
  const double *V, *lam;
};

// A pool over resident records (device pointers that stay valid while the
// pool lives).
int pool_create_from(int device, const std::vector<PoolSys> &recs, necg_pool **out);

}  // namespace necg
