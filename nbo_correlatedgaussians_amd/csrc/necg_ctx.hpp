// necg_ctx.hpp -- the single-system context behind the necg_* C ABI (shared
// by the library's translation units; not part of the public header).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <vector>

#include "../../include/necg.h"
#include "necg_abi.hpp"
#include "necg_finish.hpp"

struct necg_ctx {
  // resident state of the last evaluation on this context: bra/ket tables,
  // H, S, HC, SC and scal = (E, 1/cSc, cHc, cSc) of x at K = last_K
  int device = 0;
  int n = 0, nst = 0;
  necg::SysConst sc{};
  double *d_scvv = nullptr, *d_sml = nullptr, *d_sms = nullptr;
  int *d_flag = nullptr;
  // per-K workspace
  int Kcap = 0;
  double *bra = nullptr, *ket = nullptr, *H = nullptr, *S = nullptr, *HC = nullptr,
         *SC = nullptr, *scal = nullptr;
  double *part = nullptr, *hpart = nullptr, *spart = nullptr;
  double *hcp = nullptr, *scp = nullptr, *red = nullptr;
  size_t part_cap = 0, hpart_cap = 0;
  // properties pass: (function, term) records and per-wave partial rows
  double *prop_rec = nullptr, *prop_part = nullptr;
  size_t prop_rec_cap = 0, prop_part_cap = 0;
  // symmetry-reduced properties pass: requested mode (0 auto, 1 reduced,
  // 2 all term pairs), whether the terms were analysed and close under
  // Q P^T, the identity term, and the per-term weight table
  int prop_mode = 0, prop_analysed = 0, prop_grp_ok = 0, prop_ident = -1;
  double *prop_tab = nullptr;
  int req_kt = 0, req_ts = 0, req_w = 0;
  int last_kt = 0, last_ts = 0, last_w = 0;
  int last_K = 0;
  int ncu = 256;      // compute units that launch sizing assumes
  int dev_ncu = 256;  // compute units of the device
  // two-sided pair form: requested mode (0 auto, 1 two-sided, 2 one-sided),
  // whether the terms allow it, the per-(term, lane) target-slot words, and
  // whether the last gradient pair pass used it
  int pair_mode = 0, two_ok = 0, last_two = 0;
  unsigned *d_tperm = nullptr;
  // lower-triangle tile list (TriTile words) for K = tiles_K, kt = tiles_kt:
  // ntri entries, nsplit of its tiles split in halves; d_xr (inside the same
  // allocation) = the extra-slot range of each bra tile
  int *d_tiles = nullptr, *d_xr = nullptr;
  size_t tiles_cap = 0;
  int tiles_K = 0, tiles_kt = 0, ntri = 0, nsplit = 0;
  // folded diagonal blocks (tri_fold): requested (necg_set_fold, default on)
  // and whether the current tile list has them
  int fold_mode = 1, tiles_fold = 0;
  // the lower-triangle passes' last tiling choice (choose_tiling2): kt, ts
  // for K = tc_K, requested (tc_req_kt, tc_req_ts) and tc_ncu CUs
  int tc_K = 0, tc_req_kt = -1, tc_req_ts = -1, tc_ncu = 0, tc_kt = 0, tc_ts = 0, tc_fold = -1;
  size_t tc_budget = 0;   // the arena bytes the choice was made for (0: no arena)
  // per-(bra tile, term group) ket-side accumulator records, [t][l][slot]
  double *kpart = nullptr;
  size_t kpart_cap = 0;
  // forward tape of necg_energy_fwd: the pair partials of x at K = tape_K
  // (0: none; any other evaluation on the context clears it), named tape_id;
  // tape_two: made by the two-sided pass with ket tile tape_kt, tape_ts groups
  int tape_K = 0, tape_ntiles = 0;
  int tape_two = 0, tape_kt = 0, tape_ts = 0;
  long long tape_id = 0, tape_seq = 0;
  // state generations: `gen` counts the rebuilds of the bra/ket tables (every
  // evaluation entry point); `full_gen` is the generation whose H, S, H C,
  // S C and quotient were all formed from one x (a full evaluation at
  // last_K).  A pool reads the resident state only while they are equal.
  long long gen = 0, full_gen = -1;
  // workspace pins (necg_pin_workspace): while > 0, no call may move a
  // workspace buffer or rewrite the tile list, since a captured HIP graph
  // holds their addresses and contents
  int pins = 0;
  // allocation-free mode (necg_outprod_noalloc): while arena is set, the
  // workspace is carved from [arena, arena + arena_cap) instead of allocated,
  // and tile lists are uploaded from host copies kept here per (K, ket tile,
  // CU count) for the life of the context
  char *arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  // bytes the call's tiling choice may give the pair partials, fixed at the
  // call's first choice (later choices in the same call see the same budget)
  size_t arena_budget = 0;
  std::map<long long, std::vector<int>> arena_lists;
  // pair-kernel timing (necg_timing)
  bool timing = false;
  std::vector<hipEvent_t> ev;  // start/stop pairs, reused
  int nev = 0;
  std::mutex mu;
};

