// Per-field kernel launchers, defined in inst_f*.hip (one translation unit per field so the
// four fields compile in parallel) and called by the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace sc {

struct Launchers {
  int limbs;
  hipError_t (*fill)(hipStream_t, const FillArgs&);
  // sid: compiled composition id (STATIC_COMPS), 0 = interpreted with capacities tmax/pmax
  hipError_t (*pass)(hipStream_t, int sid, int tmax, int pmax, const PassArgs&, unsigned grid);
  hipError_t (*finalize)(hipStream_t, int pmax, const FinalizeArgs&);
  hipError_t (*tail)(hipStream_t, int sid, int tmax, int pmax, const TailArgs&);
  hipError_t (*field_op)(hipStream_t, int op, uint64_t n, const uint32_t* a, const uint32_t* b,
                         uint32_t* out, const FieldConsts* fc);
  hipError_t (*transcript)(hipStream_t, const uint32_t* g, int npts, int round,
                           const FieldConsts* fc, ProofState* st);
  // batched stages: grid = tiles of the stage / entries of the stage
  hipError_t (*batch_pass)(hipStream_t, int tmax, int pmax, const BatchArgs&, unsigned tiles);
  hipError_t (*batch_finalize)(hipStream_t, int pmax, const BatchArgs&, unsigned entries);
  // one block per entry running all of its instance's remaining rounds
  hipError_t (*batch_tail)(hipStream_t, int tmax, int pmax, const BatchArgs&, unsigned entries);
  // point-parallel rounds: fold launch (round >= 1), then an evaluation launch with pmax lanes
  // per pair
  hipError_t (*wide_fold)(hipStream_t, const PassArgs&, unsigned grid);
  hipError_t (*wide_eval)(hipStream_t, int tmax, int pmax, const PassArgs&, unsigned grid);
  // limb-major input tables -> element-major (n entries per table, a power of two)
  hipError_t (*limb_to_element)(hipStream_t, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t out_stride,
                                int num_tables);
  // point-parallel form of a batched stage: fold tiles, then evaluation tiles (args.tile_base
  // = the stage's per-pair tiles, so the evaluation partials follow theirs)
  hipError_t (*batch_wide_fold)(hipStream_t, const BatchArgs&, unsigned tiles);
  hipError_t (*batch_wide_eval)(hipStream_t, int tmax, int pmax, const BatchArgs&, unsigned tiles);
};

const Launchers& launchers_p32();
const Launchers& launchers_p64();
const Launchers& launchers_p128();
const Launchers& launchers_p256();

// Table-slot capacities compiled for the per-pair kernels (the smallest that holds the
// composition's tables is used) and point capacities (d+1 <= 4 or <= 8).
constexpr int TMAX_CHOICES[] = {4, 8, 16, 32};
constexpr int PMAX_CHOICES[] = {4, 8};

}  // namespace sc
