// Dose-influence (Dij) post-processing of a group of beamlets on the device
// (kernels in omc_k64.hip, launchers declared in omc_launch.hpp).  Regions
// are the reference's: irl = 1 + voxel, region 0 outside.  The transport
// marks every tile of TALLY_TILE regions a batch slot received energy in
// (TallyAtomic), so the passes below read only touched tiles.
//
//  k_dij_fold_finish  when a group's jobs carry all its batches: for every
//                 beamlet in order, accumEndep over its batch slots
//                 (omc_matrad.c:897-918) and the finish (accumulateResults
//                 with iout = 1, :920-1012) in one pass, in the reference's
//                 order of operations, including the carry of beamlet b-1's
//                 variance into beamlet b's sum of squares (the reference
//                 never clears accum_endep2).  Writes dose and variance of
//                 touched tiles, the per-beamlet tile map `any`, per-block
//                 maxima, the plan totals, and clears the slots it read.
//  k_accum_slots  batch statistics of one job when the batches of a group
//                 span several jobs: sum += endep/nper, sum2 += (endep/nper)^2.
//  k_dij_finish   the finish after k_accum_slots, in place (sum <- dose,
//                 sum2 <- variance); the carried variance is added to the
//                 finished sum of squares, which differs from the reference's
//                 order by one rounding.
//  k_dij_thresh   threshold = rel * max dose of the beamlet (:1608-1615)
//  k_dij_count    entries strictly above the threshold per chunk (:1617-1624)
//  k_dij_compact  order-preserving compaction into CSC rows (voxel index,
//                 0-based) and values (:1667-1680).
//
// FP64 throughout; built without FP contraction so that every value rounds
// exactly like the CPU restatement of the reference.
#pragma once

namespace omc {

constexpr int DIJ_BLOCK = 256;      // four wavefronts: four tiles of regions
constexpr int DIJ_CHUNK = 1024;     // regions per compaction chunk (four blocks' worth)
static_assert(DIJ_BLOCK == 4 * TALLY_TILE && DIJ_CHUNK % DIJ_BLOCK == 0, "tile and chunk layout");

inline long dij_ntile(long nreg) { return (nreg + TALLY_TILE - 1) / TALLY_TILE; }
inline int dij_nblk(long nreg) { return (int)((nreg + DIJ_BLOCK - 1) / DIJ_BLOCK); }
inline int dij_nchunk(long nreg) { return (int)((nreg + DIJ_CHUNK - 1) / DIJ_CHUNK); }

}  // namespace omc
