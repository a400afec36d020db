// Dose-influence (Dij) post-processing of a group of beamlets on the device
// (kernels in omc_k64.hip, launchers declared in omc_launch.hpp):
//
//  k_accum_slots  batch statistics per slot (omc_matrad.c:897-918): for
//                 slot s, e = endep/nper[s]; sum += e; sum2 += e^2; endep = 0.
//                 Slot s of a group occupies [s*stride, s*stride + n).
//  k_dij_finish   the finish of each beamlet after its batches
//                 (accumulateResults with iout = 1, omc_matrad.c:920-1012), for
//                 the slots of a group in beamlet order.  The reference never
//                 clears accum_endep2 between beamlets, so the variance of
//                 beamlet b-1 enters beamlet b's sum of squares; the kernel
//                 carries it per voxel from slot to slot and across groups.
//                 In place: sum <- dose (Gy per history), sum2 <- variance;
//                 also per-block maxima and the beamlet-summed dose/variance.
//  k_dij_thresh   threshold = rel * max dose of the beamlet (:1608-1615)
//  k_dij_count    entries strictly above the threshold per chunk (:1617-1624)
//  k_dij_compact  order-preserving compaction into CSC rows (voxel index,
//                 0-based) and values (:1667-1680).
//
// FP64 throughout; built without FP contraction so that every value rounds
// exactly like the host restatement of the reference.
#pragma once

namespace omc {

constexpr int DIJ_BLOCK = 256;
constexpr int DIJ_CHUNK = 1024;     // voxels per compaction chunk (4 per thread)

}  // namespace omc
