// Encoding shared by the host symbolic layer and the device kernels.
#pragma once

#ifndef __HIPCC__
#define __host__
#define __device__
#endif

namespace g2g {

constexpr int kLanes = 64;          // rows per wavefront (one lane per batch row)
constexpr int kMisWaves = 16;       // wavefronts per tile in the mismatch and update kernels
constexpr int kMaxC = 16;           // longest off-diagonal L column the kernels unroll for

constexpr int kFastC = 10;          // column lengths with a specialised kernel path

// stream prefetch geometry (streams.hpp): positions per chunk and ring slots
constexpr int kChunkF64 = 8;        // 4 KB chunks of a wavefront's FP64 stream
constexpr int kRingSlotsF64 = 8;    // 32 KB ring
constexpr int kChunkF32 = 16;       // 4 KB chunks of a wavefront's FP32 stream
constexpr int kRingSlotsF32 = 16;   // 64 KB ring
constexpr int kValueProducers = 6;  // wavefronts streaming a value stream (one chunk in flight each)
constexpr int kProgChunk = 512;     // program words per chunk (2 KB)
constexpr int kProgSlots = 4;       // 8 KB ring
constexpr int kStreamWaves = 2 + kValueProducers;  // walker + value producers + program producer

// Programs (one record per pivot, 4-word aligned, length fixed by c):
//   factorization  [c | c_next << 16] then 16-bit refs
//                  diag, b(k), L[c], U[c], update[c*c] (row-major), b-update[c]
//   backward       [c | c_next << 16] [perm[k]] then 16-bit refs x(j)[c], out
// c_next is the column length of the following record (0 after the last).
// A ref is a slot index; kRefFirst marks the first update of a target (store,
// not accumulate).  Slot S (the slot count) is the zero slot a missing source
// reads, S + 1 the trash slot an unneeded result goes to.
constexpr int kSlotMask = 0x7FFF;     // host-side slot index field
constexpr int kFirst = 1 << 16;       // host-side "first update" flag
constexpr unsigned kRefSlot = 0x7FFF; // 16-bit ref: slot index bits
constexpr unsigned kRefFirst = 0x8000;

__host__ __device__ constexpr int record_words(int head, int nrefs) { return (head + (nrefs + 1) / 2 + 3) / 4 * 4; }
__host__ __device__ constexpr int frec_words(int c) { return record_words(1, (c + 1) * (c + 2)); }
__host__ __device__ constexpr int brec_words(int c) { return record_words(2, c + 1); }

// row status (gpusim2grid contingency/row_status.hpp:33-43) plus "still iterating"
constexpr int kActive = -1;
constexpr int kConverged = 0;
constexpr int kMaxIter = 1;
constexpr int kDiverged = 2;
constexpr int kNotSimulated = 3;
// internal: a warm-started row whose mismatch grew (re-queued to a flat start)
constexpr int kRejected = 4;

}  // namespace g2g
