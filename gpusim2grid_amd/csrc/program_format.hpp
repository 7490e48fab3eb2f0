// Encoding shared by the host symbolic layer and the device kernels.
#pragma once

#ifndef __HIPCC__
#define __host__
#define __device__
#endif

namespace g2g {

constexpr int kLanes = 64;          // rows per wavefront (one lane per batch row)
constexpr int kMisWaves = 16;       // wavefronts per tile in the update and refinement kernels
constexpr int kMismatchWaves = 16;   // wavefronts per tile in the mismatch kernel
constexpr int kMaxC = 16;           // longest off-diagonal L column the kernels unroll for

constexpr int kFastC = 7;           // column lengths with a specialised kernel path

// stream prefetch geometry (streams.hpp): positions per chunk and ring slots
constexpr int kChunkF64 = 16;       // 8 KB chunks of a wavefront's FP64 stream
constexpr int kRingSlotsF64 = 8;    // 64 KB ring
constexpr int kChunkF32 = 32;       // 8 KB chunks of a wavefront's FP32 stream
constexpr int kRingSlotsF32 = 8;    // 64 KB ring
constexpr int kValueProducers = 6;  // wavefronts streaming a value stream (one chunk in flight each)
constexpr int kProgChunk = 512;     // program words per chunk (2 KB)
constexpr int kProgSlots = 4;       // 8 KB ring
constexpr int kProgMirror = 128;    // program words copied after the ring (the longest record read at once)

// Programs (one record per pivot, padded to 4 words so the kernels read a
// record with 16-byte loads):
//   factorization  [head | next head << 16, diag, b(k), L[c], U[c]]   sources
//                  c rows of c refs                                   rank-1 update targets (row-major)
//                  c refs                                             forward-substitution targets b(i)
//                  (2-column supernodes: see frec2_words)
//   backward       [c | c_next << 16, perm[k], x(j)[c], out]
// c_next is the column length of the following record (0 after the last).
// A ref is a 32-bit byte offset, slot * kRefUnit: the offset of the slot's
// 64-lane row of FP32 values (an FP64 slot sits at twice the offset).  Slot S
// (the slot count) is the zero slot a missing source reads, S + 1 the trash
// slot an unneeded result goes to.  The factorization reads a source slot and
// clears it; an update target accumulates (slots start at zero), so one slot
// serves a sequence of update targets without a first-update flag.
constexpr int kSlotMask = 0x7FFF;     // host-side slot index field
constexpr int kFirst = 1 << 16;       // host-side "first update" flag (bookkeeping only)
constexpr int kRefUnit = 256;         // bytes per FP32 slot row

__host__ __device__ constexpr int round4(int v) { return (v + 3) / 4 * 4; }
__host__ __device__ constexpr int frec_src_words(int c) { return 3 + 2 * c; }  // header + sources
__host__ __device__ constexpr int frec_row_words(int c) { return c; }
__host__ __device__ constexpr int frec_words(int c) { return round4(frec_src_words(c) + (c + 1) * c); }
// 2-column supernode (pivots k, k+1; c = |struct(k+1)|): header, the sources of
// pivot k (2 + 2 (c + 1)) and of pivot k+1 (2 + 2c), then c rows of c update
// targets and the c forward-substitution targets
__host__ __device__ constexpr int frec2_src_words(int c) { return 7 + 4 * c; }
__host__ __device__ constexpr int frec2_words(int c) { return round4(frec2_src_words(c) + (c + 1) * c); }
// record header field: c | (width - 1) << 15; a header word holds this
// record's field and, in its high half, the next record's
__host__ __device__ constexpr int rec_head(int c, int w) { return c | ((w - 1) << 15); }
constexpr int kFastC2 = 6;          // longest struct(k+1) of a 2-column supernode
__host__ __device__ constexpr int brec_words(int c) { return round4(3 + c); }
// backward 2-column supernode: [head, perm[k+1], perm[k], x(S)[c], out(k+1), out(k)]
__host__ __device__ constexpr int brec2_words(int c) { return round4(5 + c); }

// Walkers.  A tile's walk can be split over kMaxWalkers walker wavefronts:
// the host cuts the elimination forest into independent subtrees (each
// walker's first phase) and a small top part.  Walker 0 then waits for the
// others, merges their pending contributions to the top entries with merge
// records and walks the top; the backward substitution walks the top first
// (walker 0) and then the subtrees in parallel.  Each walker has its own
// program, A-stream and U-stream segment; segments start at multiples of
// kSegAlign positions / kProgChunk words.
constexpr int kMaxWalkers = 4;        // backward substitution
constexpr int kMaxFactorWalkers = 2;  // factorization (its slots grow with every walker)
constexpr int kSegAlign = 32;
// merge record: [kMergeHead | next head << 16, n, (dst ref, src ref) x n]
// (dst += src, src = 0); the first one also waits for the other walkers
constexpr int kMergeHead = 0x7FFE;
constexpr int kMergePairs = 31;
__host__ __device__ constexpr int mrec_words(int n) { return round4(2 + 2 * n); }
// split walk records (their n is at most kSplitRecItems):
//   dump    [kDumpHead | next << 16, n, first, ref x n]: pend[first + i] = slot ref_i
//   gather  [kGatherHead | next << 16, n, (dst ref, pend index) x n]: dst += pend[index]
//   preload [kPreloadHead | next << 16, n, (dst ref, x index) x n] (backward): dst = x[index]
constexpr int kDumpHead = 0x7FFD;
constexpr int kGatherHead = 0x7FFC;
constexpr int kPreloadHead = 0x7FFB;
constexpr int kSplitRecItems = 30;
__host__ __device__ constexpr int dump_words(int n) { return round4(3 + n); }
__host__ __device__ constexpr int pair_rec_words(int n) { return round4(2 + 2 * n); }
// walker kernels' LDS layout: per walker a 256-byte control area (value and
// program ring), then kWalkFlagBytes of cross-walker flags, the W value
// rings, the W program rings, and the slots
constexpr int kWalkerCtlBytes = 256;
constexpr int kWalkFlagBytes = 1280;
// value-stream producers per workgroup: six for one walker, two per walker
// for two (eight waves either way, which leaves 256 VGPRs per wave)
__host__ __device__ constexpr int value_producers(int W) { return W == 1 ? kValueProducers : 2 * W; }
__host__ __device__ constexpr int walk_waves(int W) { return 2 * W + value_producers(W); }
// k_factor's walk modes: the factorization; the factorization that also
// stores its L multipliers in a stream of the A stream's layout; the forward
// substitution alone over such a stream
constexpr int kFactorFull = 0, kFactorStoreL = 1, kForwardOnly = 2;
// value-stream producers of a factorization workgroup: as above, except two
// walkers, which may take more (their walk runs at up to 168 VGPRs, so twelve
// waves fit a CU), and the forward-only walk (about 40 VGPRs), which consumes
// its stream several times faster than a factorization: twelve waves
constexpr int kFactorProducers2 = 4;
__host__ __device__ constexpr int factor_value_producers(int W, int mode = kFactorFull)
{
    return mode == kForwardOnly ? 12 - 2 * W : W == 2 ? kFactorProducers2 : value_producers(W);
}
__host__ __device__ constexpr int factor_walk_waves(int W, int mode = kFactorFull)
{
    return 2 * W + factor_value_producers(W, mode);
}
// value ring geometry of a walker: positions per chunk and ring slots (four
// walkers stream half-size chunks, so their rings fit the LDS together)
__host__ __device__ constexpr int value_chunk(int real_size, int W = 1)
{
    return (real_size == 8 ? kChunkF64 : kChunkF32) / (W >= 4 ? 2 : 1);
}
__host__ __device__ constexpr int value_ring_slots(int real_size, int W)
{
    return W == 1 ? (real_size == 8 ? kRingSlotsF64 : kRingSlotsF32) : 4;
}
// (the rings of W walkers streaming with the geometry of RW walkers)
__host__ __device__ constexpr int walk_ring_bytes(int real_size, int W, int RW)
{
    return W * (kWalkerCtlBytes + value_chunk(real_size, RW) * value_ring_slots(real_size, RW) * kLanes * real_size +
                (kProgChunk * kProgSlots + kProgMirror) * 4) +
           kWalkFlagBytes;
}
__host__ __device__ constexpr int walk_ring_bytes(int real_size, int W) { return walk_ring_bytes(real_size, W, W); }
// two FP32 factorization walkers may stream half-size chunks (the four
// walkers' geometry: 16 KB rings), for slot sets too large beside full rings
constexpr int kHalfRingGeometry = 4;

// split walk: sets per tile (from kSplitMinD unknowns up), the walker
// workgroup's value producers, and its ring geometry (as for RW walkers): the
// factorization's full chunks, whose window spans a two-column step's 30
// values; the backward substitution's half chunks (a record spans at most 17)
constexpr int kSplitSets = 32;
constexpr int kSplitMinD = 2000;
constexpr int kSplitProducers = 2;
constexpr int kSplitRingF = 2;
constexpr int kSplitRingB = 4;
__host__ __device__ constexpr int split_ring_bytes(int real_size, int RW)
{
    return kWalkerCtlBytes + value_chunk(real_size, RW) * value_ring_slots(real_size, RW) * kLanes * real_size +
           (kProgChunk * kProgSlots + kProgMirror) * 4 + kWalkFlagBytes;
}

// XCDs of the device: a launch's workgroups go to them round robin
constexpr int kXcds = 8;

// mismatch program: ints per bus record
constexpr int kK1BusWords = 8;

// Extension rows.  Injection records (an extension bus's residual terms):
//   kExtInjDs    dF_P -= sa w        (index: distributed-slack record -> weight)
//   kExtInjHvdc1 dF_P -= p1 flow of line `index`;  kExtInjHvdc2: p2 flow
//   kExtInjVc    dF_Q += Q of controller `index`
// J records {A position, kind, index, add}: add = 1 adds onto a value the
// mismatch wrote (or an earlier record of the same position), 0 stores:
//   kExtJConst   the value ext_const[index] (weights, slopes, +-1 of the VC rows)
//   kExtJHvdc    dP1/dtheta1 of line index (4 * line + q: q = 0 d1/d1, 1 d1/d2,
//                2 d2/d1, 3 d2/d2)
// F records {A position, kind, index}: kExtFVc  -(|V(reg)| + sum s Q - v_set)
// of group index; kExtFShare  -(w_1 Q_k - w_k Q_1) of share row index.
constexpr int kExtInjDs = 0, kExtInjHvdc1 = 1, kExtInjHvdc2 = 2, kExtInjVc = 3;
constexpr int kExtJConst = 0, kExtJHvdc = 1;
constexpr int kExtFVc = 0, kExtFShare = 1;
constexpr int kHvdcTripped = 2;

// row status (gpusim2grid contingency/row_status.hpp:33-43) plus "still iterating"
constexpr int kActive = -1;
constexpr int kConverged = 0;
constexpr int kMaxIter = 1;
constexpr int kDiverged = 2;
constexpr int kNotSimulated = 3;
// internal: a warm-started row whose mismatch grew (re-queued to a flat start)
constexpr int kRejected = 4;

}  // namespace g2g
