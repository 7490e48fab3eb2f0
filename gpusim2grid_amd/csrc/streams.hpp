// Prefetched sequential streams for the program-driven kernels.
//
// A wavefront walks its factorization / substitution program strictly in
// order, so every global input it needs -- the program words and its tile's
// value stream -- is read sequentially.  RingStream keeps the next chunks of
// such a stream in an LDS ring: two register stages hold the chunks in flight
// (loaded with 16-byte-per-lane coalesced loads) and are written into the
// ring once the chunk two slots behind has been consumed, so a read at the
// consumption point finds its data in LDS instead of waiting on HBM.
#pragma once

#include <hip/hip_runtime.h>

#include "program_format.hpp"

namespace g2g {

// 16-byte register type (a native vector, so arrays of it stay in VGPRs)
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// T: element type.  CH: positions per chunk.  NCH: ring slots (>= 2).
// PerLane: element (pos, lane) stored [pos][64] (a tile's value stream); else
// one element per position shared by every lane (the program).
// Reverse: the stream is consumed from its end towards position 0.
// The global buffer must be readable up to nch * CH positions.
template <typename T, int CH, int NCH, bool Reverse, bool PerLane>
struct RingStream {
    static constexpr int kElems = PerLane ? CH * kLanes : CH;
    static constexpr int kBytes = kElems * static_cast<int>(sizeof(T));
    static constexpr int kV4 = kBytes / 16 / kLanes;  // 16-byte loads per lane per chunk
    static_assert(kV4 >= 1 && kBytes % (16 * kLanes) == 0, "chunk must be a whole number of wave loads");
    static_assert((CH & (CH - 1)) == 0 && (NCH & (NCH - 1)) == 0, "powers of two");
    static constexpr int kRing = CH * NCH;

    const v4u* g = nullptr;  // stream base (chunk 0)
    T* ring = nullptr;         // LDS, NCH chunks
    int nch = 0;               // chunks in the stream
    int cur = 0;               // current chunk, in consumption order
    int lane = 0;
    v4u s0[kV4], s1[kV4];

    __device__ __forceinline__ int phys(int c) const { return Reverse ? nch - 1 - c : c; }

    __device__ __forceinline__ void load(v4u (&st)[kV4], int c)
    {
        if (c < nch) {
            const v4u* src = g + static_cast<size_t>(phys(c)) * (kBytes / 16) + lane;
#pragma unroll
            for (int i = 0; i < kV4; ++i) st[i] = src[i * kLanes];
        }
    }

    __device__ __forceinline__ void store(const v4u (&st)[kV4], int c)
    {
        if (c < nch) {
            v4u* dst = reinterpret_cast<v4u*>(ring) + static_cast<size_t>(phys(c) & (NCH - 1)) * (kBytes / 16) + lane;
#pragma unroll
            for (int i = 0; i < kV4; ++i) dst[i * kLanes] = st[i];
        }
    }

    __device__ __forceinline__ void init(const void* base, int n_chunks, T* lds, int lane_id)
    {
        g = static_cast<const v4u*>(base);
        nch = n_chunks;
        ring = lds;
        lane = lane_id;
        cur = 0;
#pragma unroll
        for (int c = 0; c < NCH - 1; ++c) {
            load(s0, c);
            store(s0, c);
        }
        if ((NCH - 1) & 1) {
            load(s1, NCH - 1);
            load(s0, NCH);
        } else {
            load(s0, NCH - 1);
            load(s1, NCH);
        }
    }

    // chunk `cur` is done: the chunk NCH-1 ahead enters the ring, the one
    // after the other stage starts loading
    __device__ __forceinline__ void advance()
    {
        const int w = cur + NCH - 1;
        // the distinct markers keep the two stage paths from being merged
        // into one path over a pointer select, which would force the stages
        // into scratch memory
        if (w & 1) {
            asm volatile("; ring stage 1 begin");
            store(s1, w);
            load(s1, w + 2);
            asm volatile("; ring stage 1 end");
        } else {
            asm volatile("; ring stage 0 begin");
            store(s0, w);
            load(s0, w + 2);
            asm volatile("; ring stage 0 end");
        }
        ++cur;
    }

    // makes position p (global index) readable; p must not go backwards in
    // consumption order by more than the current chunk
    __device__ __forceinline__ void ensure(int p)
    {
        const int c = Reverse ? nch - 1 - p / CH : p / CH;
        while (cur < c) advance();
    }

    __device__ __forceinline__ T get(int p)
    {
        ensure(p);
        return PerLane ? ring[static_cast<size_t>(p & (kRing - 1)) * kLanes + lane] : ring[p & (kRing - 1)];
    }
};

// A program word, made wave-uniform.
template <typename S>
__device__ __forceinline__ int prog_at(S& s, int p)
{
    return __builtin_amdgcn_readfirstlane(s.get(p));
}

}  // namespace g2g
