// Prefetched sequential streams for the program-driven kernels.
//
// A wavefront walks its factorization / substitution program strictly in
// order, so every global input it needs -- the program words and its tile's
// value stream -- is read sequentially.  Each such stream gets an LDS ring of
// NCH chunks filled by W producer wavefronts of the same workgroup: producer
// w loads chunks w, w + W, w + 2W, ... (one chunk in flight per producer, so
// no load is carried across loop iterations and its waits are exact), writes
// each into its ring slot and tags the slot with the chunk number.  The
// consumer wavefront (the one walking the program) only reads LDS: it waits
// for the tags of the chunks it needs and publishes `lo`, the oldest chunk it
// may still read, which frees the older slots for the producers.
//
// Every wait is bounded: after kSpinLimit polls a wave sets the abort flag
// and leaves, and the other waves leave once they see it, so the grid always
// drains (the kernel then reports an error).
#pragma once

#include <hip/hip_runtime.h>

#include "program_format.hpp"

namespace g2g {

// 16-byte register type (a native vector, so arrays of it stay in VGPRs)
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));

constexpr int kSpinLimit = 1 << 22;  // polls (each after s_sleep 2) before a wave gives up

// Flow-control flags.  The LDS executes a wavefront's DS instructions in
// order and is coherent across the workgroup, so a producer's ring writes
// land before its later tag write, and a consumer's ring reads issued after
// it saw a tag see the data.  The flags are therefore relaxed atomics fenced
// against compiler reordering only: release/acquire at workgroup scope would
// also wait for every outstanding global access of the wave (the consumer's
// U-stream stores, the producer's loads), which is exactly the latency the
// rings are there to hide.
// (every flag is a wave-uniform value: read into a scalar register, so the
// polling loops branch on the scalar unit instead of masking lanes)
__device__ __forceinline__ int lds_load_acquire(const int* p)
{
    const int v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return __builtin_amdgcn_readfirstlane(v);
}

// a per-lane LDS value another wave published before a flag this wave has
// seen (relaxed, no readfirstlane: every lane keeps its own value)
__device__ __forceinline__ int lds_load_lane(const int* p)
{
    const int v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return v;
}

__device__ __forceinline__ void lds_store_release(int* p, int v)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int NCH>
struct RingCtl {
    int lo;          // oldest chunk the consumer may still read
    int abort;       // set by a wave that gave up waiting
    int tag[NCH];    // chunk number + 1 held by each slot (0: empty)
};

// T: element type.  CH: positions per chunk.  NCH: ring slots.
// PerLane: element (pos, lane) stored [pos][64] (a tile's value stream); else
// one element per position shared by every lane (the program).
// Reverse: the stream is consumed from its end towards position 0.
// MIRROR (shared forward streams): the ring is followed by a copy of its
// first MIRROR positions, written with the chunk of slot 0, so a run of up
// to MIRROR positions starting anywhere in the ring reads contiguously.
// The global buffer must be readable up to nch * CH positions.
template <typename T, int CH, int NCH, bool Reverse, bool PerLane, int MIRROR = 0>
struct Ring {
    static constexpr int kElems = PerLane ? CH * kLanes : CH;
    static constexpr int kBytes = kElems * static_cast<int>(sizeof(T));
    static constexpr int kV4 = kBytes / 16 / kLanes;  // 16-byte loads per lane per chunk
    static_assert(kV4 >= 1 && kBytes % (16 * kLanes) == 0, "chunk must be a whole number of wave loads");
    static_assert((CH & (CH - 1)) == 0 && (NCH & (NCH - 1)) == 0, "powers of two");
    static_assert(MIRROR == 0 || (!PerLane && !Reverse && MIRROR <= CH && MIRROR * sizeof(T) % 16 == 0),
                  "a mirror follows a shared forward ring and copies part of one chunk");
    static constexpr int kRing = CH * NCH;
    static constexpr int kLdsBytes = kBytes * NCH + MIRROR * static_cast<int>(sizeof(T));
    using Ctl = RingCtl<NCH>;

    // global chunk of consumption-order chunk c
    __device__ static __forceinline__ int phys(int c, int nch) { return Reverse ? nch - 1 - c : c; }
    // consumption-order chunk of global position p (p >= 0)
    __device__ static __forceinline__ int chunk_of(int p, int nch)
    {
        const int c = static_cast<int>(static_cast<unsigned>(p) / CH);
        return Reverse ? nch - 1 - c : c;
    }

    __device__ static __forceinline__ void init_ctl(Ctl* ctl, int lane)
    {
        if (lane == 0) {
            ctl->lo = 0;
            ctl->abort = 0;
        }
        for (int i = lane; i < NCH; i += kLanes) ctl->tag[i] = 0;
    }

    // ---- producer wavefront w of W ---------------------------------------
    __device__ static __forceinline__ void produce(const void* base, int nch, T* ring, Ctl* ctl, int lane, int w, int W)
    {
        const v4u* g = static_cast<const v4u*>(base);
        v4u* r = reinterpret_cast<v4u*>(ring);
        int lo = 0;
        for (int c = w; c < nch; c += W) {
            v4u st[kV4];
            const v4u* src = g + static_cast<size_t>(phys(c, nch)) * (kBytes / 16) + lane;
#pragma unroll
            for (int i = 0; i < kV4; ++i) st[i] = src[i * kLanes];
            // the slot of chunk c is free once the consumer moved past c - NCH
            int spins = 0;
            while (c - lo >= NCH) {
                lo = lds_load_acquire(&ctl->lo);
                if (c - lo < NCH) break;
                if (lds_load_acquire(&ctl->abort) || ++spins > kSpinLimit) {
                    lds_store_release(&ctl->abort, 1);
                    return;
                }
                __builtin_amdgcn_s_sleep(32);
            }
            // data slot by global chunk (what Reader::at addresses), tag by
            // consumption-order chunk (what Reader::wait checks)
            const int slot = phys(c, nch) & (NCH - 1);
            v4u* dst = r + static_cast<size_t>(slot) * (kBytes / 16) + lane;
#pragma unroll
            for (int i = 0; i < kV4; ++i) dst[i * kLanes] = st[i];
            if (MIRROR > 0 && slot == 0) {
#pragma unroll
                for (int i = 0; i < kV4; ++i)
                    if (i * kLanes + lane < MIRROR * static_cast<int>(sizeof(T)) / 16) dst[kRing * sizeof(T) / 16 + i * kLanes] = st[i];
            }
            lds_store_release(&ctl->tag[c & (NCH - 1)], c + 1);
        }
    }

    // ---- consumer wavefront ----------------------------------------------
    struct Reader {
        static constexpr int kRingLen = kRing;  // positions the ring holds
        const T* ring;   // ring base (+ this lane for per-lane streams)
        Ctl* ctl;
        int nch;
        int ready;       // chunks [lo, ready) known to be in the ring
        int lo;          // published lo
        int ready_end;   // forward streams: positions below ready * CH are readable
        int lo_next;     // forward streams: a window from (lo + 2) * CH on frees a chunk
        bool failed;

        __device__ __forceinline__ void init(T* r, Ctl* c, int n, int lane)
        {
            ring = PerLane ? r + lane : r;
            ctl = c;
            nch = n;
            ready = 0;
            lo = 0;
            ready_end = 0;
            lo_next = 2 * CH;
            failed = false;
        }

        // makes the positions from `first` to `last` (consumption order, at
        // most (NCH - 2) * CH apart) readable; frees the chunks before first's.
        // A forward stream first compares against two position thresholds: a
        // window inside the known-ready chunks that frees no chunk does nothing.
        __device__ __forceinline__ void window(int first, int last)
        {
            if constexpr (!Reverse) {
                if (last < ready_end && first < lo_next) return;
            }
            const int cf = chunk_of(first, nch);
            const int cl = chunk_of(last, nch);
            if (cf - 1 > lo) {
                lo = cf - 1;
                lds_store_release(&ctl->lo, lo);
            }
            if (ready < cf) ready = cf;  // chunks before the window are not needed
            if (ready <= cl) wait_for(cl);
            if constexpr (!Reverse) {
                ready_end = ready * CH;
                lo_next = (lo + 2) * CH;
            }
        }

        // Makes chunks up to cl readable.  One poll reads every slot's tag at
        // once and advances `ready` over all consecutive chunks already
        // written, so a poll also covers the chunks the next windows need.
        __device__ __forceinline__ void wait_for(int cl)
        {
            if (failed) return;
            int spins = 0;
            for (;;) {
                int t[NCH];
#pragma unroll
                for (int i = 0; i < NCH; ++i) t[i] = lds_load_acquire(&ctl->tag[i]);
#pragma unroll
                for (int i = 0; i < NCH; ++i) {
                    // chunk ready + i sits in slot (ready + i) & (NCH - 1)
                    int tv = t[0];
#pragma unroll
                    for (int q = 1; q < NCH; ++q) tv = ((ready & (NCH - 1)) == q) ? t[q] : tv;
                    if (tv != ready + 1 || ready >= lo + NCH) break;
                    ++ready;
                }
                if (ready > cl) return;
                if (lds_load_acquire(&ctl->abort) || ++spins > kSpinLimit) {
                    lds_store_release(&ctl->abort, 1);
                    failed = true;  // the results are discarded
                    return;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }

        __device__ __forceinline__ void wait(int c)
        {
            if (failed) return;
            int spins = 0;
            const int* tag = &ctl->tag[c & (NCH - 1)];
            while (lds_load_acquire(tag) != c + 1) {
                if (lds_load_acquire(&ctl->abort) || ++spins > kSpinLimit) {
                    lds_store_release(&ctl->abort, 1);
                    failed = true;  // the results are discarded
                    return;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }

        __device__ __forceinline__ T at(int p) const
        {
            return PerLane ? ring[static_cast<size_t>(p & (kRing - 1)) * kLanes] : ring[p & (kRing - 1)];
        }
    };
};

}  // namespace g2g
