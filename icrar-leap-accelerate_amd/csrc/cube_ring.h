// Streaming a visibility cube into LDS for the tiles that run channel
// recurrences over chunks of (row, channel) elements: the FP64 Horner
// rotate/average tiles (kernels_batched.hip) and the recomputing adjoint
// (kernels_adjoint.hip).
//
// A tile owns one baseline and walks its rows RP at a time, each row block's
// channel chunks [c0, c0 + CC) from the last channels down.  With a cube of
// 16-byte-multiple elements (complex128, or complex64 with 2 or 4 pols) at a
// 16-byte-aligned address, the chunks are copied SLOTS ahead into a ring of
// LDS slots with global_load_lds_dwordx4 (no registers held for them), and
// each is converted once to pol-0 FP64 values, zero where any pol is not
// finite (CpuLeapCalibrator.cc:201-214), for the recurrences to read.
#pragma once

#include <type_traits>

#include "batched_common.h"

#define LEAP_LDS __attribute__((address_space(3)))

namespace leap {
namespace dev {

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

constexpr int kRingThreads = 256;  // workgroup size of the tiles that stream the cube

// s_waitcnt vmcnt(0) with the export and LDS/scalar counters left free
// (gfx9 encoding: vmcnt 0, expcnt 7, lgkmcnt 15).
constexpr int kWaitVectorMem = 0x0F70;

// s_waitcnt vmcnt(n), the export and LDS/scalar counters left free; a larger
// n waits as for 8, which is safe.
__device__ __attribute__((always_inline)) inline void wait_vmem_but(uint32_t n) {
    switch (n) {
        case 0: __builtin_amdgcn_s_waitcnt(kWaitVectorMem); break;
        case 1: __builtin_amdgcn_s_waitcnt(kWaitVectorMem | 1); break;
        case 2: __builtin_amdgcn_s_waitcnt(kWaitVectorMem | 2); break;
        case 3: __builtin_amdgcn_s_waitcnt(kWaitVectorMem | 3); break;
        case 4: __builtin_amdgcn_s_waitcnt(kWaitVectorMem | 4); break;
        case 5: __builtin_amdgcn_s_waitcnt(kWaitVectorMem | 5); break;
        case 6: __builtin_amdgcn_s_waitcnt(kWaitVectorMem | 6); break;
        case 7: __builtin_amdgcn_s_waitcnt(kWaitVectorMem | 7); break;
        default: __builtin_amdgcn_s_waitcnt(kWaitVectorMem | 8); break;
    }
}

// One (row, channel) element of a cube with PP pols, read as vector units:
// a 16-byte unit per complex128 pol, pairs of complex64 pols as 16-byte units
// (8-byte units for an odd pol count).
template <typename VT, int PP>
struct StageUnit {
    typedef std::conditional_t<sizeof(VT) == 8, f64x2, std::conditional_t<PP % 2 == 0, f32x4, f32x2>> T;
    static constexpr int K = sizeof(T) / sizeof(VT);  // scalars per unit
    static constexpr int N = 2 * PP / K;               // units per element
};

// The per-tile part of a chunk copy.  Thread tid copies units tid + 256 k
// (k < K) of the chunk; with urow = CC EB / 16 units per row a power of two,
// unit g lies in row g >> lu at unit g & (urow - 1), so its cube address is
// the chunk's first byte + (tid >> lu) rowbytes + 16 (tid & (urow - 1)) (the
// lane's part, fixed for the tile) + the uniform part of 256 k.
struct DmaLane {
    uint64_t rowbytes;  // bytes between timesteps of a baseline: nbl C EB
    uint64_t lane_off;  // the lane's part (pow2 only)
    uint32_t urow;      // units per row
    int32_t lu;         // log2(urow), or -1 when urow is not a power of two
};

__device__ __attribute__((always_inline)) inline DmaLane dma_lane(uint32_t CCEB16, uint32_t nbl, uint32_t C,
                                                                  uint32_t EB) {
    DmaLane d;
    d.urow = CCEB16;
    d.rowbytes = uint64_t(nbl) * C * EB;
    d.lu = (d.urow & (d.urow - 1)) == 0 ? int32_t(__builtin_ctz(d.urow)) : -1;
    const uint32_t tid = threadIdx.x;
    d.lane_off = d.lu >= 0 ? uint64_t(tid >> d.lu) * d.rowbytes + 16u * (tid & (d.urow - 1)) : 0;
    return d;
}

// Starts the copy of rows [tb, tb + RP) x channels [c0, c0 + CC) of the cube
// into raw with global_load_lds_dwordx4 (lane i of a wave writes the 16 bytes
// at the wave's LDS address + 16 i), row r at byte r * CC * EB.  EB, the bytes
// of an element, is a multiple of 16 and the cube 16-byte aligned.  Rows past
// T read row tb, units past the chunk's cn channels the row's first unit: the
// conversion zeroes those slots.  Whole chunks of a power-of-two row take one
// address add per copy; the rest divide per unit.  The copy has landed once
// every thread has waited for its vector memory counter and passed a barrier.
template <int RP, int CC>
__device__ __attribute__((always_inline)) inline void dma_chunk(LEAP_LDS f32x4* raw,
                                                                const LEAP_GLOBAL unsigned char* vis, uint32_t EB,
                                                                const DmaLane& dl, uint32_t tb, uint32_t T,
                                                                uint32_t nbl, uint32_t bl, uint32_t C, uint32_t c0,
                                                                uint32_t cn) {
    static_assert(RP * CC % kRingThreads == 0, "whole waves of units for every wave");
    const uint32_t urow = dl.urow;
    const uint32_t total = RP * urow;  // a multiple of 256
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const LEAP_GLOBAL unsigned char* first = vis + ((size_t(tb) * nbl + bl) * C + c0) * EB;
    if (dl.lu >= 0 && tb + RP <= T && cn == uint32_t(CC)) {
        const LEAP_GLOBAL unsigned char* lane_first = first + dl.lane_off;
        for (uint32_t q = 0; q < total; q += kRingThreads) {
            const uint64_t step = uint64_t(q >> dl.lu) * dl.rowbytes + 16u * (q & (urow - 1));  // uniform
            __builtin_amdgcn_global_load_lds((const LEAP_GLOBAL void*)(lane_first + step),
                                             (LEAP_LDS void*)(raw + q + 64 * wave), 16, 0, 0);
        }
    } else {
        const uint32_t uvalid = cn * (EB / 16);
        const uint32_t lane = threadIdx.x % 64;
        for (uint32_t g0 = wave * 64; g0 < total; g0 += kRingThreads) {
            const uint32_t g = g0 + lane;
            const uint32_t r = g / urow, u = g - r * urow;
            const size_t row = size_t(tb + r < T ? tb + r : tb) * nbl + bl;
            const LEAP_GLOBAL unsigned char* src = vis + (row * C + c0) * EB + size_t(u < uvalid ? u : 0) * 16;
            __builtin_amdgcn_global_load_lds((const LEAP_GLOBAL void*)src, (LEAP_LDS void*)(raw + g0), 16, 0, 0);
        }
    }
    // keep the copies issued here, ahead of the channel loop that follows
    __builtin_amdgcn_sched_barrier(0);
}

// Converts a chunk copied by dma_chunk: element (r, c) is unit range
// [(r CC + c) N, + N) of raw; store(r, c, re, im) receives its pol 0 in FP64,
// 0 past the rows or channels or where any of its pols is not finite.
template <typename VT, int PP, int RP, int CC, typename Store>
__device__ __attribute__((always_inline)) inline void convert_chunk(const LEAP_LDS f32x4* raw, uint32_t tb,
                                                                    uint32_t T, uint32_t cn, Store&& store) {
    using SU = StageUnit<VT, PP>;
    using U = typename SU::T;
    static_assert(sizeof(U) == 16 && RP * CC % kRingThreads == 0, "16-byte units, whole elements per thread");
    const LEAP_LDS U* x = reinterpret_cast<const LEAP_LDS U*>(raw);
#pragma unroll
    for (int e = 0; e < RP * CC / kRingThreads; ++e) {
        const uint32_t i = threadIdx.x + kRingThreads * e;
        const uint32_t r = i / CC, c = i - r * CC;
        U v[SU::N];
#pragma unroll
        for (int k = 0; k < SU::N; ++k) v[k] = x[i * SU::N + k];
        bool finite = true;
#pragma unroll
        for (int k = 0; k < SU::N; ++k)
#pragma unroll
            for (int m = 0; m < SU::K; ++m) finite = finite && __builtin_isfinite(v[k][m]);
        const bool ok = tb + r < T && c < cn && finite;
        store(r, c, ok ? double(v[0][0]) : 0.0, ok ? double(v[0][1]) : 0.0);
    }
}

// A ring of SLOTS chunk copies over the rows [0, T) of baseline bl: start()
// issues the first SLOTS chunks; per chunk, in the order the tile walks them,
// next_slot() waits for this thread's copies of the oldest chunk in flight
// (later ones stay in flight) and returns its slot, which the tile reads
// after a barrier (convert()); issue() copies the next chunk into the slot
// freed last, after the barrier that follows its conversion.
template <typename VT, int RP, int CC, int SLOTS>
struct CubeRing {
    static constexpr uint32_t kSlotUnits = uint32_t(RP) * CC * 8 * sizeof(VT) / 16;  // a chunk of 4 pols
    static constexpr int kBytes = SLOTS * int(kSlotUnits) * 16;

    // Whether a cube takes the ring: 16-byte-multiple elements at a 16-byte-aligned address.
    __device__ static bool accepts(const void* vis, uint32_t P) {
        const uint32_t EB = 2 * P * uint32_t(sizeof(VT));
        return P >= 1 && P <= 4 && EB % 16 == 0 && (reinterpret_cast<uintptr_t>(vis) & 15) == 0;
    }

    LEAP_LDS f32x4* raw;
    const LEAP_GLOBAL unsigned char* vis;
    DmaLane dl;
    uint32_t P, EB, T, nbl, bl, C, n_chunks, per_wave;
    uint32_t dtb, dch, dslot, cslot, n_ahead;  // next chunk to copy, slots, chunks copied and not converted

    __device__ __attribute__((always_inline)) void start(LEAP_LDS f32x4* raw_, const LEAP_GLOBAL VT* vis_, uint32_t P_,
                                                         uint32_t T_, uint32_t nbl_, uint32_t bl_, uint32_t C_) {
        raw = raw_;
        vis = reinterpret_cast<const LEAP_GLOBAL unsigned char*>(vis_);
        P = P_;
        EB = 2 * P_ * uint32_t(sizeof(VT));
        T = T_;
        nbl = nbl_;
        bl = bl_;
        C = C_;
        n_chunks = (C_ + CC - 1) / CC;
        per_wave = uint32_t(RP) * CC * (EB / 16) / kRingThreads;  // copies per wave and chunk
        dl = dma_lane(uint32_t(CC) * (EB / 16), nbl_, C_, EB);
        dtb = 0;
        dch = n_chunks - 1;
        dslot = cslot = n_ahead = 0;
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) issue();
    }
    __device__ __attribute__((always_inline)) void issue() {
        if (dtb >= T) return;
        ++n_ahead;
        const uint32_t c0 = dch * CC;
        dma_chunk<RP, CC>(raw + dslot * kSlotUnits, vis, EB, dl, dtb, T, nbl, bl, C, c0, min(uint32_t(CC), C - c0));
        dslot = dslot + 1 == SLOTS ? 0 : dslot + 1;
        if (dch == 0) {
            dch = n_chunks - 1;
            dtb += RP;
        } else {
            --dch;
        }
    }
    __device__ __attribute__((always_inline)) const LEAP_LDS f32x4* next_slot() {
        wait_vmem_but((n_ahead - 1) * per_wave);
        --n_ahead;
        const LEAP_LDS f32x4* slot = raw + cslot * kSlotUnits;
        cslot = cslot + 1 == SLOTS ? 0 : cslot + 1;
        return slot;
    }
    // store(r, c, re, im) for every element of the chunk of row block tb with cn channels
    template <typename Store>
    __device__ __attribute__((always_inline)) void convert(const LEAP_LDS f32x4* slot, uint32_t tb, uint32_t cn,
                                                           Store&& store) const {
        if constexpr (sizeof(VT) == 8) {
            switch (P) {
                case 1: convert_chunk<VT, 1, RP, CC>(slot, tb, T, cn, store); break;
                case 2: convert_chunk<VT, 2, RP, CC>(slot, tb, T, cn, store); break;
                case 3: convert_chunk<VT, 3, RP, CC>(slot, tb, T, cn, store); break;
                default: convert_chunk<VT, 4, RP, CC>(slot, tb, T, cn, store); break;
            }
        } else {
            if (P == 2)
                convert_chunk<VT, 2, RP, CC>(slot, tb, T, cn, store);
            else
                convert_chunk<VT, 4, RP, CC>(slot, tb, T, cn, store);
        }
    }
};

}  // namespace dev
}  // namespace leap
