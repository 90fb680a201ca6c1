// Batched rotate/average for gfx950: one pass over the visibilities serves a
// block of directions.
//
// For one row (t, baseline) and direction d the reference sums, per pol p,
//   sum_c vis[c][p] * exp(i * sf_d * kappa_c),   kappa_c = (f0 + c df) / c_light,
// skipping channels whose rotated pols hold a NaN (CpuLeapCalibrator.cc:181-217).
// kappa_c is linear in c, so over a block of B channels starting at b
//   sum_k v[b + k] exp(i sf (kappa_b + k dkappa)) = exp(i sf kappa_b) * P(z),
//   P(z) = sum_k v[b + k] z^k,  z = exp(i sf dkappa),
// and P(z) is evaluated by Horner's rule: four fused multiply-adds per channel
// and direction, with the exact phasor recomputed at every block start.
//
// Precision variants:
//   FP64:  phases, sincos, Horner and sums in FP64.
//   MIXED: the phase argument sf * kappa and its reduction modulo 2 pi in FP64;
//          sincos, Horner and the sums in FP32, the block sums accumulated over
//          blocks and timesteps with Kahan compensation.
//
// Work decomposition: a workgroup owns one baseline and a block of DB
// directions (lane = direction); 256 / DB timesteps run side by side.  Each
// channel chunk of those rows is staged once into LDS (all pols read, NaN/Inf
// mask applied, the accumulated pols kept) and broadcast to every direction
// lane.  No atomics: the per-timestep partials are combined in LDS in a fixed
// order, so results are reproducible.
//
// Mask rule: a (row, channel) whose polarisations hold any non-finite value is
// skipped for every direction.  The reference tests the rotated values for NaN;
// for finite and NaN inputs the two rules agree, and for an infinite input the
// reference's result depends on the direction (inf * 0 = NaN), where here the
// channel is always skipped.
#include "cube_ring.h"

namespace leap {
namespace {

using namespace dev;

constexpr int kThreads = 256;
constexpr int kBlockCh = 32;  // Horner block length (channels)

template <typename A>
__device__ inline C2<A> phasor(double theta);
template <>
__device__ inline C2<float> phasor<float>(double theta) { return phasor_mixed(theta); }

// Running sum: plain for FP64, Kahan-compensated for FP32.
template <typename A>
struct Accum {
    C2<A> s{0, 0};
    C2<A> comp{0, 0};
    __device__ inline void add(C2<A> v) {
        if constexpr (sizeof(A) == 8) {
            s.x += v.x;
            s.y += v.y;
        } else {
            const A yx = v.x - comp.x;
            const A tx = s.x + yx;
            comp.x = (tx - s.x) - yx;
            s.x = tx;
            const A yy = v.y - comp.y;
            const A ty = s.y + yy;
            comp.y = (ty - s.y) - yy;
            s.y = ty;
        }
    }
    __device__ inline double re() const { return double(s.x) - double(comp.x); }
    __device__ inline double im() const { return double(s.y) - double(comp.y); }
};

// VT: stored component type (double: complex128, float: complex64).
// A: arithmetic type.  DB: directions per workgroup.  NP: pols accumulated
// (1 = pol 0 only, else all P <= 4).  CC: channels per LDS chunk.
template <typename VT, typename A, int DB, int NP, int CC>
__global__ __launch_bounds__(kThreads) void rotate_average_batched(
    const VT* __restrict__ vis, const double* __restrict__ uvw, const double* __restrict__ dd,
    uint32_t D, uint32_t T, uint32_t nbl, uint32_t C, uint32_t P, double kappa0, double dkappa,
    double* __restrict__ avg) {
    constexpr int RP = kThreads / DB;  // timesteps in flight
    __shared__ C2<A> s_v[RP][NP][CC];
    __shared__ double s_red[(RP > 1) ? RP : 1][DB][NP][2];

    const uint32_t bl = blockIdx.x;
    const uint32_t lane_d = threadIdx.x % DB;
    const uint32_t rsub = threadIdx.x / DB;
    const uint32_t d = blockIdx.y * DB + lane_d;
    const bool dir_ok = d < D;
    const uint32_t dc = dir_ok ? d : 0;
    // (dd * uvw).z - w = dd20 u + dd21 v + dd22 w - w (CpuLeapCalibrator.cc:190)
    const double dz0 = dd[9 * dc + 6], dz1 = dd[9 * dc + 7], dz2 = dd[9 * dc + 8];

    Accum<A> acc[NP];

    for (uint32_t tb = 0; tb < T; tb += RP) {
        const uint32_t t = tb + rsub;
        const bool row_ok = t < T;
        double sf = 0.0;
        if (row_ok) {
            const size_t row = size_t(t) * nbl + bl;
            const double u = uvw[3 * row], v = uvw[3 * row + 1], w = uvw[3 * row + 2];
            sf = -kTwoPi * ((dz0 * u + dz1 * v + dz2 * w) - w);
        }
        const C2<A> z = phasor<A>(sf * dkappa);
        for (uint32_t c0 = 0; c0 < C; c0 += CC) {
            const uint32_t cn = min(uint32_t(CC), C - c0);
            __syncthreads();
            // Stage rows [tb, tb + RP) x channels [c0, c0 + cn) into LDS.
            for (uint32_t i = threadIdx.x; i < RP * uint32_t(CC); i += kThreads) {
                const uint32_t r = i / CC, c = i - r * CC;
                const uint32_t tt = tb + r;
                C2<A> val[NP];
#pragma unroll
                for (int q = 0; q < NP; ++q) val[q] = {0, 0};
                if (tt < T && c < cn) {
                    const size_t row = size_t(tt) * nbl + bl;
                    const VT* p = vis + ((row * C + c0 + c) * P) * 2;
                    bool finite = true;
                    VT x[8];
                    for (uint32_t q = 0; q < P; ++q) {
                        x[2 * q] = p[2 * q];
                        x[2 * q + 1] = p[2 * q + 1];
                        finite = finite && __builtin_isfinite(x[2 * q]) && __builtin_isfinite(x[2 * q + 1]);
                    }
                    if (finite) {
#pragma unroll
                        for (int q = 0; q < NP; ++q) val[q] = {A(x[2 * q]), A(x[2 * q + 1])};
                    }
                }
#pragma unroll
                for (int q = 0; q < NP; ++q) s_v[r][q][c] = val[q];
            }
            __syncthreads();
            if (!row_ok || !dir_ok) continue;
            for (uint32_t b = 0; b < cn; b += kBlockCh) {
                const uint32_t len = min(uint32_t(kBlockCh), cn - b);
                const C2<A> base = phasor<A>(sf * (kappa0 + double(c0 + b) * dkappa));
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    C2<A> h = s_v[rsub][q][b + len - 1];
                    if (len == uint32_t(kBlockCh)) {
#pragma unroll
                        for (int k = kBlockCh - 2; k >= 0; --k) h = horner_step(h, z, s_v[rsub][q][b + k]);
                    } else {
                        for (int k = int(len) - 2; k >= 0; --k) h = horner_step(h, z, s_v[rsub][q][b + k]);
                    }
                    acc[q].add(cmul(base, h));
                }
            }
        }
    }
    // Combine the RP timestep lanes of each direction in a fixed order.
    if constexpr (RP > 1) {
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            s_red[rsub][lane_d][q][0] = acc[q].re();
            s_red[rsub][lane_d][q][1] = acc[q].im();
        }
        __syncthreads();
        if (rsub == 0 && dir_ok) {
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                double re = 0.0, im = 0.0;
                for (int r = 0; r < RP; ++r) {
                    re += s_red[r][lane_d][q][0];
                    im += s_red[r][lane_d][q][1];
                }
                double* o = avg + (size_t(d) * nbl + bl) * P * 2 + 2 * q;
                o[0] = re;
                o[1] = im;
            }
            if constexpr (NP == 1) {
                double* o = avg + (size_t(d) * nbl + bl) * P * 2;
                for (uint32_t q = 2; q < 2 * P; ++q) o[q] = 0.0;  // pols not accumulated
            }
        }
    } else {
        if (dir_ok) {
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                double* o = avg + (size_t(d) * nbl + bl) * P * 2 + 2 * q;
                o[0] = acc[q].re();
                o[1] = acc[q].im();
            }
            if constexpr (NP == 1) {
                double* o = avg + (size_t(d) * nbl + bl) * P * 2;
                for (uint32_t q = 2; q < 2 * P; ++q) o[q] = 0.0;  // pols not accumulated
            }
        }
    }
}

// Bytes of raw element data a thread holds in flight while staging a chunk.
constexpr int kStageBytes = 64;

// Stages the pol-0 values of a chunk's RP x CC (row, channel) elements into
// LDS as FP64 complex, 0 where the element lies past the rows (tb + r >= T)
// or channels (c >= cn) or any of its pols is not finite.  Thread i takes
// elements i, i + 256, ...; they are taken EB at a time, and every pol of the
// EB elements is loaded before any is checked, so all their loads are in
// flight together (EB * 2 * PP * sizeof(VT) <= kStageBytes where the chunk
// allows).  Slots past the rows or channels read a valid element (row 0 of the
// baseline, channel 0) and store zero.
template <typename VT, int PP, int RP, int CC, typename SV>
__device__ __attribute__((always_inline)) inline void stage_pol0(SV& s_v, const LEAP_GLOBAL VT* __restrict__ vis,
                                                                 uint32_t tb, uint32_t T, uint32_t nbl, uint32_t bl,
                                                                 uint32_t C, uint32_t c0, uint32_t cn) {
    using SU = StageUnit<VT, PP>;
    using U = typename SU::T;
    constexpr int E = RP * CC / kThreads;
    static_assert(RP * CC % kThreads == 0, "whole elements per thread");
    // EB: the largest power of two <= E whose batch fits kStageBytes (E is a power of two)
    constexpr int fit = kStageBytes / int(2 * PP * sizeof(VT));
    constexpr int EB0 = fit >= 8 ? 8 : fit >= 4 ? 4 : fit >= 2 ? 2 : 1;
    constexpr int EB = EB0 > E ? E : EB0;
    static_assert((E & (E - 1)) == 0 && E % EB == 0, "whole batches");
    // One batch at a time: a rolled loop keeps the next batch's loads from being
    // hoisted over this one (registers for one batch only).
#pragma unroll 1
    for (int e0 = 0; e0 < E; e0 += EB) {
        U raw[EB][SU::N];
#pragma unroll
        for (int e = 0; e < EB; ++e) {
            const uint32_t i = threadIdx.x + kThreads * (e0 + e);
            const uint32_t r = i / CC, c = i - r * CC;
            const bool ok = tb + r < T && c < cn;
            const size_t el = ok ? (size_t(tb + r) * nbl + bl) * C + c0 + c : size_t(bl) * C;
            const LEAP_GLOBAL U* p = reinterpret_cast<const LEAP_GLOBAL U*>(vis + el * (2 * PP));
#pragma unroll
            for (int k = 0; k < SU::N; ++k) raw[e][k] = p[k];
        }
#pragma unroll
        for (int e = 0; e < EB; ++e) {
            const uint32_t i = threadIdx.x + kThreads * (e0 + e);
            const uint32_t r = i / CC, c = i - r * CC;
            const bool ok = tb + r < T && c < cn;
            bool finite = true;
#pragma unroll
            for (int k = 0; k < SU::N; ++k)
#pragma unroll
                for (int m = 0; m < SU::K; ++m) finite = finite && __builtin_isfinite(raw[e][k][m]);
            s_v[r][0][c] = (ok && finite) ? C2<double>{double(raw[e][0][0]), double(raw[e][0][1])}
                                          : C2<double>{0.0, 0.0};
        }
    }
}

// Channels per LDS chunk of the FP64 Horner tiles with RP rows in flight:
// pol-0 tiles take 256 (row, channel) elements per chunk (at least 16
// channels), so the raw copy of a chunk (all pols, up to 64 bytes an element)
// is at most 16 KB, with kDmaSlots of them in flight; tiles keeping every pol
// take 64 channels.  The recurrences run over the chunks in order, so the
// chunk size does not change the result.
constexpr int kDmaElems = 256;  // (row, channel) elements per pol-0 chunk
constexpr int kDmaSlots = 2;    // chunks copied ahead
constexpr int horner64_chunk(int RP, int NP) {
    return NP == 1 ? (kDmaElems / RP > 16 ? kDmaElems / RP : 16) : 64;
}

// Raw copy of one chunk, as the cube holds it, in 16-byte units.
template <int BYTES>
struct Horner64Raw {
    f32x4 u[BYTES / 16];
};

// Bytes of the raw copy of a pol-0 tile's chunk for components of ES bytes
// (up to 4 pols); tiles keeping every pol stage through registers.
constexpr int horner64_raw_bytes(int RP, int CC, int NP, int ES) { return NP == 1 ? kDmaSlots * RP * CC * 8 * ES : 16; }

// Full-row FP64 Horner: per (row, direction) one exact phasor exp(i sf kappa0)
// and the step z = exp(i sf dkappa) are computed with FP64 sincos, and the
// whole channel axis is evaluated as P(z) = P_even(z^2) + z P_odd(z^2) by two
// interleaved Horner chains (independent, so the FP64 FMA pipeline stays
// busy).  Horner's error grows as n * eps (Higham, Accuracy and Stability of
// Numerical Algorithms, 5.1), about 1e-13 relative for n = 384 in FP64, so no
// re-anchoring of the phasor is needed.  On gfx950 FP64 FMA issues at the
// same rate as non-packed FP32 FMA, so this costs no more than an FP32 chain.
// Channel chunks are staged into LDS from the last to the first, and the
// chains carry across chunks.
// Each lane carries ND directions of its row, so one LDS read of a channel
// feeds 2 * ND independent chains (LDS return bandwidth, not the FMA pipe,
// limits a one-direction-per-lane layout).  A workgroup covers DB directions
// with L = DB / ND lanes per row and RP = 256 / L rows in flight.
template <int DB, int ND, int NP, int CC>
struct Horner64Smem {
    static_assert(CC % 2 == 0, "chunks hold whole channel pairs");
    static_assert(DB % ND == 0 && kThreads % (DB / ND) == 0, "lane layout");
    static constexpr int L = DB / ND;
    static constexpr int RP = kThreads / L;
    typedef double Red[(RP > 1) ? RP : 1][L][ND][NP][2];  // per-row sums of each direction
    // With fewer than 16 lanes per row a wave's 16-lane group of a 16-byte read spans two
    // rows, whose entries CC * 16 bytes apart share LDS banks: one padding element staggers them.
    static constexpr int kRowPad = L < 16 ? 1 : 0;
    C2<double> v[RP][NP][CC + kRowPad];
    // pol-0 tiles combine their row sums in the raw chunk buffer, idle once the last chunk is in
    double red[NP == 1 ? 1 : sizeof(Red) / sizeof(double)];
};


// One (baseline, direction block) tile; every thread of the workgroup calls it.
// QOUT (pol 0 only): every row's sums are also written to
// qsum[row][d] = (e0 P(z), e0 z P'(z)), the second from derivative chains run
// beside the Horner chains (E' <- E' y + E before E <- E y + v, y = z^2):
// z P'(z) = 2 y E'(y) + z (O(y) + 2 y O'(y)).  The residual-loss adjoint reads
// them instead of the cube.
// Pol-0 tiles on a cube of 16-byte-multiple elements (complex128, or complex64
// with 2 or 4 pols) at a 16-byte-aligned address copy chunks into a ring of
// kDmaSlots raw buffers with dma_chunk, kDmaSlots chunks ahead of the chains,
// so the cube streams in behind the arithmetic and takes no registers; other
// cubes are staged through registers (stage_pol0).  The pol-0 tiles combine
// their row sums in the raw buffer, which is idle once the last chunk is in.
template <typename VT, int DB, int ND, int NP, int CC, bool QOUT = false, int RB = 16>
__device__ __attribute__((always_inline)) inline void horner64_tile(
    Horner64Smem<DB, ND, NP, CC>& sm, Horner64Raw<RB>& sraw, const LEAP_GLOBAL VT* __restrict__ vis, const LEAP_GLOBAL double* __restrict__ uvw,
    const LEAP_GLOBAL double* __restrict__ dd, uint32_t D, uint32_t T, uint32_t nbl, uint32_t C, uint32_t P,
    double kappa0, double dkappa, LEAP_GLOBAL double* __restrict__ avg, uint32_t bl, uint32_t dblock,
    LEAP_GLOBAL double* __restrict__ qsum = nullptr) {
#pragma clang fp contract(off)
    static_assert(!QOUT || NP == 1, "row sums of pol 0");
    constexpr int L = DB / ND;
    constexpr int RP = kThreads / L;
    auto& s_v = sm.v;
    using Red = typename Horner64Smem<DB, ND, NP, CC>::Red;
    static_assert(NP > 1 || sizeof(Red) <= RB, "row sums fit the raw buffer");
    Red& s_red = *reinterpret_cast<Red*>(NP == 1 ? static_cast<void*>(&sraw.u[0]) : static_cast<void*>(&sm.red[0]));

    const uint32_t lane_d = threadIdx.x % L;
    const uint32_t rsub = threadIdx.x / L;
    uint32_t dir[ND];
    bool dir_ok[ND];
    double dz[ND][3];
    bool any_dir = false;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        dir[j] = dblock * DB + lane_d + j * L;
        dir_ok[j] = dir[j] < D;
        any_dir |= dir_ok[j];
        const uint32_t dc = dir_ok[j] ? dir[j] : 0;
        dz[j][0] = dd[9 * dc + 6];
        dz[j][1] = dd[9 * dc + 7];
        dz[j][2] = dd[9 * dc + 8];
    }
    const uint32_t n_chunks = (C + CC - 1) / CC;

    // Pol-0 tiles stream accepted cubes through the ring (cube_ring.h), kDmaSlots chunks ahead.
    using Ring = CubeRing<VT, RP, CC, kDmaSlots>;
    static_assert(NP > 1 || Ring::kBytes <= RB, "the ring fits the raw buffer");
    Ring ring;
    bool dma = false;
    if constexpr (NP == 1) dma = Ring::accepts((const void*)vis, P);
    // address-space cast of the shared buffer: the copies take an LDS pointer
    if constexpr (NP == 1)
        if (dma) ring.start((LEAP_LDS f32x4*)(&sraw.u[0]), vis, P, T, nbl, bl, C);

    C2<double> acc[ND][NP];
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int q = 0; q < NP; ++q) acc[j][q] = {0.0, 0.0};

    for (uint32_t tb = 0; tb < T; tb += RP) {
        const uint32_t t = tb + rsub;
        const bool row_ok = t < T;
        double u = 0.0, v = 0.0, w = 0.0;
        if (row_ok) {
            const size_t row = size_t(t) * nbl + bl;
            u = uvw[3 * row];
            v = uvw[3 * row + 1];
            w = uvw[3 * row + 2];
        }
        double sf[ND];
        C2<double> z[ND], z2[ND];
        C2<double> he[ND][NP], ho[ND][NP];
        C2<double> de[QOUT ? ND : 1], dox[QOUT ? ND : 1];  // derivative chains (QOUT)
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            sf[j] = shift_factor(dz[j], u, v, w);
            z[j] = phasor_f64(sf[j] * dkappa);
            z2[j] = cmul(z[j], z[j]);
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                he[j][q] = {0.0, 0.0};
                ho[j][q] = {0.0, 0.0};
            }
            if constexpr (QOUT) {
                de[j] = {0.0, 0.0};
                dox[j] = {0.0, 0.0};
            }
        }
        for (int ch = int(n_chunks) - 1; ch >= 0; --ch) {
            const uint32_t c0 = uint32_t(ch) * CC;
            const uint32_t cn = min(uint32_t(CC), C - c0);
            bool streamed = false;
            if constexpr (NP == 1) {
                if (dma) {
                    // This chunk's copy has landed (later copies may stay in flight) and the
                    // previous chunk's chains are done.
                    const LEAP_LDS f32x4* slot = ring.next_slot();
                    __syncthreads();
                    ring.convert(slot, tb, cn, [&](uint32_t r, uint32_t c, double re, double im) {
                        s_v[r][0][c] = C2<double>{re, im};
                    });
                    __syncthreads();
                    ring.issue();  // into the slot just converted
                    streamed = true;
                }
            }
            if (!streamed) {
                __syncthreads();
                // Pol 0 of every element of the chunk, gathered a batch at a time (stage_pol0);
                // tiles that keep several pols stage element by element.
                bool staged = false;
                if constexpr (NP == 1 && RP * CC % kThreads == 0) {
                    staged = true;
                    switch (P) {
                        case 1: stage_pol0<VT, 1, RP, CC>(s_v, vis, tb, T, nbl, bl, C, c0, cn); break;
                        case 2: stage_pol0<VT, 2, RP, CC>(s_v, vis, tb, T, nbl, bl, C, c0, cn); break;
                        case 3: stage_pol0<VT, 3, RP, CC>(s_v, vis, tb, T, nbl, bl, C, c0, cn); break;
                        case 4: stage_pol0<VT, 4, RP, CC>(s_v, vis, tb, T, nbl, bl, C, c0, cn); break;
                        default: staged = false; break;
                    }
                }
                if (!staged) {
                    for (uint32_t i = threadIdx.x; i < RP * uint32_t(CC); i += kThreads) {
                        const uint32_t r = i / CC, c = i - r * CC;
                        const uint32_t tt = tb + r;
                        C2<double> val[NP];
#pragma unroll
                        for (int q = 0; q < NP; ++q) val[q] = {0.0, 0.0};
                        if (tt < T && c < cn) {
                            const size_t row = size_t(tt) * nbl + bl;
                            const LEAP_GLOBAL VT* p = vis + ((row * C + c0 + c) * P) * 2;
                            bool finite = true;
                            VT x[8];
                            for (uint32_t q = 0; q < P; ++q) {
                                x[2 * q] = p[2 * q];
                                x[2 * q + 1] = p[2 * q + 1];
                                finite = finite && __builtin_isfinite(x[2 * q]) && __builtin_isfinite(x[2 * q + 1]);
                            }
                            if (finite) {
#pragma unroll
                                for (int q = 0; q < NP; ++q) val[q] = {double(x[2 * q]), double(x[2 * q + 1])};
                            }
                        }
#pragma unroll
                        for (int q = 0; q < NP; ++q) s_v[r][q][c] = val[q];
                    }
                }
                __syncthreads();
            }
            if (!row_ok || !any_dir) continue;
            // Every chunk runs all CC / 2 pairs.  Only a row's first chunk (the last channels)
            // can be partial; its entries past cn are zero in LDS and its chains start at +0,
            // which zero entries leave at +0, so the padding changes no bit of the result.
#pragma unroll
            for (int q = 0; q < NP; ++q) {
#pragma unroll 4
                for (int k = CC / 2 - 1; k >= 0; --k) {
                    const C2<double> ve = s_v[rsub][q][2 * k], vo = s_v[rsub][q][2 * k + 1];
#pragma unroll
                    for (int j = 0; j < ND; ++j) {
                        if constexpr (QOUT) {
                            de[j] = horner_step(de[j], z2[j], he[j][q]);
                            dox[j] = horner_step(dox[j], z2[j], ho[j][q]);
                        }
                        he[j][q] = horner_step(he[j][q], z2[j], ve);
                        ho[j][q] = horner_step(ho[j][q], z2[j], vo);
                    }
                }
            }
        }
        if (row_ok) {
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                const C2<double> e0 = phasor_f64(sf[j] * kappa0);
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    const C2<double> zo = cmul(z[j], ho[j][q]);
                    const C2<double> h = {he[j][q].x + zo.x, he[j][q].y + zo.y};
                    const C2<double> r = cmul(e0, h);
                    acc[j][q].x += r.x;
                    acc[j][q].y += r.y;
                    if constexpr (QOUT) {
                        if (dir_ok[j]) {
                            const C2<double> ye = cmul(z2[j], de[j]), yo = cmul(z2[j], dox[j]);
                            const C2<double> zi = cmul(z[j], C2<double>{ho[j][q].x + 2.0 * yo.x, ho[j][q].y + 2.0 * yo.y});
                            const C2<double> r2 = cmul(e0, C2<double>{2.0 * ye.x + zi.x, 2.0 * ye.y + zi.y});
                            LEAP_GLOBAL double* qo = qsum + ((size_t(t) * nbl + bl) * D + dir[j]) * 4;
                            qo[0] = r.x;
                            qo[1] = r.y;
                            qo[2] = r2.x;
                            qo[3] = r2.y;
                        }
                    }
                }
            }
        }
    }
    if constexpr (RP > 1) {
#pragma unroll
        for (int j = 0; j < ND; ++j)
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                s_red[rsub][lane_d][j][q][0] = acc[j][q].x;
                s_red[rsub][lane_d][j][q][1] = acc[j][q].y;
            }
        __syncthreads();
        if (rsub == 0) {
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                if (!dir_ok[j]) continue;
                LEAP_GLOBAL double* o = avg + (size_t(dir[j]) * nbl + bl) * P * 2;
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    double re = 0.0, im = 0.0;
                    for (int r = 0; r < RP; ++r) {
                        re += s_red[r][lane_d][j][q][0];
                        im += s_red[r][lane_d][j][q][1];
                    }
                    o[2 * q] = re;
                    o[2 * q + 1] = im;
                }
                for (uint32_t q = 2 * NP; q < 2 * P; ++q) o[q] = 0.0;  // pols not accumulated
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            if (!dir_ok[j]) continue;
            LEAP_GLOBAL double* o = avg + (size_t(dir[j]) * nbl + bl) * P * 2;
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                o[2 * q] = acc[j][q].x;
                o[2 * q + 1] = acc[j][q].y;
            }
            for (uint32_t q = 2 * NP; q < 2 * P; ++q) o[q] = 0.0;
        }
    }
}

template <typename VT, int DB, int ND, int NP, int CC>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NP == 1 && ND <= 2 && DB / ND >= 16 ? 4 : 1, 8))) void rotate_average_horner64(
    const VT* __restrict__ vis, const double* __restrict__ uvw, const double* __restrict__ dd,
    uint32_t D, uint32_t T, uint32_t nbl, uint32_t C, uint32_t P, double kappa0, double dkappa,
    double* __restrict__ avg) {
    __shared__ Horner64Smem<DB, ND, NP, CC> sm;
    constexpr int RB = horner64_raw_bytes(kThreads / (DB / ND), CC, NP, sizeof(VT));
    __shared__ Horner64Raw<RB> sraw;
    horner64_tile<VT, DB, ND, NP, CC, false, RB>(sm, sraw, (const LEAP_GLOBAL VT*)vis, (const LEAP_GLOBAL double*)uvw,
                                      (const LEAP_GLOBAL double*)dd, D, T, nbl, C, P, kappa0, dkappa,
                                      (LEAP_GLOBAL double*)avg, blockIdx.x, blockIdx.y);
}

// The residual-loss forward pass: avgData and the per-row sums q (QOUT above).
template <typename VT, int DB, int ND, int CC>
__global__ __launch_bounds__(kThreads) void rotate_average_horner64_q(
    const VT* __restrict__ vis, const double* __restrict__ uvw, const double* __restrict__ dd, uint32_t D, uint32_t T,
    uint32_t nbl, uint32_t C, uint32_t P, double kappa0, double dkappa, double* __restrict__ avg,
    double* __restrict__ q) {
    __shared__ Horner64Smem<DB, ND, 1, CC> sm;
    constexpr int RB = horner64_raw_bytes(kThreads / (DB / ND), CC, 1, sizeof(VT));
    __shared__ Horner64Raw<RB> sraw;
    horner64_tile<VT, DB, ND, 1, CC, true, RB>(sm, sraw, (const LEAP_GLOBAL VT*)vis, (const LEAP_GLOBAL double*)uvw,
                                           (const LEAP_GLOBAL double*)dd, D, T, nbl, C, P, kappa0, dkappa,
                                           (LEAP_GLOBAL double*)avg, blockIdx.x, blockIdx.y, (LEAP_GLOBAL double*)q);
}

// Several measurement sets in one persistent launch: workgroup g runs the
// (set, baseline, direction block) items items[off[g] .. off[g + 1]) planned on
// the host, each against its set's descriptor.
template <int DB, int ND>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(ND <= 2 ? 4 : 1, 8))) void rotate_average_multi(const MultiSetDesc* __restrict__ descs,
                                                                 const uint2* __restrict__ items,
                                                                 const uint32_t* __restrict__ off) {
    constexpr int RP = kThreads / (DB / ND);
    constexpr int CC = horner64_chunk(RP, 1);
    constexpr int RB = horner64_raw_bytes(RP, CC, 1, sizeof(double));
    __shared__ Horner64Smem<DB, ND, 1, CC> sm;
    __shared__ Horner64Raw<RB> sraw;
    const uint32_t i1 = off[blockIdx.x + 1];
    for (uint32_t i = off[blockIdx.x]; i < i1; ++i) {
        const uint2 it = items[i];
        const MultiSetDesc& m = descs[it.x];
        const uint32_t bl = it.y % m.nbl, db = it.y / m.nbl;
        const LEAP_GLOBAL double* uvw = (const LEAP_GLOBAL double*)m.uvw;
        const LEAP_GLOBAL double* dd = (const LEAP_GLOBAL double*)m.dd;
        LEAP_GLOBAL double* avg = (LEAP_GLOBAL double*)m.avg;
        if (m.c64)
            horner64_tile<float, DB, ND, 1, CC, false, RB>(sm, sraw, (const LEAP_GLOBAL float*)m.vis, uvw, dd, m.D,
                                                           m.T, m.nbl, m.C, m.P, m.kappa0, m.dkappa, avg, bl, db);
        else
            horner64_tile<double, DB, ND, 1, CC, false, RB>(sm, sraw, (const LEAP_GLOBAL double*)m.vis, uvw, dd,
                                                            m.D, m.T, m.nbl, m.C, m.P, m.kappa0, m.dkappa, avg, bl,
                                                            db);
        __syncthreads();
    }
}

// ---- packed pol-0 path ------------------------------------------------------
// s_waitcnt lgkmcnt(0) with the vector-memory and export counters left free
// (gfx9 encoding: vmcnt 63, expcnt 7, lgkmcnt 0).
constexpr int kWaitScalarMem = 0xC07F;

// The pol-0 visibility of every (row, channel) as FP64 complex, 0 where any
// polarisation of that (row, channel) is not finite (the reference skips the
// whole (row, channel) then, CpuLeapCalibrator.cc:201-214).  HBM-bound: reads
// the cube once, writes 16 B per (row, channel).  SCALED multiplies each value
// by scale[channel] (the NUFFT precorrection); element i is channel i % C.
template <typename VT, int P, bool SCALED>
__global__ __launch_bounds__(kThreads) void pack_pol0_kernel(const VT* __restrict__ vis, size_t n,
                                                             C2<double>* __restrict__ out,
                                                             const double* __restrict__ scale, uint32_t C) {
    // Four independent (row, channel)s per thread and step keep eight loads in
    // flight per thread; consecutive threads take consecutive elements.
    constexpr int U = 4;
    const size_t stride = size_t(gridDim.x) * kThreads;
    size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    // channel of element i and the channel step of one stride (SCALED only)
    uint32_t ch = 0, ch_step = 0;
    if constexpr (SCALED) {
        ch = uint32_t(i % C);
        ch_step = uint32_t(stride % C);
    }
    auto next_ch = [&](uint32_t c) {
        c += ch_step;
        return c >= C ? c - C : c;
    };
    auto value = [&](const VT* x, uint32_t c) {
        bool finite = true;
#pragma unroll
        for (int k = 0; k < 2 * P; ++k) finite = finite && __builtin_isfinite(x[k]);
        C2<double> r = finite ? C2<double>{double(x[0]), double(x[1])} : C2<double>{0.0, 0.0};
        if constexpr (SCALED) {
            const double f = scale[c];
            r.x *= f;
            r.y *= f;
        }
        return r;
    };
    for (; i + (U - 1) * stride < n; i += U * stride) {
        VT x[U][2 * P];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < 2 * P; ++k) x[u][k] = vis[(i + u * stride) * (2 * P) + k];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            out[i + u * stride] = value(x[u], ch);
            if constexpr (SCALED) ch = next_ch(ch);
        }
    }
    for (; i < n; i += stride) {
        VT x[2 * P];
#pragma unroll
        for (int k = 0; k < 2 * P; ++k) x[k] = vis[i * (2 * P) + k];
        out[i] = value(x, ch);
        if constexpr (SCALED) ch = next_ch(ch);
    }
}

// x[i] *= scale[i % C] over n FP64 complex values.
__global__ __launch_bounds__(kThreads) void scale_channels_kernel(C2<double>* __restrict__ x, size_t n,
                                                                  const double* __restrict__ scale, uint32_t C) {
    const size_t stride = size_t(gridDim.x) * kThreads;
    size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    uint32_t ch = uint32_t(i % C);
    const uint32_t step = uint32_t(stride % C);
    for (; i < n; i += stride) {
        const double f = scale[ch];
        C2<double> v = x[i];
        v.x *= f;
        v.y *= f;
        x[i] = v;
        ch += step;
        if (ch >= C) ch -= C;
    }
}

// complex64 -> complex128, for pol-0 chunks packed on the host.
__global__ __launch_bounds__(kThreads) void widen_c64_kernel(const C2<float>* __restrict__ in, size_t n,
                                                             C2<double>* __restrict__ out) {
    const size_t stride = size_t(gridDim.x) * kThreads;
    for (size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride)
        out[i] = C2<double>{double(in[i].x), double(in[i].y)};
}

// Full-row FP64 Horner over packed pol-0 rows.  One wave per row: the row's
// channel values are wave-uniform, so they are read with scalar loads into
// SGPRs and feed the FMAs directly (no LDS traffic, no VGPRs for them); each
// lane carries ND directions, a workgroup 64 * ND directions of one baseline
// over 4 rows in flight.  Same operation order as rotate_average_horner64
// with 64 lanes per row, hence the same results.
template <int ND>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(ND == 4 ? 4 : 1, 8))) void rotate_average_packed(
    const C2<double>* __restrict__ w, const double* __restrict__ uvw, const double* __restrict__ dd, uint32_t D,
    uint32_t T, uint32_t nbl, uint32_t C, uint32_t P, double kappa0, double dkappa, double* __restrict__ avg,
    int32_t accumulate) {
#pragma clang fp contract(off)
    constexpr int L = 64;
    constexpr int RP = kThreads / L;
    constexpr int DB = L * ND;
    // Per-lane state touched once per row lives in LDS (read through volatile
    // pointers so it is not promoted back to registers): the direction's dd
    // row 2 and the running row sums.  The registers hold only what the
    // channel loop needs, for more waves per SIMD.
    __shared__ double s_dz[L][ND][3];
    __shared__ double s_red[RP][L][ND][2];
    const uint32_t lane = threadIdx.x % L;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / L);
    const uint32_t bl = blockIdx.x;
    if (wave == 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            const uint32_t d = blockIdx.y * DB + lane + j * L;
            const uint32_t dc = d < D ? d : 0;
            s_dz[lane][j][0] = dd[9 * dc + 6];
            s_dz[lane][j][1] = dd[9 * dc + 7];
            s_dz[lane][j][2] = dd[9 * dc + 8];
        }
    }
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        s_red[wave][lane][j][0] = 0.0;
        s_red[wave][lane][j][1] = 0.0;
    }
    __syncthreads();
    volatile double* vdz = &s_dz[lane][0][0];
    volatile double* vacc = &s_red[wave][lane][0][0];
    const uint32_t npairs = C / 2;
    for (uint32_t t = wave; t < T; t += RP) {
        const size_t row = size_t(t) * nbl + bl;
        const double u = uvw[3 * row], v = uvw[3 * row + 1], ww = uvw[3 * row + 2];
        const C2<double>* wr = w + row * C;
        double sf[ND];
        C2<double> z[ND], z2[ND], he[ND], ho[ND];
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            const double dz[3] = {vdz[3 * j], vdz[3 * j + 1], vdz[3 * j + 2]};
            sf[j] = shift_factor(dz, u, v, ww);
            z[j] = phasor_f64(sf[j] * dkappa);
            z2[j] = cmul(z[j], z[j]);
            he[j] = {0.0, 0.0};
            ho[j] = {0.0, 0.0};
        }
        if (C & 1) {  // the top channel pairs with a zero
            const C2<double> ve = wr[C - 1];
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                he[j] = horner_step(he[j], z2[j], ve);
                ho[j] = horner_step(ho[j], z2[j], C2<double>{0.0, 0.0});
            }
        }
        // Blocks of 4 channel pairs in two SGPR buffers: wait for the current
        // block, issue the next block's scalar loads, then run the current
        // block, so the loads of one block overlap the FMAs of the other (a
        // scalar-memory wait covers every outstanding scalar load).
        uint32_t k = npairs;
        C2<double> ba[8], bb[8];
        auto load_block = [&](C2<double>* b, uint32_t top) {
#pragma unroll
            for (int i = 0; i < 8; ++i) b[i] = wr[2 * (top - 4) + i];
        };
        auto run_block = [&](const C2<double>* b) {
#pragma unroll
            for (int i = 3; i >= 0; --i) {
#pragma unroll
                for (int j = 0; j < ND; ++j) {
                    he[j] = horner_step(he[j], z2[j], b[2 * i]);
                    ho[j] = horner_step(ho[j], z2[j], b[2 * i + 1]);
                }
            }
        };
        if (k >= 4) load_block(ba, k);
        while (k >= 4) {
            __builtin_amdgcn_s_waitcnt(kWaitScalarMem);
            if (k >= 8) load_block(bb, k - 4);
            run_block(ba);
            k -= 4;
            if (k < 4) break;
            __builtin_amdgcn_s_waitcnt(kWaitScalarMem);
            if (k >= 8) load_block(ba, k - 4);
            run_block(bb);
            k -= 4;
        }
        for (; k > 0; --k) {
            const C2<double> ve = wr[2 * (k - 1)], vo = wr[2 * (k - 1) + 1];
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                he[j] = horner_step(he[j], z2[j], ve);
                ho[j] = horner_step(ho[j], z2[j], vo);
            }
        }
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            const C2<double> e0 = phasor_f64(sf[j] * kappa0);
            const C2<double> zo = cmul(z[j], ho[j]);
            const C2<double> h = {he[j].x + zo.x, he[j].y + zo.y};
            const C2<double> r = cmul(e0, h);
            vacc[2 * j] = vacc[2 * j] + r.x;
            vacc[2 * j + 1] = vacc[2 * j + 1] + r.y;
        }
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            const uint32_t dj = blockIdx.y * DB + lane + j * L;
            if (dj >= D) continue;
            double re = 0.0, im = 0.0;
            for (int r = 0; r < RP; ++r) {
                re += s_red[r][lane][j][0];
                im += s_red[r][lane][j][1];
            }
            double* o = avg + (size_t(dj) * nbl + bl) * P * 2;
            if (accumulate) {
                o[0] += re;
                o[1] += im;
            } else {
                o[0] = re;
                o[1] = im;
                for (uint32_t q = 2; q < 2 * P; ++q) o[q] = 0.0;  // pols not accumulated
            }
        }
    }
}

// A == double selects the full-row FP64 Horner kernel (ND directions per
// lane); A == float the FP32 blocked kernel (one direction per lane,
// 32-channel Horner blocks re-anchored by FP64-reduced FP32 sincos, Kahan sums).
template <typename VT, typename A, int DB, int ND, int NP>
void launch_one(hipStream_t s, const void* vis, const double* uvw, const double* dd, uint32_t D, uint32_t T,
                uint32_t nbl, uint32_t nbl_grid, uint32_t C, uint32_t P, double kappa0, double dkappa, double* avg) {
    constexpr int RP = kThreads / (DB / ND);
    constexpr int CC = sizeof(A) == 8 ? horner64_chunk(RP, NP) : (NP == 1) ? (2048 / RP < 256 ? 2048 / RP : 256) : 64;
    const dim3 grid(nbl_grid, (D + DB - 1) / DB);
    if constexpr (sizeof(A) == 8) {
        rotate_average_horner64<VT, DB, ND, NP, CC><<<grid, kThreads, 0, s>>>(
            static_cast<const VT*>(vis), uvw, dd, D, T, nbl, C, P, kappa0, dkappa, avg);
    } else {
        rotate_average_batched<VT, A, DB, NP, CC><<<grid, kThreads, 0, s>>>(
            static_cast<const VT*>(vis), uvw, dd, D, T, nbl, C, P, kappa0, dkappa, avg);
    }
    LEAP_HIP(hipGetLastError());
}

// Tile code: DB * 10 + ND (directions per workgroup, directions per lane).
template <typename VT, typename A, int NP>
void launch_db(hipStream_t s, int tile, const void* vis, const double* uvw, const double* dd, uint32_t D,
               uint32_t T, uint32_t nbl, uint32_t nbl_grid, uint32_t C, uint32_t P, double kappa0, double dkappa,
               double* avg) {
#define LEAP_TILE(DB_, ND_)                                                                                   \
    case DB_ * 10 + ND_:                                                                                      \
        launch_one<VT, A, DB_, ND_, NP>(s, vis, uvw, dd, D, T, nbl, nbl_grid, C, P, kappa0, dkappa, avg); \
        return;
    if constexpr (sizeof(A) == 8) {
        if constexpr (NP == 1) {
            // 16 directions per workgroup (internal: the residual-loss forward pass)
            if (tile == 161) {
                launch_one<VT, A, 16, 1, NP>(s, vis, uvw, dd, D, T, nbl, nbl_grid, C, P, kappa0, dkappa, avg);
                return;
            }
            if (tile == 162) {
                launch_one<VT, A, 16, 2, NP>(s, vis, uvw, dd, D, T, nbl, nbl_grid, C, P, kappa0, dkappa, avg);
                return;
            }
        }
        switch (tile) {
            LEAP_TILE(32, 1)
            LEAP_TILE(64, 1)
            LEAP_TILE(64, 2)
            LEAP_TILE(128, 1)
            LEAP_TILE(128, 2)
            LEAP_TILE(128, 4)
            LEAP_TILE(256, 1)
            LEAP_TILE(256, 2)
            LEAP_TILE(256, 4)
            default: break;
        }
    } else {
        switch (tile) {
            LEAP_TILE(32, 1)
            LEAP_TILE(64, 1)
            LEAP_TILE(128, 1)
            LEAP_TILE(256, 1)
            default: break;
        }
    }
#undef LEAP_TILE
    throw Error(LEAP_ERR_INVALID_ARGUMENT, "unsupported batched tile " + std::to_string(tile));
}

template <typename VT, typename A>
void launch_np(hipStream_t s, int db, bool all_pols, const void* vis, const double* uvw, const double* dd,
               uint32_t D, uint32_t T, uint32_t nbl, uint32_t nbl_grid, uint32_t C, uint32_t P, double kappa0, double dkappa,
               double* avg) {
    if (!all_pols || P == 1) {
        launch_db<VT, A, 1>(s, db, vis, uvw, dd, D, T, nbl, nbl_grid, C, P, kappa0, dkappa, avg);
        return;
    }
    switch (P) {
        case 2: launch_db<VT, A, 2>(s, db, vis, uvw, dd, D, T, nbl, nbl_grid, C, P, kappa0, dkappa, avg); break;
        case 3: launch_db<VT, A, 3>(s, db, vis, uvw, dd, D, T, nbl, nbl_grid, C, P, kappa0, dkappa, avg); break;
        case 4: launch_db<VT, A, 4>(s, db, vis, uvw, dd, D, T, nbl, nbl_grid, C, P, kappa0, dkappa, avg); break;
        default: throw Error(LEAP_ERR_NOT_IMPLEMENTED, "unsupported polarisation count");
    }
}

}  // namespace

void launch_pack_pol0(hipStream_t s, const void* vis, int32_t vis_dtype, size_t n, uint32_t P, double* out,
                      const double* scale, uint32_t C) {
    if (n == 0) return;
    require(scale == nullptr || C > 0, LEAP_ERR_INTERNAL, "scaled pack needs the channel count");
    // 256 workgroups per CU: read/write streams reach ~5.2 TB/s on MI355X
    // (a device-to-device copy reaches 4.9 TB/s).
    const size_t want = (n + 4 * kThreads - 1) / (4 * kThreads);
    const uint32_t grid = uint32_t(std::min<size_t>(want, 256 * 256));
    C2<double>* o = reinterpret_cast<C2<double>*>(out);
#define LEAP_PACK(VT, NP_)                                                                                    \
    do {                                                                                                      \
        if (scale)                                                                                            \
            pack_pol0_kernel<VT, NP_, true><<<grid, kThreads, 0, s>>>(static_cast<const VT*>(vis), n, o, scale, C); \
        else                                                                                                  \
            pack_pol0_kernel<VT, NP_, false><<<grid, kThreads, 0, s>>>(static_cast<const VT*>(vis), n, o, nullptr, 0); \
    } while (0)
    const bool c64 = vis_dtype == LEAP_VIS_C64;
    switch (P) {
        case 1: if (c64) LEAP_PACK(float, 1); else LEAP_PACK(double, 1); break;
        case 2: if (c64) LEAP_PACK(float, 2); else LEAP_PACK(double, 2); break;
        case 3: if (c64) LEAP_PACK(float, 3); else LEAP_PACK(double, 3); break;
        case 4: if (c64) LEAP_PACK(float, 4); else LEAP_PACK(double, 4); break;
        default: throw Error(LEAP_ERR_NOT_IMPLEMENTED, "polarisation counts above 4 are not supported");
    }
#undef LEAP_PACK
    LEAP_HIP(hipGetLastError());
}

void launch_scale_channels(hipStream_t s, double* x, size_t n, const double* scale, uint32_t C) {
    if (n == 0) return;
    const uint32_t grid = uint32_t(std::min<size_t>((n + kThreads - 1) / kThreads, 256 * 64));
    scale_channels_kernel<<<grid, kThreads, 0, s>>>(reinterpret_cast<C2<double>*>(x), n, scale, C);
    LEAP_HIP(hipGetLastError());
}

void launch_widen_c64(hipStream_t s, const void* in, size_t n, double* out) {
    if (n == 0) return;
    const uint32_t grid = uint32_t(std::min<size_t>((n + kThreads - 1) / kThreads, 256 * 64));
    widen_c64_kernel<<<grid, kThreads, 0, s>>>(static_cast<const C2<float>*>(in), n, reinterpret_cast<C2<double>*>(out));
    LEAP_HIP(hipGetLastError());
}

bool packed_tile_supported(int32_t tile) { return tile == 641 || tile == 1282 || tile == 2564; }

void launch_rotate_packed(hipStream_t s, int32_t tile, const double* w, const double* uvw, const double* dd,
                          uint32_t D, uint32_t T, uint32_t nbl, uint32_t nbl_grid, uint32_t C, uint32_t P, double f0,
                          double df, double* avg, int32_t accumulate) {
    if (nbl_grid == 0 || D == 0) return;
    const double kappa0 = f0 / kSpeedOfLight;
    const double dkappa = df / kSpeedOfLight;
    const C2<double>* wp = reinterpret_cast<const C2<double>*>(w);
    const uint32_t db = uint32_t(tile / 10);
    const dim3 grid(nbl_grid, (D + db - 1) / db);
    switch (tile) {
        case 641: rotate_average_packed<1><<<grid, kThreads, 0, s>>>(wp, uvw, dd, D, T, nbl, C, P, kappa0, dkappa, avg, accumulate); break;
        case 1282: rotate_average_packed<2><<<grid, kThreads, 0, s>>>(wp, uvw, dd, D, T, nbl, C, P, kappa0, dkappa, avg, accumulate); break;
        case 2564: rotate_average_packed<4><<<grid, kThreads, 0, s>>>(wp, uvw, dd, D, T, nbl, C, P, kappa0, dkappa, avg, accumulate); break;
        default: throw Error(LEAP_ERR_INVALID_ARGUMENT, "unsupported packed tile " + std::to_string(tile));
    }
    LEAP_HIP(hipGetLastError());
}

void launch_rotate_multi(hipStream_t s, int32_t tile, const MultiSetDesc* descs, const uint2* items,
                         const uint32_t* off, uint32_t n_blocks) {
    if (n_blocks == 0) return;
    switch (tile) {
        case 2562: rotate_average_multi<256, 2><<<n_blocks, kThreads, 0, s>>>(descs, items, off); break;
        case 1282: rotate_average_multi<128, 2><<<n_blocks, kThreads, 0, s>>>(descs, items, off); break;
        case 1281: rotate_average_multi<128, 1><<<n_blocks, kThreads, 0, s>>>(descs, items, off); break;
        case 641: rotate_average_multi<64, 1><<<n_blocks, kThreads, 0, s>>>(descs, items, off); break;
        case 321: rotate_average_multi<32, 1><<<n_blocks, kThreads, 0, s>>>(descs, items, off); break;
        default: throw Error(LEAP_ERR_INVALID_ARGUMENT, "unsupported multi-set tile " + std::to_string(tile));
    }
    LEAP_HIP(hipGetLastError());
}

int32_t multi_blocks_per_cu(int32_t tile) {
    int n = 0;
    switch (tile) {
        case 2562: LEAP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, rotate_average_multi<256, 2>, kThreads, 0)); break;
        case 1282: LEAP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, rotate_average_multi<128, 2>, kThreads, 0)); break;
        case 1281: LEAP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, rotate_average_multi<128, 1>, kThreads, 0)); break;
        case 641: LEAP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, rotate_average_multi<64, 1>, kThreads, 0)); break;
        default: LEAP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, rotate_average_multi<32, 1>, kThreads, 0)); break;
    }
    return n > 0 ? n : 1;
}

void launch_rotate_batched(hipStream_t s, int32_t precision, int32_t direction_block, bool all_pols,
                           const void* vis, int32_t vis_dtype, const double* uvw, const double* dd, uint32_t D,
                           uint32_t T, uint32_t nbl, uint32_t nbl_grid, uint32_t C, uint32_t P, double f0,
                           double df, double* avg, int32_t fp32_blocks) {
    if (nbl_grid == 0 || D == 0) return;
    const double kappa0 = f0 / kSpeedOfLight;
    const double dkappa = df / kSpeedOfLight;
    const bool c64 = vis_dtype == LEAP_VIS_C64;
    if (precision == LEAP_PRECISION_MIXED && fp32_blocks) {
        if (c64)
            launch_np<float, float>(s, direction_block, all_pols, vis, uvw, dd, D, T, nbl, nbl_grid, C, P, kappa0, dkappa, avg);
        else
            launch_np<double, float>(s, direction_block, all_pols, vis, uvw, dd, D, T, nbl, nbl_grid, C, P, kappa0, dkappa, avg);
    } else {
        if (c64)
            launch_np<float, double>(s, direction_block, all_pols, vis, uvw, dd, D, T, nbl, nbl_grid, C, P, kappa0, dkappa, avg);
        else
            launch_np<double, double>(s, direction_block, all_pols, vis, uvw, dd, D, T, nbl, nbl_grid, C, P, kappa0, dkappa, avg);
    }
}

void launch_rotate_forward_q(hipStream_t s, int32_t tile, const void* vis, int32_t vis_dtype, const double* uvw,
                             const double* dd, uint32_t D, uint32_t T, uint32_t nbl, uint32_t C, uint32_t P, double f0,
                             double df, double* avg, double* q) {
    if (nbl == 0 || D == 0) return;
    const double kappa0 = f0 / kSpeedOfLight;
    const double dkappa = df / kSpeedOfLight;
    const bool c64 = vis_dtype == LEAP_VIS_C64;
#define LEAP_FQ(DB_, ND_)                                                                                         \
    case DB_ * 10 + ND_: {                                                                                        \
        constexpr int RP = kThreads / (DB_ / ND_);                                                                \
        constexpr int CC = horner64_chunk(RP, 1);                                                                 \
        const dim3 grid(nbl, (D + DB_ - 1) / DB_);                                                                \
        if (c64)                                                                                                  \
            rotate_average_horner64_q<float, DB_, ND_, CC><<<grid, kThreads, 0, s>>>(                             \
                static_cast<const float*>(vis), uvw, dd, D, T, nbl, C, P, kappa0, dkappa, avg, q);                \
        else                                                                                                      \
            rotate_average_horner64_q<double, DB_, ND_, CC><<<grid, kThreads, 0, s>>>(                            \
                static_cast<const double*>(vis), uvw, dd, D, T, nbl, C, P, kappa0, dkappa, avg, q);               \
        LEAP_HIP(hipGetLastError());                                                                              \
        return;                                                                                                   \
    }
    switch (tile) {
        LEAP_FQ(16, 2)
        LEAP_FQ(32, 1)
        LEAP_FQ(64, 1)
        LEAP_FQ(128, 2)
        LEAP_FQ(256, 2)
        default: break;
    }
#undef LEAP_FQ
    throw Error(LEAP_ERR_INVALID_ARGUMENT, "unsupported forward tile " + std::to_string(tile));
}

}  // namespace leap
