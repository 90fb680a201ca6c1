// Batched pseudo-inverse products cal[d] = Ad * dphi[d] + cal1[d] for D
// directions: the reference's per-direction GEMV (CpuLeapCalibrator.cc:178,
// cuBLAS call at CudaLeapCalibrator.cu:259) becomes one GEMM
//   cal^T [D x N] = dphi [D x K] * AdT [K x N]   (+ cal1).
//
// FP64: split-K over baselines, each workgroup holds a block of 16 directions'
// dphi in LDS and streams its K-chunk of AdT rows once (coalesced over the
// antenna index), so AdT is read D / 16 times instead of D times.
//
// Mixed: split-3 x BF16 on MFMA with FP32 accumulation.  Every operand x is
// split as x = hi + mid + lo with hi = bf16(x), mid = bf16(x - hi),
// lo = bf16(x - hi - mid) (about 24 significant bits), and the product keeps
// the six terms hi*hi, hi*mid, mid*hi, hi*lo, lo*hi, mid*mid, the ones above
// 2^-24 relative.  The split planes are stored in the operand layout of
// v_mfma_f32_16x16x32_bf16 (8 consecutive k of one row or column per lane), so
// each lane loads 16 contiguous bytes per plane and K-step.
#include "cube_ring.h"
#include "inversion_common.h"
#include "leap_internal.h"

namespace leap {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kDirTile = 16;   // FP64 GEMM: directions per workgroup
constexpr int kKChunk = 256;   // FP64 GEMM: baselines per workgroup

__global__ __launch_bounds__(256) void gemm_f64_partial(const double* __restrict__ AdT, int32_t K, int32_t N,
                                                        const double* __restrict__ delta, uint32_t D,
                                                        double* __restrict__ partial) {
    __shared__ double s_x[kDirTile][kKChunk];
    const int32_t chunk = blockIdx.x;
    const uint32_t d0 = blockIdx.y * kDirTile;
    const int32_t k0 = chunk * kKChunk;
    const int32_t kn = min(kKChunk, K - k0);
    for (int32_t i = threadIdx.x; i < kDirTile * kKChunk; i += 256) {
        const int32_t dd = i / kKChunk, k = i - dd * kKChunk;
        s_x[dd][k] = (d0 + dd < D && k < kn) ? delta[size_t(d0 + dd) * K + k0 + k] : 0.0;
    }
    __syncthreads();
    for (int32_t n = threadIdx.x; n < N; n += 256) {
        double acc[kDirTile];
#pragma unroll
        for (int j = 0; j < kDirTile; ++j) acc[j] = 0.0;
        for (int32_t k = 0; k < kn; ++k) {
            const double b = AdT[size_t(k0 + k) * N + n];
#pragma unroll
            for (int j = 0; j < kDirTile; ++j) acc[j] = fma(s_x[j][k], b, acc[j]);
        }
#pragma unroll
        for (int j = 0; j < kDirTile; ++j)
            if (d0 + j < D) partial[(size_t(chunk) * D + d0 + j) * N + n] = acc[j];
    }
}

__global__ void gemm_reduce(const double* __restrict__ partial, int32_t n_chunks, uint32_t D, int32_t N,
                            const double* __restrict__ cal1, double* __restrict__ cal) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(D) * N) return;
    double s = 0.0;
    for (int32_t c = 0; c < n_chunks; ++c) s += partial[size_t(c) * D * N + i];
    cal[i] = s + cal1[i];
}

// ---- split-3 x BF16 -------------------------------------------------------
__device__ inline void split3(double x, __bf16& hi, __bf16& mid, __bf16& lo) {
    hi = __bf16(float(x));
    const double r1 = x - double(float(hi));
    mid = __bf16(float(r1));
    const double r2 = r1 - double(float(mid));
    lo = __bf16(float(r2));
}

// Planes of a row-major M x K matrix X (the A operand, rows = output rows):
// out[p][mt][ks][lane][j] = split_p(X[mt*16 + lane%16][ks*32 + (lane/16)*8 + j]).
__global__ void split_a_planes(const double* __restrict__ X, uint32_t M, int32_t K, uint32_t Mt, int32_t Ks,
                               __bf16* __restrict__ out) {
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;  // one (mt, ks, lane)
    const size_t total = size_t(Mt) * Ks * 64;
    if (idx >= total) return;
    const uint32_t lane = idx % 64;
    const size_t t = idx / 64;
    const int32_t ks = int32_t(t % Ks);
    const uint32_t mt = uint32_t(t / Ks);
    const uint32_t m = mt * 16 + lane % 16;
    const int32_t kb = ks * 32 + int32_t(lane / 16) * 8;
    for (int j = 0; j < 8; ++j) {
        const int32_t k = kb + j;
        const double x = (m < M && k < K) ? X[size_t(m) * K + k] : 0.0;
        __bf16 h, md, l;
        split3(x, h, md, l);
        out[(0 * total + idx) * 8 + j] = h;
        out[(1 * total + idx) * 8 + j] = md;
        out[(2 * total + idx) * 8 + j] = l;
    }
}

// The A operand's planes straight from avgData: X[m][k] = delta_phase(d = m,
// n = k) (inversion_common.h), split like split_a_planes, without the FP64
// delta matrix in between.  Each thread's 8 consecutive k of one row are one
// 16-byte store per plane.
__global__ void delta_split_a_planes(const double* __restrict__ avg, uint32_t nbl, uint32_t P,
                                     const int2* __restrict__ pairs, int32_t N, const double* __restrict__ cal1,
                                     uint32_t M, int32_t K, uint32_t Mt, int32_t Ks, bf16x8* __restrict__ out) {
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;  // one (mt, ks, lane)
    const size_t total = size_t(Mt) * Ks * 64;
    if (idx >= total) return;
    const uint32_t lane = idx % 64;
    const size_t t = idx / 64;
    const int32_t ks = int32_t(t % Ks);
    const uint32_t mt = uint32_t(t / Ks);
    const uint32_t m = mt * 16 + lane % 16;
    const int32_t kb = ks * 32 + int32_t(lane / 16) * 8;
    // The 8 elements' loads (antenna pairs, then cal1 and avgData) all go out
    // before any element is computed; indices past the matrix are clamped to
    // valid ones and their results replaced by 0.
    const uint32_t mc = m < M ? m : M - 1;
    double acal1[8], re[8], im[8];
    if (K > 1) {  // uniform; with K = 1 there are no baselines and every element is 0
        int2 pr[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) pr[j] = pairs[min(kb + j, K - 2)];  // K - 1 pairs, K - 1 <= nbl
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int32_t a1 = min(max(pr[j].x, 0), N - 1), a2 = min(max(pr[j].y, 0), N - 1);
            acal1[j] = cal1[size_t(mc) * N + a1] - cal1[size_t(mc) * N + a2];
            const double* a = avg + (size_t(mc) * nbl + min(kb + j, K - 2)) * (2 * P);
            re[j] = a[0];
            im[j] = a[1];
        }
    }
    bf16x8 h, md, l;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int32_t k = kb + j;
        const double x = (m < M && k < K - 1) ? dev::delta_phase_of(acal1[j], re[j], im[j]) : 0.0;
        __bf16 a, b, c;
        split3(x, a, b, c);
        h[j] = a;
        md[j] = b;
        l[j] = c;
    }
    out[0 * total + idx] = h;
    out[1 * total + idx] = md;
    out[2 * total + idx] = l;
}

// Planes of a row-major K x N matrix B (the B operand, columns = output columns):
// out[p][nt][ks][lane][j] = split_p(B[ks*32 + (lane/16)*8 + j][nt*16 + lane%16]).
__global__ void split_b_planes(const double* __restrict__ B, int32_t K, int32_t N, int32_t Nt, int32_t Ks,
                               __bf16* __restrict__ out) {
    const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t total = size_t(Nt) * Ks * 64;
    if (idx >= total) return;
    const uint32_t lane = idx % 64;
    const size_t t = idx / 64;
    const int32_t ks = int32_t(t % Ks);
    const int32_t nt = int32_t(t / Ks);
    const int32_t n = nt * 16 + int32_t(lane % 16);
    const int32_t kb = ks * 32 + int32_t(lane / 16) * 8;
    for (int j = 0; j < 8; ++j) {
        const int32_t k = kb + j;
        const double x = (n < N && k < K) ? B[size_t(k) * N + n] : 0.0;
        __bf16 h, md, l;
        split3(x, h, md, l);
        out[(0 * total + idx) * 8 + j] = h;
        out[(1 * total + idx) * 8 + j] = md;
        out[(2 * total + idx) * 8 + j] = l;
    }
}

// One wave per (16-row m-tile, 16-column n-tile, K-split); 4 waves per
// workgroup take 4 consecutive m-tiles of one n-tile and K-split.
// partial[split][m][n] (FP32) for m < Mt*16, n < Nt*16.
__global__ __launch_bounds__(256) void gemm_bf16x3_mfma(const bf16x8* __restrict__ Ap, const bf16x8* __restrict__ Bp,
                                                       uint32_t Mt, int32_t Nt, int32_t Ks, int32_t ks_per_split,
                                                       float* __restrict__ partial) {
    const uint32_t lane = threadIdx.x % 64;
    const uint32_t wave = threadIdx.x / 64;
    const int32_t nt = blockIdx.x;
    const uint32_t mt = blockIdx.y * 4 + wave;
    const int32_t split = blockIdx.z;
    if (mt >= Mt) return;  // whole wave exits together
    const size_t a_total = size_t(Mt) * Ks * 64;
    const size_t b_total = size_t(Nt) * Ks * 64;
    const int32_t ks0 = split * ks_per_split;
    const int32_t ks1 = min(Ks, ks0 + ks_per_split);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int32_t ks = ks0; ks < ks1; ++ks) {
        const size_t ai = (size_t(mt) * Ks + ks) * 64 + lane;
        const size_t bi = (size_t(nt) * Ks + ks) * 64 + lane;
        const bf16x8 ah = Ap[ai], am = Ap[a_total + ai], al = Ap[2 * a_total + ai];
        const bf16x8 bh = Bp[bi], bm = Bp[b_total + bi], bl = Bp[2 * b_total + bi];
        // smallest terms first
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
    }
    // D layout: lane holds column lane%16, rows (lane/16)*4 + i.
    const uint32_t M16 = Mt * 16;
    const uint32_t N16 = uint32_t(Nt) * 16;
    const uint32_t col = uint32_t(nt) * 16 + lane % 16;
    for (int i = 0; i < 4; ++i) {
        const uint32_t row = mt * 16 + (lane / 16) * 4 + i;
        partial[(size_t(split) * M16 + row) * N16 + col] = acc[i];
    }
}

// Whole-output variant for M, N <= 256 (Mt, Nt <= 16 tiles): a workgroup of 8
// waves takes one K-split for all output tiles, so every operand byte comes
// from HBM once per split.  Wave w owns m-tiles 2w, 2w + 1 and all n-tiles (32
// tiles, 128 FP32 accumulators per lane).  The B operand of a K-step (all
// n-tiles, 3 planes, 48 KB) is copied into LDS with global_load_lds, two
// K-steps ahead, through three stages (144 KB, so one workgroup per CU); each
// wave holds its A tiles in registers, loaded one K-step ahead.  Per tile and
// K-step the six products go in the order of gemm_bf16x3_mfma.
constexpr int kFullTM = 2, kFullWaves = 8, kFullThreads = 64 * kFullWaves;
constexpr int kFullMaxTiles = 16;                         // per dimension: 256 rows or columns
constexpr int kFullBChunks = kFullMaxTiles * 3 * 64;      // 16-byte chunks of one K-step's B
constexpr int kFullBCopies = kFullBChunks / kFullThreads;  // 1 KB copies per wave and K-step
constexpr int kFullStages = 3;                            // the K-step in use and two copied ahead
constexpr int kFullSplits = 256;                          // K-splits of the whole-output kernel (one per CU)
static_assert(kFullBChunks % kFullThreads == 0 && kFullWaves * kFullTM == kFullMaxTiles, "tiling");

__global__ __launch_bounds__(kFullThreads) void gemm_bf16x3_full(const bf16x8* __restrict__ Ap,
                                                                 const bf16x8* __restrict__ Bp, uint32_t Mt,
                                                                 int32_t Nt, int32_t Ks, int32_t ks_per_split,
                                                                 float* __restrict__ partial) {
    __shared__ bf16x8 s_b[kFullStages][kFullBChunks];  // [stage][(n-tile * 3 + plane) * 64 + lane]
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid % 64;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const uint32_t m0 = wave * kFullTM;
    const bool rows = m0 < Mt;  // wave-uniform; a wave without rows still copies B and meets the barriers
    const int32_t split = blockIdx.x;
    const size_t a_total = size_t(Mt) * Ks * 64;
    const size_t b_total = size_t(Nt) * Ks * 64;
    const int32_t ks0 = split * ks_per_split;
    const int32_t ks1 = min(Ks, ks0 + ks_per_split);
    f32x4 acc[kFullTM][kFullMaxTiles];
#pragma unroll
    for (int i = 0; i < kFullTM; ++i)
#pragma unroll
        for (int j = 0; j < kFullMaxTiles; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // B of K-step ks into stage st: wave w copies the 64-chunk groups g = w + 8 r (n-tile g / 3,
    // plane g % 3), 1 KB each.  Groups of absent n-tiles, and K-steps past the split (whose stage is
    // not read again), copy valid data (n-tile Nt - 1, K-step ks1 - 1) so that every wave issues
    // exactly kFullBCopies copies per K-step and the waits below count them exactly.
    auto copy_b = [&](int32_t ks, int st) {
        const int32_t kc = ks < ks1 ? ks : ks1 - 1;
#pragma unroll
        for (int r = 0; r < kFullBCopies; ++r) {
            const uint32_t g = wave + kFullWaves * r;
            const uint32_t nt = g / 3, p = g % 3;
            const uint32_t ntc = nt < uint32_t(Nt) ? nt : uint32_t(Nt) - 1;
            const bf16x8* src = Bp + p * b_total + (size_t(ntc) * Ks + kc) * 64 + lane;
            __builtin_amdgcn_global_load_lds((const LEAP_GLOBAL void*)src, (LEAP_LDS void*)(&s_b[st][g * 64]), 16, 0,
                                             0);
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    // A of K-step ks (clamped to the split, as above) into a; a wave without rows loads m-tiles
    // 0 and 1 and does not use them, so every wave issues the same loads
    auto load_a = [&](int32_t ks, bf16x8 (*a)[3]) {
        const int32_t kc = ks < ks1 ? ks : ks1 - 1;
        const uint32_t ma = rows ? m0 : 0;
#pragma unroll
        for (int i = 0; i < kFullTM; ++i) {
            const bf16x8* t = Ap + (size_t(ma + i) * Ks + kc) * 64;  // ma + kFullTM <= Mt (a multiple of 4)
            a[i][0] = t[lane];
            a[i][1] = t[a_total + lane];
            a[i][2] = t[2 * a_total + lane];
        }
    };
    bf16x8 a[kFullTM][3], an[kFullTM][3];
    if (ks0 < ks1) {  // uniform
        copy_b(ks0, 0);
        copy_b(ks0 + 1, 1);
        load_a(ks0, a);
        __builtin_amdgcn_s_waitcnt(dev::kWaitVectorMem);
    }
    for (int32_t ks = ks0; ks < ks1; ++ks) {
        const int st = (ks - ks0) % kFullStages;
        // Past the barrier: this K-step's B is in stage st (every wave waited for its copies), and
        // every wave is done with the previous K-step, whose stage the copy below refills.
        __syncthreads();
        load_a(ks + 1, an);
        copy_b(ks + 2, (ks - ks0 + 2) % kFullStages);
        if (rows) {
#pragma unroll
            for (int j = 0; j < kFullMaxTiles; ++j) {
                if (j < Nt) {  // uniform
                    const bf16x8 bh = s_b[st][(j * 3) * 64 + lane];
                    const bf16x8 bm = s_b[st][(j * 3 + 1) * 64 + lane];
                    const bf16x8 bl = s_b[st][(j * 3 + 2) * 64 + lane];
                    // the kFullTM tiles' accumulation chains interleaved
#pragma unroll
                    for (int i = 0; i < kFullTM; ++i)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][1], bm, acc[i][j], 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < kFullTM; ++i)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][0], bl, acc[i][j], 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < kFullTM; ++i)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][2], bh, acc[i][j], 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < kFullTM; ++i)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][0], bm, acc[i][j], 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < kFullTM; ++i)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][1], bh, acc[i][j], 0, 0, 0);
#pragma unroll
                    for (int i = 0; i < kFullTM; ++i)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][0], bh, acc[i][j], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kFullTM; ++i)
#pragma unroll
            for (int q = 0; q < 3; ++q) a[i][q] = an[i][q];
        // the next K-step's A and B (issued before the copies of the K-step after it) have landed
        dev::wait_vmem_but(kFullBCopies);
    }
    if (!rows) return;
    const uint32_t M16 = Mt * 16;
    const uint32_t N16 = uint32_t(Nt) * 16;
#pragma unroll
    for (int i = 0; i < kFullTM; ++i) {
#pragma unroll
        for (int j = 0; j < kFullMaxTiles; ++j) {
            if (j < Nt) {
                const uint32_t col = uint32_t(j) * 16 + lane % 16;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t row = (m0 + i) * 16 + (lane / 16) * 4 + r;
                    partial[(size_t(split) * M16 + row) * N16 + col] = acc[i][j][r];
                }
            }
        }
    }
}

__global__ void mfma_reduce(const float* __restrict__ partial, int32_t n_split, uint32_t D, int32_t N, uint32_t M16,
                            uint32_t N16, const double* __restrict__ cal1, double* __restrict__ cal) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(D) * N) return;
    const uint32_t d = uint32_t(i / N), n = uint32_t(i % N);
    double s = 0.0;
#pragma unroll 8
    for (int32_t c = 0; c < n_split; ++c) s += double(partial[(size_t(c) * M16 + d) * N16 + n]);  // split order
    cal[i] = s + cal1[i];
}

}  // namespace

size_t gemm_f64_partial_doubles(int32_t K, int32_t N, uint32_t D) {
    return size_t((K + kKChunk - 1) / kKChunk) * D * N;
}

void launch_gemm_f64(hipStream_t s, const double* AdT, int32_t K, int32_t N, const double* delta, uint32_t D,
                     const double* cal1, double* partial, double* cal) {
    const int32_t chunks = (K + kKChunk - 1) / kKChunk;
    gemm_f64_partial<<<dim3(chunks, (D + kDirTile - 1) / kDirTile), 256, 0, s>>>(AdT, K, N, delta, D, partial);
    LEAP_HIP(hipGetLastError());
    const size_t n = size_t(D) * N;
    gemm_reduce<<<dim3(uint32_t((n + 255) / 256)), 256, 0, s>>>(partial, chunks, D, N, cal1, cal);
    LEAP_HIP(hipGetLastError());
}

size_t bf16_planes_bytes(uint32_t rows, int32_t K) {
    const size_t tiles = (rows + 15) / 16;
    const size_t ks = size_t((K + 31) / 32);
    return 3 * tiles * ks * 64 * 8 * sizeof(__bf16);
}

void launch_split_b(hipStream_t s, const double* B, int32_t K, int32_t N, void* planes) {
    const int32_t Nt = (N + 15) / 16, Ks = (K + 31) / 32;
    const size_t total = size_t(Nt) * Ks * 64;
    split_b_planes<<<dim3(uint32_t((total + 255) / 256)), 256, 0, s>>>(B, K, N, Nt, Ks,
                                                                       static_cast<__bf16*>(planes));
    LEAP_HIP(hipGetLastError());
}

bool mfma_full_output(uint32_t D, int32_t N) {
    return (D + 63) / 64 * 4 <= uint32_t(kFullMaxTiles) && (N + 15) / 16 <= kFullMaxTiles;
}

int32_t mfma_n_split(int32_t K, uint32_t D, int32_t N) {
    const int32_t Ks = (K + 31) / 32;
    if (mfma_full_output(D, N)) return std::min(Ks, kFullSplits);  // about one workgroup per CU
    const int64_t tiles = int64_t((D + 63) / 64) * ((N + 15) / 16);
    int32_t split = int32_t(std::max<int64_t>(1, 2048 / std::max<int64_t>(1, tiles)));
    return std::min(split, Ks);
}

size_t mfma_partial_floats(int32_t K, uint32_t D, int32_t N) {
    const uint32_t Mt = (D + 63) / 64 * 4;
    const int32_t Nt = (N + 15) / 16;
    return size_t(mfma_n_split(K, D, N)) * Mt * 16 * size_t(Nt) * 16;
}

void launch_gemm_bf16x3(hipStream_t s, const void* b_planes, int32_t K, int32_t N, const double* delta, uint32_t D,
                        void* a_planes, const double* cal1, float* partial, double* cal, const DeltaSource* from) {
    const uint32_t Mt = (D + 63) / 64 * 4;  // whole 4-wave groups
    const int32_t Nt = (N + 15) / 16, Ks = (K + 31) / 32;
    const size_t a_total = size_t(Mt) * Ks * 64;
    if (from)
        delta_split_a_planes<<<dim3(uint32_t((a_total + 255) / 256)), 256, 0, s>>>(
            from->avg, from->nbl, from->pols, from->pairs, N, cal1, D, K, Mt, Ks, static_cast<bf16x8*>(a_planes));
    else
        split_a_planes<<<dim3(uint32_t((a_total + 255) / 256)), 256, 0, s>>>(delta, D, K, Mt, Ks,
                                                                             static_cast<__bf16*>(a_planes));
    LEAP_HIP(hipGetLastError());
    int32_t n_split = mfma_n_split(K, D, N);
    const int32_t ks_per = (Ks + n_split - 1) / n_split;
    n_split = (Ks + ks_per - 1) / ks_per;  // no empty splits
    if (mfma_full_output(D, N))
        gemm_bf16x3_full<<<dim3(n_split), kFullThreads, 0, s>>>(static_cast<const bf16x8*>(a_planes),
                                                                    static_cast<const bf16x8*>(b_planes), Mt, Nt, Ks,
                                                                    ks_per, partial);
    else
        gemm_bf16x3_mfma<<<dim3(Nt, Mt / 4, n_split), 256, 0, s>>>(static_cast<const bf16x8*>(a_planes),
                                                                   static_cast<const bf16x8*>(b_planes), Mt, Nt, Ks,
                                                                   ks_per, partial);
    LEAP_HIP(hipGetLastError());
    const size_t n = size_t(D) * N;
    mfma_reduce<<<dim3(uint32_t((n + 255) / 256)), 256, 0, s>>>(partial, n_split, D, N, Mt * 16, uint32_t(Nt) * 16,
                                                                cal1, cal);
    LEAP_HIP(hipGetLastError());
}

}  // namespace leap
