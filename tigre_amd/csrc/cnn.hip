// 3D convolutions of the learned proximal network, BF16 MFMA on gfx950.
//
// The network (a residual CNN that stands in for the TV step of ASD-POCS) is
// six 3x3x3 'same' convolutions with 32 channels between them.  All of them,
// forward and backward, run as implicit GEMMs on v_mfma_f32_16x16x32_bf16
// (A: 16 rows x 32 k, B: 32 k x 16 columns, FP32 accumulate; lane l holds
// row / column l % 16 and k = 8 (l / 16) .. + 7, and result register r of lane
// l is row 4 (l / 16) + r, column l % 16):
//
//  * conv_zm_kernel (32 -> 32 channels): rows 16 output channels, columns 16
//    voxels along x, k one tap's 32 input channels; the weights in registers,
//    input planes staged in LDS once each while the workgroup marches along z.
//  * conv_taps_kernel (one FP32 channel -> 32): k is the tap instead, so one
//    MFMA covers all 27 taps of 16 voxels.
//  * conv_out_kernel (32 -> one FP32 channel added to a residual input): the
//    taps are the columns instead, P[u][tap] for every input voxel, then
//    in-plane tap sums in LDS.
//  * conv_kernel: the plain gather form (one 16-byte load per lane per tap is
//    an A fragment), for the single-channel input with the residual epilogue.
//  * wgrad_kernel: dw[tap][o][i] = sum_v x[v + tap][i] dy[v][o], GEMM k =
//    32 voxels of an x-row.  The x and dy rows are staged in LDS transposed
//    (channel-major), so fragments are again 16-byte reads.  Tap 27 is the
//    bias: x replaced by 1 in channel 0, so dw[27][o][0] = sum_v dy[v][o].
//
// Volumes are N x D x H x W (x fastest); taps index (dz, dy, dx) in
// {-1, 0, 1}^3 as 9 (dz + 1) + 3 (dy + 1) + (dx + 1); outside the volume the
// input is zero.
#include <mutex>
#include <type_traits>

#include "common.hpp"

namespace tigre {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kC = 32;            // channels between layers
constexpr int kTilesPerWave = 4;  // 16-voxel tiles per wavefront (64 voxels along x)

// Round to nearest even; infinities keep their bits and NaNs stay NaN (quiet
// bit set when the dropped half of the significand was not zero).  Both
// results are computed and one selected, so a store loop has no branches.
__device__ __forceinline__ unsigned short bf16_bits(float f) {
    const unsigned u = __float_as_uint(f);
    const unsigned special = (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u);
    const unsigned rounded = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    return (unsigned short)((u & 0x7f800000u) == 0x7f800000u ? special : rounded);
}

// A fragment of 8 channels (8 q .. 8 q + 7) of voxel `vox`, or zeros.
template <int kIn>
__device__ __forceinline__ bf16x8 load_frag(const void* __restrict__ src, long long vox, int q, bool ok) {
    bf16x8 f;
    if (kIn == TIGRE_CNN_IN_BF16X32) {
        uint4 raw = make_uint4(0u, 0u, 0u, 0u);
        if (ok) raw = *reinterpret_cast<const uint4*>(static_cast<const unsigned short*>(src) + vox * kC + 8 * q);
        f = __builtin_bit_cast(bf16x8, raw);
    } else {
        float v = 0.0f;
        if (ok && q == 0) v = static_cast<const float*>(src)[vox];
        uint4 raw = make_uint4((unsigned)bf16_bits(v), 0u, 0u, 0u);
        f = __builtin_bit_cast(bf16x8, raw);
    }
    return f;
}

template <int kIn, int kEpi>
__global__ __launch_bounds__(256) void conv_kernel(const void* __restrict__ x, const unsigned short* __restrict__ w,
                                                   const float* __restrict__ bias, const void* __restrict__ aux,
                                                   void* __restrict__ y, int N, int D, int H, int W) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, q = lane >> 4;
    const long long row = (long long)blockIdx.y * 4 + wave;  // (n, z, y) row of the wavefront
    const long long rows = (long long)N * D * H;
    if (row >= rows) return;  // uniform per wavefront
    const int yy0 = (int)(row % H), zz0 = (int)((row / H) % D);
    const long long n = row / ((long long)H * D);
    const int x0 = blockIdx.x * (16 * kTilesPerWave);

    f32x4 acc[kTilesPerWave][2];
#pragma unroll
    for (int t = 0; t < kTilesPerWave; ++t) acc[t][0] = acc[t][1] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int tap = 0; tap < 27; ++tap) {
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        const int zz = zz0 + dz, yy = yy0 + dy;
        const bool row_ok = zz >= 0 && zz < D && yy >= 0 && yy < H;
        const long long base = ((n * D + zz) * H + yy) * (long long)W;
        const unsigned short* wt = w + (long long)tap * kC * kC;
        const bf16x8 b0 = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(wt + m * kC + 8 * q));
        const bf16x8 b1 = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(wt + (16 + m) * kC + 8 * q));
#pragma unroll
        for (int t = 0; t < kTilesPerWave; ++t) {
            const int xx = x0 + 16 * t + m + dx;
            const bf16x8 a = load_frag<kIn>(x, base + xx, q, row_ok && xx >= 0 && xx < W);
            acc[t][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc[t][0], 0, 0, 0);
            acc[t][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc[t][1], 0, 0, 0);
        }
    }

    const long long obase = row * (long long)W;
#pragma unroll
    for (int t = 0; t < kTilesPerWave; ++t) {
#pragma unroll
        for (int hb = 0; hb < 2; ++hb) {
            const int o = hb * 16 + m;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int xo = x0 + 16 * t + 4 * q + r;
                if (xo >= W) continue;
                const long long vox = obase + xo;
                float v = acc[t][hb][r];
                if (kEpi == TIGRE_CNN_EPI_BIAS_RELU) {
                    v += bias[o];
                    v = v > 0.0f ? v : 0.0f;
                    static_cast<unsigned short*>(y)[vox * kC + o] = bf16_bits(v);
                } else if (kEpi == TIGRE_CNN_EPI_RELU_MASK) {
                    const unsigned short a = static_cast<const unsigned short*>(aux)[vox * kC + o];
                    v = (a != 0 && (a & 0x8000u) == 0) ? v : 0.0f;  // activation > 0
                    static_cast<unsigned short*>(y)[vox * kC + o] = bf16_bits(v);
                } else {  // residual output: channel 0 only, FP32
                    if (o == 0) static_cast<float*>(y)[vox] = static_cast<const float*>(aux)[vox] + (v + bias[0]);
                }
            }
        }
    }
}

// Staged input blocks.  A workgroup computes 4 output rows; the 6-row block
// of an input plane its taps read (with one halo voxel on either side along x)
// is copied into LDS, voxels outside the volume as zeros.
constexpr int kLdsTiles = 8;                    // 16-voxel tiles per workgroup row (single channel)
constexpr int kLdsBX = 16 * kLdsTiles + 2;      // staged voxels along x (with the halo)
constexpr int kLdsRows = 4 + 2;                 // staged rows: the workgroup's 4 and the halo

__device__ uint4 g_zero_chunk[4];  // zero-initialised, never written: the source of the zero halo

// The 32 -> 32 channel convolution (forward with bias and ReLU; data gradient
// with the ReLU mask), implicit GEMM on v_mfma_f32_16x16x32_bf16: rows = 16
// output channels (one half of the 32, fixed per wavefront), columns = 16
// voxels along x, k = one tap's 32 input channels.  The A fragments of all 27
// taps for the wavefront's half (108 VGPRs) are loaded once, so LDS carries
// only the staged activations: a B fragment is one 16-byte read of 8 channels
// of one voxel, with the channel groups stored in the order q ^ ((x >> 2) & 3)
// (conflict free).  Result register r of lane (m, q) is output channel
// 16 hb + 4 q + r of voxel m: four consecutive channels, stored as one 8-byte
// write.  A workgroup owns a 4 x (16 kZmTiles) voxel tile for a chunk of
// kZmChunk output planes and marches through its input planes z0 - 1 ..
// z0 + kZmChunk once: each input plane is staged once (direct global-to-LDS
// loads, the zero halo from a zero chunk), into one of two LDS buffers while
// the previous plane is computed, and feeds the accumulators of the three
// output planes it contributes to (z + 1 through its dz = -1 taps, z through
// dz = 0, z - 1 through dz = +1); the output plane that has received all three
// is written.  The staged row sr feeds output row sr - dy for each dy, so one
// fragment read serves up to 3 dz x 3 dy MFMAs.  The wavefronts are the
// kZmTiles x-tiles times the two channel halves; the two halves of a tile
// read the same staged fragments.  The two buffers are separate arrays, so
// the compiler can tell the copies into one from the reads of the other and
// does not wait for the copies before the reads.  Every output voxel sums its
// taps in the order dz = -1, 0, +1, then dx, then dy.  (With the weights
// staged in LDS as well and 32 output channels per wavefront, every plane
// read 72 fragments per 216 MFMAs instead of 18 per 108: 0.88 ms per 256^3
// layer against 0.77 ms here; staging the three input planes of each output
// plane in turn, with a full wait between copy and use, 1.39 ms.)
constexpr int kZmTiles = 2;                        // 16-voxel x-tiles per workgroup
constexpr int kZmChunk = 32;                       // output planes per workgroup
constexpr int kZmBX = 16 * kZmTiles + 2;           // staged voxels along x (with the halo)
constexpr int kZmChunks = kLdsRows * kZmBX * 4;    // 16-byte chunks of one staged plane
constexpr int kZmThreads = 128 * kZmTiles;

// Output plane `plane` (pointer offsets of 32-bit lane parts from a uniform
// plane base): row orow of the tile at lane offset loff + orow W kC.
template <int kEpi, bool full>
__device__ __forceinline__ void zm_store(const f32x4 (&acc)[4], const float (&bias)[4], const void* __restrict__ aux,
                                         void* __restrict__ y, long long plane, int H, int W, int y0, int xv,
                                         int loff) {
    const unsigned short* ap = static_cast<const unsigned short*>(aux) + plane;
    unsigned short* yp = static_cast<unsigned short*>(y) + plane;
#pragma unroll
    for (int orow = 0; orow < 4; ++orow) {
        if (y0 + orow >= H) break;
        if (!full && xv >= W) continue;
        const int off = loff + orow * W * kC;
        float v[4] = {acc[orow][0], acc[orow][1], acc[orow][2], acc[orow][3]};
        if (kEpi == TIGRE_CNN_EPI_BIAS_RELU) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] += bias[r];
                v[r] = v[r] > 0.0f ? v[r] : 0.0f;
            }
        } else {
            const uint2 a = *reinterpret_cast<const uint2*>(ap + off);
            const unsigned ah[4] = {a.x & 0xffffu, a.x >> 16, a.y & 0xffffu, a.y >> 16};
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = (ah[r] != 0 && (ah[r] & 0x8000u) == 0) ? v[r] : 0.0f;  // activation > 0
        }
        const uint2 out = make_uint2((unsigned)bf16_bits(v[0]) | ((unsigned)bf16_bits(v[1]) << 16),
                                     (unsigned)bf16_bits(v[2]) | ((unsigned)bf16_bits(v[3]) << 16));
        *reinterpret_cast<uint2*>(yp + off) = out;
    }
}

template <int kEpi>
__global__ __launch_bounds__(kZmThreads) __attribute__((amdgpu_waves_per_eu(2, 2))) void conv_zm_kernel(
    const unsigned short* __restrict__ x, const unsigned short* __restrict__ w, const float* __restrict__ bias,
    const void* __restrict__ aux, void* __restrict__ y, int N, int D, int H, int W) {
    __shared__ uint4 plane_a[kZmChunks], plane_b[kZmChunks];  // [row][voxel][channel group, swizzled]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, q = lane >> 4;
    const int tile = wave % kZmTiles, hb = wave / kZmTiles;
    const int x0 = blockIdx.x * (16 * kZmTiles), y0 = blockIdx.y * 4;
    const int nchunk = (D + kZmChunk - 1) / kZmChunk;
    const int zc = (int)(blockIdx.z % (unsigned)nchunk);
    const long long n = blockIdx.z / (unsigned)nchunk;
    const int z_lo = zc * kZmChunk, z_hi = min(z_lo + kZmChunk, D);  // output planes z_lo .. z_hi - 1
    const int xv = x0 + 16 * tile + m;                                // this lane's output voxel along x
    const int och = 16 * hb + 4 * q;                                  // its first output channel

    // A fragments: row = output channel 16 hb + m, k = input channels 8 q .. 8 q + 7, per tap
    bf16x8 wf[27];
#pragma unroll
    for (int t = 0; t < 27; ++t)
        wf[t] = __builtin_bit_cast(bf16x8,
                                   *reinterpret_cast<const uint4*>(w + ((long long)(t * kC + 16 * hb + m) * kC + 8 * q)));
    float bias_lane[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (kEpi == TIGRE_CNN_EPI_BIAS_RELU) {
#pragma unroll
        for (int r = 0; r < 4; ++r) bias_lane[r] = bias[och + r];
    }
    unsigned boff[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int xi = 16 * tile + m + d;
        boff[d] = (unsigned)((xi * 4 + (q ^ ((xi >> 2) & 3))) * 16);
    }

    // the chunks this thread copies of every staged plane: element offsets inside the plane, or -1
    // for the zero halo (a plane has fewer than 2^31 elements)
    constexpr int kStageIters = (kZmChunks + kZmThreads - 1) / kZmThreads;
    int soff[kStageIters];
#pragma unroll
    for (int it = 0; it < kStageIters; ++it) {
        const int i = it * kZmThreads + tid;
        const int yi = i / (4 * kZmBX), r = i - yi * 4 * kZmBX, xi = r >> 2, sg = r & 3;
        const int gy = y0 - 1 + yi, gx = x0 - 1 + xi;
        soff[it] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? (gy * W + gx) * kC + 8 * (sg ^ ((xi >> 2) & 3)) : -1;
    }
    const long long plane_elems = (long long)H * W * kC;
    auto stage = [&](int zi, auto buf) {  // issue the copies of input plane zi into buffer buf
        const unsigned short* xp = x + (n * D + zi) * plane_elems;
        uint4* dst = decltype(buf)::value ? plane_b : plane_a;
#pragma unroll
        for (int it = 0; it < kStageIters; ++it) {
            if (it * kZmThreads + tid < kZmChunks) {
                const void* src = soff[it] >= 0 ? static_cast<const void*>(xp + soff[it]) : g_zero_chunk;
                __builtin_amdgcn_global_load_lds(src, &dst[it * kZmThreads + wave * 64], 16, 0, 0);
            }
        }
    };
    const int loff = (y0 * W + xv) * kC + och;  // this lane's output offset in row y0 of a plane

    f32x4 acc[3][4];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[s][r] = f32x4{0.f, 0.f, 0.f, 0.f};

    if (z_lo - 1 >= 0) stage(z_lo - 1, std::integral_constant<int, 0>{});
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    // one input plane, staged in buffer kB: the accumulator slots of output planes zi + 1, zi, zi - 1
    // are kN, kC0, kP
    auto plane_step = [&](int zi, auto sn, auto sc, auto sp, auto sb) {
        constexpr int kN = decltype(sn)::value, kC0 = decltype(sc)::value, kP = decltype(sp)::value;
        constexpr int kB = decltype(sb)::value;
        // addresses are formed from the plane index afresh in every step (a scalar base plus the
        // lane's 32-bit offset), instead of 64-bit pointers per unrolled step held in VGPRs
        asm volatile("" : "+s"(zi));
        if (zi + 1 <= z_hi && zi + 1 < D) stage(zi + 1, std::integral_constant<int, 1 - kB>{});
        if (zi >= 0 && zi < D) {  // uniform per workgroup
            const char* lds = reinterpret_cast<const char*>(kB ? plane_b : plane_a);
            const bool on[3] = {zi + 1 < z_hi, zi >= z_lo && zi < z_hi, zi - 1 >= z_lo};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                bf16x8 b[kLdsRows];
#pragma unroll
                for (int sr = 0; sr < kLdsRows; ++sr)
                    b[sr] = __builtin_bit_cast(bf16x8,
                                               *reinterpret_cast<const uint4*>(lds + boff[d] + sr * (kZmBX * 64)));
#pragma unroll
                for (int dz = 0; dz < 3; ++dz) {
                    if (!on[dz]) continue;
                    f32x4(&A)[4] = acc[dz == 0 ? kN : (dz == 1 ? kC0 : kP)];
#pragma unroll
                    for (int sr = 0; sr < kLdsRows; ++sr)
#pragma unroll
                        for (int dy = 0; dy < 3; ++dy) {
                            const int orow = sr - dy;
                            if (orow < 0 || orow > 3) continue;  // resolved at compile time
                            A[orow] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[dz * 9 + dy * 3 + d], b[sr], A[orow],
                                                                             0, 0, 0);
                        }
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();  // the next plane is in LDS; this plane's buffer may be overwritten
        const int zo = zi - 1;  // complete: it has received its dz = -1, 0 and +1 taps
        if (zo >= z_lo && zo < z_hi) {
            const long long plane = (n * D + zo) * plane_elems;
            if (x0 + 16 * kZmTiles <= W)
                zm_store<kEpi, true>(acc[kP], bias_lane, aux, y, plane, H, W, y0, xv, loff);
            else
                zm_store<kEpi, false>(acc[kP], bias_lane, aux, y, plane, H, W, y0, xv, loff);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[kP][r] = f32x4{0.f, 0.f, 0.f, 0.f};
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    for (int zi = z_lo - 1; zi <= z_hi; zi += 6) {
        plane_step(zi, I0{}, I1{}, I2{}, I0{});
        if (zi + 1 > z_hi) break;
        plane_step(zi + 1, I2{}, I0{}, I1{}, I1{});
        if (zi + 2 > z_hi) break;
        plane_step(zi + 2, I1{}, I2{}, I0{}, I0{});
        if (zi + 3 > z_hi) break;
        plane_step(zi + 3, I0{}, I1{}, I2{}, I1{});
        if (zi + 4 > z_hi) break;
        plane_step(zi + 4, I2{}, I0{}, I1{}, I0{});
        if (zi + 5 > z_hi) break;
        plane_step(zi + 5, I1{}, I2{}, I0{}, I1{});
    }
}

// The residual output layer (32 bf16 channels -> one FP32 channel added to
// aux) with the taps as the GEMM columns.  P[u][tap] = sum_i x[u][i] w[tap][0][i]
// for every input voxel u is one MFMA pair per 16 voxels (columns = taps
// 0..31, 27 used), and y[v] = aux[v] + (sum_tap P[v + tap][tap] + bias[0]).
// The 32 -> 32 kernel would spend 27 MFMAs per 16 outputs on this layer with 15 of its
// 16 columns idle; here it is 2 per 16 input voxels.  A workgroup owns a
// 64 x 4 tile of outputs and a chunk of kOutZ planes, and marches through the
// input planes z0 - 1 .. z0 + kOutZ: for each it computes P over the 66 x 6
// halo'd tile into LDS; then every lane (one output column) adds the 9-tap
// in-plane sum of each dz group to the accumulator of the output plane that
// group feeds (input plane z feeds z + 1 with the dz = -1 taps, z with dz = 0,
// z - 1 with dz = +1), and writes the output plane that has received all three.
// Every output sums its groups in the order dz = -1, 0, +1 and each group in
// (dy, dx) order: the result does not depend on the launch shape.  The next
// plane's A fragments are loaded while the current plane is summed.
constexpr int kOutTX = 64, kOutTY = 4, kOutZ = 16;
constexpr int kOutHX = kOutTX + 2, kOutHY = kOutTY + 2;
constexpr int kOutU = kOutHX * kOutHY;           // halo'd voxels per plane (396)
constexpr int kOutBlocks = (kOutU + 15) / 16;     // MFMA row blocks per plane (25)
constexpr int kOutPerWave = (kOutBlocks + 3) / 4; // row blocks per wavefront, at most (7)
constexpr int kOutStride = kOutBlocks * 16 + 1;   // LDS row stride of one tap (odd: conflict-free writes)

__global__ __launch_bounds__(256) void conv_out_kernel(const unsigned short* __restrict__ x,
                                                       const unsigned short* __restrict__ w,
                                                       const float* __restrict__ bias, const float* __restrict__ aux,
                                                       float* __restrict__ y, int N, int D, int H, int W) {
    __shared__ float P[27 * kOutStride];  // [tap][halo'd voxel]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, q = lane >> 4;
    const int x0 = blockIdx.x * kOutTX, y0 = blockIdx.y * kOutTY;
    const int nchunk = (D + kOutZ - 1) / kOutZ;
    const int zc = (int)(blockIdx.z % (unsigned)nchunk);
    const long long n = blockIdx.z / (unsigned)nchunk;
    const int z0 = zc * kOutZ, z1 = min(z0 + kOutZ, D);  // output planes z0 .. z1 - 1

    // B fragments: column 16 hb + m is the tap, k = 8 q .. 8 q + 7 the input channel
    bf16x8 b[2];
#pragma unroll
    for (int hb = 0; hb < 2; ++hb) {
        const int tap = 16 * hb + m;
        uint4 raw = make_uint4(0u, 0u, 0u, 0u);
        if (tap < 27) raw = *reinterpret_cast<const uint4*>(w + (long long)tap * kC * kC + 8 * q);
        b[hb] = __builtin_bit_cast(bf16x8, raw);
    }
    // A fragments of plane z: row block wave + 4 k, voxel 16 (wave + 4 k) + m, channels 8 q ..
    auto load_plane = [&](int z, bf16x8* a) {
#pragma unroll
        for (int k = 0; k < kOutPerWave; ++k) {
            const int u = 16 * (wave + 4 * k) + m;
            const int ux = u % kOutHX, uy = u / kOutHX;
            const int gx = x0 - 1 + ux, gy = y0 - 1 + uy;
            const bool ok = u < kOutU && gx >= 0 && gx < W && gy >= 0 && gy < H && z >= 0 && z < D;
            const long long vox = ((n * D + z) * H + gy) * (long long)W + gx;
            a[k] = load_frag<TIGRE_CNN_IN_BF16X32>(x, vox, q, ok);
        }
    };
    bf16x8 a[kOutPerWave];
    load_plane(z0 - 1, a);

    const int ox = x0 + lane, oy = y0 + wave;
    const bool out_ok = ox < W && oy < H;
    // accumulators of output planes zi - 1, zi, zi + 1 while input plane zi is processed
    float a_prev = 0.0f, a_cur = 0.0f, a_next = 0.0f;
    for (int zi = z0 - 1; zi <= z1; ++zi) {
        if (zi >= 0 && zi < D) {  // uniform per workgroup
#pragma unroll
            for (int k = 0; k < kOutPerWave; ++k) {
                const int blk = wave + 4 * k;
                if (blk >= kOutBlocks) break;  // uniform per wavefront
                const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
                const f32x4 r0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[k], b[0], zero, 0, 0, 0);
                const f32x4 r1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[k], b[1], zero, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int u = 16 * blk + 4 * q + r;
                    P[m * kOutStride + u] = r0[r];
                    if (16 + m < 27) P[(16 + m) * kOutStride + u] = r1[r];
                }
            }
            __syncthreads();  // P complete
            if (zi + 1 <= z1) load_plane(zi + 1, a);
            float s[3];
            const int c = (wave + 1) * kOutHX + lane + 1;  // this lane's voxel in the halo'd tile
#pragma unroll
            for (int g = 0; g < 3; ++g) {
                float t = 0.0f;
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int tap = 9 * g + 3 * (dy + 1) + (dx + 1);
                        t += P[tap * kOutStride + c + dy * kOutHX + dx];
                    }
                s[g] = t;
            }
            a_next += s[0];  // dz = -1: output plane zi + 1
            a_cur += s[1];   // dz = 0:  output plane zi
            a_prev += s[2];  // dz = +1: output plane zi - 1
            __syncthreads();  // P is read before the next plane overwrites it
        } else if (zi + 1 <= z1) {
            load_plane(zi + 1, a);
        }
        const int zo = zi - 1;  // complete: it has its dz = -1, 0 and +1 groups
        if (zo >= z0 && zo < z1 && out_ok) {
            const long long vox = ((n * D + zo) * H + oy) * (long long)W + ox;
            y[vox] = aux[vox] + (a_prev + bias[0]);
        }
        a_prev = a_cur;
        a_cur = a_next;
        a_next = 0.0f;
    }
}

// Convolution of a single FP32 channel (the image, or the loss gradient) into
// 32 channels.  With one input channel the GEMM k dimension of conv_kernel
// holds 1 useful channel of 32; here k is the tap instead: A[v][k] =
// x[v + tap k] (k < 27, zero above), B[k][o] = w[k][o][0], so one MFMA covers
// all taps of 16 voxels and 16 output channels.  The workgroup's 3 x 6 x 130
// block of the input is staged in LDS (zero outside the volume) and every
// lane gathers its 8 taps of each tile from there.
template <int kEpi>
__global__ __launch_bounds__(256) void conv_taps_kernel(const float* __restrict__ x,
                                                        const unsigned short* __restrict__ w,
                                                        const float* __restrict__ bias, const void* __restrict__ aux,
                                                        void* __restrict__ y, int N, int D, int H, int W) {
    constexpr int kPlane = kLdsRows * kLdsBX;
    __shared__ unsigned short blk[3 * kPlane];  // [dz][row][voxel], rounded to bf16 once when staged
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, q = lane >> 4;
    const int x0 = blockIdx.x * (16 * kLdsTiles), y0 = blockIdx.y * 4;
    const int zz0 = (int)(blockIdx.z % (unsigned)D);
    const long long n = blockIdx.z / (unsigned)D;
    const int yo = y0 + wave;

    for (int i = tid; i < 3 * kPlane; i += 256) {
        const int pz = i / kPlane, r = i - pz * kPlane, yi = r / kLdsBX, xi = r - yi * kLdsBX;
        const int gz = zz0 - 1 + pz, gy = y0 - 1 + yi, gx = x0 - 1 + xi;
        float v = 0.0f;
        if (gz >= 0 && gz < D && gy >= 0 && gy < H && gx >= 0 && gx < W)
            v = x[((n * D + gz) * H + gy) * (long long)W + gx];
        blk[i] = bf16_bits(v);
    }
    // this lane's output channels m and 16 + m
    float bias_lane[2] = {0.0f, 0.0f};
    if (kEpi == TIGRE_CNN_EPI_BIAS_RELU) {
        bias_lane[0] = bias[m];
        bias_lane[1] = bias[16 + m];
    }
    // B fragments: k = 8 q + j is the tap, column o = 16 hb + m
    bf16x8 b[2];
#pragma unroll
    for (int hb = 0; hb < 2; ++hb) {
        unsigned short bits[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int tap = 8 * q + j;
            bits[j] = tap < 27 ? w[((long long)tap * kC + 16 * hb + m) * kC] : (unsigned short)0;
        }
        const uint4 raw = make_uint4(bits[0] | ((unsigned)bits[1] << 16), bits[2] | ((unsigned)bits[3] << 16),
                                     bits[4] | ((unsigned)bits[5] << 16), bits[6] | ((unsigned)bits[7] << 16));
        b[hb] = __builtin_bit_cast(bf16x8, raw);
    }
    // this lane's taps: offsets in the block relative to (row wave, voxel m) of tile 0
    int toff[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int tap = 8 * q + j;
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;  // each + 1
        toff[j] = tap < 27 ? (dz * kLdsRows + wave + dy) * kLdsBX + m + dx : -1;
    }
    __syncthreads();

    if (yo >= H) return;
    const long long obase = ((n * D + zz0) * H + yo) * (long long)W;
    // one tile at a time (its MFMA pair, then its stores), so few registers are live and several
    // workgroups share a CU; a workgroup whose tiles lie inside the volume along x stores without
    // per-voxel tests
    auto run = [&](auto full_tag) {
        constexpr bool kFull = decltype(full_tag)::value;
#pragma unroll
        for (int t = 0; t < kLdsTiles; ++t) {
            unsigned short bits[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) bits[j] = toff[j] >= 0 ? blk[toff[j] + 16 * t] : (unsigned short)0;
            const uint4 raw = make_uint4(bits[0] | ((unsigned)bits[1] << 16), bits[2] | ((unsigned)bits[3] << 16),
                                         bits[4] | ((unsigned)bits[5] << 16), bits[6] | ((unsigned)bits[7] << 16));
            const bf16x8 a = __builtin_bit_cast(bf16x8, raw);
            f32x4 acc[2];
            acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[0], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[1], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
            for (int hb = 0; hb < 2; ++hb) {
                const int o = hb * 16 + m;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int xo = x0 + 16 * t + 4 * q + r;
                    if (!kFull && xo >= W) continue;
                    const long long vox = obase + xo;
                    float v = acc[hb][r];
                    if (kEpi == TIGRE_CNN_EPI_BIAS_RELU) {
                        v += bias_lane[hb];
                        v = v > 0.0f ? v : 0.0f;
                    } else {
                        const unsigned short am = static_cast<const unsigned short*>(aux)[vox * kC + o];
                        v = (am != 0 && (am & 0x8000u) == 0) ? v : 0.0f;  // activation > 0
                    }
                    static_cast<unsigned short*>(y)[vox * kC + o] = bf16_bits(v);
                }
            }
        }
    };
    if (x0 + 16 * kLdsTiles <= W)
        run(std::true_type{});
    else
        run(std::false_type{});
}

// Weight gradient.  Block: one tap (blockIdx.x, 27 = bias) and one chunk of
// x-row segments of 32 voxels (blockIdx.y); wavefront w owns the 16 x 16
// block (input channels 16 (w & 1) .., output channels 16 (w >> 1) ..).
constexpr int kLdsStride = 40;  // bf16 per LDS row: 32 voxels + 8 padding (80 bytes)

template <int kXin, int kDyin>
__global__ __launch_bounds__(256) void wgrad_kernel(const void* __restrict__ x, const void* __restrict__ dy,
                                                    float* __restrict__ part, int N, int D, int H, int W,
                                                    long long segs_per_chunk) {
    __shared__ __attribute__((aligned(16))) unsigned short xt[kC * kLdsStride];
    __shared__ __attribute__((aligned(16))) unsigned short dt[kC * kLdsStride];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, q = lane >> 4;
    const int tap = blockIdx.x;
    const int dz = tap / 9 - 1, dyy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int segW = (W + 31) / 32;
    const long long nseg = (long long)N * D * H * segW;
    const long long s0 = (long long)blockIdx.y * segs_per_chunk;
    const long long s1 = s0 + segs_per_chunk < nseg ? s0 + segs_per_chunk : nseg;
    const int ib = 16 * (wave & 1), ob = 16 * (wave >> 1);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    // staging role of this thread: tensor (x or dy), voxel of the segment, 8-channel group
    const bool stage_x = tid < 128;
    const int sv = (tid & 127) >> 2, sq = tid & 3;
    for (long long s = s0; s < s1; ++s) {
        const int xs = (int)(s % segW);
        const long long row = s / segW;
        const int y0 = (int)(row % H), z0 = (int)((row / H) % D);
        const long long n = row / ((long long)H * D);
        const int xv = xs * 32 + sv;
        unsigned short vals[8];
        if (stage_x) {
            if (tap == 27) {
                for (int j = 0; j < 8; ++j) vals[j] = 0;
                if (sq == 0 && xv < W) vals[0] = 0x3f80;  // 1.0
            } else {
                const int zz = z0 + dz, yy = y0 + dyy, xx = xv + dx;
                const bool ok = xv < W && zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 && xx < W;
                const long long vox = ((n * D + zz) * H + yy) * (long long)W + xx;
                const bf16x8 f = load_frag<kXin>(x, vox, sq, ok);
                const uint4 raw = __builtin_bit_cast(uint4, f);
                const unsigned r4[4] = {raw.x, raw.y, raw.z, raw.w};
                for (int j = 0; j < 4; ++j) {
                    vals[2 * j] = (unsigned short)(r4[j] & 0xffffu);
                    vals[2 * j + 1] = (unsigned short)(r4[j] >> 16);
                }
            }
            for (int j = 0; j < 8; ++j) xt[(8 * sq + j) * kLdsStride + sv] = vals[j];
        } else {
            const long long vox = row * (long long)W + xv;
            const bf16x8 f = load_frag<kDyin>(dy, vox, sq, xv < W);
            const uint4 raw = __builtin_bit_cast(uint4, f);
            const unsigned r4[4] = {raw.x, raw.y, raw.z, raw.w};
            for (int j = 0; j < 4; ++j) {
                vals[2 * j] = (unsigned short)(r4[j] & 0xffffu);
                vals[2 * j + 1] = (unsigned short)(r4[j] >> 16);
            }
            for (int j = 0; j < 8; ++j) dt[(8 * sq + j) * kLdsStride + sv] = vals[j];
        }
        __syncthreads();
        const bf16x8 a = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(xt + (ib + m) * kLdsStride + 8 * q));
        const bf16x8 b = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(dt + (ob + m) * kLdsStride + 8 * q));
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        __syncthreads();
    }
    // D[i][o]: row i = ib + 4 q + r, column o = ob + m; stored as part[chunk][tap][o][i]
    float* p = part + ((long long)blockIdx.y * 28 + tap) * kC * kC;
#pragma unroll
    for (int r = 0; r < 4; ++r) p[(ob + m) * kC + ib + 4 * q + r] = acc[r];
}

// out[i] = sum over chunks of part[c][i], in chunk order.
__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const float* __restrict__ part, int nchunk,
                                                           float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 28 * kC * kC) return;
    float s = 0.0f;
    for (int c = 0; c < nchunk; ++c) s += part[(long long)c * 28 * kC * kC + i];
    out[i] = s;
}

long long wgrad_chunks(long long nseg) {
    long long c = (nseg + 15) / 16;  // at least 16 segments per block
    if (c > 256) c = 256;
    return c < 1 ? 1 : c;
}

}  // namespace

size_t cnn_wgrad_workspace(int N, int D, int H, int W) {
    const long long nseg = (long long)N * D * H * ((W + 31) / 32);
    return (size_t)wgrad_chunks(nseg) * 28 * kC * kC * sizeof(float);
}

int launch_cnn_conv(const void* x, int x_mode, const unsigned short* w, const float* bias, const void* aux, void* y,
                    int epi, int N, int D, int H, int W, hipStream_t stream) {
    TIGRE_REQUIRE(N > 0 && D > 0 && H > 0 && W > 0, "empty volume");
    TIGRE_REQUIRE(x_mode == TIGRE_CNN_IN_BF16X32 || x_mode == TIGRE_CNN_IN_F32X1, "unknown input mode");
    TIGRE_REQUIRE(epi == TIGRE_CNN_EPI_BIAS_RELU || epi == TIGRE_CNN_EPI_RELU_MASK || epi == TIGRE_CNN_EPI_RESIDUAL,
                  "unknown epilogue");
    TIGRE_REQUIRE(epi == TIGRE_CNN_EPI_RELU_MASK || bias != nullptr, "missing bias");
    TIGRE_REQUIRE(epi == TIGRE_CNN_EPI_BIAS_RELU || aux != nullptr, "missing mask / residual input");
    TIGRE_REQUIRE((reinterpret_cast<uintptr_t>(x) & 15) == 0 && (reinterpret_cast<uintptr_t>(w) & 15) == 0,
                  "activations and weights must be 16-byte aligned");
    const long long rows = (long long)N * D * H;
    TIGRE_REQUIRE((rows + 3) / 4 <= 2147483647LL, "volume too large");
    dim3 grid((unsigned)((W + 16 * kTilesPerWave - 1) / (16 * kTilesPerWave)), (unsigned)((rows + 3) / 4));
#define TIGRE_CONV(I, E) hipLaunchKernelGGL((conv_kernel<I, E>), grid, dim3(256), 0, stream, x, w, bias, aux, y, N, D, H, W)
    if (x_mode == TIGRE_CNN_IN_BF16X32) {
        TIGRE_REQUIRE((long long)N * D <= 65535, "too many planes for one launch");
        const unsigned short* xa = static_cast<const unsigned short*>(x);
        const dim3 lgrid((unsigned)((W + 16 * kLdsTiles - 1) / (16 * kLdsTiles)), (unsigned)((H + 3) / 4),
                         (unsigned)(N * D));
        const long long zmz = (long long)N * ((D + kZmChunk - 1) / kZmChunk);
        TIGRE_REQUIRE(zmz <= 65535, "too many planes for one launch");
        const dim3 zgrid((unsigned)((W + 16 * kZmTiles - 1) / (16 * kZmTiles)), lgrid.y, (unsigned)zmz);
        TIGRE_REQUIRE((long long)H * W * kC <= 2147483647LL, "plane too large");
        if (epi != TIGRE_CNN_EPI_RESIDUAL)
            TIGRE_REQUIRE((reinterpret_cast<uintptr_t>(y) & 7) == 0 && (reinterpret_cast<uintptr_t>(aux) & 7) == 0,
                          "32-channel outputs and masks must be 8-byte aligned");
        if (epi == TIGRE_CNN_EPI_BIAS_RELU) {
            hipLaunchKernelGGL((conv_zm_kernel<TIGRE_CNN_EPI_BIAS_RELU>), zgrid, dim3(kZmThreads), 0, stream, xa, w, bias,
                               aux, y, N, D, H, W);
        } else if (epi == TIGRE_CNN_EPI_RELU_MASK) {
            hipLaunchKernelGGL((conv_zm_kernel<TIGRE_CNN_EPI_RELU_MASK>), zgrid, dim3(kZmThreads), 0, stream, xa, w, bias,
                               aux, y, N, D, H, W);
        } else {
            const long long zgrid = (long long)N * ((D + kOutZ - 1) / kOutZ);
            TIGRE_REQUIRE(zgrid <= 65535 && (H + kOutTY - 1) / kOutTY <= 65535, "volume too large for one launch");
            const dim3 ogrid((unsigned)((W + kOutTX - 1) / kOutTX), (unsigned)((H + kOutTY - 1) / kOutTY),
                             (unsigned)zgrid);
            hipLaunchKernelGGL(conv_out_kernel, ogrid, dim3(256), 0, stream, xa, w, bias,
                               static_cast<const float*>(aux), static_cast<float*>(y), N, D, H, W);
        }
    } else {
        TIGRE_REQUIRE((long long)N * D <= 65535, "too many planes for one launch");
        const float* xf = static_cast<const float*>(x);
        const dim3 lgrid((unsigned)((W + 16 * kLdsTiles - 1) / (16 * kLdsTiles)), (unsigned)((H + 3) / 4),
                         (unsigned)(N * D));
        if (epi == TIGRE_CNN_EPI_BIAS_RELU)
            hipLaunchKernelGGL((conv_taps_kernel<TIGRE_CNN_EPI_BIAS_RELU>), lgrid, dim3(256), 0, stream, xf, w, bias,
                               aux, y, N, D, H, W);
        else if (epi == TIGRE_CNN_EPI_RELU_MASK)
            hipLaunchKernelGGL((conv_taps_kernel<TIGRE_CNN_EPI_RELU_MASK>), lgrid, dim3(256), 0, stream, xf, w, bias,
                               aux, y, N, D, H, W);
        else TIGRE_CONV(TIGRE_CNN_IN_F32X1, TIGRE_CNN_EPI_RESIDUAL);
    }
#undef TIGRE_CONV
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

int launch_cnn_wgrad(const void* x, int x_mode, const void* dy, int dy_mode, float* dw, int N, int D, int H, int W,
                     void* ws, size_t ws_bytes, hipStream_t stream) {
    TIGRE_REQUIRE(N > 0 && D > 0 && H > 0 && W > 0, "empty volume");
    TIGRE_REQUIRE(x_mode == TIGRE_CNN_IN_BF16X32 || x_mode == TIGRE_CNN_IN_F32X1, "unknown input mode");
    TIGRE_REQUIRE(dy_mode == TIGRE_CNN_IN_BF16X32 || dy_mode == TIGRE_CNN_IN_F32X1, "unknown gradient mode");
    TIGRE_REQUIRE(ws != nullptr && ws_bytes >= cnn_wgrad_workspace(N, D, H, W), "workspace too small");
    const long long nseg = (long long)N * D * H * ((W + 31) / 32);
    const long long nchunk = wgrad_chunks(nseg);
    const long long per = (nseg + nchunk - 1) / nchunk;
    float* part = static_cast<float*>(ws);
    dim3 grid(28, (unsigned)nchunk);
#define TIGRE_WG(XI, DI) hipLaunchKernelGGL((wgrad_kernel<XI, DI>), grid, dim3(256), 0, stream, x, dy, part, N, D, H, W, per)
    if (x_mode == TIGRE_CNN_IN_BF16X32) {
        if (dy_mode == TIGRE_CNN_IN_BF16X32) TIGRE_WG(TIGRE_CNN_IN_BF16X32, TIGRE_CNN_IN_BF16X32);
        else TIGRE_WG(TIGRE_CNN_IN_BF16X32, TIGRE_CNN_IN_F32X1);
    } else {
        if (dy_mode == TIGRE_CNN_IN_BF16X32) TIGRE_WG(TIGRE_CNN_IN_F32X1, TIGRE_CNN_IN_BF16X32);
        else TIGRE_WG(TIGRE_CNN_IN_F32X1, TIGRE_CNN_IN_F32X1);
    }
#undef TIGRE_WG
    TIGRE_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((28 * kC * kC + 255) / 256), dim3(256), 0, stream, part, (int)nchunk,
                       dw);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

// ---------------------------------------------------------------------------
// Weight gradient without allocation.  The per-block partial sums go to a
// library-held device array sized for the most any volume needs (wgrad_chunks
// caps the chunks at 256: 256 x 28 x 32 x 32 floats, 28 MiB), one copy per GPU
// for the life of the process.  Every partial the reduction reads is written
// by the same call, so nothing needs clearing; calls that overlap on one GPU
// (two streams) share the array and race.  The array's device address is
// looked up once per device, with that device current; host code passes that
// address, never the symbol itself.
// ---------------------------------------------------------------------------
static __device__ __align__(256) float scratch[256 * 28 * kC * kC];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};

int scratch_address(void** out) {
    int dev = 0;
    TIGRE_HIP_TRY(hipGetDevice(&dev));
    TIGRE_REQUIRE(dev >= 0 && dev < kScratchDevices, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        TIGRE_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        TIGRE_REQUIRE(p != nullptr, "no device address for the weight-gradient scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TIGRE_OK;
}
}  // namespace

size_t cnn_wgrad_static_bytes() { return sizeof(scratch); }

int launch_cnn_wgrad_noalloc(const void* x, int x_mode, const void* dy, int dy_mode, float* dw, int N, int D, int H,
                             int W, hipStream_t stream) {
    TIGRE_REQUIRE(N > 0 && D > 0 && H > 0 && W > 0, "empty volume");
    const size_t need = cnn_wgrad_workspace(N, D, H, W);
    TIGRE_REQUIRE(need <= sizeof(scratch), "weight-gradient partials exceed the library's static scratch");
    void* ws = nullptr;
    TIGRE_CHECK(scratch_address(&ws));
    return launch_cnn_wgrad(x, x_mode, dy, dy_mode, dw, N, D, H, W, ws, sizeof(scratch), stream);
}

}  // namespace tigre
