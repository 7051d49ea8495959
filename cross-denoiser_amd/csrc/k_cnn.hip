// Bandwidth CNN of the online training step on gfx950 MFMA: five 3x3 'same' convolutions
// 15 -> 32 -> 32 -> 32 -> 32 -> 6 with ReLU between them, BF16 operands, FP32 accumulation.
//
// The convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16 (M = 32 pixels of one row,
// N = 32 output channels, K = 16 input channels; 16x16x32 where a side has only 16 channels):
//   conv3x3_l1        layer 1 forward: the five FP32 3-channel input images (the concatenation
//                     and the bf16 rounding fused into the staging) -> bias + ReLU -> bf16.
//                     Tile: 16 rows x 32 px per 256-lane workgroup plus a 1-px halo in LDS;
//                     each B fragment read from LDS feeds up to three output rows. Persistent,
//                     software-pipelined blocks; the kernel fragments sit in LDS.
//   cnn5_fwd_stream   layers 2..5 forward in one launch: a block walks a strip of rows, each of
//                     its four waves computing one layer's row per step from 4-row LDS rings
//                     of the layer below (the hidden activations are written once and never
//                     read back); kernel fragments in registers.
//   conv3x3_bwd       backward of layers 5..3: the data gradient into the layer below (conv3x3
//                     scheme on the flipped, transposed kernel) and the weight gradient of one
//                     16x32-px tile from one staging of dY and the layer input X, eight waves per
//                     block. The ReLU derivative of the data gradient is read off X itself
//                     (x > 0 in bf16), so the forward stores no mask. Weight gradients take both
//                     operands staged pixel-major in LDS as they lie in HBM and read them
//                     channel-major with the transposing ds_read_b64_tr_b16.
//   conv3x3_bwd12     backward of layer 2 the same way, with layer 1's weight gradient computed
//                     on the same tile from layer 1's pre-activation gradient kept in LDS and the
//                     five input images (so that gradient is never written to HBM).
//   wgrad_reduce1/2   deterministic two-stage reduction of the per-block weight-gradient partials.
//   pack_weights      FP32 master weights -> bf16 MFMA B fragments for the forward and the data
//                     gradient (flipped/transposed) convolutions.
#include "xdn_common.h"
#include "xdn_cnn.h"

namespace xdn {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;

typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;

// Two floats -> packed bf16 {a, b}, round to nearest even (one v_cvt_pk_bf16_f32; finite inputs).
__device__ __forceinline__ uint32_t pack2(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}

__device__ __forceinline__ f32x16 mfma(u32x4 a, u32x4 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// ReLU of two packed bf16 values: a negative bf16 (sign bit set) reads as a negative int16, so a
// packed signed 16-bit max with 0 turns it into +0 and keeps every other value (one v_pk_max_i16
// instead of two FP32 maxima before the conversion; rounding is monotonic, so the result is the same).
__device__ __forceinline__ uint32_t relu_bf2(uint32_t w)
{
    typedef __attribute__((ext_vector_type(2))) short s16x2;
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, w), s16x2{0, 0}));
}

// Raw buffer accesses: a load outside [0, bytes) of its buffer returns zero and a store there is
// dropped, so guarded accesses (image borders, partial tiles, optional outputs) need no branch. With
// no branch around any vector memory access, the compiler knows exactly how many of them were issued
// after a prefetch and waits for the prefetch alone, not for the stores issued behind it.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef __attribute__((ext_vector_type(2))) unsigned int u32x2;
typedef __attribute__((ext_vector_type(3))) unsigned int u32x3;
constexpr uint32_t kOob = 0x80000000u;  // offset past every buffer (all are smaller than 2 GiB)
__device__ __forceinline__ rsrc_t make_rsrc(const void *p, size_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, p ? (int)bytes : 0, 0x00020000);
}
__device__ __forceinline__ u32x4 bld16(rsrc_t r, uint32_t off) { return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0); }
__device__ __forceinline__ u32x3 bld12(rsrc_t r, uint32_t off) { return __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, 0); }
__device__ __forceinline__ u32x2 bld8(rsrc_t r, uint32_t off) { return __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0); }
__device__ __forceinline__ uint32_t bld4(rsrc_t r, uint32_t off) { return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0); }
__device__ __forceinline__ void bst16(u32x4 v, rsrc_t r, uint32_t off) { __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, 0); }
__device__ __forceinline__ void bst8(u32x2 v, rsrc_t r, uint32_t off) { __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, 0); }
__device__ __forceinline__ void bst4(uint32_t v, rsrc_t r, uint32_t off) { __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, 0); }

// ---------------------------------------------------------------------------------------------
// First layer forward
// ---------------------------------------------------------------------------------------------
struct ConvBatch {
    TileMap map;
    ConvFrame f[XDN_MAX_FRAMES];
};
static_assert(sizeof(ConvBatch) <= 4096, "kernel argument too large");

constexpr int CT_W = 32, CT_H = 16;

// Layer 1: the five FP32 3-channel images of the network input (15 channels; the concatenation and
// the bf16 rounding are fused into the staging, channel 15 is zero) -> bias + ReLU -> bf16 32
// channels. Tile: 16 rows x 32 px per 256-lane workgroup, the tile plus a 1-px halo staged in LDS as
// NHWC bf16 with a padded pixel stride (conflict-free ds_read_b128).
// Persistent and software-pipelined: a block walks the tiles it, it + gridDim.x, ... (gridDim.x a
// multiple of 8, so a block stays on one XCD and xcd_tile keeps each XCD on a contiguous run of
// tiles). The global loads of the next tile are issued into registers before the MFMAs of the
// current one and land in LDS at the end of the iteration. The kernel fragments of the tile's
// frame sit in LDS (reloaded when the frame changes) and are the MFMA A operands; the B operands
// are 32 pixels x 16 input channels read from the LDS tile, so the accumulators hold C[channel][pixel]
// and a lane owns 16 channels c(k) = 8(k/4) + 4h + k%4 of one pixel: the epilogue works in
// registers, and a permlane32 swap between the two half-waves gives each lane 2 x 8 consecutive
// channels for 16-B stores. Each wave owns 4 output rows; every B fragment loaded from LDS feeds
// the MFMAs of up to three output rows (the three kernel rows).
__global__ __launch_bounds__(256, 2) void conv3x3_l1(ConvBatch b)
{
    constexpr int PS = 16 + 8;  // LDS pixel stride in bf16 (12 dwords: conflict-free b128 reads)
    constexpr int LW = CT_W + 2, LH = CT_H + 2;
    __shared__ __attribute__((aligned(16))) uint16_t lds[LH * LW * PS];
    __shared__ u32x4 wl[9 * 64];  // kernel fragments of the current frame: [tap][lane]
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6, px = lane & 31, h = lane >> 5;
    const int total = b.map.total;

    // ---- input tile (+1 px halo, zero outside the image): fetch to registers, commit to LDS ----
    constexpr int NV = (LH * LW + 255) / 256;
    float fv[NV][15];
    // every load is issued (zero outside the image: offset kOob), so the loop's memory accesses
    // carry no branches
    auto fetch = [&](int it) {
        const TileLoc T = locate(b.map, xcd_tile(it, total));
        const ConvFrame &G = b.f[T.f];
        const int gH = G.H, gW = G.W, tx0 = T.tx * CT_W, ty0 = T.ty * CT_H;
        const size_t npx = (size_t)gH * gW;
        rsrc_t rin[5];
#pragma unroll
        for (int im = 0; im < 5; ++im) rin[im] = make_rsrc(G.in[im], npx * 12);
#pragma unroll
        for (int u = 0; u < NV; ++u)
        {
            const int pix = tid + 256 * u;
            const int ly = pix / LW, lx = pix - ly * LW;
            const int gy = ty0 + ly - 1, gx = tx0 + lx - 1;
            const bool ok = pix < LH * LW && gy >= 0 && gy < gH && gx >= 0 && gx < gW;
            const uint32_t p = (uint32_t)(gy * gW + gx);
#pragma unroll
            for (int im = 0; im < 5; ++im)
            {
                const u32x3 v = bld12(rin[im], ok ? p * 12u : kOob);
                fv[u][3 * im] = __uint_as_float(v.x);
                fv[u][3 * im + 1] = __uint_as_float(v.y);
                fv[u][3 * im + 2] = __uint_as_float(v.z);
            }
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int u = 0; u < NV; ++u)
        {
            const int pix = tid + 256 * u;
            if (pix < LH * LW)
            {
                const float *v = fv[u];
                u32x4 a, c;
                a.x = pack2(v[0], v[1]); a.y = pack2(v[2], v[3]); a.z = pack2(v[4], v[5]); a.w = pack2(v[6], v[7]);
                c.x = pack2(v[8], v[9]); c.y = pack2(v[10], v[11]); c.z = pack2(v[12], v[13]);
                c.w = pack2(v[14], 0.f);
                *reinterpret_cast<u32x4 *>(&lds[pix * PS]) = a;
                *reinterpret_cast<u32x4 *>(&lds[pix * PS + 8]) = c;
            }
        }
    };

    // The frame's kernel fragments and bias go to LDS when the frame changes. A tile's staging is
    // committed at the end of the previous iteration, behind its stores: waiting for the prefetch
    // there does not wait for those stores.
    __shared__ float4 wb[8];  // bias of the current frame, [32]
    auto frame_setup = [&](int fi) {
        const ConvFrame &G = b.f[fi];
        for (int e = tid; e < 9 * 64; e += 256) wl[e] = static_cast<const u32x4 *>(G.wfrag)[e];
        if (tid < 32) reinterpret_cast<float *>(wb)[tid] = G.bias[tid];
    };
    int cur_f = -1;
    if ((int)blockIdx.x < total)
    {
        fetch(blockIdx.x);
        cur_f = locate(b.map, xcd_tile(blockIdx.x, total)).f;
        frame_setup(cur_f);
        commit();
    }
    for (int it = blockIdx.x; it < total; it += gridDim.x)
    {
    const TileLoc L = locate(b.map, xcd_tile(it, total));
    const ConvFrame &F = b.f[L.f];
    const int H = F.H, W = F.W;
    const int x0 = L.tx * CT_W, y0 = L.ty * CT_H;
    __syncthreads();  // tile `it` and its frame's fragments are in LDS
    const int nxt = it + (int)gridDim.x;
    // the next tile's loads (the last iteration reloads its own tile: no branch around the loads),
    // kept ahead of the MFMAs: left to itself the scheduler sinks them behind the whole tile, and
    // the commit at the end of the iteration then waits out their full latency
    fetch(nxt < total ? nxt : it);
    __builtin_amdgcn_sched_barrier(0);
    float bias[16];  // bias of the lane's channels c(k) = 8(k/4) + 4h + k%4
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        const float4 bv = wb[2 * j + h];
        bias[4 * j] = bv.x; bias[4 * j + 1] = bv.y; bias[4 * j + 2] = bv.z; bias[4 * j + 3] = bv.w;
    }

    f32x16 acc[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int k = 0; k < 16; ++k) acc[r][k] = 0.f;
    // wave wv owns output rows 4wv..4wv+3, which read LDS rows 4wv..4wv+5: each B fragment (LDS row
    // lr, column shift dx) is read once and feeds the kernel rows dy with output row lr - dy in 0..3
#pragma unroll
    for (int dx = 0; dx < 3; ++dx)
    {
        u32x4 wk[3], bx[6];
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) wk[dy] = wl[(dy * 3 + dx) * 64 + lane];
#pragma unroll
        for (int lr = 0; lr < 6; ++lr)
            bx[lr] = *reinterpret_cast<const u32x4 *>(&lds[((4 * wv + lr) * LW + px + dx) * PS + 8 * h]);
#pragma unroll
        for (int lr = 0; lr < 6; ++lr)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
            {
                const int orow = lr - dy;
                if (orow >= 0 && orow < 4) acc[orow] = mfma(wk[dy], bx[lr], acc[orow]);
            }
    }

    // ---- epilogue, per output row: lane (px, h) holds channels c(k) of pixel (row, x0 + px) ----
    const rsrc_t rout = make_rsrc(F.out, (size_t)H * W * 64);
#pragma unroll
    for (int orow = 0; orow < 4; ++orow)
    {
        const int gy = y0 + 4 * wv + orow, gx = x0 + px;
        const bool in = gy < H && gx < W;
        const uint32_t p = (uint32_t)(gy * W + gx);
        uint32_t d[8];  // d[2j + e] = channels 8j + 4h + 2e, +1
#pragma unroll
        for (int q = 0; q < 8; ++q)
            d[q] = relu_bf2(pack2(acc[orow][2 * q] + bias[2 * q], acc[orow][2 * q + 1] + bias[2 * q + 1]));
        // lane h gathers channels 16g + 8h .. 16g + 8h + 7 (g = 0, 1): the low half-wave keeps its
        // group 2g and takes the high half's; the high half keeps group 2g + 1 and takes the low's
        u32x4 v[2];
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int e = 0; e < 2; ++e)
            {
                const auto sw = __builtin_amdgcn_permlane32_swap(d[4 * g + e], d[4 * g + 2 + e], false, false);
                v[g][e] = (uint32_t)sw[0];
                v[g][2 + e] = (uint32_t)sw[1];
            }
        const uint32_t o = in ? p * 64u + 16u * h : kOob;
        bst16(v[0], rout, o);
        bst16(v[1], rout, in ? o + 32u : kOob);
    }
    __syncthreads();  // every wave is done reading tile `it`
    if (nxt < total)
    {
        const int nf = locate(b.map, xcd_tile(nxt, total)).f;
        if (nf != cur_f)
        {
            frame_setup(nf);
            cur_f = nf;
        }
        commit();
    }
    }  // tile loop
}

// ---------------------------------------------------------------------------------------------
// Layers 2..5 of the forward, streamed down the rows
// ---------------------------------------------------------------------------------------------
constexpr int FS_VALID = 26;         // output columns per strip (32-px windows lose 1 px a side per layer)
constexpr int FS_PX = CT_W + 2;      // ring pixels: window columns -1 .. 32
constexpr int FS_PS = 40;            // ring pixel stride in bf16 (conflict-free b128 reads)
constexpr int FS_SKEW = 9;           // steps from a strip's first layer-2 row to its first output row

struct FwdStreamBatch {
    int n;
    int blk_begin[XDN_MAX_FRAMES + 1];
    int nstrip[XDN_MAX_FRAMES], seg_rows[XDN_MAX_FRAMES];
    FwdStreamFrame f[XDN_MAX_FRAMES];
};
static_assert(sizeof(FwdStreamBatch) <= 4096, "kernel argument too large");

// A block owns a strip of FS_VALID output columns [x0, x0 + 26) and a segment of rows [ys, ye).
// Every layer runs on the same 32-px window [x0 - 3, x0 + 29); layer k's row is exact on
// [x0 - 5 + k, x0 + 31 - k), so layer 5 is exact on the strip. Wave w computes layer w + 2: at step
// s its row is r = ys - 3 + s - 2w, read from the previous layer's rows r - 1 .. r + 1 in a
// 4-row LDS ring (written in earlier steps; the previous layer writes row r + 2 in this step), and
// writes it to its own ring (zero outside the image: the next layer's 'same' padding). Layer 1's
// output rows stream in from HBM two steps ahead through registers. Each layer stores the rows of
// [ys, ye) and the columns of the strip, with the MFMAs of an output row in a fixed order (column
// shift, channel chunk, kernel row), so a row is the same whichever strip or segment computes it.
// One barrier per step. Kernel fragments and bias of the wave's layer sit in registers.
__global__ __launch_bounds__(256, 3) void cnn5_fwd_stream(FwdStreamBatch b)
{
    __shared__ __attribute__((aligned(16))) uint16_t ring[4][4][FS_PX * FS_PS];  // [input of layer w + 2][slot]
    const int tid = threadIdx.x, lane = tid & 63, px = lane & 31, h = lane >> 5;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    int f = 0;
    while (f + 1 < b.n && (int)blockIdx.x >= b.blk_begin[f + 1]) ++f;
    const FwdStreamFrame &F = b.f[f];
    const int li = blockIdx.x - b.blk_begin[f], nblk = b.blk_begin[f + 1] - b.blk_begin[f];
    const int nstrip = b.nstrip[f];
    const int tile = xcd_tile(li, nblk);  // strip-fastest: neighbouring strips share an XCD
    const int strip = tile % nstrip, seg = tile / nstrip;
    const int H = F.H, W = F.W;
    const int x0 = strip * FS_VALID, xw = x0 - 3;
    const int ys = seg * b.seg_rows[f], ye = min(H, ys + b.seg_rows[f]);

    for (int i = tid; i < 4 * 4 * FS_PX * FS_PS / 8; i += 256)
        reinterpret_cast<u32x4 *>(&ring[0][0][0])[i] = u32x4{0u, 0u, 0u, 0u};

    u32x4 wreg[18];  // [tap][chunk] fragments of this wave's layer
    {
        const u32x4 *wf = static_cast<const u32x4 *>(F.wfrag[wv]);
#pragma unroll
        for (int i = 0; i < 18; ++i) wreg[i] = wf[i * 64 + lane];
    }
    float bias[16];  // channels c(k) = 8(k/4) + 4h + k%4
#pragma unroll
    for (int k = 0; k < 16; ++k) bias[k] = F.bias[wv][8 * (k >> 2) + 4 * h + (k & 3)];

    const size_t npx = (size_t)H * W;
    const rsrc_t rx1 = make_rsrc(F.x1, npx * 64);
    const rsrc_t rout = make_rsrc(wv < 3 ? static_cast<const void *>(F.x[wv < 3 ? wv : 0]) : F.out,
                                  npx * (wv < 3 ? 64 : 24));
    // layer 1's rows: chunk tid (< 136) = pixel tid / 4 of the ring (column xw - 1 + tid / 4), quarter tid % 4
    const int lpx = tid >> 2, lq = tid & 3, lcol = xw - 1 + lpx;
    const bool lrow = tid < FS_PX * 4 && lcol >= 0 && lcol < W;
    auto xoff = [&](int y) -> uint32_t {
        return lrow && y >= 0 && y < H ? (uint32_t)((y * W + lcol) * 32 + lq * 8) * 2u : kOob;
    };
    auto put = [&](int y, u32x4 v) {
        if (tid < FS_PX * 4) *reinterpret_cast<u32x4 *>(&ring[0][y & 3][lpx * FS_PS + lq * 8]) = v;
    };
    __syncthreads();
    {
        const u32x4 a = bld16(rx1, xoff(ys - 4)), c = bld16(rx1, xoff(ys - 3)), d = bld16(rx1, xoff(ys - 2));
        put(ys - 4, a);
        put(ys - 3, c);
        put(ys - 2, d);
    }
    // Two rows in flight, each followed by four (dropped) stores as in a step: the steps' waits for a
    // prefetch then count the same accesses behind it whether it was issued here or in a step.
    u32x4 pf0 = bld16(rx1, xoff(ys - 1));
    const u32x4 zero4 = {0u, 0u, 0u, 0u};
    bst16(zero4, rout, kOob); bst16(zero4, rout, kOob + 32u); bst8(u32x2{0u, 0u}, rout, kOob + 64u);
    u32x4 pf1 = bld16(rx1, xoff(ys));
    bst16(zero4, rout, kOob + 96u); bst16(zero4, rout, kOob + 128u); bst8(u32x2{0u, 0u}, rout, kOob + 160u);
    __syncthreads();

    // the lane's B-fragment address in slot 0 of its input ring: a slot adds a wave-uniform offset
    // and the column shift / channel chunk an immediate one
    const uint16_t *in = &ring[wv][0][0] + px * FS_PS + 8 * h;
    const int gx = xw + px;
    const bool cin = gx >= 0 && gx < W;
    const bool ccol = cin && px >= 3 && px < 3 + FS_VALID;  // a column this strip stores
    auto step = [&](int s, u32x4 &pf) {
        put(ys - 1 + s, pf);                // layer 1's row ys - 1 + s, loaded two steps ago
        pf = bld16(rx1, xoff(ys + 1 + s));  // ... and the one two steps ahead
        const int r = ys - 3 + s - 2 * wv;
        const uint16_t *rows[3];
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) rows[dy] = in + ((r - 1 + dy) & 3) * (FS_PX * FS_PS);
        f32x16 acc;
#pragma unroll
        for (int k = 0; k < 16; ++k) acc[k] = 0.f;
#pragma unroll
        for (int dx = 0; dx < 3; ++dx)
#pragma unroll
            for (int kc = 0; kc < 2; ++kc)
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
                {
                    const u32x4 bfr = *reinterpret_cast<const u32x4 *>(rows[dy] + dx * FS_PS + kc * 16);
                    acc = mfma(wreg[(dy * 3 + dx) * 2 + kc], bfr, acc);
                }
        const bool rin = r >= 0 && r < H;
        const bool st = rin && ccol && r >= ys && r < ye;
        const uint32_t p = (uint32_t)(r * W + gx);
        // Every wave issues the same four stores (the unused ones at kOob), so the number of memory
        // accesses behind a prefetch is the same on every path and its wait excludes the stores.
        // Outside the image the row is zero (the next layer's padding): ReLU and that mask are one
        // median, med3(x, 0, +inf) = max(x, 0) and med3(x, 0, 0) = 0.
        const float lim = rin && cin ? __builtin_inff() : 0.f;
        float v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = __builtin_amdgcn_fmed3f(acc[k] + bias[k], 0.f, lim);
        uint32_t d[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) d[q] = pack2(v[2 * q], v[2 * q + 1]);
        u32x4 o2[2];
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int e = 0; e < 2; ++e)
            {
                const auto sw = __builtin_amdgcn_permlane32_swap(d[4 * g + e], d[4 * g + 2 + e], false, false);
                o2[g][e] = (uint32_t)sw[0];
                o2[g][2 + e] = (uint32_t)sw[1];
            }
        const bool l5 = wv == 3;
        // layer 5: the network output, channels 0..3 on the h = 0 lane (16 B), 4..5 on h = 1 (8 B)
        const float f0 = acc[0] + bias[0], f1 = acc[1] + bias[1], f2 = acc[2] + bias[2], f3 = acc[3] + bias[3];
        const u32x4 sA = l5 ? u32x4{__float_as_uint(f0), __float_as_uint(f1), __float_as_uint(f2), __float_as_uint(f3)}
                            : o2[0];
        const uint32_t oA = !st ? kOob : (l5 ? (h == 0 ? p * 24u : kOob) : p * 64u + 16u * h);
        const uint32_t oB = st && !l5 ? p * 64u + 16u * h + 32u : kOob;
        const uint32_t oC = st && l5 && h == 1 ? p * 24u + 16u : kOob;
        bst16(sA, rout, oA);
        bst16(o2[1], rout, oB);
        bst8(u32x2{__float_as_uint(f0), __float_as_uint(f1)}, rout, oC);
        if (!l5)
        {
            uint16_t *nr = &ring[wv + 1][r & 3][(px + 1) * FS_PS + 8 * h];
            *reinterpret_cast<u32x4 *>(nr) = o2[0];
            *reinterpret_cast<u32x4 *>(nr + 16) = o2[1];
        }
        __syncthreads();
    };
    // an even number of steps (an extra last step computes rows no layer stores): the loop body is
    // two steps with no branch, so the accesses behind each prefetch are the same every time
    const int S = (ye - ys + FS_SKEW + 1) & ~1;
    for (int s = 0; s < S; s += 2)
    {
        step(s, pf0);
        step(s + 1, pf1);
    }
}

// ---------------------------------------------------------------------------------------------
// Weight gradient
// ---------------------------------------------------------------------------------------------
typedef short v4s __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v4s lds_v4s;

// Transposing LDS read (ds_read_b64_tr_b16). In each group of 16 lanes, lane L receives element
// L%4 of the 8-byte blocks addressed by the lanes 4j + L/4 (j = 0..3). So when lane s of a group
// addresses pixel p0 + s/4, channels c0 + 4(s%4) .. +3 of a pixel-major tile, lane L receives
// channel c0 + L of the pixels p0 .. p0+3: four consecutive K values of an MFMA operand.
__device__ __forceinline__ uint2 ld_tr(const uint16_t *p)
{
    return __builtin_bit_cast(uint2, __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s *)p));
}
// The same on an LDS-address-space pointer: offsets added to it stay 32-bit and constant ones go
// into the instruction's offset field.
typedef __attribute__((address_space(3))) uint16_t lds_u16;
__device__ __forceinline__ uint2 ld_tr(lds_u16 *p)
{
    return __builtin_bit_cast(uint2, __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s *)p));
}

// acc + the two bf16 halves of w, in FP32 (v_dot2_f32_bf16 against (1, 1)).
__device__ __forceinline__ float add_bf2(uint32_t w, float acc)
{
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, w), bf16x2{(__bf16)1.f, (__bf16)1.f}, acc, false);
}
__device__ __forceinline__ float add_bf8(u32x4 w, float acc)
{
    return add_bf2(w.w, add_bf2(w.z, add_bf2(w.y, add_bf2(w.x, acc))));
}

// 0xFFFF in each 16-bit half of w that holds a nonzero bf16 (+0 and -0 count as zero), else 0:
// the half's magnitude plus 0x7FFF sets its bit 15 exactly when the magnitude is nonzero, and an
// arithmetic shift by 15 spreads that bit (packed 16-bit integer ops).
__device__ __forceinline__ uint32_t nz_mask(uint32_t w)
{
    typedef __attribute__((ext_vector_type(2))) unsigned short u16x2;
    typedef __attribute__((ext_vector_type(2))) short s16x2;
    const u16x2 t = __builtin_bit_cast(u16x2, w & 0x7FFF7FFFu) + u16x2{0x7FFF, 0x7FFF};
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(s16x2, t) >> s16x2{15, 15});
}

template <int N> struct IntC { static constexpr int value = N; };

// ---------------------------------------------------------------------------------------------
// Fused backward of a 32 -> 32 layer
// ---------------------------------------------------------------------------------------------
struct BwdBatch {
    int n;
    int blk_begin[XDN_MAX_FRAMES + 1];
    BwdFrame f[XDN_MAX_FRAMES];
};
static_assert(sizeof(BwdBatch) <= 4096, "kernel argument too large");

// One staging of a 16x32-px tile of dY and of the layer input X (both with their 1-px halo, zero
// outside the image) feeds both products of the layer's backward pass:
//   data gradient  dX = conv(dY, flipped/transposed W), kept where X's bf16 value is nonzero (the
//                  ReLU derivative, read from the staged X tile); B operands read from the dY tile
//                  with ds_read_b128, as conv3x3_l1 reads its input tile;
//   weight gradient dW[t] += X(p + shift_t)^T dY(p) over the tile's pixels: transposing reads of
//                  both tiles (M = input channel, N = output channel, K = pixels), accumulators
//                  kept in registers across the block's tiles.
// dY is read from HBM once instead of once per product. Eight waves share the work of a tile:
// wave w computes the data gradient of tile rows 2w, 2w+1 and the weight gradient of taps
// 0..4 (w < 4) or 5..8 (w >= 4) over tile rows 4(w%4) .. 4(w%4)+3, so each wave holds 80 + 32
// accumulators and the next tile's prefetch, which fits two waves per SIMD (one 512-lane block per
// CU: the two staged tiles and the kernel fragments take 117 KB of LDS; both tiles have an 80-B
// pixel stride, so a lane row reading one pixel each is free of bank conflicts beyond 2-way). The block walks a fixed
// set of its frame's tiles (xcd_strips over tiles) and writes one partial, so a frame's gradient is
// the same alone and in a batch.
// L5 (the last layer, 32 -> 6): dY is the FP32 6-channel gradient of the network output, staged as
// bf16 16 channels (zero padded); the data gradient has one 16-channel K chunk, and the weight
// gradient runs on v_mfma_f32_16x16x32_bf16 (M = 16 input channels, two tiles; N = the 16 output
// channels; K = the 32 px of one tile row).
template <bool L5>
__global__ __launch_bounds__(512, 1) void conv3x3_bwd(BwdBatch b)
{
    constexpr int PS = L5 ? 24 : 40;       // dY pixel stride in bf16 (conflict-free b128 reads)
    constexpr int XS = 40;                 // X pixel stride in bf16 (conflict-free reads of a pixel row)
    constexpr int KC = L5 ? 1 : 2;
    constexpr int LW = CT_W + 2, LH = CT_H + 2;
    constexpr int NCH = LH * LW * 4;       // 16-B chunks of one staged X tile (and dY tile, middle layers)
    constexpr int NT = 512;
    constexpr int NV = (NCH + NT - 1) / NT;
    constexpr int NVD = L5 ? (LH * LW + NT - 1) / NT : NV;  // dY staging items per thread (L5: pixels)
    __shared__ __attribute__((aligned(16))) uint16_t dys[LH * LW * PS];
    __shared__ __attribute__((aligned(16))) uint16_t xs[LH * LW * XS];
    __shared__ u32x4 wl[9 * KC * 64];
    __shared__ float dbs[4][32];
    static_assert(LH * LW * PS * 2 >= 4 * 1024 * 4, "partial reduction space");
    int f = 0;
    while (f + 1 < b.n && (int)blockIdx.x >= b.blk_begin[f + 1]) ++f;
    const BwdFrame &F = b.f[f];
    const int li = blockIdx.x - b.blk_begin[f], nblk = b.blk_begin[f + 1] - b.blk_begin[f];
    const int H = F.H, W = F.W;
    const int ntx = (W + CT_W - 1) / CT_W, nty = (H + CT_H - 1) / CT_H;
    const int tid = threadIdx.x, lane = tid & 63, px = lane & 31, h = lane >> 5, ch = lane & 31;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int rg = wv & 3, tg = wv >> 2;  // weight-gradient row group and tap group
    const int grp = (lane >> 4) & 1, s16 = lane & 15, jq = s16 >> 2, qq = s16 & 3, kg = lane >> 4;
    const int cb = 16 * grp + 4 * qq;      // channel offset addressed by this lane in transposing reads
    const uint16_t *gx = static_cast<const uint16_t *>(F.x);

    for (int e = tid; e < 9 * KC * 64; e += NT) wl[e] = static_cast<const u32x4 *>(F.wfrag)[e];

    // The rest of the kernel once per tap group, a compile-time constant on each side of a
    // wave-uniform branch: every tap's shift is then an immediate LDS offset, and each group keeps
    // its own weight-gradient accumulators across the tile loop (one shared loop would have to
    // merge the two groups' registers after every tile). Both sides run the same barriers.
    auto run = [&](auto tgc) {
    constexpr int TG = decltype(tgc)::value;
    typedef __attribute__((ext_vector_type(4))) float f32x4;
    f32x16 wacc[L5 ? 1 : 5];         // taps 5 TG + j (middle layers)
    f32x4 wacc16[L5 ? 5 : 1][2];     // L5: taps 5 TG + j, input-channel tiles m
#pragma unroll
    for (int j = 0; j < (L5 ? 1 : 5); ++j)
#pragma unroll
        for (int k = 0; k < 16; ++k) wacc[j][k] = 0.f;
#pragma unroll
    for (int j = 0; j < (L5 ? 5 : 1); ++j)
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int k = 0; k < 4; ++k) wacc16[j][m][k] = 0.f;
    float db = 0.f;

    // Branch-free buffer accesses (zero / dropped outside the image), so the wait for a prefetch
    // at its commit, placed behind the tile's stores, does not wait for those stores.
    const size_t npx = (size_t)H * W;
    const rsrc_t rdy = make_rsrc(F.dy, npx * (L5 ? 24 : 64)), rxx = make_rsrc(gx, npx * 64);
    const rsrc_t rdx = make_rsrc(F.dx, npx * 64);
    u32x4 rd[L5 ? 1 : NV], rx[NV];
    float fd[L5 ? NVD : 1][6];
    auto fetch = [&](int t) {
        const int y0 = (t / ntx) * CT_H, x0 = (t % ntx) * CT_W;
#pragma unroll
        for (int u = 0; u < NV; ++u)
        {
            const int i = tid + NT * u;
            const int pix = i >> 2, q = i & 3;
            const int ly = pix / LW, lx = pix - ly * LW;
            const int gy = y0 + ly - 1, gxx = x0 + lx - 1;
            const bool ok = i < NCH && gy >= 0 && gy < H && gxx >= 0 && gxx < W;
            const uint32_t o = ok ? (uint32_t)((gy * W + gxx) * 32 + q * 8) * 2u : kOob;
            if constexpr (!L5) rd[u] = bld16(rdy, o);
            rx[u] = bld16(rxx, o);
        }
        if constexpr (L5)
#pragma unroll
            for (int u = 0; u < NVD; ++u)
            {
                const int pix = tid + NT * u;
                const int ly = pix / LW, lx = pix - ly * LW;
                const int gy = y0 + ly - 1, gxx = x0 + lx - 1;
                const bool ok = pix < LH * LW && gy >= 0 && gy < H && gxx >= 0 && gxx < W;
                const uint32_t p = (uint32_t)(gy * W + gxx);
#pragma unroll
                for (int k = 0; k < 3; ++k)
                {
                    const u32x2 v = bld8(rdy, ok ? p * 24u + 8u * k : kOob);
                    fd[u][2 * k] = __uint_as_float(v.x);
                    fd[u][2 * k + 1] = __uint_as_float(v.y);
                }
            }
    };
    auto commit = [&]() {
#pragma unroll
        for (int u = 0; u < NV; ++u)
        {
            const int i = tid + NT * u;
            if (i < NCH)
            {
                const int pix = i >> 2, q = i & 3;
                if constexpr (!L5) *reinterpret_cast<u32x4 *>(&dys[pix * PS + q * 8]) = rd[u];
                *reinterpret_cast<u32x4 *>(&xs[pix * XS + q * 8]) = rx[u];
            }
        }
        if constexpr (L5)
#pragma unroll
            for (int u = 0; u < NVD; ++u)
            {
                const int pix = tid + NT * u;
                if (pix < LH * LW)
                {
                    const u32x4 a = {pack2(fd[u][0], fd[u][1]), pack2(fd[u][2], fd[u][3]), pack2(fd[u][4], fd[u][5]), 0u};
                    *reinterpret_cast<u32x4 *>(&dys[pix * PS]) = a;
                    *reinterpret_cast<u32x4 *>(&dys[pix * PS + 8]) = u32x4{0u, 0u, 0u, 0u};
                }
            }
    };

    const StripWalk sw = xcd_strips(li, nblk, ntx * nty);
    if (sw.first < sw.end)
    {
        fetch(sw.first);
        commit();
    }
    for (int t = sw.first; t < sw.end; t += sw.step)
    {
        const int y0 = (t / ntx) * CT_H, x0 = (t % ntx) * CT_W;
        __syncthreads();  // tile t is in LDS
        const int nxt = t + sw.step;
        fetch(nxt < sw.end ? nxt : t);  // the last tile reloads itself: no branch around the loads
        // The MFMA phases run at a raised wave priority, so a SIMD issues the MFMAs of one wave
        // ahead of the other wave's epilogue, staging and address work, which then fills the
        // cycles the matrix pipe is busy (measured: cnn5 backward -24 us with conv3x3_bwd12).
        __builtin_amdgcn_s_setprio(1);

        if constexpr (!L5)
        {
            // ---- weight gradient: taps 5 TG .. over rows 4 rg + 2 pr + h, 8-px K chunks c ----
            {
                constexpr int NJ = TG == 0 ? 5 : 4;
#pragma unroll 1
                for (int pr = 0; pr < 2; ++pr)
                {
                    const int rr = 4 * rg + 2 * pr + h;
                    lds_u16 *dp = (lds_u16 *)dys + ((rr + 1) * LW + jq + 1) * PS + cb;
                    lds_u16 *xp = (lds_u16 *)xs + (rr * LW + jq) * XS + cb;
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                    {
                        const uint2 b0 = ld_tr(dp + 8 * c * PS), b1 = ld_tr(dp + (8 * c + 4) * PS);
                        const u32x4 B = {b0.x, b0.y, b1.x, b1.y};
                        if constexpr (TG == 0) db = add_bf8(B, db);
#pragma unroll
                        for (int j = 0; j < NJ; ++j)
                        {
                            const int tap = 5 * TG + j, dy = tap / 3, sh = tap - 3 * dy;
                            const int o = (dy * LW + 8 * c + sh) * XS;
                            const uint2 a0 = ld_tr(xp + o), a1 = ld_tr(xp + o + 4 * XS);
                            const u32x4 A = {a0.x, a0.y, a1.x, a1.y};
                            wacc[j] = mfma(A, B, wacc[j]);
                        }
                    }
                }
            }
        }
        else
        {
            // ---- weight gradient on 16x16x32: tile rows 4 rg + r, K = the row's 32 px ----
            {
                constexpr int NJ = TG == 0 ? 5 : 4;
#pragma unroll 1
                for (int r = 0; r < 4; ++r)
                {
                    const int row = 4 * rg + r;
                    lds_u16 *dp = (lds_u16 *)dys + ((row + 1) * LW + 8 * kg + jq + 1) * PS + 4 * qq;
                    lds_u16 *xp = (lds_u16 *)xs + (row * LW + 8 * kg + jq) * XS + 4 * qq;
                    const uint2 b0 = ld_tr(dp), b1 = ld_tr(dp + 4 * PS);
                    const u32x4 B = {b0.x, b0.y, b1.x, b1.y};
                    if constexpr (TG == 0) db = add_bf8(B, db);
#pragma unroll
                    for (int j = 0; j < NJ; ++j)
                    {
                        const int tap = 5 * TG + j, dy = tap / 3, sh = tap - 3 * dy;
#pragma unroll
                        for (int m = 0; m < 2; ++m)
                        {
                            const int o = (dy * LW + sh) * XS + 16 * m;
                            const uint2 a0 = ld_tr(xp + o), a1 = ld_tr(xp + o + 4 * XS);
                            wacc16[j][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                __builtin_bit_cast(bf16x8, u32x4{a0.x, a0.y, a1.x, a1.y}),
                                __builtin_bit_cast(bf16x8, B), wacc16[j][m], 0, 0, 0);
                        }
                    }
                }
            }
        }

        // ---- data gradient: output rows 2wv, 2wv+1 (the conv3x3 loop on the dY tile) ----
        f32x16 acc[2];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int k = 0; k < 16; ++k) acc[r][k] = 0.f;
#pragma unroll
        for (int dx = 0; dx < 3; ++dx)
#pragma unroll
            for (int kc = 0; kc < KC; ++kc)
            {
                u32x4 wk[3], bx[4];
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) wk[dy] = wl[((dy * 3 + dx) * KC + kc) * 64 + lane];
#pragma unroll
                for (int lr = 0; lr < 4; ++lr)
                    bx[lr] = *reinterpret_cast<const u32x4 *>(&dys[((2 * wv + lr) * LW + px + dx) * PS + kc * 16 + 8 * h]);
#pragma unroll
                for (int lr = 0; lr < 4; ++lr)
#pragma unroll
                    for (int dy = 0; dy < 3; ++dy)
                    {
                        const int orow = lr - dy;
                        if (orow >= 0 && orow < 2) acc[orow] = mfma(wk[dy], bx[lr], acc[orow]);
                    }
            }
        __builtin_amdgcn_s_setprio(0);
        // epilogue: keep the gradient where X's bf16 value is > 0, 16-B stores after permlane32 swaps
#pragma unroll
        for (int orow = 0; orow < 2; ++orow)
        {
            const int gy = y0 + 2 * wv + orow, gxx = x0 + px;
            uint32_t d[8];  // d[2j + e] = channels 8j + 4h + 2e, +1
#pragma unroll
            for (int q = 0; q < 8; ++q) d[q] = pack2(acc[orow][2 * q], acc[orow][2 * q + 1]);
            // ReLU derivative from the staged X tile: keep a channel where X's bf16 value is nonzero
            // (X >= 0; +-0 both count as zero)
            const uint16_t *xp = &xs[((2 * wv + orow + 1) * LW + px + 1) * XS + 4 * h];
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const u32x2 xw = *reinterpret_cast<const u32x2 *>(xp + 8 * j);
#pragma unroll
                for (int e = 0; e < 2; ++e) d[2 * j + e] &= nz_mask(xw[e]);
            }
            u32x4 v[2];
#pragma unroll
            for (int g = 0; g < 2; ++g)
#pragma unroll
                for (int e = 0; e < 2; ++e)
                {
                    const auto s = __builtin_amdgcn_permlane32_swap(d[4 * g + e], d[4 * g + 2 + e], false, false);
                    v[g][e] = (uint32_t)s[0];
                    v[g][2 + e] = (uint32_t)s[1];
                }
            const bool in = gy < H && gxx < W;
            const uint32_t o = in ? (uint32_t)(gy * W + gxx) * 64u + 16u * h : kOob;
            bst16(v[0], rdx, o);
            bst16(v[1], rdx, in ? o + 32u : kOob);
        }
        __syncthreads();  // every wave is done reading tile t
        if (nxt < sw.end) commit();
    }

    // partial: [t][ci][co] then [co] (L5: compact, [t][32 ci][6 co] then [6]); tap t's four row-group
    // waves are added in LDS
    float *out = F.part + (size_t)li * (L5 ? WG_PART_L5 : WG_PART);
    constexpr int CO = L5 ? 6 : 32;
    float *red = reinterpret_cast<float *>(dys);  // [4 row groups][1024]
#pragma unroll
    for (int t = 0; t < 9; ++t)
    {
        const int g = t < 5 ? 0 : 1, j = t - 5 * g;
        __syncthreads();
        if (TG == g)
        {
            if constexpr (!L5)
#pragma unroll
                for (int k = 0; k < 16; ++k)
                {
                    const int ci = 8 * (k >> 2) + 4 * h + (k & 3);
                    red[rg * 1024 + ci * 32 + ch] = wacc[j][k];
                }
            else
                // 16x16 accumulators: column co = lane % 16, rows ci = 16m + 4kg + i
#pragma unroll
                for (int m = 0; m < 2; ++m)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        red[rg * 1024 + (16 * m + 4 * kg + i) * 32 + (lane & 15)] = wacc16[j][m][i];
        }
        __syncthreads();
        for (int e = tid; e < 32 * CO; e += NT)
        {
            const int r = L5 ? (e / CO) * 32 + e % CO : e;  // [ci][32] slot of entry (ci, co)
            out[t * 32 * CO + e] = (red[r] + red[1024 + r]) + (red[2048 + r] + red[3072 + r]);
        }
    }
    if constexpr (!L5)
    {
        db += __shfl_xor(db, 32);
        if (TG == 0 && h == 0) dbs[rg][ch] = db;
    }
    else
    {
        db += __shfl_xor(db, 16);
        db += __shfl_xor(db, 32);
        if (TG == 0 && lane < 32) dbs[rg][lane] = lane < 16 ? db : 0.f;
    }
    __syncthreads();
    if (tid < CO) out[9 * 32 * CO + tid] = (dbs[0][tid] + dbs[1][tid]) + (dbs[2][tid] + dbs[3][tid]);
    };
    if (tg == 0) run(IntC<0>{});
    else run(IntC<1>{});
}

// ---------------------------------------------------------------------------------------------
// Backward of layer 2 fused with layer 1's weight gradient
// ---------------------------------------------------------------------------------------------
struct Bwd12Batch {
    int n;
    int blk_begin[XDN_MAX_FRAMES + 1];
    Bwd12Frame f[XDN_MAX_FRAMES];
};
static_assert(sizeof(Bwd12Batch) <= 4096, "kernel argument too large");

// Layer 2's backward on one staging of a 16x32-px tile of its dY and its input X (both with the
// 1-px halo, as in conv3x3_bwd) and of layer 1's five input images (1-px halo, bf16, 15 channels
// and a zero one):
//   layer 2 weight gradient: wave w takes tap w over the whole tile (32 MFMAs; it holds the block's
//     only accumulator of that tap) and tap 8 over tile rows 2w, 2w+1 (4 MFMAs; the eight waves'
//     sums of tap 8 are added when the block writes its partial);
//   layer 2 data gradient: wave w computes tile rows 2w, 2w+1 as conv3x3_bwd does and keeps it
//     where X is nonzero: that is layer 1's pre-activation gradient dY1, which goes into LDS over
//     the dY tile (read by then) instead of to HBM;
//   layer 1 weight gradient (v_mfma_f32_16x16x32_bf16: M = 16 input channels, N = 16 output
//     channels, two tiles, K = one tile row's 32 px): wave w takes tap w over the tile and tap 8
//     over rows 2w, 2w+1, A read from the image tile and B from dY1, both with transposing reads.
// The tap-major split keeps 48 accumulator registers per wave for the two weight gradients (a
// row-group split would need 80 + 72), so the three products and the prefetch of the three staged
// tiles fit two waves per SIMD. dY1 never reaches HBM, and layer 1's weight gradient needs no
// launch of its own. Four barriers per tile: the tile is staged; every wave is done with the dY
// tile; dY1 is in LDS; every wave is done with all tiles.
__global__ __launch_bounds__(512, 1) void conv3x3_bwd12(Bwd12Batch b)
{
    constexpr int PS = 40, XS = 40;   // dY / X pixel strides in bf16 (as conv3x3_bwd)
    constexpr int IS = 16;            // image-tile pixel stride: 15 channels + 1 zero
    constexpr int D1S = 40;           // dY1 pixel stride ([16 rows][32 px], over the dY tile)
    constexpr int LW = CT_W + 2, LH = CT_H + 2;
    constexpr int NCH = LH * LW * 4;  // 16-B chunks of a staged dY or X tile
    constexpr int NT = 512;
    constexpr int NV = (NCH + NT - 1) / NT;
    constexpr int NI = (LH * LW + NT - 1) / NT;  // image-tile pixels per thread
    __shared__ __attribute__((aligned(16))) uint16_t dys[LH * LW * PS];
    __shared__ __attribute__((aligned(16))) uint16_t xs[LH * LW * XS];
    __shared__ __attribute__((aligned(16))) uint16_t ims[LH * LW * IS];
    __shared__ u32x4 wl[9 * 2 * 64];
    __shared__ float dbs[8][32], dbs1[8][32];
    static_assert(LH * LW * PS >= CT_H * CT_W * D1S, "dY1 tile space");
    static_assert(LH * LW * PS * 2 >= 8 * 1024 * 4, "partial reduction space");
    static_assert(LH * LW * XS * 2 >= 8 * 512 * 4, "layer-1 partial reduction space");
    int f = 0;
    while (f + 1 < b.n && (int)blockIdx.x >= b.blk_begin[f + 1]) ++f;
    const Bwd12Frame &F = b.f[f];
    const int li = blockIdx.x - b.blk_begin[f], nblk = b.blk_begin[f + 1] - b.blk_begin[f];
    const int H = F.H, W = F.W;
    const int ntx = (W + CT_W - 1) / CT_W, nty = (H + CT_H - 1) / CT_H;
    const int tid = threadIdx.x, lane = tid & 63, px = lane & 31, h = lane >> 5, ch = lane & 31;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: this wave's tap
    const int grp = (lane >> 4) & 1, s16 = lane & 15, jq = s16 >> 2, qq = s16 & 3, kg = lane >> 4;
    const int cb = 16 * grp + 4 * qq;  // channel offset addressed by this lane in transposing reads
    const int tdy = wv / 3, tsh = wv - 3 * tdy;

    for (int e = tid; e < 9 * 2 * 64; e += NT) wl[e] = static_cast<const u32x4 *>(F.wfrag)[e];

    // per-lane LDS bases of the transposing reads; a row, chunk or tap-8 shift adds an immediate
    lds_u16 *const dys3 = (lds_u16 *)dys, *const xs3 = (lds_u16 *)xs, *const ims3 = (lds_u16 *)ims;
    lds_u16 *const dB = dys3 + ((h + 1) * LW + jq + 1) * PS + cb;                 // layer 2 dY, rows 2p + h
    lds_u16 *const xA = xs3 + ((h + tdy) * LW + jq + tsh) * XS + cb;              // layer 2 X, tap wv
    lds_u16 *const dB8 = dys3 + ((2 * wv + h + 1) * LW + jq + 1) * PS + cb;       // rows 2wv + h
    lds_u16 *const xA8 = xs3 + ((2 * wv + h + 2) * LW + jq + 2) * XS + cb;        // tap 8 there
    lds_u16 *const d1B = dys3 + (8 * kg + jq) * D1S + 4 * qq;                    // dY1, row 0
    lds_u16 *const iA = ims3 + ((tdy * LW) + 8 * kg + jq + tsh) * IS + 4 * qq;    // images, tap wv
    lds_u16 *const iA8 = ims3 + ((2 * wv + 2) * LW + 8 * kg + jq + 2) * IS + 4 * qq;  // tap 8, row 2wv

    typedef __attribute__((ext_vector_type(4))) float f32x4;
    f32x16 wa2, wa28;    // layer 2: tap wv over the tile; tap 8 over rows 2wv, 2wv + 1
    f32x4 wa1[2], wa18[2];  // layer 1 (two output-channel tiles): the same split
#pragma unroll
    for (int k = 0; k < 16; ++k) wa2[k] = wa28[k] = 0.f;
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
        for (int k = 0; k < 4; ++k) wa1[n][k] = wa18[n][k] = 0.f;
    float db2 = 0.f, db1[2] = {0.f, 0.f};

    // staging: branch-free buffer loads (zero outside the image)
    const size_t npx = (size_t)H * W;
    const rsrc_t rdy = make_rsrc(F.dy, npx * 64), rxx = make_rsrc(F.x, npx * 64);
    rsrc_t rim[5];
#pragma unroll
    for (int im = 0; im < 5; ++im) rim[im] = make_rsrc(F.img[im], npx * 12);
    u32x4 rd[NV], rx[NV];
    u32x3 fi[NI][5];
    auto fetch = [&](int t) {
        const int y0 = (t / ntx) * CT_H, x0 = (t % ntx) * CT_W;
#pragma unroll
        for (int u = 0; u < NV; ++u)
        {
            const int i = tid + NT * u;
            const int pix = i >> 2, q = i & 3;
            const int ly = pix / LW, lx = pix - ly * LW;
            const int gy = y0 + ly - 1, gxx = x0 + lx - 1;
            const bool ok = i < NCH && gy >= 0 && gy < H && gxx >= 0 && gxx < W;
            const uint32_t o = ok ? (uint32_t)((gy * W + gxx) * 32 + q * 8) * 2u : kOob;
            rd[u] = bld16(rdy, o);
            rx[u] = bld16(rxx, o);
        }
#pragma unroll
        for (int u = 0; u < NI; ++u)
        {
            const int pix = tid + NT * u;
            const int ly = pix / LW, lx = pix - ly * LW;
            const int gy = y0 + ly - 1, gxx = x0 + lx - 1;
            const bool ok = pix < LH * LW && gy >= 0 && gy < H && gxx >= 0 && gxx < W;
            const uint32_t o = ok ? (uint32_t)(gy * W + gxx) * 12u : kOob;
#pragma unroll
            for (int im = 0; im < 5; ++im) fi[u][im] = bld12(rim[im], o);
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int u = 0; u < NV; ++u)
        {
            const int i = tid + NT * u;
            if (i < NCH)
            {
                const int pix = i >> 2, q = i & 3;
                *reinterpret_cast<u32x4 *>(&dys[pix * PS + q * 8]) = rd[u];
                *reinterpret_cast<u32x4 *>(&xs[pix * XS + q * 8]) = rx[u];
            }
        }
#pragma unroll
        for (int u = 0; u < NI; ++u)
        {
            const int pix = tid + NT * u;
            if (pix < LH * LW)
            {
                float v[15];
#pragma unroll
                for (int im = 0; im < 5; ++im)
                {
                    v[3 * im] = __uint_as_float(fi[u][im].x);
                    v[3 * im + 1] = __uint_as_float(fi[u][im].y);
                    v[3 * im + 2] = __uint_as_float(fi[u][im].z);
                }
                const u32x4 a = {pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]), pack2(v[6], v[7])};
                const u32x4 c = {pack2(v[8], v[9]), pack2(v[10], v[11]), pack2(v[12], v[13]), pack2(v[14], 0.f)};
                *reinterpret_cast<u32x4 *>(&ims[pix * IS]) = a;
                *reinterpret_cast<u32x4 *>(&ims[pix * IS + 8]) = c;
            }
        }
    };

    const StripWalk sw = xcd_strips(li, nblk, ntx * nty);
    if (sw.first < sw.end)
    {
        fetch(sw.first);
        commit();
    }
    for (int t = sw.first; t < sw.end; t += sw.step)
    {
        __syncthreads();  // tile t is in LDS
        const int nxt = t + sw.step;
        fetch(nxt < sw.end ? nxt : t);  // the last tile reloads itself: no branch around the loads
        __builtin_amdgcn_s_setprio(1);  // MFMA phases at a raised wave priority (as in conv3x3_bwd)

        // ---- layer 2 weight gradient: tap wv over rows 2p + h, 8-px K chunks c; tap 8 over 2wv + h ----
#pragma unroll 2
        for (int p = 0; p < 8; ++p)
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
                const int od = (2 * p * LW + 8 * c) * PS, ox = (2 * p * LW + 8 * c) * XS;
                const uint2 b0 = ld_tr(dB + od), b1 = ld_tr(dB + od + 4 * PS);
                const uint2 a0 = ld_tr(xA + ox), a1 = ld_tr(xA + ox + 4 * XS);
                wa2 = mfma(u32x4{a0.x, a0.y, a1.x, a1.y}, u32x4{b0.x, b0.y, b1.x, b1.y}, wa2);
            }
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
            const uint2 b0 = ld_tr(dB8 + 8 * c * PS), b1 = ld_tr(dB8 + (8 * c + 4) * PS);
            const u32x4 B = {b0.x, b0.y, b1.x, b1.y};
            db2 = add_bf8(B, db2);
            const uint2 a0 = ld_tr(xA8 + 8 * c * XS), a1 = ld_tr(xA8 + (8 * c + 4) * XS);
            wa28 = mfma(u32x4{a0.x, a0.y, a1.x, a1.y}, B, wa28);
        }

        // ---- layer 2 data gradient: rows 2wv, 2wv + 1, kept where X is nonzero -> dY1 ----
        f32x16 acc[2];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int k = 0; k < 16; ++k) acc[r][k] = 0.f;
#pragma unroll
        for (int dx = 0; dx < 3; ++dx)
#pragma unroll
            for (int kc = 0; kc < 2; ++kc)
            {
                u32x4 wk[3], bx[4];
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) wk[dy] = wl[((dy * 3 + dx) * 2 + kc) * 64 + lane];
#pragma unroll
                for (int lr = 0; lr < 4; ++lr)
                    bx[lr] = *reinterpret_cast<const u32x4 *>(&dys[((2 * wv + lr) * LW + px + dx) * PS + kc * 16 + 8 * h]);
#pragma unroll
                for (int lr = 0; lr < 4; ++lr)
#pragma unroll
                    for (int dy = 0; dy < 3; ++dy)
                    {
                        const int orow = lr - dy;
                        if (orow >= 0 && orow < 2) acc[orow] = mfma(wk[dy], bx[lr], acc[orow]);
                    }
            }
        __builtin_amdgcn_s_setprio(0);
        u32x4 g1[2][2];  // dY1 of rows 2wv + orow: lane (px, h) holds channels 8h .. 8h + 7 and 16 + 8h ..
#pragma unroll
        for (int orow = 0; orow < 2; ++orow)
        {
            uint32_t d[8];  // d[2j + e] = channels 8j + 4h + 2e, +1
#pragma unroll
            for (int q = 0; q < 8; ++q) d[q] = pack2(acc[orow][2 * q], acc[orow][2 * q + 1]);
            const uint16_t *xp = &xs[((2 * wv + orow + 1) * LW + px + 1) * XS + 4 * h];
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const u32x2 xw = *reinterpret_cast<const u32x2 *>(xp + 8 * j);
#pragma unroll
                for (int e = 0; e < 2; ++e) d[2 * j + e] &= nz_mask(xw[e]);
            }
#pragma unroll
            for (int g = 0; g < 2; ++g)
#pragma unroll
                for (int e = 0; e < 2; ++e)
                {
                    const auto sp = __builtin_amdgcn_permlane32_swap(d[4 * g + e], d[4 * g + 2 + e], false, false);
                    g1[orow][g][e] = (uint32_t)sp[0];
                    g1[orow][g][2 + e] = (uint32_t)sp[1];
                }
        }
        __syncthreads();  // every wave is done with the dY tile
#pragma unroll
        for (int orow = 0; orow < 2; ++orow)
        {
            uint16_t *dp = &dys[((2 * wv + orow) * CT_W + px) * D1S + 8 * h];
            *reinterpret_cast<u32x4 *>(dp) = g1[orow][0];
            *reinterpret_cast<u32x4 *>(dp + 16) = g1[orow][1];
        }
        __syncthreads();  // dY1 is in LDS

        // ---- layer 1 weight gradient: tap wv over the 16 rows; tap 8 over rows 2wv, 2wv + 1 ----
        __builtin_amdgcn_s_setprio(1);
#pragma unroll 4
        for (int r = 0; r < CT_H; ++r)
        {
            const uint2 a0 = ld_tr(iA + r * LW * IS), a1 = ld_tr(iA + r * LW * IS + 4 * IS);
            const bf16x8 A = __builtin_bit_cast(bf16x8, u32x4{a0.x, a0.y, a1.x, a1.y});
#pragma unroll
            for (int n = 0; n < 2; ++n)
            {
                const int od = r * CT_W * D1S + 16 * n;
                const uint2 b0 = ld_tr(d1B + od), b1 = ld_tr(d1B + od + 4 * D1S);
                wa1[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    A, __builtin_bit_cast(bf16x8, u32x4{b0.x, b0.y, b1.x, b1.y}), wa1[n], 0, 0, 0);
            }
        }
#pragma unroll
        for (int r = 0; r < 2; ++r)
        {
            const uint2 a0 = ld_tr(iA8 + r * LW * IS), a1 = ld_tr(iA8 + r * LW * IS + 4 * IS);
            const bf16x8 A = __builtin_bit_cast(bf16x8, u32x4{a0.x, a0.y, a1.x, a1.y});
#pragma unroll
            for (int n = 0; n < 2; ++n)
            {
                lds_u16 *const bp = d1B + (2 * wv + r) * CT_W * D1S + 16 * n;
                const uint2 b0 = ld_tr(bp), b1 = ld_tr(bp + 4 * D1S);
                const u32x4 B = {b0.x, b0.y, b1.x, b1.y};
                db1[n] = add_bf8(B, db1[n]);
                wa18[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, __builtin_bit_cast(bf16x8, B), wa18[n], 0, 0, 0);
            }
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();  // every wave is done reading tile t
        if (nxt < sw.end) commit();
    }

    // partials: layer 2 [t][ci][co] then [co]; layer 1 compact [t][15 ci][32 co] then [32 co]. Taps
    // 0..7 come whole from wave t; tap 8 is the sum of the eight waves' row pairs, added in LDS.
    float *out = F.part + (size_t)li * WG_PART;
    float *out1 = F.part1 + (size_t)li * WG_PART_L1;
    float *red = reinterpret_cast<float *>(dys);   // [8 waves][1024]
    float *red1 = reinterpret_cast<float *>(xs);   // [8 waves][512]
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k)
    {
        const int ci = 8 * (k >> 2) + 4 * h + (k & 3);
        out[wv * 1024 + ci * 32 + ch] = wa2[k];
        red[wv * 1024 + ci * 32 + ch] = wa28[k];
    }
    // 16x16 accumulators: column co = 16n + lane % 16, rows ci = 4kg + i (input channel 15 is padding)
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            const int ci = 4 * kg + i, co = 16 * n + (lane & 15);
            if (ci < 15) out1[wv * 15 * 32 + ci * 32 + co] = wa1[n][i];
            red1[wv * 512 + ci * 32 + co] = wa18[n][i];
        }
    db2 += __shfl_xor(db2, 32);
    if (h == 0) dbs[wv][ch] = db2;
#pragma unroll
    for (int n = 0; n < 2; ++n)
    {
        float v = db1[n];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (lane < 16) dbs1[wv][16 * n + lane] = v;
    }
    __syncthreads();
    for (int e = tid; e < 1024; e += NT)
        out[8 * 1024 + e] = ((red[e] + red[1024 + e]) + (red[2048 + e] + red[3072 + e])) +
                            ((red[4096 + e] + red[5120 + e]) + (red[6144 + e] + red[7168 + e]));
    if (tid < 15 * 32)
        out1[8 * 15 * 32 + tid] = ((red1[tid] + red1[512 + tid]) + (red1[1024 + tid] + red1[1536 + tid])) +
                                  ((red1[2048 + tid] + red1[2560 + tid]) + (red1[3072 + tid] + red1[3584 + tid]));
    if (tid < 32)
        out[9 * 1024 + tid] = ((dbs[0][tid] + dbs[1][tid]) + (dbs[2][tid] + dbs[3][tid])) +
                              ((dbs[4][tid] + dbs[5][tid]) + (dbs[6][tid] + dbs[7][tid]));
    else if (tid < 64)
    {
        const int c = tid - 32;
        out1[9 * 15 * 32 + c] = ((dbs1[0][c] + dbs1[1][c]) + (dbs1[2][c] + dbs1[3][c])) +
                                ((dbs1[4][c] + dbs1[5][c]) + (dbs1[6][c] + dbs1[7][c]));
    }
}

// Two-stage deterministic reduction of the per-block partials, batched over jobs (a job is one
// layer of one frame). Stage 1: stage[j][split][e] = sum of the partials split*P/S .. of job j;
// stage 2 writes the flat gradient of the layer ([9][CI][CO] then [CO], unpadded).
struct RedBatch {
    int n, splits;
    RedJob j[WG_MAX_JOBS];
};
static_assert(sizeof(RedBatch) <= 4096, "kernel argument too large");

__global__ __launch_bounds__(256) void wgrad_reduce1(RedBatch b)
{
    const RedJob &J = b.j[blockIdx.z];
    const int sp = blockIdx.y;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= J.pstride) return;
    const int nb = J.nblk;
    const int lo = (int)((long)nb * sp / b.splits), hi = (int)((long)nb * (sp + 1) / b.splits);
    // the loads of eight partials are issued together; the sum keeps the order lo, lo + 1, ...
    float s = 0.f;
    const float *p = J.part + e;
    for (int i0 = lo; i0 < hi; i0 += 8)
    {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = i0 + k < hi ? p[(size_t)(i0 + k) * J.pstride] : 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (i0 + k < hi) s += v[k];
    }
    J.stage[sp * J.pstride + e] = s;
}

__global__ __launch_bounds__(256) void wgrad_reduce2(RedBatch b)
{
    const RedJob &J = b.j[blockIdx.y];
    const int e = blockIdx.x * 256 + threadIdx.x;
    const int cc = J.ci_cnt < 0 ? J.ci : J.ci_cnt;
    const int nw = J.taps * cc * J.co;
    if (e >= nw + (J.bias ? J.co : 0)) return;
    int src, dst;
    if (e < nw)
    {
        const int t = e / (cc * J.co), r = e - t * cc * J.co, cl = r / J.co, co = r - cl * J.co;
        src = J.compact ? e : t * 1024 + cl * 32 + co;
        dst = J.transpose ? co * J.ci + J.ci_off + cl : (t * J.ci + J.ci_off + cl) * J.co + co;
    }
    else
    {
        src = J.compact ? e : 9 * 1024 + (e - nw);
        dst = J.taps * J.ci * J.co + (e - nw);
    }
    float v[WG_SPLITS];
#pragma unroll
    for (int sp = 0; sp < WG_SPLITS; ++sp) v[sp] = sp < b.splits ? J.stage[sp * J.pstride + src] : 0.f;
    float s = 0.f;
#pragma unroll
    for (int sp = 0; sp < WG_SPLITS; ++sp)
        if (sp < b.splits) s += v[sp];
    J.grad[dst] = s;
}

// ---------------------------------------------------------------------------------------------
// FP32 master [9][ci][co] + [co]  ->  bf16 B fragments:
//   fwd  frag[t][kc][lane] = W[t][kc*16 + 8h + j][lane%32]        (KC = ceil(ci/16))
//   dgrad frag[t][kc][lane] = W[8-t][lane%32][kc*16 + 8h + j]     (KC = ceil(co/16))
// zero-padded outside ci x co; bias32[c] = b[c] (0 padded).
// ---------------------------------------------------------------------------------------------
struct PackBatch {
    int n;
    PackJob j[XDN_MAX_FRAMES * 5];
};

__global__ __launch_bounds__(64) void pack_weights(PackBatch b)
{
    const PackJob &J = b.j[blockIdx.y];
    const int lane = threadIdx.x, h = lane >> 5, n = lane & 31;
    const int kcf = (J.ci + 15) / 16, kcd = (J.co + 15) / 16;
    const int item = blockIdx.x;  // t * 2 + kc for fwd (first 18), then 18 + t * 2 + kc for dgrad
    const float *Wm = J.w;
    if (item < 18)
    {
        const int t = item >> 1, kc = item & 1;
        if (kc >= kcf) return;
        uint32_t v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            float e[2];
#pragma unroll
            for (int s = 0; s < 2; ++s)
            {
                const int ci = kc * 16 + 8 * h + 2 * q + s;
                e[s] = (ci < J.ci && n < J.co) ? Wm[(t * J.ci + ci) * J.co + n] : 0.f;
            }
            v[q] = pack2(e[0], e[1]);
        }
        reinterpret_cast<u32x4 *>(J.frag_fwd)[(t * kcf + kc) * 64 + lane] = u32x4{v[0], v[1], v[2], v[3]};
        if (t == 0 && kc == 0 && h == 0) J.bias32[n] = n < J.co ? J.b[n] : 0.f;
    }
    else if (J.frag_dgrad)
    {
        const int it = item - 18, t = it >> 1, kc = it & 1;
        if (kc >= kcd) return;
        uint32_t v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            float e[2];
#pragma unroll
            for (int s = 0; s < 2; ++s)
            {
                const int co = kc * 16 + 8 * h + 2 * q + s;
                e[s] = (co < J.co && n < J.ci) ? Wm[((8 - t) * J.ci + n) * J.co + co] : 0.f;
            }
            v[q] = pack2(e[0], e[1]);
        }
        reinterpret_cast<u32x4 *>(J.frag_dgrad)[(t * kcd + kc) * 64 + lane] = u32x4{v[0], v[1], v[2], v[3]};
    }
}

// ---------------------------------------------------------------------------------------------
static int div_up(int a, int b) { return (a + b - 1) / b; }

hipError_t launch_conv_l1(int n, const ConvFrame *fr, hipStream_t st)
{
    ConvBatch b{};
    int tx[XDN_MAX_FRAMES], ty[XDN_MAX_FRAMES];
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        tx[i] = div_up(fr[i].W, CT_W);
        ty[i] = div_up(fr[i].H, CT_H);
    }
    for (int i = 0; i < n; ++i)  // buffer offsets are 32-bit with kOob = 2^31 past every buffer
        if ((size_t)fr[i].H * fr[i].W * 64 >= ((size_t)1 << 31)) return hipErrorInvalidValue;
    b.map = make_tilemap(n, tx, ty);
    if (b.map.total == 0) return hipSuccess;
    // persistent grid: two resident blocks per CU, a multiple of the 8 XCDs
    constexpr int kGrid = 256 * 2;
    const int grid = b.map.total < kGrid ? b.map.total : kGrid;
    hipLaunchKernelGGL(conv3x3_l1, dim3(grid), dim3(256), 0, st, b);
    return hipGetLastError();
}

int bwd_blocks(int H, int W)
{
    const int tiles = div_up(W, CT_W) * div_up(H, CT_H);
    // ~16 tiles per block: one block per CU at 1920x1080 (4080 tiles); a smaller image still gets a
    // block per CU where it has the tiles (1280x720: 256 blocks of about 7 tiles instead of 113 of 16)
    int nb = div_up(tiles, 16);
    if (nb < 256) nb = 256;
    if (nb > tiles) nb = tiles;
    return nb < 1 ? 1 : nb;
}

hipError_t launch_bwd(int n, const BwdFrame *fr, bool last, hipStream_t st)
{
    for (int i = 0; i < n; ++i)  // 32-bit buffer offsets, kOob = 2^31 past every buffer
        if ((size_t)fr[i].H * fr[i].W * 64 >= ((size_t)1 << 31)) return hipErrorInvalidValue;
    BwdBatch b{};
    b.n = n;
    int acc = 0;
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        b.blk_begin[i] = acc;
        acc += fr[i].nblk;
    }
    b.blk_begin[n] = acc;
    if (acc == 0) return hipSuccess;
    if (last)
        hipLaunchKernelGGL(conv3x3_bwd<true>, dim3(acc), dim3(512), 0, st, b);
    else
        hipLaunchKernelGGL(conv3x3_bwd<false>, dim3(acc), dim3(512), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_bwd12(int n, const Bwd12Frame *fr, hipStream_t st)
{
    for (int i = 0; i < n; ++i)  // 32-bit buffer offsets, kOob = 2^31 past every buffer
        if ((size_t)fr[i].H * fr[i].W * 64 >= ((size_t)1 << 31)) return hipErrorInvalidValue;
    Bwd12Batch b{};
    b.n = n;
    int acc = 0;
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        b.blk_begin[i] = acc;
        acc += fr[i].nblk;
    }
    b.blk_begin[n] = acc;
    if (acc == 0) return hipSuccess;
    hipLaunchKernelGGL(conv3x3_bwd12, dim3(acc), dim3(512), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_fwd_stream(int n, const FwdStreamFrame *fr, hipStream_t st)
{
    FwdStreamBatch b{};
    b.n = n;
    int acc = 0;
    for (int i = 0; i < n; ++i)
    {
        const int H = fr[i].H, W = fr[i].W;
        if ((size_t)H * W * 64 >= ((size_t)1 << 31)) return hipErrorInvalidValue;  // 32-bit buffer offsets
        b.f[i] = fr[i];
        b.nstrip[i] = div_up(W, FS_VALID);
        // about three blocks per CU (all resident at once: 43.5 KB of LDS, <= 168 VGPRs) for one
        // 1920x1080 frame; a segment re-runs FS_SKEW rows of its neighbour, so segments stay long.
        // Measured: three resident blocks per CU instead of two, layers 2-5 -10 us.
        int nseg = (768 + b.nstrip[i] / 2) / b.nstrip[i];
        nseg = nseg < 1 ? 1 : (nseg > div_up(H, 32) ? div_up(H, 32) : nseg);
        b.seg_rows[i] = div_up(H, nseg);
        nseg = div_up(H, b.seg_rows[i]);
        b.blk_begin[i] = acc;
        acc += b.nstrip[i] * nseg;
    }
    b.blk_begin[n] = acc;
    if (acc == 0) return hipSuccess;
    hipLaunchKernelGGL(cnn5_fwd_stream, dim3(acc), dim3(256), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_wgrad_reduce(int n, const RedJob *jobs, hipStream_t st)
{
    for (int j0 = 0; j0 < n; j0 += WG_MAX_JOBS)
    {
        const int m = n - j0 < WG_MAX_JOBS ? n - j0 : WG_MAX_JOBS;
        RedBatch b{};
        b.n = m;
        b.splits = WG_SPLITS;
        int maxw = 0, maxp = 0;
        for (int i = 0; i < m; ++i)
        {
            const RedJob &J = jobs[j0 + i];
            b.j[i] = J;
            const int w = J.taps * (J.ci_cnt < 0 ? J.ci : J.ci_cnt) * J.co + J.co;
            maxw = w > maxw ? w : maxw;
            maxp = J.pstride > maxp ? J.pstride : maxp;
        }
        hipLaunchKernelGGL(wgrad_reduce1, dim3(div_up(maxp, 256), WG_SPLITS, m), dim3(256), 0, st, b);
        hipLaunchKernelGGL(wgrad_reduce2, dim3(div_up(maxw, 256), m), dim3(256), 0, st, b);
        if (hipError_t e = hipGetLastError()) return e;
    }
    return hipSuccess;
}

hipError_t launch_pack(int n, const PackJob *jobs, hipStream_t st)
{
    PackBatch b{};
    b.n = n;
    for (int i = 0; i < n; ++i) b.j[i] = jobs[i];
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pack_weights, dim3(36, n), dim3(64), 0, st, b);
    return hipGetLastError();
}

}  // namespace xdn
