// Kernels of the reference's bandwidth U-Net (scripts/main.py:83-118) on gfx950.
//
//   uconv3x3<SRC, CA, CB, CO, EPI>  3x3 'same' convolution as an implicit GEMM on
//     v_mfma_f32_32x32x16_bf16 (M = 32 px of a row, N = 32 output channels per N-tile, K = 16
//     input channels per chunk) over concat[a (CA ch), b (CB ch)] staged in LDS with zero padding
//     to whole K chunks; 16x32-px tiles, every A fragment feeds up to three output rows. Forward
//     (bias + ReLU -> bf16) and data gradient (flipped, transposed fragments; ReLU-masked, or
//     plain for concat and pooled inputs). Channel counts are the stored ones (8, 16, 24, 32, 48).
//   uwgrad<SRC, CA, CB, CO, MT>  weight gradient over 8x32-px strips, operands from pixel-major
//     LDS tiles through ds_read_b64_tr_b16, one 32-channel input slice (MT) per instantiation,
//     software-pipelined strip loop.
//   maxpool2 / pool_bwd          2x2 max pool and its gradient (to the first maximum, plus the
//     skip-connection gradient, ReLU-masked).
//   convT1x1s2 / convT_bwd       transposed 1x1 stride-2 convolution and its gradients.
//   head_bwd                     gradients of the 1x1 8->6 output convolution (its forward runs in
//                                the epilogue of the last 3x3 layer, uconv3x3<..., HEAD>).
//   upack                        FP32 master weights -> bf16 fragments.
#include "xdn_common.h"
#include "xdn_unet.h"

namespace xdn {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;
typedef short v4s __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v4s lds_v4s;

__device__ __forceinline__ uint32_t u_f2bf(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ uint32_t u_pack2(float a, float b) { return u_f2bf(a) | (u_f2bf(b) << 16); }
__device__ __forceinline__ float u_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float u_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }
__device__ __forceinline__ f32x16 u_mfma(u32x4 a, u32x4 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}
// transposing LDS read, see ld_tr of k_cnn.hip
__device__ __forceinline__ uint2 u_ld_tr(const uint16_t *p)
{
    return __builtin_bit_cast(uint2, __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s *)p));
}
// bf16 word pair: keep each half where the mask half is > 0
__device__ __forceinline__ uint32_t u_keep(uint32_t mw, uint32_t vw)
{
    const uint32_t lo = mw & 0xFFFFu, hi = mw >> 16;
    const uint32_t klo = ((lo & 0x8000u) == 0 && lo != 0) ? 0xFFFFu : 0u;
    const uint32_t khi = ((hi & 0x8000u) == 0 && hi != 0) ? 0xFFFF0000u : 0u;
    return vw & (klo | khi);
}
__device__ __forceinline__ float u_bf(const uint16_t *p, int i) { return __uint_as_float((uint32_t)p[i] << 16); }

static int udiv_up(int a, int b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------------------------
// 3x3 convolution (forward / data gradient)
// ---------------------------------------------------------------------------------------------
struct UcBatch {
    TileMap map;
    UcFrame f[XDN_MAX_FRAMES];
};
static_assert(sizeof(UcBatch) <= 4096, "kernel argument too large");

constexpr int UT_W = 32, UT_H = 16;

template <int SRC, int CA, int CB, int CO, int EPI, bool HEAD = false>
__global__ __launch_bounds__(256, (CB == 0 && CA <= 16 && CO <= 24) ? 4 : 1) void uconv3x3(UcBatch b)
{
    static_assert(!HEAD || (CO == 8 && EPI == UC_EPI_RELU), "the head follows the 8-channel ReLU layer");
    constexpr int CI = SRC == UC_SRC_IMAGES ? 15 : CA + CB;
    constexpr int KC = (CI + 15) / 16, CIP = KC * 16, PS = CIP + 8;
    constexpr int NT = (CO + 31) / 32;
    constexpr int LW = UT_W + 2, LH = UT_H + 2;
    constexpr int OS = CO + 8;  // epilogue pixel stride (bf16)
    constexpr int IN_ELEMS = LH * LW * PS, OUT_ELEMS = UT_H * UT_W * OS;
    __shared__ __attribute__((aligned(16))) uint16_t lds[IN_ELEMS > OUT_ELEMS ? IN_ELEMS : OUT_ELEMS];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, px = lane & 31, h = lane >> 5;
    const TileLoc L = locate(b.map, xcd_tile(blockIdx.x, b.map.total));
    const UcFrame &F = b.f[L.f];
    const int H = F.H, W = F.W;
    const int x0 = L.tx * UT_W, y0 = L.ty * UT_H;
    // kernel fragments and bias first: their loads are in flight together with the tile's
    u32x4 wf[9][KC][NT];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                wf[t][kc][nt] = static_cast<const u32x4 *>(F.wfrag)[((t * KC + kc) * NT + nt) * 64 + lane];
    float bias[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) bias[nt] = EPI == UC_EPI_RELU ? F.bias[nt * 32 + px] : 0.f;
    if constexpr (SRC == UC_SRC_BF16)
    {
        constexpr int QA = CA / 8, QB = CB / 8, QP = CIP / 8;
        for (int i = tid; i < LH * LW * QP; i += 256)
        {
            const int pix = i / QP, q = i - pix * QP;
            const int ly = pix / LW, lx = pix - ly * LW;
            const int gy = y0 + ly - 1, gx = x0 + lx - 1;
            u32x4 v = {0u, 0u, 0u, 0u};
            if (gy >= 0 && gy < H && gx >= 0 && gx < W)
            {
                const size_t p = (size_t)gy * W + gx;
                if (q < QA)
                    v = *reinterpret_cast<const u32x4 *>(static_cast<const uint16_t *>(F.a) + p * CA + q * 8);
                else if (q < QA + QB)
                    v = *reinterpret_cast<const u32x4 *>(static_cast<const uint16_t *>(F.b) + p * CB + (q - QA) * 8);
            }
            *reinterpret_cast<u32x4 *>(&lds[pix * PS + q * 8]) = v;
        }
    }
    else
    {
        for (int pix = tid; pix < LH * LW; pix += 256)
        {
            const int ly = pix / LW, lx = pix - ly * LW;
            const int gy = y0 + ly - 1, gx = x0 + lx - 1;
            float v[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = 0.f;
            if (gy >= 0 && gy < H && gx >= 0 && gx < W)
            {
                const size_t p = (size_t)gy * W + gx;
#pragma unroll
                for (int im = 0; im < 5; ++im)
                {
                    const float *s = F.img[im] + 3 * p;
                    v[3 * im] = s[0]; v[3 * im + 1] = s[1]; v[3 * im + 2] = s[2];
                }
            }
            u32x4 a, c;
            a.x = u_pack2(v[0], v[1]); a.y = u_pack2(v[2], v[3]); a.z = u_pack2(v[4], v[5]); a.w = u_pack2(v[6], v[7]);
            c.x = u_pack2(v[8], v[9]); c.y = u_pack2(v[10], v[11]); c.z = u_pack2(v[12], v[13]);
            c.w = u_pack2(v[14], v[15]);
            *reinterpret_cast<u32x4 *>(&lds[pix * PS]) = a;
            *reinterpret_cast<u32x4 *>(&lds[pix * PS + 8]) = c;
        }
    }
    __syncthreads();

    f32x16 acc[4][NT];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int k = 0; k < 16; ++k) acc[r][nt][k] = 0.f;
#pragma unroll
    for (int ir = 0; ir < 6; ++ir)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx)
#pragma unroll
            for (int kc = 0; kc < KC; ++kc)
            {
                const u32x4 a = *reinterpret_cast<const u32x4 *>(&lds[((4 * wv + ir) * LW + px + dx) * PS + kc * 16 + 8 * h]);
#pragma unroll
                for (int orow = 0; orow < 4; ++orow)
                {
                    const int dy = ir - orow;
                    if (dy < 0 || dy > 2) continue;
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) acc[orow][nt] = u_mfma(a, wf[dy * 3 + dx][kc][nt], acc[orow][nt]);
                }
            }

    // epilogue: lane owns output channel nt*32 + px of 16 pixels x = 8(k/4) + 4h + k%4 per row;
    // through LDS to 16-byte stores of whole pixels
    {
        __syncthreads();  // the input tile is dead
#pragma unroll
        for (int orow = 0; orow < 4; ++orow)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
            {
                const int n = nt * 32 + px;
                if (n >= CO) continue;
#pragma unroll
                for (int k = 0; k < 16; ++k)
                {
                    const int x = 8 * (k >> 2) + 4 * h + (k & 3);
                    const float v = acc[orow][nt][k];
                    lds[((4 * wv + orow) * UT_W + x) * OS + n] =
                        (uint16_t)u_f2bf(EPI == UC_EPI_RELU ? fmaxf(v + bias[nt], 0.f) : v);
                }
            }
        __syncthreads();
        // head weights [8][6] + bias [6], read once (the stores below could alias them)
        float hw[HEAD ? 54 : 1];
        if constexpr (HEAD)
#pragma unroll
            for (int j = 0; j < 54; ++j) hw[j] = F.head_w[j];
        constexpr int QO = CO / 8;
        for (int i = tid; i < UT_H * UT_W * QO; i += 256)
        {
            const int pix = i / QO, q = i - pix * QO;
            const int gy = y0 + pix / UT_W, gx = x0 + (pix & (UT_W - 1));
            if (gy >= H || gx >= W) continue;
            u32x4 v = *reinterpret_cast<const u32x4 *>(&lds[pix * OS + q * 8]);
            const size_t o = ((size_t)gy * W + gx) * CO + q * 8;
            if constexpr (EPI == UC_EPI_MASK)
            {
                const u32x4 m = *reinterpret_cast<const u32x4 *>(F.mask + o);
                v.x = u_keep(m.x, v.x); v.y = u_keep(m.y, v.y); v.z = u_keep(m.z, v.z); v.w = u_keep(m.w, v.w);
            }
            *reinterpret_cast<u32x4 *>(static_cast<uint16_t *>(F.out) + o) = v;
            if constexpr (HEAD)
            {
                // the head on the stored bf16 activation: bias, then the 8 inputs in order
                const float in[8] = {u_lo(v.x), u_hi(v.x), u_lo(v.y), u_hi(v.y),
                                     u_lo(v.z), u_hi(v.z), u_lo(v.w), u_hi(v.w)};
                float *po = F.head_out + ((size_t)gy * W + gx) * 6;
#pragma unroll
                for (int k = 0; k < 6; ++k)
                {
                    float sum = hw[48 + k];
#pragma unroll
                    for (int c = 0; c < 8; ++c) sum += in[c] * hw[c * 6 + k];
                    po[k] = sum;
                }
            }
        }
    }
}

template <int SRC, int CA, int CB, int CO, int EPI, bool HEAD = false>
static hipError_t uconv_t(int n, const UcFrame *fr, hipStream_t st)
{
    UcBatch b{};
    int tx[XDN_MAX_FRAMES], ty[XDN_MAX_FRAMES];
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        tx[i] = udiv_up(fr[i].W, UT_W);
        ty[i] = udiv_up(fr[i].H, UT_H);
    }
    b.map = make_tilemap(n, tx, ty);
    if (b.map.total == 0) return hipSuccess;
    hipLaunchKernelGGL((uconv3x3<SRC, CA, CB, CO, EPI, HEAD>), dim3(b.map.total), dim3(256), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_uconv(int kind, int n, const UcFrame *fr, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (n > XDN_MAX_FRAMES) return hipErrorInvalidValue;
    switch (kind)
    {
    case UC_F_0: return uconv_t<UC_SRC_IMAGES, 0, 0, 8, UC_EPI_RELU>(n, fr, st);
    case UC_F_8_8: return uconv_t<UC_SRC_BF16, 8, 0, 8, UC_EPI_RELU>(n, fr, st);
    case UC_F_8_16: return uconv_t<UC_SRC_BF16, 8, 0, 16, UC_EPI_RELU>(n, fr, st);
    case UC_F_16_16: return uconv_t<UC_SRC_BF16, 16, 0, 16, UC_EPI_RELU>(n, fr, st);
    case UC_F_16_32: return uconv_t<UC_SRC_BF16, 16, 0, 32, UC_EPI_RELU>(n, fr, st);
    case UC_F_32_32: return uconv_t<UC_SRC_BF16, 32, 0, 32, UC_EPI_RELU>(n, fr, st);
    case UC_F_48_16: return uconv_t<UC_SRC_BF16, 32, 16, 16, UC_EPI_RELU>(n, fr, st);
    case UC_F_24_8: return uconv_t<UC_SRC_BF16, 16, 8, 8, UC_EPI_RELU>(n, fr, st);
    case UC_D_8_8: return uconv_t<UC_SRC_BF16, 8, 0, 8, UC_EPI_MASK>(n, fr, st);
    case UC_D_16_8: return uconv_t<UC_SRC_BF16, 16, 0, 8, UC_EPI_PLAIN>(n, fr, st);
    case UC_D_16_16: return uconv_t<UC_SRC_BF16, 16, 0, 16, UC_EPI_MASK>(n, fr, st);
    case UC_D_32_16: return uconv_t<UC_SRC_BF16, 32, 0, 16, UC_EPI_PLAIN>(n, fr, st);
    case UC_D_32_32: return uconv_t<UC_SRC_BF16, 32, 0, 32, UC_EPI_MASK>(n, fr, st);
    case UC_D_16_48: return uconv_t<UC_SRC_BF16, 16, 0, 48, UC_EPI_PLAIN>(n, fr, st);
    case UC_D_8_24: return uconv_t<UC_SRC_BF16, 8, 0, 24, UC_EPI_PLAIN>(n, fr, st);
    case UC_F_24_8_HEAD: return uconv_t<UC_SRC_BF16, 16, 8, 8, UC_EPI_RELU, true>(n, fr, st);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------------------------
// Weight gradient of a 3x3 layer: input slice [32 MT, 32 MT + 32) of concat[a, b] (or the images)
// ---------------------------------------------------------------------------------------------
struct UwBatch {
    int n;
    int blk_begin[XDN_MAX_FRAMES + 1];
    UwFrame f[XDN_MAX_FRAMES];
};
static_assert(sizeof(UwBatch) <= 4096, "kernel argument too large");

constexpr int US_W = 32, US_H = 8;                // strip: 8 rows x 32 px
constexpr int UX_C = US_W + 2, UX_R = US_H + 2;   // staged X: 10 rows x 34 px (1-px halo)

template <int SRC, int CA, int CB, int CO, int MT>
__global__ __launch_bounds__(256, 2) void uwgrad(UwBatch b)
{
    // input-slice channels of this instantiation; staged channel counts (pixel strides) 16 or 32,
    // zero-padded: with 16 the lane groups 0 and 1 read the same channels, so the accumulator rows
    // (or columns) 16-31 repeat 0-15 and are not written out
    constexpr int CS = SRC == UC_SRC_IMAGES ? 15 : ((CA + CB - 32 * MT) < 32 ? (CA + CB - 32 * MT) : 32);
    constexpr int XS = CS <= 16 ? 16 : 32, DS = CO <= 16 ? 16 : 32;
    constexpr int X_ELEMS = UX_R * UX_C * XS, D_ELEMS = US_H * US_W * DS;
    constexpr int LDS_ELEMS = (X_ELEMS + D_ELEMS) > 8192 ? (X_ELEMS + D_ELEMS) : 8192;  // >= 4 x 1024 floats
    __shared__ __attribute__((aligned(16))) uint16_t lds[LDS_ELEMS];
    __shared__ float dbs[4][32];
    uint16_t *xs = lds, *ds = lds + X_ELEMS;
    int f = 0;
    while (f + 1 < b.n && (int)blockIdx.x >= b.blk_begin[f + 1]) ++f;
    const UwFrame &F = b.f[f];
    const int li = blockIdx.x - b.blk_begin[f], nblk = b.blk_begin[f + 1] - b.blk_begin[f];
    const int H = F.H, W = F.W;
    const int sx = (W + US_W - 1) / US_W, sy = (H + US_H - 1) / US_H;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, h = lane >> 5, ch = lane & 31;
    const int rr = 2 * wv + h;
    const int grp = (lane >> 4) & 1, s16 = lane & 15, jq = s16 >> 2, qq = s16 & 3;
    const int cbx = (XS > 16 ? 16 * grp : 0) + 4 * qq, cbd = (DS > 16 ? 16 * grp : 0) + 4 * qq;

    constexpr int QX = XS / 8, QD = DS / 8;  // 16-byte chunks per staged pixel
    constexpr int NX = SRC == UC_SRC_BF16 ? (UX_R * UX_C * QX + 255) / 256 : (UX_R * UX_C + 255) / 256;
    constexpr int ND = US_H * US_W * QD / 256;
    u32x4 rx[SRC == UC_SRC_BF16 ? NX : 1], rd[ND];
    float fx[SRC == UC_SRC_BF16 ? 1 : NX][15];
    auto fetch = [&](int s) {
        const int y0 = (s / sx) * US_H, x0 = (s % sx) * US_W;
#pragma unroll
        for (int u = 0; u < NX; ++u)
        {
            if constexpr (SRC == UC_SRC_BF16)
            {
                const int i = tid + 256 * u, pix = i / QX, q = i % QX;
                const int r = pix / UX_C, col = pix - r * UX_C;
                const int gy = y0 - 1 + r, gx = x0 - 1 + col;
                const int g = 32 * MT + 8 * q;  // first concat channel of this chunk
                rx[u] = u32x4{0u, 0u, 0u, 0u};
                if (i < UX_R * UX_C * QX && gy >= 0 && gy < H && gx >= 0 && gx < W)
                {
                    const size_t p = (size_t)gy * W + gx;
                    if (g < CA)
                        rx[u] = *reinterpret_cast<const u32x4 *>(static_cast<const uint16_t *>(F.a) + p * CA + g);
                    else if (g < CA + CB)
                        rx[u] = *reinterpret_cast<const u32x4 *>(static_cast<const uint16_t *>(F.b) + p * CB + (g - CA));
                }
            }
            else
            {
                const int pix = tid + 256 * u;
                const int r = pix / UX_C, col = pix - r * UX_C;
                const int gy = y0 - 1 + r, gx = x0 - 1 + col;
#pragma unroll
                for (int j = 0; j < 15; ++j) fx[u][j] = 0.f;
                if (pix < UX_R * UX_C && gy >= 0 && gy < H && gx >= 0 && gx < W)
                {
                    const size_t p = (size_t)gy * W + gx;
#pragma unroll
                    for (int im = 0; im < 5; ++im)
                    {
                        const float *sp = F.img[im] + 3 * p;
                        fx[u][3 * im] = sp[0]; fx[u][3 * im + 1] = sp[1]; fx[u][3 * im + 2] = sp[2];
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < ND; ++u)
        {
            const int i = tid + 256 * u, pix = i / QD, q = i % QD;
            const int r = pix / US_W, col = pix - r * US_W;
            const int gy = y0 + r, gx = x0 + col;
            rd[u] = u32x4{0u, 0u, 0u, 0u};
            if (q < CO / 8 && gy < H && gx < W)
                rd[u] = *reinterpret_cast<const u32x4 *>(F.dy + ((size_t)gy * W + gx) * CO + q * 8);
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int u = 0; u < NX; ++u)
        {
            if constexpr (SRC == UC_SRC_BF16)
            {
                const int i = tid + 256 * u;
                if (i < UX_R * UX_C * QX) *reinterpret_cast<u32x4 *>(&xs[(i / QX) * XS + (i % QX) * 8]) = rx[u];
            }
            else
            {
                const int pix = tid + 256 * u;
                if (pix < UX_R * UX_C)
                {
                    u32x4 a, c, z = {0u, 0u, 0u, 0u};
                    a.x = u_pack2(fx[u][0], fx[u][1]); a.y = u_pack2(fx[u][2], fx[u][3]);
                    a.z = u_pack2(fx[u][4], fx[u][5]); a.w = u_pack2(fx[u][6], fx[u][7]);
                    c.x = u_pack2(fx[u][8], fx[u][9]); c.y = u_pack2(fx[u][10], fx[u][11]);
                    c.z = u_pack2(fx[u][12], fx[u][13]); c.w = u_pack2(fx[u][14], 0.f);
                    *reinterpret_cast<u32x4 *>(&xs[pix * XS]) = a;
                    *reinterpret_cast<u32x4 *>(&xs[pix * XS + 8]) = c;
                    if constexpr (XS > 16)
                    {
                        *reinterpret_cast<u32x4 *>(&xs[pix * XS + 16]) = z;
                        *reinterpret_cast<u32x4 *>(&xs[pix * XS + 24]) = z;
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < ND; ++u)
        {
            const int i = tid + 256 * u;
            *reinterpret_cast<u32x4 *>(&ds[(i / QD) * DS + (i % QD) * 8]) = rd[u];
        }
    };

    // CO > 16: 32x32x16 MFMA, M = 32 input channels, N = 32 output channels, K = 16 px (8 px of row
    // 2wv + h per lane half). CO <= 16: 16x16x32 MFMA, M = 16 input channels (MTI tiles), N = 16
    // output channels, K = 32 px = one strip row (8 px per lane group kg = lane / 16), a quarter of
    // the MFMA work and of the accumulator registers.
    constexpr bool N16 = CO <= 16;
    constexpr int MTI = XS / 16;
    typedef __attribute__((ext_vector_type(4))) float f32x4;
    f32x16 acc[N16 ? 1 : 9];
    f32x4 acc16[N16 ? 9 : 1][N16 ? MTI : 1];
    if constexpr (N16)
    {
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int m = 0; m < MTI; ++m)
#pragma unroll
                for (int k = 0; k < 4; ++k) acc16[t][m][k] = 0.f;
    }
    else
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int k = 0; k < 16; ++k) acc[t][k] = 0.f;
    float db = 0.f;
    const int kg = lane >> 4;
    const int nstrip = sx * sy;
    const StripWalk sw = xcd_strips(li, nblk, nstrip);
    if (sw.first < sw.end) fetch(sw.first);
    for (int s = sw.first; s < sw.end; s += sw.step)
    {
        __syncthreads();
        commit();
        __syncthreads();
        if (s + sw.step < sw.end) fetch(s + sw.step);
        if constexpr (N16)
        {
            // lane group kg takes pixels 8kg .. 8kg+7 of the row; lane L of the group reads the
            // transposed channel L (4 pixels per ds_read_b64_tr_b16)
#pragma unroll
            for (int r = 0; r < 2; ++r)
            {
                const int row = 2 * wv + r;
                const uint16_t *dp = &ds[(row * US_W + 8 * kg + jq) * DS + 4 * qq];
                const uint2 b0 = u_ld_tr(dp), b1 = u_ld_tr(dp + 4 * DS);
                const u32x4 B = {b0.x, b0.y, b1.x, b1.y};
                db += (u_lo(B.x) + u_hi(B.x)) + (u_lo(B.y) + u_hi(B.y)) + (u_lo(B.z) + u_hi(B.z)) +
                      (u_lo(B.w) + u_hi(B.w));
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int sh = 0; sh < 3; ++sh)
#pragma unroll
                        for (int m = 0; m < MTI; ++m)
                        {
                            const uint16_t *xp = &xs[((row + dy) * UX_C + 8 * kg + sh + jq) * XS + 16 * m + 4 * qq];
                            const uint2 a0 = u_ld_tr(xp), a1 = u_ld_tr(xp + 4 * XS);
                            acc16[dy * 3 + sh][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                __builtin_bit_cast(bf16x8, u32x4{a0.x, a0.y, a1.x, a1.y}),
                                __builtin_bit_cast(bf16x8, B), acc16[dy * 3 + sh][m], 0, 0, 0);
                        }
            }
            continue;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
            const uint16_t *dp = &ds[(rr * US_W + 8 * c + jq) * DS + cbd];
            const uint2 b0 = u_ld_tr(dp), b1 = u_ld_tr(dp + 4 * DS);
            const u32x4 B = {b0.x, b0.y, b1.x, b1.y};
            db += (u_lo(B.x) + u_hi(B.x)) + (u_lo(B.y) + u_hi(B.y)) + (u_lo(B.z) + u_hi(B.z)) + (u_lo(B.w) + u_hi(B.w));
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int sh = 0; sh < 3; ++sh)
                {
                    const uint16_t *xp = &xs[((rr + dy) * UX_C + 8 * c + sh + jq) * XS + cbx];
                    const uint2 a0 = u_ld_tr(xp), a1 = u_ld_tr(xp + 4 * XS);
                    acc[dy * 3 + sh] = u_mfma(u32x4{a0.x, a0.y, a1.x, a1.y}, B, acc[dy * 3 + sh]);
                }
        }
    }
    // compact partial [9][CS][CO] + [CO] of this input slice
    float *out = F.part + (size_t)li * compact_part(9, CS, CO);
    float *red = reinterpret_cast<float *>(lds);  // [4 waves][1024]
    static_assert(sizeof(lds) >= 4 * 1024 * sizeof(float), "reduction buffer");
#pragma unroll
    for (int t = 0; t < 9; ++t)
    {
        __syncthreads();
        if constexpr (N16)
        {
            // 16x16 accumulators: column co = lane % 16, rows ci = 16m + 4kg + i
#pragma unroll
            for (int m = 0; m < MTI; ++m)
#pragma unroll
                for (int i = 0; i < 4; ++i) red[wv * 1024 + (16 * m + 4 * kg + i) * 32 + (lane & 15)] = acc16[t][m][i];
        }
        else
#pragma unroll
            for (int k = 0; k < 16; ++k)
            {
                const int ci = 8 * (k >> 2) + 4 * h + (k & 3);
                red[wv * 1024 + ci * 32 + ch] = acc[t][k];
            }
        __syncthreads();
        for (int e = tid; e < CS * CO; e += 256)
        {
            const int ci = e / CO, co = e - ci * CO, r = ci * 32 + co;
            out[t * CS * CO + e] = (red[r] + red[1024 + r]) + (red[2048 + r] + red[3072 + r]);
        }
    }
    if constexpr (N16)
    {
        db += __shfl_xor(db, 16);
        db += __shfl_xor(db, 32);
        if (lane < 16) dbs[wv][lane] = db;
    }
    else
    {
        db += __shfl_xor(db, 32);
        if (h == 0) dbs[wv][ch] = db;
    }
    __syncthreads();
    if (tid < CO) out[9 * CS * CO + tid] = (dbs[0][tid] + dbs[1][tid]) + (dbs[2][tid] + dbs[3][tid]);
}

// up to 512 blocks per image (two per CU): the quarter- and half-resolution layers have few strips
int ublocks(int H, int W)
{
    const int strips = udiv_up(W, US_W) * udiv_up(H, US_H);
    return strips < 1 ? 1 : (strips < 512 ? strips : 512);
}

template <int SRC, int CA, int CB, int CO, int MT>
static hipError_t uwgrad_t(int n, const UwFrame *fr, hipStream_t st)
{
    UwBatch b{};
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
    hipLaunchKernelGGL((uwgrad<SRC, CA, CB, CO, MT>), dim3(acc), dim3(256), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_uwgrad(int kind, int n, const UwFrame *fr, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (n > XDN_MAX_FRAMES) return hipErrorInvalidValue;
    switch (kind)
    {
    case UW_0: return uwgrad_t<UC_SRC_IMAGES, 0, 0, 8, 0>(n, fr, st);
    case UW_8_8: return uwgrad_t<UC_SRC_BF16, 8, 0, 8, 0>(n, fr, st);
    case UW_8_16: return uwgrad_t<UC_SRC_BF16, 8, 0, 16, 0>(n, fr, st);
    case UW_16_16: return uwgrad_t<UC_SRC_BF16, 16, 0, 16, 0>(n, fr, st);
    case UW_16_32: return uwgrad_t<UC_SRC_BF16, 16, 0, 32, 0>(n, fr, st);
    case UW_32_32: return uwgrad_t<UC_SRC_BF16, 32, 0, 32, 0>(n, fr, st);
    case UW_48_16_M0: return uwgrad_t<UC_SRC_BF16, 32, 16, 16, 0>(n, fr, st);
    case UW_48_16_M1: return uwgrad_t<UC_SRC_BF16, 32, 16, 16, 1>(n, fr, st);
    case UW_24_8: return uwgrad_t<UC_SRC_BF16, 16, 8, 8, 0>(n, fr, st);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------------------------
// Max pool and its gradient
// ---------------------------------------------------------------------------------------------
struct PoolBatch {
    int n, total;
    int begin[XDN_MAX_FRAMES + 1];  // first item of each frame
    PoolFrame f[XDN_MAX_FRAMES];
};
struct PoolBwBatch {
    int n, total;
    int begin[XDN_MAX_FRAMES + 1];
    PoolBwFrame f[XDN_MAX_FRAMES];
};

__device__ __forceinline__ int item_frame(const int *begin, int n, int i)
{
    int f = 0;
    while (f + 1 < n && i >= begin[f + 1]) ++f;
    return f;
}

// one item = one pooled pixel x one 8-channel chunk
template <int C>
__global__ __launch_bounds__(256) void maxpool2(PoolBatch b)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= b.total) return;
    const int f = item_frame(b.begin, b.n, i);
    const PoolFrame &F = b.f[f];
    const int W2 = F.W / 2;
    const int li = i - b.begin[f], q = li % (C / 8), p2 = li / (C / 8);
    const int y = p2 / W2, x = p2 - y * W2;
    u32x4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        v[k] = *reinterpret_cast<const u32x4 *>(F.in + ((size_t)(2 * y + (k >> 1)) * F.W + 2 * x + (k & 1)) * C + q * 8);
    uint32_t o[4];
#pragma unroll
    for (int w = 0; w < 4; ++w)
    {
        float lo = -__builtin_inff(), hi = -__builtin_inff();
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            const uint32_t word = w == 0 ? v[k].x : (w == 1 ? v[k].y : (w == 2 ? v[k].z : v[k].w));
            lo = fmaxf(lo, u_lo(word));
            hi = fmaxf(hi, u_hi(word));
        }
        o[w] = (__float_as_uint(lo) >> 16) | (__float_as_uint(hi) & 0xFFFF0000u);
    }
    *reinterpret_cast<u32x4 *>(F.out + (size_t)p2 * C + q * 8) = u32x4{o[0], o[1], o[2], o[3]};
}

// one item = one 2x2 window x one 8-channel chunk: vector loads of the window's four activations,
// skip gradients and outputs
template <int C>
__global__ __launch_bounds__(256) void pool_bwd(PoolBwBatch b)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= b.total) return;
    const int f = item_frame(b.begin, b.n, i);
    const PoolBwFrame &F = b.f[f];
    const int W = F.W, W2 = W / 2;
    const int li = i - b.begin[f], q = li % (C / 8), p2 = li / (C / 8);
    const int y2 = p2 / W2, x2 = p2 - y2 * W2;
    size_t pix[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) pix[k] = (size_t)(2 * y2 + (k >> 1)) * W + 2 * x2 + (k & 1);
    u32x4 a[4], sk[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        a[k] = *reinterpret_cast<const u32x4 *>(F.act + pix[k] * C + q * 8);
        sk[k] = *reinterpret_cast<const u32x4 *>(F.dskip + pix[k] * F.sstride + F.soff + q * 8);
    }
    const u32x4 dpw = *reinterpret_cast<const u32x4 *>(F.dpool + (size_t)p2 * C + q * 8);
    auto half = [](const u32x4 &w, int c) {
        const uint32_t x = c < 2 ? w.x : (c < 4 ? w.y : (c < 6 ? w.z : w.w));
        return (c & 1) ? u_hi(x) : u_lo(x);
    };
    float r[4][8];
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = half(a[k], c);
        // first maximum of the window in row-major order
        int arg = 0;
        float best = v[0];
#pragma unroll
        for (int k = 1; k < 4; ++k)
            if (v[k] > best) best = v[k], arg = k;
        const float dp = half(dpw, c);
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            const float g = (arg == k ? dp : 0.f) + half(sk[k], c);
            r[k][c] = v[k] > 0.f ? g : 0.f;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        *reinterpret_cast<u32x4 *>(F.dy + pix[k] * C + q * 8) =
            u32x4{u_pack2(r[k][0], r[k][1]), u_pack2(r[k][2], r[k][3]), u_pack2(r[k][4], r[k][5]),
                  u_pack2(r[k][6], r[k][7])};
}

template <int C>
static hipError_t maxpool_t(int n, const PoolFrame *fr, hipStream_t st)
{
    PoolBatch b{};
    b.n = n;
    int acc = 0;
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        b.begin[i] = acc;
        acc += (fr[i].H / 2) * (fr[i].W / 2) * (C / 8);
    }
    b.begin[n] = b.total = acc;
    if (acc == 0) return hipSuccess;
    hipLaunchKernelGGL(maxpool2<C>, dim3(udiv_up(acc, 256)), dim3(256), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_maxpool(int C, int n, const PoolFrame *fr, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (n > XDN_MAX_FRAMES) return hipErrorInvalidValue;
    if (C == 8) return maxpool_t<8>(n, fr, st);
    if (C == 16) return maxpool_t<16>(n, fr, st);
    return hipErrorInvalidValue;
}

template <int C>
static hipError_t pool_bwd_t(int n, const PoolBwFrame *fr, hipStream_t st)
{
    PoolBwBatch b{};
    b.n = n;
    int acc = 0;
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        b.begin[i] = acc;
        acc += (fr[i].H / 2) * (fr[i].W / 2) * (C / 8);
    }
    b.begin[n] = b.total = acc;
    if (acc == 0) return hipSuccess;
    hipLaunchKernelGGL(pool_bwd<C>, dim3(udiv_up(acc, 256)), dim3(256), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_pool_bwd(int C, int n, const PoolBwFrame *fr, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (n > XDN_MAX_FRAMES) return hipErrorInvalidValue;
    if (C == 8) return pool_bwd_t<8>(n, fr, st);
    if (C == 16) return pool_bwd_t<16>(n, fr, st);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------------------------
// Transposed 1x1 stride-2 convolution
// ---------------------------------------------------------------------------------------------
struct CtBatch {
    int n;
    int blk_begin[XDN_MAX_FRAMES + 1];
    CtFrame f[XDN_MAX_FRAMES];
};
static_assert(sizeof(CtBatch) <= 4096, "kernel argument too large");

// forward: out(2y,2x) = W in(y,x) + b, the other three outputs of the 2x2 block = b. FP32 MFMA
// (v_mfma_f32_32x32x2_f32: M = output channels, N = 32 coarse pixels, K = 2 input channels per
// step; a wave takes four such pixel groups with the weights in registers) on the FP32 master
// weights, so the products are those of FP32 arithmetic;
// a lane ends with 16 (C = 32) or 8 (C = 16) channels c(i) = 8(i/4) + 4h + i%4 of its pixel, and a
// permlane32 swap gives it 8 consecutive channels per 16-B store.
template <int C>
__global__ __launch_bounds__(256) void convT1x1s2(CtBatch b)
{
    constexpr int NI = C == 32 ? 16 : 8;  // accumulator rows holding output channels
    constexpr int NG = NI / 8;            // 16-B stores per output pixel and lane
    constexpr int GPW = 4;                // 32-pixel groups per wave
    int f = 0;
    while (f + 1 < b.n && (int)blockIdx.x >= b.blk_begin[f + 1]) ++f;
    const CtFrame &F = b.f[f];
    const int li = blockIdx.x - b.blk_begin[f];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, px = lane & 31, h = lane >> 5;
    const int w = F.w_, np = F.h * w, W2 = 2 * w;
    // A operand of step s: W[co = px][ci = 2s + h]; the bias of the lane's outputs
    float wa[C / 2], bo[NI];
#pragma unroll
    for (int s2 = 0; s2 < C / 2; ++s2) wa[s2] = px < C ? F.w[px * C + 2 * s2 + h] : 0.f;
#pragma unroll
    for (int i = 0; i < NI; ++i) bo[i] = F.b[8 * (i >> 2) + 4 * h + (i & 3)];
    u32x4 vb[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g)
    {
        const int cb = 16 * g + 8 * h;
        vb[g] = u32x4{u_pack2(F.b[cb], F.b[cb + 1]), u_pack2(F.b[cb + 2], F.b[cb + 3]),
                      u_pack2(F.b[cb + 4], F.b[cb + 5]), u_pack2(F.b[cb + 6], F.b[cb + 7])};
    }
    for (int gi = 0; gi < GPW; ++gi)
    {
        const int p = ((li * 4 + wv) * GPW + gi) * 32 + px;
        if (p - px >= np) break;  // uniform across the wave
        // B operand: in[p][ci = 2s + h]
        float xb[C / 2];
#pragma unroll
        for (int s2 = 0; s2 < C / 2; ++s2) xb[s2] = 0.f;
        if (p < np)
#pragma unroll
            for (int c = 0; c < C; c += 8)
            {
                const u32x4 v = *reinterpret_cast<const u32x4 *>(F.in + (size_t)p * C + c);
                const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) xb[c / 2 + j] = h ? u_hi(wd[j]) : u_lo(wd[j]);
            }
        f32x16 acc;
#pragma unroll
        for (int k = 0; k < 16; ++k) acc[k] = 0.f;
#pragma unroll
        for (int s2 = 0; s2 < C / 2; ++s2) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(wa[s2], xb[s2], acc, 0, 0, 0);
        uint32_t d[NI / 2];
#pragma unroll
        for (int q = 0; q < NI / 2; ++q) d[q] = u_pack2(acc[2 * q] + bo[2 * q], acc[2 * q + 1] + bo[2 * q + 1]);
        // lane h gathers channels 16g + 8h .. +7 (see conv3x3_l1 of k_cnn.hip)
        u32x4 v[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
            for (int e = 0; e < 2; ++e)
            {
                const auto sw = __builtin_amdgcn_permlane32_swap(d[4 * g + e], d[4 * g + 2 + e], false, false);
                v[g][e] = (uint32_t)sw[0];
                v[g][2 + e] = (uint32_t)sw[1];
            }
        if (p < np)
        {
            const int y = p / w, x = p - y * w;
#pragma unroll
            for (int k = 0; k < 4; ++k)
            {
                uint16_t *o = F.out + ((size_t)(2 * y + (k >> 1)) * W2 + 2 * x + (k & 1)) * C + 8 * h;
#pragma unroll
                for (int g = 0; g < NG; ++g) *reinterpret_cast<u32x4 *>(o + 16 * g) = k == 0 ? v[g] : vb[g];
            }
        }
    }
}

// backward: blocks walk coarse pixels in chunks of 256, one per thread for the loads: the chunk's
// inputs and dOut(2y,2x) go to LDS as they are (bf16, pixel-major) and each thread adds the sum of
// dOut over its 2x2 output block to its bias-gradient sums. Each wave then takes 64 pixels of the
// chunk: dIn = W^T dOut(2y,2x) on FP32 MFMA for 32 channels (FP32 VALU for 16), FP32 master
// weights, masked by the input's ReLU; dW[ci][co] = sum in (x) dOut(2y,2x) on bf16 MFMA (M = ci,
// N = co, K = pixels; operands through ds_read_b64_tr_b16). The waves' accumulators and the
// threads' bias sums are reduced through LDS in a fixed order at the end.
template <int C>
__global__ __launch_bounds__(256) void convT_bwd(CtBatch b)
{
    static_assert(C == 16 || C == 32, "channel count");
    constexpr int Q = C / 8;
    __shared__ __attribute__((aligned(16))) uint16_t tile[2][256 * C];  // inputs, dOut(2y,2x)
    __shared__ float red[4][C * C + C];
    __shared__ __attribute__((aligned(16))) float ws[C == 16 ? C * C : 4];  // [co][ci] (VALU dIn)
    int f = 0;
    while (f + 1 < b.n && (int)blockIdx.x >= b.blk_begin[f + 1]) ++f;
    const CtFrame &F = b.f[f];
    const int li = blockIdx.x - b.blk_begin[f], nblk = b.blk_begin[f + 1] - b.blk_begin[f];
    const int h = F.h, w = F.w_, np = h * w, W2 = 2 * w;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int NI = C == 32 ? 16 : 8;  // dIn accumulator rows holding input channels
    // C = 32: A operand of the dIn product, step s: W^T[ci = lane%32][co = 2s + lane/32]
    float wt[C == 32 ? C / 2 : 1];
    if constexpr (C == 32)
#pragma unroll
        for (int s2 = 0; s2 < C / 2; ++s2) wt[s2] = F.w[(2 * s2 + (lane >> 5)) * C + (lane & 31)];
    else
        for (int e = tid; e < C * C; e += 256) ws[e] = F.w[e];
    float bacc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) bacc[c] = 0.f;
    typedef __attribute__((ext_vector_type(C == 32 ? 16 : 4))) float accv;
    accv acc;
#pragma unroll
    for (int k = 0; k < (C == 32 ? 16 : 4); ++k) acc[k] = 0.f;
    // transposing-read addressing (see ld_tr of k_cnn.hip): lane s of a 16-lane group addresses
    // pixel +s/4, channels +4(s%4)
    const int s16 = lane & 15, jq = s16 >> 2, qq = s16 & 3;
    for (int base = li * 256; base < np; base += nblk * 256)
    {
        __syncthreads();  // the previous chunk's tiles are read
        const int p = base + tid;
        u32x4 xv[Q], gv[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) xv[k] = gv[k] = u32x4{0u, 0u, 0u, 0u};
        float gs[C];
#pragma unroll
        for (int c = 0; c < C; ++c) gs[c] = 0.f;
        if (p < np)
        {
            const int y = p / w, xx = p - y * w;
#pragma unroll
            for (int k = 0; k < Q; ++k) xv[k] = *reinterpret_cast<const u32x4 *>(F.in + (size_t)p * C + 8 * k);
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                const uint16_t *gp = F.dout + ((size_t)(2 * y + (t >> 1)) * W2 + 2 * xx + (t & 1)) * F.dstride;
#pragma unroll
                for (int k = 0; k < Q; ++k)
                {
                    const u32x4 v = *reinterpret_cast<const u32x4 *>(gp + 8 * k);
                    if (t == 0) gv[k] = v;
                    const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                    {
                        gs[8 * k + 2 * j] += u_lo(wd[j]);
                        gs[8 * k + 2 * j + 1] += u_hi(wd[j]);
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < Q; ++k)
        {
            *reinterpret_cast<u32x4 *>(&tile[0][tid * C + 8 * k]) = xv[k];
            *reinterpret_cast<u32x4 *>(&tile[1][tid * C + 8 * k]) = gv[k];
        }
#pragma unroll
        for (int c = 0; c < C; ++c) bacc[c] += gs[c];
        __syncthreads();  // this chunk's tiles are visible
        // dIn. C = 32: FP32 MFMA (A = W^T from registers, B = dOut(2y,2x) of the wave's 32-pixel group
        // from the tile), masked by the staged input and stored through the permlane32 swap. C = 16
        // (half the MFMA rows would be padding): FP32 VALU per thread, output channels in order.
        if constexpr (C == 16)
        {
            if (p < np)
            {
                float x[C], g[C];
#pragma unroll
                for (int k = 0; k < Q; ++k)
                {
                    const uint32_t xw[4] = {xv[k].x, xv[k].y, xv[k].z, xv[k].w};
                    const uint32_t gw[4] = {gv[k].x, gv[k].y, gv[k].z, gv[k].w};
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                    {
                        x[8 * k + 2 * j] = u_lo(xw[j]); x[8 * k + 2 * j + 1] = u_hi(xw[j]);
                        g[8 * k + 2 * j] = u_lo(gw[j]); g[8 * k + 2 * j + 1] = u_hi(gw[j]);
                    }
                }
                float d[C];
#pragma unroll
                for (int c = 0; c < C; ++c) d[c] = 0.f;
#pragma unroll
                for (int co = 0; co < C; ++co)
#pragma unroll
                    for (int c = 0; c < C; c += 4)
                    {
                        const float4 wq = *reinterpret_cast<const float4 *>(&ws[co * C + c]);
                        d[c] += wq.x * g[co]; d[c + 1] += wq.y * g[co]; d[c + 2] += wq.z * g[co]; d[c + 3] += wq.w * g[co];
                    }
#pragma unroll
                for (int c = 0; c < C; c += 8)
                    *reinterpret_cast<u32x4 *>(F.din + (size_t)p * C + c) =
                        u32x4{u_pack2(x[c] > 0.f ? d[c] : 0.f, x[c + 1] > 0.f ? d[c + 1] : 0.f),
                              u_pack2(x[c + 2] > 0.f ? d[c + 2] : 0.f, x[c + 3] > 0.f ? d[c + 3] : 0.f),
                              u_pack2(x[c + 4] > 0.f ? d[c + 4] : 0.f, x[c + 5] > 0.f ? d[c + 5] : 0.f),
                              u_pack2(x[c + 6] > 0.f ? d[c + 6] : 0.f, x[c + 7] > 0.f ? d[c + 7] : 0.f)};
            }
        }
        else
#pragma unroll
        for (int gq = 0; gq < 2; ++gq)
        {
            const int pl = 64 * wv + 32 * gq + (lane & 31), hh = lane >> 5;
            float xb[C / 2];
#pragma unroll
            for (int k = 0; k < Q; ++k)
            {
                const u32x4 v = *reinterpret_cast<const u32x4 *>(&tile[1][pl * C + 8 * k]);
                const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) xb[4 * k + j] = hh ? u_hi(wd[j]) : u_lo(wd[j]);
            }
            f32x16 dacc;
#pragma unroll
            for (int k = 0; k < 16; ++k) dacc[k] = 0.f;
#pragma unroll
            for (int s2 = 0; s2 < C / 2; ++s2) dacc = __builtin_amdgcn_mfma_f32_32x32x2f32(wt[s2], xb[s2], dacc, 0, 0, 0);
            // rows c(i) = 8(i/4) + 4hh + i%4: four bf16 inputs at 8j + 4hh per j
            uint32_t d[NI / 2];
#pragma unroll
            for (int j = 0; j < NI / 4; ++j)
            {
                const uint2 xm = *reinterpret_cast<const uint2 *>(&tile[0][pl * C + 8 * j + 4 * hh]);
                const float xs[4] = {u_lo(xm.x), u_hi(xm.x), u_lo(xm.y), u_hi(xm.y)};
                d[2 * j] = u_pack2(xs[0] > 0.f ? dacc[4 * j] : 0.f, xs[1] > 0.f ? dacc[4 * j + 1] : 0.f);
                d[2 * j + 1] = u_pack2(xs[2] > 0.f ? dacc[4 * j + 2] : 0.f, xs[3] > 0.f ? dacc[4 * j + 3] : 0.f);
            }
            u32x4 v[NI / 8];
#pragma unroll
            for (int g = 0; g < NI / 8; ++g)
#pragma unroll
                for (int e = 0; e < 2; ++e)
                {
                    const auto sw = __builtin_amdgcn_permlane32_swap(d[4 * g + e], d[4 * g + 2 + e], false, false);
                    v[g][e] = (uint32_t)sw[0];
                    v[g][2 + e] = (uint32_t)sw[1];
                }
            const int pp = base + pl;
            if (pp < np)
#pragma unroll
                for (int g = 0; g < NI / 8; ++g)
                    *reinterpret_cast<u32x4 *>(F.din + (size_t)pp * C + 16 * g + 8 * hh) = v[g];
        }
        // dW: pixels past the end of the image were staged as zeros
        if constexpr (C == 32)
        {
            // 32x32x16: lane l holds row ci = l%32, k = pixels 8(l/32) + 0..7 of the K step
            const int grp = (lane >> 4) & 1, hh = lane >> 5;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
            {
                const int pb = 64 * wv + 16 * ks + 8 * hh + jq;
                const uint16_t *xp = &tile[0][pb * C + 16 * grp + 4 * qq];
                const uint16_t *gp = &tile[1][pb * C + 16 * grp + 4 * qq];
                const uint2 a0 = u_ld_tr(xp), a1 = u_ld_tr(xp + 4 * C);
                const uint2 b0 = u_ld_tr(gp), b1 = u_ld_tr(gp + 4 * C);
                acc = u_mfma(u32x4{a0.x, a0.y, a1.x, a1.y}, u32x4{b0.x, b0.y, b1.x, b1.y}, acc);
            }
        }
        else
        {
            // 16x16x32: lane l holds row ci = l%16, k = pixels 8(l/16) + 0..7 of the K step
            const int kg = lane >> 4;
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
            {
                const int pb = 64 * wv + 32 * ks + 8 * kg + jq;
                const uint16_t *xp = &tile[0][pb * C + 4 * qq];
                const uint16_t *gp = &tile[1][pb * C + 4 * qq];
                const uint2 a0 = u_ld_tr(xp), a1 = u_ld_tr(xp + 4 * C);
                const uint2 b0 = u_ld_tr(gp), b1 = u_ld_tr(gp + 4 * C);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(bf16x8, u32x4{a0.x, a0.y, a1.x, a1.y}),
                    __builtin_bit_cast(bf16x8, u32x4{b0.x, b0.y, b1.x, b1.y}), acc, 0, 0, 0);
            }
        }
    }
    // bias sums: wave reduction in a fixed order, then the four waves
#pragma unroll
    for (int c = 0; c < C; ++c)
        for (int off = 32; off > 0; off >>= 1) bacc[c] += __shfl_down(bacc[c], off);
    // accumulators: C[row = ci][col = co]
#pragma unroll
    for (int k = 0; k < (C == 32 ? 16 : 4); ++k)
    {
        const int ci = C == 32 ? 8 * (k >> 2) + 4 * (lane >> 5) + (k & 3) : 4 * (lane >> 4) + k;
        const int co = C == 32 ? (lane & 31) : (lane & 15);
        red[wv][ci * C + co] = acc[k];
    }
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < C; ++c) red[wv][C * C + c] = bacc[c];
    __syncthreads();
    // compact partial: dW [ci][co], then db
    float *out = F.part + (size_t)li * compact_part(1, C, C);
    for (int e = tid; e < C * C + C; e += 256) out[e] = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
}

template <int C, bool BWD>
static hipError_t convT_t(int n, const CtFrame *fr, hipStream_t st)
{
    CtBatch b{};
    b.n = n;
    int acc = 0;
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        b.blk_begin[i] = acc;
        acc += BWD ? fr[i].nblk : udiv_up(fr[i].h * fr[i].w_, 4 * 4 * 32);
    }
    b.blk_begin[n] = acc;
    if (acc == 0) return hipSuccess;
    if (BWD)
        hipLaunchKernelGGL(convT_bwd<C>, dim3(acc), dim3(256), 0, st, b);
    else
        hipLaunchKernelGGL(convT1x1s2<C>, dim3(acc), dim3(256), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_convT(int C, int n, const CtFrame *fr, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (n > XDN_MAX_FRAMES) return hipErrorInvalidValue;
    if (C == 16) return convT_t<16, false>(n, fr, st);
    if (C == 32) return convT_t<32, false>(n, fr, st);
    return hipErrorInvalidValue;
}

hipError_t launch_convT_bwd(int C, int n, const CtFrame *fr, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (n > XDN_MAX_FRAMES) return hipErrorInvalidValue;
    if (C == 16) return convT_t<16, true>(n, fr, st);
    if (C == 32) return convT_t<32, true>(n, fr, st);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------------------------
// Head backward: 1x1 convolution 8 -> 6
// ---------------------------------------------------------------------------------------------
struct HeadBatch {
    int n;
    int blk_begin[XDN_MAX_FRAMES + 1];
    HeadFrame f[XDN_MAX_FRAMES];
};
static_assert(sizeof(HeadBatch) <= 4096, "kernel argument too large");

__global__ __launch_bounds__(256) void head_bwd(HeadBatch b)
{
    __shared__ float red[4][54];
    int f = 0;
    while (f + 1 < b.n && (int)blockIdx.x >= b.blk_begin[f + 1]) ++f;
    const HeadFrame &F = b.f[f];
    const int li = blockIdx.x - b.blk_begin[f], nblk = b.blk_begin[f + 1] - b.blk_begin[f];
    const int np = F.H * F.W, tid = threadIdx.x;
    float w[48];
#pragma unroll
    for (int e = 0; e < 48; ++e) w[e] = F.w[e];
    float g[54];
#pragma unroll
    for (int e = 0; e < 54; ++e) g[e] = 0.f;
    for (int p = li * 256 + tid; p < np; p += nblk * 256)
    {
        const u32x4 v = *reinterpret_cast<const u32x4 *>(F.in + (size_t)p * 8);
        const float in[8] = {u_lo(v.x), u_hi(v.x), u_lo(v.y), u_hi(v.y), u_lo(v.z), u_hi(v.z), u_lo(v.w), u_hi(v.w)};
        float dp[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) dp[k] = F.dparam[(size_t)p * 6 + k];
        uint32_t o[4];
#pragma unroll
        for (int c = 0; c < 8; c += 2)
        {
            float s0 = 0.f, s1 = 0.f;
#pragma unroll
            for (int k = 0; k < 6; ++k)
            {
                s0 += w[c * 6 + k] * dp[k];
                s1 += w[(c + 1) * 6 + k] * dp[k];
            }
            o[c / 2] = u_pack2(in[c] > 0.f ? s0 : 0.f, in[c + 1] > 0.f ? s1 : 0.f);
        }
        *reinterpret_cast<u32x4 *>(F.din + (size_t)p * 8) = u32x4{o[0], o[1], o[2], o[3]};
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int k = 0; k < 6; ++k) g[c * 6 + k] += in[c] * dp[k];
#pragma unroll
        for (int k = 0; k < 6; ++k) g[48 + k] += dp[k];
    }
#pragma unroll
    for (int e = 0; e < 54; ++e)
    {
        float v = g[e];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        g[e] = v;
    }
    if ((tid & 63) == 0)
#pragma unroll
        for (int e = 0; e < 54; ++e) red[tid >> 6][e] = g[e];
    __syncthreads();
    float *out = F.part + (size_t)li * compact_part(1, 8, 6);  // [8][6] then [6]
    if (tid < 54) out[tid] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
}

hipError_t launch_head_bwd(int n, const HeadFrame *fr, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (n > XDN_MAX_FRAMES) return hipErrorInvalidValue;
    HeadBatch b{};
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
    hipLaunchKernelGGL(head_bwd, dim3(acc), dim3(256), 0, st, b);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Weight packing: fwd frag[t][kc][nt][lane] = W[t][kc*16 + 8h + j][nt*32 + lane%32],
// dgrad frag[t][kc][nt][lane] = W[8-t][nt*32 + lane%32][kc*16 + 8h + j]; zero outside ci x co
// ---------------------------------------------------------------------------------------------
constexpr int UPACK_MAX = 24;  // jobs per launch
struct UPackBatch {
    int n;
    UPackJob j[UPACK_MAX];
};
static_assert(sizeof(UPackBatch) <= 4096, "kernel argument too large");

__host__ __device__ inline int upack_items(const UPackJob &J)
{
    if (J.copy_n > 0) return (J.copy_n + 63) / 64;
    return 9 * ((J.ci + 15) / 16) * ((J.co + 31) / 32) + (J.frag_dgrad ? 9 * ((J.co + 15) / 16) * ((J.ci + 31) / 32) : 0);
}

__global__ __launch_bounds__(64) void upack(UPackBatch b)
{
    const UPackJob &J = b.j[blockIdx.y];
    const int lane = threadIdx.x, h = lane >> 5, n = lane & 31;
    const int item = blockIdx.x;
    if (J.copy_n > 0)
    {
        const int e = item * 64 + lane;
        if (e < J.copy_n) static_cast<float *>(J.frag_fwd)[e] = J.w[e];
        return;
    }
    const int kcf = (J.ci + 15) / 16, ntf = (J.co + 31) / 32;
    const int kcd = (J.co + 15) / 16, ntd = (J.ci + 31) / 32;
    const int nf = 9 * kcf * ntf, nd = J.frag_dgrad ? 9 * kcd * ntd : 0;
    if (item < nf)
    {
        const int t = item / (kcf * ntf), r = item - t * kcf * ntf, kc = r / ntf, nt = r - kc * ntf;
        const int co = nt * 32 + n;
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 8; j += 2)
        {
            const int c0 = kc * 16 + 8 * h + j, c1 = c0 + 1;
            const float a = (c0 < J.ci && co < J.co) ? J.w[(t * J.ci + c0) * J.co + co] : 0.f;
            const float bb = (c1 < J.ci && co < J.co) ? J.w[(t * J.ci + c1) * J.co + co] : 0.f;
            w[j >> 1] = u_pack2(a, bb);
        }
        static_cast<u32x4 *>(J.frag_fwd)[item * 64 + lane] = u32x4{w[0], w[1], w[2], w[3]};
        if (t == 0 && kc == 0 && h == 0) J.bias32[nt * 32 + n] = co < J.co ? J.b[co] : 0.f;
    }
    else if (item < nf + nd)
    {
        const int it = item - nf;
        const int t = it / (kcd * ntd), r = it - t * kcd * ntd, kc = r / ntd, nt = r - kc * ntd;
        const int ci = nt * 32 + n;
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 8; j += 2)
        {
            const int o0 = kc * 16 + 8 * h + j, o1 = o0 + 1;
            const float a = (o0 < J.co && ci < J.ci) ? J.w[((8 - t) * J.ci + ci) * J.co + o0] : 0.f;
            const float bb = (o1 < J.co && ci < J.ci) ? J.w[((8 - t) * J.ci + ci) * J.co + o1] : 0.f;
            w[j >> 1] = u_pack2(a, bb);
        }
        static_cast<u32x4 *>(J.frag_dgrad)[it * 64 + lane] = u32x4{w[0], w[1], w[2], w[3]};
    }
}

hipError_t launch_upack(int n, const UPackJob *jobs, hipStream_t st)
{
    for (int c0 = 0; c0 < n; c0 += UPACK_MAX)
    {
        const int m = n - c0 < UPACK_MAX ? n - c0 : UPACK_MAX;
        UPackBatch b{};
        b.n = m;
        int maxi = 0;
        for (int i = 0; i < m; ++i)
        {
            b.j[i] = jobs[c0 + i];
            const int items = upack_items(b.j[i]);
            maxi = items > maxi ? items : maxi;
        }
        if (maxi == 0) continue;
        hipLaunchKernelGGL(upack, dim3(maxi, m), dim3(64), 0, st, b);
        if (hipError_t e = hipGetLastError()) return e;
    }
    return hipSuccess;
}

}  // namespace xdn
