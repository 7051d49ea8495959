// Shared device helpers for the xdn gfx950 kernels.
//
// Every kernel in this library works on a *frame batch*: a by-value table of up to
// XDN_MAX_FRAMES frames (each with its own H, W and buffer pointers) that one launch tiles
// together. A single op call is a batch of the B images of an NHWC tensor; the serving loop
// batches frames of different streams (and resolutions) into the same launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define XDN_MAX_FRAMES 16
#define XDN_NUM_XCD 8

namespace xdn {

// Reflect-101 ("mirror") index used by the regression and the spatial filter
// (regression.cu.cc:128-129, 413-414). Valid for -n < x < 2n-1.
__device__ __forceinline__ int mirror(int x, int n) { return x >= n ? 2 * n - 2 - x : (x < 0 ? -x : x); }

// iDivUp with C integer division (cuda_utils.cuh:275-278); not a ceiling for negative a.
__device__ __forceinline__ int idivup(int a, int b) { return (a % b != 0) ? (a / b + 1) : (a / b); }

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

constexpr float kLog2e = 1.4426950408889634f;

// Tiling of one launch over the frames of a batch.
struct TileMap {
    int n;                             // frames in the batch
    int total;                         // total tiles
    int begin[XDN_MAX_FRAMES + 1];     // first tile of each frame
    int tiles_x[XDN_MAX_FRAMES];       // tiles along x per frame
};

// XCD-aware block order: blocks are dealt round-robin to the 8 XCDs; give each XCD a
// contiguous run of tiles so that neighbouring tiles (which share halos) hit the same L2.
__device__ __forceinline__ int xcd_tile(int bid, int total)
{
    const int per = total / XDN_NUM_XCD, rem = total % XDN_NUM_XCD;
    const int xcd = bid % XDN_NUM_XCD, local = bid / XDN_NUM_XCD;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// XCD-aware strip walk of a weight-gradient block: the frame's blocks are li = 0 .. nblk - 1 and
// blocks dealt round-robin to the 8 XCDs, so the blocks with one li mod 8 share an XCD wherever the
// frame's blocks start in the launch. Each such group takes a contiguous band of the frame's nstrip
// strips (row-major, so vertically adjacent strips, which share halo rows, meet in one L2) and its
// blocks take the band's strips round-robin: first, first + step, ... below end. The walk depends
// on li only, so a frame's partial sums are the same alone and in a batch. With fewer than 8
// blocks the plain round-robin over all blocks is used.
struct StripWalk {
    int first, step, end;
};
__device__ __forceinline__ StripWalk xcd_strips(int li, int nblk, int nstrip)
{
    if (nblk < XDN_NUM_XCD) return StripWalk{li, nblk, nstrip};
    const int g = li % XDN_NUM_XCD, lb = li / XDN_NUM_XCD, nb = (nblk - g + XDN_NUM_XCD - 1) / XDN_NUM_XCD;
    const int s0 = (int)((long)nstrip * g / XDN_NUM_XCD), s1 = (int)((long)nstrip * (g + 1) / XDN_NUM_XCD);
    return StripWalk{s0 + lb, nb, s1};
}

struct TileLoc {
    int f, tx, ty;
};

__device__ __forceinline__ TileLoc locate(const TileMap &m, int tile)
{
    TileLoc l;
    int f = 0;
    while (f + 1 < m.n && tile >= m.begin[f + 1]) ++f;
    const int local = tile - m.begin[f];
    l.f = f;
    l.ty = local / m.tiles_x[f];
    l.tx = local - l.ty * m.tiles_x[f];
    return l;
}

__device__ __forceinline__ float3 ld3(const float *p, size_t i)
{
    return make_float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

__device__ __forceinline__ void st3(float *p, size_t i, float3 v)
{
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

__device__ __forceinline__ float dist2(float3 a, float3 b)
{
    const float x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
    return x * x + y * y + z * z;
}

}  // namespace xdn

// Host-side helper to build a TileMap for frames of given sizes (in tiles).
namespace xdn {
inline TileMap make_tilemap(int n, const int *tiles_x, const int *tiles_y)
{
    TileMap m{};
    m.n = n;
    int acc = 0;
    for (int i = 0; i < n; ++i)
    {
        m.begin[i] = acc;
        m.tiles_x[i] = tiles_x[i];
        acc += tiles_x[i] * tiles_y[i];
    }
    m.begin[n] = acc;
    m.total = acc;
    return m;
}
}  // namespace xdn
