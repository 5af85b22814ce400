// gfx950 kernels of the soft-shadow visibility path (one frame).
//
// Stage map (reference -> kernel):
//   sampleMaxZcal / triangleMaxZcal / sampleRectangleAABBcal /
//   verticesRectangleAABBcal + AABBReduce_kernel  -> k_reduce (one pass)
//   lightPlaneParamCal host part                 -> k_finalize
//   writeSampleAndBinPair + maxZpreBinReduce      -> k_rx_key
//   thrust::sort_by_key (receivers) + FindStart    -> k_scan_* + k_rx_scatter
//   triangleVertexSetup + countPairNumForPerTri    -> k_tri_count
//   exclusive_scan + writeTriBinPair + sort_by_key -> k_scan_* + k_tri_write + k_tri_scatter_wide
//   findTriangleStartEnd                           -> k_tables
//   shadowCal_per_sample_kernel                    -> k_occlude
//
// Every kernel reads the frame status first, so the whole frame is one
// fixed launch sequence with no host round trip (graph-capturable).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fd_device.h"
#include "fd_kernels.h"

using namespace fdk;

namespace {

constexpr int RED_BLOCK = 256;
constexpr int RED_ITEMS = 8;           // receivers / vertices per thread in k_reduce
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_CHUNK = SCAN_BLOCK * SCAN_ITEMS;

// --------------------------------------------------------------------------
// k_init: reset the frame state (a kernel, not a memset, so that replays of
// a captured graph start from the same state).
// --------------------------------------------------------------------------
__device__ __forceinline__ void init_body(const FdArgs& a, const int blk) {
    if (threadIdx.x != 0 || blk != 0) return;
    FdState* s = a.st;
    const uint32_t encMax = fenc(FD_F32_MAX), encNegMax = fenc(-FD_F32_MAX);
    s->rxZmaxAll = encNegMax;
    s->rxMin[0] = s->rxMin[1] = s->rxMin[2] = encMax;
    s->rxMaxA[0] = s->rxMaxA[1] = encNegMax;
    s->rxMaxZ = encNegMax;
    s->vxZmax = encNegMax;
    s->vxMin[0] = s->vxMin[1] = s->vxMin[2] = encMax;
    s->vxMaxA[0] = s->vxMaxA[1] = encNegMax;
    s->vxMaxZ = encNegMax;
    s->validRect = 0;
    s->status = FD_EXIT_OK;
    s->validSamples = 0;
    s->validTris = 0;
    s->keyOutOfRange = 0;
    s->capacityHit = 0;
    s->pairs = 0;
    s->rxTotal = 0;
    s->totalChunks = 0;
    s->triOrder = FD_TRI_ORDER_LIST;
}

// Bin statistics reset (maxZ, minRange) for the bins of this frame.
__device__ __forceinline__ void init_bins_body(const FdArgs& a, const int blk) {
    int b = blk * blockDim.x + threadIdx.x;
    const int nb = 1 << (2 * a.gridLog2);
    if (b >= nb) return;
    const uint32_t encMax = fenc(FD_F32_MAX), encNegMax = fenc(-FD_F32_MAX);
    a.binStat[5 * b + 0] = encNegMax;  // maxZ
    a.binStat[5 * b + 1] = encNegMax;  // max of rect min x
    a.binStat[5 * b + 2] = encNegMax;  // max of rect min y
    a.binStat[5 * b + 3] = encMax;     // min of rect max x
    a.binStat[5 * b + 4] = encMax;     // min of rect max y
}

// Block-level reduction of NV uint32 values with op (0 = min, 1 = max,
// 2 = sum); the block's results go to part[k * P + blk] (no atomics: many
// blocks updating the few words of the frame state would serialise).
template <int NV>
__device__ void block_reduce_store(uint32_t (&v)[NV], const int (&op)[NV], uint32_t* part, int P, int blk) {
    __shared__ uint32_t sh[NV][RED_BLOCK / 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const uint32_t r = op[k] == 2 ? (uint32_t)wave_sum_i32((int)v[k]) : op[k] ? wave_max_u32(v[k]) : wave_min_u32(v[k]);
        if (l == 0) sh[k][w] = r;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int k = threadIdx.x;
        uint32_t r = sh[k][0];
        for (int i = 1; i < RED_BLOCK / 64; ++i)
            r = op[k] == 2 ? r + sh[k][i] : op[k] ? max(r, sh[k][i]) : min(r, sh[k][i]);
        part[(size_t)k * P + blk] = r;
    }
}

// --------------------------------------------------------------------------
// k_reduce: one pass over receivers (blocks [0, nbR)) and vertices (blocks
// [nbR, nbR + nbV)).  Receivers: z for the far plane
// (Step2_1_samplerRectangleAABB.inl:1-34), back-face test against the four
// light corners and the light-plane rectangle (:53-137).  The rectangle's max
// corner x' + F/z - S is reduced as max(x' + F/z) - S after the frame's S is
// known: rounding of (a - S) is monotone in a, so both give the same float.
// Vertices: Step2_2_modelRectangleAABB.inl:1-99.
// --------------------------------------------------------------------------
__device__ __forceinline__ void reduce_body(const FdArgs& a, const int blk) {
    const int nbR = a.nbR;
    const float* m = a.L.m;
    const uint32_t encMax = fenc(FD_F32_MAX), encNegMax = fenc(-FD_F32_MAX);
    if ((int)blk < nbR) {
        uint32_t v[7] = {encNegMax, encMax, encMax, encMax, encNegMax, encNegMax, encNegMax};
        int cnt = 0;
        const int base = blk * RED_BLOCK * RED_ITEMS + threadIdx.x;
#pragma unroll
        for (int it = 0; it < RED_ITEMS; ++it) {
            const int i = base + it * RED_BLOCK;
            if (i >= a.WH) break;
            const float4 p = a.pos[i];
            float4 out = make_float4(0.f, 0.f, 0.f, __uint_as_float(FD_TAG_NONE));
            if (!(p.x != p.x)) {
                const float z = getZ(m, p.x, p.y, p.z);
                if (!(z != z)) v[0] = max(v[0], fenc(z));
                const float4 n = a.nrm[i];
                const float* c[4] = {a.L.ur, a.L.ul, a.L.dl, a.L.dr};
                bool back = true;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    back = back && (dot3(c[k][0] - p.x, c[k][1] - p.y, c[k][2] - p.z, n.x, n.y, n.z) <= 0);
                if (back) {
                    out.w = __uint_as_float(FD_TAG_BACK);
                } else {
                    float rx, ry, rz;
                    makeRect(m, p.x, p.y, p.z, rx, ry, rz);
                    const float ax = rx + a.L.F[0] / rz;
                    const float ay = ry + a.L.F[1] / rz;
                    if (!(rx != rx)) v[1] = min(v[1], fenc(rx));
                    if (!(ry != ry)) v[2] = min(v[2], fenc(ry));
                    if (!(rz != rz)) { v[3] = min(v[3], fenc(rz)); v[6] = max(v[6], fenc(rz)); }
                    if (!(ax != ax)) v[4] = max(v[4], fenc(ax));
                    if (!(ay != ay)) v[5] = max(v[5], fenc(ay));
                    ++cnt;
                    out = make_float4(rx, ry, rz, __uint_as_float(FD_TAG_VALID));
                }
            }
            a.rec[i] = out;
        }
        uint32_t v8[8] = {v[0], v[1], v[2], v[3], v[4], v[5], v[6], (uint32_t)cnt};
        const int op[8] = {1, 0, 0, 0, 1, 1, 1, 2};
        block_reduce_store<8>(v8, op, a.part, nbR + a.nbV, blk);
    } else {
        uint32_t v[7] = {encNegMax, encMax, encMax, encMax, encNegMax, encNegMax, encNegMax};
        const int base = (blk - nbR) * RED_BLOCK * RED_ITEMS + threadIdx.x;
#pragma unroll
        for (int it = 0; it < RED_ITEMS; ++it) {
            const int vi = base + it * RED_BLOCK;
            if (vi >= a.nverts) break;
            const float* p = a.verts + (size_t)vi * a.vstride;
            const float px = p[0], py = p[1], pz = p[2];
            float rx, ry, rz;
            makeRect(m, px, py, pz, rx, ry, rz);
            const float ax = rx + a.L.F[0] / rz;
            const float ay = ry + a.L.F[1] / rz;
            if (!(rz != rz)) { v[0] = max(v[0], fenc(rz)); v[3] = min(v[3], fenc(rz)); v[6] = max(v[6], fenc(rz)); }
            if (!(rx != rx)) v[1] = min(v[1], fenc(rx));
            if (!(ry != ry)) v[2] = min(v[2], fenc(ry));
            if (!(ax != ax)) v[4] = max(v[4], fenc(ax));
            if (!(ay != ay)) v[5] = max(v[5], fenc(ay));
        }
        const int op[7] = {1, 0, 0, 0, 1, 1, 1};
        block_reduce_store<7>(v, op, a.part, nbR + a.nbV, blk);
    }
}

// --------------------------------------------------------------------------
// k_finalize: the host part of lightPlaneParamCal
// (Step2_3_defineFDbuffer.inl:5-111) on the device, with its early exits.
// --------------------------------------------------------------------------
__device__ __forceinline__ void finalize_body(const FdArgs& a, const int blk) {
    if (blk != 0) return;
    FdState* s = a.st;
    {
        // the k_reduce block partials: receiver blocks [0, nbR), vertex
        // blocks [nbR, nbR + nbV); min/max of order-preserving encodings
        __shared__ uint32_t red[2][8][FD_FIN_THREADS / 64];
        const int nbR = a.nbR, P = a.nbR + a.nbV;
        const int op[8] = {1, 0, 0, 0, 1, 1, 1, 2};
        const uint32_t encMax = fenc(FD_F32_MAX), encNegMax = fenc(-FD_F32_MAX);
        uint32_t r[2][8];
#pragma unroll
        for (int k = 0; k < 8; ++k) r[0][k] = r[1][k] = op[k] == 2 ? 0u : op[k] ? encNegMax : encMax;
        for (int i = threadIdx.x; i < P; i += blockDim.x) {
            const int g = i < nbR ? 0 : 1;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (g == 1 && k == 7) continue;
                const uint32_t x = a.part[(size_t)k * P + i];
                r[g][k] = op[k] == 2 ? r[g][k] + x : op[k] ? max(r[g][k], x) : min(r[g][k], x);
            }
        }
        const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t x = op[k] == 2 ? (uint32_t)wave_sum_i32((int)r[g][k])
                                              : op[k] ? wave_max_u32(r[g][k]) : wave_min_u32(r[g][k]);
                if (l == 0) red[g][k][w] = x;
            }
        __syncthreads();
        if (threadIdx.x != 0) return;
        const int nw = (int)(blockDim.x + 63) >> 6;
        uint32_t o[2][8];
        for (int g = 0; g < 2; ++g)
            for (int k = 0; k < 8; ++k) {
                uint32_t x = red[g][k][0];
                for (int j = 1; j < nw; ++j)
                    x = op[k] == 2 ? x + red[g][k][j] : op[k] ? max(x, red[g][k][j]) : min(x, red[g][k][j]);
                o[g][k] = x;
            }
        s->rxZmaxAll = o[0][0];
        s->rxMin[0] = o[0][1]; s->rxMin[1] = o[0][2]; s->rxMin[2] = o[0][3];
        s->rxMaxA[0] = o[0][4]; s->rxMaxA[1] = o[0][5];
        s->rxMaxZ = o[0][6];
        s->validRect = (int)o[0][7];
        s->vxZmax = o[1][0];
        s->vxMin[0] = o[1][1]; s->vxMin[1] = o[1][2]; s->vxMin[2] = o[1][3];
        s->vxMaxA[0] = o[1][4]; s->vxMaxA[1] = o[1][5];
        s->vxMaxZ = o[1][6];
    }
    // sampleMaxZcal / triangleMaxZcal: thrust::reduce(max) seeded with -1
    const float zs = fmax_(-1.0f, fdec(s->rxZmaxAll));
    const float zt = fmax_(-1.0f, fdec(s->vxZmax));
    s->sampleMaxZ = zs;
    s->triMaxZ = zt;
    if (zs < 0.0f) { s->status = FD_EXIT_SAMPLE_MAXZ; return; }
    if (zt < 0.0f) { s->status = FD_EXIT_TRI_MAXZ; return; }
    const float f = fmax_(zs, zt) + 0.1f;
    s->f = f;
    const float Sx = a.L.F[0] / f, Sy = a.L.F[1] / f;
    s->S[0] = Sx;
    s->S[1] = Sy;
    float sa[8] = {fdec(s->rxMin[0]), fdec(s->rxMin[1]), fdec(s->rxMin[2]), 0.0f,
                   fdec(s->rxMaxA[0]) - Sx, fdec(s->rxMaxA[1]) - Sy, fdec(s->rxMaxZ), 0.0f};
    float ma[8] = {fdec(s->vxMin[0]), fdec(s->vxMin[1]), fdec(s->vxMin[2]), 0.0f,
                   fdec(s->vxMaxA[0]) - Sx, fdec(s->vxMaxA[1]) - Sy, fdec(s->vxMaxZ), 0.0f};
    // the reductions started from +-FD_F32_MAX; (-MAX) - S rounds back to -MAX
    sa[4] = fmax_(sa[4], -FD_F32_MAX); sa[5] = fmax_(sa[5], -FD_F32_MAX);
    ma[4] = fmax_(ma[4], -FD_F32_MAX); ma[5] = fmax_(ma[5], -FD_F32_MAX);
    for (int k = 0; k < 8; ++k) { s->sAABB[k] = sa[k]; s->mAABB[k] = ma[k]; }
    if (s->validRect == 0) { s->status = FD_EXIT_NO_VALID_SAMPLE; return; }
    if (ma[0] > ma[4] || ma[1] > ma[5] || sa[0] > sa[4] || sa[1] > sa[5]) { s->status = FD_EXIT_EMPTY_AABB; return; }
    float B0 = fmax_(sa[0], ma[0]), B1 = fmax_(sa[1], ma[1]);
    float E0 = fmin_(sa[4], ma[4]), E1 = fmin_(sa[5], ma[5]);
    s->B[0] = B0; s->B[1] = B1; s->B[2] = ma[2];
    s->E[0] = E0; s->E[1] = E1; s->E[2] = sa[6];
    if (B0 >= E0 || B1 >= E1) { s->status = FD_EXIT_EMPTY_LP; return; }
    const int G = 1 << a.gridLog2;
    s->G = G;
    s->log2G = a.gridLog2;
    s->nb = G * G;
    s->factor[0] = (float)G / (E0 - B0);
    s->factor[1] = (float)G / (E1 - B1);
}

// --------------------------------------------------------------------------
// Receiver key (writeSampleAndBinPair, Step3_bindSampleToBin.inl:5-59) and
// the per-bin receiver statistics (maxZpreBinReduce_kernel, :61-104), which
// are min/max reductions and so are folded into the same pass with atomics.
// Each block owns RXB consecutive receivers and writes its key histogram.
// --------------------------------------------------------------------------
struct RxRect { float bx, by, ex, ey, z; };

__device__ __forceinline__ RxRect rx_rect(const float4 r, const FdState* s, const FdLight& L) {
    RxRect o;
    o.bx = r.x;
    o.by = r.y;
    o.ex = r.x + L.F[0] / r.z - s->S[0];
    o.ey = r.y + L.F[1] / r.z - s->S[1];
    o.z = r.z;
    return o;
}

// bin key of a valid receiver, or nb when outside the light plane
__device__ __forceinline__ int rx_key(const RxRect& q, const FdState* s, int nb, int log2G, bool& oor) {
    oor = false;
    const float B0 = s->B[0], B1 = s->B[1], E0 = s->E[0], E1 = s->E[1];
    if ((q.bx > q.ex || q.by > q.ey) || (q.bx < B0 || q.by < B1 || q.ex > E0 || q.ey > E1)) return nb;
    const float cx = (q.ex + q.bx) / 2.0f;
    const float cy = (q.ey + q.by) / 2.0f;
    const int bx = cvt_sat(floorf((cx - B0) * s->factor[0]));
    const int by = cvt_sat(floorf((cy - B1) * s->factor[1]));
    const int k = (int)(((unsigned)by << log2G) | (unsigned)bx);
    if (k < 0 || k >= nb) { oor = true; return nb; }
    return k;
}

// Receivers are traversed in pixel tiles of 64 (tile-major): TW x (64/TW)
// with TW = 1 << twl (8x8 by default), so that each 64-lane wave of the sort
// -- and later of the occlusion -- holds a compact tile rather than a
// 64-pixel strip.  The tile shape changes only the order of receivers within
// a bin, never a result.  Position p -> dense id, or -1 for padding outside
// the view.
__device__ __forceinline__ int tile_id(int p, int W, int H, int twl) {
    const int TW = 1 << twl;
    const int WT = (W + TW - 1) >> twl;
    const int t = p >> 6, j = p & 63;
    const int x = (t % WT) * TW + (j & (TW - 1));
    const int y = (t / WT) * (64 >> twl) + (j >> twl);
    return (x < W && y < H) ? y * W + x : -1;
}

__device__ __forceinline__ void rx_key_body(const FdArgs& a, const int blk) {
    extern __shared__ uint32_t hist[];  // nb + 1 counters, then (grids up to FD_LDS_BINS) 5 x nb bin statistics
    __shared__ uint32_t blkCount[2];    // valid receivers, keys out of range
    const FdState* s = a.st;
    if (s->status != FD_EXIT_OK) return;
    const int nb = s->nb, log2G = s->log2G;
    // bin statistics of the block in LDS, merged into the global ones once per
    // bin the block touched (else one set of global atomics per wave and key)
    const bool ldsStat = a.binLds && nb <= FD_LDS_BINS;
    uint32_t* bstat = hist + nb + 1;
    for (int k = threadIdx.x; k <= nb; k += blockDim.x) hist[k] = 0;
    if (ldsStat)
        for (int k = threadIdx.x; k < 5 * nb; k += blockDim.x) bstat[k] = (k % 5) < 3 ? 0u : 0xffffffffu;
    if (threadIdx.x < 2) blkCount[threadIdx.x] = 0;
    __syncthreads();
    int nvalid = 0, noor = 0;
    const int base = blk * FD_RXB;
    for (int it = 0; it < FD_RXB / 256; ++it) {
        const int p = base + it * 256 + threadIdx.x;
        const int i = p < a.npos ? tile_id(p, a.W, a.H, a.tileLog2) : -1;
        int key = nb;
        RxRect q{};
        const bool in = i >= 0;
        if (in) {
            const float4 r = a.rec[i];
            if (__float_as_uint(r.w) == FD_TAG_VALID) {
                q = rx_rect(r, s, a.L);
                bool oor;
                key = rx_key(q, s, nb, log2G, oor);
                if (key < nb || oor) ++nvalid;
                if (oor) ++noor;
            }
            a.rxKey[i] = (uint16_t)key;
        }
        // Segmented wave reduction: one pass per distinct key among the
        // wave's lanes (receivers of a wave are neighbours, so usually 1-3),
        // giving the key histogram and the per-bin receiver statistics with
        // one LDS add and five global atomics per (wave, key).
        const uint32_t v0 = (in && q.z == q.z) ? fenc(q.z) : 0u;
        const uint32_t v1 = (in && q.bx == q.bx) ? fenc(q.bx) : 0u;
        const uint32_t v2 = (in && q.by == q.by) ? fenc(q.by) : 0u;
        const uint32_t v3 = (in && q.ex == q.ex) ? fenc(q.ex) : 0xffffffffu;
        const uint32_t v4 = (in && q.ey == q.ey) ? fenc(q.ey) : 0xffffffffu;
        uint64_t pending = __ballot(in);
        while (pending) {
            const int lead = __ffsll((unsigned long long)pending) - 1;
            const int k0 = uniform(__shfl(key, lead, 64));
            const bool mine = in && key == k0;
            const uint64_t m = __ballot(mine);
            pending &= ~m;
            if (lane_id() == 0) atomicAdd(&hist[k0], (uint32_t)__popcll(m));
            if (k0 < nb) {
                const uint32_t r0 = wave_max_u32(mine ? v0 : 0u);
                const uint32_t r1 = wave_max_u32(mine ? v1 : 0u);
                const uint32_t r2 = wave_max_u32(mine ? v2 : 0u);
                const uint32_t r3 = wave_min_u32(mine ? v3 : 0xffffffffu);
                const uint32_t r4 = wave_min_u32(mine ? v4 : 0xffffffffu);
                if (lane_id() == 0) {
                    uint32_t* bs = (ldsStat ? bstat : a.binStat) + 5 * k0;
                    atomicMax(bs + 0, r0); atomicMax(bs + 1, r1); atomicMax(bs + 2, r2);
                    atomicMin(bs + 3, r3); atomicMin(bs + 4, r4);
                }
            }
        }
    }
    nvalid = wave_sum_i32(nvalid);
    noor = wave_sum_i32(noor);
    if (lane_id() == 0) {
        if (nvalid) atomicAdd(&blkCount[0], (uint32_t)nvalid);
        if (noor) atomicAdd(&blkCount[1], (uint32_t)noor);
    }
    __syncthreads();
    // valid receivers = binned ones (the scanned histogram) + out-of-range
    // keys, totalled in k_post_pairs; only the rare out-of-range count is
    // accumulated here
    if (threadIdx.x == 0 && blkCount[1]) atomicAdd(&a.st->keyOutOfRange, (int)blkCount[1]);
    for (int k = threadIdx.x; k <= nb; k += blockDim.x) {
        const uint32_t h = hist[k];
        a.histR[(size_t)k * a.nblkR + blk] = h;
        if (ldsStat && k < nb && h) {
            const uint32_t* l = bstat + 5 * k;
            uint32_t* bs = a.binStat + 5 * k;
            atomicMax(bs + 0, l[0]); atomicMax(bs + 1, l[1]); atomicMax(bs + 2, l[2]);
            atomicMin(bs + 3, l[3]); atomicMin(bs + 4, l[4]);
        }
    }
}

// --------------------------------------------------------------------------
// Device-wide exclusive scan of uint32 counts (thrust::exclusive_scan /
// the offsets of a stable counting sort) in two launches: chunk sums, then
// each chunk's scan offset by the sums before it.  Totals are kept in 64 bits.
// --------------------------------------------------------------------------
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t* tot) {
    __shared__ uint64_t wsum[SCAN_BLOCK / 64];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t y = __shfl_up(x, o, 64);
        if (l >= o) x += y;
    }
    if (l == 63) wsum[w] = x;
    __syncthreads();
    uint64_t off = 0, all = 0;
    for (int i = 0; i < SCAN_BLOCK / 64; ++i) {
        if (i < w) off += wsum[i];
        all += wsum[i];
    }
    __syncthreads();
    *tot = all;
    return off + x - v;
}

__device__ __forceinline__ void scan_reduce_body(const uint32_t* __restrict__ d, long n, uint64_t* __restrict__ part,
                                                 const int* status, const int blk) {
    if (status && *status != FD_EXIT_OK) return;
    const long base = (long)blk * SCAN_CHUNK + threadIdx.x * SCAN_ITEMS;
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) sum += d[base + k];
    uint64_t tot;
    block_excl_scan_u64(sum, &tot);
    if (threadIdx.x == 0) part[blk] = tot;
}

// Second pass: each chunk adds the sum of the chunk sums before it (read
// from part[], at most a few hundred values, by the whole block) to its own
// exclusive scan; block 0 also writes the grand total.
__device__ __forceinline__ void scan_apply_body(uint32_t* __restrict__ d, long n, const uint64_t* __restrict__ part,
                                                unsigned long long* total, const int* status, const int blk) {
    if (status && *status != FD_EXIT_OK) return;
    const int np = (int)((n + SCAN_CHUNK - 1) / SCAN_CHUNK);
    uint64_t pre = 0, all = 0;
    for (int i = threadIdx.x; i < np; i += SCAN_BLOCK) {
        const uint64_t x = part[i];
        if (i < blk) pre += x;
        all += x;
    }
    uint64_t preTot, allTot;
    block_excl_scan_u64(pre, &preTot);
    block_excl_scan_u64(all, &allTot);
    if (blk == 0 && threadIdx.x == 0 && total) *total = allTot;
    const long base = (long)blk * SCAN_CHUNK + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? d[base + k] : 0u;
        sum += v[k];
    }
    uint64_t tot;
    uint64_t off = block_excl_scan_u64(sum, &tot) + preTot;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) d[base + k] = (uint32_t)off;
        off += v[k];
    }
}

// --------------------------------------------------------------------------
// Stable receiver scatter: one wave per block walks its RXB receivers in
// order, 64 at a time; lanes holding equal keys are matched with ballots
// over the key bits and ranked with mbcnt, so the result equals a stable
// sort_by_key (Step3_bindSampleToBin.inl:125-127).  Binned receivers also
// get their 32-byte occlusion record written in sorted order.
// --------------------------------------------------------------------------
__device__ __forceinline__ void rx_scatter_body(const FdArgs& a, const int blk) {
    extern __shared__ int cnt[];  // nb + 1, -1 = not loaded yet
    const FdState* s = a.st;
    if (s->status != FD_EXIT_OK) return;
    const int nb = s->nb;
    int kbits = 1;
    while ((1 << kbits) <= nb) ++kbits;
    for (int k = threadIdx.x; k <= nb; k += 64) cnt[k] = -1;
    __syncthreads();
    const int base = blk * FD_RXB;
    for (int it = 0; it < FD_RXB / 64; ++it) {
        const int pp = base + it * 64 + lane_id();
        const int i = pp < a.npos ? tile_id(pp, a.W, a.H, a.tileLog2) : -1;
        const bool act = i >= 0;
        const int key = act ? (int)a.rxKey[i] : 0;
        uint64_t eq = __ballot(act);
        for (int b = 0; b < kbits; ++b) {
            const bool bit = (key >> b) & 1;
            const uint64_t bal = __ballot(act && bit);
            eq &= bit ? bal : ~bal;
        }
        if (act) {
            const int rank = mbcnt(eq);
            const int gcount = __popcll(eq);
            int b0 = cnt[key];
            if (b0 < 0) b0 = (int)a.histR[(size_t)key * a.nblkR + blk];
            const int p = b0 + rank;
            if (rank == gcount - 1) cnt[key] = b0 + gcount;
            if ((unsigned)p < (unsigned)a.WH) a.sortedId[p] = i;
            if (key < nb && (unsigned)p < (unsigned)a.WH) {
                const float4 r = a.rec[i];
                const RxRect q = rx_rect(r, s, a.L);
                const float4 ps = a.pos[i];
                FdRxRec o;
                o.q[0] = ps.x; o.q[1] = ps.y; o.q[2] = ps.z;
                o.rminx = q.bx; o.rminy = q.by; o.rmaxx = q.ex; o.rmaxy = q.ey; o.z = q.z;
                a.sortedRec[p] = o;
            }
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------------------
// Triangle setup (triangleVertexSetup_kernel, Step0_triangleSetup.inl:56-191)
// fused with the pair count (countPairNumForPerTri_kernel,
// Step4_RasterizeTriangleToBin.inl:5-113).  The bin loop walks the union of
// the wave's bin rectangles with a wave-uniform bin index, so the bin
// statistics are scalar loads and the block histogram is updated once per
// (wave, bin).
// --------------------------------------------------------------------------
struct BinStat { float maxZ, mx, my, mz, mw; };

__device__ __forceinline__ BinStat load_bin(const uint32_t* bs) {
    BinStat b;
    b.maxZ = fdec(bs[0]); b.mx = fdec(bs[1]); b.my = fdec(bs[2]); b.mz = fdec(bs[3]); b.mw = fdec(bs[4]);
    return b;
}

// the bin test of the count pass (:94), on the LP-clamped AABB
__device__ __forceinline__ bool bin_test(const BinStat& b, float cbx, float cby, float cbz, float cex, float cey) {
    return b.maxZ >= cbz && cbx < b.mx && cby < b.my && b.mz < cex && b.mw < cey;
}

// Bin statistics staged in LDS by the pair passes (when the grid has at most
// FD_LDS_BINS bins): float4 (maxZ, mx, my, mz) and mw per bin, decoded once
// per block instead of five L2 loads per (triangle, bin) test.
__device__ __forceinline__ void stage_bins(const FdArgs& a, int nb, float4* q, float* w) {
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        const uint32_t* bs = a.binStat + 5 * k;
        q[k] = make_float4(fdec(bs[0]), fdec(bs[1]), fdec(bs[2]), fdec(bs[3]));
        w[k] = fdec(bs[4]);
    }
}
__device__ __forceinline__ BinStat bin_at(const FdArgs& a, bool staged, const float4* q, const float* w, int bp) {
    if (staged) {
        const float4 v = q[bp];
        return BinStat{v.x, v.y, v.z, v.w, w[bp]};
    }
    return load_bin(a.binStat + 5 * bp);
}

struct TriClamp { float bx, by, bz, ex, ey; int sx, sy, ex_, ey_; bool ok; };

// the count pass's triangle rejection, LP clamp and bin range (:41-83)
__device__ __forceinline__ TriClamp tri_clamp(const float4 mn, const float4 mx, const FdState* s) {
    TriClamp c;
    const float B0 = s->B[0], B1 = s->B[1], E0 = s->E[0], E1 = s->E[1];
    c.ok = !(mn.x > mx.x || mn.y > mx.y || mn.z > mx.z || mn.x > E0 || mn.y > E1 || mx.x < B0 || mx.y < B1);
    c.bx = fmax_(mn.x, B0);
    c.by = fmax_(mn.y, B1);
    c.bz = mn.z;
    c.ex = fmin_(mx.x, E0);
    c.ey = fmin_(mx.y, E1);
    const float G = (float)s->G;
    c.sx = cvt_sat_u16(fmax_(floorf((c.bx - B0) * s->factor[0]), 0.0f));
    c.sy = cvt_sat_u16(fmax_(floorf((c.by - B1) * s->factor[1]), 0.0f));
    c.ex_ = cvt_sat_u16(fmin_(ceilf((c.ex - B0) * s->factor[0]), G));
    c.ey_ = cvt_sat_u16(fmin_(ceilf((c.ey - B1) * s->factor[1]), G));
    if (!c.ok) { c.sx = c.sy = 0; c.ex_ = c.ey_ = 0; }
    return c;
}

// Visits, in ascending bin order, every bin of the light-plane grid that the
// bin rectangle [sx, ex) x [sy, ey) of at least one lane covers, calling
// body(bp, i, j, mine) with a wave-uniform bin bp = j << log2G | i and
// mine = the lane's rectangle holds it.  Two schedules with the same visiting
// order (so the per-bin ballot ranking of the callers stays stable):
//  * the union rectangle of the wave, row by row (cheap per bin; best when
//    the lanes' rectangles overlap);
//  * "next covered bin": every lane walks its own rectangle and the wave
//    takes the smallest next bin of any lane (a wave minimum per bin; best
//    when the rectangles are scattered, where the union is mostly empty).
// The cheaper one is picked per wave from the union area and the summed
// rectangle areas.
template <typename F>
__device__ __forceinline__ void for_covered_bins(bool has, int sx, int sy, int ex, int ey, int log2G, F&& body) {
    const int ux0 = uniform(wave_min_i32(has ? sx : 1 << 20));
    const int uy0 = uniform(wave_min_i32(has ? sy : 1 << 20));
    const int ux1 = uniform(wave_max_i32(has ? ex : 0));
    const int uy1 = uniform(wave_max_i32(has ? ey : 0));
    if (ux1 <= ux0 || uy1 <= uy0) return;
    const int unionArea = (ux1 - ux0) * (uy1 - uy0);
    const int sumArea = uniform(wave_sum_i32(has ? (ex - sx) * (ey - sy) : 0));
    if (unionArea <= 2 * sumArea) {
        for (int j = uy0; j < uy1; ++j)
            for (int i = ux0; i < ux1; ++i)
                body(j << log2G | i, i, j, has && j >= sy && j < ey && i >= sx && i < ex);
        return;
    }
    constexpr int DONE = 1 << 30;
    int ci = sx, cj = sy;
    int cur = has ? (cj << log2G | ci) : DONE;
    for (;;) {
        const int bp = uniform(wave_min_i32(cur));
        if (bp == DONE) break;
        const bool mine = cur == bp;
        body(bp, bp & ((1 << log2G) - 1), bp >> log2G, mine);
        if (mine) {
            if (++ci == ex) { ci = sx; ++cj; }
            cur = cj < ey ? (cj << log2G | ci) : DONE;
        }
    }
}

// Pair passes: a lane walks its own bin rectangle when it holds at most
// FD_BIG_AREA bins; larger rectangles (a triangle spanning much of the light
// plane) are walked by the whole wave, one rectangle at a time, so that one
// lane's long walk does not hold up its wave and block.
#define FD_BIG_AREA 64

// The bins of a wave-uniform rectangle [sx, ex) x [sy, ey), 64 per step:
// body(valid, i, j) on every lane.  Steps follow raster order and, within a
// step, lanes do too (rows of up to 64 bins share a step when narrower).
template <typename F>
__device__ __forceinline__ void walk_rect_wave(int sx, int sy, int ex, int ey, F&& body) {
    const int w = ex - sx;
    const int lane = lane_id();
    if (w <= 64) {
        int wl = 0;
        while ((1 << wl) < w) ++wl;
        const int rows = 64 >> wl;
        const int di = lane & ((1 << wl) - 1), dj = lane >> wl;
        for (int j0 = sy; j0 < ey; j0 += rows) body(di < w && j0 + dj < ey, sx + di, j0 + dj);
    } else {
        for (int j = sy; j < ey; ++j)
            for (int i0 = sx; i0 < ex; i0 += 64) body(i0 + lane < ex, i0 + lane, j);
    }
}

// Area class of a triangle projected on the light plane (its three
// makeRectangle points): 7 - floor(log2(area in bin cells)) clamped to
// [0, FD_TRI_CLASSES - 1]; 0 holds the largest triangles (128 cells and up,
// infinite areas too), the last class those under 1/128 of a cell and the
// degenerate ones (area zero or NaN).  The area is
// |(x1 - x0)(y2 - y0) - (y1 - y0)(x2 - x0)| / 2 in float, in that order.
__device__ __forceinline__ int area_class(const float (&px)[3], const float (&py)[3], const FdState* s) {
    const float twice = fabsf((px[1] - px[0]) * (py[2] - py[0]) - (py[1] - py[0]) * (px[2] - px[0]));
    const float cells = ((twice * 0.5f) * s->factor[0]) * s->factor[1];
    if (!(cells > 0.0f)) return FD_TRI_CLASSES - 1;
    const int e = (int)(__float_as_uint(cells) >> 23) - 127;   // floor(log2), -127 for denormals
    return min(max(7 - e, 0), FD_TRI_CLASSES - 1);
}

// Mesh order: the largest area class that comes last.  A class is
// 7 - floor(log2(area in bin cells)) and a bin cell is 1/4^log2G of the
// light plane, so 1/512 of the plane is class 16 - 2 log2G (8 at 16x16
// bins); classes up to it go last.
__device__ __forceinline__ int mesh_big_class(int log2G) { return 16 - 2 * log2G; }

// Walk key of an area class: the class itself (area order, big < 0), or
// for the mesh order the classes above big first (keys 0..14-big, larger
// triangles first) and the classes up to big last (one key, 15 - big).
__device__ __forceinline__ int walk_key(int cls, int big) {
    if (big < 0) return cls;
    return cls > big ? cls - big - 1 : FD_TRI_CLASSES - 1 - big;
}

__device__ __forceinline__ void tri_count_body(const FdArgs& a, const int blk) {
    // nb counters, then the staged bin statistics (float4 + float per bin);
    // 16-byte aligned behind the block's static counters, for the float4s
    extern __shared__ __align__(16) uint32_t thist[];
    __shared__ uint32_t blkValid, blkPairs, blkOverlaps;
    const FdState* s = a.st;
    if (s->status != FD_EXIT_OK) return;
    const int nb = s->nb, log2G = s->log2G;
    const bool staged = a.binLds && nb <= FD_LDS_BINS;
    float4* binQ = (float4*)(thist + nb);
    float* binW = (float*)(binQ + nb);
    for (int k = threadIdx.x; k < nb; k += blockDim.x) thist[k] = 0;
    if (threadIdx.x == 0) blkValid = blkPairs = blkOverlaps = 0;
    if (staged) stage_bins(a, nb, binQ, binW);
    __syncthreads();
    const int t = blk * FD_TRB + threadIdx.x;
    const bool act = t < a.ntris;
    const float* m = a.L.m;
    float4 mn = make_float4(FD_F32_MAX, FD_F32_MAX, FD_F32_MAX, 1.0f);
    float4 mx = make_float4(-FD_F32_MAX, -FD_F32_MAX, -FD_F32_MAX, 1.0f);
    int valid = 0;
    if (act) {
        int i0 = a.idx[3 * t], i1 = a.idx[3 * t + 1], i2 = a.idx[3 * t + 2];
        // an index outside the vertex buffer makes the triangle unusable
        // (the reference reads out of bounds); it is treated as culled
        const bool badIdx = (unsigned)i0 >= (unsigned)a.nverts || (unsigned)i1 >= (unsigned)a.nverts ||
                            (unsigned)i2 >= (unsigned)a.nverts;
        if (badIdx) i0 = i1 = i2 = 0;
        const float* p0 = a.verts + (size_t)i0 * a.vstride;
        const float* p1 = a.verts + (size_t)i1 * a.vstride;
        const float* p2 = a.verts + (size_t)i2 * a.vstride;
        const float v0[3] = {p0[0], p0[1], p0[2]}, v1[3] = {p1[0], p1[1], p1[2]}, v2[3] = {p2[0], p2[1], p2[2]};
        // n = cross(v0 - v2, v1 - v0)   (:96-98)
        const float e0x = v1[0] - v0[0], e0y = v1[1] - v0[1], e0z = v1[2] - v0[2];
        const float e1x = v0[0] - v2[0], e1y = v0[1] - v2[1], e1z = v0[2] - v2[2];
        const float nx = e1y * e0z - e1z * e0y, ny = e1z * e0x - e1x * e0z, nz = e1x * e0y - e1y * e0x;
        bool culled = badIdx;
        if (a.cull && !badIdx) {
            const float* c[4] = {a.L.ur, a.L.ul, a.L.dr, a.L.dl};
            const float* vv[3] = {v0, v1, v2};
            culled = true;
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    culled = culled && (dot3(nx, ny, nz, c[k][0] - vv[j][0], c[k][1] - vv[j][1], c[k][2] - vv[j][2]) <= 0);
        }
        float px[3] = {0.0f, 0.0f, 0.0f}, py[3] = {0.0f, 0.0f, 0.0f};
        if (!culled) {
            const float* vv[3] = {v0, v1, v2};
#pragma unroll
            for (int k = 0; k < 3; ++k) {  // getTriAABB (:37-54)
                float rx, ry, rz;
                makeRect(m, vv[k][0], vv[k][1], vv[k][2], rx, ry, rz);
                px[k] = rx;
                py[k] = ry;
                mn.x = fmin_(mn.x, rx);
                mn.y = fmin_(mn.y, ry);
                mn.z = fmin_(mn.z, rz);
                mx.z = fmax_(mx.z, rz);
                mx.x = fmax_(mx.x, rx + a.L.F[0] / rz - s->S[0]);
                mx.y = fmax_(mx.y, ry + a.L.F[1] / rz - s->S[1]);
            }
            if (!(mn.x > s->E[0] || mn.y > s->E[1] || mx.x < s->B[0] || mx.y < s->B[1])) {
                valid = 1;
                FdTri d;
                for (int k = 0; k < 3; ++k) { d.v0[k] = v0[k]; d.v1[k] = v1[k]; d.v2[k] = v2[k]; }
                d.bminx = mn.x; d.bminy = mn.y; d.bmaxx = mx.x; d.bmaxy = mx.y;
                d.minz = mn.z; d.maxz = mx.z; d.pad0 = 0.0f;
                // occlusion-side triangle constants, same expressions as
                // Step5_2_shadowCalPerSample.inl:130 and :7
                const float ax = v1[0] - v0[0], ay = v1[1] - v0[1], az = v1[2] - v0[2];
                const float bx = v2[0] - v1[0], by = v2[1] - v1[1], bz = v2[2] - v1[2];
                d.n[0] = ay * bz - az * by; d.n[1] = az * bx - ax * bz; d.n[2] = ax * by - ay * bx;
                for (int k = 0; k < 3; ++k) {
                    d.e01[k] = v0[k] - v1[k];
                    d.e12[k] = v1[k] - v2[k];
                    d.e20[k] = v2[k] - v0[k];
                }
                d.pad1[0] = d.pad1[1] = d.pad1[2] = d.pad1[3] = 0.0f;
                a.tri[t] = d;
            }
        }
        a.triAABB[2 * t] = mn;
        a.triAABB[2 * t + 1] = mx;
        a.triClass[t] = (uint8_t)(valid ? area_class(px, py, s) : FD_TRI_CLASSES - 1);
    }
    {
        const int nv = wave_sum_i32(valid);
        if (lane_id() == 0 && nv) atomicAdd(&blkValid, (uint32_t)nv);
        // list coherence: how often a valid triangle's culling box overlaps
        // the next triangle's (the next lane's; the wave's last lane has none)
        const float nx0 = __shfl_down(mn.x, 1, 64), ny0 = __shfl_down(mn.y, 1, 64);
        const float nx1 = __shfl_down(mx.x, 1, 64), ny1 = __shfl_down(mx.y, 1, 64);
        const int nvld = __shfl_down(valid, 1, 64);
        const bool pair = valid && nvld && lane_id() < 63;
        const bool over = pair && mn.x <= nx1 && nx0 <= mx.x && mn.y <= ny1 && ny0 <= mx.y;
        const int np = __popcll(__ballot(pair)), no = __popcll(__ballot(over));
        if (lane_id() == 0 && np) {
            atomicAdd(&blkPairs, (uint32_t)np);
            atomicAdd(&blkOverlaps, (uint32_t)no);
        }
    }
    // count pass
    TriClamp c = tri_clamp(mn, mx, s);
    if (!act) { c.ok = false; c.sx = c.sy = c.ex_ = c.ey_ = 0; }
    int count = 0, tsx = 6000, tsy = 6000, tex = 0, tey = 0;
    const bool has = c.ok && c.sx < c.ex_ && c.sy < c.ey_;
    const bool big = has && (c.ex_ - c.sx) * (c.ey_ - c.sy) > FD_BIG_AREA;
    // small rectangles: every lane walks its own; per-block bin counts by
    // LDS atomics
    if (has && !big) {
        for (int j = c.sy; j < c.ey_; ++j)
            for (int i = c.sx; i < c.ex_; ++i) {
                const int bp = j << log2G | i;
                const BinStat b = bin_at(a, staged, binQ, binW, bp);
                if (bin_test(b, c.bx, c.by, c.bz, c.ex, c.ey)) {
                    ++count;
                    tsx = min(tsx, i); tsy = min(tsy, j);
                    tex = max(tex, i + 1); tey = max(tey, j + 1);
                    atomicAdd(&thist[bp], 1u);
                }
            }
    }
    // large rectangles: the wave walks them one by one
    for (uint64_t bigs = __ballot(big); bigs; bigs &= bigs - 1) {
        const int L = __ffsll((unsigned long long)bigs) - 1;
        const float bx = __shfl(c.bx, L, 64), by = __shfl(c.by, L, 64), bz = __shfl(c.bz, L, 64);
        const float ex = __shfl(c.ex, L, 64), ey = __shfl(c.ey, L, 64);
        const int sx = uniform(__shfl(c.sx, L, 64)), sy = uniform(__shfl(c.sy, L, 64));
        const int ex_ = uniform(__shfl(c.ex_, L, 64)), ey_ = uniform(__shfl(c.ey_, L, 64));
        int n = 0, mnx = 6000, mny = 6000, mxx = 0, mxy = 0;
        walk_rect_wave(sx, sy, ex_, ey_, [&](bool v, int i, int j) {
            const int bp = j << log2G | i;
            const bool pass = v && bin_test(bin_at(a, staged, binQ, binW, v ? bp : 0), bx, by, bz, ex, ey);
            if (pass) {
                atomicAdd(&thist[bp], 1u);
                mnx = min(mnx, i); mny = min(mny, j);
                mxx = max(mxx, i + 1); mxy = max(mxy, j + 1);
            }
            n += __popcll(__ballot(pass));
        });
        mnx = wave_min_i32(mnx); mny = wave_min_i32(mny);
        mxx = wave_max_i32(mxx); mxy = wave_max_i32(mxy);
        if (lane_id() == L) { count = n; tsx = mnx; tsy = mny; tex = mxx; tey = mxy; }
    }
    if (act) {
        a.triCount[t] = count;
        a.triOff[t] = (uint32_t)count;
        a.triRange[t] = make_uint2((uint32_t)tsx | ((uint32_t)tsy << 16), (uint32_t)tex | ((uint32_t)tey << 16));
    }
    __syncthreads();
    if (threadIdx.x == 0) {   // totalled in k_post_pairs
        a.part[blk] = blkValid;
        a.part[a.nblkT + blk] = blkPairs;
        a.part[2 * a.nblkT + blk] = blkOverlaps;
    }
    for (int k = threadIdx.x; k < nb; k += blockDim.x) a.histT[(size_t)k * a.nblkT + blk] = thist[k];
}

// The large rectangles of the write pass, walked by the whole wave; a pair's
// place in its triangle's run is its rank among the passing bins in raster
// order, as in the per-lane walk.
__device__ __forceinline__ void tri_write_big(const FdArgs& a, int log2G, bool staged, const float4* binQ,
                                              const float* binW, bool big, int sx, int sy, int ex_, int ey_,
                                              const TriClamp& c, unsigned long long o0) {
    const unsigned long long cap = a.pairCap;
    for (uint64_t bigs = __ballot(big); bigs; bigs &= bigs - 1) {
        const int L = __ffsll((unsigned long long)bigs) - 1;
        const float bx = __shfl(c.bx, L, 64), by = __shfl(c.by, L, 64), bz = __shfl(c.bz, L, 64);
        const float ex = __shfl(c.ex, L, 64), ey = __shfl(c.ey, L, 64);
        const int rx0 = uniform(__shfl(sx, L, 64)), ry0 = uniform(__shfl(sy, L, 64));
        const int rx1 = uniform(__shfl(ex_, L, 64)), ry1 = uniform(__shfl(ey_, L, 64));
        unsigned long long o = __shfl(o0, L, 64);
        const uint32_t tag = (uint32_t)((threadIdx.x & ~63u) + (unsigned)L) << 16;
        walk_rect_wave(rx0, ry0, rx1, ry1, [&](bool v, int i, int j) {
            const int bp = j << log2G | i;
            const bool pass = v && bin_test(bin_at(a, staged, binQ, binW, v ? bp : 0), bx, by, bz, ex, ey);
            const uint64_t m = __ballot(pass);
            if (pass) {
                const unsigned long long d = o + (unsigned)mbcnt(m);
                if (d < cap) a.pairKT[d] = tag | (uint32_t)bp;
            }
            o += (unsigned)__popcll(m);
        });
    }
}

// Stable pair scatter (writeTriBinPair_kernel + sort_by_key,
// Step4_RasterizeTriangleToBin.inl:161-223): per bin, triangles ascending.
// Pair write: every triangle writes its (bin) pairs at its triangle-major
// offset (scanned pair counts), bins ascending within the triangle -- the
// order in which the reference's writeTriBinPair_kernel emits them
// (Step4_RasterizeTriangleToBin.inl:161-204), with the count pass's test.
__device__ __forceinline__ void tri_write_body(const FdArgs& a, const int blk) {
    extern __shared__ uint32_t lds[];   // staged bin statistics
    const FdState* s = a.st;
    if (s->status != FD_EXIT_OK) return;
    const int nb = s->nb, log2G = s->log2G;
    const bool staged = a.binLds && nb <= FD_LDS_BINS;
    float4* binQ = (float4*)lds;
    float* binW = (float*)(binQ + nb);
    if (staged) {
        stage_bins(a, nb, binQ, binW);
        __syncthreads();
    }
    const int t = blk * FD_TRB + threadIdx.x;
    if (uniform(blk * FD_TRB + (int)(threadIdx.x & ~63u)) >= a.ntris) return;   // whole wave past the end
    // the walk over the pair range (bins with pairs) of the count pass
    bool has = t < a.ntris;
    int sx = 0, sy = 0, ex = 0, ey = 0;
    TriClamp c{};
    unsigned long long o = 0;
    if (has) {
        const uint2 rg = a.triRange[t];
        sx = rg.x & 0xffff; sy = rg.x >> 16; ex = rg.y & 0xffff; ey = rg.y >> 16;
        has = sx < ex && sy < ey;
        if (has) {
            c = tri_clamp(a.triAABB[2 * t], a.triAABB[2 * t + 1], s);
            has = c.ok;
            o = a.triOff[t];
        }
    }
    const bool big = has && (ex - sx) * (ey - sy) > FD_BIG_AREA;
    if (has && !big) {
        const unsigned long long cap = a.pairCap;
        const uint32_t tag = (uint32_t)threadIdx.x << 16;
        for (int j = sy; j < ey; ++j)
            for (int i = sx; i < ex; ++i) {
                const int bp = j << log2G | i;
                if (bin_test(bin_at(a, staged, binQ, binW, bp), c.bx, c.by, c.bz, c.ex, c.ey)) {
                    if (o < cap) a.pairKT[o] = tag | (uint32_t)bp;
                    ++o;
                }
            }
    }
    tri_write_big(a, log2G, staged, binQ, binW, big, sx, sy, ex, ey, c, o);
}

// Stable counting scatter of one triangle block's pairs (the reference's
// stable sort_by_key by bin, Step4_RasterizeTriangleToBin.inl:206-223), by
// the block's 16 waves at once: wave w owns triangles t0 + 64w .. t0 + 64w + 63 and with them a contiguous slice
// of the triangle-major pair list.  A triangle has at most one pair per bin,
// so a wave holds at most 64 pairs of a bin and the 16 per-wave counts of a
// bin fit 16-bit halves of LDS words (two waves per word, counted with 32-bit
// LDS adds).  After a barrier every bin's counts become exclusive prefixes
// over the waves (at most 1024, still 16 bits), and each wave ranks its
// slice from base + prefix: lanes with equal bins are matched by ballots
// over the key bits and ranked with mbcnt, the last lane of a group advances
// the running offset.  Result: bin-major, ascending triangle id within a
// bin.  LDS: FD_SCATTER_WIDE_LDS bytes per bin (bins <= 4096: grids up to
// 64 x 64).
__device__ __forceinline__ int wide_half(int w, int k, int nb) { return (((w >> 1) * nb + k) << 1) | (w & 1); }

__device__ __forceinline__ void tri_scatter_wide_body(const FdArgs& a, const int blk) {
    extern __shared__ uint32_t wlds[];
    const FdState* s = a.st;
    if (s->status != FD_EXIT_OK) return;
    const int nb = s->nb;
    uint32_t* cnt32 = wlds;                                   // 8 * nb words
    uint16_t* cnt16 = (uint16_t*)wlds;
    uint32_t* base = wlds + 8 * nb;                           // nb block offsets
    int kbits = 1;
    while ((1 << kbits) < nb) ++kbits;
    for (int k = threadIdx.x; k < 8 * nb; k += FD_TRB) cnt32[k] = 0u;
    for (int k = threadIdx.x; k < nb; k += FD_TRB) base[k] = a.histT[(size_t)k * a.nblkT + blk];
    __syncthreads();
    const int w = (int)(threadIdx.x >> 6);
    const int ta = blk * FD_TRB + w * 64;
    const int tb = min(ta + 64, a.ntris);
    unsigned long long p0 = 0, p1 = 0;
    if (ta < a.ntris) {
        p0 = a.triOff[ta];
        p1 = tb < a.ntris ? (unsigned long long)a.triOff[tb] : s->pairs;
    }
    const unsigned long long cap = a.pairCap;
    const unsigned long long e = p1 < cap ? p1 : cap;
    const uint32_t inc = (w & 1) ? 0x10000u : 1u;
    for (unsigned long long p = p0 + lane_id(); p < e; p += 64)
        atomicAdd(&cnt32[(w >> 1) * nb + (int)(a.pairKT[p] & 0xffffu)], inc);
    __syncthreads();
    for (int k = threadIdx.x; k < nb; k += FD_TRB) {
        uint32_t run = 0;
        for (int v = 0; v < 16; ++v) {
            const int h = wide_half(v, k, nb);
            const uint32_t c = cnt16[h];
            cnt16[h] = (uint16_t)run;
            run += c;
        }
    }
    __syncthreads();
    const int t0 = blk * FD_TRB;
    for (unsigned long long c0 = p0; c0 < p1; c0 += 64) {
        const unsigned long long p = c0 + lane_id();
        const bool act = p < e;
        const uint32_t v = act ? a.pairKT[p] : 0u;
        const int key = (int)(v & 0xffffu);
        uint64_t eq = __ballot(act);
        for (int b = 0; b < kbits; ++b) {
            const bool bit = (key >> b) & 1;
            const uint64_t bal = __ballot(act && bit);
            eq &= bit ? bal : ~bal;
        }
        if (act) {
            const int rank = mbcnt(eq);
            const int gcount = __popcll(eq);
            const int h = wide_half(w, key, nb);
            const uint32_t b0 = cnt16[h];
            const unsigned long long dst = (unsigned long long)base[key] + b0 + (unsigned)rank;
            if (rank == gcount - 1) cnt16[h] = (uint16_t)(b0 + gcount);
            if (dst < cap) a.sortedTri[dst] = t0 + (int)(v >> 16);
        }
    }
}

// --------------------------------------------------------------------------
// Per-bin tables (FindStart_kernel, binRaster.inl:5-59) and the occlusion
// work list: one work item per 64 sorted receivers of a bin (the invalid key
// included, so every receiver's output is written exactly once).  Work items
// are issued bin by bin, longest triangle list first (by power-of-two class
// of the list length; the invalid key, with no list, last), so that the
// launch does not end on a few long-running waves.  The order only schedules
// work: every receiver's result is the same in any order.
// --------------------------------------------------------------------------
__device__ __forceinline__ void tables_body(const FdArgs& a, const int blk) {
    __shared__ int wsum[16];
    __shared__ int classCount[33], classBase[33];
    FdState* s = a.st;
    const int status = s->status;
    // the receivers' start / end table exists once they are sorted: FDLaunch
    // binds the samples before it pairs the triangles (FD/api.cpp:505-522),
    // so a frame that ends without pairs (or over the pair arena) has it too
    if (status != FD_EXIT_OK && status != FD_EXIT_NO_PAIRS && status != FD_EXIT_CAPACITY) return;
    const int nb = s->nb;
    if (status != FD_EXIT_OK) {
        for (int k = threadIdx.x; k <= nb; k += blockDim.x) {
            a.rxStart[k] = (int)a.histR[(size_t)k * a.nblkR];
            a.rxEnd[k] = (k < nb) ? (int)a.histR[(size_t)(k + 1) * a.nblkR] : a.WH;
        }
        return;
    }
    if (threadIdx.x < 33) classCount[threadIdx.x] = 0;
    __syncthreads();
    // bin tables and the cost class of every bin (0 = longest lists)
    for (int k = threadIdx.x; k <= nb; k += blockDim.x) {
        const int rs = (int)a.histR[(size_t)k * a.nblkR];
        const int re = (k < nb) ? (int)a.histR[(size_t)(k + 1) * a.nblkR] : a.WH;
        a.rxStart[k] = rs;
        a.rxEnd[k] = re;
        int cls = 32;
        if (k < nb) {
            const int ts = (int)a.histT[(size_t)k * a.nblkT];
            const int te = (k < nb - 1) ? (int)a.histT[(size_t)(k + 1) * a.nblkT] : (int)s->pairs;
            a.triStart[k] = ts;
            a.triEnd[k] = te;
            cls = __clz(te - ts + 1);   // 31 - floor(log2(len + 1)): longer lists, smaller class
        }
        atomicAdd(&classCount[cls], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int c = 0;
        for (int i = 0; i < 33; ++i) { classBase[i] = c; c += classCount[i]; }
    }
    __syncthreads();
    // stable within a class (ascending bin index, so that neighbouring bins,
    // which share triangles, stay close in the issue order): ranks from wave
    // ballots, then a prefix over the waves of each 1024-bin slice
    __shared__ int waveCount[16][33];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k0 = 0; k0 <= nb; k0 += 1024) {
        const int k = k0 + threadIdx.x;
        const bool in = k <= nb;
        const int cls = !in ? 0 : (k < nb) ? __clz(a.triEnd[k] - a.triStart[k] + 1) : 32;
        for (int i = lane; i < 33; i += 64) waveCount[w][i] = 0;
        int rankInWave = 0;
        for (uint64_t pending = __ballot(in); pending;) {
            const int lead = __ffsll((unsigned long long)pending) - 1;
            const int c = __shfl(cls, lead, 64);
            const uint64_t m = __ballot(in && cls == c);
            if (in && cls == c) rankInWave = mbcnt(m);
            if (lane == lead) waveCount[w][c] = __popcll(m);
            pending &= ~m;
        }
        __syncthreads();
        if (threadIdx.x < 33) {   // per class: exclusive prefix over the waves, then advance the class base
            int run = classBase[threadIdx.x];
            for (int j = 0; j < 16; ++j) {
                const int c = waveCount[j][threadIdx.x];
                waveCount[j][threadIdx.x] = run;
                run += c;
            }
            classBase[threadIdx.x] = run;
        }
        __syncthreads();
        if (in) a.binAtRank[waveCount[w][cls] + rankInWave] = k;
        __syncthreads();
    }
    // work-item prefix in rank order
    int carry = 0;
    for (int k0 = 0; k0 <= nb; k0 += 1024) {
        const int r = k0 + threadIdx.x;
        int nch = 0;
        if (r <= nb) {
            const int k = a.binAtRank[r];
            nch = (a.rxEnd[k] - a.rxStart[k] + 63) >> 6;
        }
        // block exclusive scan of nch
        int x = nch;
        const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(x, o, 64);
            if (l >= o) x += y;
        }
        if (l == 63) wsum[w] = x;
        __syncthreads();
        int off = carry, all = 0;
        for (int i = 0; i < 16; ++i) {
            if (i < w) off += wsum[i];
            all += wsum[i];
        }
        if (r <= nb) a.chunkStart[r] = off + x - nch;
        carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.chunkStart[nb + 1] = carry;
        s->totalChunks = carry;
    }
}

// Area order of one bin's triangle list (frames whose walk order is
// FD_TRI_ORDER_AREA): a stable counting sort of triStart..triEnd of the
// reference-order list by the triangles' area classes, into the same range
// of pairKT (free once the pairs are scattered).  One block of
// FD_ORDER_THREADS per bin.  A lane finds the lanes of its wave with its own key from one ballot per key
// bit (the AND of the matching ballots), its rank within them by mbcnt; the
// first such lane adds their number.  Pass 1 totals the keys (and keeps
// them in LDS for lists of up to FD_ORDER_CACHE entries), pass 2 ranks every
// entry and offsets it by an exclusive prefix over the waves of each slice.
__device__ __forceinline__ void tri_order_body(const FdArgs& a, const int bin) {
    constexpr int NWV = FD_ORDER_THREADS / 64, K = FD_TRI_CLASSES;
    __shared__ int waveCount[NWV][K];
    __shared__ int keyBase[K];
    __shared__ uint8_t keyCache[FD_ORDER_CACHE];
    const FdState* s = a.st;
    if (s->status != FD_EXIT_OK || s->triOrder == FD_TRI_ORDER_LIST || bin >= s->nb) return;
    const int ts = a.triStart[bin], te = a.triEnd[bin];
    if (te <= ts) return;
    const int lane = lane_id(), w = threadIdx.x >> 6;
    int* out = (int*)a.pairKT;
    const int big = s->triOrder == FD_TRI_ORDER_MESH ? mesh_big_class(s->log2G) : -1;
    auto key_of = [&](int i, int& tri) -> int {
        tri = a.sortedTri[i];
        const int cls = ((unsigned)tri < (unsigned)a.ntris) ? (int)a.triClass[tri] : K - 1;
        return walk_key(cls, big);
    };
    // lanes of the wave with the same key (key K: no entry)
    auto same_key = [](int key) -> uint64_t {
        uint64_t m = ~0ull;
#pragma unroll
        for (int j = 0; j < FD_WALK_KEY_BITS; ++j) {
            const uint64_t b = __ballot((key >> j) & 1);
            m &= ((key >> j) & 1) ? b : ~b;
        }
        return m;
    };
    for (int k = threadIdx.x; k < K; k += FD_ORDER_THREADS) keyBase[k] = 0;
    __syncthreads();
    const bool cached = te - ts <= FD_ORDER_CACHE;
    // pass 1: key totals
    for (int i0 = ts; i0 < te; i0 += FD_ORDER_THREADS) {
        const int i = i0 + threadIdx.x;
        int tri = 0;
        const int key = i < te ? key_of(i, tri) : K;
        if (cached && i < te) keyCache[i - ts] = (uint8_t)key;
        const uint64_t m = same_key(key);
        if (key < K && mbcnt(m) == 0) atomicAdd(&keyBase[key], __popcll(m));
    }
    __syncthreads();
    if (threadIdx.x == 0) {   // exclusive prefix over the keys
        int run = 0;
        for (int k = 0; k < K; ++k) {
            const int x = keyBase[k];
            keyBase[k] = run;
            run += x;
        }
    }
    __syncthreads();
    // pass 2: stable scatter, one slice of FD_ORDER_THREADS entries at a time
    for (int i0 = ts; i0 < te; i0 += FD_ORDER_THREADS) {
        const int i = i0 + threadIdx.x;
        int tri = 0, key = K;
        if (i < te) {
            if (cached) {
                tri = a.sortedTri[i];
                key = keyCache[i - ts];
            } else {
                key = key_of(i, tri);
            }
        }
        for (int k = lane; k < K; k += 64) waveCount[w][k] = 0;
        const uint64_t m = same_key(key);
        const int rank = mbcnt(m);
        if (key < K && rank == 0) waveCount[w][key] = __popcll(m);
        __syncthreads();
        for (int k = threadIdx.x; k < K; k += FD_ORDER_THREADS) {   // per key: prefix over the waves
            int run = keyBase[k];
            for (int j = 0; j < NWV; ++j) {
                const int x = waveCount[j][k];
                waveCount[j][k] = run;
                run += x;
            }
            keyBase[k] = run;
        }
        __syncthreads();
        if (key < K) out[ts + waveCount[w][key] + rank] = tri;
        __syncthreads();
    }
}

// Exit bookkeeping after the scans (setupTriangleVertex returns 1 when no
// triangle reached the light plane, Step0_triangleSetup.inl:224-225;
// FDLaunch skips the rest when there are no pairs, FD/api.cpp:522).
__device__ __forceinline__ void post_pairs_body(const FdArgs& a, const int blk) {
    if (blk != 0) return;
    FdState* s = a.st;
    if (s->status != FD_EXIT_OK) return;
    // valid triangles: the pair-count blocks' totals (no count pass without
    // triangles)
    __shared__ int wsum[3][FD_FIN_THREADS / 64];
    int v = 0, np = 0, no = 0;
    if (a.ntris > 0)
        for (int i = threadIdx.x; i < a.nblkT; i += blockDim.x) {
            v += (int)a.part[i];
            np += (int)a.part[a.nblkT + i];
            no += (int)a.part[2 * a.nblkT + i];
        }
    v = wave_sum_i32(v);
    np = wave_sum_i32(np);
    no = wave_sum_i32(no);
    if ((threadIdx.x & 63) == 0) {
        wsum[0][threadIdx.x >> 6] = v;
        wsum[1][threadIdx.x >> 6] = np;
        wsum[2][threadIdx.x >> 6] = no;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    int nv = 0;
    long pairsAdj = 0, overlaps = 0;
    for (int j = 0; j < (int)(blockDim.x + 63) >> 6; ++j) {
        nv += wsum[0][j];
        pairsAdj += wsum[1][j];
        overlaps += wsum[2][j];
    }
    s->validTris = nv;
    // walk order of the bins' lists.  A triangle soup (neighbouring
    // triangles mostly apart): largest projected triangles first, big
    // occluders fill receivers early.  A connected mesh: the mesh order,
    // largest first except those covering 1/512 of the light plane or more,
    // which come last in list order (mostly triangles near the light; first,
    // they would scatter each object's triangles, which a receiver often
    // needs together) -- unless the lists are long (many objects overlap:
    // a receiver fills from whichever object comes first in the list, so
    // the list order, which keeps every object together, wins)
    const bool mesh = 2 * overlaps >= pairsAdj;
    const bool dense = s->pairs > (unsigned long long)FD_MESH_MAX_LIST16 * 256ull;
    s->triOrder = a.triOrder != FD_TRI_ORDER_AUTO ? a.triOrder
                  : !mesh ? FD_TRI_ORDER_AREA : dense ? FD_TRI_ORDER_LIST : FD_TRI_ORDER_MESH;
    // valid receivers: binned (keys < nb, the scanned histogram at key nb)
    // plus keys out of range
    s->validSamples = (int)a.histR[(size_t)s->nb * a.nblkR] + s->keyOutOfRange;
    if (s->validTris == 0) s->status = FD_EXIT_NO_VALID_TRI;
    else if (s->pairs == 0) s->status = FD_EXIT_NO_PAIRS;
    else if (s->pairs > a.pairCap || s->pairs >= 0x7fffffffull) { s->status = FD_EXIT_CAPACITY; s->capacityHit = 1; }
}

// --------------------------------------------------------------------------
// Occlusion (shadowCal_per_sample_kernel, Step5_2_shadowCalPerSample.inl:37-234).
// A work item is 64 sorted receivers of one bin, one receiver per lane, run
// by several waves that take interleaved triangles of the bin's list.  Each
// wave walks its triangles with a wave-uniform index, so triangle records
// arrive through scalar loads; the receivers' light rows live in LDS
// (row-major by row, lane-minor: conflict-free), shared by the waves.
// --------------------------------------------------------------------------
template <typename T>
struct Bits;
template <>
struct Bits<uint32_t> {
    static constexpr int width = 32;
    __device__ static int ctz(uint32_t v) { return __builtin_ctz(v); }
    __device__ static int popc(uint32_t v) { return __popc(v); }
};
template <>
struct Bits<uint64_t> {
    static constexpr int width = 64;
    __device__ static int ctz(uint64_t v) { return __builtin_ctzll(v); }
    __device__ static int popc(uint64_t v) { return __popcll(v); }
};

// getRowMaskForPlaneDiscriminant (:14-33): the mask of light samples along
// one row (or of rows) whose plane value D + B*c is non-negative, as
//   f1 = D >= 0, f2 = B*n + D >= 0;  f1 == f2 -> all / none,
//   else k = (int)ceil(-1*D/B), mask = (1 << k) - 1, complemented if !f1.
// Only min(k, R) is observable once the mask is clipped to the R-sample
// grid, and k outside [0, width) (shift to zero, then -1) is all ones, so k
// is reduced to kk in [0, R] with every out-of-grid value mapped to R.
//
// The division is the hot operation.  q' = -D * rcp(B) (v_rcp_f32, 1 ulp)
// is within |q| * 2^-22 of the correctly rounded quotient q, i.e. within
// 2e-5 for q <= 80.  Whenever q' is at least 2^-12 away from every integer,
// ceil(q') == ceil(q); q' >= R + 1 implies q > R.  Only the remaining cases
// (near an integer, near zero, negative, or |B| outside [2^-100, 2^100])
// take the IEEE division.
__device__ __forceinline__ int k_exact(float D, float B, int R) {
    const float c = ceilf(-1.0f * D / B);
    if (c != c) return 0;                       // cvt.rzi of NaN
    if (c < 0.0f || c > (float)R) return R;     // shift out of range / beyond the grid
    return (int)c;
}

// One edge plane A*row + B*col + C over the R x R sample grid, negated when
// the receiver sees the triangle's back (see the set-up in occlude_body).
struct Edge {
    float A, B, C;
    float nRowN;      // -(A * (R-1)):  f2 of the row-range mask is DN >= nRowN
    float nColN;      // -(B * (R-1)):  f2 of a row mask is D >= nColN
    float rA, rB;     // rcp(A), rcp(B); NaN when |A| / |B| is outside [2^-100, 2^100]
};

__device__ __forceinline__ bool rcp_ok(float x) {
    const float ax = fabsf(x);
    return ax >= 7.8886091e-31f && ax <= 1.2676506e+30f;
}
__device__ __forceinline__ float rcp_or_nan(float x) {
    return rcp_ok(x) ? __builtin_amdgcn_rcpf(x) : __builtin_nanf("");
}

// kk for one (D, divisor) in the division branch; sets slow when the exact
// quotient is needed.  With |divisor| in [2^-100, 2^100] the reciprocal
// quotient q is within 2e-5 of the exact one for q <= R + 1, and the exact
// quotient is >= 0 in this branch, so q > -2e-5: anything within 2^-12 of an
// integer (0 included) goes to the exact path.  q > R + 1 means all columns.
// A NaN reciprocal (divisor out of range) makes q NaN: always exact.
__device__ __forceinline__ int k_fast(float D, float r, float Rf, bool& slow) {
    const float q = (-D) * r;
    const float dist = fabsf(q - rintf(q));
    const bool fine = dist > 2.44140625e-4f || q > Rf + 1.0f;
    slow = !fine;
    return (int)__builtin_amdgcn_fmed3f(ceilf(q), 0.0f, Rf);   // NaN -> 0, then in [0, R]
}

template <typename T>
__device__ __forceinline__ T low_mask(int kk) {
    return (kk >= Bits<T>::width) ? ~(T)0 : (((T)1 << kk) - (T)1);
}

// AND of the three edge masks at plane offsets D[k] (row masks: divisor B,
// row-range masks: divisor A).  Each edge mask is all (f1 && f2), none
// (!f1 && !f2), a prefix [0, kk) (f1 && !f2) or a suffix [kk, R) (!f1 && f2);
// with v = (f1 != f2) ? kk : R the AND is the interval
// [max over !f1 edges of v, min over f1 edges of v).
// General form: valid for every input, including infinite limits B*(R-1).
template <typename T, bool ROWS>
__device__ __forceinline__ T three_edges(const Edge (&e)[3], const float (&D)[3], float Rf, int R) {
    bool f1[3], f2[3], slow[3];
    int kk[3];
    bool anySlow = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        f1[k] = D[k] >= 0.0f;
        f2[k] = (D[k] - (ROWS ? e[k].nRowN : e[k].nColN)) >= 0.0f;   // fl(B*n + D) >= 0
        kk[k] = k_fast(D[k], ROWS ? e[k].rA : e[k].rB, Rf, slow[k]);
        slow[k] = slow[k] && (f1[k] != f2[k]);
        anySlow |= slow[k];
    }
    if (anySlow) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (slow[k]) kk[k] = k_exact(D[k], ROWS ? e[k].A : e[k].B, R);
    }
    int v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = (f1[k] != f2[k]) ? kk[k] : R;
    const int hi = min(min(f1[0] ? v[0] : R, f1[1] ? v[1] : R), f1[2] ? v[2] : R);
    const int lo = max(max(f1[0] ? 0 : v[0], f1[1] ? 0 : v[1]), f1[2] ? 0 : v[2]);
    return low_mask<T>(hi) & ~low_mask<T>(lo);
}

// Bits [lo, lo + w) of a width-bit word, w + lo <= width.
template <typename T>
__device__ __forceinline__ T interval_bits(unsigned w, unsigned lo) {
    if (Bits<T>::width == 32) {
        unsigned m;
        asm("v_bfm_b32 %0, %1, %2" : "=v"(m) : "v"(w), "v"(lo));   // ((1 << w[4:0]) - 1) << lo[4:0]
        return (w >= 32u) ? ~(T)0 : (T)m;
    }
    return (w >= 64u) ? ~(T)0 : (T)((((uint64_t)1 << w) - 1ull) << (lo & 63u));
}

// Integer max3 / min3 as single instructions (kept opaque so that the
// selects feeding them are not folded into longer select chains).
__device__ __forceinline__ int max3i(int a, int b, int c) {
    int r;
    asm("v_max3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ int min3i(int a, int b, int c) {
    int r;
    asm("v_min3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Fast form of three_edges for limits nN = -(divisor * (R-1)) that are finite
// in every lane of the wave (the caller checks |A|, |B| <= 2^100).
//  * f2: for finite a and b, fl(a + b) >= 0 exactly when a + b >= 0 (a sum
//    of two floats is 0 or at least one denormal step away from it), so
//    fl(B*n + D) >= 0  <=>  D >= -fl(B*n)  -- also for infinite D.
//  * kk: only read when f1 != f2.  Then the exact quotient q lies in
//    [0, n + ulp] (f1 && !f2: D >= 0 > B; !f1 && f2: D < 0 < B, both with
//    |D| <= |B*n|).  X = D * rcp(B) + R is within |q| 2^-22 + 2^-19 < 2^-15
//    of Y = R - q.  The kernel forms x = fma(D, rcp(B), R + 2^-12) (Rs is
//    R + 2^-12, exact), which rounds X + 2^-12 by at most 2^-18.  When
//    fract(x) >= 2^-11, with k = floor(x):  X lies in
//    [k + 2^-12 - 2^-18, k + 1 - 2^-12 + 2^-18), so Y lies strictly inside
//    (k, k + 1) and kk = ceil(q) = R - floor(Y) = R - k, in [1, R-1] (no
//    clamp).  The lanes left (fract below 2^-11, or NaN from a NaN
//    reciprocal) divide exactly.
//  * with c = R - kk the AND is [R - min over suffix edges of c,
//    R - max over prefix edges of c), "none" edges giving the empty set.
template <typename T>
__device__ __forceinline__ T three_edges_fast(const float (&D)[3], const float (&nN)[3], const float (&r)[3],
                                              const float (&div)[3], float Rs, int R) {
    bool f1[3], f2[3], slow[3];
    int c[3];
    bool anySlow = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        f1[k] = D[k] >= 0.0f;
        f2[k] = D[k] >= nN[k];
        const float x = __builtin_fmaf(D[k], r[k], Rs);
        asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(c[k]) : "v"(x));   // floor(x), saturating, NaN -> 0
        slow[k] = !(__builtin_amdgcn_fractf(x) >= 4.8828125e-4f) && (f1[k] != f2[k]);
        anySlow |= slow[k];
    }
    if (anySlow) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (slow[k]) c[k] = R - k_exact(D[k], div[k], R);
    }
    const int hiC = max3i((f1[0] && !f2[0]) ? c[0] : 0, (f1[1] && !f2[1]) ? c[1] : 0, (f1[2] && !f2[2]) ? c[2] : 0);
    const int loC = min3i((!f1[0] && f2[0]) ? c[0] : R, (!f1[1] && f2[1]) ? c[1] : R, (!f1[2] && f2[2]) ? c[2] : R);
    const bool none = (!f1[0] && !f2[0]) || (!f1[1] && !f2[1]) || (!f1[2] && !f2[2]);
    const unsigned w = __builtin_elementwise_sub_sat((unsigned)loC, (unsigned)hiC);
    const T m = interval_bits<T>(w, (unsigned)(R - loC));
    return none ? (T)0 : m;
}

// Rows whose three-edge mask can be non-empty.  The reference visits every
// row of the row-range mask (each edge on its own); most of those rows end
// with an empty AND because a prefix edge (B < 0) and a suffix edge (B > 0)
// exclude each other.  Skipping a row whose AND is empty changes no output,
// so the window below only has to contain every row with a non-empty AND.
//
// Bound.  For a row r and column c in [0, R-1], a set bit of one edge's mask
// implies A r + B c + C >= -tau with tau = 4u (|A| R + |B| R + |C|)
// (u = 2^-24: the rounding of A*r, of + C, of the quotient and of B*(R-1);
// the set bit means c < q (prefix), c >= q (suffix) or both end columns
// pass).  The code uses one tau' = 2^-20 (2 M R + Mc) + 2^-126 for the three
// edges (M = max |A|, |B|; Mc = max |C|) and C' = fl(C + tau').
// A prefix edge m and a suffix edge p meet in some real column only if
//   |B_m| (A_p r + C'_p) + B_p (A_m r + C'_m) >= 0,
// which for a pair (i, j) of opposite-signed B is  s G(r) >= 0  with
// G(r) = (B_i A_j - B_j A_i) r + (B_i C'_j - B_j C'_i) = al r + be, s = sign(B_i).
// al and be are computed in float; their rounding is at most 3u (Sa R + Sb)
// (Sa, Sb: sums of the absolute products, Sa <= 2 M^2, Sb <= 2 M Mc'), and
// E = 2^-19 M (M R + Mc') + 2^-126 bounds it, so every row with a non-empty
// mask has  s (al r + be) + E >= 0, i.e.
//   r >= t0 - E/|al|  (s al > 0)   or   r <= t0 + E/|al|  (s al < 0),  t0 = -be/al.
// t0 and E/|al| come from v_rcp_f32 (2^-22 relative); a margin of 2^-10
// covers that while |t0| <= 2^11, and a pair is used only when
// E/|al| <= 2^10, so that beyond |t0| = 2^11 the window is far outside
// [0, R-1] either way.  The bounds t0 -/+ (E/|al| + 2^-10) are formed with
// one fma each (t0 - m = -(be * ra + m), t0 + m = -be * ra + m), which only
// rounds less than a product followed by a sum.  E is raised to at least
// 2^-30 (a larger E only widens the window), so m <= 2^10 also keeps
// |ra| <= 2^40, i.e. |al| >= 2^-40: al is a normal number, and a zero or
// denormal al (infinite or huge reciprocal) is never used.  Lanes with any
// |A|, |B|, |C| above 2^40 (where the products could overflow) are not
// used; every bound that is used is then finite.
template <typename T>
__device__ __forceinline__ T row_window(const Edge (&e)[3], float Rf, float mAB) {
    const float u20 = 9.5367431640625e-07f;   // 2^-20
    const float tiny = 1.17549435e-38f;       // 2^-126
    const float lim = 1.09951163e+12f;        // 2^40
    // mAB = max |A|, |B| (NaN-free maximum); a NaN coefficient makes the
    // lane's pair terms NaN, which fail the |al| test and are not used
    const float mC = fmaxf(fmaxf(fabsf(e[0].C), fabsf(e[1].C)), fabsf(e[2].C));
    const bool ok = fmaxf(mAB, mC) <= lim;
    const float tau = __builtin_fmaf(__builtin_fmaf(mAB, 2.0f * Rf, mC), u20, tiny);
    const float Cp[3] = {e[0].C + tau, e[1].C + tau, e[2].C + tau};
    const float E = fmaxf(__builtin_fmaf(__builtin_fmaf(mAB, Rf, mC + tau), mAB * 1.9073486328125e-06f, tiny),
                          9.31322575e-10f);   // >= 2^-30
    bool pos[3], neg[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { pos[k] = e[k].B > 0.0f; neg[k] = e[k].B < 0.0f; }
    float lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int i = k, j = (k + 1) % 3;
        const float pA = e[i].B * e[j].A, qA = e[j].B * e[i].A;
        const float pB = e[i].B * Cp[j], qB = e[j].B * Cp[i];
        const float al = pA - qA, be = pB - qB;
        const float ra = __builtin_amdgcn_rcpf(al);
        const float m = __builtin_fmaf(E, fabsf(ra), 9.765625e-4f);   // E/|al| + 2^-10
        const bool opp = (pos[i] && neg[j]) || (neg[i] && pos[j]);
        const bool use = ok && opp && m <= 1024.0f;
        const bool lower = pos[i] == (al > 0.0f);
        lo[k] = (use && lower) ? -__builtin_fmaf(be, ra, m) : 0.0f;
        hi[k] = (use && !lower) ? __builtin_fmaf(-be, ra, m) : Rf - 1.0f;
    }
    const float l = __builtin_amdgcn_fmed3f(fmaxf(fmaxf(lo[0], lo[1]), lo[2]), 0.0f, Rf);
    const float h = __builtin_amdgcn_fmed3f(fminf(fminf(hi[0], hi[1]), hi[2]), -1.0f, Rf - 1.0f);
    int nlc, hc;   // -ceil(l) = floor(-l), floor(h)
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(nlc) : "v"(-l));
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(hc) : "v"(h));
    const unsigned w = (unsigned)max(hc + 1 + nlc, 0);
    return interval_bits<T>(w, (unsigned)(-nlc));
}

// test3DPointInTriangle, Step5_2_shadowCal.inl:213-222
__device__ __forceinline__ bool point_in_tri(const float* A, const float* B, const float* C, float px, float py, float pz) {
    auto cross = [](float ax, float ay, float az, float bx, float by, float bz, float* o) {
        o[0] = ay * bz - az * by; o[1] = az * bx - ax * bz; o[2] = ax * by - ay * bx;
    };
    float N1[3], N2[3], N3[3];
    cross(B[0] - A[0], B[1] - A[1], B[2] - A[2], px - B[0], py - B[1], pz - B[2], N1);
    cross(C[0] - B[0], C[1] - B[1], C[2] - B[2], px - C[0], py - C[1], pz - C[2], N2);
    if (dot3(N1[0], N1[1], N1[2], N2[0], N2[1], N2[2]) < 0.0f) return false;
    cross(A[0] - C[0], A[1] - C[1], A[2] - C[2], px - A[0], py - A[1], pz - A[2], N3);
    if (dot3(N2[0], N2[1], N2[2], N3[0], N3[1], N3[2]) < 0.0f) return false;
    return true;
}

// The work item's rows go out; LDS holds row r's low words at [r][lane],
// the high words (64-bit rows) R rows further.  The waves of the work item
// share the write-out.
__device__ __forceinline__ void store_masks(const FdArgs& a, const uint32_t* lds, int id, int n, int R, int wpr,
                                            int wave = 0, int waves = 1) {
    // coalesced write-out: consecutive lanes write consecutive words of one
    // receiver's mask (rows x wpr words); the receiver id comes from its
    // lane by a cross-lane read with every lane active
    const int per = R * wpr;
    const int total = n * per;
    for (int j0 = wave * 64; j0 < total; j0 += 64 * waves) {
        const int j = j0 + lane_id();
        const int l = min(j, total - 1) / per, w = j - l * per;
        const int idl = __shfl(id, l, 64);
        if (j < total) {
            const int r = w / wpr, h = w - r * wpr;
            a.masks[(size_t)idl * per + w] = lds[(h * R + r) * 64 + l];
        }
    }
}

// Masks of n receivers (lane l holds the id of receiver l) set to all
// occluded (back) or none, written straight to the output.
__device__ __forceinline__ void fill_masks(const FdArgs& a, int id, bool back, int n, int R, int wpr) {
    const int per = R * wpr;
    const int total = n * per;
    const uint32_t w0 = (R >= 32) ? 0xffffffffu : ((1u << R) - 1u);
    const uint32_t w1 = (R >= 64) ? 0xffffffffu : (R > 32 ? (1u << (R - 32)) - 1u : 0u);
    for (int j0 = 0; j0 < total; j0 += 64) {
        const int j = j0 + lane_id();
        const int l = min(j, total - 1) / per, w = j - l * per;
        const int idl = __shfl(id, l, 64);
        const bool bk = __shfl((int)back, l, 64) != 0;
        if (j < total) a.masks[(size_t)idl * per + w] = bk ? ((w % wpr) == 0 ? w0 : w1) : 0u;
    }
}

// Merge of a segment's partial masks: atomic OR of the non-zero words into
// the (zeroed) output; same word order as store_masks.
__device__ __forceinline__ void or_masks(const FdArgs& a, const uint32_t* lds, int id, int n, int R, int wpr,
                                         int wave = 0, int waves = 1) {
    const int per = R * wpr;
    const int total = n * per;
    for (int j0 = wave * 64; j0 < total; j0 += 64 * waves) {
        const int j = j0 + lane_id();
        const int l = min(j, total - 1) / per, w = j - l * per;
        const int idl = __shfl(id, l, 64);
        if (j < total) {
            const int r = w / wpr, h = w - r * wpr;
            const uint32_t v = lds[(h * R + r) * 64 + l];
            if (v) atomicOr(a.masks + (size_t)idl * per + w, v);
        }
    }
}

// Work item -> rank of its bin in the occlusion order: binary search over
// the work-item prefix (the bin is binAtRank[rank]).
__device__ __forceinline__ int chunk_rank(const FdArgs& a, int c, int nb) {
    int lo = 0, hi = nb + 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.chunkStart[mid] <= c) lo = mid; else hi = mid;
    }
    return uniform(lo);
}

// --------------------------------------------------------------------------
// Bin-queue occlusion variant (shadowCal_CTA_kernel, Step5_2_shadowCal.inl:
// 408-1100, dead code in the reference).  Its arithmetic is not the
// per-sample kernel's: edge planes from dot(cross(edge, axis), S - v), a
// fourth plane (the triangle's own) in the row masks, other boundary tests,
// and mask rows along y_delt (bit j of row i: upRight + j x_delt + i y_delt).
// Not part of the per-sample bit-exact contract; selected per context.
//
// Work items and outputs are the per-sample kernel's (64 sorted receivers of
// a bin, the bin's triangle list or a stretch of it).  The lanes take 64
// consecutive triangles, two groups of 32 (the reference's warps: a group
// applies the triangle plane when any of its triangles reaches a row and is
// not entirely in front of the receiver, :876/:929), and the wave walks the
// receivers one by one, their data broadcast from the owning lane.  Row
// masks are merged with LDS atomic ORs into the receivers' rows; rows that
// become full are flagged and skipped (as the reference's flag word, :162-164).
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T cta_shl(int k) { return (k >= 0 && k < Bits<T>::width) ? ((T)1 << k) : (T)0; }

// min(max(ceil(-D / div), 0), R) with the IEEE quotient (NaN -> R), from
// the reciprocal r = rcp_or_nan(div) where k_fast allows it (its bound holds
// for this clamp as well: beyond R + 1 both give R, below 0 both give 0).
__device__ __forceinline__ int cta_k(float D, float div, float r, float Rf) {
    bool slow;
    int k = k_fast(D, r, Rf, slow);
    if (slow) k = (int)fmaxf(0.0f, fminf(Rf, ceilf(-1.0f * D / div)));
    return k;
}

// calLightRowRange (:302-357); bits >= W are cleared by the caller.
// rB = rcp_or_nan(B).
template <typename T>
__device__ __forceinline__ T cta_row_range(float B, float D, int W, float rB) {
    if (B <= 0.0f && D <= 0.0f) return (T)0;
    if (B == 0.0f || D == 0.0f) return cta_shl<T>(W) - (T)1;
    const T m = cta_shl<T>(cta_k(D, B, rB, (float)W)) - (T)1;
    return B < 0.0f ? m : ~m;
}

// calLightHeightRange (:225-300), H = W = R; rA = rcp_or_nan(A)
template <typename T>
__device__ __forceinline__ T cta_height_range(float A, float B, float C, int R, T fullH, float rA) {
    const float D = fmaxf(B * (float)(R - 1) + C, C);
    if (A <= 0.0f && D <= 0.0f) return (T)0;
    if (A == 0.0f || D == 0.0f) return fullH;
    const T m = cta_shl<T>(cta_k(D, A, rA, (float)R)) - (T)1;
    return A < 0.0f ? m : (~m & fullH);
}

__device__ __forceinline__ float bcast(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// The triangle loop of the variant for one work item: n receivers (lane s
// holds receiver s in q), triangle list entries [ts, te).  Rows go to lds
// (row r of receiver s at lds[r * 64 + s], high words R * 64 further).
template <typename T>
__device__ __forceinline__ void cta_tris(const FdArgs& a, uint32_t* lds, T* fullRows, int ts, int te, int n,
                                         const FdRxRec& q, int R) {
    const int lane = lane_id();
    const T FULL = (R >= Bits<T>::width) ? ~(T)0 : (((T)1 << R) - (T)1);   // lightWidthFullBit
    // (T)(1 << H) - 1 with the int shift of the reference (:431): all ones of T from H = 32 on
    const T fullH = (R >= 32) ? ~(T)0 : (((T)1 << R) - (T)1);
    const float xd0 = a.L.xd[0], xd1 = a.L.xd[1], xd2 = a.L.xd[2];
    const float yd0 = a.L.yd[0], yd1 = a.L.yd[1], yd2 = a.L.yd[2];
    const float ur0 = a.L.ur[0], ur1 = a.L.ur[1], ur2 = a.L.ur[2];
    const float eps = a.eps;
    const uint64_t half = lane < 32 ? 0x00000000ffffffffull : 0xffffffff00000000ull;
    for (int tb = ts; tb < te; tb += 64) {
        const int ti = tb + lane;
        int tid = ti < te ? a.sortedTri[ti] : -1;
        const bool has = (unsigned)tid < (unsigned)a.ntris;
        float v[3][3] = {}, bminx = 3.402823466e+38f, bminy = 3.402823466e+38f, bmaxx = -3.402823466e+38f,
              bmaxy = -3.402823466e+38f, minz = 3.402823466e+38f, maxz = -3.402823466e+38f;
        if (has) {
            const FdTri& t = a.tri[tid];
            for (int k = 0; k < 3; ++k) { v[0][k] = t.v0[k]; v[1][k] = t.v1[k]; v[2][k] = t.v2[k]; }
            bminx = t.bminx; bminy = t.bminy; bmaxx = t.bmaxx; bmaxy = t.bmaxy;
            minz = t.minz; maxz = t.maxz;
        }
        // receiver-independent parts (:667-674, and the cross products of :720-742):
        // edge k runs from v_k to v_{k+1} (v0_v1, v1_v2, v2_v0), its plane is
        // dot(cross(edge, axis), S - v_{k+1})
        float cy[3][3], cx[3][3], cl[3][3];
        for (int k = 0; k < 3; ++k) {
            const int j = (k + 1) % 3;
            const float ex = v[k][0] - v[j][0], ey = v[k][1] - v[j][1], ez = v[k][2] - v[j][2];
            const float lx = ur0 - v[j][0], ly = ur1 - v[j][1], lz = ur2 - v[j][2];
            cy[k][0] = ey * yd2 - ez * yd1; cy[k][1] = ez * yd0 - ex * yd2; cy[k][2] = ex * yd1 - ey * yd0;
            cx[k][0] = ey * xd2 - ez * xd1; cx[k][1] = ez * xd0 - ex * xd2; cx[k][2] = ex * xd1 - ey * xd0;
            cl[k][0] = ey * lz - ez * ly;   cl[k][1] = ez * lx - ex * lz;   cl[k][2] = ex * ly - ey * lx;
        }
        // n = cross(v1 - v2, v2 - v0); A3, B3, C3 = n . y_delt, n . x_delt, n . (upRight - v0)
        float nn[3];
        {
            const float ax = v[1][0] - v[2][0], ay = v[1][1] - v[2][1], az = v[1][2] - v[2][2];
            const float bx = v[2][0] - v[0][0], by = v[2][1] - v[0][1], bz = v[2][2] - v[0][2];
            nn[0] = ay * bz - az * by; nn[1] = az * bx - ax * bz; nn[2] = ax * by - ay * bx;
        }
        const float A3o = dot3(nn[0], nn[1], nn[2], yd0, yd1, yd2);
        const float B3o = dot3(nn[0], nn[1], nn[2], xd0, xd1, xd2);
        const float C3o = dot3(nn[0], nn[1], nn[2], ur0 - v[0][0], ur1 - v[0][1], ur2 - v[0][2]);
        if (!__any(has)) break;
        for (int s = 0; s < n; ++s) {
            const T done = fullRows[s];
            if (done == FULL) continue;   // every row of this receiver is full
            const float rminx = bcast(q.rminx, s), rminy = bcast(q.rminy, s);
            const float rmaxx = bcast(q.rmaxx, s), rmaxy = bcast(q.rmaxy, s), rz = bcast(q.z, s);
            // :686-689
            const bool in = bminx <= rminx && bminy <= rminy && rmaxx <= bmaxx && rmaxy <= bmaxy && minz <= rz;
            if (!__any(in)) continue;
            const float S0 = bcast(q.q[0], s), S1 = bcast(q.q[1], s), S2 = bcast(q.q[2], s);
            T rows = 0;
            float A[4], B[4], C[4], rB[4];
            if (in) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int j = (k + 1) % 3;
                    const float sx = S0 - v[j][0], sy = S1 - v[j][1], sz = S2 - v[j][2];
                    A[k] = dot3(cy[k][0], cy[k][1], cy[k][2], sx, sy, sz);
                    B[k] = dot3(cx[k][0], cx[k][1], cx[k][2], sx, sy, sz);
                    C[k] = dot3(cl[k][0], cl[k][1], cl[k][2], sx, sy, sz);
                }
                A[3] = A3o; B[3] = B3o; C[3] = C3o;
                const float temp = dot3(nn[0], nn[1], nn[2], S0 - v[0][0], S1 - v[0][1], S2 - v[0][2]);   // :772
                const bool skip = fabsf(temp) < eps && point_in_tri(v[0], v[1], v[2], S0, S1, S2);   // :783
                if (!skip) {
                    if (temp > 0.0f) {   // :785-802
#pragma unroll
                        for (int k = 0; k < 4; ++k) { A[k] = -A[k]; B[k] = -B[k]; C[k] = -C[k]; }
                    }
                    rows = fullH;   // :821-842
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        rows &= cta_height_range<T>(A[k], B[k], C[k], R, fullH, rcp_or_nan(A[k]));
                        rB[k] = rcp_or_nan(B[k]);
                    }
                }
            }
            // the group's triangle-plane decision (:876, :929), on the rows
            // before the removal of full rows
            const bool needDepth = rows != (T)0 && !(maxz < rz);
            const bool depth = (__ballot(needDepth) & half) != 0ull;
            rows &= FULL & ~done;   // rows >= H are not part of the grid
            while (rows) {
                const int i = Bits<T>::ctz(rows);
                const float fi = (float)i;
                T m = FULL;
#pragma unroll
                for (int k = 0; k < 3; ++k) m &= cta_row_range<T>(B[k], A[k] * fi + C[k], R, rB[k]);
                if (depth) m &= cta_row_range<T>(B[3], A[3] * fi + C[3], R, rB[3]);
                m &= FULL;
                if (m != (T)0) {
                    T cur;
                    if (Bits<T>::width == 32) {
                        cur = (T)(atomicOr(lds + i * 64 + s, (uint32_t)m) | (uint32_t)m);
                    } else {
                        const uint64_t mm = (uint64_t)m;
                        const uint32_t lo = atomicOr(lds + i * 64 + s, (uint32_t)mm) | (uint32_t)mm;
                        const uint32_t hi = atomicOr(lds + (R + i) * 64 + s, (uint32_t)(mm >> 32)) | (uint32_t)(mm >> 32);
                        cur = (T)(((uint64_t)hi << 32) | lo);
                    }
                    if (cur == FULL) atomicOr(fullRows + s, (T)1 << i);
                }
                rows &= rows - (T)1;
            }
        }
    }
}

// NW waves per work item (1 or 2): with 2, the waves share the work item's
// receivers and LDS rows and take alternate triangles of the list (rows
// merged with LDS atomics), so a CU holds twice the waves for the same LDS.
template <typename T, bool COUNT, bool CTA = false, int NW = 1>
__device__ __forceinline__ void occlude_body(const FdArgs& a, const int blkIn) {
    // wpr * R * 64 row words, then 64 full-row masks (T, 64-bit atomics for
    // 64-bit rows: the array is aligned for them)
    extern __shared__ __align__(16) uint32_t lds[];
    const FdState* s = a.st;
    const int status = s->status;
    const int R = a.L.res;                  // rows == columns
    const int wpr = a.L.wpr;
    const int lane = lane_id();
    const int wv = NW > 1 ? uniform((int)(threadIdx.x >> 6)) : 0;   // wave of the work item
    const T FULL = (R >= Bits<T>::width) ? ~(T)0 : (((T)1 << R) - (T)1);
    // work item = (triangle segment seg, chunk c); segment-major so that
    // neighbouring work items share a stretch of the same triangle list
    const int S = a.segs;
    const int seg = blkIn / a.chunkSlots;
    const int c = blkIn - seg * a.chunkSlots;

    if (status != FD_EXIT_OK) {
        if (seg != 0 || wv != 0) return;
        // early exit of the frame: every receiver shadow 0, except back-facing
        // receivers once the back-face stage has run (FD/api.cpp:489,
        // Step2_1_samplerRectangleAABB.inl:97)
        const int i0 = c * 64;
        if (i0 >= a.WH) return;
        const int n = min(64, a.WH - i0);
        const bool backStage = status >= FD_EXIT_NO_VALID_SAMPLE;
        bool back = false;
        if (lane < n) {
            back = backStage && __float_as_uint(a.rec[i0 + lane].w) == FD_TAG_BACK;
            a.shadow[i0 + lane] = back ? 1.0f : 0.0f;
        }
        fill_masks(a, i0 + lane, back, n, R, wpr);
        return;
    }
    if (c >= s->totalChunks) return;
    const int nb = s->nb;
    const int rank = chunk_rank(a, c, nb);
    const int bin = uniform(a.binAtRank[rank]);
    if (bin == nb && seg != 0) return;
    // this work item's stretch of the bin's triangle list (bins 0..nb-1)
    int ts = 0, te = 0;
    if (bin != nb) {
        ts = a.triStart[bin];
        te = a.triEnd[bin];
        if (S > 1) {
            int per = (te - ts + S - 1) / S;
            if (CTA) per = (per + 31) & ~31;   // stretches keep the groups of 32
            ts += seg * per;
            te = min(te, ts + per);
            if (ts >= te) return;   // nothing to add; the finishing pass counts the merged masks
        }
    }
    const int s0 = a.rxStart[bin] + (c - a.chunkStart[rank]) * 64;
    const int n = min(64, a.rxEnd[bin] - s0);
    const bool act = lane < n;
    int id = 0;
    if (act) id = a.sortedId[s0 + lane];
    if ((unsigned)id >= (unsigned)a.WH) id = 0;   // defensive: ids form a permutation

    if (bin == nb) {
        // receivers outside the light plane, invalid or back-facing
        if (wv != 0) return;
        bool back = false;
        if (act) {
            back = __float_as_uint(a.rec[id].w) == FD_TAG_BACK;
            a.shadow[id] = back ? 1.0f : 0.0f;
        }
        fill_masks(a, id, back, n, R, wpr);
        return;
    }

    FdRxRec q{};
    if (act) q = a.sortedRec[s0 + lane];
    for (int r = wv; r < R; r += NW) {
        lds[r * 64 + lane] = 0u;
        if (wpr == 2) lds[(R + r) * 64 + lane] = 0u;
    }
    // rows shared by several waves: the rows each receiver has full, as any
    // wave found them (after the rows)
    constexpr bool SHARE_FULL = NW > 1 && !CTA;
    T* const fullSh = (T*)(lds + wpr * R * 64);
    if (SHARE_FULL && wv == 0) fullSh[lane] = (T)0;
    if (NW > 1) __syncthreads();
    if (CTA) {
        T* fullRows = (T*)(lds + wpr * R * 64);
        fullRows[lane] = (T)0;
        __syncthreads();
        cta_tris<T>(a, lds, fullRows, ts, te, n, q, R);
        __syncthreads();
        if (S == 1) {
            int cnt = 0;
            for (int r = 0; r < R; ++r) {
                cnt += __popc(lds[r * 64 + lane]);
                if (wpr == 2) cnt += __popc(lds[(R + r) * 64 + lane]);
            }
            if (act) a.shadow[id] = (float)cnt / (float)a.L.resNum;   // :1080-1088
            store_masks(a, lds, id, n, R, wpr);
        } else {
            or_masks(a, lds, id, n, R, wpr);
        }
        return;
    }
    const float Lx = a.L.ur[0] - q.q[0], Ly = a.L.ur[1] - q.q[1], Lz = a.L.ur[2] - q.q[2];  // L0 - Q
    const float xd0 = a.L.xd[0], xd1 = a.L.xd[1], xd2 = a.L.xd[2];
    const float yd0 = a.L.yd[0], yd1 = a.L.yd[1], yd2 = a.L.yd[2];
    // light axes along x and y (x_delt = (c, 0, 0), y_delt = (0, c', 0), zeros
    // of either sign): then dot(n, x_delt) = n.x * c exactly, up to the sign
    // of a zero result, whenever n is finite (the other two products are
    // zeros).  No later step can tell +0 from -0 (comparisons with 0, the
    // exact-division fallback maps both infinities of a zero divisor to R, and
    // the row window only uses magnitudes).
    const bool axial = xd1 == 0.0f && xd2 == 0.0f && yd0 == 0.0f && yd2 == 0.0f;
    const float nsub = (float)(unsigned)(R - 1);
    const float Rf = (float)R;
    const float Rs = Rf + 2.44140625e-4f;   // R + 2^-12, exact
    const float eps = a.eps;
    T full = 0;
    unsigned long long nCand = 0, nPass = 0, nRows = 0, nWaveIt = 0, nWaveRows = 0, nRefRows = 0, nPlanes = 0;
    // read-only for this kernel: the constant address space lets the uniform
    // per-triangle loads go through the scalar unit (s_load)
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(4))) int* cint_p;
    typedef const __attribute__((address_space(4))) FdTri* ctri_p;
#else
    typedef const int* cint_p;   // host pass: the kernel body is never run
    typedef const FdTri* ctri_p;
#endif
    // the bin's list in the frame's walk order (the masks do not depend on it);
    // the order is read once into a scalar register, so the list address
    // stays scalar
    const int walkOrder = uniform(s->triOrder);
    const cint_p triList = (cint_p)(walkOrder != FD_TRI_ORDER_LIST ? (const int*)a.pairKT : a.sortedTri);
    const ctri_p tris = (ctri_p)(a.tri);
    bool live = act;   // receiver in this wave and not yet fully shadowed
    // Prefilter (timed kernel): the wave reads its next 64 list entries and
    // their culling boxes with vector loads and keeps the triangles that pass
    // the containment + depth test against bounds of all its receivers (the
    // largest lower corner, the smallest upper corner, the largest depth).  A
    // triangle that fails them fails the test of every lane, which then skips
    // it: dropping it early changes nothing.  The counting kernel walks every
    // entry, as the reference does.
    const int* listV = walkOrder != FD_TRI_ORDER_LIST ? (const int*)a.pairKT : a.sortedTri;
    float gx0 = 0.0f, gy0 = 0.0f, gx1 = 0.0f, gy1 = 0.0f, gz = 0.0f;
    if (!COUNT) {
        auto su = [](float v) { return __int_as_float(uniform(__float_as_int(v))); };   // into a scalar register
        gx0 = su(wave_max_f32(act ? q.rminx : -INFINITY));
        gy0 = su(wave_max_f32(act ? q.rminy : -INFINITY));
        gx1 = su(wave_min_f32(act ? q.rmaxx : INFINITY));
        gy1 = su(wave_min_f32(act ? q.rmaxy : INFINITY));
        gz = su(wave_max_f32(act ? q.z : -INFINITY));
    }
    // one pair (triangle tid, this wave's receivers): cull, planes, rows
    auto step = [&](const int tid) {
        if ((unsigned)tid >= (unsigned)a.ntris) return;   // defensive: pair entries are triangle ids
        if (SHARE_FULL) {   // rows the other waves filled
            full |= (T)fullSh[lane];
            if (full == FULL) live = false;
        }
        const FdTri t = tris[tid];
        // the whole record in one round trip (the prefiltered pairs mostly
        // pass the cull below, which needs only part of it)
        if (!COUNT) asm volatile("" ::"s"(t.v0[0]), "s"(t.bminx), "s"(t.minz), "s"(t.n[0]), "s"(t.e20[0]));
        // containment + depth cull (:120-127)
        const bool pass = live && t.bminx <= q.rminx && t.bminy <= q.rminy && q.rmaxx <= t.bmaxx &&
                          q.rmaxy <= t.bmaxy && t.minz <= q.z;
        if (!pass) return;
        if (COUNT) ++nPass;
        // planes (Q,v0,v1), (Q,v1,v2), (Q,v2,v0) (:150-172) start from
        // a_k = v_{k+1} - Q; a_2 = v0 - Q also gives
        // dir = dot(Q - v0, n) (:130-132) = -dot(v0 - Q, n), exactly (every
        // rounding commutes with the negation)
        const float* vb[3] = {t.v1, t.v2, t.v0};
        const float* ev[3] = {t.e01, t.e12, t.e20};
        float av[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            av[k][0] = vb[k][0] - q.q[0];
            av[k][1] = vb[k][1] - q.q[1];
            av[k][2] = vb[k][2] - q.q[2];
        }
        const float ndir = dot3(av[2][0], av[2][1], av[2][2], t.n[0], t.n[1], t.n[2]);   // -dir
        // Q on the triangle: the reference skips the pair (:134-135).  The
        // counting variant tests here, in the reference's order; the timed
        // variant tests below, only for pairs that can still set a bit (the
        // outcome is the same: a skipped pair sets nothing).
        const bool onTri = fabsf(ndir) < eps;
        if (COUNT && onTri && point_in_tri(t.v0, t.v1, t.v2, q.q[0], q.q[1], q.q[2])) return;
        // The reference negates t0, t1, L0_Q when dir < 0.  The sign is put
        // into L0 - Q and the light axes before the products (a sign flip
        // commutes with every rounding): A, B, C come out negated, identical
        // up to the sign of zero results, which no later step can observe.
        const uint32_t sg = (ndir > 0.0f) ? 0x80000000u : 0u;   // dir < 0
        auto flip = [sg](float v) { return __uint_as_float(__float_as_uint(v) ^ sg); };
        const float Lxs = flip(Lx), Lys = flip(Ly), Lzs = flip(Lz);
        Edge e[3];
        float nv[3][3];
        bool finiteC = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float ax = av[k][0], ay = av[k][1], az = av[k][2];
            const float bx = ev[k][0], by = ev[k][1], bz = ev[k][2];
            nv[k][0] = ay * bz - az * by;
            nv[k][1] = az * bx - ax * bz;
            nv[k][2] = ax * by - ay * bx;
            e[k].C = dot3(nv[k][0], nv[k][1], nv[k][2], Lxs, Lys, Lzs);
            finiteC = finiteC && __builtin_isfinite(e[k].C);
        }
        // a finite C = dot(n, L0 - Q) implies a finite n (an infinite
        // component makes a product infinite or NaN)
        if (axial && !__any(!finiteC)) {
            const float xs = flip(xd0), ys = flip(yd1);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                e[k].A = nv[k][0] * xs;
                e[k].B = nv[k][1] * ys;
            }
        } else {
            const float x0 = flip(xd0), x1 = flip(xd1), x2 = flip(xd2);
            const float y0 = flip(yd0), y1 = flip(yd1), y2 = flip(yd2);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                e[k].A = dot3(nv[k][0], nv[k][1], nv[k][2], x0, x1, x2);
                e[k].B = dot3(nv[k][0], nv[k][1], nv[k][2], y0, y1, y2);
            }
        }
        // |A|, |B| <= 2^100 for the fast mask form (a NaN dropped by fmaxf is
        // harmless there: it makes both f1 and f2 false, as in the reference)
        const float mAB = fmaxf(fmaxf(fmaxf(fabsf(e[0].A), fabsf(e[1].A)), fabsf(e[2].A)),
                                fmaxf(fmaxf(fabsf(e[0].B), fabsf(e[1].B)), fabsf(e[2].B)));
        const bool big = !(mAB <= 1.2676506e+30f);
        if (COUNT) {
            // the rows the reference visits for this pair: its row-range mask
            // minus the rows already full (general form, before any window)
            ++nPlanes;
            Edge g[3];
            float DNg[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                g[k] = e[k];
                g[k].nRowN = -(e[k].A * nsub);
                g[k].nColN = -(e[k].B * nsub);
                g[k].rA = rcp_or_nan(e[k].A);
                g[k].rB = rcp_or_nan(e[k].B);
                DNg[k] = (e[k].B > 0 ? -g[k].nColN : 0.0f) + e[k].C;
            }
            nRefRows += Bits<T>::popc((~full) & FULL & three_edges<T, true>(g, DNg, Rf, R));
        }
        // row by row, lowest row first (:204-222)
        auto rows = [&](T valid, auto rowMask) {
            while (valid) {
                if (COUNT) ++nRows;
                const int r = Bits<T>::ctz(valid);
                const float fr = (float)r;
                float D[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) D[k] = e[k].A * fr + e[k].C;
                const T M = rowMask(D);
                uint32_t* w0 = lds + r * 64 + lane;
                T cur;
                if (Bits<T>::width == 32) {
                    if (NW > 1) {
                        cur = (T)(atomicOr(w0, (uint32_t)M) | (uint32_t)M);
                    } else {
                        cur = (T)(*w0 | (uint32_t)M);
                        *w0 = (uint32_t)cur;
                    }
                } else {
                    uint32_t* w1 = lds + (R + r) * 64 + lane;
                    const uint64_t mm = (uint64_t)M;
                    uint32_t lo32, hi32;
                    if (NW > 1) {
                        lo32 = atomicOr(w0, (uint32_t)mm) | (uint32_t)mm;
                        hi32 = atomicOr(w1, (uint32_t)(mm >> 32)) | (uint32_t)(mm >> 32);
                    } else {
                        lo32 = *w0 | (uint32_t)mm;
                        hi32 = *w1 | (uint32_t)(mm >> 32);
                        *w0 = lo32;
                        *w1 = hi32;
                    }
                    cur = (T)(((uint64_t)hi32 << 32) | lo32);
                }
                if (cur == FULL) {
                    full |= (T)1 << r;
                    if (SHARE_FULL) {
                        if (Bits<T>::width == 32) atomicOr((uint32_t*)(fullSh + lane), (uint32_t)1 << r);
                        else atomicOr((unsigned long long*)(fullSh + lane), 1ull << r);
                    }
                }
                valid &= valid - (T)1;
            }
        };
        auto wave_rows = [&](T valid) {
            if (COUNT) {
                // wave row iterations = max over lanes of this triangle's rows
                int nr = Bits<T>::popc(valid);
                for (int o = 32; o >= 1; o >>= 1) nr = max(nr, __shfl_xor(nr, o, 64));
                if (lane == __ffsll((unsigned long long)__ballot(1)) - 1) nWaveRows += nr;
            }
        };
        if (!__any(big)) {
            // fast form: finite limits in every lane of the wave.
            // |A|, |B| <= 2^100 here, so only the lower end of the reciprocal
            // range is checked (NaN coefficients fail it as well);
            // DN = C + max(B (R-1), 0) equals (B > 0 ? B (R-1) : 0) + C up to
            // the sign of a zero, B being finite or NaN
            float nR[3], rA[3], dA[3], bn[3], DN[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                dA[k] = e[k].A;
                nR[k] = -(e[k].A * nsub);
                bn[k] = e[k].B * nsub;
                rA[k] = fabsf(e[k].A) >= 7.8886091e-31f ? __builtin_amdgcn_rcpf(e[k].A) : __builtin_nanf("");
                DN[k] = e[k].C + fmaxf(bn[k], 0.0f);
            }
            // rows that may hold occluded samples (:176-191), then the row
            // window; a wave none of whose lanes keeps a row skips the
            // triangle at either step (the cheaper test first)
            T valid = (~full) & FULL & three_edges_fast<T>(DN, nR, rA, dA, Rs, R);
            if (!__any(valid != (T)0)) return;
            valid &= row_window<T>(e, Rf, mAB);
            if (!__any(valid != (T)0)) return;   // no lane can gain a bit from this triangle
            if (!COUNT && onTri && valid != (T)0 && point_in_tri(t.v0, t.v1, t.v2, q.q[0], q.q[1], q.q[2])) return;
            float nC[3], rB[3], dB[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                dB[k] = e[k].B;
                nC[k] = -bn[k];
                rB[k] = fabsf(e[k].B) >= 7.8886091e-31f ? __builtin_amdgcn_rcpf(e[k].B) : __builtin_nanf("");
                asm("" : "+v"(rB[k]));   // keep the select out of the row loop
            }
            wave_rows(valid);
            rows(valid, [&](const float (&D)[3]) { return three_edges_fast<T>(D, nC, rB, dB, Rs, R); });
        } else {
            if (!COUNT && onTri && point_in_tri(t.v0, t.v1, t.v2, q.q[0], q.q[1], q.q[2])) return;
            float DN[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                e[k].nRowN = -(e[k].A * nsub);
                e[k].nColN = -(e[k].B * nsub);
                e[k].rA = rcp_or_nan(e[k].A);
                e[k].rB = rcp_or_nan(e[k].B);
                DN[k] = (e[k].B > 0 ? -e[k].nColN : 0.0f) + e[k].C;
            }
            const T valid = (~full) & FULL & three_edges<T, true>(e, DN, Rf, R);
            wave_rows(valid);
            rows(valid, [&](const float (&D)[3]) { return (T)(FULL & three_edges<T, false>(e, D, Rf, R)); });
        }
        if (full == FULL) live = false;
    };
    if (COUNT) {
        for (int ti = ts + wv; ti < te; ti += NW) {
            const uint64_t liveMask = __ballot(live);
            if (liveMask == 0) break;
            nCand += __popcll(liveMask);
            ++nWaveIt;
            step(triList[ti]);
        }
    } else {
        // blocks of 64 of this wave's entries; the candidate loop reads the
        // block's ids from the lanes that loaded them (written once per
        // block, so any subset of lanes that runs an iteration reads the
        // same id).  The next block's ids load while this block's boxes are
        // tested and its pairs run; the box fields load unconditionally, in
        // one round trip.
        int tidNext = (ts + wv + lane * NW < te) ? listV[ts + wv + lane * NW] : -1;
        for (int blk0 = ts + wv; blk0 < te; blk0 += 64 * NW) {
            if (__ballot(live) == 0) break;
            const int tj = blk0 + lane * NW;
            const int tidv = tidNext;
            bool keep = false;
            float4 bmn = make_float4(0.f, 0.f, 0.f, 0.f), bmx = bmn;
            const bool okv = (unsigned)tidv < (unsigned)a.ntris;
            if (okv) {
                bmn = a.triAABB[2 * tidv];
                bmx = a.triAABB[2 * tidv + 1];
            }
            const int tn = tj + 64 * NW;
            tidNext = tn < te ? listV[tn] : -1;
            if (okv) keep = bmn.x <= gx0 && bmn.y <= gy0 && gx1 <= bmx.x && gy1 <= bmx.y && bmn.z <= gz;
            for (uint64_t cmask = __ballot(keep); cmask; cmask &= cmask - 1) {
                if (__ballot(live) == 0) break;
                step(__builtin_amdgcn_readlane(tidv, uniform(__ffsll((unsigned long long)cmask) - 1)));
            }
        }
    }
    if (COUNT) {
        // per-lane counts: nCand is wave-wide already
        unsigned long long p = nPass, r = nRows, rr = nRefRows, pl = nPlanes;
        for (int o = 32; o >= 1; o >>= 1) {
            p += __shfl_xor(p, o, 64);
            r += __shfl_xor(r, o, 64);
            rr += __shfl_xor(rr, o, 64);
            pl += __shfl_xor(pl, o, 64);
        }
        unsigned long long wr = nWaveRows;
        for (int o = 32; o >= 1; o >>= 1) wr += __shfl_xor(wr, o, 64);
        if (lane == 0) {
            atomicAdd(a.opCount + 0, nCand);
            atomicAdd(a.opCount + 1, p);
            atomicAdd(a.opCount + 2, r);
            if (wv == 0) atomicAdd(a.opCount + 3, 1ull);
            atomicAdd(a.opCount + 4, nWaveIt);
            atomicAdd(a.opCount + 5, wr);
            atomicAdd(a.opCount + 6, rr);
            atomicAdd(a.opCount + 7, pl);
        }
    }
    __syncthreads();
    // fraction = popcount / (W*H) (:226-231); split lists: after the merge.
    // The write-out is shared by the waves.
    if (wv == 0 && S == 1) {
        int cnt = 0;
        for (int r = 0; r < R; ++r) {
            cnt += __popc(lds[r * 64 + lane]);
            if (wpr == 2) cnt += __popc(lds[(R + r) * 64 + lane]);
        }
        if (act) a.shadow[id] = (float)cnt / (float)a.L.resNum;
    }
    if (S == 1) store_masks(a, lds, id, n, R, wpr, wv, NW);
    else or_masks(a, lds, id, n, R, wpr, wv, NW);
}

// Split lists only: zeroes the mask output before the occlusion merges into
// it (a fill kernel, so that the sequence stays graph-capturable).
constexpr int ZERO_WORDS = 4096;   // words per block
__device__ __forceinline__ void zero_masks_body(const FdArgs& a, const int blk) {
    const size_t n = (size_t)a.WH * a.L.res * a.L.wpr;
    const size_t base = (size_t)blk * ZERO_WORDS;
    for (int k = threadIdx.x; k < ZERO_WORDS; k += blockDim.x)
        if (base + k < n) a.masks[base + k] = 0u;
}

// Split lists: shadow value of every receiver of a
// real bin from its merged mask (popcount / samples, :226-231).
__device__ __forceinline__ void occlude_finish_body(const FdArgs& a, const int c) {
    __shared__ int cnt[64];
    const FdState* s = a.st;
    if (s->status != FD_EXIT_OK || c >= s->totalChunks) return;
    const int nb = s->nb;
    const int rank = chunk_rank(a, c, nb);
    const int bin = a.binAtRank[rank];
    if (bin == nb) return;
    const int s0 = a.rxStart[bin] + (c - a.chunkStart[rank]) * 64;
    const int n = min(64, a.rxEnd[bin] - s0);
    const int lane = lane_id();
    int id = lane < n ? a.sortedId[s0 + lane] : 0;
    if ((unsigned)id >= (unsigned)a.WH) id = 0;   // defensive: ids form a permutation
    cnt[lane] = 0;
    __syncthreads();
    // each receiver's words are contiguous: consecutive lanes read
    // consecutive words, and the counts gather per receiver in LDS
    const int per = a.L.res * a.L.wpr;
    const int total = n * per;
    for (int j0 = 0; j0 < total; j0 += 64) {
        const int j = j0 + lane;
        const int l = min(j, total - 1) / per, w = j - l * per;
        const int idl = __shfl(id, l, 64);
        if (j < total) {
            const int b = __popc(a.masks[(size_t)idl * per + w]);
            if (b) atomicAdd(&cnt[l], b);
        }
    }
    __syncthreads();
    if (lane < n) a.shadow[id] = (float)cnt[lane] / (float)a.L.resNum;
}

// ==========================================================================
// Kernel entry points.  Single frame: the FdArgs block travels as the kernel
// argument.  Batched: every stage is ONE launch over all instances; block b
// belongs to instance i with map[i] <= b < map[i+1] (host-built prefix), and
// the instance's FdArgs is read from the device-resident instance table.
// ==========================================================================
__global__ void k_init(FdArgs a) { init_body(a, blockIdx.x); }
__global__ void k_init_bins(FdArgs a) { init_bins_body(a, blockIdx.x); }
__global__ __launch_bounds__(RED_BLOCK) void k_reduce(FdArgs a) { reduce_body(a, blockIdx.x); }
__global__ __launch_bounds__(FD_FIN_THREADS) void k_finalize(FdArgs a) { finalize_body(a, blockIdx.x); }
__global__ __launch_bounds__(256) void k_rx_key(FdArgs a) { rx_key_body(a, blockIdx.x); }
__global__ __launch_bounds__(64) void k_rx_scatter(FdArgs a) { rx_scatter_body(a, blockIdx.x); }
__global__ __launch_bounds__(FD_TRB) void k_tri_count(FdArgs a) { tri_count_body(a, blockIdx.x); }
__global__ __launch_bounds__(FD_TRB) void k_tri_write(FdArgs a) { tri_write_body(a, blockIdx.x); }
__global__ __launch_bounds__(FD_TRB) void k_tri_scatter_wide(FdArgs a) { tri_scatter_wide_body(a, blockIdx.x); }
__global__ __launch_bounds__(1024) void k_tables(FdArgs a) { tables_body(a, blockIdx.x); }
__global__ __launch_bounds__(FD_FIN_THREADS) void k_post_pairs(FdArgs a) { post_pairs_body(a, blockIdx.x); }
__global__ __launch_bounds__(FD_ORDER_THREADS) void k_tri_order(FdArgs a) { tri_order_body(a, blockIdx.x); }
// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD
// b % 8).  Work items of one bin share that bin's triangle list, so the
// occlusion kernels renumber blocks so that each XCD takes runs of 16
// consecutive work items (runs still alternate between XCDs, which keeps the
// load balanced): neighbouring chunks of a bin then read the list from the
// same L2.  A bijection of [0, n); the tail past the last full group of
// 8 x 16 blocks keeps its order.
__device__ __forceinline__ int xcd_block(int b, int n) {
    constexpr int G = 16;
    if (b >= n - n % (8 * G)) return b;
    const int j = b % (8 * G);
    return b - j + (j & 7) * G + (j >> 3);
}

__global__ __launch_bounds__(256) void k_zero_masks(FdArgs a) { zero_masks_body(a, blockIdx.x); }
__global__ __launch_bounds__(64) void k_occlude_finish(FdArgs a) { occlude_finish_body(a, blockIdx.x); }

// NW waves per occlusion work item (they share its receivers and LDS rows):
// 4 (the default) or 2 per frame (FdArgs::occWaves).  The counting variant
// runs one wave per work item: each receiver then meets its triangles in
// list order with the full-row state of all earlier ones, as in the
// reference, so its work counters do not depend on the schedule.
// 8 waves per SIMD (the most a SIMD holds): the target makes the compiler
// keep the scalar registers within 8 waves' share (else 99-103 SGPRs, 7-8).
template <typename T, bool COUNT, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(8))) void k_occlude(FdArgs a) {
    occlude_body<T, COUNT, false, NW>(a, xcd_block(blockIdx.x, gridDim.x));
}
template <typename T>
__global__ __launch_bounds__(64) void k_occlude_cta(FdArgs a) {
    occlude_body<T, false, true>(a, xcd_block(blockIdx.x, gridDim.x));
}
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const uint32_t* __restrict__ d, long n,
                                                            uint64_t* __restrict__ part, const int* status) {
    scan_reduce_body(d, n, part, status, blockIdx.x);
}
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(uint32_t* __restrict__ d, long n,
                                                           const uint64_t* __restrict__ part,
                                                           unsigned long long* total, const int* status) {
    scan_apply_body(d, n, part, total, status, blockIdx.x);
}

// instance of block b: map[i] <= b < map[i+1] (wave-uniform binary search)
__device__ __forceinline__ int inst_of(const int* __restrict__ map, int n, int b) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (map[mid] <= b) lo = mid; else hi = mid;
    }
    return uniform(lo);
}

#define FD_BATCHED(NAME, BOUND, BODY)                                                          \
    __global__ __launch_bounds__(BOUND) void NAME(const FdArgs* __restrict__ inst,              \
                                                  const int* __restrict__ map, int n) {        \
        const int b = blockIdx.x;                                                              \
        const int i = inst_of(map, n, b);                                                      \
        const FdArgs a = inst[i];                                                              \
        BODY(a, b - map[i]);                                                                   \
    }
FD_BATCHED(kb_init, 64, init_body)
FD_BATCHED(kb_init_bins, 256, init_bins_body)
FD_BATCHED(kb_reduce, RED_BLOCK, reduce_body)
FD_BATCHED(kb_finalize, FD_FIN_THREADS, finalize_body)
FD_BATCHED(kb_rx_key, 256, rx_key_body)
FD_BATCHED(kb_rx_scatter, 64, rx_scatter_body)
FD_BATCHED(kb_tri_count, FD_TRB, tri_count_body)
FD_BATCHED(kb_tri_write, FD_TRB, tri_write_body)
FD_BATCHED(kb_tri_scatter_wide, FD_TRB, tri_scatter_wide_body)
FD_BATCHED(kb_tables, 1024, tables_body)
FD_BATCHED(kb_post_pairs, FD_FIN_THREADS, post_pairs_body)
FD_BATCHED(kb_tri_order, FD_ORDER_THREADS, tri_order_body)
#undef FD_BATCHED

template <typename T, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(8))) void kb_occlude(const FdArgs* __restrict__ inst, const int* __restrict__ map, int n) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int i = inst_of(map, n, b);
    const FdArgs a = inst[i];
    occlude_body<T, false, false, NW>(a, b - map[i]);
}
__global__ __launch_bounds__(256) void kb_zero_masks(const FdArgs* __restrict__ inst, const int* __restrict__ map,
                                                     int n) {
    const int b = blockIdx.x;
    const int i = inst_of(map, n, b);
    zero_masks_body(inst[i], b - map[i]);
}
__global__ __launch_bounds__(64) void kb_occlude_finish(const FdArgs* __restrict__ inst, const int* __restrict__ map,
                                                        int n) {
    const int b = blockIdx.x;
    const int i = inst_of(map, n, b);
    occlude_finish_body(inst[i], b - map[i]);
}

// which = 0: receiver key histogram, 1: pair histogram
// The three scanned arrays of a frame: 0 receiver histogram, 1 pair
// histogram, 2 per-triangle pair counts (the last two have the same total).
__device__ __forceinline__ uint32_t* scan_data(const FdArgs& a, int which) {
    return which == 0 ? a.histR : which == 1 ? a.histT : a.triOff;
}
__device__ __forceinline__ long scan_len(const FdArgs& a, int which) {
    return which == 0 ? a.nR : which == 1 ? a.nT : (long)a.ntris;
}
__global__ __launch_bounds__(SCAN_BLOCK) void kb_scan_reduce(const FdArgs* __restrict__ inst, const int* __restrict__ map,
                                                             int n, int which) {
    const int b = blockIdx.x;
    const int i = inst_of(map, n, b);
    const FdArgs a = inst[i];
    scan_reduce_body(scan_data(a, which), scan_len(a, which), a.scanPart, &a.st->status, b - map[i]);
}
__global__ __launch_bounds__(SCAN_BLOCK) void kb_scan_apply(const FdArgs* __restrict__ inst, const int* __restrict__ map,
                                                            int n, int which) {
    const int b = blockIdx.x;
    const int i = inst_of(map, n, b);
    const FdArgs a = inst[i];
    scan_apply_body(scan_data(a, which), scan_len(a, which), a.scanPart, which ? &a.st->pairs : &a.st->rxTotal,
                    &a.st->status, b - map[i]);
}

// Verification of the row-mask primitive on given (D, B) pairs (n = R - 1),
// through the same fast / exact paths as the occlusion kernel.  mode 0: the
// general form for every pair; mode 1: the fast form wherever its
// precondition |B| <= 2^100 holds, the general form elsewhere.
__global__ void k_selftest_rowmask(const float* D, const float* B, int n, int R, int mode, unsigned long long* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Edge e[3];
    float DD[3], nC[3], rB[3], dB[3];
    for (int k = 0; k < 3; ++k) {
        e[k].A = 0.0f; e[k].B = B[i]; e[k].C = 0.0f;
        e[k].nRowN = 0.0f;
        e[k].nColN = -(B[i] * (float)(unsigned)(R - 1));
        e[k].rA = __builtin_nanf("");
        e[k].rB = rcp_or_nan(B[i]);
        DD[k] = D[i];
        nC[k] = e[k].nColN; rB[k] = e[k].rB; dB[k] = B[i];
    }
    const bool fast = mode == 1 && fabsf(B[i]) <= 1.2676506e+30f;
    const float Rs = (float)R + 2.44140625e-4f;   // R + 2^-12
    const uint64_t FULL = (R >= 64) ? ~0ull : ((1ull << R) - 1ull);
    if (R <= 32) {
        out[i] = fast ? (uint64_t)three_edges_fast<uint32_t>(DD, nC, rB, dB, Rs, R)
                      : (uint64_t)(three_edges<uint32_t, false>(e, DD, (float)R, R)) & FULL;
    } else {
        out[i] = fast ? three_edges_fast<uint64_t>(DD, nC, rB, dB, Rs, R)
                      : three_edges<uint64_t, false>(e, DD, (float)R, R) & FULL;
    }
}

// Verification of the row window: for edge-plane triple i (coef[9i + 3k ..]
// = A_k, B_k, C_k) over an R x R grid, out[4i] = the rows whose three-edge
// AND (general form, clipped to the grid) is non-empty, out[4i + 1] = the
// same rows by the fast form, out[4i + 2] = row_window's rows, out[4i + 3] =
// the row-range mask (the rows the reference visits, general form).  Where
// the occlusion kernel would not take its fast branch (max |A|, |B| > 2^100,
// the only branch that applies the window) the fast rows repeat the general
// ones and the window is every row.
template <typename T>
__device__ __forceinline__ void selftest_rowwindow_one(const float* coef, int R, unsigned long long* out) {
    Edge e[3];
    float nC[3], rB[3], dB[3];
    const float nsub = (float)(unsigned)(R - 1);
    const float Rf = (float)R;
    const float Rs = Rf + 2.44140625e-4f;   // R + 2^-12
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e[k].A = coef[3 * k];
        e[k].B = coef[3 * k + 1];
        e[k].C = coef[3 * k + 2];
        e[k].nRowN = -(e[k].A * nsub);
        e[k].nColN = -(e[k].B * nsub);
        e[k].rA = rcp_or_nan(e[k].A);
        e[k].rB = rcp_or_nan(e[k].B);
        nC[k] = e[k].nColN;
        dB[k] = e[k].B;
        rB[k] = fabsf(e[k].B) >= 7.8886091e-31f ? __builtin_amdgcn_rcpf(e[k].B) : __builtin_nanf("");
    }
    const float mAB = fmaxf(fmaxf(fmaxf(fabsf(e[0].A), fabsf(e[1].A)), fabsf(e[2].A)),
                            fmaxf(fmaxf(fabsf(e[0].B), fabsf(e[1].B)), fabsf(e[2].B)));
    const bool fast = mAB <= 1.2676506e+30f;
    const T FULL = (R >= Bits<T>::width) ? ~(T)0 : (((T)1 << R) - (T)1);
    uint64_t gen = 0, fst = 0;
    for (int r = 0; r < R; ++r) {
        float D[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) D[k] = e[k].A * (float)r + e[k].C;
        if ((FULL & three_edges<T, false>(e, D, Rf, R)) != (T)0) gen |= 1ull << r;
        if (fast && three_edges_fast<T>(D, nC, rB, dB, Rs, R) != (T)0) fst |= 1ull << r;
    }
    float DN[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) DN[k] = (e[k].B > 0 ? -e[k].nColN : 0.0f) + e[k].C;
    out[0] = gen;
    out[1] = fast ? fst : gen;
    out[2] = fast ? (uint64_t)(row_window<T>(e, Rf, mAB) & FULL) : (uint64_t)FULL;
    out[3] = (uint64_t)(FULL & three_edges<T, true>(e, DN, Rf, R));
}

__global__ void k_selftest_rowwindow(const float* coef, int n, int R, unsigned long long* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (R <= 32) selftest_rowwindow_one<uint32_t>(coef + 9 * (size_t)i, R, out + 4 * (size_t)i);
    else selftest_rowwindow_one<uint64_t>(coef + 9 * (size_t)i, R, out + 4 * (size_t)i);
}

// Stage read-back of the receiver rectangles in the reference's 2-layer
// layout (Step2_1_samplerRectangleAABB.inl:122-123); verification only.
__global__ void k_debug_rect(FdArgs a, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.WH) return;
    const float4 r = a.rec[i];
    float v[8] = {FD_F32_MAX, FD_F32_MAX, FD_F32_MAX, 1.0f, -FD_F32_MAX, -FD_F32_MAX, -FD_F32_MAX, 1.0f};
    if (__float_as_uint(r.w) == FD_TAG_VALID) {
        const RxRect q = rx_rect(r, a.st, a.L);
        v[0] = q.bx; v[1] = q.by; v[2] = q.z; v[4] = q.ex; v[5] = q.ey; v[6] = q.z;
    }
    for (int k = 0; k < 8; ++k) out[(size_t)i * 8 + k] = v[k];
}

}  // namespace

// ===========================================================================
// Host-side launch sequence of one frame.
// ===========================================================================
namespace fdk {

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

static hipError_t scan_u32(uint32_t* d, long n, uint64_t* part, unsigned long long* total, const int* status,
                           hipStream_t st) {
    const int nch = cdiv(n, SCAN_CHUNK);
    k_scan_reduce<<<nch, SCAN_BLOCK, 0, st>>>(d, n, part, status);
    k_scan_apply<<<nch, SCAN_BLOCK, 0, st>>>(d, n, part, total, status);
    return hipGetLastError();
}

long scan_partials_needed(long n) { return cdiv(n, SCAN_CHUNK); }
int reduce_blocks(long n) { return cdiv(n, RED_BLOCK * RED_ITEMS); }
int scan_chunks(long n) { return cdiv(n, SCAN_CHUNK); }
int zero_blocks(const FdArgs& a) { return (int)cdiv((long)a.WH * a.L.res * a.L.wpr, ZERO_WORDS); }

hipError_t enqueue_batch(const FdBatchPlan& p, hipStream_t st) {
    const int n = p.n;
    auto L = [&](int m) { return dim3((unsigned)(p.total[m] > 0 ? p.total[m] : 1)); };
    const int nb = p.maxNb;
    const size_t binLds = nb <= FD_LDS_BINS ? (size_t)nb * FD_BIN_LDS_BYTES : 0;
    kb_init<<<L(FD_MAP_ONE), 64, 0, st>>>(p.inst, p.map[FD_MAP_ONE], n);
    kb_init_bins<<<L(FD_MAP_BINS), 256, 0, st>>>(p.inst, p.map[FD_MAP_BINS], n);
    kb_reduce<<<L(FD_MAP_RED), RED_BLOCK, 0, st>>>(p.inst, p.map[FD_MAP_RED], n);
    kb_finalize<<<L(FD_MAP_ONE), FD_FIN_THREADS, 0, st>>>(p.inst, p.map[FD_MAP_ONE], n);
    kb_rx_key<<<L(FD_MAP_RX), 256, (nb + 1) * sizeof(uint32_t) + binLds, st>>>(p.inst, p.map[FD_MAP_RX], n);
    kb_scan_reduce<<<L(FD_MAP_SCANR), SCAN_BLOCK, 0, st>>>(p.inst, p.map[FD_MAP_SCANR], n, 0);
    kb_scan_apply<<<L(FD_MAP_SCANR), SCAN_BLOCK, 0, st>>>(p.inst, p.map[FD_MAP_SCANR], n, 0);
    kb_rx_scatter<<<L(FD_MAP_RX), 64, (nb + 1) * sizeof(int), st>>>(p.inst, p.map[FD_MAP_RX], n);
    if (p.anyTris) {
        kb_tri_count<<<L(FD_MAP_TRI), FD_TRB, nb * sizeof(uint32_t) + binLds, st>>>(p.inst, p.map[FD_MAP_TRI], n);
        kb_scan_reduce<<<L(FD_MAP_SCANT), SCAN_BLOCK, 0, st>>>(p.inst, p.map[FD_MAP_SCANT], n, 1);
        kb_scan_apply<<<L(FD_MAP_SCANT), SCAN_BLOCK, 0, st>>>(p.inst, p.map[FD_MAP_SCANT], n, 1);
    }
    kb_post_pairs<<<L(FD_MAP_ONE), FD_FIN_THREADS, 0, st>>>(p.inst, p.map[FD_MAP_ONE], n);
    if (p.anyTris) {
        kb_scan_reduce<<<L(FD_MAP_SCANO), SCAN_BLOCK, 0, st>>>(p.inst, p.map[FD_MAP_SCANO], n, 2);
        kb_scan_apply<<<L(FD_MAP_SCANO), SCAN_BLOCK, 0, st>>>(p.inst, p.map[FD_MAP_SCANO], n, 2);
        kb_tri_write<<<L(FD_MAP_TRI), FD_TRB, binLds, st>>>(p.inst, p.map[FD_MAP_TRI], n);
        kb_tri_scatter_wide<<<L(FD_MAP_TRI), FD_TRB, (size_t)nb * FD_SCATTER_WIDE_LDS, st>>>(p.inst, p.map[FD_MAP_TRI], n);
    }
    kb_tables<<<L(FD_MAP_ONE), 1024, 0, st>>>(p.inst, p.map[FD_MAP_ONE], n);
    if (p.anyTris) kb_tri_order<<<L(FD_MAP_NB), FD_ORDER_THREADS, 0, st>>>(p.inst, p.map[FD_MAP_NB], n);
    if (p.total[FD_MAP_ZERO] > 0) kb_zero_masks<<<L(FD_MAP_ZERO), 256, 0, st>>>(p.occ[0], p.map[FD_MAP_ZERO], n);
    for (int c = 0; c < FD_OCC_CLASSES; ++c) {   // fd_occ_class: 32 / 64-bit rows x 2 / 4 / 8 waves
        if (p.nocc[c] == 0) continue;
        const int wpr = c / 3 + 1, nw = 2 << (c % 3);
        const size_t shm = (size_t)wpr * p.maxRes[c] * 64 * sizeof(uint32_t) + 64 * wpr * sizeof(uint32_t);
        const dim3 g = L(FD_MAP_OCC0 + c);
        const int* mp = p.map[FD_MAP_OCC0 + c];
        const FdArgs* in = p.occ[c];
        const int k = p.nocc[c];
        switch (c) {
            case 0: kb_occlude<uint32_t, 2><<<g, 64 * nw, shm, st>>>(in, mp, k); break;
            case 1: kb_occlude<uint32_t, 4><<<g, 64 * nw, shm, st>>>(in, mp, k); break;
            case 2: kb_occlude<uint32_t, 8><<<g, 64 * nw, shm, st>>>(in, mp, k); break;
            case 3: kb_occlude<uint64_t, 2><<<g, 64 * nw, shm, st>>>(in, mp, k); break;
            case 4: kb_occlude<uint64_t, 4><<<g, 64 * nw, shm, st>>>(in, mp, k); break;
            default: kb_occlude<uint64_t, 8><<<g, 64 * nw, shm, st>>>(in, mp, k); break;
        }
    }
    if (p.total[FD_MAP_FIN] > 0) kb_occlude_finish<<<L(FD_MAP_FIN), 64, 0, st>>>(p.occ[0], p.map[FD_MAP_FIN], n);
    return hipGetLastError();
}

// Number of differing 32-bit words of two device buffers (autotuner check).
__global__ __launch_bounds__(256) void k_count_diff(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                                    size_t n, unsigned long long* out) {
    int c = 0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) c += a[i] != b[i];
    c = wave_sum_i32(c);
    if (lane_id() == 0 && c) atomicAdd(out, (unsigned long long)c);
}

hipError_t count_diff(const uint32_t* a, const uint32_t* b, size_t n, unsigned long long* out, hipStream_t st) {
    hipError_t e = hipMemsetAsync(out, 0, sizeof(unsigned long long), st);
    if (e != hipSuccess) return e;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    k_count_diff<<<(unsigned)std::max<size_t>(blocks, 1), 256, 0, st>>>(a, b, n, out);
    return hipGetLastError();
}

hipError_t selftest_rowmask(const float* D, const float* B, int n, int R, int mode, unsigned long long* out) {
    k_selftest_rowmask<<<cdiv(n, 256), 256>>>(D, B, n, R, mode, out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e;
}

hipError_t selftest_rowwindow(const float* coef, int n, int R, unsigned long long* out) {
    k_selftest_rowwindow<<<cdiv(n, 256), 256>>>(coef, n, R, out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e;
}

hipError_t debug_rect(const FdArgs& a, float* out8, hipStream_t st) {
    k_debug_rect<<<cdiv(a.WH, 256), 256, 0, st>>>(a, out8);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

hipError_t enqueue_frame(const FdArgs& a, hipStream_t st, hipEvent_t* ev) {
    const int nb = 1 << (2 * a.gridLog2);
    const size_t binLds = a.binLds ? (size_t)nb * FD_BIN_LDS_BYTES : 0;
    const int nbR = a.nbR, nbV = a.nbV;
    FdState* s = a.st;
    auto mark = [&](int i) { if (ev) (void)hipEventRecord(ev[i], st); };
    mark(0);
    k_init<<<1, 64, 0, st>>>(a);
    k_init_bins<<<cdiv(nb, 256), 256, 0, st>>>(a);
    k_reduce<<<nbR + nbV, RED_BLOCK, 0, st>>>(a);
    k_finalize<<<1, FD_FIN_THREADS, 0, st>>>(a);
    mark(1);
    // receivers: key + bin statistics, stable counting sort
    k_rx_key<<<a.nblkR, 256, (nb + 1) * sizeof(uint32_t) + binLds, st>>>(a);
    hipError_t e = scan_u32(a.histR, a.nR, a.scanPart, &s->rxTotal, &s->status, st);
    if (e != hipSuccess) return e;
    k_rx_scatter<<<a.nblkR, 64, (nb + 1) * sizeof(int), st>>>(a);
    mark(2);
    // triangles: setup + pair count, stable pair scatter
    if (a.ntris > 0) {
        k_tri_count<<<a.nblkT, FD_TRB, nb * sizeof(uint32_t) + binLds, st>>>(a);
        e = scan_u32(a.histT, a.nT, a.scanPart, &s->pairs, &s->status, st);
        if (e != hipSuccess) return e;
        k_post_pairs<<<1, FD_FIN_THREADS, 0, st>>>(a);
        e = scan_u32(a.triOff, a.ntris, a.scanPart, &s->pairs, &s->status, st);
        if (e != hipSuccess) return e;
        k_tri_write<<<a.nblkT, FD_TRB, binLds, st>>>(a);
        k_tri_scatter_wide<<<a.nblkT, FD_TRB, (size_t)nb * FD_SCATTER_WIDE_LDS, st>>>(a);
    } else {
        k_post_pairs<<<1, FD_FIN_THREADS, 0, st>>>(a);
    }
    mark(3);
    k_tables<<<1, 1024, 0, st>>>(a);
    if (a.ntris > 0) k_tri_order<<<nb, FD_ORDER_THREADS, 0, st>>>(a);
    mark(4);
    // occlusion (or the early-exit fill); split lists zero the masks first
    // and count them after the merge
    const int wpr = a.L.wpr;
    if (a.segs > 1) k_zero_masks<<<zero_blocks(a), 256, 0, st>>>(a);
    if (a.occVariant == FD_OCC_BIN_QUEUE) {
        const int chunks = a.chunkSlots * a.segs;
        const size_t shmCta = (size_t)wpr * a.L.res * 64 * sizeof(uint32_t) + 64 * sizeof(uint64_t);   // + full-row flags
        if (wpr == 1) k_occlude_cta<uint32_t><<<chunks, 64, shmCta, st>>>(a);
        else k_occlude_cta<uint64_t><<<chunks, 64, shmCta, st>>>(a);
    } else {
        const FdArgs& ao = a;
        const int chunks = a.chunkSlots * a.segs;
        const size_t shm = (size_t)wpr * a.L.res * 64 * sizeof(uint32_t) + 64 * wpr * sizeof(uint32_t);
        if (a.opCount) {
            if (wpr == 1) k_occlude<uint32_t, true, 1><<<chunks, 64, shm, st>>>(ao);
            else k_occlude<uint64_t, true, 1><<<chunks, 64, shm, st>>>(ao);
        } else {
            const int nw = fd_valid_waves(a.occWaves) ? a.occWaves : fd_default_waves(wpr, a.chunkSlots);
            const int c = fd_occ_class(wpr, nw);
            switch (c) {
                case 0: k_occlude<uint32_t, false, 2><<<chunks, 64 * nw, shm, st>>>(ao); break;
                case 1: k_occlude<uint32_t, false, 4><<<chunks, 64 * nw, shm, st>>>(ao); break;
                case 2: k_occlude<uint32_t, false, 8><<<chunks, 64 * nw, shm, st>>>(ao); break;
                case 3: k_occlude<uint64_t, false, 2><<<chunks, 64 * nw, shm, st>>>(ao); break;
                case 4: k_occlude<uint64_t, false, 4><<<chunks, 64 * nw, shm, st>>>(ao); break;
                default: k_occlude<uint64_t, false, 8><<<chunks, 64 * nw, shm, st>>>(ao); break;
            }
        }
    }
    if (a.segs > 1) k_occlude_finish<<<a.chunkSlots, 64, 0, st>>>(a);
    mark(5);
    return hipGetLastError();
}

}  // namespace fdk
