// Launch interface of the gfx950 kernels (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fd_types.h"

#define FD_RXB 1024   // receivers per block in the key / scatter passes
#define FD_TRB 256    // triangles per block in the count / scatter passes

// Everything one frame's kernels read, passed by value (kernarg segment).
struct FdArgs {
    FdLight L;
    float eps;
    int cull;
    int gridLog2;
    int pad0;
    // occluders
    const float* verts;
    int vstride;          // floats between consecutive vertices
    int nverts;
    const int* idx;       // int3 per triangle
    int ntris;
    // receivers
    const float4* pos;
    const float4* nrm;
    int W, H, WH;
    int npos;             // receiver traversal positions: 8x8 tiles covering the view
    // outputs (dense y*W + x)
    float* shadow;
    uint32_t* masks;      // WH * res * wpr words
    // workspace
    FdState* st;
    float4* rec;          // WH: (x', y', z, tag)
    uint16_t* rxKey;      // WH
    uint32_t* histR;      // (nb + 1) * nblkR, scanned in place
    int nblkR;
    int nblkT;
    int* sortedId;        // WH
    FdRxRec* sortedRec;   // WH
    uint32_t* binStat;    // nb * 5 (order-preserving encodings)
    float4* triAABB;      // 2 * ntris
    FdTri* tri;           // ntris
    int* triCount;        // ntris
    uint2* triRange;      // ntris: (sx | sy << 16, ex | ey << 16)
    uint32_t* histT;      // nb * nblkT, scanned in place
    int* sortedTri;       // pairCap
    unsigned long long pairCap;
    int* rxStart;         // nb + 1
    int* rxEnd;           // nb + 1
    int* triStart;        // nb
    int* triEnd;          // nb
    int* chunkStart;      // nb + 2
    uint64_t* scanPart;   // scan partial sums
    unsigned long long* opCount;  // optional occlusion counters (FD_NOPS), or null
};

// Occlusion counters: lane candidate tests, lane cull passes, lane row
// visits, work items, wave triangle iterations, wave row iterations.
#define FD_NOPS 6

// Stage boundaries recorded when profiling (events[0..FD_NSTAGES]).
#define FD_NSTAGES 5   // light plane, receiver binning, triangle pairing, tables, occlusion

namespace fdk {
hipError_t enqueue_frame(const FdArgs& a, hipStream_t st, hipEvent_t* stageEvents = nullptr);
long scan_partials_needed(long n);
}  // namespace fdk
