// Plain-old-data records shared by the host orchestration and the gfx950
// kernels of the soft-shadow visibility path.  Nothing here depends on HIP.
#pragma once
#include <stdint.h>

#define FD_F32_MAX 3.402823466e+38f   // FD_F32_MAX, 4DZB/FDZB_Soft_Shadow/util/Util.h:59

// Light block computed on the host once per frame (FDsetAreaLight,
// 4DZB/FDZB_Soft_Shadow/api.cpp:303-400).
struct FdLight {
    float m[16];          // upRight MVP, column-major (upRightMat)
    float xd[3], yd[3];   // x_delt, y_delt: light-sample steps
    float ur[3], ul[3], dl[3], dr[3];  // corner sample positions
    float F[2];           // RectangleFactor
    int res;              // light grid side (res_w == res_h)
    int resNum;           // res * res
    int wpr;              // 32-bit words per mask row: 1 (res <= 32) or 2
    int pad;
};

// Frame exit codes: FD_EXIT_* of the public header (the order of the early
// returns in FD::FDLaunch, 4DZB/FDZB_Soft_Shadow/api.cpp:495-559).
#include "../../include/fd4d.h"

// Device-resident state of one frame (written by the kernels, read back
// only for stats).  Float reductions are kept as order-preserving uint32
// encodings so they can use integer atomics; NaN never enters them.
struct FdState {
    uint32_t rxZmaxAll;     // max z over non-NaN receivers
    uint32_t rxMin[3];      // min x', y', z over valid receivers
    uint32_t rxMaxA[2];     // max x' + F/z, y' + F/z
    uint32_t rxMaxZ;        // max z over valid receivers
    uint32_t vxZmax;        // max z over vertices
    uint32_t vxMin[3];
    uint32_t vxMaxA[2];
    uint32_t vxMaxZ;
    int validRect;          // front-facing non-NaN receivers
    int status;             // FdExit
    int validSamples;       // receivers bound to a bin
    int validTris;          // triangles that reached the light plane
    int keyOutOfRange;      // valid receivers whose key left the grid
    int capacityHit;
    unsigned long long pairs;  // total (bin, triangle) pairs
    unsigned long long rxTotal;
    int totalChunks;        // occlusion work items
    int triOrder;           // order of each bin's list in the occlusion: 1 reference (ascending id), 2 area classes, 3 mesh
    // light-plane parameters (lightPlaneParamCal, Step2_3_defineFDbuffer.inl:5-111)
    float sampleMaxZ, triMaxZ, f;
    float S[2];             // RectangleFactor / f
    float sAABB[8];         // receiver-rect AABB (minx,miny,minz,0,maxx,maxy,maxz,0)
    float mAABB[8];         // vertex-rect AABB
    float B[3], E[3];       // light-plane begin / end
    float factor[2];
    int G, log2G, nb;       // bins per side, log2, number of bins
    int pad1;
};

// Per-triangle record staged for the occlusion kernel (128 bytes, read with
// uniform scalar loads).  Fields are the reference's TriangleData
// (privateDefs.h:44-54) plus triangle-only subexpressions of the occlusion
// arithmetic, computed with the same expressions.
struct FdTri {
    float v0[3], v1[3], v2[3];
    float bminx, bminy, bmaxx, bmaxy;   // projected AABB (aabb2Dmin/max)
    float minz, maxz;
    float pad0;
    float n[3];      // cross(v1 - v0, v2 - v1)
    float e01[3];    // v0 - v1
    float e12[3];    // v1 - v2
    float e20[3];    // v2 - v0
    float pad1[4];
};

// Sorted receiver record consumed by the occlusion kernel (32 bytes).
struct FdRxRec {
    float q[3];      // world position
    float rminx, rminy, rmaxx, rmaxy;   // light-plane rectangle
    float z;         // rectangle depth
};

// Receiver tags stored in rec.w
#define FD_TAG_NONE 0u     // pos.x is NaN: no receiver
#define FD_TAG_BACK 1u     // faces away from all four light corners
#define FD_TAG_VALID 2u
