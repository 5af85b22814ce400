// Launch interface of the gfx950 kernels (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fd_types.h"

#define FD_RXB 1024   // receivers per block in the key / scatter passes
#define FD_TRB 1024   // triangles per block in the count / scatter passes
#define FD_TRI_CLASSES 16      // area classes of the area walk order
// Automatic walk order of a connected mesh: the mesh order up to this many
// pairs per 256 bins (the mean list length in 16x16 bins: a small triangle
// touches about one bin at any grid), the list order above (dense scenes,
// where a receiver fills from whichever object comes first in the list)
#define FD_MESH_MAX_LIST16 7168
#define FD_WALK_KEY_BITS 5     // bits of an area class or of the no-entry value FD_TRI_CLASSES
#define FD_ORDER_CACHE 32768   // list entries whose walk keys the area-order pass keeps in LDS
#define FD_ORDER_THREADS 1024  // threads per bin in the area-order pass

// Everything one frame's kernels read, passed by value (kernarg segment).
struct FdArgs {
    FdLight L;
    float eps;
    int cull;
    int gridLog2;
    int tileLog2;         // receiver tiles of (1 << tileLog2) x (64 >> tileLog2) pixels, 3..6
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
    int nblkR;            // blocks of the receiver key / scatter passes
    int nblkT;            // blocks of the triangle count / scatter passes
    int nbR, nbV;         // k_reduce blocks over receivers / vertices
    long nR, nT;          // lengths of the receiver / pair histograms
    int* sortedId;        // WH
    FdRxRec* sortedRec;   // WH
    uint32_t* binStat;    // nb * 5 (order-preserving encodings)
    float4* triAABB;      // 2 * ntris
    FdTri* tri;           // ntris
    int* triCount;        // ntris
    uint2* triRange;      // ntris: (sx | sy << 16, ex | ey << 16)
    uint32_t* histT;      // nb * nblkT, scanned in place
    uint32_t* triOff;     // ntris: pair counts, scanned in place into triangle-major offsets
    uint32_t* pairKT;     // pairCap: triangle-major pairs, (index in block << 16) | bin; after the
                          // scatter, the bins' lists in area-class order (when the frame uses it)
    uint8_t* triClass;    // ntris: area class of each light-plane triangle's culling box (0 = largest)
    int* sortedTri;       // pairCap
    unsigned long long pairCap;
    int* rxStart;         // nb + 1
    int* rxEnd;           // nb + 1
    int* triStart;        // nb
    int* triEnd;          // nb
    int* chunkStart;      // nb + 2: work-item prefix over the bins in occlusion order (by rank)
    int* binAtRank;       // nb + 1: occlusion order of the bins, longest triangle list first
    uint64_t* scanPart;   // scan partial sums
    uint32_t* part;       // per-block partials: 8 per k_reduce block (value-major), then 1 per pair-count block
    unsigned long long* opCount;  // optional occlusion counters (FD_NOPS), or null
    // occlusion work split: chunkSlots = receiver chunk slots (64 each, an upper
    // bound of the work items); each bin's triangle list is cut into segs
    // segments, one work item per (chunk, segment); with segs > 1 the partial
    // masks merge by atomic OR into zeroed masks and a finishing pass counts them
    int chunkSlots;
    int segs;
    // pair passes: bin statistics staged in LDS (the launch allocates
    // FD_BIN_LDS_BYTES per bin besides the counters); set when nb <= FD_LDS_BINS
    int binLds;
    // occlusion kernel: FD_OCC_PER_SAMPLE (the contract) or FD_OCC_BIN_QUEUE
    int occVariant;
    // order of each bin's triangle list in the per-sample occlusion:
    // FD_TRI_ORDER_AUTO, FD_TRI_ORDER_LIST, FD_TRI_ORDER_AREA or FD_TRI_ORDER_MESH
    int triOrder;
    // waves per occlusion work item (per-sample kernel): 2, 4 or 8
    int occWaves;
    // stretches per bin list asked for (fd_set_list_split; 0 = the automatic
    // rule, which a batch that fills the GPU overrides with no split)
    int splitReq;
};

// Stretches of a bin's triangle list (fd_set_list_split): 1 = whole lists,
// up to FD_LIST_SPLIT_MAX (fd4d.h).  The automatic rule: a frame of at
// most FD_SPLIT_ITEMS work items and fewer than FD_SPLIT_MIN_WAVES waves in
// all (work items x waves per item) splits its lists into about
// FD_SPLIT_WAVES waves, at most 16 stretches; larger frames keep whole
// lists.  Each stretch merges its masks into the output by atomic OR, which
// only pays where the GPU would otherwise idle.
#define FD_SPLIT_ITEMS 2048
#define FD_SPLIT_MIN_WAVES 6144
#define FD_SPLIT_WAVES 49152

// Waves per occlusion work item (per-sample kernel) unless a frame asks for
// another count (2, 4 or 8): the waves share a work item's receivers, LDS
// rows and full-row masks and split its triangle list.  8 for 64-bit rows
// (32 KiB of rows per work item: LDS holds 4 work items per CU, so 8 waves
// each fill it) and for frames of up to FD_WAVES8_ITEMS work items (too few
// to fill the GPU with 4 waves each); 4 for larger frames of 32-bit rows.
#define FD_WAVES8_ITEMS 8192
inline int fd_default_waves(int wpr, long items) { return (wpr == 2 || items <= FD_WAVES8_ITEMS) ? 8 : 4; }
inline bool fd_valid_waves(int w) { return w == 2 || w == 4 || w == 8; }

// Occlusion classes of a batched launch: 32-bit / 64-bit rows x 2 / 4 / 8
// waves per work item (one launch each).
#define FD_OCC_CLASSES 6
inline int fd_occ_class(int wpr, int occWaves) { return (wpr - 1) * 3 + (occWaves == 2 ? 0 : occWaves == 4 ? 1 : 2); }


#define FD_LDS_BINS 1024
#define FD_BIN_LDS_BYTES 20

// Pair scatter with all 16 waves of a triangle block: 16 x 16-bit per-wave
// counts and a 32-bit block offset per bin in LDS (144 KiB at the largest
// grid, 64 x 64 = 4096 bins).
#define FD_SCATTER_WIDE_LDS 36

// Occlusion counters, in the order documented at fd_get_profile.
#define FD_NOPS FD_PROFILE_COUNTERS

// Stage boundaries recorded when profiling (events[0..FD_NSTAGES]).
#define FD_NSTAGES 5   // light plane, receiver binning, triangle pairing, tables, occlusion

#define FD_FIN_THREADS 1024   // k_finalize / k_post_pairs block size

// Block maps of a batched launch (prefix sums over instances, n + 1 each).
enum FdMap { FD_MAP_ONE, FD_MAP_BINS, FD_MAP_RED, FD_MAP_RX, FD_MAP_SCANR, FD_MAP_TRI, FD_MAP_SCANT, FD_MAP_OCC0,
             FD_MAP_OCC1, FD_MAP_OCC2, FD_MAP_OCC3, FD_MAP_OCC4, FD_MAP_OCC5, FD_MAP_ZERO, FD_MAP_FIN, FD_MAP_SCANO,
             FD_MAP_NB, FD_NMAPS };
static_assert(FD_MAP_OCC0 + FD_OCC_CLASSES == FD_MAP_ZERO, "one block map per occlusion class");
static_assert(FD_TRI_CLASSES < (1 << FD_WALK_KEY_BITS), "area classes and the no-entry value fit the key bits");

struct FdBatchPlan {
    const FdArgs* inst;       // device: instance table (n)
    const FdArgs* occ[FD_OCC_CLASSES];   // device: instances of each occlusion class (fd_occ_class), largest first
    int n, nocc[FD_OCC_CLASSES];
    const int* map[FD_NMAPS]; // device prefix arrays
    int total[FD_NMAPS];      // blocks per stage launch
    int maxNb;                // largest bin count (dynamic LDS sizes)
    int maxRes[FD_OCC_CLASSES];   // largest light grid per occlusion class
    bool anyTris;
};

namespace fdk {
hipError_t enqueue_frame(const FdArgs& a, hipStream_t st, hipEvent_t* stageEvents = nullptr);
hipError_t enqueue_batch(const FdBatchPlan& p, hipStream_t st);
long scan_partials_needed(long n);
int reduce_blocks(long n);
int scan_chunks(long n);
int zero_blocks(const FdArgs& a);   // blocks of the mask-zeroing pass
}  // namespace fdk
