// Host side of libfd4d: the context that replaces the reference's process
// globals (FDc/FourD.cu:26-135), the light set-up (FD::FDsetAreaLight,
// FD/api.cpp:303-400), workspace management and the C ABI of include/fd4d.h.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include <algorithm>

#include "../../include/fd4d.h"
#include "fd_kernels.h"
#include "fd_types.h"
#include "fd_host.h"

namespace fdh {

// ---------------------------------------------------------------------------
// Light set-up on the host (float arithmetic in the reference's order;
// compiled with -ffp-contract=off).  glm 0.9.8 semantics for lookAt,
// infinitePerspective (fovy in radians) and mat4 products
// (4DZB/glm/gtc/matrix_transform.inl:378-393,521-546,
// 4DZB/glm/detail/type_mat4x4.inl:595-611).
// ---------------------------------------------------------------------------
struct V3 { float x, y, z; };
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 scl(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dotp(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 crs(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 nrmz(V3 v) { float inv = 1.0f / std::sqrt(dotp(v, v)); return scl(v, inv); }

void light_setup(const fd_area_light& d, FdLight& L) {
    std::memset(&L, 0, sizeof(L));
    const int W = d.res_w, H = d.res_h;
    V3 pos{d.position[0], d.position[1], d.position[2]};
    V3 view = nrmz({d.view_dir[0], d.view_dir[1], d.view_dir[2]});
    V3 up = nrmz({d.up_dir[0], d.up_dir[1], d.up_dir[2]});
    V3 right = nrmz(crs(view, up));
    V3 top = nrmz(crs(right, view));
    V3 xd = scl(right, d.right_radius * -2.0f / (float)W);
    V3 yd = scl(top, d.top_radius * -2.0f / (float)H);
    const float xf = ((float)W - 1.0f) / (float)W;
    const float yf = ((float)H - 1.0f) / (float)H;
    V3 nright = scl(right, -1.0f), ntop = scl(top, -1.0f);
    V3 ur = add(add(pos, scl(scl(right, xf), d.right_radius)), scl(scl(top, yf), d.top_radius));
    V3 ul = add(add(pos, scl(scl(nright, xf), d.right_radius)), scl(scl(top, yf), d.top_radius));
    V3 dl = add(add(pos, scl(scl(nright, xf), d.right_radius)), scl(scl(ntop, yf), d.top_radius));
    V3 dr = add(add(pos, scl(scl(right, xf), d.right_radius)), scl(scl(ntop, yf), d.top_radius));
    // view matrix at the up-right sample, looking along view_dir
    V3 center = add(ur, scl(view, 10.0f));
    V3 f = nrmz(sub(center, ur));
    V3 s = nrmz(crs(f, up));
    V3 u = crs(s, f);
    float MV[4][4] = {{s.x, u.x, -f.x, 0.0f}, {s.y, u.y, -f.y, 0.0f}, {s.z, u.z, -f.z, 0.0f},
                      {-dotp(s, ur), -dotp(u, ur), dotp(f, ur), 1.0f}};
    // infinite perspective, fovy = 120 (radians, glm 0.9.8), aspect 1, near 0.1
    const float zNear = 0.1f;
    const float range = std::tan(120.0f / 2.0f) * zNear;
    const float left = -range * 1.0f, rightp = range * 1.0f, bottom = -range, topp = range;
    float P[4][4] = {};
    P[0][0] = (2.0f * zNear) / (rightp - left);
    P[1][1] = (2.0f * zNear) / (topp - bottom);
    P[2][2] = -1.0f;
    P[2][3] = -1.0f;
    P[3][2] = -2.0f * zNear;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
            L.m[4 * j + i] = ((P[0][i] * MV[j][0] + P[1][i] * MV[j][1]) + P[2][i] * MV[j][2]) + P[3][i] * MV[j][3];
    const V3* src[6] = {&xd, &yd, &ur, &ul, &dl, &dr};
    float* dst[6] = {L.xd, L.yd, L.ur, L.ul, L.dl, L.dr};
    for (int k = 0; k < 6; ++k) { dst[k][0] = src[k]->x; dst[k][1] = src[k]->y; dst[k][2] = src[k]->z; }
    L.F[0] = d.right_radius * 2.0f * xf * P[0][0] * 0.5f;
    L.F[1] = d.top_radius * 2.0f * yf * P[1][1] * 0.5f;
    L.res = W;
    L.resNum = W * H;
    L.wpr = (W <= 32) ? 1 : 2;
}

}  // namespace fdh

using namespace fdh;

struct fd_ctx {
    int device = 0;
    // configuration (reference defaults: FDc/FourD.cu:61-62, constants.h:13)
    int cull = 1;
    float eps = 1e-6f;
    int gridLog2 = 4;
    int tileLog2 = 3;     // 8x8 receiver tiles
    int occVariant = FD_OCC_PER_SAMPLE;
    int triOrder = FD_TRI_ORDER_AUTO;
    int occWaves = 0;     // 0: fd_default_waves of the frame's row width and work items
    int listSplit = 0;    // 0: the automatic rule
    size_t maxTris = 0, maxVerts = 0;
    int W = 0, H = 0;
    std::vector<Model> models;
    const float4* pos = nullptr;
    const float4* nrm = nullptr;
    float* shadow = nullptr;
    uint32_t* masks = nullptr;
    bool lightSet = false;
    fd_area_light lightDesc{};
    FdLight light{};
    unsigned long long pairCap = 64ull << 20;
    // workspace (one arena, laid out per frame by layout())
    DevBuf arena;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    FdArgs last{};
    bool haveLast = false;
    // profiling: stage events and occlusion op counters
    bool profile = false;
    bool countOps = false;
    hipEvent_t stageEv[FD_NSTAGES + 1] = {};
};

namespace fdh {

int hip_status(hipError_t e) {
    if (e == hipSuccess) return FD_OK;
    std::fprintf(stderr, "fd4d: HIP error %d (%s)\n", (int)e, hipGetErrorString(e));
    return FD_ERR_HIP;
}

// Frame arguments that do not depend on workspace placement.
void fill_args(FdArgs& a, const Model& md, const float4* pos, const float4* nrm, int W, int H, const FdLight& L,
               float eps, int cull, int gridLog2, int tileLog2, float* shadow, uint32_t* masks,
               unsigned long long pairCap) {
    std::memset(&a, 0, sizeof(a));
    a.L = L;
    a.eps = eps;
    a.cull = cull;
    a.gridLog2 = gridLog2;
    a.tileLog2 = tileLog2;
    a.verts = md.verts;
    a.vstride = md.vstride;
    a.nverts = md.nverts;
    a.idx = md.idx;
    a.ntris = md.ntris;
    a.pos = pos;
    a.nrm = nrm;
    a.W = W;
    a.H = H;
    a.WH = W * H;
    a.shadow = shadow;
    a.masks = masks;
    a.pairCap = pairCap;
    const int nb = 1 << (2 * gridLog2);
    const int TW = 1 << tileLog2, TH = 64 >> tileLog2;
    a.npos = ((W + TW - 1) / TW) * ((H + TH - 1) / TH) * 64;
    a.nblkR = cdiv(a.npos, FD_RXB);
    a.nblkT = cdiv(std::max(md.ntris, 1), FD_TRB);
    a.nR = (long)(nb + 1) * a.nblkR;
    a.nT = (long)nb * a.nblkT;
    a.nbR = fdk::reduce_blocks(a.WH);
    a.nbV = fdk::reduce_blocks(a.nverts);
    a.chunkSlots = cdiv(a.WH, 64) + nb + 1;
    a.segs = 1;   // apply_split, once the waves per work item are known
    a.splitReq = 0;
    a.binLds = nb <= FD_LDS_BINS;
}

// Stretches per bin list: req (1..FD_LIST_SPLIT_MAX), or for 0 the
// automatic rule.  Small frames (up to FD_SPLIT_ITEMS work items and fewer
// than FD_SPLIT_MIN_WAVES waves) leave most of the GPU idle while each work
// item walks its bin's whole list: their lists split into about
// FD_SPLIT_WAVES waves in all.  Larger frames fill the GPU already, and
// splitting would only cost the merge and the early exit on full rows.  A
// frame without triangles has no lists to split.  Needs a.occWaves.
void apply_split(FdArgs& a, int req) {
    a.splitReq = req > 0 ? req : 0;
    if (a.ntris <= 0) {
        a.segs = 1;
    } else if (req > 0) {
        a.segs = std::min(req, FD_LIST_SPLIT_MAX);
    } else {
        const long waves = (long)a.chunkSlots * std::max(a.occWaves, 1);
        a.segs = (a.chunkSlots <= FD_SPLIT_ITEMS && waves < FD_SPLIT_MIN_WAVES)
                     ? (int)std::max(1L, std::min(16L, FD_SPLIT_WAVES / waves)) : 1;
    }
}

// Place the frame's workspace at `base` (nullptr: size only); returns bytes.
size_t layout(FdArgs& a, char* base, bool countOps) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> void* {
        off = (off + 255) & ~(size_t)255;
        void* p = base ? base + off : nullptr;
        off += bytes ? bytes : 16;
        return p;
    };
    const size_t WH = (size_t)a.WH, T = (size_t)std::max(a.ntris, 1);
    const int nb = 1 << (2 * a.gridLog2);
    a.st = (FdState*)take(sizeof(FdState));
    a.rec = (float4*)take(WH * sizeof(float4));
    a.rxKey = (uint16_t*)take(WH * sizeof(uint16_t));
    a.histR = (uint32_t*)take(a.nR * sizeof(uint32_t));
    a.sortedId = (int*)take(WH * sizeof(int));
    a.sortedRec = (FdRxRec*)take(WH * sizeof(FdRxRec));
    a.binStat = (uint32_t*)take((size_t)nb * 5 * sizeof(uint32_t));
    a.triAABB = (float4*)take(2 * T * sizeof(float4));
    a.tri = (FdTri*)take(T * sizeof(FdTri));
    a.triCount = (int*)take(T * sizeof(int));
    a.triRange = (uint2*)take(T * sizeof(uint2));
    a.histT = (uint32_t*)take(a.nT * sizeof(uint32_t));
    a.triOff = (uint32_t*)take(T * sizeof(uint32_t));
    a.triClass = (uint8_t*)take(T);
    a.pairKT = (uint32_t*)take((size_t)a.pairCap * sizeof(uint32_t));
    a.sortedTri = (int*)take((size_t)a.pairCap * sizeof(int));
    int* tb = (int*)take((size_t)(6 * (nb + 2)) * sizeof(int));
    a.rxStart = tb;
    a.rxEnd = tb ? tb + (nb + 2) : nullptr;
    a.triStart = tb ? tb + 2 * (nb + 2) : nullptr;
    a.triEnd = tb ? tb + 3 * (nb + 2) : nullptr;
    a.chunkStart = tb ? tb + 4 * (nb + 2) : nullptr;
    a.binAtRank = tb ? tb + 5 * (nb + 2) : nullptr;
    const long np = std::max(std::max(fdk::scan_partials_needed(a.nR), fdk::scan_partials_needed(a.nT)),
                             fdk::scan_partials_needed(T)) + 1;
    a.scanPart = (uint64_t*)take(np * sizeof(uint64_t));
    a.part = (uint32_t*)take((size_t)std::max(8 * (a.nbR + a.nbV), 3 * a.nblkT) * sizeof(uint32_t));
    a.opCount = countOps ? (unsigned long long*)take(FD_NOPS * sizeof(unsigned long long)) : nullptr;
    if (!a.masks) a.masks = (uint32_t*)take(WH * a.L.res * a.L.wpr * sizeof(uint32_t));
    return off;
}

// Size the workspace for the current view / scene / grid and fill FdArgs.
}  // namespace fdh

namespace {

// The frame's arguments from the context's settings (no workspace yet).
int frame_args(fd_ctx* c, FdArgs& a) {
    if (c->models.empty() || !c->pos || !c->nrm || !c->shadow || c->W <= 0 || c->H <= 0 || !c->lightSet)
        return FD_ERR_STATE;
    if (c->lightDesc.res_w != c->lightDesc.res_h || c->lightDesc.res_w < 1 || c->lightDesc.res_w > 64)
        return FD_ERR_UNSUPPORTED_LIGHT;
    fill_args(a, c->models[0], c->pos, c->nrm, c->W, c->H, c->light, c->eps, c->cull, c->gridLog2, c->tileLog2, c->shadow,
              c->masks, c->pairCap);
    a.occVariant = c->occVariant;
    a.triOrder = c->triOrder;
    a.occWaves = c->occWaves ? c->occWaves : fd_default_waves(a.L.wpr, a.chunkSlots);
    apply_split(a, c->listSplit);
    return FD_OK;
}

int prepare(fd_ctx* c, FdArgs& a) {
    int rc = frame_args(c, a);
    if (rc != FD_OK) return rc;
    FdArgs probe = a;
    const size_t bytes = layout(probe, nullptr, c->countOps);
    hipError_t e = c->arena.ensure(bytes);
    if (e != hipSuccess) return hip_status(e);
    layout(a, (char*)c->arena.p, c->countOps);
    return FD_OK;
}

}  // namespace

namespace fdh {

void fill_stats(const FdState& s, fd_stats* o) {
    std::memset(o, 0, sizeof(*o));
    o->exit_code = s.status;
    o->valid_receivers = s.validSamples;
    o->front_receivers = s.validRect;
    o->valid_triangles = s.validTris;
    o->pairs = s.pairs;
    o->key_out_of_range = s.keyOutOfRange;
    o->work_items = s.totalChunks;
    o->light_plane_begin[0] = s.B[0];
    o->light_plane_begin[1] = s.B[1];
    o->light_plane_end[0] = s.E[0];
    o->light_plane_end[1] = s.E[1];
    o->far_plane = s.f;
}

}  // namespace fdh

namespace {

struct M16 { float m[16]; };

__global__ void k_move_model(float* vbo, int vstride, int nv, int nthreads, M16 mm) {
    const float* m = mm.m;
    // cudaModeTransform's transform kernel (FDc/moveModel.inl:6-30)
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < (unsigned)nv; t += nthreads) {
        float* p = vbo + (size_t)t * vstride;
        const float x = p[0], y = p[1], z = p[2], w = 1.0f;
        const float rx = m[0] * x + m[4] * y + m[8] * z + m[12] * w;
        const float ry = m[1] * x + m[5] * y + m[9] * z + m[13] * w;
        const float rz = m[2] * x + m[6] * y + m[10] * z + m[14] * w;
        p[0] = rx;
        p[1] = ry;
        p[2] = rz;
    }
}

}  // namespace

namespace fdh {
hipError_t move_vertices(float* verts, int vstride, int nverts, const float m[16], hipStream_t st) {
    M16 mm;
    std::memcpy(mm.m, m, sizeof(mm.m));
    const int grid = std::min(65536, std::max(nverts, 1));
    hipLaunchKernelGGL(k_move_model, dim3(grid), dim3(256), 0, st, verts, vstride, nverts, grid * 256, mm);
    return hipGetLastError();
}
}  // namespace fdh

// ---------------------------------------------------------------------------
// Allocation-free launch (fd_launch_async_noalloc).  The frame workspace lives
// in this file-scope device array: every GPU the library runs on holds its own
// copy until the process ends.  On the host the array's name is only a
// stand-in, never a device address: fd_init resolves the current device's copy
// once with hipGetSymbolAddress and keeps that address per device, and only
// that address reaches kernels (inside FdArgs).
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[FD_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
void* g_scratch[kMaxDevices] = {};   // device address of `scratch` per device (null: not resolved)

// Resolve the current device's copy of the scratch array (once per device).
int scratch_resolve(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return FD_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (g_scratch[dev]) return FD_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_status(e);
    if (!ptr) return FD_ERR_HIP;
    g_scratch[dev] = ptr;
    return FD_OK;
}

void* scratch_of(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratchMu);
    return g_scratch[dev];
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int fd_init(fd_ctx** out) {
    if (!out) return FD_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return FD_ERR_NO_DEVICE;
    fd_ctx* c = new fd_ctx();
    (void)hipGetDevice(&c->device);
    // the allocation-free launch's scratch array on this device
    int rc = scratch_resolve(c->device);
    if (rc != FD_OK) {
        delete c;
        return rc;
    }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return FD_ERR_HIP;
    }
    *out = c;
    return FD_OK;
}

int fd_destroy(fd_ctx* c) {
    if (!c) return FD_ERR_ARG;
    (void)hipDeviceSynchronize();
    c->arena.release();
    for (hipEvent_t& ev : c->stageEv)
        if (ev) (void)hipEventDestroy(ev);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    delete c;
    return FD_OK;
}

int fd_set_backface_cull(fd_ctx* c, int enable) {
    if (!c) return FD_ERR_ARG;
    c->cull = enable ? 1 : 0;
    return FD_OK;
}

int fd_set_occlusion_variant(fd_ctx* c, int variant) {
    if (!c || (variant != FD_OCC_PER_SAMPLE && variant != FD_OCC_BIN_QUEUE)) return FD_ERR_ARG;
    c->occVariant = variant;
    return FD_OK;
}

int fd_set_list_split(fd_ctx* c, int segs) {
    if (!c || segs < 0 || segs > FD_LIST_SPLIT_MAX) return FD_ERR_ARG;
    c->listSplit = segs;
    return FD_OK;
}

int fd_set_occlusion_waves(fd_ctx* c, int waves) {
    if (!c || (waves != 0 && !fd_valid_waves(waves))) return FD_ERR_ARG;
    c->occWaves = waves;
    return FD_OK;
}

int fd_set_triangle_order(fd_ctx* c, int order) {
    if (!c || order < FD_TRI_ORDER_AUTO || order > FD_TRI_ORDER_MESH)
        return FD_ERR_ARG;
    c->triOrder = order;
    return FD_OK;
}

int fd_add_model(fd_ctx* c, const float* v, size_t vsz, int nv, const int* idx, int nt) {
    if (!c || !v || !idx || nv <= 0 || nt < 0 || vsz < 12 || vsz % 4) return FD_ERR_ARG;
    if (c->maxTris && (size_t)nt > c->maxTris) return FD_ERR_ARG;
    if (c->maxVerts && (size_t)nv > c->maxVerts) return FD_ERR_ARG;
    c->models.push_back(Model{v, (int)(vsz / 4), nv, idx, nt});
    return FD_OK;
}

int fd_clear_models(fd_ctx* c) {
    if (!c) return FD_ERR_ARG;
    c->models.clear();
    return FD_OK;
}

int fd_set_sample_buffers(fd_ctx* c, const float* p4, const float* n4) {
    if (!c || !p4 || !n4) return FD_ERR_ARG;
    c->pos = (const float4*)p4;
    c->nrm = (const float4*)n4;
    return FD_OK;
}

int fd_set_output_buffers(fd_ctx* c, float* shadow, uint32_t* masks) {
    if (!c || !shadow) return FD_ERR_ARG;
    c->shadow = shadow;
    c->masks = masks;
    return FD_OK;
}

int fd_register_res(fd_ctx* c) {
    if (!c) return FD_ERR_ARG;
    return (c->pos && c->nrm && c->shadow && !c->models.empty()) ? FD_OK : FD_ERR_STATE;
}

int fd_unregister_res(fd_ctx* c) {
    if (!c) return FD_ERR_ARG;
    c->pos = nullptr;
    c->nrm = nullptr;
    c->shadow = nullptr;
    c->masks = nullptr;
    return FD_OK;
}

int fd_set_bin_num(fd_ctx* c, size_t n) {
    if (!c) return FD_ERR_ARG;
    int l = 0;
    while ((size_t(1) << l) < n) ++l;
    if ((size_t(1) << l) != n || l < 3 || l > 6) return -1;  // the reference's -1 for a bad size
    c->gridLog2 = l;
    return FD_OK;
}

static int tile_log2(int tw) {
    for (int l = 3; l <= 6; ++l)
        if (tw == (1 << l)) return l;
    return -1;
}

int fd_set_tile_width(fd_ctx* c, int tw) {
    if (!c) return FD_ERR_ARG;
    const int l = tile_log2(tw);
    if (l < 0) return FD_ERR_ARG;
    c->tileLog2 = l;
    return FD_OK;
}

int fd_set_view(fd_ctx* c, int w, int h) {
    if (!c || w <= 0 || h <= 0 || (long)w * h > (1l << 28)) return FD_ERR_ARG;
    c->W = w;
    c->H = h;
    return FD_OK;
}

int fd_set_scenes_esp(fd_ctx* c, float eps) {
    if (!c) return FD_ERR_ARG;
    c->eps = eps;
    return FD_OK;
}

int fd_set_scene_params(fd_ctx* c, size_t maxTris, size_t maxVerts) {
    if (!c) return FD_ERR_ARG;
    c->maxTris = maxTris;
    c->maxVerts = maxVerts;
    return FD_OK;
}

int fd_set_area_light(fd_ctx* c, const fd_area_light* d) {
    if (!c || !d) return FD_ERR_ARG;
    // the light replaces the previous one even when its grid is unsupported:
    // launches then fail with FD_ERR_UNSUPPORTED_LIGHT instead of running the
    // old light (the reference's kernel reports the grid and writes nothing,
    // Step5_2_shadowCalPerSample.inl:251-254)
    c->lightDesc = *d;
    c->lightSet = true;
    if (d->res_w != d->res_h || d->res_w < 1 || d->res_w > 64) return FD_ERR_UNSUPPORTED_LIGHT;
    light_setup(*d, c->light);
    return FD_OK;
}

int fd_launch_async(fd_ctx* c, void* stream) {
    if (!c) return FD_ERR_ARG;
    FdArgs a;
    int rc = prepare(c, a);
    if (rc != FD_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    if (a.opCount) e = hipMemsetAsync(a.opCount, 0, FD_NOPS * sizeof(unsigned long long), st);  // profiling only
    if (e == hipSuccess) e = fdk::enqueue_frame(a, st, c->profile ? c->stageEv : nullptr);
    c->last = a;
    c->haveLast = true;
    return hip_status(e);
}

int fd_launch_async_noalloc(fd_ctx* c, void* stream) {
    if (!c) return FD_ERR_ARG;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return FD_ERR_HIP;
    char* base = (char*)scratch_of(dev);
    if (!base) return FD_ERR_STATE;   // no fd_init on this device
    FdArgs a;
    int rc = frame_args(c, a);
    if (rc != FD_OK) return rc;
    // the pair arena takes what the rest of the workspace leaves, up to the
    // context's pair capacity (a frame with more pairs exits with
    // FD_EXIT_CAPACITY, as with a small fd_set_pair_capacity)
    FdArgs probe = a;
    probe.pairCap = 0;
    const size_t fixed = layout(probe, nullptr, false);
    const size_t slack = 1024;   // alignment of the two pair arrays
    if (fixed + slack + 8 > (size_t)FD_NOALLOC_SCRATCH_BYTES) {
        std::fprintf(stderr, "fd4d: fd_launch_async_noalloc: the frame needs %zu bytes of workspace before its pair "
                     "arena, the scratch holds %zu\n", fixed, (size_t)FD_NOALLOC_SCRATCH_BYTES);
        return FD_ERR_SCRATCH;
    }
    a.pairCap = std::min<unsigned long long>(c->pairCap, (FD_NOALLOC_SCRATCH_BYTES - fixed - slack) / 8);
    probe = a;
    if (layout(probe, nullptr, false) > (size_t)FD_NOALLOC_SCRATCH_BYTES) return FD_ERR_SCRATCH;
    layout(a, base, false);
    hipError_t e = fdk::enqueue_frame(a, (hipStream_t)stream, nullptr);
    c->last = a;
    c->haveLast = true;
    return hip_status(e);
}

int fd_get_stats(fd_ctx* c, void* stream, fd_stats* out) {
    if (!c || !c->haveLast) return FD_ERR_STATE;
    FdState s;
    hipError_t e = hipMemcpyAsync(&s, c->last.st, sizeof(s), hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hip_status(e);
    if (out) fill_stats(s, out);
    return s.status == FD_EXIT_CAPACITY ? FD_ERR_CAPACITY : FD_OK;
}

int fd_launch(fd_ctx* c, void* stream, fd_stats* out) {
    if (!c) return FD_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    (void)hipEventRecord(c->ev0, st);
    int rc = fd_launch_async(c, stream);
    if (rc != FD_OK) return rc;
    (void)hipEventRecord(c->ev1, st);
    FdState s;
    hipError_t e = hipMemcpyAsync(&s, c->last.st, sizeof(s), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_status(e);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    if (out) {
        fill_stats(s, out);
        out->elapsed_ms = ms;
    }
    return s.status == FD_EXIT_CAPACITY ? FD_ERR_CAPACITY : FD_OK;
}

int fd_set_pair_capacity(fd_ctx* c, size_t n) {
    if (!c || n == 0 || n >= 0x7fffffffull) return FD_ERR_ARG;
    c->pairCap = n;
    return FD_OK;
}

int fd_move_model(fd_ctx* c, const float m[16], void* stream) {
    if (!c || !m) return FD_ERR_ARG;
    if (c->models.empty()) return FD_ERR_STATE;
    const Model& md = c->models[0];
    hipError_t e = fdh::move_vertices((float*)md.verts, md.vstride, md.nverts, m, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    return hip_status(e);
}

int fd_set_profiling(fd_ctx* c, int stage_events, int count_ops) {
    if (!c) return FD_ERR_ARG;
    if (stage_events && !c->stageEv[0])
        for (hipEvent_t& ev : c->stageEv)
            if (hipEventCreate(&ev) != hipSuccess) return FD_ERR_HIP;
    c->profile = stage_events != 0;
    c->countOps = count_ops != 0;
    return FD_OK;
}

int fd_get_profile(fd_ctx* c, float* stage_ms, int nstages, unsigned long long* ops) {
    if (!c || !c->haveLast) return FD_ERR_STATE;
    if (stage_ms && c->profile) {
        hipError_t e = hipEventSynchronize(c->stageEv[FD_NSTAGES]);
        if (e != hipSuccess) return hip_status(e);
        for (int i = 0; i < nstages && i < FD_NSTAGES; ++i)
            (void)hipEventElapsedTime(stage_ms + i, c->stageEv[i], c->stageEv[i + 1]);
    }
    if (ops) {
        if (!c->last.opCount) return FD_ERR_STATE;
        hipError_t e = hipDeviceSynchronize();   // the frame may run on a non-blocking stream
        if (e == hipSuccess)
            e = hipMemcpy(ops, c->last.opCount, FD_NOPS * sizeof(unsigned long long), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_status(e);
    }
    return FD_OK;
}

// ---- device buffers and streams for callers without their own HIP plumbing
int fd_device_count(int* n) {
    if (!n) return FD_ERR_ARG;
    *n = 0;
    if (hipGetDeviceCount(n) != hipSuccess) { *n = 0; return FD_ERR_NO_DEVICE; }
    return *n > 0 ? FD_OK : FD_ERR_NO_DEVICE;
}
int fd_dev_alloc(size_t bytes, void** out) {
    if (!out) return FD_ERR_ARG;
    *out = nullptr;
    return hip_status(hipMalloc(out, bytes ? bytes : 16));
}
int fd_dev_free(void* p) { return p ? hip_status(hipFree(p)) : FD_OK; }
int fd_memcpy_htod(void* dst, const void* src, size_t bytes) {
    if (!dst || (!src && bytes)) return FD_ERR_ARG;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}
int fd_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    if (!dst || (!src && bytes)) return FD_ERR_ARG;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}
int fd_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || (!src && bytes)) return FD_ERR_ARG;
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
}
int fd_stream_create(void** out) {
    if (!out) return FD_ERR_ARG;
    return hip_status(hipStreamCreateWithFlags((hipStream_t*)out, hipStreamNonBlocking));
}
int fd_stream_destroy(void* s) { return s ? hip_status(hipStreamDestroy((hipStream_t)s)) : FD_OK; }
int fd_stream_sync(void* s) { return hip_status(hipStreamSynchronize((hipStream_t)s)); }
int fd_device_sync(void) { return hip_status(hipDeviceSynchronize()); }

int fd_mask_words(int res) {
    if (res < 1 || res > 64) return FD_ERR_ARG;
    return res * (res <= 32 ? 1 : 2);
}

const char* fd_status_string(int code) {
    switch (code) {
        case FD_OK: return "ok";
        case FD_ERR_ARG: return "invalid argument";
        case FD_ERR_STATE: return "missing set-up call";
        case FD_ERR_HIP: return "HIP runtime error";
        case FD_ERR_CAPACITY: return "pair arena too small";
        case FD_ERR_UNSUPPORTED_LIGHT: return "unsupported light grid";
        case FD_ERR_NO_DEVICE: return "no GPU device";
        case FD_ERR_SCRATCH: return "frame workspace larger than the static scratch";
        default: return "unknown";
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched form: many frames (views / lights of uneven sizes) over the
// context's scene, one launch per stage for all of them, optionally captured
// into a HIP graph.
// ---------------------------------------------------------------------------
namespace fdh {

size_t plan_bytes(int n) {
    // instance table + two occlusion tables + maps (n + 1 ints each), 256-aligned
    return ((sizeof(FdArgs) * (size_t)(2 * n) + 255) & ~(size_t)255) + (size_t)FD_NMAPS * (n + 1) * sizeof(int) + 256;
}

size_t build_plan(const std::vector<FdArgs>& args, bool anyTris, char* devBase, std::vector<char>& image,
                  FdBatchPlan& P) {
    const int n = (int)args.size();
    P = FdBatchPlan{};
    P.n = n;
    P.maxNb = 0;
    P.anyTris = anyTris;
    std::vector<int> occIdx[FD_OCC_CLASSES];
    for (int i = 0; i < n; ++i) {
        P.maxNb = std::max(P.maxNb, 1 << (2 * args[i].gridLog2));
        occIdx[fd_occ_class(args[i].L.wpr, args[i].occWaves)].push_back(i);
    }
    for (int cl = 0; cl < FD_OCC_CLASSES; ++cl) {  // largest work first: shorter tail
        auto cost = [&](int i) { return (double)args[i].WH * args[i].L.res * args[i].L.res; };
        std::stable_sort(occIdx[cl].begin(), occIdx[cl].end(), [&](int x, int y) { return cost(x) > cost(y); });
        P.nocc[cl] = (int)occIdx[cl].size();
        P.maxRes[cl] = 1;
        for (int i : occIdx[cl]) P.maxRes[cl] = std::max(P.maxRes[cl], args[i].L.res);
    }
    const size_t argBytes = (sizeof(FdArgs) * (size_t)(2 * n) + 255) & ~(size_t)255;
    image.assign(argBytes + (size_t)FD_NMAPS * (n + 1) * sizeof(int), 0);
    FdArgs* hArgs = (FdArgs*)image.data();
    // a batch that fills the GPU with whole-list work items does not split the
    // triangle lists of its small instances (the merge would only cost)
    long slots = 0;
    for (int i = 0; i < n; ++i) slots += args[i].chunkSlots;
    for (int i = 0; i < n; ++i) {
        hArgs[i] = args[i];
        if (slots > FD_SPLIT_ITEMS && !hArgs[i].splitReq) hArgs[i].segs = 1;
        hArgs[i].binLds = P.maxNb <= FD_LDS_BINS;   // the batched launch sizes LDS for the largest grid
    }
    int k = n;
    for (int cl = 0; cl < FD_OCC_CLASSES; ++cl)
        for (int i : occIdx[cl]) hArgs[k++] = hArgs[i];
    int* hMaps = (int*)(image.data() + argBytes);
    const FdArgs* dArgs = (const FdArgs*)devBase;
    const int* dMaps = (const int*)(devBase + argBytes);
    P.inst = dArgs;
    // the occlusion table holds the classes one after the other
    for (int cl = 0, base = n; cl < FD_OCC_CLASSES; base += P.nocc[cl++]) P.occ[cl] = dArgs + base;
    for (int m = 0; m < FD_NMAPS; ++m) {
        const bool occCls = m >= FD_MAP_OCC0 && m < FD_MAP_OCC0 + FD_OCC_CLASSES;
        const bool occ = occCls || m == FD_MAP_ZERO || m == FD_MAP_FIN;
        const std::vector<int>* order = occCls ? &occIdx[m - FD_MAP_OCC0] : nullptr;
        std::vector<int> all;   // zero / finish maps: every instance, in occlusion-table order
        if (m == FD_MAP_ZERO || m == FD_MAP_FIN) {
            for (int cl = 0; cl < FD_OCC_CLASSES; ++cl) all.insert(all.end(), occIdx[cl].begin(), occIdx[cl].end());
            order = &all;
        }
        const int cnt = occ ? (int)order->size() : n;
        int* mp = hMaps + (size_t)m * (n + 1);
        P.map[m] = dMaps + (size_t)m * (n + 1);
        int acc = 0;
        mp[0] = 0;
        for (int q = 0; q < cnt; ++q) {
            const FdArgs& a = hArgs[occ ? (*order)[q] : q];
            const int nb = 1 << (2 * a.gridLog2);
            int blocks = 1;
            switch (m) {
                case FD_MAP_ONE: blocks = 1; break;
                case FD_MAP_BINS: blocks = cdiv(nb, 256); break;
                case FD_MAP_RED: blocks = a.nbR + a.nbV; break;
                case FD_MAP_RX: blocks = a.nblkR; break;
                case FD_MAP_SCANR: blocks = fdk::scan_chunks(a.nR); break;
                case FD_MAP_TRI: blocks = a.nblkT; break;
                case FD_MAP_SCANT: blocks = fdk::scan_chunks(a.nT); break;
                case FD_MAP_SCANO: blocks = fdk::scan_chunks(std::max(a.ntris, 1)); break;
                case FD_MAP_NB: blocks = nb; break;
                case FD_MAP_ZERO: blocks = a.segs > 1 ? fdk::zero_blocks(a) : 0; break;
                case FD_MAP_FIN: blocks = a.segs > 1 ? a.chunkSlots : 0; break;
                default: blocks = a.chunkSlots * a.segs; break;
            }
            acc += blocks;
            mp[q + 1] = acc;
        }
        P.total[m] = acc;
    }
    return image.size();
}

}  // namespace fdh

struct fd_batch {
    fd_ctx* ctx = nullptr;
    int n = 0;
    std::vector<FdArgs> args;
    DevBuf arena, tables;
    FdBatchPlan plan{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" int fd_batch_create(fd_ctx* c, const fd_instance* in, int n, fd_batch** out) {
    if (!c || !in || n <= 0 || !out) return FD_ERR_ARG;
    if (c->models.empty()) return FD_ERR_STATE;
    const Model& md = c->models[0];
    fd_batch* b = new fd_batch();
    b->ctx = c;
    b->n = n;
    b->args.resize(n);
    std::vector<size_t> off(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        const fd_instance& q = in[i];
        if (!q.d_position4 || !q.d_normal4 || !q.d_shadow || q.width <= 0 || q.height <= 0 ||
            (long)q.width * q.height > (1l << 28)) { delete b; return FD_ERR_ARG; }
        if (q.light.res_w != q.light.res_h || q.light.res_w < 1 || q.light.res_w > 64) {
            delete b;
            return FD_ERR_UNSUPPORTED_LIGHT;
        }
        int g = c->gridLog2;
        if (q.bins_per_side) {
            int l = 0;
            while ((1 << l) < q.bins_per_side) ++l;
            if ((1 << l) != q.bins_per_side || l < 3 || l > 6) { delete b; return FD_ERR_ARG; }
            g = l;
        }
        int tl = c->tileLog2;
        if (q.tile_width) {
            tl = tile_log2(q.tile_width);
            if (tl < 0) { delete b; return FD_ERR_ARG; }
        }
        FdLight L;
        light_setup(q.light, L);
        if (q.occ_waves != 0 && !fd_valid_waves(q.occ_waves)) { delete b; return FD_ERR_ARG; }
        if (q.tri_order < FD_TRI_ORDER_AUTO || q.tri_order > FD_TRI_ORDER_MESH) {
            delete b;
            return FD_ERR_ARG;
        }
        if (q.list_split < 0 || q.list_split > FD_LIST_SPLIT_MAX) { delete b; return FD_ERR_ARG; }
        const unsigned long long cap = q.max_pairs ? q.max_pairs : (16ull << 20);
        fill_args(b->args[i], md, (const float4*)q.d_position4, (const float4*)q.d_normal4, q.width, q.height, L,
                  c->eps, c->cull, g, tl, q.d_shadow, q.d_masks, cap);
        b->args[i].triOrder = q.tri_order ? q.tri_order : c->triOrder;
        b->args[i].occWaves = q.occ_waves ? q.occ_waves : c->occWaves ? c->occWaves
                                                     : fd_default_waves(L.wpr, b->args[i].chunkSlots);
        apply_split(b->args[i], q.list_split ? q.list_split : c->listSplit);   // after occWaves
        FdArgs probe = b->args[i];
        off[i + 1] = off[i] + ((layout(probe, nullptr, false) + 255) & ~(size_t)255);
    }
    hipError_t e = b->arena.ensure(off[n]);
    if (e != hipSuccess) { delete b; return hip_status(e); }
    for (int i = 0; i < n; ++i) layout(b->args[i], (char*)b->arena.p + off[i], false);
    e = b->tables.ensure(fdh::plan_bytes(n));
    if (e == hipSuccess) {
        std::vector<char> image;
        fdh::build_plan(b->args, md.ntris > 0, (char*)b->tables.p, image, b->plan);
        e = hipMemcpy(b->tables.p, image.data(), image.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) { delete b; return hip_status(e); }
    *out = b;
    return FD_OK;
}

extern "C" int fd_batch_capture(fd_batch* b, void* stream) {
    if (!b) return FD_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    if (!st) return FD_ERR_ARG;  // capture needs a created stream
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_status(e);
    hipError_t e2 = fdk::enqueue_batch(b->plan, st);
    e = hipStreamEndCapture(st, &b->graph);
    if (e2 != hipSuccess) return hip_status(e2);
    if (e != hipSuccess) return hip_status(e);
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    return hip_status(e);
}

extern "C" int fd_batch_launch(fd_batch* b, void* stream) {
    if (!b) return FD_ERR_ARG;
    if (b->exec) return hip_status(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return hip_status(fdk::enqueue_batch(b->plan, (hipStream_t)stream));
}

extern "C" int fd_batch_get_stats(fd_batch* b, int i, fd_stats* out) {
    if (!b || i < 0 || i >= b->n || !out) return FD_ERR_ARG;
    hipError_t e = hipDeviceSynchronize();
    FdState s;
    if (e == hipSuccess) e = hipMemcpy(&s, b->args[i].st, sizeof(s), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_status(e);
    fill_stats(s, out);
    return s.status == FD_EXIT_CAPACITY ? FD_ERR_CAPACITY : FD_OK;
}

extern "C" int fd_batch_destroy(fd_batch* b) {
    if (!b) return FD_ERR_ARG;
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->arena.release();
    b->tables.release();
    delete b;
    return FD_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: bin grid and receiver tile shape of one instance class.  Each
// candidate runs as a one-instance batch (the batched form's own launch
// sequence) into private output buffers, timed with events on a private
// stream, and is compared word by word with the default layout's output.
// ---------------------------------------------------------------------------
namespace fdk {
hipError_t count_diff(const uint32_t* a, const uint32_t* b, size_t n, unsigned long long* out, hipStream_t st);
}

extern "C" int fd_autotune(fd_ctx* c, const fd_instance* inst, fd_tune_entry* cands, int n, int reps, int* best) {
    if (!c || !inst || !cands || n <= 0 || !best || reps <= 0 || reps > 100) return FD_ERR_ARG;
    if (c->models.empty()) return FD_ERR_STATE;
    const int R = inst->light.res_w;
    if (R < 1 || R > 64 || inst->light.res_h != R) return FD_ERR_UNSUPPORTED_LIGHT;
    if (inst->width <= 0 || inst->height <= 0 || (long)inst->width * inst->height > (1l << 28)) return FD_ERR_ARG;
    const size_t WH = (size_t)inst->width * inst->height;
    const size_t words = WH + WH * (size_t)fd_mask_words(R);   // shadow values, then masks
    DevBuf ref, out, cnt;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipError_t e = ref.ensure(words * 4);
    if (e == hipSuccess) e = out.ensure(words * 4);
    if (e == hipSuccess) e = cnt.ensure(sizeof(unsigned long long));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ev0);
    if (e == hipSuccess) e = hipEventCreate(&ev1);
    // one layout: returns a status, the median device time and the outputs
    auto run = [&](int bins, int tw, int waves, int order, int split, uint32_t* dst, float* ms) -> int {
        fd_instance q = *inst;
        q.bins_per_side = bins;
        q.tile_width = tw;
        q.occ_waves = waves;
        q.tri_order = order;
        q.list_split = split;
        q.d_shadow = (float*)dst;
        q.d_masks = dst + WH;
        fd_batch* b = nullptr;
        int rc = fd_batch_create(c, &q, 1, &b);
        if (rc != FD_OK) return rc;
        hipError_t he = fdk::enqueue_batch(b->plan, st);
        std::vector<float> t;
        for (int r = 0; r < reps && he == hipSuccess; ++r) {
            he = hipEventRecord(ev0, st);
            if (he == hipSuccess) he = fdk::enqueue_batch(b->plan, st);
            if (he == hipSuccess) he = hipEventRecord(ev1, st);
            if (he == hipSuccess) he = hipEventSynchronize(ev1);
            float x = 0.0f;
            if (he == hipSuccess) he = hipEventElapsedTime(&x, ev0, ev1);
            t.push_back(x);
        }
        fd_stats s{};
        if (he == hipSuccess) rc = fd_batch_get_stats(b, 0, &s);
        else rc = hip_status(he);
        fd_batch_destroy(b);
        if (rc != FD_OK) return rc;
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
        return FD_OK;
    };
    int rc = hip_status(e);
    float refMs = 0.0f;
    if (rc == FD_OK) rc = run(16, 8, 0, 0, 0, (uint32_t*)ref.p, &refMs);
    *best = -1;
    for (int i = 0; i < n && rc == FD_OK; ++i) {
        float ms = 0.0f;
        rc = run(cands[i].bins_per_side, cands[i].tile_width, cands[i].occ_waves, cands[i].tri_order,
                 cands[i].list_split, (uint32_t*)out.p, &ms);
        if (rc == FD_ERR_CAPACITY) {   // this grid needs more pairs than the instance's arena
            cands[i].ms = -2.0f;
            rc = FD_OK;
            continue;
        }
        if (rc != FD_OK) break;
        unsigned long long diff = 0;
        e = fdk::count_diff((const uint32_t*)ref.p, (const uint32_t*)out.p, words, (unsigned long long*)cnt.p, st);
        if (e == hipSuccess) e = hipMemcpyAsync(&diff, cnt.p, sizeof(diff), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) { rc = hip_status(e); break; }
        cands[i].ms = diff ? -1.0f : ms;
        if (!diff && (*best < 0 || ms < cands[*best].ms)) *best = i;
    }
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

// ---------------------------------------------------------------------------
// Stage read-back (verification only; synchronises)
// ---------------------------------------------------------------------------
namespace fdk {
hipError_t debug_rect(const FdArgs& a, float* out8, hipStream_t st);
hipError_t selftest_rowmask(const float* D, const float* B, int n, int R, int mode, unsigned long long* out);
hipError_t selftest_rowwindow(const float* coef, int n, int R, unsigned long long* out);
}

extern "C" int fd_selftest_rowwindow(const float* hcoef, int n, int R, unsigned long long* hout) {
    if (!hcoef || !hout || n <= 0 || R < 1 || R > 64) return FD_ERR_ARG;
    float* dC = nullptr;
    unsigned long long* dO = nullptr;
    hipError_t e = hipMalloc(&dC, (size_t)n * 9 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dO, (size_t)n * 4 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemcpy(dC, hcoef, (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = fdk::selftest_rowwindow(dC, n, R, dO);
    if (e == hipSuccess) e = hipMemcpy(hout, dO, (size_t)n * 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (dC) (void)hipFree(dC);
    if (dO) (void)hipFree(dO);
    return hip_status(e);
}

extern "C" int fd_selftest_rowmask(const float* hD, const float* hB, int n, int R, int mode, unsigned long long* hout) {
    if (!hD || !hB || !hout || n <= 0 || R < 1 || R > 64 || mode < 0 || mode > 1) return FD_ERR_ARG;
    float *dD = nullptr, *dB = nullptr;
    unsigned long long* dO = nullptr;
    hipError_t e = hipMalloc(&dD, n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dB, n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dO, n * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemcpy(dD, hD, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB, hB, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = fdk::selftest_rowmask(dD, dB, n, R, mode, dO);
    if (e == hipSuccess) e = hipMemcpy(hout, dO, n * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (dD) (void)hipFree(dD);
    if (dB) (void)hipFree(dB);
    if (dO) (void)hipFree(dO);
    return hip_status(e);
}

extern "C" int fd_debug_read(fd_ctx* c, const char* name, void* dst, size_t maxb, size_t* bytes) {
    if (!c || !name) return FD_ERR_ARG;
    if (!c->haveLast) return FD_ERR_STATE;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_status(e);
    const FdArgs& a = c->last;
    FdState s;
    (void)hipMemcpy(&s, a.st, sizeof(s), hipMemcpyDeviceToHost);
    const int nb = 1 << (2 * a.gridLog2);
    std::vector<char> out;
    auto get = [&](const void* src, size_t nbytes) {
        out.resize(nbytes);
        if (nbytes) e = hipMemcpy(out.data(), src, nbytes, hipMemcpyDeviceToHost);
    };
    const std::string n(name);
    // the receiver tables exist once the receivers are sorted (frames that
    // reach the pairing stage)
    const bool pipelineRan = s.status == FD_EXIT_OK || s.status == FD_EXIT_NO_PAIRS || s.status == FD_EXIT_CAPACITY;
    if (n == "state") {
        out.resize(sizeof(s));
        std::memcpy(out.data(), &s, sizeof(s));
    } else if (n == "light") {
        out.resize(sizeof(FdLight));
        std::memcpy(out.data(), &a.L, sizeof(FdLight));
    } else if (n == "rect") {
        float* d = nullptr;
        e = hipMalloc(&d, (size_t)a.WH * 8 * sizeof(float));
        if (e == hipSuccess) e = fdk::debug_rect(a, d, nullptr);
        if (e == hipSuccess) get(d, (size_t)a.WH * 8 * sizeof(float));
        if (d) (void)hipFree(d);
    } else if (n == "key") {
        std::vector<uint16_t> k(a.WH);
        e = hipMemcpy(k.data(), a.rxKey, a.WH * sizeof(uint16_t), hipMemcpyDeviceToHost);
        out.resize(a.WH * sizeof(int));
        int* o = (int*)out.data();
        for (int i = 0; i < a.WH; ++i) o[i] = k[i];
    } else if (n == "sorted_id") {
        // The kernels order receivers inside a bin by 8x8 pixel tiles (lane
        // coherence); the reference's stable sort orders them by ascending
        // id.  The export canonicalises each key segment to ascending ids, so
        // it compares as the reference's permutation; the per-key sets are
        // exactly what the kernels produced.
        get(a.sortedId, (size_t)a.WH * sizeof(int));
        if (e == hipSuccess && pipelineRan) {
            std::vector<int> rs(nb + 1), re(nb + 1);
            e = hipMemcpy(rs.data(), a.rxStart, (nb + 1) * sizeof(int), hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(re.data(), a.rxEnd, (nb + 1) * sizeof(int), hipMemcpyDeviceToHost);
            int* ids = (int*)out.data();
            for (int k = 0; k <= nb && e == hipSuccess; ++k)
                if (rs[k] < re[k] && re[k] <= a.WH) std::sort(ids + rs[k], ids + re[k]);
        }
    } else if (n == "bin_start" || n == "bin_end" || n == "bin_tri_start" || n == "bin_tri_end") {
        const bool tri = n.find("tri") != std::string::npos;
        const int cnt = tri ? nb : nb + 1;
        std::vector<int> st(cnt), en(cnt);
        if (pipelineRan && !(tri && s.status != FD_EXIT_OK)) {
            e = hipMemcpy(st.data(), tri ? a.triStart : a.rxStart, cnt * sizeof(int), hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(en.data(), tri ? a.triEnd : a.rxEnd, cnt * sizeof(int), hipMemcpyDeviceToHost);
        }
        // FindStart_kernel leaves absent keys at (0, 0) (binRaster.inl:38-45)
        for (int k = 0; k < cnt; ++k)
            if (st[k] == en[k]) st[k] = en[k] = 0;
        out.resize(cnt * sizeof(int));
        std::memcpy(out.data(), (n == "bin_start" || n == "bin_tri_start") ? st.data() : en.data(), cnt * sizeof(int));
    } else if (n == "bin_maxz" || n == "bin_minrange") {
        std::vector<uint32_t> raw((size_t)nb * 5);
        e = hipMemcpy(raw.data(), a.binStat, raw.size() * 4, hipMemcpyDeviceToHost);
        auto dec = [](uint32_t v) {
            uint32_t u = (v & 0x80000000u) ? (v & 0x7fffffffu) : ~v;
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        };
        if (n == "bin_maxz") {
            out.resize(nb * sizeof(float));
            float* o = (float*)out.data();
            for (int b = 0; b < nb; ++b) o[b] = dec(raw[5 * b]);
        } else {
            out.resize(nb * 4 * sizeof(float));
            float* o = (float*)out.data();
            for (int b = 0; b < nb; ++b)
                for (int k = 0; k < 4; ++k) o[4 * b + k] = dec(raw[5 * b + 1 + k]);
        }
    } else if (n == "tri_aabb") {
        get(a.triAABB, (size_t)a.ntris * 2 * sizeof(float4));
    } else if (n == "tri_data") {
        get(a.tri, (size_t)a.ntris * sizeof(FdTri));
    } else if (n == "tri_count") {
        get(a.triCount, (size_t)a.ntris * sizeof(int));
    } else if (n == "tri_range") {
        get(a.triRange, (size_t)a.ntris * sizeof(uint2));
    } else if (n == "pair_tri") {
        const size_t np = std::min<unsigned long long>(s.pairs, a.pairCap);
        get(a.sortedTri, np * sizeof(int));
    } else if (n == "pair_tri_walk") {
        const size_t np = std::min<unsigned long long>(s.pairs, a.pairCap);
        get(s.status == FD_EXIT_OK && s.triOrder != FD_TRI_ORDER_LIST ? (const void*)a.pairKT : (const void*)a.sortedTri,
            np * sizeof(int));
    } else if (n == "tri_class") {
        get(a.triClass, (size_t)a.ntris);
    } else {
        return FD_ERR_ARG;
    }
    if (e != hipSuccess) return hip_status(e);
    if (bytes) *bytes = out.size();
    if (dst && maxb) std::memcpy(dst, out.data(), std::min(maxb, out.size()));
    return FD_OK;
}
