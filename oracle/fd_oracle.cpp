// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A host C++ restatement of galaxy-wolf/4DZB's soft-shadow visibility path
// (FD::FDLaunch and the stages it drives), written stage by stage in the
// reference's expression order so that every intermediate can be compared
// bit for bit with the HIP implementation.  It is used only by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
// checker (or as the timed CPU baseline), never as part of the product path.
//
// Build: g++ -O2 -ffp-contract=off -fno-fast-math -fopenmp (see oracle/Makefile).
// IEEE single precision throughout, no FMA contraction, denormals preserved.
//
// Pinned semantics where the reference is undefined, racy or unportable
// (each is counted in the stats record):
//   * float->int conversion saturates, NaN -> 0   (PTX cvt.rzi semantics)
//   * (T)1 << k with k outside [0, width) is 0     (PTX shl clamps)
//   * row / column masks are clipped to the light grid: bits >= W dropped,
//     rows >= H ignored                            (reference writes them into
//     other threads' shared memory / counts them in popc)
//   * receiver keys outside [0, numBins) are mapped to the invalid key
//   * receivers sorted under the invalid key are never occluded
//   * the pair writer uses the count pass's (LP-clamped) bin test, so
//     written == counted
//   * outputs are dense y*W + x (the reference writes at y << ceil(log2 W) | x)
//   * min/max ignore NaN (fminf/fmaxf) and order -0 below +0
//
// Parity status per function: see the "pins" list in tests/test_oracle.py.
// One decision is parity unpinned: the bin-queue (CTA) variant's
// group-level triangle-plane decision (occludeCTA below) has no check
// beyond the GPU kernel agreeing with this restatement -- no ray-cast
// geometry in which it changes a bit exists; its row primitives and masks
// are pinned (tests/test_oracle_cta.py).  That variant is outside the
// bit-exact contract of the per-sample path.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

const float F32_MAX = 3.402823466e+38f;  // FD_F32_MAX, FD/util/Util.h:59

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

// helper_math.h (CUDA Samples) float3 operators, as used by the reference.
inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// host normalize: v * rsqrtf(dot(v,v)), host rsqrtf = 1/sqrtf (helper_math.h non-CUDA path);
// identical to glm's normalize = v * inversesqrt(dot(v,v)), inversesqrt = 1/sqrt
// (4DZB/glm/detail/func_geometric.inl:88-95, func_exponential.inl:130-133).
inline f3 normalize(f3 v) { float inv = 1.0f / std::sqrt(dot(v, v)); return v * inv; }

// min/max: NaN-ignoring (fminf/fmaxf), with -0 < +0 so results are order-free.
inline float fmin_(float a, float b) {
    if (std::isnan(a)) return b;
    if (std::isnan(b)) return a;
    if (a == b) return std::signbit(a) ? a : b;
    return a < b ? a : b;
}
inline float fmax_(float a, float b) {
    if (std::isnan(a)) return b;
    if (std::isnan(b)) return a;
    if (a == b) return std::signbit(a) ? b : a;
    return a > b ? a : b;
}

// float -> int32, round toward zero, saturating, NaN -> 0 (PTX cvt.rzi.s32.f32).
inline int cvt_rzi(float x) {
    if (std::isnan(x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int)0x80000000;
    return (int)x;
}
// float -> uint16, saturating, NaN -> 0 (PTX cvt.rzi.u16.f32).
inline int cvt_rzi_u16(float x) {
    if (std::isnan(x) || x <= 0.0f) return 0;
    if (x >= 65535.0f) return 65535;
    return (int)x;
}

// ---------------------------------------------------------------------------
// Light  (FD/AreaLightDesc.h:3-13, FDc/privateDefs.h:66-101)
// ---------------------------------------------------------------------------
struct LightDesc {
    float rightRadius, topRadius;
    float position[3], viewDir[3], upDir[3];
    int resW, resH;
};

struct Light {
    f3 position, viewDir, upDir, rightDir, topDir, x_delt, y_delt;
    f3 upRight, upLeft, downLeft, downRight;
    float rightRadius, topRadius;
    int resW, resH, resNum;
    float upRightMat[16];
    f2 RectangleFactor;
};

// glm 0.9.8 column-major 4x4 (m[col][row]).
struct M4 { float m[4][4]; };

// glm::lookAtRH, 4DZB/glm/gtc/matrix_transform.inl:521-546
static M4 lookAtRH(f3 eye, f3 center, f3 up) {
    f3 f = normalize(center - eye);
    f3 s = normalize(cross(f, up));
    f3 u = cross(s, f);
    M4 R{};
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) R.m[c][r] = (c == r) ? 1.0f : 0.0f;
    R.m[0][0] = s.x; R.m[1][0] = s.y; R.m[2][0] = s.z;
    R.m[0][1] = u.x; R.m[1][1] = u.y; R.m[2][1] = u.z;
    R.m[0][2] = -f.x; R.m[1][2] = -f.y; R.m[2][2] = -f.z;
    R.m[3][0] = -dot(s, eye);
    R.m[3][1] = -dot(u, eye);
    R.m[3][2] = dot(f, eye);
    return R;
}

// glm::infinitePerspectiveRH, 4DZB/glm/gtc/matrix_transform.inl:378-393
// (glm 0.9.8: fovy in radians, so the reference's 120.0f is 120 rad).
static M4 infinitePerspectiveRH(float fovy, float aspect, float zNear) {
    float range = std::tan(fovy / 2.0f) * zNear;
    float left = -range * aspect, right = range * aspect, bottom = -range, top = range;
    M4 R{};
    R.m[0][0] = (2.0f * zNear) / (right - left);
    R.m[1][1] = (2.0f * zNear) / (top - bottom);
    R.m[2][2] = -1.0f;
    R.m[2][3] = -1.0f;
    R.m[3][2] = -2.0f * zNear;
    return R;
}

// glm mat4 * mat4, 4DZB/glm/detail/type_mat4x4.inl:595-611
static M4 mul(const M4& A, const M4& B) {
    M4 R{};
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
            R.m[j][i] = ((A.m[0][i] * B.m[j][0] + A.m[1][i] * B.m[j][1]) + A.m[2][i] * B.m[j][2]) + A.m[3][i] * B.m[j][3];
    return R;
}

// FD::FDsetAreaLight, FD/api.cpp:303-400
Light setAreaLight(const LightDesc& d) {
    Light L{};
    L.position = {d.position[0], d.position[1], d.position[2]};
    L.viewDir = {d.viewDir[0], d.viewDir[1], d.viewDir[2]};
    L.upDir = {d.upDir[0], d.upDir[1], d.upDir[2]};
    L.rightRadius = d.rightRadius;
    L.topRadius = d.topRadius;
    L.resW = d.resW;
    L.resH = d.resH;
    L.resNum = d.resW * d.resH;
    L.upDir = normalize(L.upDir);                       // :328
    L.viewDir = normalize(L.viewDir);                   // :329
    L.rightDir = normalize(cross(L.viewDir, L.upDir));  // :330-331
    L.topDir = normalize(cross(L.rightDir, L.viewDir)); // :332-333
    L.x_delt = L.rightDir * (L.rightRadius * -2.0f / (float)L.resW);  // :335
    L.y_delt = L.topDir * (L.topRadius * -2.0f / (float)L.resH);      // :336
    float xf = ((float)L.resW - 1.0f) / (float)L.resW;                 // :340
    float yf = ((float)L.resH - 1.0f) / (float)L.resH;                 // :341
    L.upRight = L.position + L.rightDir * xf * L.rightRadius + L.topDir * yf * L.topRadius;                      // :343
    L.upLeft = L.position + (-1.0f * L.rightDir) * xf * L.rightRadius + (L.topDir) * yf * L.topRadius;          // :344
    L.downLeft = L.position + (-1.0f * L.rightDir) * xf * L.rightRadius + (-1.0f * L.topDir) * yf * L.topRadius; // :345
    L.downRight = L.position + (L.rightDir) * xf * L.rightRadius + (-1.0f * L.topDir) * yf * L.topRadius;       // :346
    // :384-391  upRight MVP = infinitePerspective(120, 1, 0.1) * lookAt(upRight, upRight + 10*view, up)
    f3 center = L.upRight + 10.0f * L.viewDir;
    M4 MV = lookAtRH(L.upRight, center, L.upDir);
    M4 P = infinitePerspectiveRH(120.0f, 1.0f, 0.1f);
    M4 MVP = mul(P, MV);
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) L.upRightMat[4 * c + r] = MVP.m[c][r];
    // :394-395
    L.RectangleFactor.x = L.rightRadius * 2.0f * xf * P.m[0][0] * 0.5f;
    L.RectangleFactor.y = L.topRadius * 2.0f * yf * P.m[1][1] * 0.5f;
    return L;
}

// FD::makeRectangle / FD::getZ, FDc/FDprojection.inl:6-19
inline f3 makeRectangle(const float* m, f3 t) {
    f3 r;
    r.z = m[3] * t.x + m[7] * t.y + m[11] * t.z + m[15] * 1.0f;
    r.x = m[0] * t.x + m[4] * t.y + m[8] * t.z + m[12] * 1.0f;
    r.x = r.x / r.z * 0.5f + 0.5f;
    r.y = m[1] * t.x + m[5] * t.y + m[9] * t.z + m[13] * 1.0f;
    r.y = r.y / r.z * 0.5f + 0.5f;
    return r;
}
inline float getZ(const float* m, f3 t) { return m[3] * t.x + m[7] * t.y + m[11] * t.z + m[15] * 1.0f; }

// ---------------------------------------------------------------------------
// Inputs / outputs
// ---------------------------------------------------------------------------
struct Input {
    const float* verts; int vstride;  // floats between vertices (>= 3)
    int nverts;
    const int* idx; int ntris;
    const float* pos4; const float* nrm4; int W, H;
    LightDesc light;
    float eps;
    int backfaceCull;
    int gridLog2;
    const int* subset; int nsubset;   // occlusion only for these dense receivers (nullptr = all)
    int nthreads;
    int variant;                      // occlusion: 0 per-sample kernel, 1 bin-queue (CTA) kernel
};

enum Exit {
    EXIT_OK = 0, EXIT_SAMPLE_MAXZ = 1, EXIT_TRI_MAXZ = 2, EXIT_NO_VALID_SAMPLE = 3,
    EXIT_EMPTY_AABB = 4, EXIT_EMPTY_LP = 5, EXIT_NO_VALID_TRI = 6, EXIT_NO_PAIRS = 7,
    EXIT_UNSUPPORTED_LIGHT = 8,
};

struct TriData { f3 v0, v1, v2; float maxz, minz; f2 aabbMin, aabbMax; };  // FDc/privateDefs.h:44-54
struct BinRange { int sx, sy, ex, ey; float minZ, maxZ; };                    // FDc/privateDefs.h:37-42

struct Result {
    int exitCode = 0;
    Light light{};
    float sampleMaxZ = 0, triMaxZ = 0, f = 0;
    f2 S{0, 0};
    float sampleAABB[8] = {0};  // (minx,miny,minz,-, maxx,maxy,maxz,-)
    float modelAABB[8] = {0};
    float lpBegin[3] = {0}, lpEnd[3] = {0}, lpFactor[2] = {0};
    int widthBins = 0, widthBinsLog2 = 0, heightBins = 0, numBins = 0;
    int validSampleRect = 0;   // valid receivers of the AABB stage (front-facing, non-NaN)
    int validSampleNum = 0;    // receivers bound to a bin
    int validTriNum = 0;
    long long pairNum = 0;
    // per receiver (dense)
    std::vector<float> rect;          // 8 floats: layer0 (minx,miny,z,1), layer1 (maxx,maxy,z,1)
    std::vector<int> backface;        // 1 if back-facing (output forced to 1.0)
    std::vector<int> key;             // bin key per receiver (dense order, before sort)
    std::vector<int> sortedKey, sortedId;  // stable sort_by_key; ids dense
    std::vector<int> binStart, binEnd;     // numBins+1
    std::vector<float> binMaxZ, binMinRange;  // numBins, numBins*4
    // per triangle
    std::vector<float> triAABB;       // 8 floats per triangle
    std::vector<int> triValid;
    std::vector<TriData> triData;
    std::vector<int> triCount;
    std::vector<long long> triPrefix;
    std::vector<BinRange> triRange;
    std::vector<int> pairBin, pairTri;     // sorted by bin (stable)
    std::vector<int> binTriStart, binTriEnd;
    // outputs
    std::vector<float> shadow;
    std::vector<uint32_t> masks;  // per receiver: H rows x wpr words
    std::vector<uint8_t> computed;  // receiver had its occlusion evaluated
    int wpr = 1;
    // counters
    long long evColClip = 0, evRowClip = 0, evKeyOutOfRange = 0, evValidInvalidKey = 0;
    long long candTests = 0, cullPass = 0, planeTris = 0, rowVisits = 0;
    double msStages[6] = {0};
};

// ---------------------------------------------------------------------------
// Bit helpers for the occlusion stage (U32/U64 rows)
// ---------------------------------------------------------------------------
// getRowMaskForPlaneDiscriminant, FDc/Step5_2_shadowCalPerSample.inl:14-33,
// with the shift and conversion pinned as described in the file header and
// the result clipped to FULL.  `clip` is set when bits outside FULL were set.
inline uint64_t rowMask(float D, float B, unsigned n, uint64_t FULL, int width) {
    bool f1 = (D >= 0.0f);
    bool f2 = (B * (float)n + D >= 0.0f);
    uint64_t m;
    if (f1 && f2) m = FULL;
    else if (!f1 && !f2) m = 0;
    else {
        int k = cvt_rzi(std::ceil(-1.0f * D / B));
        // width-bit word: shl by k outside [0,width) gives 0
        uint64_t one_shl = (k >= 0 && k < width) ? (1ull << k) : 0ull;
        m = one_shl - 1ull;
        if (!f1) m = ~m;
        if (width < 64) m &= (1ull << width) - 1ull;
    }
    return m;  // unclipped: the caller ANDs the three edge masks, then clips to FULL
}

// test3DPointInTriangle, FDc/Step5_2_shadowCal.inl:213-222
inline bool pointInTri(f3 A, f3 B, f3 C, f3 P) {
    f3 N1 = cross((B - A), (P - B));
    f3 N2 = cross((C - B), (P - C));
    if (dot(N1, N2) < 0.0f) return false;
    f3 N3 = cross((A - C), (P - A));
    if (dot(N2, N3) < 0.0f) return false;
    return true;
}

// getPlaneDiscriminant, FDc/Step5_2_shadowCalPerSample.inl:2-12
inline void planeDisc(float& A, float& B, float& C, f3 v0, f3 v1, f3 Q, f3 t0, f3 t1, f3 L0_Q) {
    f3 n = cross(v1 - Q, v0 - v1);
    A = dot(n, t0);
    B = dot(n, t1);
    C = dot(n, L0_Q);
}

struct OccStats { long long colClip = 0, rowClip = 0, cand = 0, pass = 0, plane = 0, rows = 0; };

// One receiver of shadowCal_per_sample_kernel, FDc/Step5_2_shadowCalPerSample.inl:37-234.
// Q: receiver position; sp: (rect minx, miny, maxx, maxy); spz: rect z.
// Writes H rows of masks (as uint64) and returns the popcount.
int occludeOne(const Light& L, float eps, f3 Q, f4 sp, float spz, const int* tris, int ntri,
               const std::vector<TriData>& td, uint64_t* mask, OccStats& st) {
    const int W = L.resW, H = L.resH;
    const unsigned wSub1 = (unsigned)(W - 1), hSub1 = (unsigned)(H - 1);
    const uint64_t FULL = (W >= 64) ? ~0ull : ((1ull << W) - 1ull);
    const uint64_t FULLH = FULL;  // lightHeightFullBit = lightWidthFullBit (:59-60), W == H
    const int width = (W <= 32 && H <= 32) ? 32 : 64;  // U32 rows up to 32, U64 beyond
    for (int r = 0; r < H; ++r) mask[r] = 0;
    uint64_t full = 0;
    for (int ti = 0; ti < ntri && full != FULLH; ++ti) {
        const TriData& t = td[tris[ti]];
        ++st.cand;
        // :120-127 containment + depth cull
        if (!(t.aabbMin.x <= sp.x && t.aabbMin.y <= sp.y && sp.z <= t.aabbMax.x && sp.w <= t.aabbMax.y &&
              t.minz <= spz))
            continue;
        ++st.pass;
        f3 v0 = t.v0, v1 = t.v1, v2 = t.v2;
        f3 n = cross((v1 - v0), (v2 - v1));  // :130
        float dir = dot(Q - v0, n);          // :132
        if (std::fabs(dir) < eps && pointInTri(v0, v1, v2, Q)) continue;  // :134-135
        ++st.plane;
        f3 t0 = L.x_delt, t1 = L.y_delt, L0_Q = L.upRight - Q;  // :137-141
        if (dir < 0) { t0 = t0 * -1.0f; t1 = t1 * -1.0f; L0_Q = L0_Q * -1.0f; }  // :143-148
        float A0, B0, C0, A1, B1, C1, A2, B2, C2;
        planeDisc(A0, B0, C0, v0, v1, Q, t0, t1, L0_Q);
        planeDisc(A1, B1, C1, v1, v2, Q, t0, t1, L0_Q);
        planeDisc(A2, B2, C2, v2, v0, Q, t0, t1, L0_Q);
        // :178-188 row-range masks
        uint64_t N0 = rowMask((B0 > 0 ? B0 * (float)wSub1 : 0.0f) + C0, A0, hSub1, FULLH, width);
        uint64_t N1 = rowMask((B1 > 0 ? B1 * (float)wSub1 : 0.0f) + C1, A1, hSub1, FULLH, width);
        uint64_t N2 = rowMask((B2 > 0 ? B2 * (float)wSub1 : 0.0f) + C2, A2, hSub1, FULLH, width);
        uint64_t nrows = N0 & N1 & N2;
        if (nrows & ~FULLH) ++st.rowClip;   // rows >= H would be visited by the reference
        uint64_t valid = (~full) & nrows & FULLH;  // :191-200
        bool cclip = false;
        while (valid) {  // :204-222, lowest set row first
            int r = __builtin_ctzll(valid);
            ++st.rows;
            uint64_t M0 = rowMask(A0 * (float)r + C0, B0, wSub1, FULL, width);
            uint64_t M1 = rowMask(A1 * (float)r + C1, B1, wSub1, FULL, width);
            uint64_t M2 = rowMask(A2 * (float)r + C2, B2, wSub1, FULL, width);
            uint64_t mm = M0 & M1 & M2;
            if (mm & ~FULL) cclip = true;   // bits >= W the reference would count
            mask[r] |= mm & FULL;
            if (mask[r] == FULL) full |= (1ull << r);
            valid ^= (1ull << r);
        }
        if (cclip) ++st.colClip;
    }
    int cnt = 0;
    for (int r = 0; r < H; ++r) cnt += __builtin_popcountll(mask[r]);
    return cnt;
}

// ---------------------------------------------------------------------------
// Bin-queue (CTA) occlusion variant: shadowCal_CTA_kernel,
// FDc/Step5_2_shadowCal.inl:408-1100 (dead code in the reference, whose
// FDLaunch runs the per-sample kernel).  Its arithmetic differs from the
// per-sample kernel's: other plane products, a fourth (triangle) plane, other
// boundary tests, and mask rows along y_delt (bit j of row i is the sample
// upRight + j x_delt + i y_delt).  Pinned semantics besides the header's:
//   * triangles are taken in groups of 32 consecutive entries of the bin's
//     list (the reference's warps); a group applies the triangle plane to
//     all its rows when any of its triangles has a non-empty row range and
//     maxz >= the receiver's z (:876, :929).  The reference decides this on
//     row ranges from which the rows already full were removed, which other
//     warps set concurrently; here the ranges before that removal are used.
//   * rows >= H (possible for 32 < H < 64, where the reference's int shift
//     makes the height mask all 64 bits) are ignored and counted.
// ---------------------------------------------------------------------------
inline uint64_t wordMask(int width) { return width >= 64 ? ~0ull : (1ull << width) - 1ull; }
inline uint64_t shlW(int k, int width) { return (k >= 0 && k < width) ? (1ull << k) : 0ull; }

// calLightHeightRange, FDc/Step5_2_shadowCal.inl:225-300
inline uint64_t ctaHeightRange(float A, float B, float C, int W, int H, int width, uint64_t fullH) {
    const float D = std::fmax(B * (float)(W - 1) + C, C);
    if (A <= 0.0f && D <= 0.0f) return 0;
    if (A == 0.0f || D == 0.0f) return fullH;
    const int k = (int)std::fmax(std::fmin(std::ceil(-1.0f * D / A), (float)H), 0.0f);   // NaN -> H
    const uint64_t a = (shlW(k, width) - 1ull) & wordMask(width);
    return A < 0.0f ? a : (~a & fullH);
}

// calLightRowRange, FDc/Step5_2_shadowCal.inl:302-357
inline uint64_t ctaRowRange(float B, float D, int W, int width) {
    if (B <= 0.0f && D <= 0.0f) return 0;
    if (B == 0.0f || D == 0.0f) return (shlW(W, width) - 1ull) & wordMask(width);
    const int k = (int)std::fmax(0.0f, std::fmin((float)W, std::ceil(-1.0f * D / B)));   // NaN -> W
    const uint64_t a = (shlW(k, width) - 1ull) & wordMask(width);
    return B < 0.0f ? a : (~a & wordMask(width));
}

// One receiver of shadowCal_CTA_kernel (the body of the sample loop,
// :682-1046, for every triangle group of the bin's list).
int occludeCTA(const Light& L, float eps, f3 S, f4 sp, float spz, const int* tris, int ntri,
               const std::vector<TriData>& td, uint64_t* mask, OccStats& st) {
    const int W = L.resW, H = L.resH;
    const int width = (W <= 32) ? 32 : 64;                       // U32 / U64 (:1120-1125)
    const uint64_t FULL = wordMask(W);                           // U32/U64lightWidthBit (FD/api.cpp:318,325)
    const uint64_t fullH = (H >= 32) ? wordMask(width) : (1ull << H) - 1ull;   // (T)(1 << H) - 1 (:431)
    const uint64_t rowsH = wordMask(H);
    const f3 xd = L.x_delt, yd = L.y_delt, ur = L.upRight;
    for (int r = 0; r < H; ++r) mask[r] = 0;
    struct Lane { float A[4], B[4], C[4]; uint64_t rows; };
    Lane lane[32];
    for (int g0 = 0; g0 < ntri; g0 += 32) {
        const int n = std::min(32, ntri - g0);
        bool depth = false;
        for (int l = 0; l < n; ++l) {
            const TriData& t = td[tris[g0 + l]];
            Lane& c = lane[l];
            c.rows = 0;
            ++st.cand;
            const bool inAABB = t.aabbMin.x <= sp.x && t.aabbMin.y <= sp.y && sp.z <= t.aabbMax.x &&
                                sp.w <= t.aabbMax.y && t.minz <= spz;   // :686-689
            if (!inAABB) continue;   // initHeightBit = 0 (:783-784): no rows
            ++st.pass;
            const f3 v0 = t.v0, v1 = t.v1, v2 = t.v2;
            // triangle plane (:667-674)
            const f3 v1_v2 = v1 - v2, v2_v0 = v2 - v0;
            const f3 nn = cross(v1_v2, v2_v0);
            const f3 L_v0 = ur - v0;
            c.A[3] = dot(nn, yd); c.B[3] = dot(nn, xd); c.C[3] = dot(nn, L_v0);
            // edge planes (:720-742)
            const f3 v0_v1 = v0 - v1;
            const f3 L_v1 = ur - v1, L_v2 = ur - v2;
            const f3 s_v1 = S - v1;
            c.A[0] = dot(cross(v0_v1, yd), s_v1); c.B[0] = dot(cross(v0_v1, xd), s_v1); c.C[0] = dot(cross(v0_v1, L_v1), s_v1);
            const f3 s_v2 = S - v2;
            c.A[1] = dot(cross(v1_v2, yd), s_v2); c.B[1] = dot(cross(v1_v2, xd), s_v2); c.C[1] = dot(cross(v1_v2, L_v2), s_v2);
            const f3 s_v0 = S - v0;
            c.A[2] = dot(cross(v2_v0, yd), s_v0); c.B[2] = dot(cross(v2_v0, xd), s_v0); c.C[2] = dot(cross(v2_v0, L_v0), s_v0);
            const float temp = dot(nn, s_v0);   // :772
            if (std::fabs(temp) < eps && pointInTri(v0, v1, v2, S)) continue;   // :783-784
            ++st.plane;
            if (temp > 0.0f)   // :785-802
                for (int k = 0; k < 4; ++k) { c.A[k] *= -1.0f; c.B[k] *= -1.0f; c.C[k] *= -1.0f; }
            uint64_t rows = fullH;   // :821-842
            for (int k = 0; k < 4; ++k) rows &= ctaHeightRange(c.A[k], c.B[k], c.C[k], W, H, width, fullH);
            c.rows = rows;
            if (rows != 0 && !(t.maxz < spz)) depth = true;   // :876, :929
        }
        for (int l = 0; l < n; ++l) {
            const Lane& c = lane[l];
            if (c.rows & ~rowsH) ++st.rowClip;
            uint64_t rows = c.rows & rowsH;
            while (rows) {   // :880-899 / :933-953 (same rows either way)
                const int i = __builtin_ctzll(rows);
                ++st.rows;
                const float fi = (float)i;
                uint64_t m = FULL;   // calLightRowShadowWith(Out)Depth, :359-397
                for (int k = 0; k < 3; ++k) m &= ctaRowRange(c.B[k], c.A[k] * fi + c.C[k], W, width);
                if (depth) m &= ctaRowRange(c.B[3], c.A[3] * fi + c.C[3], W, width);
                mask[i] |= m & FULL;
                rows &= rows - 1ull;
            }
        }
    }
    int cnt = 0;
    for (int r = 0; r < H; ++r) cnt += __builtin_popcountll(mask[r]);
    return cnt;
}

// ---------------------------------------------------------------------------
// The frame (FD::FDLaunch, FD/api.cpp:402-582)
// ---------------------------------------------------------------------------
static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

void run(const Input& in, Result& R) {
    const int W = in.W, H = in.H, WH = W * H;
    const int T = in.ntris, V = in.nverts;
    const float eps = in.eps;
#ifdef _OPENMP
    if (in.nthreads > 0) omp_set_num_threads(in.nthreads);
#endif
    R = Result();
    R.light = setAreaLight(in.light);
    const Light& L = R.light;
    const int LW = L.resW, LH = L.resH;
    R.wpr = (LW <= 32) ? 1 : 2;
    R.shadow.assign(WH, 0.0f);                    // FDresult.clear(), FD/api.cpp:489
    R.masks.assign((size_t)WH * LH * R.wpr, 0u);
    R.computed.assign(WH, 0);
    R.backface.assign(WH, 0);
    R.rect.assign((size_t)WH * 8, 0.0f);
    if (LW != LH || LW < 1 || LW > 64) { R.exitCode = EXIT_UNSUPPORTED_LIGHT; return; }
    const float* m = L.upRightMat;
    auto P = [&](int i) { return f3{in.pos4[4 * i], in.pos4[4 * i + 1], in.pos4[4 * i + 2]}; };
    auto N = [&](int i) { return f3{in.nrm4[4 * i], in.nrm4[4 * i + 1], in.nrm4[4 * i + 2]}; };
    auto VX = [&](int v) { const float* p = in.verts + (size_t)v * in.vstride; return f3{p[0], p[1], p[2]}; };
    double t0 = now_ms();

    // ---- lightPlaneParamCal, FDc/Step2_3_defineFDbuffer.inl:5-111 ------------
    // sampleMaxZcal, FDc/Step2_1_samplerRectangleAABB.inl:1-50
    float zs = -F32_MAX;
    for (int i = 0; i < WH; ++i) {
        f3 p = P(i);
        if (!std::isnan(p.x)) zs = fmax_(zs, getZ(m, p));
    }
    R.sampleMaxZ = fmax_(-1.0f, zs);
    if (R.sampleMaxZ < 0.0f) { R.exitCode = EXIT_SAMPLE_MAXZ; return; }
    // triangleMaxZcal, FDc/Step2_2_modelRectangleAABB.inl:1-46
    float zt = -F32_MAX;
    for (int v = 0; v < V; ++v) zt = fmax_(zt, getZ(m, VX(v)));
    R.triMaxZ = fmax_(-1.0f, zt);
    if (R.triMaxZ < 0.0f) { R.exitCode = EXIT_TRI_MAXZ; return; }
    R.f = fmax_(R.sampleMaxZ, R.triMaxZ) + 0.1f;          // :22
    const f2 F = L.RectangleFactor;
    R.S = {F.x / R.f, F.y / R.f};                         // :24-25 (recomputed per frame)
    const f2 S = R.S;

    // sampleRectangleAABBcal, FDc/Step2_1_samplerRectangleAABB.inl:53-191
    float sa[8] = {F32_MAX, F32_MAX, F32_MAX, 0, -F32_MAX, -F32_MAX, -F32_MAX, 0};
    int validRect = 0;
    for (int i = 0; i < WH; ++i) {
        float xmn = F32_MAX, xmx = -F32_MAX, ymn = F32_MAX, ymx = -F32_MAX, zmn = F32_MAX, zmx = -F32_MAX;
        f3 sp = P(i), n = N(i);
        if (!std::isnan(sp.x)) {
            if (dot(L.upRight - sp, n) <= 0 && dot(L.upLeft - sp, n) <= 0 && dot(L.downLeft - sp, n) <= 0 &&
                dot(L.downRight - sp, n) <= 0) {
                R.backface[i] = 1;   // shadowResult[sid] = 1.0f  (:97)
            } else {
                ++validRect;
                f3 rc = makeRectangle(m, sp);
                xmn = rc.x;
                xmx = rc.x + F.x / rc.z - S.x;
                ymn = rc.y;
                ymx = rc.y + F.y / rc.z - S.y;
                zmn = rc.z;
                zmx = rc.z;
            }
        }
        float* r = &R.rect[(size_t)i * 8];
        r[0] = xmn; r[1] = ymn; r[2] = zmn; r[3] = 1.0f;  // layer 0 (:122)
        r[4] = xmx; r[5] = ymx; r[6] = zmx; r[7] = 1.0f;  // layer 1 (:123)
        sa[0] = fmin_(sa[0], xmn); sa[1] = fmin_(sa[1], ymn); sa[2] = fmin_(sa[2], zmn);
        sa[4] = fmax_(sa[4], xmx); sa[5] = fmax_(sa[5], ymx); sa[6] = fmax_(sa[6], zmx);
    }
    for (int i = 0; i < WH; ++i)
        if (R.backface[i]) { R.shadow[i] = 1.0f; }
    std::memcpy(R.sampleAABB, sa, sizeof(sa));
    R.validSampleRect = validRect;
    // back-facing receivers are fully shadowed from here on: all mask bits set
    auto fillBackfaceMasks = [&]() {
        const uint64_t FULL = (LW >= 64) ? ~0ull : ((1ull << LW) - 1ull);
        for (int i = 0; i < WH; ++i)
            if (R.backface[i]) {
                uint32_t* mk = &R.masks[(size_t)i * LH * R.wpr];
                for (int r = 0; r < LH; ++r) {
                    if (R.wpr == 1) mk[r] = (uint32_t)FULL;
                    else { mk[2 * r] = (uint32_t)FULL; mk[2 * r + 1] = (uint32_t)(FULL >> 32); }
                }
            }
    };
    fillBackfaceMasks();
    if (validRect == 0) { R.exitCode = EXIT_NO_VALID_SAMPLE; return; }  // :178-181

    // verticesRectangleAABBcal, FDc/Step2_2_modelRectangleAABB.inl:48-124
    float ma[8] = {F32_MAX, F32_MAX, F32_MAX, 0, -F32_MAX, -F32_MAX, -F32_MAX, 0};
    for (int v = 0; v < V; ++v) {
        f3 rc = makeRectangle(m, VX(v));
        ma[0] = fmin_(ma[0], rc.x);
        ma[4] = fmax_(ma[4], rc.x + F.x / rc.z - S.x);
        ma[1] = fmin_(ma[1], rc.y);
        ma[5] = fmax_(ma[5], rc.y + F.y / rc.z - S.y);
        ma[2] = fmin_(ma[2], rc.z);
        ma[6] = fmax_(ma[6], rc.z);
    }
    std::memcpy(R.modelAABB, ma, sizeof(ma));
    // :45-47
    if (ma[0] > ma[4] || ma[1] > ma[5] || sa[0] > sa[4] || sa[1] > sa[5]) { R.exitCode = EXIT_EMPTY_AABB; return; }
    // :49-63
    float B[3], E[3];
    B[0] = fmax_(sa[0], ma[0]);
    B[1] = fmax_(sa[1], ma[1]);
    E[0] = fmin_(sa[4], ma[4]);
    E[1] = fmin_(sa[5], ma[5]);
    B[2] = ma[2];
    E[2] = sa[6];
    std::memcpy(R.lpBegin, B, sizeof(B));
    std::memcpy(R.lpEnd, E, sizeof(E));
    if (B[0] >= E[0] || B[1] >= E[1]) { R.exitCode = EXIT_EMPTY_LP; return; }
    // :96-104  (bin grid: fixed 16x16 in the reference; gridLog2 generalises it)
    const int G = 1 << in.gridLog2;
    R.widthBins = G; R.heightBins = G;
    R.widthBinsLog2 = (int)std::ceil(std::log2((float)G));
    R.numBins = G * G;
    R.lpFactor[0] = (float)G / (E[0] - B[0]);
    R.lpFactor[1] = (float)G / (E[1] - B[1]);
    const f2 fac{R.lpFactor[0], R.lpFactor[1]};
    const int NB = R.numBins, LOG2 = R.widthBinsLog2, INVALID = NB;
    double t1 = now_ms();
    R.msStages[0] = t1 - t0;

    // ---- setupTriangleVertex, FDc/Step0_triangleSetup.inl:37-231 ------------
    R.triAABB.assign((size_t)T * 8, 0.0f);
    R.triValid.assign(T, 0);
    R.triData.assign(T, TriData{});
    const bool cull = in.backfaceCull != 0;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < T; ++t) {
        int i0 = in.idx[3 * t], i1 = in.idx[3 * t + 1], i2 = in.idx[3 * t + 2];
        f3 v0 = VX(i0), v1 = VX(i1), v2 = VX(i2);
        f4 mn{F32_MAX, F32_MAX, F32_MAX, 1.0f}, mx{-F32_MAX, -F32_MAX, -F32_MAX, 1.0f};
        f3 e0 = v1 - v0;
        f3 e1 = v0 - v2;
        f3 n = cross(e1, e0);
        if (cull && dot(n, L.upRight - v0) <= 0 && dot(n, L.upRight - v1) <= 0 && dot(n, L.upRight - v2) <= 0 &&
            dot(n, L.upLeft - v0) <= 0 && dot(n, L.upLeft - v1) <= 0 && dot(n, L.upLeft - v2) <= 0 &&
            dot(n, L.downRight - v0) <= 0 && dot(n, L.downRight - v1) <= 0 && dot(n, L.downRight - v2) <= 0 &&
            dot(n, L.downLeft - v0) <= 0 && dot(n, L.downLeft - v1) <= 0 && dot(n, L.downLeft - v2) <= 0) {
            // back-face culled (:115)
        } else {
            f3 vv[3] = {v0, v1, v2};
            for (int k = 0; k < 3; ++k) {  // getTriAABB (:37-54)
                f3 rc = makeRectangle(m, vv[k]);
                mn.x = fmin_(mn.x, rc.x);
                mn.y = fmin_(mn.y, rc.y);
                mn.z = fmin_(mn.z, rc.z);
                mx.z = fmax_(mx.z, rc.z);
                mx.x = fmax_(mx.x, rc.x + F.x / rc.z - S.x);
                mx.y = fmax_(mx.y, rc.y + F.y / rc.z - S.y);
            }
            if (mn.x > E[0] || mn.y > E[1] || mx.x < B[0] || mx.y < B[1]) {
                // outside the light plane (:122-124)
            } else {
                TriData d;
                d.v0 = v0; d.v1 = v1; d.v2 = v2;
                d.minz = mn.z; d.maxz = mx.z;
                d.aabbMax = {mx.x, mx.y};
                d.aabbMin = {mn.x, mn.y};
                R.triData[t] = d;
                R.triValid[t] = 1;
            }
        }
        float* a = &R.triAABB[(size_t)t * 8];
        a[0] = mn.x; a[1] = mn.y; a[2] = mn.z; a[3] = mn.w;
        a[4] = mx.x; a[5] = mx.y; a[6] = mx.z; a[7] = mx.w;
    }
    int vt = 0;
    for (int t = 0; t < T; ++t) vt += R.triValid[t];
    R.validTriNum = vt;
    double t2 = now_ms();
    R.msStages[1] = t2 - t1;
    if (vt == 0) { R.exitCode = EXIT_NO_VALID_TRI; return; }  // :224-225

    // ---- bindSampleToBin, FDc/Step3_bindSampleToBin.inl:5-148 ---------------
    R.key.assign(WH, INVALID);
    int validSamples = 0;
    for (int i = 0; i < WH; ++i) {
        const float* r = &R.rect[(size_t)i * 8];
        f4 rb{r[0], r[1], r[2], r[3]}, re{r[4], r[5], r[6], r[7]};
        int result = INVALID;
        if ((rb.x > re.x || rb.y > re.y) || (rb.x < B[0] || rb.y < B[1] || re.x > E[0] || re.y > E[1])) {
        } else {
            f2 c{(re.x + rb.x) / 2.0f, (re.y + rb.y) / 2.0f};
            int bx = cvt_rzi(std::floor((c.x - B[0]) * fac.x));
            int by = cvt_rzi(std::floor((c.y - B[1]) * fac.y));
            result = (int)((unsigned)by << LOG2) | bx;
            ++validSamples;
            if (result < 0 || result >= NB) { ++R.evKeyOutOfRange; result = INVALID; ++R.evValidInvalidKey; }
        }
        R.key[i] = result;
    }
    R.validSampleNum = validSamples;
    // thrust::sort_by_key (stable radix sort) of (key, id)
    {
        std::vector<int> order(WH);
        for (int i = 0; i < WH; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return R.key[a] < R.key[b]; });
        R.sortedKey.resize(WH);
        R.sortedId.resize(WH);
        for (int i = 0; i < WH; ++i) { R.sortedKey[i] = R.key[order[i]]; R.sortedId[i] = order[i]; }
    }
    // FindStart_kernel, FDc/binRaster.inl:5-33 (start/end cleared to 0)
    R.binStart.assign(NB + 1, 0);
    R.binEnd.assign(NB + 1, 0);
    for (int i = 0; i < WH; ++i) {
        int k = R.sortedKey[i];
        if (i == 0 || k != R.sortedKey[i - 1]) {
            R.binStart[k] = i;
            if (i > 0) R.binEnd[R.sortedKey[i - 1]] = i;
        }
        if (i == WH - 1) R.binEnd[k] = i + 1;
    }
    // maxZpreBinReduce_kernel, FDc/Step3_bindSampleToBin.inl:61-104
    R.binMaxZ.assign(NB, -F32_MAX);
    R.binMinRange.assign((size_t)NB * 4, 0.0f);
    for (int b = 0; b < NB; ++b) {
        float mz = -F32_MAX;
        f4 mr{-F32_MAX, -F32_MAX, F32_MAX, F32_MAX};
        for (int s = R.binStart[b]; s < R.binEnd[b]; ++s) {
            const float* r = &R.rect[(size_t)R.sortedId[s] * 8];
            mz = fmax_(mz, r[2]);
            mr.x = fmax_(mr.x, r[0]);
            mr.y = fmax_(mr.y, r[1]);
            mr.z = fmin_(mr.z, r[4]);
            mr.w = fmin_(mr.w, r[5]);
        }
        R.binMaxZ[b] = mz;
        R.binMinRange[4 * b] = mr.x; R.binMinRange[4 * b + 1] = mr.y;
        R.binMinRange[4 * b + 2] = mr.z; R.binMinRange[4 * b + 3] = mr.w;
    }
    double t3 = now_ms();
    R.msStages[2] = t3 - t2;

    // ---- countTriBinPairNum / bindTriToBin, FDc/Step4_RasterizeTriangleToBin.inl:5-234
    R.triCount.assign(T, 0);
    R.triRange.assign(T, BinRange{6000, 6000, 0, 0, 0, 0});
    std::vector<f4> clampB(T), clampE(T);
    auto binTest = [&](int bp, f4 b, f4 e) {  // :94
        float mz = R.binMaxZ[bp];
        const float* mr = &R.binMinRange[4 * (size_t)bp];
        return mz >= b.z && b.x < mr[0] && b.y < mr[1] && mr[2] < e.x && mr[3] < e.y;
    };
#pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < T; ++t) {
        const float* a = &R.triAABB[(size_t)t * 8];
        f4 b{a[0], a[1], a[2], a[3]}, e{a[4], a[5], a[6], a[7]};
        int result = 0;
        int bsx = 6000, bsy = 6000, bex = 0, bey = 0;
        if (b.x > e.x || b.y > e.y || b.z > e.z || b.x > E[0] || b.y > E[1] || e.x < B[0] || e.y < B[1]) {
        } else {
            b.x = fmax_(b.x, B[0]);  // :74-77
            b.y = fmax_(b.y, B[1]);
            e.x = fmin_(e.x, E[0]);
            e.y = fmin_(e.y, E[1]);
            int cbx = cvt_rzi_u16(fmax_(std::floor((b.x - B[0]) * fac.x), 0.0f));  // :80-83
            int cby = cvt_rzi_u16(fmax_(std::floor((b.y - B[1]) * fac.y), 0.0f));
            int cex = cvt_rzi_u16(fmin_(std::ceil((e.x - B[0]) * fac.x), (float)G));
            int cey = cvt_rzi_u16(fmin_(std::ceil((e.y - B[1]) * fac.y), (float)G));
            for (int j = cby; j < cey; ++j)
                for (int i = cbx; i < cex; ++i) {
                    int bp = j << LOG2 | i;
                    if (binTest(bp, b, e)) {
                        ++result;
                        bsx = std::min(bsx, i); bsy = std::min(bsy, j);
                        bex = std::max(bex, i + 1); bey = std::max(bey, j + 1);
                    }
                }
        }
        R.triCount[t] = result;
        R.triRange[t] = BinRange{bsx, bsy, bex, bey, b.z, e.z};
        clampB[t] = b;
        clampE[t] = e;
    }
    R.triPrefix.assign(T, 0);
    long long acc = 0;
    for (int t = 0; t < T; ++t) { R.triPrefix[t] = acc; acc += R.triCount[t]; }  // thrust::exclusive_scan
    R.pairNum = acc;
    double t4 = now_ms();
    if (acc == 0) { R.msStages[3] = t4 - t3; R.exitCode = EXIT_NO_PAIRS; return; }  // FD/api.cpp:522
    std::vector<int> pb(acc), pt(acc);
#pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < T; ++t) {  // writeTriBinPair_kernel (:161-204), count pass's clamped test
        const BinRange& br = R.triRange[t];
        long long pos = R.triPrefix[t];
        int cnt = 0;
        for (int j = br.sy; j < br.ey; ++j)
            for (int i = br.sx; i < br.ex; ++i) {
                int bp = j << LOG2 | i;
                if (binTest(bp, clampB[t], clampE[t])) { pb[pos + cnt] = bp; pt[pos + cnt] = t; ++cnt; }
            }
    }
    {  // thrust::sort_by_key (stable) by bin
        std::vector<long long> order(acc);
        for (long long i = 0; i < acc; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](long long a, long long b) { return pb[a] < pb[b]; });
        R.pairBin.resize(acc);
        R.pairTri.resize(acc);
        for (long long i = 0; i < acc; ++i) { R.pairBin[i] = pb[order[i]]; R.pairTri[i] = pt[order[i]]; }
    }
    R.binTriStart.assign(NB, 0);  // findTriangleStartEnd, FDc/binRaster.inl:49-59
    R.binTriEnd.assign(NB, 0);
    for (long long i = 0; i < acc; ++i) {
        int k = R.pairBin[i];
        if (i == 0 || k != R.pairBin[i - 1]) {
            R.binTriStart[k] = (int)i;
            if (i > 0) R.binTriEnd[R.pairBin[i - 1]] = (int)i;
        }
        if (i == acc - 1) R.binTriEnd[k] = (int)(i + 1);
    }
    double t5 = now_ms();
    R.msStages[3] = t5 - t3;

    // ---- shadowCal_perSample, FDc/Step5_2_shadowCalPerSample.inl:236-259 -----
    // Every sorted receiver whose key is a real bin (the reference launches
    // validSampleNum threads over the sorted arrays; identical unless a valid
    // receiver's key fell outside the grid).
    std::vector<int> todo;
    if (in.subset) {
        std::vector<uint8_t> want(WH, 0);
        for (int k = 0; k < in.nsubset; ++k)
            if (in.subset[k] >= 0 && in.subset[k] < WH) want[in.subset[k]] = 1;
        for (int s = 0; s < WH; ++s)
            if (R.sortedKey[s] < NB && want[R.sortedId[s]]) todo.push_back(s);
    } else {
        for (int s = 0; s < WH; ++s)
            if (R.sortedKey[s] < NB) todo.push_back(s);
    }
    long long cc = 0, cr = 0, ca = 0, cp = 0, cpl = 0, cro = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : cc, cr, ca, cp, cpl, cro)
    for (long k = 0; k < (long)todo.size(); ++k) {
        int s = todo[k];
        int id = R.sortedId[s];
        int bin = R.sortedKey[s];
        const float* r = &R.rect[(size_t)id * 8];
        f4 sp{r[0], r[1], r[4], r[5]};
        float spz = r[6];
        OccStats st;
        uint64_t mk[64];
        int ts = R.binTriStart[bin], te = R.binTriEnd[bin];
        int cnt = in.variant == 1
                      ? occludeCTA(L, eps, P(id), sp, spz, R.pairTri.data() + ts, te - ts, R.triData, mk, st)
                      : occludeOne(L, eps, P(id), sp, spz, R.pairTri.data() + ts, te - ts, R.triData, mk, st);
        R.shadow[id] = (float)cnt / (float)L.resNum;  // :231 (Step5_2_shadowCal.inl:1087 for the variant)
        uint32_t* out = &R.masks[(size_t)id * LH * R.wpr];
        for (int rr = 0; rr < LH; ++rr) {
            if (R.wpr == 1) out[rr] = (uint32_t)mk[rr];
            else { out[2 * rr] = (uint32_t)mk[rr]; out[2 * rr + 1] = (uint32_t)(mk[rr] >> 32); }
        }
        R.computed[id] = 1;
        cc += st.colClip; cr += st.rowClip; ca += st.cand; cp += st.pass; cpl += st.plane; cro += st.rows;
    }
    R.evColClip = cc; R.evRowClip = cr; R.candTests = ca; R.cullPass = cp; R.planeTris = cpl; R.rowVisits = cro;
    R.msStages[4] = now_ms() - t5;
    R.exitCode = EXIT_OK;
}

}  // namespace orc

// ===========================================================================
// C interface (ctypes)
// ===========================================================================
extern "C" {

struct or_light_desc {
    float right_radius, top_radius;
    float position[3], view_dir[3], up_dir[3];
    int res_w, res_h;
};

struct or_input {
    const float* verts; int vstride_floats; int nverts;
    const int* idx; int ntris;
    const float* pos4; const float* nrm4; int w, h;
    or_light_desc light;
    float eps;
    int backface_cull;
    int grid_log2;
    const int* subset; int nsubset;
    int nthreads;
    int variant;
};

void* or_run(const or_input* ci) {
    orc::Input in{};
    in.verts = ci->verts; in.vstride = ci->vstride_floats; in.nverts = ci->nverts;
    in.idx = ci->idx; in.ntris = ci->ntris;
    in.pos4 = ci->pos4; in.nrm4 = ci->nrm4; in.W = ci->w; in.H = ci->h;
    in.light.rightRadius = ci->light.right_radius; in.light.topRadius = ci->light.top_radius;
    for (int k = 0; k < 3; ++k) {
        in.light.position[k] = ci->light.position[k];
        in.light.viewDir[k] = ci->light.view_dir[k];
        in.light.upDir[k] = ci->light.up_dir[k];
    }
    in.light.resW = ci->light.res_w; in.light.resH = ci->light.res_h;
    in.eps = ci->eps; in.backfaceCull = ci->backface_cull; in.gridLog2 = ci->grid_log2;
    in.subset = ci->subset; in.nsubset = ci->nsubset; in.nthreads = ci->nthreads;
    in.variant = ci->variant;
    auto* R = new orc::Result();
    orc::run(in, *R);
    return R;
}

void or_free(void* h) { delete (orc::Result*)h; }

// Scalars: exit code, counts, reductions, LP params, light.
int or_scalars(void* h, double* out, int n) {
    auto* R = (orc::Result*)h;
    const orc::Light& L = R->light;
    std::vector<double> v = {
        (double)R->exitCode, R->sampleMaxZ, R->triMaxZ, R->f, R->S.x, R->S.y,
        R->lpBegin[0], R->lpBegin[1], R->lpBegin[2], R->lpEnd[0], R->lpEnd[1], R->lpEnd[2],
        R->lpFactor[0], R->lpFactor[1], (double)R->widthBins, (double)R->widthBinsLog2, (double)R->numBins,
        (double)R->validSampleRect, (double)R->validSampleNum, (double)R->validTriNum, (double)R->pairNum,
        (double)R->evColClip, (double)R->evRowClip, (double)R->evKeyOutOfRange, (double)R->evValidInvalidKey,
        (double)R->candTests, (double)R->cullPass, (double)R->planeTris, (double)R->rowVisits,
        R->msStages[0], R->msStages[1], R->msStages[2], R->msStages[3], R->msStages[4], (double)R->wpr};
    for (int k = 0; k < 8; ++k) v.push_back(R->sampleAABB[k]);
    for (int k = 0; k < 8; ++k) v.push_back(R->modelAABB[k]);
    for (int k = 0; k < 16; ++k) v.push_back(L.upRightMat[k]);
    const orc::f3* vs[] = {&L.rightDir, &L.topDir, &L.x_delt, &L.y_delt, &L.upRight, &L.upLeft, &L.downLeft, &L.downRight};
    for (auto* p : vs) { v.push_back(p->x); v.push_back(p->y); v.push_back(p->z); }
    v.push_back(L.RectangleFactor.x);
    v.push_back(L.RectangleFactor.y);
    int k = 0;
    for (; k < n && k < (int)v.size(); ++k) out[k] = v[k];
    return (int)v.size();
}

// Raw float32 bit patterns of the light block (for bit-exact comparison).
int or_light_bits(void* h, float* out) {
    auto* R = (orc::Result*)h;
    const orc::Light& L = R->light;
    int k = 0;
    for (int i = 0; i < 16; ++i) out[k++] = L.upRightMat[i];
    const orc::f3* vs[] = {&L.rightDir, &L.topDir, &L.x_delt, &L.y_delt, &L.upRight, &L.upLeft, &L.downLeft, &L.downRight};
    for (auto* p : vs) { out[k++] = p->x; out[k++] = p->y; out[k++] = p->z; }
    out[k++] = L.RectangleFactor.x;
    out[k++] = L.RectangleFactor.y;
    out[k++] = R->f;
    out[k++] = R->S.x;
    out[k++] = R->S.y;
    for (int i = 0; i < 3; ++i) out[k++] = R->lpBegin[i];
    for (int i = 0; i < 3; ++i) out[k++] = R->lpEnd[i];
    out[k++] = R->lpFactor[0];
    out[k++] = R->lpFactor[1];
    for (int i = 0; i < 8; ++i) out[k++] = R->sampleAABB[i];
    for (int i = 0; i < 8; ++i) out[k++] = R->modelAABB[i];
    return k;
}

// Array accessors: returns element count; copies when dst != nullptr.
long long or_array(void* h, const char* name, void* dst) {
    auto* R = (orc::Result*)h;
    std::string n(name);
    auto cp = [&](const void* src, size_t count, size_t esz) -> long long {
        if (dst && count) std::memcpy(dst, src, count * esz);
        return (long long)count;
    };
    if (n == "shadow") return cp(R->shadow.data(), R->shadow.size(), 4);
    if (n == "masks") return cp(R->masks.data(), R->masks.size(), 4);
    if (n == "computed") return cp(R->computed.data(), R->computed.size(), 1);
    if (n == "rect") return cp(R->rect.data(), R->rect.size(), 4);
    if (n == "backface") return cp(R->backface.data(), R->backface.size(), 4);
    if (n == "key") return cp(R->key.data(), R->key.size(), 4);
    if (n == "sorted_key") return cp(R->sortedKey.data(), R->sortedKey.size(), 4);
    if (n == "sorted_id") return cp(R->sortedId.data(), R->sortedId.size(), 4);
    if (n == "bin_start") return cp(R->binStart.data(), R->binStart.size(), 4);
    if (n == "bin_end") return cp(R->binEnd.data(), R->binEnd.size(), 4);
    if (n == "bin_maxz") return cp(R->binMaxZ.data(), R->binMaxZ.size(), 4);
    if (n == "bin_minrange") return cp(R->binMinRange.data(), R->binMinRange.size(), 4);
    if (n == "tri_aabb") return cp(R->triAABB.data(), R->triAABB.size(), 4);
    if (n == "tri_valid") return cp(R->triValid.data(), R->triValid.size(), 4);
    if (n == "tri_data") return cp(R->triData.data(), R->triData.size(), sizeof(orc::TriData));
    if (n == "tri_count") return cp(R->triCount.data(), R->triCount.size(), 4);
    if (n == "tri_prefix") return cp(R->triPrefix.data(), R->triPrefix.size(), 8);
    if (n == "tri_range") return cp(R->triRange.data(), R->triRange.size(), sizeof(orc::BinRange));
    if (n == "pair_bin") return cp(R->pairBin.data(), R->pairBin.size(), 4);
    if (n == "pair_tri") return cp(R->pairTri.data(), R->pairTri.size(), 4);
    if (n == "bin_tri_start") return cp(R->binTriStart.data(), R->binTriStart.size(), 4);
    if (n == "bin_tri_end") return cp(R->binTriEnd.data(), R->binTriEnd.size(), 4);
    return -1;
}

// Occlusion of the given receivers against EVERY triangle that reached the
// light plane, with no binning (a brute-force cross-check of the binning).
int or_occlude_all(void* h, const or_input* ci, const int* ids, int n, uint32_t* out) {
    auto* R = (orc::Result*)h;
    if (R->exitCode != orc::EXIT_OK) return -1;
    std::vector<int> all;
    for (int t = 0; t < (int)R->triValid.size(); ++t)
        if (R->triValid[t]) all.push_back(t);
    const int LH = R->light.resH;
#pragma omp parallel for schedule(dynamic, 16)
    for (int k = 0; k < n; ++k) {
        int id = ids[k];
        const float* r = &R->rect[(size_t)id * 8];
        const float* p = ci->pos4 + 4 * (size_t)id;
        orc::OccStats st;
        uint64_t mk[64];
        orc::occludeOne(R->light, ci->eps, orc::f3{p[0], p[1], p[2]}, orc::f4{r[0], r[1], r[4], r[5]}, r[6],
                        all.data(), (int)all.size(), R->triData, mk, st);
        uint32_t* o = out + (size_t)k * LH * R->wpr;
        for (int rr = 0; rr < LH; ++rr) {
            if (R->wpr == 1) o[rr] = (uint32_t)mk[rr];
            else { o[2 * rr] = (uint32_t)mk[rr]; o[2 * rr + 1] = (uint32_t)(mk[rr] >> 32); }
        }
    }
    return 0;
}

// The bin-queue variant's range primitives on their own (closed-form tests):
// kind 0 = calLightRowRange(B = a, D = b), kind 1 = calLightHeightRange(A = a, B = b, C = c).
unsigned long long or_cta_range(int kind, float a, float b, float c, int lightW) {
    const int width = lightW <= 32 ? 32 : 64;
    if (kind == 0) return orc::ctaRowRange(a, b, lightW, width);
    const uint64_t fullH = (lightW >= 32) ? orc::wordMask(width) : (1ull << lightW) - 1ull;
    return orc::ctaHeightRange(a, b, c, lightW, lightW, width, fullH);
}

// The light-row mask primitive on its own (for closed-form tests).
unsigned long long or_rowmask(float D, float B, unsigned n, int lightW, int* clip) {
    uint64_t FULL = (lightW >= 64) ? ~0ull : ((1ull << lightW) - 1ull);
    int width = lightW <= 32 ? 32 : 64;
    uint64_t m = orc::rowMask(D, B, n, FULL, width);
    if (clip) *clip = (m & ~FULL) ? 1 : 0;
    return m & FULL;
}

// makeRectangle on its own.
void or_make_rectangle(const float* mat16, const float* p3, float* out3) {
    orc::f3 r = orc::makeRectangle(mat16, orc::f3{p3[0], p3[1], p3[2]});
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}

// The shading pass of the reference application (deferred Phong over the
// light samples, 4DZB/Resources/shader/phoneShading.shf:23-61), in float:
// normalize(v) = v / sqrt(dot(v, v)), pow = powf, max = fmaxf, expressions in
// the shader's order.  g0..g3: float4 x n (gatherData.shf:41-44 packing);
// samples: ns x 3; light: La, Ld, Ls (3 each); cam: 3.  out: float4 x n; a
// texel whose intensity 1 - shadow is NaN keeps (0, 0, 0, 0) (discard).
void or_shade(int n, const float* shadow, const float* g0, const float* g1, const float* g2, const float* g3,
              int ns, const float* samples, const float* cam, const float* La, const float* Ld, const float* Ls,
              float* out) {
    auto dot = [](const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    auto normalize = [&](float* v) {
        const float l = std::sqrt(dot(v, v));
        v[0] = v[0] / l; v[1] = v[1] / l; v[2] = v[2] / l;
    };
    for (int i = 0; i < n; ++i) {
        float* o = out + 4 * (size_t)i;
        const float intensity = 1.0f - shadow[i];                  // :25
        if (std::isnan(intensity)) { o[0] = o[1] = o[2] = o[3] = 0.0f; continue; }   // :26-27
        const float* a0 = g0 + 4 * (size_t)i;
        const float* a1 = g1 + 4 * (size_t)i;
        const float* a2 = g2 + 4 * (size_t)i;
        const float* a3 = g3 + 4 * (size_t)i;
        const float P[3] = {a0[0], a0[1], a0[2]};                  // :34-39
        const float ka[3] = {a2[0], a2[1], a2[2]}, kd[3] = {a3[0], a3[1], a3[2]}, ks[3] = {a0[3], a1[3], a2[3]};
        const float shin = a3[3];
        float N[3] = {a1[0], a1[1], a1[2]};
        normalize(N);                                               // :41
        float V[3] = {cam[0] - P[0], cam[1] - P[1], cam[2] - P[2]};
        normalize(V);                                               // :42
        float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = 0; k < ns; ++k) {                              // :45-58
            float L[3] = {samples[3 * k] - P[0], samples[3 * k + 1] - P[1], samples[3 * k + 2] - P[2]};
            normalize(L);
            const float LdN = dot(L, N);
            const float t2 = 2.0f * LdN;
            const float R[3] = {t2 * N[0] - L[0], t2 * N[1] - L[1], t2 * N[2] - L[2]};
            const float dif = std::fmax(LdN, 0.0f);
            const float spw = std::pow(std::fmax(dot(R, V), 0.0f), shin);
            for (int j = 0; j < 3; ++j) {
                const float amb = La[j] * ka[j];
                const float dd = Ld[j] * kd[j] * dif;
                const float sc = Ls[j] * ks[j] * spw;
                c[j] += amb + (dd + sc) * intensity;
            }
            c[3] += 1.0f;
        }
        const float cnt = (float)(unsigned)ns;                     // :60
        for (int j = 0; j < 4; ++j) o[j] = c[j] / cnt;
    }
}

// FD::moveModel (4DZB/FDZB_Soft_Shadow/api.cpp:17-20) -> cudaModeTransform ->
// the transform kernel (FDc/moveModel.inl:6-30): each vertex position, with
// w set to 1, is multiplied by the column-major matrix m (mul, :6-13: each
// output is ((m[i]*x + m[4+i]*y) + m[8+i]*z) + m[12+i]*w, left to right, no
// FMA) and x, y, z are written back in place; every other float of the
// vertex (stride vstride floats) is left as it was.
void or_move_model(float* verts, int vstride, int nverts, const float* m) {
    for (int t = 0; t < nverts; ++t) {
        float* p = verts + (size_t)t * vstride;
        const float x = p[0], y = p[1], z = p[2], w = 1.0f;   // pos.w = 1.0f (:21)
        const float rx = m[0] * x + m[4] * y + m[8] * z + m[12] * w;
        const float ry = m[1] * x + m[5] * y + m[9] * z + m[13] * w;
        const float rz = m[2] * x + m[6] * y + m[10] * z + m[14] * w;
        p[0] = rx;
        p[1] = ry;
        p[2] = rz;
    }
}

int or_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
