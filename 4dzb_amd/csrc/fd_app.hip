// The application stages on either side of the visibility path, on the
// device:
//   * G-buffer producer -- replaces Render::pass0 + gatherData.shf
//     (4DZB/Render/Render.cpp:390-431, 4DZB/Resources/shader/gatherData.shf:22-46)
//     with a double-precision ray cast of an analytic scene (ground plane +
//     axis-aligned boxes) into the four RGBA32F G-buffer targets;
//   * shading consumer -- replaces Render::pass2 + phoneShading.shf
//     (4DZB/Render/Render.cpp:510-565, 4DZB/Resources/shader/phoneShading.shf:23-61).
// Both are one thread per texel, stream-ordered, no host synchronisation.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include "fd_host.h"

namespace {

// ---------------------------------------------------------------------------
// G-buffer producer.  The per-pixel arithmetic repeats the host generator
// (csrc/scenes.cpp, fdgen_gbuffer_boxes) operation for operation in double
// precision (IEEE division and square root, no contraction), so both give the
// same bits.
// ---------------------------------------------------------------------------
struct GbArgs {
    double eye[3], f[3], s[3], u[3];
    double tanh_, aspect, ground_half;
    int w, h, nbox, hasMat;
    const double* boxes;   // nbox x 6
    const float* mat;      // (nbox + 1) x 10: ka, kd, ks, shininess
    float4 *g0, *g1, *g2, *g3;
};

struct DV { double x, y, z; };
__device__ __forceinline__ DV dadd(DV a, DV b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ DV dmul(DV a, double k) { return {a.x * k, a.y * k, a.z * k}; }
__device__ __forceinline__ DV dnormalize(DV a) {
    const double l = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return dmul(a, 1.0 / l);
}

__global__ __launch_bounds__(256) void k_gbuffer_boxes(GbArgs a) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)a.w * a.h) return;
    const int y = (int)(i / a.w), x = (int)(i - (long)y * a.w);
    const DV eye{a.eye[0], a.eye[1], a.eye[2]};
    const DV f{a.f[0], a.f[1], a.f[2]}, s{a.s[0], a.s[1], a.s[2]}, u{a.u[0], a.u[1], a.u[2]};
    const double ndcx = (2.0 * (x + 0.5) / a.w - 1.0) * a.tanh_ * a.aspect;
    const double ndcy = (1.0 - 2.0 * (y + 0.5) / a.h) * a.tanh_;
    const DV d = dnormalize(dadd(dadd(f, dmul(s, ndcx)), dmul(u, ndcy)));
    double tbest = 1e300;
    DV nbest{0, 0, 0};
    int obj = -1;
    if (d.z < 0.0) {
        const double t = -eye.z / d.z;
        const DV p = dadd(eye, dmul(d, t));
        if (fabs(p.x) <= a.ground_half && fabs(p.y) <= a.ground_half) { tbest = t; nbest = {0, 0, 1}; obj = 0; }
    }
    const double e[3] = {eye.x, eye.y, eye.z}, dd[3] = {d.x, d.y, d.z};
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(4))) double* cdbl_p;   // uniform box records: scalar loads
#else
    typedef const double* cdbl_p;
#endif
    const cdbl_p boxes = (cdbl_p)a.boxes;
    for (int b = 0; b < a.nbox; ++b) {
        const cdbl_p bmin = boxes + 6 * b;
        const cdbl_p bmax = bmin + 3;
        double t0 = -1e300, t1 = 1e300;
        int axis = -1;
        double sgn = 0;
        bool miss = false;
        for (int k = 0; k < 3; ++k) {
            if (fabs(dd[k]) < 1e-300) {
                if (e[k] < bmin[k] || e[k] > bmax[k]) { miss = true; break; }
                continue;
            }
            double ta = (bmin[k] - e[k]) / dd[k], tb = (bmax[k] - e[k]) / dd[k];
            double sa = -1.0;
            if (ta > tb) { const double tmp = ta; ta = tb; tb = tmp; sa = 1.0; }
            if (ta > t0) { t0 = ta; axis = k; sgn = sa; }
            if (tb < t1) t1 = tb;
        }
        if (miss || t0 > t1 || t0 <= 0.0 || axis < 0) continue;
        if (t0 < tbest) {
            tbest = t0;
            nbest = {0, 0, 0};
            if (axis == 0) nbest.x = sgn; else if (axis == 1) nbest.y = sgn; else nbest.z = sgn;
            obj = 1 + b;
        }
    }
    if (tbest < 1e299) {
        const DV p = dadd(eye, dmul(d, tbest));
        const float* m = a.hasMat ? a.mat + 10 * obj : nullptr;
        a.g0[i] = make_float4((float)p.x, (float)p.y, (float)p.z, m ? m[6] : 1.0f);
        a.g1[i] = make_float4((float)nbest.x, (float)nbest.y, (float)nbest.z, m ? m[7] : 1.0f);
        if (m) {
            if (a.g2) a.g2[i] = make_float4(m[0], m[1], m[2], m[8]);
            if (a.g3) a.g3[i] = make_float4(m[3], m[4], m[5], m[9]);
        }
    } else {
        const float qn = __builtin_nanf("");
        const float4 nn = make_float4(qn, qn, qn, qn);
        a.g0[i] = nn;
        a.g1[i] = nn;
        if (a.g2) a.g2[i] = nn;
        if (a.g3) a.g3[i] = nn;
    }
}

// ---------------------------------------------------------------------------
// Shading consumer (phoneShading.shf:23-61), in float with IEEE division and
// square root: normalize(v) = v / sqrt(dot(v, v)), pow = powf, max = fmaxf;
// expressions evaluated in the shader's order.
// ---------------------------------------------------------------------------
struct ShadeArgs {
    int n, ns;
    float cam[3], La[3], Ld[3], Ls[3];
    const float* samples;   // ns x 3
    const float* shadow;
    const float4 *g0, *g1, *g2, *g3;
    float4* out;
};

struct F3 { float x, y, z; };
__device__ __forceinline__ F3 fsub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float fdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 fnormalize(F3 v) {
    const float l = sqrtf(fdot(v, v));
    return {v.x / l, v.y / l, v.z / l};
}

__global__ __launch_bounds__(256) void k_shade(ShadeArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float intensity = 1.0f - a.shadow[i];            // :25
    if (intensity != intensity) {                            // :26-27 discard: the cleared colour stays
        a.out[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float4 g0 = a.g0[i], g1 = a.g1[i], g2 = a.g2[i], g3 = a.g3[i];
    const F3 P{g0.x, g0.y, g0.z}, Nw{g1.x, g1.y, g1.z};
    const F3 ka{g2.x, g2.y, g2.z}, kd{g3.x, g3.y, g3.z}, ks{g0.w, g1.w, g2.w};
    const float shin = g3.w;
    const F3 N = fnormalize(Nw);                              // :41
    const F3 V = fnormalize(fsub(F3{a.cam[0], a.cam[1], a.cam[2]}, P));
    float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int k = 0; k < a.ns; ++k) {                          // :45-58
        const F3 sp{a.samples[3 * k], a.samples[3 * k + 1], a.samples[3 * k + 2]};
        const F3 L = fnormalize(fsub(sp, P));
        const float LdN = fdot(L, N);
        const float t2 = 2.0f * LdN;
        const F3 R{t2 * N.x - L.x, t2 * N.y - L.y, t2 * N.z - L.z};
        const float dif = fmaxf(LdN, 0.0f);
        const float spw = powf(fmaxf(fdot(R, V), 0.0f), shin);
        const F3 amb{a.La[0] * ka.x, a.La[1] * ka.y, a.La[2] * ka.z};
        const F3 dd{a.Ld[0] * kd.x * dif, a.Ld[1] * kd.y * dif, a.Ld[2] * kd.z * dif};
        const F3 sc{a.Ls[0] * ks.x * spw, a.Ls[1] * ks.y * spw, a.Ls[2] * ks.z * spw};
        c.x += amb.x + (dd.x + sc.x) * intensity;
        c.y += amb.y + (dd.y + sc.y) * intensity;
        c.z += amb.z + (dd.z + sc.z) * intensity;
        c.w += 1.0f;
    }
    const float cnt = (float)(unsigned)a.ns;                  // :60
    a.out[i] = make_float4(c.x / cnt, c.y / cnt, c.z / cnt, c.w / cnt);
}

struct H3 { double x, y, z; };
H3 hsub(H3 a, H3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
H3 hcross(H3 a, H3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
H3 hnormalize(H3 a) {
    const double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    const double k = 1.0 / l;
    return {a.x * k, a.y * k, a.z * k};
}

// A small stream-ordered upload of host tables (freed in stream order).
struct Upload {
    void* p = nullptr;
    hipError_t put(const void* src, size_t bytes, hipStream_t st) {
        hipError_t e = hipMallocAsync(&p, bytes ? bytes : 16, st);
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, st);
        return e;
    }
    void free(hipStream_t st) { if (p) (void)hipFreeAsync(p, st); p = nullptr; }
};

}  // namespace

extern "C" {

int fd_gbuffer_boxes(const fd_gbuffer_scene* sc, float* d_g0, float* d_g1, float* d_g2, float* d_g3, void* stream) {
    if (!sc || !d_g0 || !d_g1 || sc->width <= 0 || sc->height <= 0 || sc->nbox < 0 || sc->nbox > 256 ||
        (sc->nbox > 0 && !sc->boxes) || (long)sc->width * sc->height > (1L << 30))
        return FD_ERR_ARG;
    if ((d_g2 || d_g3) && !sc->materials) return FD_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    GbArgs a{};
    // camera basis on the host, as the host generator forms it
    const H3 eye{sc->eye[0], sc->eye[1], sc->eye[2]}, at{sc->at[0], sc->at[1], sc->at[2]};
    const H3 up{sc->up[0], sc->up[1], sc->up[2]};
    const H3 f = hnormalize(hsub(at, eye));
    const H3 s = hnormalize(hcross(f, up));
    const H3 u = hcross(s, f);
    const double e3[3] = {eye.x, eye.y, eye.z}, f3[3] = {f.x, f.y, f.z}, s3[3] = {s.x, s.y, s.z}, u3[3] = {u.x, u.y, u.z};
    std::memcpy(a.eye, e3, sizeof(e3));
    std::memcpy(a.f, f3, sizeof(f3));
    std::memcpy(a.s, s3, sizeof(s3));
    std::memcpy(a.u, u3, sizeof(u3));
    a.tanh_ = std::tan(0.5 * sc->fovy_deg * 3.141592653589793 / 180.0);
    a.aspect = (double)sc->width / sc->height;
    a.ground_half = sc->ground_half;
    a.w = sc->width;
    a.h = sc->height;
    a.nbox = sc->nbox;
    a.hasMat = sc->materials != nullptr;
    a.g0 = (float4*)d_g0;
    a.g1 = (float4*)d_g1;
    a.g2 = (float4*)d_g2;
    a.g3 = (float4*)d_g3;
    Upload ub, um;
    hipError_t e = ub.put(sc->boxes, (size_t)sc->nbox * 6 * sizeof(double), st);
    if (e == hipSuccess && sc->materials) e = um.put(sc->materials, (size_t)(sc->nbox + 1) * sizeof(fd_material), st);
    if (e == hipSuccess) {
        a.boxes = (const double*)ub.p;
        a.mat = (const float*)um.p;
        const long n = (long)sc->width * sc->height;
        hipLaunchKernelGGL(k_gbuffer_boxes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
        e = hipGetLastError();
    }
    ub.free(st);
    um.free(st);
    return fdh::hip_status(e);
}

int fd_shade(const fd_shade_params* p, const float* d_shadow, const float* d_g0, const float* d_g1,
             const float* d_g2, const float* d_g3, float* d_rgba, void* stream) {
    if (!p || !d_shadow || !d_g0 || !d_g1 || !d_g2 || !d_g3 || !d_rgba || p->width <= 0 || p->height <= 0 ||
        p->nsamples <= 0 || p->nsamples > 4096 || !p->samples || (long)p->width * p->height > (1L << 30))
        return FD_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    ShadeArgs a{};
    a.n = p->width * p->height;
    a.ns = p->nsamples;
    for (int k = 0; k < 3; ++k) {
        a.cam[k] = p->camera[k];
        a.La[k] = p->La[k];
        a.Ld[k] = p->Ld[k];
        a.Ls[k] = p->Ls[k];
    }
    a.shadow = d_shadow;
    a.g0 = (const float4*)d_g0;
    a.g1 = (const float4*)d_g1;
    a.g2 = (const float4*)d_g2;
    a.g3 = (const float4*)d_g3;
    a.out = (float4*)d_rgba;
    Upload us;
    hipError_t e = us.put(p->samples, (size_t)p->nsamples * 3 * sizeof(float), st);
    if (e == hipSuccess) {
        a.samples = (const float*)us.p;
        hipLaunchKernelGGL(k_shade, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a);
        e = hipGetLastError();
    }
    us.free(st);
    return fdh::hip_status(e);
}

}  // extern "C"
