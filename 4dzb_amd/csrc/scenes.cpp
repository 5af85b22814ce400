// Deterministic synthetic scene and receiver generators for the soft-shadow
// visibility workloads (host C++, no GPU).  Every generator is seeded with a
// SplitMix64 stream and builds geometry in double precision before rounding
// to float, so the same seed gives the same bytes on every x86-64 host that
// runs this image's libm.
//
// Layouts produced (what the visibility library consumes):
//   receivers : two float4 arrays of W*H texels, row-major (y*W + x); xyz is the
//               world position / normal, w = 1.  A receiver with pos.x = NaN is
//               "no receiver" (G-buffer background), as in the reference's
//               G-buffer (4DZB/Render/Render.cpp:423-426).
//   occluders : float4 vertices (w = 1, 16-byte stride) + int3 triangle indices.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0,1) with 53 random bits
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * u01(); }
    double normal() {
        double u1 = 1.0 - u01();  // (0,1]
        double u2 = u01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct D3 { double x, y, z; };
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 normalize(D3 a) { double l = std::sqrt(dot(a, a)); return a * (1.0 / l); }

inline void put4(float* p, D3 v) { p[0] = (float)v.x; p[1] = (float)v.y; p[2] = (float)v.z; p[3] = 1.0f; }

// An equilateral triangle with the given centroid, edge length, unit normal
// and in-plane rotation.  Winding is counter-clockwise about the normal.
void equilateral(float* v4, int* idx, long base, D3 c, double edge, D3 n, double theta) {
    D3 a = std::fabs(n.x) < 0.9 ? D3{1, 0, 0} : D3{0, 1, 0};
    D3 e1 = normalize(cross(n, a));
    D3 e2 = cross(n, e1);
    double R = edge / std::sqrt(3.0);
    for (int k = 0; k < 3; ++k) {
        double ang = theta + 2.0943951023931953 * k;
        D3 p = c + e1 * (R * std::cos(ang)) + e2 * (R * std::sin(ang));
        put4(v4 + 4 * (base + k), p);
        idx[k] = (int)(base + k);
    }
}

D3 unit_sphere_dir(Rng& r) {
    double z = r.uni(-1.0, 1.0);
    double phi = r.uni(0.0, 6.283185307179586);
    double s = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    return {s * std::cos(phi), s * std::sin(phi), z};
}

}  // namespace

extern "C" {

// Receivers on the plane z = 0 at pixel centres over [-half, half]^2, normal +z.
int fdgen_receivers_plane(int w, int h, float half, float* pos4, float* nrm4) {
    if (w <= 0 || h <= 0 || !pos4 || !nrm4) return -1;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            long i = (long)y * w + x;
            double px = -half + (x + 0.5) * (2.0 * half / w);
            double py = -half + (y + 0.5) * (2.0 * half / h);
            put4(pos4 + 4 * i, {px, py, 0.0});
            put4(nrm4 + 4 * i, {0.0, 0.0, 1.0});
        }
    return 0;
}

// Receivers on the heightfield z = amp*sin(x)*cos(y) at pixel centres over
// [-hx,hx] x [-hy,hy], with the analytic normal.
int fdgen_receivers_heightfield(int w, int h, float hx, float hy, float amp, float* pos4, float* nrm4) {
    if (w <= 0 || h <= 0 || !pos4 || !nrm4) return -1;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            long i = (long)y * w + x;
            double px = -hx + (x + 0.5) * (2.0 * hx / w);
            double py = -hy + (y + 0.5) * (2.0 * hy / h);
            double z = amp * std::sin(px) * std::cos(py);
            D3 n = normalize({-amp * std::cos(px) * std::cos(py), amp * std::sin(px) * std::sin(py), 1.0});
            put4(pos4 + 4 * i, {px, py, z});
            put4(nrm4 + 4 * i, n);
        }
    return 0;
}

// Box layout of the boxes G-buffer: `nbox` axis-aligned boxes standing on
// the ground, centre U[-7,7]^2, half extents U[0.3,1.0], height U[0.5,2.5].
// out: nbox x {min x, y, z, max x, y, z}.
int fdgen_boxes_layout(int nbox, uint64_t seed, double* out) {
    if (nbox < 0 || nbox > 256 || (nbox > 0 && !out)) return -1;
    Rng r(seed);
    for (int b = 0; b < nbox; ++b) {
        double cx = r.uni(-7, 7), cy = r.uni(-7, 7);
        double sx = r.uni(0.3, 1.0), sy = r.uni(0.3, 1.0), hz = r.uni(0.5, 2.5);
        double* o = out + 6 * b;
        o[0] = cx - sx; o[1] = cy - sy; o[2] = 0.0;
        o[3] = cx + sx; o[4] = cy + sy; o[5] = hz;
    }
    return 0;
}

// A G-buffer ray-cast on the host: a pinhole camera (eye, at, up, vertical
// field of view) over the ground plane z = 0 (|x|,|y| <= ground_half) with
// `nbox` axis-aligned boxes standing on it.  Writes the four RGBA32F targets
// of the reference's G-buffer pass (gatherData.shf:41-44): g0 = (position,
// ks.r), g1 = (normal, ks.g), g2 = (ka, ks.b), g3 = (kd, shininess), with
// mat = nbox + 1 materials of 10 floats (ka, kd, ks, shininess; [0] the
// ground).  Background texels are NaN in every channel (the reference clears
// to NaN, Render.cpp:423-426).  g2 / g3 / mat may be null (then only g0, g1
// are written, with w = 1).  Double precision, in a fixed operation order
// (the device ray cast repeats it bit for bit).
int fdgen_gbuffer_boxes(int w, int h, int nbox, const double* boxes, const double* eye3, const double* at3,
                        const double* up3, double fovy_deg, double ground_half, const float* mat, float* g0, float* g1,
                        float* g2, float* g3) {
    if (w <= 0 || h <= 0 || nbox < 0 || nbox > 256 || (nbox > 0 && !boxes) || !g0 || !g1) return -1;
    const D3 eye{eye3[0], eye3[1], eye3[2]}, at{at3[0], at3[1], at3[2]}, upv{up3[0], up3[1], up3[2]};
    D3 f = normalize(at - eye);
    D3 s = normalize(cross(f, upv));
    D3 u = cross(s, f);
    double tanh_ = std::tan(0.5 * fovy_deg * 3.141592653589793 / 180.0);
    double aspect = (double)w / h;
    const float qnan = std::numeric_limits<float>::quiet_NaN();
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            long i = (long)y * w + x;
            double ndcx = (2.0 * (x + 0.5) / w - 1.0) * tanh_ * aspect;
            double ndcy = (1.0 - 2.0 * (y + 0.5) / h) * tanh_;
            D3 d = normalize(f + s * ndcx + u * ndcy);
            double tbest = 1e300;
            D3 nbest{0, 0, 0};
            int obj = -1;
            if (d.z < 0.0) {
                double t = -eye.z / d.z;
                D3 p = eye + d * t;
                if (std::fabs(p.x) <= ground_half && std::fabs(p.y) <= ground_half) { tbest = t; nbest = {0, 0, 1}; obj = 0; }
            }
            const double e[3] = {eye.x, eye.y, eye.z}, dd[3] = {d.x, d.y, d.z};
            for (int b = 0; b < nbox; ++b) {
                const double* bmin = boxes + 6 * b;
                const double* bmax = bmin + 3;
                double t0 = -1e300, t1 = 1e300;
                int axis = -1;
                double sgn = 0;
                bool miss = false;
                for (int a = 0; a < 3; ++a) {
                    if (std::fabs(dd[a]) < 1e-300) {
                        if (e[a] < bmin[a] || e[a] > bmax[a]) { miss = true; break; }
                        continue;
                    }
                    double ta = (bmin[a] - e[a]) / dd[a], tb = (bmax[a] - e[a]) / dd[a];
                    double sa = -1.0;
                    if (ta > tb) { double tmp = ta; ta = tb; tb = tmp; sa = 1.0; }
                    if (ta > t0) { t0 = ta; axis = a; sgn = sa; }
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
                const float* m = mat ? mat + 10 * obj : nullptr;
                put4(g0 + 4 * i, eye + d * tbest);
                put4(g1 + 4 * i, nbest);
                if (m) {
                    g0[4 * i + 3] = m[6];
                    g1[4 * i + 3] = m[7];
                    if (g2) { g2[4 * i] = m[0]; g2[4 * i + 1] = m[1]; g2[4 * i + 2] = m[2]; g2[4 * i + 3] = m[8]; }
                    if (g3) { g3[4 * i] = m[3]; g3[4 * i + 1] = m[4]; g3[4 * i + 2] = m[5]; g3[4 * i + 3] = m[9]; }
                }
            } else {
                for (int k = 0; k < 4; ++k) {
                    g0[4 * i + k] = qnan; g1[4 * i + k] = qnan;
                    if (g2) g2[4 * i + k] = qnan;
                    if (g3) g3[4 * i + k] = qnan;
                }
            }
        }
    return 0;
}

// The boxes receivers of the workloads: the G-buffer above with 20 boxes by
// default, camera at (0, -13, 9) looking at the origin, 50 degrees, ground
// half-size 10; position and normal only (w = 1).
int fdgen_receivers_boxes(int w, int h, int nbox, uint64_t seed, float* pos4, float* nrm4) {
    if (w <= 0 || h <= 0 || nbox < 0 || nbox > 256 || !pos4 || !nrm4) return -1;
    double boxes[256 * 6];
    fdgen_boxes_layout(nbox, seed, boxes);
    const double eye[3] = {0.0, -13.0, 9.0}, at[3] = {0.0, 0.0, 0.0}, up[3] = {0.0, 0.0, 1.0};
    return fdgen_gbuffer_boxes(w, h, nbox, boxes, eye, at, up, 50.0, 10.0, nullptr, pos4, nrm4, nullptr, nullptr);
}

// `ntri` unshared equilateral triangles: centroid uniform in the box, edge
// uniform in [e0,e1], normal uniform on the sphere, in-plane angle uniform.
// Writes 3*ntri float4 vertices and ntri int3 indices.
long fdgen_tris_uniform(long ntri, uint64_t seed, const float* box6, float e0, float e1, float* verts4, int* idx3) {
    if (ntri < 0 || !box6 || !verts4 || !idx3) return -1;
    Rng r(seed);
    for (long t = 0; t < ntri; ++t) {
        D3 c{r.uni(box6[0], box6[1]), r.uni(box6[2], box6[3]), r.uni(box6[4], box6[5])};
        double edge = r.uni(e0, e1);
        D3 n = unit_sphere_dir(r);
        double th = r.uni(0.0, 6.283185307179586);
        equilateral(verts4, idx3 + 3 * t, 3 * t, c, edge, n, th);
    }
    return ntri;
}

// `ntri` unshared equilateral triangles with log-normal edge length
// exp(N(mu, sigma)) around `ncluster` Gaussian clusters (std `csig`) whose
// centres are uniform in the box; centroids are clamped into the box.
long fdgen_tris_clusters(long ntri, uint64_t seed, int ncluster, float csig, const float* box6,
                         float mu, float sigma, float* verts4, int* idx3) {
    if (ntri < 0 || ncluster <= 0 || ncluster > 1024 || !box6 || !verts4 || !idx3) return -1;
    Rng r(seed);
    D3 cc[1024];
    for (int k = 0; k < ncluster; ++k)
        cc[k] = {r.uni(box6[0], box6[1]), r.uni(box6[2], box6[3]), r.uni(box6[4], box6[5])};
    for (long t = 0; t < ntri; ++t) {
        int k = (int)(r.next() % (uint64_t)ncluster);
        D3 c = cc[k] + D3{r.normal() * csig, r.normal() * csig, r.normal() * csig};
        c.x = std::fmin(std::fmax(c.x, (double)box6[0]), (double)box6[1]);
        c.y = std::fmin(std::fmax(c.y, (double)box6[2]), (double)box6[3]);
        c.z = std::fmin(std::fmax(c.z, (double)box6[4]), (double)box6[5]);
        double edge = std::exp(mu + sigma * r.normal());
        D3 n = unit_sphere_dir(r);
        double th = r.uni(0.0, 6.283185307179586);
        equilateral(verts4, idx3 + 3 * t, 3 * t, c, edge, n, th);
    }
    return ntri;
}

// UV sphere mesh size for `slices` x `stacks`.
long fdgen_sphere_counts(int slices, int stacks, long* nverts_per, long* ntris_per) {
    if (slices < 3 || stacks < 2) return -1;
    *nverts_per = 2 + (long)(stacks - 1) * slices;
    *ntris_per = 2L * slices * (stacks - 1);
    return 0;
}

// `nsph` UV spheres (outward counter-clockwise winding), radius uniform in
// [r0,r1], centre uniform in the box.  Vertices are shared within a sphere.
long fdgen_spheres(int nsph, int slices, int stacks, uint64_t seed, const float* box6, float r0, float r1,
                   float* verts4, int* idx3) {
    long nvp, ntp;
    if (nsph < 0 || fdgen_sphere_counts(slices, stacks, &nvp, &ntp) != 0 || !box6 || !verts4 || !idx3) return -1;
    Rng r(seed);
    const double PI = 3.141592653589793;
    long vt = 0, tt = 0;
    for (int s = 0; s < nsph; ++s) {
        double rad = r.uni(r0, r1);
        D3 c{r.uni(box6[0], box6[1]), r.uni(box6[2], box6[3]), r.uni(box6[4], box6[5])};
        long base = vt;
        put4(verts4 + 4 * vt++, c + D3{0, 0, rad});  // north pole
        for (int i = 1; i < stacks; ++i) {
            double th = PI * i / stacks;
            for (int j = 0; j < slices; ++j) {
                double ph = 2.0 * PI * j / slices;
                D3 d{std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
                put4(verts4 + 4 * vt++, c + d * rad);
            }
        }
        put4(verts4 + 4 * vt++, c + D3{0, 0, -rad});  // south pole
        long south = vt - 1;
        auto ring = [&](int i, int j) { return base + 1 + (long)(i - 1) * slices + (j % slices); };
        for (int j = 0; j < slices; ++j) {
            int* t = idx3 + 3 * tt++;
            t[0] = (int)base; t[1] = (int)ring(1, j); t[2] = (int)ring(1, j + 1);
        }
        for (int i = 1; i < stacks - 1; ++i)
            for (int j = 0; j < slices; ++j) {
                int* t = idx3 + 3 * tt++;
                t[0] = (int)ring(i, j); t[1] = (int)ring(i + 1, j); t[2] = (int)ring(i + 1, j + 1);
                t = idx3 + 3 * tt++;
                t[0] = (int)ring(i, j); t[1] = (int)ring(i + 1, j + 1); t[2] = (int)ring(i, j + 1);
            }
        for (int j = 0; j < slices; ++j) {
            int* t = idx3 + 3 * tt++;
            t[0] = (int)ring(stacks - 1, j); t[1] = (int)south; t[2] = (int)ring(stacks - 1, j + 1);
        }
    }
    return tt;
}

// Uniform draw helper for instance poses (exposed so workload tables are
// reproducible from one seed without Python's RNG).
int fdgen_uniform(uint64_t seed, long n, double a, double b, double* out) {
    if (n < 0 || !out) return -1;
    Rng r(seed);
    for (long i = 0; i < n; ++i) out[i] = r.uni(a, b);
    return 0;
}

}  // extern "C"
