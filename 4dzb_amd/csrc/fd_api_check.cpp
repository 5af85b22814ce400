// fd_api_check — runs one frame through the C++ mirror of the reference's
// FD:: interface (include/fd_api.hpp), in the call order of the reference
// application: Render::Render (4DZB/Render/Render.cpp:179-189), the model
// upload (4DZB/ResourceManager/MeshManager.cpp:87), FDregisterGLres
// (4DZB/main.cpp:56) and the per-frame light update + launch
// (Render.cpp:473-482).
//
//   fd_api_check <frame.bin> <out.bin>
//
// frame.bin (little endian): int32 W, H, nverts, ntris, res, cull; float32
// eps, rightRadius, topRadius, position[3], viewDir[3], upDir[3]; then
// nverts float4 vertices, ntris int3 indices, W*H float4 positions, W*H
// float4 normals.  out.bin: int32 exit code, float32 elapsed ms, W*H float32
// shadow values, W*H*fd_mask_words(res) uint32 mask words.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fd_api.hpp"

namespace {

template <typename T>
bool read_n(FILE* f, T* p, size_t n) {
    return fread(p, sizeof(T), n, f) == n;
}

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) {
        if (fd_dev_alloc(bytes, &p) != FD_OK) throw FD::Error("fd_dev_alloc", FD_ERR_HIP);
    }
    ~DevBuf() {
        if (p) fd_dev_free(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s frame.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) {
        perror(argv[1]);
        return 2;
    }
    int hdr[6];
    float fl[12];
    if (!read_n(f, hdr, 6) || !read_n(f, fl, 12)) {
        fprintf(stderr, "short header\n");
        return 2;
    }
    const int W = hdr[0], H = hdr[1], nv = hdr[2], nt = hdr[3], res = hdr[4], cull = hdr[5];
    if (W <= 0 || H <= 0 || nv <= 0 || nt <= 0 || res < 1 || res > 64) {
        fprintf(stderr, "bad header\n");
        return 2;
    }
    const size_t WH = (size_t)W * H;
    std::vector<float> verts((size_t)nv * 4), pos(WH * 4), nrm(WH * 4);
    std::vector<int> idx((size_t)nt * 3);
    if (!read_n(f, verts.data(), verts.size()) || !read_n(f, idx.data(), idx.size()) ||
        !read_n(f, pos.data(), pos.size()) || !read_n(f, nrm.data(), nrm.size())) {
        fprintf(stderr, "short body\n");
        return 2;
    }
    fclose(f);
    const int words = fd_mask_words(res);
    try {
        DevBuf dV(verts.size() * 4), dI(idx.size() * 4), dP(pos.size() * 4), dN(nrm.size() * 4);
        DevBuf dS(WH * 4), dM(WH * (size_t)words * 4);
        fd_memcpy_htod(dV.p, verts.data(), verts.size() * 4);
        fd_memcpy_htod(dI.p, idx.data(), idx.size() * 4);
        fd_memcpy_htod(dP.p, pos.data(), pos.size() * 4);
        fd_memcpy_htod(dN.p, nrm.data(), nrm.size() * 4);

        FD::FDinit();
        FD::FDsetBackfaceCull(cull != 0);
        FD::FDsetSceneParams(1 << 21, 1 << 21);
        FD::FDsetScenesEsp(fl[0]);
        FD::FDsetSampleGLTex((const float*)dP.p, (const float*)dN.p);
        FD::FDsetOutputGLBuffer((float*)dS.p, (uint32_t*)dM.p);
        FD::FDsetView(W, H);
        FD::FDaddModel((const float*)dV.p, nv, (const int*)dI.p, nt, 4 * sizeof(float), "first model");
        FD::FDregisterGLres();

        FD::AreaLightDes light;
        light.rightRadius = fl[1];
        light.topRadius = fl[2];
        for (int k = 0; k < 3; ++k) {
            light.position[k] = fl[3 + k];
            light.viewDir[k] = fl[6 + k];
            light.upDir[k] = fl[9 + k];
        }
        light.lightResWidth = light.lightResHeight = res;
        FD::FDsetAreaLight(light);
        const float ms = FD::FDLaunch();
        const int exitCode = FD::FDlastStats().exit_code;

        std::vector<float> shadow(WH);
        std::vector<uint32_t> masks(WH * (size_t)words);
        fd_memcpy_dtoh(shadow.data(), dS.p, WH * 4);
        fd_memcpy_dtoh(masks.data(), dM.p, masks.size() * 4);
        FD::FDunregisterGLres();
        FD::FDshutdown();

        FILE* o = fopen(argv[2], "wb");
        if (!o) {
            perror(argv[2]);
            return 2;
        }
        fwrite(&exitCode, 4, 1, o);
        fwrite(&ms, 4, 1, o);
        fwrite(shadow.data(), 4, WH, o);
        fwrite(masks.data(), 4, masks.size(), o);
        fclose(o);
        printf("exit %d, %.3f ms\n", exitCode, ms);
    } catch (const FD::Error& e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
