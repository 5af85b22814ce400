/*
 * fd_api.hpp — C++ mirror of the reference's FDZB_Soft_Shadow interface
 * (namespace FD, 4DZB/FDZB_Soft_Shadow/api.h:10-56) over the C ABI in fd4d.h.
 *
 * Same function names, call order and argument meaning as the reference, so
 * a caller of FD::FD* switches by changing the include and the buffer
 * arguments.  What follows from the boundary:
 *   - GL handles become device pointers with explicit counts (FDaddModel,
 *     FDsetSampleGLTex, FDsetOutputGLBuffer); the caller owns the buffers;
 *   - FDsetOutputGLBuffer also takes an optional light-sample mask buffer;
 *   - a failing call throws FD::Error (the reference's checkCudaErrors ends
 *     the process instead); FDsetBinNum keeps the reference's int return
 *     (0, or -1 for an unsupported size);
 *   - FDLaunch returns the frame's device time in ms, like the reference;
 *     FDlastStats() returns the full record of the last frame.
 * Like the reference, the interface keeps one process-wide context (here: one
 * per thread, created by FDinit).
 */
#ifndef FD_API_HPP
#define FD_API_HPP

#include <stdexcept>
#include <string>

#include "fd4d.h"

namespace FD {

/* AreaLightDes, 4DZB/FDZB_Soft_Shadow/AreaLightDesc.h:3-13 (glm::vec3
 * members as float[3]). */
struct AreaLightDes {
    float rightRadius = 0.0f, topRadius = 0.0f;
    float position[3] = {0, 0, 0};
    float viewDir[3] = {0, 0, -1};
    float upDir[3] = {0, 1, 0};
    int lightResWidth = 16, lightResHeight = 16;
};

class Error : public std::runtime_error {
public:
    Error(const std::string& what, int code)
        : std::runtime_error(what + ": " + fd_status_string(code)), code(code) {}
    int code;
};

namespace detail {
inline fd_ctx*& ctx() {
    static thread_local fd_ctx* c = nullptr;
    return c;
}
inline fd_stats& last() {
    static thread_local fd_stats s{};
    return s;
}
inline void check(int rc, const char* what) {
    if (rc != FD_OK) throw Error(what, rc);
}
inline fd_ctx* need() {
    if (!ctx()) throw Error("FD: FDinit not called", FD_ERR_STATE);
    return ctx();
}
}  // namespace detail

/* FD::moveModel, FD/api.cpp:17-20. */
inline void moveModel(float m[16]) { detail::check(fd_move_model(detail::need(), m, nullptr), "moveModel"); }

/* FD::FDinit, FD/api.cpp:48-144 (a second call starts a fresh context). */
inline void FDinit() {
    if (detail::ctx()) fd_destroy(detail::ctx());
    detail::ctx() = nullptr;
    detail::check(fd_init(&detail::ctx()), "FDinit");
}

/* No counterpart: releases the context (the reference keeps its globals
 * until the process ends). */
inline void FDshutdown() {
    if (detail::ctx()) fd_destroy(detail::ctx());
    detail::ctx() = nullptr;
}

/* FD::FDsetBackfaceCull, FD/api.cpp:146-150. */
inline void FDsetBackfaceCull(bool enable) {
    detail::check(fd_set_backface_cull(detail::need(), enable ? 1 : 0), "FDsetBackfaceCull");
}

/* FD::FDaddModel, FD/api.cpp:152-160: VBO/IBO become device pointers with
 * their element counts; modelName is kept for the caller's bookkeeping only,
 * as in the reference. */
inline void FDaddModel(const float* d_vertices, int numVertices, const int* d_indices, int numTriangles,
                       size_t vertexSizeBytes, const std::string& modelName) {
    (void)modelName;
    detail::check(fd_add_model(detail::need(), d_vertices, vertexSizeBytes, numVertices, d_indices, numTriangles),
                  "FDaddModel");
}

/* FD::FDsetSampleGLTex, FD/api.cpp:162-166: float4 position / normal images. */
inline void FDsetSampleGLTex(const float* d_position4, const float* d_normal4) {
    detail::check(fd_set_sample_buffers(detail::need(), d_position4, d_normal4), "FDsetSampleGLTex");
}

/* FD::FDsetOutputGLBuffer, FD/api.cpp:168-171: the shadow-value buffer, plus
 * (optional) the light-sample masks. */
inline void FDsetOutputGLBuffer(float* d_shadow, uint32_t* d_masks = nullptr) {
    detail::check(fd_set_output_buffers(detail::need(), d_shadow, d_masks), "FDsetOutputGLBuffer");
}

/* FD::FDregisterGLres / FDunregisterGLres, FD/api.cpp:173-192. */
inline void FDregisterGLres() { detail::check(fd_register_res(detail::need()), "FDregisterGLres"); }
inline void FDunregisterGLres() { detail::check(fd_unregister_res(detail::need()), "FDunregisterGLres"); }

/* FD::FDsetBinNum, FD/api.h:36: 0 on success, -1 for an unsupported size. */
inline int FDsetBinNum(size_t binNum) { return fd_set_bin_num(detail::need(), binNum) == FD_OK ? 0 : -1; }

/* FD::FDsetView, FD/api.cpp:205-237. */
inline void FDsetView(int viewWidth, int viewHeight) {
    detail::check(fd_set_view(detail::need(), viewWidth, viewHeight), "FDsetView");
}

/* FD::FDsetScenesEsp, FD/api.cpp:238-241. */
inline void FDsetScenesEsp(float scenes_esp) {
    detail::check(fd_set_scenes_esp(detail::need(), scenes_esp), "FDsetScenesEsp");
}

/* FD::FDsetSceneParams, FD/api.cpp:243-294. */
inline void FDsetSceneParams(size_t maxTriangleNum, size_t maxVerticesNum) {
    detail::check(fd_set_scene_params(detail::need(), maxTriangleNum, maxVerticesNum), "FDsetSceneParams");
}

/* FD::FDsetAreaLight, FD/api.cpp:303-400. */
inline void FDsetAreaLight(const AreaLightDes& d) {
    fd_area_light L{};
    L.right_radius = d.rightRadius;
    L.top_radius = d.topRadius;
    for (int k = 0; k < 3; ++k) {
        L.position[k] = d.position[k];
        L.view_dir[k] = d.viewDir[k];
        L.up_dir[k] = d.upDir[k];
    }
    L.res_w = d.lightResWidth;
    L.res_h = d.lightResHeight;
    detail::check(fd_set_area_light(detail::need(), &L), "FDsetAreaLight");
}

/* FD::FDLaunch, FD/api.cpp:402-582: one frame; returns its device time (ms). */
inline float FDLaunch() {
    fd_stats s{};
    detail::check(fd_launch(detail::need(), nullptr, &s), "FDLaunch");
    detail::last() = s;
    return s.elapsed_ms;
}

/* No counterpart: the record of the last FDLaunch (exit code, counts). */
inline const fd_stats& FDlastStats() { return detail::last(); }

}  // namespace FD

#endif  // FD_API_HPP
