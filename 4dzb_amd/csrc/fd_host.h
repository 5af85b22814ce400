// Host-side helpers shared by the C ABI, the batched form and the server.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <vector>
#include "../../include/fd4d.h"
#include "fd_kernels.h"
#include "fd_types.h"

namespace fdh {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; cap = 0; }
        size_t want = bytes ? bytes : 256;
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; cap = 0; }
};

struct Model {
    const float* verts;
    int vstride;   // floats between vertices
    int nverts;
    const int* idx;
    int ntris;
};

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }
// In-place vertex transform of a resident model (FD::moveModel,
// FDc/moveModel.inl:6-30), enqueued on st.
hipError_t move_vertices(float* verts, int vstride, int nverts, const float m[16], hipStream_t st);
int hip_status(hipError_t e);
// FD::FDsetAreaLight (4DZB/FDZB_Soft_Shadow/api.cpp:303-400)
void light_setup(const fd_area_light& d, FdLight& L);
void fill_args(FdArgs& a, const Model& md, const float4* pos, const float4* nrm, int W, int H, const FdLight& L,
               float eps, int cull, int gridLog2, int tileLog2, float* shadow, uint32_t* masks,
               unsigned long long pairCap);
// an explicit list split (fd_set_list_split), 0 = the automatic rule
void apply_split(FdArgs& a, int req);
// place the frame workspace at base (nullptr: size only); returns bytes
size_t layout(FdArgs& a, char* base, bool countOps);
void fill_stats(const FdState& s, fd_stats* o);
// device bytes of a batch plan for n instances
size_t plan_bytes(int n);
// host image of a batch plan whose device copy will live at devBase
size_t build_plan(const std::vector<FdArgs>& args, bool anyTris, char* devBase, std::vector<char>& image,
                  FdBatchPlan& P);

}  // namespace fdh
