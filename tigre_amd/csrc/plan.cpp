// Projection plans: the per-view records of one geometry prepared and
// uploaded once, so that forward and backward projections launch with no
// host synchronisation, no allocation and no record upload.  Scratch comes
// from a caller workspace, which makes every plan launch capturable in a
// HIP graph (the kernels issue no memset; the output buffers are written in
// full).  The tunables are the views per beam workgroup (1, 2 or 4), the
// LDS slab depth of the beam kernel (1..8 principal planes) and the slices
// per lane of the backprojector (4, 8 or 16).
#include <cstring>

#include "common.hpp"

using namespace tigre;

struct TigrePlan_ {
    int device = 0;
    HostGeometry g;
    ForwardRecords fwd[2];              // [0] Siddon, [1] interpolated
    void* d_fwd[2] = {nullptr, nullptr};
    void* d_groups[2] = {nullptr, nullptr};
    void* d_bwd = nullptr;
    bool axial = false;
    int max_planes = 8;
    int zper = 16;
    ~TigrePlan_() {
        for (int k = 0; k < 2; ++k) {
            if (d_fwd[k]) (void)hipFree(d_fwd[k]);
            if (d_groups[k]) (void)hipFree(d_groups[k]);
        }
        if (d_bwd) (void)hipFree(d_bwd);
    }
};

namespace {

int upload_sync(void** d, const void* h, size_t bytes) {
    if (bytes == 0) return TIGRE_OK;
    TIGRE_HIP_TRY(hipMalloc(d, bytes));
    TIGRE_HIP_TRY(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
    return TIGRE_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int plan_on_device(const TigrePlan_* p) {
    TIGRE_REQUIRE(p != nullptr, "null plan");
    int dev = -1;
    TIGRE_HIP_TRY(hipGetDevice(&dev));
    TIGRE_REQUIRE(dev == p->device, "plan used on another device than the one it was created on");
    return TIGRE_OK;
}

size_t plan_workspace(const TigrePlan_* p, int dtype) {
    const HostGeometry& g = p->g;
    const size_t nproj = (size_t)g.nDetec[0] * g.nDetec[1] * g.nangles;
    const size_t nvox = (size_t)g.nVoxel[0] * g.nVoxel[1] * g.nVoxel[2];
    const size_t ptmp = dtype == TIGRE_DTYPE_F32 ? 0 : align256(nproj * sizeof(float));
    const size_t vtmp = dtype == TIGRE_DTYPE_F32 ? 0 : align256(nvox * sizeof(float));
    size_t b = atb_workspace_bytes(g) + vtmp;
    for (int k = 0; k < 2; ++k) b = std::max(b, ax_workspace_bytes(g, p->fwd[k], k == 1) + ptmp);
    return b;
}

}  // namespace

extern "C" {

int tigre_plan_create(const TigreGeometry* geo, const float* angles, int nangles, int views_per_group,
                      int max_slab_planes, int slices_per_thread, TigrePlan* out) {
    set_error("");
    TIGRE_REQUIRE(out != nullptr, "null plan handle");
    *out = nullptr;
    TIGRE_REQUIRE(views_per_group == 1 || views_per_group == 2 || views_per_group == 4,
                  "views per group must be 1, 2 or 4");
    TIGRE_REQUIRE(max_slab_planes >= 1 && max_slab_planes <= 8, "slab planes must be in 1..8");
    TIGRE_REQUIRE(slices_per_thread == 4 || slices_per_thread == 8 || slices_per_thread == 16,
                  "slices per thread must be 4, 8 or 16");
    TigrePlan_* p = new TigrePlan_();
    int st = [&]() -> int {
        TIGRE_HIP_TRY(hipGetDevice(&p->device));
        TIGRE_CHECK(import_geometry(geo, angles, nangles, &p->g));
        p->max_planes = max_slab_planes;
        p->zper = slices_per_thread;
        for (int k = 0; k < 2; ++k) {
            build_forward_records(p->g, k == 1, views_per_group, &p->fwd[k]);
            TIGRE_CHECK(upload_sync(&p->d_fwd[k], p->fwd[k].ordered.data(),
                                    p->fwd[k].ordered.size() * sizeof(FwdView)));
            TIGRE_CHECK(upload_sync(&p->d_groups[k], p->fwd[k].groups.data(),
                                    p->fwd[k].groups.size() * sizeof(int)));
        }
        std::vector<BwdView> bwd;
        backward_views(p->g, &bwd);
        p->axial = backward_is_axial(p->g, bwd);
        TIGRE_CHECK(upload_sync(&p->d_bwd, bwd.data(), bwd.size() * sizeof(BwdView)));
        return TIGRE_OK;
    }();
    if (st != TIGRE_OK) {
        delete p;
        return st;
    }
    *out = p;
    return TIGRE_OK;
}

int tigre_plan_workspace_size(TigrePlan plan, int dtype, size_t* bytes) {
    set_error("");
    TIGRE_REQUIRE(plan != nullptr && bytes != nullptr, "null argument");
    TIGRE_REQUIRE(valid_dtype(dtype), "unknown storage type");
    *bytes = plan_workspace(plan, dtype);
    return TIGRE_OK;
}

int tigre_plan_ax(TigrePlan plan, const void* d_img, void* d_proj, int projection_type, int layout, int dtype,
                  void* d_workspace, size_t workspace_bytes, void* stream) {
    set_error("");
    TIGRE_CHECK(plan_on_device(plan));
    TIGRE_REQUIRE(d_img != nullptr && d_proj != nullptr, "null device pointer");
    TIGRE_REQUIRE(projection_type == TIGRE_PROJ_SIDDON || projection_type == TIGRE_PROJ_INTERPOLATED,
                  "unknown projection type");
    TIGRE_REQUIRE(layout == TIGRE_LAYOUT_PYTHON || layout == TIGRE_LAYOUT_MATLAB, "unknown detector layout");
    TIGRE_REQUIRE(valid_dtype(dtype), "unknown storage type");
    TIGRE_REQUIRE(workspace_bytes >= plan_workspace(plan, dtype), "workspace smaller than tigre_plan_workspace_size");
    TIGRE_REQUIRE(d_workspace != nullptr && (reinterpret_cast<uintptr_t>(d_workspace) & 255) == 0,
                  "workspace must be 256-byte aligned");
    const HostGeometry& g = plan->g;
    const int k = projection_type == TIGRE_PROJ_INTERPOLATED ? 1 : 0;
    hipStream_t s = (hipStream_t)stream;
    char* ws = static_cast<char*>(d_workspace);
    float* out = static_cast<float*>(d_proj);
    const long long nproj = (long long)g.nDetec[0] * g.nDetec[1] * g.nangles;
    if (dtype != TIGRE_DTYPE_F32) {
        out = reinterpret_cast<float*>(ws);
        ws += align256((size_t)nproj * sizeof(float));
    }
    TIGRE_CHECK(ax_launch(d_img, dtype, g, plan->fwd[k], (const FwdView*)plan->d_fwd[k],
                          (const int*)plan->d_groups[k], projection_type, layout, out, ws, plan->max_planes, s));
    if (dtype != TIGRE_DTYPE_F32) TIGRE_CHECK(launch_store_converted(out, d_proj, nproj, dtype, s));
    return TIGRE_OK;
}

int tigre_plan_atb(TigrePlan plan, const void* d_proj, void* d_vol, int backprojection_type, int layout, int dtype,
                   void* d_workspace, size_t workspace_bytes, void* stream) {
    set_error("");
    TIGRE_CHECK(plan_on_device(plan));
    TIGRE_REQUIRE(d_proj != nullptr && d_vol != nullptr, "null device pointer");
    TIGRE_REQUIRE(backprojection_type == TIGRE_BACKPROJ_FDK || backprojection_type == TIGRE_BACKPROJ_MATCHED,
                  "unknown backprojection type");
    TIGRE_REQUIRE(layout == TIGRE_LAYOUT_PYTHON || layout == TIGRE_LAYOUT_MATLAB, "unknown detector layout");
    TIGRE_REQUIRE(valid_dtype(dtype), "unknown storage type");
    TIGRE_REQUIRE(workspace_bytes >= plan_workspace(plan, dtype), "workspace smaller than tigre_plan_workspace_size");
    TIGRE_REQUIRE(d_workspace != nullptr && (reinterpret_cast<uintptr_t>(d_workspace) & 255) == 0,
                  "workspace must be 256-byte aligned");
    const HostGeometry& g = plan->g;
    hipStream_t s = (hipStream_t)stream;
    char* ws = static_cast<char*>(d_workspace);
    float* out = static_cast<float*>(d_vol);
    const long long nvox = (long long)g.nVoxel[0] * g.nVoxel[1] * g.nVoxel[2];
    if (dtype != TIGRE_DTYPE_F32) {
        out = reinterpret_cast<float*>(ws);
        ws += align256((size_t)nvox * sizeof(float));
    }
    TIGRE_CHECK(atb_launch(d_proj, dtype, g, (const BwdView*)plan->d_bwd, plan->axial, backprojection_type, layout,
                           out, ws, plan->zper, AtbUpdate{}, s));
    if (dtype != TIGRE_DTYPE_F32) TIGRE_CHECK(launch_store_converted(out, d_vol, nvox, dtype, s));
    return TIGRE_OK;
}

int tigre_plan_sart_update(TigrePlan plan, const float* d_residual, float* d_x, const float* d_inv_v,
                           int backprojection_type, int layout, int noneg, void* d_workspace, size_t workspace_bytes,
                           void* stream) {
    set_error("");
    TIGRE_CHECK(plan_on_device(plan));
    TIGRE_REQUIRE(d_residual != nullptr && d_x != nullptr && d_inv_v != nullptr, "null device pointer");
    TIGRE_REQUIRE(backprojection_type == TIGRE_BACKPROJ_FDK || backprojection_type == TIGRE_BACKPROJ_MATCHED,
                  "unknown backprojection type");
    TIGRE_REQUIRE(layout == TIGRE_LAYOUT_PYTHON || layout == TIGRE_LAYOUT_MATLAB, "unknown detector layout");
    TIGRE_REQUIRE(workspace_bytes >= plan_workspace(plan, TIGRE_DTYPE_F32),
                  "workspace smaller than tigre_plan_workspace_size");
    TIGRE_REQUIRE(d_workspace != nullptr && (reinterpret_cast<uintptr_t>(d_workspace) & 255) == 0,
                  "workspace must be 256-byte aligned");
    AtbUpdate ep;
    ep.x = d_x;
    ep.inv_v = d_inv_v;
    ep.noneg = noneg ? 1 : 0;
    TIGRE_CHECK(atb_launch(d_residual, TIGRE_DTYPE_F32, plan->g, (const BwdView*)plan->d_bwd, plan->axial,
                           backprojection_type, layout, d_x, static_cast<char*>(d_workspace), plan->zper, ep,
                           (hipStream_t)stream));
    return TIGRE_OK;
}

int tigre_sart_residual(const float* d_proj, const float* d_w, float* d_ax, long long n, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_proj != nullptr && d_w != nullptr && d_ax != nullptr, "null device pointer");
    TIGRE_REQUIRE(n >= 0, "negative size");
    return launch_sart_residual(d_proj, d_w, d_ax, n, (hipStream_t)stream);
}

int tigre_cnn_conv(const void* d_x, int x_mode, const void* d_w, const float* d_bias, const void* d_aux, void* d_y,
                   int epilogue, int n, int d, int h, int w, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_x != nullptr && d_w != nullptr && d_y != nullptr, "null device pointer");
    return launch_cnn_conv(d_x, x_mode, static_cast<const unsigned short*>(d_w), d_bias, d_aux, d_y, epilogue, n, d,
                           h, w, (hipStream_t)stream);
}

int tigre_cnn_wgrad_workspace_size(int n, int d, int h, int w, size_t* bytes) {
    set_error("");
    TIGRE_REQUIRE(bytes != nullptr, "null argument");
    TIGRE_REQUIRE(n > 0 && d > 0 && h > 0 && w > 0, "empty volume");
    *bytes = cnn_wgrad_workspace(n, d, h, w);
    return TIGRE_OK;
}

int tigre_cnn_wgrad(const void* d_x, int x_mode, const void* d_dy, int dy_mode, float* d_dw, int n, int d, int h,
                    int w, void* d_workspace, size_t workspace_bytes, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_x != nullptr && d_dy != nullptr && d_dw != nullptr, "null device pointer");
    return launch_cnn_wgrad(d_x, x_mode, d_dy, dy_mode, d_dw, n, d, h, w, d_workspace, workspace_bytes,
                            (hipStream_t)stream);
}

int tigre_plan_destroy(TigrePlan plan) {
    set_error("");
    if (plan == nullptr) return TIGRE_OK;
    int prev = -1;
    TIGRE_HIP_TRY(hipGetDevice(&prev));
    TIGRE_HIP_TRY(hipSetDevice(plan->device));
    delete plan;
    TIGRE_HIP_TRY(hipSetDevice(prev));
    return TIGRE_OK;
}

}  // extern "C"
