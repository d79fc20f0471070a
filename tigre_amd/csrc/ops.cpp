// Operator drivers on device-resident data: build the per-view records,
// upload them, stage the padded operand and launch the kernel, all on the
// caller's stream with stream-ordered scratch allocations.
#include "common.hpp"

namespace tigre {

int ax_device(const float* d_img, const HostGeometry& g, float* d_proj, int projection_type, int layout,
              hipStream_t stream) {
    TIGRE_REQUIRE(projection_type == TIGRE_PROJ_SIDDON || projection_type == TIGRE_PROJ_INTERPOLATED,
                  "unknown projection type");
    TIGRE_REQUIRE(layout == TIGRE_LAYOUT_PYTHON || layout == TIGRE_LAYOUT_MATLAB, "unknown detector layout");
    const bool interp = projection_type == TIGRE_PROJ_INTERPOLATED;
    std::vector<FwdView> views;
    forward_views(g, interp, &views);
    void* d_views = nullptr;
    TIGRE_CHECK(upload_async(&d_views, views.data(), views.size() * sizeof(FwdView), stream));
    const BrickGrid b = interp ? interp_bricks(g) : siddon_bricks(g);
    float* bricked = nullptr;
    TIGRE_HIP_TRY(hipMallocAsync((void**)&bricked, (size_t)b.elements() * sizeof(float), stream));
    TIGRE_CHECK(launch_brick_volume(d_img, g, b, bricked, stream));
    if (interp)
        TIGRE_CHECK(launch_ax_interp(bricked, g, b, (const FwdView*)d_views, layout, d_proj, stream));
    else
        TIGRE_CHECK(launch_ax_siddon(bricked, g, b, (const FwdView*)d_views, layout, d_proj, stream));
    TIGRE_HIP_TRY(hipFreeAsync(bricked, stream));
    TIGRE_HIP_TRY(hipFreeAsync(d_views, stream));
    return TIGRE_OK;
}

int atb_device(const float* d_proj, const HostGeometry& g, float* d_vol, int backprojection_type, int layout,
               hipStream_t stream) {
    TIGRE_REQUIRE(backprojection_type == TIGRE_BACKPROJ_FDK || backprojection_type == TIGRE_BACKPROJ_MATCHED,
                  "unknown backprojection type");
    TIGRE_REQUIRE(layout == TIGRE_LAYOUT_PYTHON || layout == TIGRE_LAYOUT_MATLAB, "unknown detector layout");
    std::vector<BwdView> views;
    backward_views(g, &views);
    const bool axial = backward_is_axial(g, views);
    void* d_views = nullptr;
    TIGRE_CHECK(upload_async(&d_views, views.data(), views.size() * sizeof(BwdView), stream));
    const size_t quad_bytes =
        (size_t)quad_det(g.nDetec[0]) * quad_det(g.nDetec[1]) * (size_t)g.nangles * 4 * sizeof(float);
    float* quads = nullptr;
    TIGRE_HIP_TRY(hipMallocAsync((void**)&quads, quad_bytes, stream));
    TIGRE_CHECK(launch_quad_projections(d_proj, g.nDetec[0], g.nDetec[1], g.nangles, layout, quads, stream));
    TIGRE_CHECK(launch_atb(quads, g, (const BwdView*)d_views, backprojection_type == TIGRE_BACKPROJ_MATCHED, axial,
                           d_vol, stream));
    TIGRE_HIP_TRY(hipFreeAsync(quads, stream));
    TIGRE_HIP_TRY(hipFreeAsync(d_views, stream));
    return TIGRE_OK;
}

}  // namespace tigre
