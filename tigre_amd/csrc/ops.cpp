// Operator drivers on device-resident data: build the per-view records,
// upload them, stage the padded operand and launch the kernel, all on the
// caller's stream with stream-ordered scratch allocations.
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace tigre {

namespace {

// Beam eligibility of a view for the beam-staged Siddon projector: every ray
// of the view must advance along the principal axis P in one direction.  The
// ray direction is linear in the pixel coordinates, so checking the four
// detector corners covers the whole detector.
struct BeamClass {
    int principal, dir;
    bool ok;
};

BeamClass beam_class(const HostGeometry& g, const FwdView& w) {
    const int nU = g.nDetec[0], nV = g.nDetec[1];
    auto ray = [&](double u, double pv, int c) {
        double s[3];
        ray_source(w, u, pv, s);
        return w.org[c] + u * w.du[c] + pv * w.dv[c] - s[c];
    };
    const double cu = 0.5 * (nU - 1), cv = 0.5 * (nV - 1);
    BeamClass k;
    k.principal = std::fabs(ray(cu, cv, 0)) >= std::fabs(ray(cu, cv, 1)) ? 0 : 1;
    k.dir = ray(cu, cv, k.principal) >= 0.0 ? 1 : -1;
    k.ok = true;
    for (int cu_i = 0; cu_i < 2; ++cu_i)
        for (int cv_i = 0; cv_i < 2; ++cv_i) {
            const double u = cu_i ? nU - 1 : 0, pv = cv_i ? nV - 1 : 0;
            const double d[3] = {ray(u, pv, 0), ray(u, pv, 1), ray(u, pv, 2)};
            const double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            // a ray that barely advances along P would make the slab boxes huge
            if (!(d[k.principal] * k.dir > 0.1 * n)) k.ok = false;
        }
    return k;
}

}  // namespace

void build_forward_records(const HostGeometry& g, bool interp, int views_per_group, ForwardRecords* rec) {
    std::vector<FwdView> views;
    forward_views(g, interp, &views);
    // Beam-eligible views grouped by (principal, direction), in rotation-angle
    // order inside a class, up to views_per_group per workgroup; the others go
    // through the gather-driven kernels.  pad[0] keeps each view's output index.
    std::vector<int> cls(views.size());
    std::vector<int> beam, rest;
    for (size_t i = 0; i < views.size(); ++i) {
        BeamClass k = beam_class(g, views[i]);
        // parallel-beam Siddon walks a ray tilted in z at a fixed height, outside
        // the slab boxes planned from the tilted beam: such views take the
        // gather kernel (horizontal rays, the usual parallel geometry, do not)
        if (g.parallel && !interp) {
            double s[3];
            ray_source(views[i], 0.0, 0.0, s);
            if (std::fabs(views[i].org[2] - s[2]) >= 1e-3) k.ok = false;
        }
        cls[i] = k.ok ? k.principal * 2 + (k.dir > 0 ? 1 : 0) : -1;
        (k.ok ? beam : rest).push_back((int)i);
    }
    auto angle = [&](int i) {
        const double a = std::fmod(g.angles[3 * i], 2.0 * M_PI);
        return a < 0.0 ? a + 2.0 * M_PI : a;
    };
    std::stable_sort(beam.begin(), beam.end(), [&](int a, int b) {
        if (cls[a] != cls[b]) return cls[a] < cls[b];
        return angle(a) < angle(b);
    });
    // Views of one workgroup must be close in angle, or their beams spread
    // apart and the slab boxes grow: a group spans at most 3.5 typical view
    // spacings (a class can wrap around 0 and leave a gap inside its list),
    // and at most the angle over which the beam of a detector tile moves by
    // 2.5 tile widths at the edge of the volume (sparse views, spread subsets:
    // groups of views 18 degrees apart made boxes several times a beam's).
    double tile_mm = INFINITY, reach_mm = 0.0;
    for (int i = 0; i < g.nangles; ++i) {
        const double mag = g.parallel ? 1.0 : g.DSO[i] / g.DSD[i];
        tile_mm = std::min(tile_mm, 16.0 * g.dDetec[0] * mag);  // the beam kernel's 16-pixel tile at the axis
        reach_mm = std::max(reach_mm, std::hypot(g.offOrig[3 * i], g.offOrig[3 * i + 1]));
    }
    reach_mm += 0.5 * std::hypot(g.sVoxel[0], g.sVoxel[1]);
    const double geo_spread = reach_mm > 0.0 ? 2.5 * tile_mm / reach_mm : INFINITY;
    std::vector<double> sorted_angles;
    for (int i : beam) sorted_angles.push_back(angle(i));
    std::sort(sorted_angles.begin(), sorted_angles.end());
    std::vector<double> gaps;
    for (size_t i = 1; i < sorted_angles.size(); ++i) gaps.push_back(sorted_angles[i] - sorted_angles[i - 1]);
    double spacing = 0.0;
    if (!gaps.empty()) {
        std::nth_element(gaps.begin(), gaps.begin() + gaps.size() / 2, gaps.end());
        spacing = gaps[gaps.size() / 2];
    }
    const double max_spread = std::min(std::max(3.5 * spacing, 0.05), geo_spread);
    const size_t per = (size_t)std::min(4, std::max(1, views_per_group));
    rec->ordered.clear();
    rec->groups.clear();
    rec->ordered.reserve(views.size());
    for (size_t k = 0; k < beam.size();) {
        const int c = cls[beam[k]];
        size_t e = k;
        while (e < beam.size() && e - k < per && cls[beam[e]] == c && angle(beam[e]) - angle(beam[k]) <= max_spread)
            ++e;
        rec->groups.insert(rec->groups.end(), {(int)rec->ordered.size(), (int)(e - k), c / 2, (c & 1) ? 1 : -1});
        rec->stage_used[c / 2] = true;
        for (size_t j = k; j < e; ++j) rec->ordered.push_back(views[beam[j]]);
        k = e;
    }
    rec->nbeam = (int)rec->ordered.size();
    for (int i : rest) rec->ordered.push_back(views[i]);
    // Groups of fewer than four views leave view slots of a workgroup empty:
    // when no group holds more than two views, every workgroup traces two
    // tiles of each of its views, and when every group holds one view, 2 x 2
    // tiles of it, provided the launch keeps a workgroup for every CU of the
    // 256-CU MI355X (a single view of a 256 x 256 detector stacked into 64
    // workgroups measured 0.116 against 0.058 ms unstacked; of a 512 x 512
    // one, into 256, 0.221 against 0.250 ms).
    int most = 0;
    for (int k = 0; k < rec->ngroups(); ++k) most = std::max(most, rec->groups[4 * k + 1]);
    rec->stack = 1;
    if (most == 1 || most == 2) {
        const int st = most == 1 ? 4 : 2, edge_v = st == 4 ? 32 : 16;
        const long long tiles = (long long)((g.nDetec[0] + 31) / 32) * ((g.nDetec[1] + edge_v - 1) / edge_v);
        if (tiles * rec->ngroups() >= 256) rec->stack = st;
    }
}

namespace {
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

size_t ax_workspace_bytes(const HostGeometry& g, const ForwardRecords& rec, bool interp) {
    size_t bytes = 0;
    for (int a = 0; a < 2; ++a)
        if (rec.stage_used[a]) bytes += align256((size_t)stage_elements(g, a) * sizeof(float));
    if (rec.nrest() > 0) {
        const BrickGrid b = interp ? interp_bricks(g) : siddon_bricks(g);
        bytes += align256((size_t)b.elements() * sizeof(float));
    }
    return bytes;
}

int ax_launch(const void* d_img, int dtype, const HostGeometry& g, const ForwardRecords& rec, const FwdView* d_views,
              const int* d_groups, int projection_type, int layout, float* d_proj, char* ws, int max_planes,
              hipStream_t stream) {
    const bool interp = projection_type == TIGRE_PROJ_INTERPOLATED;
    float* stage[2] = {nullptr, nullptr};
    for (int a = 0; a < 2; ++a) {
        if (!rec.stage_used[a]) continue;
        stage[a] = reinterpret_cast<float*>(ws);
        ws += align256((size_t)stage_elements(g, a) * sizeof(float));
        TIGRE_CHECK(launch_stage_volume(d_img, dtype, g, a, stage[a], stream));
    }
    if (rec.ngroups() > 0)
        TIGRE_CHECK(launch_ax_beam(stage[0], stage[1], g, d_views, d_groups, rec.ngroups(), rec.stack,
                                   projection_type, layout, d_proj, max_planes, stream));
    if (rec.nrest() > 0) {
        const BrickGrid b = interp ? interp_bricks(g) : siddon_bricks(g);
        float* bricked = reinterpret_cast<float*>(ws);
        TIGRE_CHECK(launch_brick_volume(d_img, dtype, g, b, bricked, stream));
        if (interp)
            TIGRE_CHECK(launch_ax_interp(bricked, g, b, d_views + rec.nbeam, rec.nrest(), layout, d_proj, stream));
        else
            TIGRE_CHECK(launch_ax_siddon(bricked, g, b, d_views + rec.nbeam, rec.nrest(), layout, d_proj, stream));
    }
    return TIGRE_OK;
}

int ax_device(const void* d_img, const HostGeometry& g, void* d_proj_out, int projection_type, int layout,
              int dtype, hipStream_t stream) {
    TIGRE_REQUIRE(projection_type == TIGRE_PROJ_SIDDON || projection_type == TIGRE_PROJ_INTERPOLATED,
                  "unknown projection type");
    TIGRE_REQUIRE(layout == TIGRE_LAYOUT_PYTHON || layout == TIGRE_LAYOUT_MATLAB, "unknown detector layout");
    TIGRE_REQUIRE(valid_dtype(dtype), "unknown storage type");
    const bool interp = projection_type == TIGRE_PROJ_INTERPOLATED;
    StreamScratch scratch(stream);
    // narrow storage: the kernels write float projections, rounded at the end
    const long long nproj = (long long)g.nDetec[0] * g.nDetec[1] * g.nangles;
    float* d_proj = static_cast<float*>(d_proj_out);
    if (dtype != TIGRE_DTYPE_F32) TIGRE_CHECK(scratch.alloc(&d_proj, (size_t)nproj * sizeof(float)));
    ForwardRecords rec;
    build_forward_records(g, interp, 4, &rec);
    void* d_views = nullptr;
    void* d_groups = nullptr;
    TIGRE_CHECK(scratch.upload(&d_views, rec.ordered.data(), rec.ordered.size() * sizeof(FwdView)));
    if (rec.ngroups() > 0) TIGRE_CHECK(scratch.upload(&d_groups, rec.groups.data(), rec.groups.size() * sizeof(int)));
    char* ws = nullptr;
    const size_t wsb = ax_workspace_bytes(g, rec, interp);
    if (wsb) TIGRE_CHECK(scratch.alloc(&ws, wsb));
    TIGRE_CHECK(ax_launch(d_img, dtype, g, rec, (const FwdView*)d_views, (const int*)d_groups, projection_type, layout,
                          d_proj, ws, 8, stream));
    if (dtype != TIGRE_DTYPE_F32) TIGRE_CHECK(launch_store_converted(d_proj, d_proj_out, nproj, dtype, stream));
    return TIGRE_OK;
}

size_t atb_workspace_bytes(const HostGeometry& g) {
    const size_t QU = (size_t)quad_det(g.nDetec[0]);
    return align256((QU * quad_det(g.nDetec[1]) * (size_t)g.nangles + QU * kQuadSlackRows) * 4 * sizeof(float));
}

int atb_launch(const void* d_proj, int dtype, const HostGeometry& g, const BwdView* d_views, bool axial,
               int backprojection_type, int layout, float* d_vol, char* ws, int zper, const AtbUpdate& ep,
               hipStream_t stream) {
    float* quads = reinterpret_cast<float*>(ws);
    TIGRE_CHECK(launch_quad_projections(d_proj, dtype, g.nDetec[0], g.nDetec[1], g.nangles, layout, quads, stream));
    TIGRE_CHECK(launch_atb(quads, g, d_views, backprojection_type == TIGRE_BACKPROJ_MATCHED, axial, d_vol, zper, ep,
                           stream));
    return TIGRE_OK;
}

int atb_device(const void* d_proj, const HostGeometry& g, void* d_vol_out, int backprojection_type, int layout,
               int dtype, hipStream_t stream) {
    TIGRE_REQUIRE(backprojection_type == TIGRE_BACKPROJ_FDK || backprojection_type == TIGRE_BACKPROJ_MATCHED,
                  "unknown backprojection type");
    TIGRE_REQUIRE(layout == TIGRE_LAYOUT_PYTHON || layout == TIGRE_LAYOUT_MATLAB, "unknown detector layout");
    TIGRE_REQUIRE(valid_dtype(dtype), "unknown storage type");
    StreamScratch scratch(stream);
    const long long nvox = (long long)g.nVoxel[0] * g.nVoxel[1] * g.nVoxel[2];
    float* d_vol = static_cast<float*>(d_vol_out);
    if (dtype != TIGRE_DTYPE_F32) TIGRE_CHECK(scratch.alloc(&d_vol, (size_t)nvox * sizeof(float)));
    std::vector<BwdView> views;
    backward_views(g, &views);
    const bool axial = backward_is_axial(g, views);
    void* d_views = nullptr;
    TIGRE_CHECK(scratch.upload(&d_views, views.data(), views.size() * sizeof(BwdView)));
    char* ws = nullptr;
    TIGRE_CHECK(scratch.alloc(&ws, atb_workspace_bytes(g)));
    TIGRE_CHECK(atb_launch(d_proj, dtype, g, (const BwdView*)d_views, axial, backprojection_type, layout, d_vol, ws,
                           16, AtbUpdate{}, stream));
    if (dtype != TIGRE_DTYPE_F32) TIGRE_CHECK(launch_store_converted(d_vol, d_vol_out, nvox, dtype, stream));
    return TIGRE_OK;
}

}  // namespace tigre
