// Projection plans: the per-view records of one geometry prepared and
// uploaded once, so that forward and backward projections launch with no
// host synchronisation, no allocation and no record upload.  Scratch comes
// from a caller workspace, which makes every plan launch capturable in a
// HIP graph (the kernels issue no memset; the output buffers are written in
// full).  The tunables are the views per beam workgroup (1, 2 or 4), the
// LDS slab depth of the beam kernel (1..8 principal planes) and the slices
// per lane of the backprojector (4, 8 or 16).
#include <algorithm>
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

// A batch: one SART subset update of each of many independent problems, as
// five multi-job launches (staging copies, beam projections, residuals, quad
// copies, FDK backprojections with the update epilogue).  The job tables live
// on the device from creation on, so a run only launches kernels.  Views a
// beam workgroup does not take (gather kernels) and backprojections the staged
// multi-job kernel does not cover run as per-problem launches in between.
struct TigreBatch_ {
    int device = 0;
    void* d_tables = nullptr;
    const StageJob* d_stage = nullptr;
    const BeamJob* d_beam = nullptr;
    const ResidualJob* d_res = nullptr;
    const QuadJob* d_quad = nullptr;
    const AtbJob* d_atb[3] = {nullptr, nullptr, nullptr};  // slices per lane 4, 8, 16
    int n_stage = 0, n_beam = 0, n_res = 0, n_quad = 0, n_atb[3] = {0, 0, 0};
    long long b_stage = 0, b_beam = 0, b_res = 0, b_quad = 0, b_atb[3] = {0, 0, 0};
    std::vector<TigreBatchStep> steps;
    std::vector<int> rest_steps;  // steps with views for the gather kernels
    std::vector<int> atb_steps;   // steps whose backprojection runs on its own
    ~TigreBatch_() {
        if (d_tables) (void)hipFree(d_tables);
    }
};

namespace {

int zslot(int zper) { return zper == 4 ? 0 : (zper == 8 ? 1 : 2); }

// Byte offset of the staging copy for principal axis a in a forward workspace
// (the layout ax_launch uses), and of the bricked copy after them.
size_t stage_offset(const HostGeometry& g, const ForwardRecords& rec, int a) {
    size_t off = 0;
    for (int k = 0; k < a; ++k)
        if (rec.stage_used[k]) off += align256((size_t)stage_elements(g, k) * sizeof(float));
    return off;
}

bool atb_multi_ok(const TigrePlan_* p) {
    // the staged kernel: FDK weight, axial views, row pairs in 16 bits
    return p->axial && !p->g.parallel && p->g.nDetec[1] + 1 < 65536;
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

int tigre_cnn_wgrad_noalloc(const void* d_x, int x_mode, const void* d_dy, int dy_mode, float* d_dw, int n, int d,
                            int h, int w, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_x != nullptr && d_dy != nullptr && d_dw != nullptr, "null device pointer");
    return launch_cnn_wgrad_noalloc(d_x, x_mode, d_dy, dy_mode, d_dw, n, d, h, w, (hipStream_t)stream);
}

int tigre_cnn_wgrad_noalloc_limit(size_t* bytes) {
    set_error("");
    TIGRE_REQUIRE(bytes != nullptr, "null argument");
    *bytes = cnn_wgrad_static_bytes();
    return TIGRE_OK;
}

int tigre_batch_create(const TigreBatchStep* steps, int nsteps, TigreBatch* out) {
    set_error("");
    TIGRE_REQUIRE(out != nullptr, "null batch handle");
    *out = nullptr;
    TIGRE_REQUIRE(steps != nullptr && nsteps > 0, "a batch needs at least one step");
    int dev = 0;
    TIGRE_HIP_TRY(hipGetDevice(&dev));
    std::vector<StageJob> stage;
    std::vector<BeamJob> beam;
    std::vector<ResidualJob> res;
    std::vector<QuadJob> quad;
    std::vector<AtbJob> atb[3];
    long long b_stage = 0, b_beam = 0, b_res = 0, b_quad = 0, b_atb[3] = {0, 0, 0};
    std::vector<int> rest_steps, atb_steps;
    std::vector<const void*> owned;  // buffers a step writes: no two steps may share one
    for (int i = 0; i < nsteps; ++i) {
        const TigreBatchStep& st = steps[i];
        const TigrePlan_* p = st.plan;
        TIGRE_CHECK(plan_on_device(p));
        TIGRE_REQUIRE(!p->g.parallel, "batched SART steps are cone beam");
        TIGRE_REQUIRE(st.d_x && st.d_proj && st.d_w && st.d_ax && st.d_inv_v && st.d_workspace,
                      "null pointer in a batch step");
        TIGRE_REQUIRE((reinterpret_cast<uintptr_t>(st.d_workspace) & 255) == 0, "workspace must be 256-byte aligned");
        TIGRE_REQUIRE(st.workspace_bytes >= plan_workspace(p, TIGRE_DTYPE_F32),
                      "workspace smaller than tigre_plan_workspace_size");
        TIGRE_REQUIRE(((reinterpret_cast<uintptr_t>(st.d_proj) | reinterpret_cast<uintptr_t>(st.d_w) |
                        reinterpret_cast<uintptr_t>(st.d_ax)) & 15) == 0,
                      "projection buffers must be 16-byte aligned");
        for (const void* q : {(const void*)st.d_x, (const void*)st.d_ax, (const void*)st.d_workspace}) {
            for (const void* o : owned) TIGRE_REQUIRE(o != q, "two batch steps share an image, scratch or workspace");
            owned.push_back(q);
        }
        const HostGeometry& g = p->g;
        const ForwardRecords& rec = p->fwd[0];
        TIGRE_REQUIRE(stage_elements(g, 0) < (1LL << 32) && stage_elements(g, 1) < (1LL << 32),
                      "volume too large for 32-bit staging offsets");
        char* ws = static_cast<char*>(st.d_workspace);
        float* stg[2] = {nullptr, nullptr};
        for (int a = 0; a < 2; ++a) {
            if (!rec.stage_used[a]) continue;
            stg[a] = reinterpret_cast<float*>(ws + stage_offset(g, rec, a));
            StageJob j;
            j.vol = st.d_x;
            j.out = stg[a];
            j.nx = g.nVoxel[0];
            j.ny = g.nVoxel[1];
            j.nz = g.nVoxel[2];
            j.principal = a;
            j.na_pad = (a == 0 ? (g.nVoxel[1] + 3) & ~3 : (g.nVoxel[0] + 3) & ~3);
            j.begin = (int)b_stage;
            b_stage += stage_job_blocks(g, a, &j.gx, &j.gy);
            TIGRE_REQUIRE(j.gy <= 65535 && g.nVoxel[2] + 2 <= 65535, "volume too large for the staging copy");
            stage.push_back(j);
        }
        if (rec.ngroups() > 0) {
            BeamJob j;
            j.stage_x = stg[0];
            j.stage_y = stg[1];
            j.views = static_cast<const FwdView*>(p->d_fwd[0]);
            j.groups = static_cast<const int*>(p->d_groups[0]);
            j.out = st.d_ax;
            j.acc = g.accuracy;
            j.dvx = g.dVoxel[0];
            j.dvy = g.dVoxel[1];
            j.dvz = g.dVoxel[2];
            j.nx = g.nVoxel[0];
            j.ny = g.nVoxel[1];
            j.nz = g.nVoxel[2];
            j.na_x = (g.nVoxel[1] + 3) & ~3;
            j.na_y = (g.nVoxel[0] + 3) & ~3;
            j.ngroups = rec.ngroups();
            j.nU = g.nDetec[0];
            j.nV = g.nDetec[1];
            j.max_planes = p->max_planes;
            j.begin = (int)b_beam;
            j.stack = rec.stack;
            b_beam += beam_job_blocks(g, rec.ngroups(), rec.stack);
            beam.push_back(j);
        }
        if (rec.nrest() > 0) rest_steps.push_back(i);
        const long long nproj = (long long)g.nDetec[0] * g.nDetec[1] * g.nangles;
        {
            ResidualJob j;
            j.proj = st.d_proj;
            j.w = st.d_w;
            j.ax = st.d_ax;
            j.n = nproj;
            j.blocks = residual_job_blocks(nproj);
            j.begin = (int)b_res;
            b_res += j.blocks;
            res.push_back(j);
        }
        {
            QuadJob j;
            j.proj = st.d_ax;
            j.quads = reinterpret_cast<float*>(ws);
            j.nU = g.nDetec[0];
            j.nV = g.nDetec[1];
            j.nviews = g.nangles;
            j.blocks = quad_job_blocks(j.nU, j.nV, j.nviews);
            j.begin = (int)b_quad;
            j.pad = 0;
            b_quad += j.blocks;
            quad.push_back(j);
        }
        TIGRE_REQUIRE(quad_det(g.nDetec[0]) * 16 < (1LL << 24) &&
                          quad_det(g.nDetec[0]) * quad_det(g.nDetec[1]) * 16 < (1LL << 32),
                      "detector too large for one backprojection launch");
        if (atb_multi_ok(p)) {
            const int k = zslot(p->zper);
            AtbJob j;
            j.quads = reinterpret_cast<const float*>(ws);
            j.views = static_cast<const BwdView*>(p->d_bwd);
            j.vol = st.d_x;
            j.x = st.d_x;
            j.inv_v = st.d_inv_v;
            j.scale = 1.0f;
            j.nU = g.nDetec[0];
            j.nV = g.nDetec[1];
            j.nviews = g.nangles;
            j.nx = g.nVoxel[0];
            j.ny = g.nVoxel[1];
            j.nz = g.nVoxel[2];
            j.noneg = st.noneg ? 1 : 0;
            j.gx = (j.nx + 15) / 16;
            j.gy = (j.ny + 15) / 16;
            const int gz = (j.nz + p->zper - 1) / p->zper;
            TIGRE_REQUIRE(j.gy <= 65535 && gz <= 65535, "volume too large for one launch");
            j.begin = (int)b_atb[k];
            b_atb[k] += (long long)j.gx * j.gy * gz;
            atb[k].push_back(j);
        } else {
            atb_steps.push_back(i);
        }
    }
    TIGRE_REQUIRE(b_stage < (1LL << 31) && b_beam < (1LL << 31) && b_res < (1LL << 31) && b_quad < (1LL << 31) &&
                      b_atb[0] < (1LL << 31) && b_atb[1] < (1LL << 31) && b_atb[2] < (1LL << 31),
                  "batch too large for one launch per stage");
    TigreBatch_* b = new TigreBatch_();
    int st = [&]() -> int {
        b->device = dev;
        b->steps.assign(steps, steps + nsteps);
        b->rest_steps = rest_steps;
        b->atb_steps = atb_steps;
        // one table allocation: every array at a 256-byte boundary
        const size_t sizes[7] = {stage.size() * sizeof(StageJob), beam.size() * sizeof(BeamJob),
                                 res.size() * sizeof(ResidualJob), quad.size() * sizeof(QuadJob),
                                 atb[0].size() * sizeof(AtbJob), atb[1].size() * sizeof(AtbJob),
                                 atb[2].size() * sizeof(AtbJob)};
        const void* src[7] = {stage.data(), beam.data(), res.data(), quad.data(), atb[0].data(), atb[1].data(),
                              atb[2].data()};
        size_t off[7], total = 0;
        for (int k = 0; k < 7; ++k) {
            off[k] = total;
            total += align256(sizes[k]);
        }
        std::vector<char> host(total, 0);
        for (int k = 0; k < 7; ++k)
            if (sizes[k]) std::memcpy(host.data() + off[k], src[k], sizes[k]);
        TIGRE_CHECK(upload_sync(&b->d_tables, host.data(), total));
        const char* base = static_cast<const char*>(b->d_tables);
        b->d_stage = reinterpret_cast<const StageJob*>(base + off[0]);
        b->d_beam = reinterpret_cast<const BeamJob*>(base + off[1]);
        b->d_res = reinterpret_cast<const ResidualJob*>(base + off[2]);
        b->d_quad = reinterpret_cast<const QuadJob*>(base + off[3]);
        for (int k = 0; k < 3; ++k) b->d_atb[k] = reinterpret_cast<const AtbJob*>(base + off[4 + k]);
        b->n_stage = (int)stage.size();
        b->n_beam = (int)beam.size();
        b->n_res = (int)res.size();
        b->n_quad = (int)quad.size();
        for (int k = 0; k < 3; ++k) {
            b->n_atb[k] = (int)atb[k].size();
            b->b_atb[k] = b_atb[k];
        }
        b->b_stage = b_stage;
        b->b_beam = b_beam;
        b->b_res = b_res;
        b->b_quad = b_quad;
        return TIGRE_OK;
    }();
    if (st != TIGRE_OK) {
        delete b;
        return st;
    }
    *out = b;
    return TIGRE_OK;
}

int tigre_batch_sart(TigreBatch batch, void* stream) {
    set_error("");
    TIGRE_REQUIRE(batch != nullptr, "null batch");
    int dev = -1;
    TIGRE_HIP_TRY(hipGetDevice(&dev));
    TIGRE_REQUIRE(dev == batch->device, "batch used on another device than the one it was created on");
    hipStream_t s = (hipStream_t)stream;
    TigreBatch_& b = *batch;
    if (b.n_stage) TIGRE_CHECK(launch_stage_multi(b.d_stage, b.n_stage, b.b_stage, s));
    if (b.n_beam) TIGRE_CHECK(launch_ax_beam_multi(b.d_beam, b.n_beam, b.b_beam, s));
    for (int i : b.rest_steps) {  // views for the gather-driven Siddon kernel
        const TigreBatchStep& st = b.steps[i];
        const TigrePlan_* p = st.plan;
        const HostGeometry& g = p->g;
        const ForwardRecords& rec = p->fwd[0];
        const BrickGrid bg = siddon_bricks(g);
        float* bricked = reinterpret_cast<float*>(static_cast<char*>(st.d_workspace) + stage_offset(g, rec, 2));
        TIGRE_CHECK(launch_brick_volume(st.d_x, TIGRE_DTYPE_F32, g, bg, bricked, s));
        TIGRE_CHECK(launch_ax_siddon(bricked, g, bg, static_cast<const FwdView*>(p->d_fwd[0]) + rec.nbeam,
                                     rec.nrest(), TIGRE_LAYOUT_PYTHON, st.d_ax, s));
    }
    TIGRE_CHECK(launch_residual_multi(b.d_res, b.n_res, b.b_res, s));
    TIGRE_CHECK(launch_quad_multi(b.d_quad, b.n_quad, b.b_quad, s));
    const int zpers[3] = {4, 8, 16};
    for (int k = 0; k < 3; ++k)
        if (b.n_atb[k]) TIGRE_CHECK(launch_atb_multi(b.d_atb[k], b.n_atb[k], b.b_atb[k], zpers[k], s));
    for (int i : b.atb_steps) {  // backprojections outside the staged multi-job kernel
        const TigreBatchStep& st = b.steps[i];
        AtbUpdate ep;
        ep.x = st.d_x;
        ep.inv_v = st.d_inv_v;
        ep.noneg = st.noneg ? 1 : 0;
        TIGRE_CHECK(launch_atb(static_cast<const float*>(st.d_workspace), st.plan->g,
                               static_cast<const BwdView*>(st.plan->d_bwd), false, st.plan->axial, st.d_x,
                               st.plan->zper, ep, s));
    }
    return TIGRE_OK;
}

int tigre_batch_destroy(TigreBatch batch) {
    set_error("");
    if (batch == nullptr) return TIGRE_OK;
    int prev = -1;
    TIGRE_HIP_TRY(hipGetDevice(&prev));
    TIGRE_HIP_TRY(hipSetDevice(batch->device));
    delete batch;
    TIGRE_HIP_TRY(hipSetDevice(prev));
    return TIGRE_OK;
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
