// Internal definitions shared by the host core and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tigre_amd.h"

namespace tigre {

// ---------------------------------------------------------------------------
// Error handling: every failure becomes a status code plus a thread-local
// message; nothing exits the process.
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
const char* last_error();

struct Status {
    int code = TIGRE_OK;
    bool ok() const { return code == TIGRE_OK; }
};

#define TIGRE_HIP_TRY(expr)                                                              \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            ::tigre::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
            return TIGRE_ERR_HIP;                                                        \
        }                                                                                \
    } while (0)

#define TIGRE_REQUIRE(cond, msg)                                                         \
    do {                                                                                 \
        if (!(cond)) {                                                                   \
            ::tigre::set_error(msg);                                                     \
            return TIGRE_ERR_INVALID_ARGUMENT;                                           \
        }                                                                                \
    } while (0)

#define TIGRE_CHECK(call)                                                                \
    do {                                                                                 \
        int _s = (call);                                                                 \
        if (_s != TIGRE_OK) return _s;                                                   \
    } while (0)

// ---------------------------------------------------------------------------
// Per-view parameter records, computed on the host in FP64 from the float32
// geometry and uploaded once per call.
// ---------------------------------------------------------------------------

// Forward projectors.  Positions in voxel-index units.  For the Siddon
// projector the volume corner is at 0 (voxel i spans [i, i+1)); for the
// interpolated projector voxel centres sit at integers.
struct FwdView {
    double src[3];     // source
    double org[3];     // centre of detector pixel (u = 0, row counted from the top = 0)
    double du[3];      // step per detector column u
    double dv[3];      // step per row counted from the top (pv = nV - 1 - v)
    double first;      // interpolated: first sample index floor(crop / accuracy)
    double cap;        // interpolated: (2 DSO / dmin + crop) / accuracy
    double pad[2];     // pad[0]: index of this view's projection in the output;
                       // pad[1]: 1 for a parallel beam (every pixel has its own source), else 0
};
static_assert(sizeof(FwdView) == 16 * sizeof(double), "FwdView layout");

// Start of the ray of pixel (u, pv): the view's source (cone beam) or, for a
// parallel beam, the source moved with the pixel (Siddon_projection_parallel.cu
// :144-146), so all rays of the view are parallel.  For a cone beam the shift
// is exactly zero.
__host__ __device__ inline void ray_source(const FwdView& w, double u, double pv, double s[3]) {
    const double k = w.pad[1];
    for (int c = 0; c < 3; ++c) s[c] = w.src[c] + k * (u * w.du[c] + pv * w.dv[c]);
}

// Source-to-pixel vector of pixel (u, pv); for a parallel beam the pixel
// shifts of source and pixel cancel, so it is taken as org - src exactly.
__host__ __device__ inline double ray_dir(const FwdView& w, double u, double pv, int c) {
    return w.pad[1] != 0.0 ? w.org[c] - w.src[c] : w.org[c] + u * w.du[c] + pv * w.dv[c] - w.src[c];
}

// Interpolated projector's sample count ceil(|d| / accuracy)
// (ray_interpolated_projection.cu:164-166).  A parallel beam's |d| is DSD in
// voxel units, often an exact multiple of the accuracy, where the count hangs
// on the last bit of the geometry; values within 1e-9 of an integer count as
// that integer.
__host__ __device__ inline double sample_count(const FwdView& w, double len, double acc) {
    return ceil(len / acc - (w.pad[1] != 0.0 ? 1e-9 : 0.0));
}

// Backprojectors.  x in mm, y and z in detector-pixel units (the frame of
// the reference's computeDeltasCube).
struct BwdView {
    float org[3];      // voxel (0,0,0) centre minus source, with the COR shift in y
    float dX[3];
    float dY[3];
    float dZ[3];
    float src[3];      // source
    float tnum;        // DSO - DSD - src.x
    float uoff;        // nDetecU/2 - 0.5 + 1 (texture coordinate -> pixel index + 1)
    float voff;        // nDetecV/2 - 0.5 + 1
    // FDK weight (voxel_backprojection.cu:243-249): W = 1/U^2 with
    // U = (DSO + ry sin - rx cos)/DSO, affine in the voxel index:
    // U = wfdk[0] + ix wfdk[1] + iy wfdk[2]
    float wfdk[3];
    float DSD;
    // matched weight (voxel_backprojection2.cu:201-279), W = L^3 / (DSD |S - x|^2),
    // L = |S - D(u, v)|; with every length divided by DSD, W = l2^(3/2) / s2 where
    // l2 = 1 + (Day/DSD)^2 + (Dz/DSD)^2 and s2 = |S - x|^2 / DSD^2
    float mrv0[3];          // real position of voxel (0,0,0) / DSD
    float mdv[3];           // voxel size / DSD
    float mS[2];            // real source position (x, y) / DSD
    float mdet[4];          // dU / DSD, dV / DSD, detector offsets day0 / DSD, dz0 / DSD (coordinates pixel + 1)
    float pad[6];
};
static_assert(sizeof(BwdView) == 40 * sizeof(float), "BwdView layout");

// ---------------------------------------------------------------------------
// Host-side geometry (FP64 copy of the caller's float32 geometry).
// ---------------------------------------------------------------------------
struct HostGeometry {
    int nVoxel[3];          // X, Y, Z
    double sVoxel[3];
    double dVoxel[3];
    int nDetec[2];          // U, V
    double sDetec[2];
    double dDetec[2];
    double accuracy;
    int nangles;
    std::vector<double> angles;    // 3 per view
    std::vector<double> DSO, DSD, COR;
    std::vector<double> offOrig;   // 3 per view (X, Y, Z)
    std::vector<double> offDetec;  // 2 per view (U, V)
    std::vector<double> rot;       // 3 per view (roll, pitch, yaw)
    bool parallel = false;         // parallel-beam operators (the reference's *_parallel functions)
};

int import_geometry(const TigreGeometry* geo, const float* angles, int nangles, HostGeometry* out);

// Per-view records (geometry.cpp).
void forward_views(const HostGeometry& g, bool voxel_centre_origin, std::vector<FwdView>* out);
void backward_views(const HostGeometry& g, std::vector<BwdView>* out);
bool backward_is_axial(const HostGeometry& g, const std::vector<BwdView>& v);

// ---------------------------------------------------------------------------
// Device helpers (runtime.cpp)
// ---------------------------------------------------------------------------
int select_device(const int* gpuids, int ngpus);
int upload_async(void** d_ptr, const void* host, size_t bytes, hipStream_t stream);

// Stream-ordered scratch of one call: every block allocated through it is
// released with hipFreeAsync on the same stream when the guard goes out of
// scope, on success and on every early error return alike.
class StreamScratch {
  public:
    explicit StreamScratch(hipStream_t stream) : stream_(stream) {}
    ~StreamScratch() {
        for (void* p : blocks_) (void)hipFreeAsync(p, stream_);
    }
    StreamScratch(const StreamScratch&) = delete;
    StreamScratch& operator=(const StreamScratch&) = delete;
    template <typename T>
    int alloc(T** p, size_t bytes) {
        void* q = nullptr;
        TIGRE_HIP_TRY(hipMallocAsync(&q, bytes, stream_));
        blocks_.push_back(q);
        *p = static_cast<T*>(q);
        return TIGRE_OK;
    }
    int upload(void** p, const void* host, size_t bytes) {
        TIGRE_CHECK(upload_async(p, host, bytes, stream_));
        blocks_.push_back(*p);
        return TIGRE_OK;
    }

  private:
    hipStream_t stream_;
    std::vector<void*> blocks_;
};

// Optional HIP-event timing of kernel launches (TIGRE_KERNEL_* ids).  The
// scope records a start event on construction and a stop event on
// destruction; both are no-ops while timing is disabled.
class KernelTimerScope {
  public:
    KernelTimerScope(int kernel, hipStream_t stream);
    ~KernelTimerScope();
    KernelTimerScope(const KernelTimerScope&) = delete;
    KernelTimerScope& operator=(const KernelTimerScope&) = delete;

  private:
    int kernel_;
    hipStream_t stream_;
    hipEvent_t start_ = nullptr;
};
int kernel_timing_enable(int on);
int kernel_time_collect(int kernel, double* total_ms, int* launches);

// Bricked copies of a volume for the forward projectors.
struct BrickGrid {
    int nbx, nby, nbz;   // bricks per axis
    int per_brick;       // floats per brick: 64 (4^3 bricks) or 125 (5^3 apron bricks)
    long long elements() const { return (long long)nbx * nby * nbz * per_brick; }
};
BrickGrid siddon_bricks(const HostGeometry& g);
BrickGrid interp_bricks(const HostGeometry& g);

// Launchers (one per operator; all enqueue on `stream`).
int launch_brick_volume(const void* vol, int dtype, const HostGeometry& g, const BrickGrid& b, float* bricked,
                        hipStream_t stream);
int launch_ax_siddon(const float* bricked, const HostGeometry& g, const BrickGrid& b, const FwdView* d_views,
                     int nviews, int layout, float* proj, hipStream_t stream);
// Beam-staged Siddon (ax_beam.hip).  Staging copy for principal axis P
// (0 = x, 1 = y): [p][z][a] with a zero border, a = the other horizontal axis.
long long stage_elements(const HostGeometry& g, int principal);
int launch_stage_volume(const void* vol, int dtype, const HostGeometry& g, int principal, float* stage,
                        hipStream_t stream);
// Beam-staged forward projection (Siddon or interpolated); d_groups: ngroups
// records {first view, view count (1..4), principal, direction}; stack: the
// detector tiles a workgroup traces per view (1; 2 when no group holds more
// than two views; 4, as 2 x 2, when every group holds one).
int launch_ax_beam(const float* stage_x, const float* stage_y, const HostGeometry& g, const FwdView* d_views,
                   const int* d_groups, int ngroups, int stack, int projection_type, int layout, float* proj,
                   int max_planes, hipStream_t stream);
int launch_ax_interp(const float* bricked, const HostGeometry& g, const BrickGrid& b, const FwdView* d_views,
                     int nviews, int layout, float* proj, hipStream_t stream);
int launch_quad_projections(const void* proj, int dtype, int nU, int nV, int nviews, int layout, float* quads,
                            hipStream_t stream);
// Optional SART epilogue of the backprojector: vol = max?(x + inv_v * Atb, 0),
// inv_v an nY x nX image broadcast over z (x == nullptr: plain store).
struct AtbUpdate {
    const float* x = nullptr;
    const float* inv_v = nullptr;
    int noneg = 0;
};
// zper: slices per lane, 4, 8 or 16.
int launch_atb(const float* quads, const HostGeometry& g, const BwdView* d_views, bool matched,
               bool axial, float* vol, int zper, const AtbUpdate& ep, hipStream_t stream);
int run_pocs_tv(const float* img, float* dst, float alpha, long nx, long ny, long nz, int maxiter,
                hipStream_t stream);
int run_aw_pocs_tv(const float* img, float* dst, float alpha, float delta, long nx, long ny, long nz, int maxiter,
                   hipStream_t stream);
int run_piccs_tv(const float* img, const float* prior, float* dst, float alpha, float ratio, long nx, long ny, long nz,
                 int maxiter, hipStream_t stream);
int run_tvdenoise(const float* src, float* dst, float lambda, const float* spacing, long nx, long ny,
                  long nz, int maxiter, hipStream_t stream);

// Forward-projection records: view records (beam-grouped views first, the
// gather kernels' views after them), workgroup groups {first, count,
// principal, direction}, and which staging copies the groups read.
struct ForwardRecords {
    std::vector<FwdView> ordered;
    std::vector<int> groups;
    int nbeam = 0;
    bool stage_used[2] = {false, false};
    int stack = 1;  // detector tiles per view of a beam workgroup: 1, 2 or 4
    int ngroups() const { return (int)groups.size() / 4; }
    int nrest() const { return (int)ordered.size() - nbeam; }
};
// views_per_group: 1..4 views per beam workgroup.
void build_forward_records(const HostGeometry& g, bool interp, int views_per_group, ForwardRecords* rec);
// Scratch of one forward projection (staging copies, bricks), bytes; the
// layout used by ax_launch: each piece 256-byte aligned, in that order.
size_t ax_workspace_bytes(const HostGeometry& g, const ForwardRecords& rec, bool interp);
int ax_launch(const void* d_img, int dtype, const HostGeometry& g, const ForwardRecords& rec, const FwdView* d_views,
              const int* d_groups, int projection_type, int layout, float* d_proj, char* ws, int max_planes,
              hipStream_t stream);
// Scratch of one backprojection (the quad projections), bytes.
size_t atb_workspace_bytes(const HostGeometry& g);
int atb_launch(const void* d_proj, int dtype, const HostGeometry& g, const BwdView* d_views, bool axial,
               int backprojection_type, int layout, float* d_vol, char* ws, int zper, const AtbUpdate& ep,
               hipStream_t stream);

// r = W (p - Ax) over n elements, written over ax (sart.hip).
int launch_sart_residual(const float* proj, const float* w, float* ax, long long n, hipStream_t stream);

// ---------------------------------------------------------------------------
// Multi-job launches: one launch runs the same stage of many independent
// problems.  Job j owns the blocks [begin, begin + its block count) of the
// 1-D grid; jobs are sorted by begin, the first begins at 0.
// ---------------------------------------------------------------------------
struct StageJob {    // staging copy of an FP32 volume for principal axis `principal`
    const float* vol;
    float* out;
    int nx, ny, nz, principal, na_pad, gx, gy, begin;  // grid (gx, gy, nz + 2)
};
struct BeamJob {     // beam-staged Siddon projection of one view list
    const float* stage_x;
    const float* stage_y;
    const FwdView* views;
    const int* groups;
    float* out;
    double acc, dvx, dvy, dvz;
    int nx, ny, nz, na_x, na_y, ngroups, nU, nV, max_planes, begin, stack;
};
struct ResidualJob { // r = W (p - Ax), n elements
    const float* proj;
    const float* w;
    float* ax;
    long long n;
    int blocks, begin;
};
struct QuadJob {     // quad copy of FP32 projections, Python layout
    const float* proj;
    float* quads;
    int nU, nV, nviews, blocks, begin, pad;
};
struct AtbJob {      // FDK backprojection (axial views) with the SART update epilogue
    const float* quads;
    const BwdView* views;
    float* vol;
    const float* x;
    const float* inv_v;
    float scale;
    int nU, nV, nviews, nx, ny, nz, noneg, gx, gy, begin;
};

// Index of the job that owns block b (binary search over the begins; every
// lane of a workgroup computes the same index).
template <class Job>
__device__ __forceinline__ int find_job(const Job* __restrict__ jobs, int njobs, int b) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].begin <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

int launch_stage_multi(const StageJob* d_jobs, int njobs, long long blocks, hipStream_t stream);
int launch_ax_beam_multi(const BeamJob* d_jobs, int njobs, long long blocks, hipStream_t stream);
int launch_residual_multi(const ResidualJob* d_jobs, int njobs, long long blocks, hipStream_t stream);
int launch_quad_multi(const QuadJob* d_jobs, int njobs, long long blocks, hipStream_t stream);
int launch_atb_multi(const AtbJob* d_jobs, int njobs, long long blocks, int zper, hipStream_t stream);
// Block counts of one job of each kind (host side).
int stage_job_blocks(const HostGeometry& g, int principal, int* gx, int* gy);
long long beam_job_blocks(const HostGeometry& g, int ngroups, int stack);
int residual_job_blocks(long long n);
int quad_job_blocks(int nU, int nV, int nviews);

// Convolutions of the learned proximal network (cnn.hip).
size_t cnn_wgrad_workspace(int N, int D, int H, int W);
int launch_cnn_conv(const void* x, int x_mode, const unsigned short* w, const float* bias, const void* aux, void* y,
                    int epi, int N, int D, int H, int W, hipStream_t stream);
int launch_cnn_wgrad(const void* x, int x_mode, const void* dy, int dy_mode, float* dw, int N, int D, int H, int W,
                     void* ws, size_t ws_bytes, hipStream_t stream);
// The same without a workspace: partials in the library's static device array
// of cnn_wgrad_static_bytes() bytes (cnn.hip).
size_t cnn_wgrad_static_bytes();
int launch_cnn_wgrad_noalloc(const void* x, int x_mode, const void* dy, int dy_mode, float* dw, int N, int D, int H,
                             int W, hipStream_t stream);

// Operator drivers on device pointers (ops.cpp).
int ax_device(const void* d_img, const HostGeometry& g, void* d_proj, int projection_type, int layout, int dtype,
              hipStream_t stream);
int atb_device(const void* d_proj, const HostGeometry& g, void* d_vol, int backprojection_type, int layout,
               int dtype, hipStream_t stream);

// Storage types (TIGRE_DTYPE_*): element i of a buffer of `dtype`, widened to
// float.  Used by the re-layout kernels, so every operator computes in float.
__device__ __forceinline__ float load_storage(const void* __restrict__ p, long long i, int dtype) {
    if (dtype == TIGRE_DTYPE_BF16) return __uint_as_float((unsigned)static_cast<const unsigned short*>(p)[i] << 16);
    if (dtype == TIGRE_DTYPE_F16) return (float)static_cast<const _Float16*>(p)[i];
    return static_cast<const float*>(p)[i];
}
inline size_t storage_bytes(int dtype) { return dtype == TIGRE_DTYPE_F32 ? 4 : 2; }
inline bool valid_dtype(int dtype) {
    return dtype == TIGRE_DTYPE_F32 || dtype == TIGRE_DTYPE_BF16 || dtype == TIGRE_DTYPE_F16;
}
// float -> dtype, round to nearest even (storage.hip)
int launch_store_converted(const float* src, void* dst, long long n, int dtype, hipStream_t stream);

// Quad projection dimensions: tap origins -1 .. n on each detector axis.
__host__ __device__ inline long quad_det(int n) { return (long)n + 2; }
// Rows of quads allocated after the last view of a quad copy: the staged
// backprojector copies whole 4 x 16-quad blocks of a detector window, which
// may run up to 23 rows past a view's last row and 15 quads past a row's end
// (into the next view or row, never read; past the last view, into this slack).
constexpr long kQuadSlackRows = 25;

}  // namespace tigre
