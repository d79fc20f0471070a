/*
 * tigre_amd -- C ABI of the MI355X (gfx950) cone-beam projector pair and TV step.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (0 = success, the codes of the reference's Common/CUDA/errors.hpp:1-7
 * otherwise) and never exits the process or resets the device.  The message of
 * the last failure on the calling thread is available from tigre_last_error().
 *
 * Two families of entry points:
 *  - host-pointer entry points keep the reference contract (host memory in,
 *    host memory out, blocking) so an existing Cython/MEX gateway can call them;
 *  - device-pointer entry points (suffix _device) take device pointers and a
 *    hipStream_t (passed as void*), enqueue work on that stream and return
 *    without synchronising; inputs and outputs stay resident in HBM.
 */
#ifndef TIGRE_AMD_H
#define TIGRE_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: identical numbering to Common/CUDA/errors.hpp:1-7, plus one
 * code for arguments the library rejects before touching the device. */
#define TIGRE_OK                     0
#define TIGRE_ERR_HIP                1
#define TIGRE_ERR_NO_CAPABLE_DEVICES 2
#define TIGRE_ERR_NO_FREE_DEVICES    3
#define TIGRE_ERR_BAD_ASSERT         4
#define TIGRE_ERR_ASSERT_FAIL        5
#define TIGRE_ERR_INVALID_ARGUMENT   6

/* Detector layouts: the reference chooses one at compile time
 * (IS_FOR_PYTIGRE, Common/CUDA/TIGRE_common.hpp:10-23); here it is a flag. */
#define TIGRE_LAYOUT_PYTHON 0   /* projection index view*nV*nU + v*nU + u */
#define TIGRE_LAYOUT_MATLAB 1   /* projection index view*nV*nU + u*nV + v */

/* Projector / backprojector selectors (names as in Ax.py / Atb.py). */
#define TIGRE_PROJ_SIDDON       0
#define TIGRE_PROJ_INTERPOLATED 1
#define TIGRE_BACKPROJ_FDK      0
#define TIGRE_BACKPROJ_MATCHED  1

/* Storage types of volumes and projections (accumulation is always FP32). */
#define TIGRE_DTYPE_F32  0   /* float */
#define TIGRE_DTYPE_BF16 1   /* bfloat16 (upper half of an IEEE float) */
#define TIGRE_DTYPE_F16  2   /* IEEE binary16 */

/*
 * Per-projection geometry.  Field-for-field and byte-for-byte the layout of
 * the reference's struct Geometry (Common/CUDA/types_TIGRE.hpp:48-85), so a
 * gateway that fills the reference struct can pass it here unchanged.  The
 * per-view arrays hold one float per projection; alpha, theta, psi, the
 * unitX/Y/Z fields and maxLength are scratch in the reference and ignored here.
 */
typedef struct TigreGeometry {
    int   nVoxelX, nVoxelY, nVoxelZ;
    float sVoxelX, sVoxelY, sVoxelZ;
    float dVoxelX, dVoxelY, dVoxelZ;
    float *offOrigX, *offOrigY, *offOrigZ;
    float *DSO;
    int   nDetecU, nDetecV;
    float sDetecU, sDetecV;
    float dDetecU, dDetecV;
    float *offDetecU, *offDetecV;
    float *DSD;
    float *dRoll;
    float *dPitch;
    float *dYaw;
    float unitX, unitY, unitZ;
    float alpha, theta, psi;
    float *COR;
    float maxLength;
    float accuracy;
} TigreGeometry;

/* ---------------------------------------------------------------------- */
/* Host-pointer entry points (reference contract)                          */
/* ---------------------------------------------------------------------- */

/* Replaces siddon_ray_projection (Common/CUDA/Siddon_projection.hpp:54) and
 * interpolation_projection (Common/CUDA/ray_interpolated_projection.hpp:55).
 * img: nVoxelZ x nVoxelY x nVoxelX floats, x fastest.  result[i] points at
 * the nDetecU*nDetecV floats of projection i (the reference's float**).
 * angles: nangles x 3 floats (alpha, theta, psi).  gpuids/ngpus select the
 * device (only the first id is used: one device per call). */
int tigre_ax(const float* img, const TigreGeometry* geo, float** result,
             const float* angles, int nangles, int projection_type, int layout,
             const int* gpuids, int ngpus);

/* Replaces voxel_backprojection (Common/CUDA/voxel_backprojection.hpp:52)
 * and voxel_backprojection2 (Common/CUDA/voxel_backprojection2.hpp:57).
 * projections: nangles projections of nDetecU*nDetecV floats; result is
 * overwritten with nVoxelZ x nVoxelY x nVoxelX floats. */
int tigre_atb(const float* projections, const TigreGeometry* geo, float* result,
              const float* angles, int nangles, int backprojection_type, int layout,
              const int* gpuids, int ngpus);

/* Parallel beam.  Replace siddon_ray_projection_parallel
 * (Common/CUDA/Siddon_projection_parallel.hpp:56) and
 * interpolation_projection_parallel (ray_interpolated_projection_parallel.hpp:57),
 * same arguments as tigre_ax: every pixel has its own source, so the rays of a
 * view are parallel (geometry mode 'parallel', _Ax.pyx:67-87). */
int tigre_ax_parallel(const float* img, const TigreGeometry* geo, float** result,
                      const float* angles, int nangles, int projection_type, int layout,
                      const int* gpuids, int ngpus);

/* Replaces voxel_backprojection_parallel (voxel_backprojection_parallel.hpp:53):
 * each voxel is projected orthogonally onto the detector and sampled
 * bilinearly, weight 1.  The reference uses it for 'FDK' and 'matched' alike
 * (_Atb.pyx:56-76), so there is no backprojector argument. */
int tigre_atb_parallel(const float* projections, const TigreGeometry* geo, float* result,
                       const float* angles, int nangles, int layout, const int* gpuids, int ngpus);

/* Replaces pocs_tv (Common/CUDA/GD_TV.hpp:58).  image_size = {X, Y, Z}. */
int tigre_pocs_tv(const float* img, float* dst, float alpha, const long* image_size,
                  int maxIter, const int* gpuids, int ngpus);

/* Replaces aw_pocs_tv (Common/CUDA/GD_AwTV.hpp:58): adaptive-weighted TV
 * steps, difference weights exp(-(d / delta)^2). */
int tigre_aw_pocs_tv(const float* img, float* dst, float alpha, const long* image_size,
                     int maxIter, float delta, const int* gpuids, int ngpus);

/* Replaces piccs_tv (Common/CUDA/PICCS.hpp:58): steps along
 * (1 - ratio) grad TV(img) + ratio grad TV(img - prior). */
int tigre_piccs_tv(const float* img, const float* prior, float* dst, float alpha, float ratio,
                   const long* image_size, int maxIter, const int* gpuids, int ngpus);

/* Replaces tvdenoising (Common/CUDA/tv_proximal.hpp:52).  spacing = {dx, dy, dz}. */
int tigre_tvdenoising(const float* src, float* dst, float lambda, const float* spacing,
                      const long* image_size, int maxIter, const int* gpuids, int ngpus);

/* ---------------------------------------------------------------------- */
/* Device-pointer entry points (resident data, caller's stream)            */
/* ---------------------------------------------------------------------- */

/* Forward projection of a device volume into device projections laid out as
 * in tigre_ax (contiguous, nangles * nDetecU * nDetecV floats). */
int tigre_ax_device(const float* d_img, const TigreGeometry* geo, float* d_proj,
                    const float* angles, int nangles, int projection_type, int layout,
                    void* stream);

/* Backprojection of device projections into a device volume (overwritten). */
int tigre_atb_device(const float* d_proj, const TigreGeometry* geo, float* d_vol,
                     const float* angles, int nangles, int backprojection_type, int layout,
                     void* stream);

/* The two operators above on volumes and projections stored as `dtype`
 * (TIGRE_DTYPE_*; both operands share it).  Values are widened to float on
 * load, all arithmetic and accumulation is float, results are rounded to
 * `dtype` (round to nearest even) on store.  No counterpart in the reference,
 * which stores float only. */
int tigre_ax_device_typed(const void* d_img, const TigreGeometry* geo, void* d_proj,
                          const float* angles, int nangles, int projection_type, int layout,
                          int dtype, void* stream);
int tigre_atb_device_typed(const void* d_proj, const TigreGeometry* geo, void* d_vol,
                           const float* angles, int nangles, int backprojection_type, int layout,
                           int dtype, void* stream);

/* Device-pointer forms of tigre_ax_parallel / tigre_atb_parallel (storage type dtype). */
int tigre_ax_parallel_device(const void* d_img, const TigreGeometry* geo, void* d_proj,
                             const float* angles, int nangles, int projection_type, int layout,
                             int dtype, void* stream);
int tigre_atb_parallel_device(const void* d_proj, const TigreGeometry* geo, void* d_vol,
                              const float* angles, int nangles, int layout, int dtype, void* stream);

/* maxIter TV steepest-descent steps on a device volume (X fastest); d_dst may
 * equal d_img.  image_size = {X, Y, Z}. */
int tigre_pocs_tv_device(const float* d_img, float* d_dst, float alpha, const long* image_size,
                         int maxIter, void* stream);

/* Device-pointer forms of tigre_aw_pocs_tv and tigre_piccs_tv. */
int tigre_aw_pocs_tv_device(const float* d_img, float* d_dst, float alpha, const long* image_size,
                            int maxIter, float delta, void* stream);
int tigre_piccs_tv_device(const float* d_img, const float* d_prior, float* d_dst, float alpha,
                          float ratio, const long* image_size, int maxIter, void* stream);

/* ROF TV denoise of a device volume (d_dst may equal d_src). */
int tigre_tvdenoising_device(const float* d_src, float* d_dst, float lambda, const float* spacing,
                             const long* image_size, int maxIter, void* stream);

/* ---------------------------------------------------------------------- */
/* Devices and diagnostics                                                 */
/* ---------------------------------------------------------------------- */

/* Replace GetGpuCount / GetGpuName (Common/CUDA/gpuUtils.hpp:5-15). */
int tigre_get_gpu_count(void);
int tigre_get_gpu_name(int device, char* name, int len);

/* Kernel timing.  No counterpart in the reference: a measurement hook for
 * benchmarks.  While enabled, every launch of the kernels below is bracketed
 * by HIP events on the stream it is launched on; tigre_kernel_time waits for
 * the recorded launches of one kernel, returns the summed duration and the
 * launch count, and forgets them. */
#define TIGRE_KERNEL_AX_SIDDON       0
#define TIGRE_KERNEL_AX_INTERPOLATED 1
#define TIGRE_KERNEL_ATB_FDK         2
#define TIGRE_KERNEL_ATB_MATCHED     3
#define TIGRE_KERNEL_RELAYOUT        4   /* volume bricking / projection quads */
#define TIGRE_KERNEL_POCS_TV         5   /* both passes of one TV iteration */
#define TIGRE_KERNEL_TVDENOISE       6
#define TIGRE_KERNEL_PROBE           7   /* tigre_valu_probe */
#define TIGRE_KERNEL_COUNT           8
int tigre_kernel_timing(int enable);
int tigre_kernel_time(int kernel, double* total_ms, int* launches);

/* VALU issue-rate probe (no counterpart in the reference; a measurement aid):
 * `blocks` workgroups of 256 lanes, each lane running `iters` rounds of eight
 * independent VALU operations with no memory access, then one store to d_out
 * (blocks x 256 floats).  Mode 0: FP32 FMAs on all 64 lanes of each
 * wavefront; 1: the same with only lanes 0..15 active; 2: integer add and
 * compare/select chains; 3: FMA chains interleaved with a ring of integer adds.  Its instruction count is known, so its duration
 * gives the VALU issue rate against which the operator kernels' instruction
 * counts are priced. */
int tigre_valu_probe(float* d_out, int blocks, int iters, int mode, void* stream);

/* ---------------------------------------------------------------------- */
/* Projection plans                                                       */
/* ---------------------------------------------------------------------- */

/* A plan holds the per-view records of one geometry and view list, built
 * and uploaded once on the current device; its launches allocate nothing,
 * synchronise nothing and upload nothing, so a caller can capture them in a
 * HIP graph.  They replace the per-call setup of Ax/Atb inside an iterative
 * method (the subset loop of iterative_recon_alg.py:326-345, whose Ax/Atb
 * calls rebuild the geometry on every subset).  Tunables: views_per_group
 * (1, 2, 4) beam-workgroup views, max_slab_planes (1..8) LDS slab depth,
 * slices_per_thread (4, 8, 16) backprojector slices per lane.  Results do not
 * depend on views_per_group or max_slab_planes; slices_per_thread changes the
 * backprojection by rounding only. */
typedef struct TigrePlan_* TigrePlan;
int tigre_plan_create(const TigreGeometry* geo, const float* angles, int nangles, int views_per_group,
                      int max_slab_planes, int slices_per_thread, TigrePlan* out);
/* Scratch bytes every launch of this plan needs for storage type dtype
 * (the workspace must be 256-byte aligned, device memory of the plan's device). */
int tigre_plan_workspace_size(TigrePlan plan, int dtype, size_t* bytes);
/* tigre_ax_device_typed / tigre_atb_device_typed on the plan's views. */
int tigre_plan_ax(TigrePlan plan, const void* d_img, void* d_proj, int projection_type, int layout, int dtype,
                  void* d_workspace, size_t workspace_bytes, void* stream);
int tigre_plan_atb(TigrePlan plan, const void* d_proj, void* d_vol, int backprojection_type, int layout, int dtype,
                   void* d_workspace, size_t workspace_bytes, void* stream);
/* SART image update of one subset, fused into the backprojector
 * (iterative_recon_alg.py:389-399 and the clip at :344):
 * x = x + inv_v * Atb(residual), then max(x, 0) when noneg; inv_v = lambda / V,
 * an nVoxelY x nVoxelX image broadcast over z.  FP32 storage. */
int tigre_plan_sart_update(TigrePlan plan, const float* d_residual, float* d_x, const float* d_inv_v,
                           int backprojection_type, int layout, int noneg, void* d_workspace, size_t workspace_bytes,
                           void* stream);
/* SART residual (iterative_recon_alg.py:395): ax = W * (proj - ax), n elements,
 * 16-byte aligned buffers. */
int tigre_sart_residual(const float* d_proj, const float* d_w, float* d_ax, long long n, void* stream);
int tigre_plan_destroy(TigrePlan plan);

/* Batches: one SART subset update (tigre_plan_ax Siddon, tigre_sart_residual,
 * tigre_plan_sart_update FDK) of each of many independent problems, run as one
 * launch per stage (staging copies, beam projections, residuals, quad copies,
 * backprojections) over all of them, so small problems fill the device next
 * to large ones.  Replaces the reference's one-problem-at-a-time subset loop
 * (iterative_recon_alg.py:326-345) when many reconstructions run together; no
 * counterpart in the reference.  The job tables are uploaded at creation;
 * tigre_batch_sart only launches kernels (no allocation, synchronisation or
 * upload), so it can be captured in a HIP graph.  Results equal the three plan
 * calls per step bit for bit.  Steps must not share an image, a scratch
 * projection buffer or a workspace; FP32, cone beam, Python layout. */
typedef struct TigreBatchStep {
    TigrePlan plan;         /* the subset's plan */
    float* d_x;             /* the problem's image, updated in place */
    const float* d_proj;    /* measured projections of the subset (16-byte aligned) */
    const float* d_w;       /* W weights of the subset (16-byte aligned) */
    float* d_ax;            /* scratch projections of the subset's size (16-byte aligned) */
    const float* d_inv_v;   /* lambda / V of the subset, nVoxelY x nVoxelX */
    void* d_workspace;      /* tigre_plan_workspace_size(plan, TIGRE_DTYPE_F32) bytes, 256-byte aligned */
    size_t workspace_bytes;
    int noneg;              /* clip the updated image at 0 */
} TigreBatchStep;
typedef struct TigreBatch_* TigreBatch;
int tigre_batch_create(const TigreBatchStep* steps, int nsteps, TigreBatch* out);
int tigre_batch_sart(TigreBatch batch, void* stream);
int tigre_batch_destroy(TigreBatch batch);

/* ---------------------------------------------------------------------- */
/* Learned proximal step: 3x3x3 convolutions (BF16 MFMA, FP32 accumulate)  */
/* ---------------------------------------------------------------------- */

/* No counterpart in the reference.  Building blocks of the residual CNN that
 * can replace minimizeTV (GD_TV.hpp:58) inside ASD-POCS: forward, data-
 * gradient and weight-gradient passes of 'same' 3x3x3 convolutions with 32
 * channels between layers.  Volumes are n x d x h x w (x fastest);
 * activations are bf16 with 32 channels per voxel, channel fastest; a
 * single-channel FP32 volume (the image, or the loss gradient) can stand in
 * for an activation input (it becomes channel 0).  Taps index (dz, dy, dx) in
 * {-1,0,1}^3 as 9(dz+1) + 3(dy+1) + (dx+1); outside the volume inputs are 0. */
#define TIGRE_CNN_IN_BF16X32 0   /* bf16 activations, 32 channels */
#define TIGRE_CNN_IN_F32X1   1   /* one FP32 channel */
#define TIGRE_CNN_EPI_BIAS_RELU 0  /* y = bf16(max(conv + bias, 0)), 32 channels */
#define TIGRE_CNN_EPI_RELU_MASK 1  /* y = bf16(conv * (aux > 0)), aux bf16 activations (data gradient) */
#define TIGRE_CNN_EPI_RESIDUAL  2  /* y = aux + conv[channel 0] + bias[0], FP32, one channel */
/* y[v][o] = sum_tap sum_i x[v + tap][i] * w[tap][o][i], w: 27 x 32 x 32 bf16
 * (16-byte aligned), bias: 32 floats (unused by RELU_MASK). */
int tigre_cnn_conv(const void* d_x, int x_mode, const void* d_w, const float* d_bias, const void* d_aux, void* d_y,
                   int epilogue, int n, int d, int h, int w, void* stream);
/* dw[tap][o][i] = sum_v x[v + tap][i] * dy[v][o] (FP32, 28 x 32 x 32; row 27
 * holds the bias gradient sum_v dy[v][o] at [27][o][0]).  Deterministic: a
 * fixed-order reduction of per-block partial sums kept in the workspace. */
int tigre_cnn_wgrad_workspace_size(int n, int d, int h, int w, size_t* bytes);
int tigre_cnn_wgrad(const void* d_x, int x_mode, const void* d_dy, int dy_mode, float* d_dw, int n, int d, int h,
                    int w, void* d_workspace, size_t workspace_bytes, void* stream);
/* tigre_cnn_wgrad without a workspace argument and without any allocation
 * (no counterpart in the reference: for callers that cannot allocate, such as
 * code under HIP graph capture or a training loop that forbids hipMalloc).
 * The per-block partial sums go to a static device array the library holds on
 * each GPU it runs on until the process ends; its size, the largest
 * tigre_cnn_wgrad_workspace_size of any volume (28 MiB), is returned by
 * tigre_cnn_wgrad_noalloc_limit, and a volume needing more is rejected.  The
 * result equals tigre_cnn_wgrad's bit for bit.  Calls that overlap on one GPU
 * (different streams) share the array and race. */
int tigre_cnn_wgrad_noalloc(const void* d_x, int x_mode, const void* d_dy, int dy_mode, float* d_dw, int n, int d,
                            int h, int w, void* stream);
int tigre_cnn_wgrad_noalloc_limit(size_t* bytes);

/* Message of the last failed call on this thread ("" if none). */
const char* tigre_last_error(void);

/* Library version string. */
const char* tigre_version(void);

#ifdef __cplusplus
}
#endif
#endif /* TIGRE_AMD_H */
