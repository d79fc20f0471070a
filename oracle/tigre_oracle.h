/*
 * TEST INFRASTRUCTURE ONLY -- the FP64 CPU oracle.
 *
 * A double-precision, OpenMP restatement of the cone-beam operators of
 * CERN/TIGRE v3.1.3 (Common/CUDA).  It exists so that tests, the smoke check
 * and the CPU-baseline leg of the benchmark can check the HIP library against
 * an independent implementation of the same ray/voxel-weight algorithms.  The
 * product library never links, loads or calls anything in this directory.
 *
 * Geometry is passed as doubles whose values the caller has already rounded
 * to float32 (the Python front end's cast_to_single, geometry.py:121-140), so
 * the oracle sees the same numbers as the device; everything after that is
 * FP64.
 */
#ifndef TIGRE_ORACLE_H
#define TIGRE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct og_geo {
    int nVoxelX, nVoxelY, nVoxelZ;
    double sVoxelX, sVoxelY, sVoxelZ;
    double dVoxelX, dVoxelY, dVoxelZ;
    int nDetecU, nDetecV;
    double sDetecU, sDetecV;
    double dDetecU, dDetecV;
    double accuracy;
    int nangles;
    const double* angles;   /* nangles x 3 : (alpha, theta, psi), ZYZ Euler  */
    const double* DSO;      /* nangles */
    const double* DSD;
    const double* COR;
    const double* offOrigX; /* volume offsets, one per view */
    const double* offOrigY;
    const double* offOrigZ;
    const double* offDetecU;
    const double* offDetecV;
    const double* dRoll;
    const double* dPitch;
    const double* dYaw;
} og_geo;

/* Parallel-beam forms (the reference's *_parallel functions). */
void og_forward_frame_parallel(const og_geo* g, int view, int voxel_centre_origin, double out[12]);
void og_ax_siddon_parallel(const float* img, const og_geo* g, int layout_matlab, double* out);
void og_ax_interp_parallel(const float* img, const og_geo* g, int layout_matlab, double* out);
void og_atb_parallel(const float* proj, const og_geo* g, int layout_matlab, double* out);

/* Per-view frames (exposed so tests can pin them against closed forms). */
void og_forward_frame(const og_geo* g, int view, int voxel_centre_origin,
                      double out[12]);  /* uvOrigin, deltaU, deltaV, source */
double og_maxdistance_cuboid(const og_geo* g, int view);
void og_backward_frame(const og_geo* g, int view,
                       double out[15]); /* origin, dX, dY, dZ, source */

/* Forward projection, out: nangles x nV x nU (layout_matlab=0) or
 * nangles x nU x nV (layout_matlab=1).  img: X fastest, nZ x nY x nX. */
void og_ax_siddon(const float* img, const og_geo* g, int layout_matlab, double* out);
void og_ax_interp(const float* img, const og_geo* g, int layout_matlab, double* out);

/* Backprojection, proj in the same layouts, out: nZ x nY x nX. */
void og_atb_fdk(const float* proj, const og_geo* g, int layout_matlab, double* out);
void og_atb_matched(const float* proj, const og_geo* g, int layout_matlab, double* out);

/* TV steepest descent (GD_TV.cu pocs_tv), FP64 throughout.
 * img/dst: nz x ny x nx, x fastest. */
void og_pocs_tv(const double* img, double* dst, double alpha,
                long nx, long ny, long nz, int maxiter);
/* One TV sub-gradient evaluation (exposed for the closed-form tests). */
void og_tv_gradient(const double* img, double* grad, long nx, long ny, long nz);

/* Generalised TV sub-gradient (GD_TV / GD_AwTV / PICCS kernels): pad_z = 1
 * for the zero-padded z of pocs_tv / aw_pocs_tv, inv_delta != 0 for the AwTV
 * weights, eps 1e-8 (TV, AwTV) or 1e-6 (PICCS). */
void og_tv_gradient_ex(const double* img, double* grad, long nx, long ny, long nz, int pad_z,
                       double inv_delta, double eps);
/* aw_pocs_tv (GD_AwTV.cu) and piccs_tv (PICCS.cu), FP64. */
void og_aw_pocs_tv(const double* img, double* dst, double alpha, long nx, long ny, long nz,
                   int maxiter, double delta);
void og_piccs_tv(const double* img, const double* prior, double* dst, double alpha, double ratio,
                 long nx, long ny, long nz, int maxiter);

/* ROF primal-dual denoise (tv_proximal.cu tvdenoising), FP64. */
void og_tvdenoise(const double* src, double* dst, double lambda,
                  const double spacing[3], long nx, long ny, long nz, int maxiter);

void og_set_threads(int n);
int og_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
