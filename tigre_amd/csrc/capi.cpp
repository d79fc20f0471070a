// extern "C" entry points (declared in include/tigre_amd.h).
//
// Host-pointer entry points reproduce the reference's call contract: host
// buffers in, host buffers out, blocking, output overwritten.  They stage the
// data through device memory on a private stream and synchronise before they
// return.  Device-pointer entry points enqueue on the caller's stream.
#include <cstring>

#include "common.hpp"

using namespace tigre;

namespace {

// RAII stream + device buffers for the host-pointer entry points.
struct HostCall {
    hipStream_t stream = nullptr;
    std::vector<void*> bufs;
    ~HostCall() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            for (void* b : bufs) (void)hipFreeAsync(b, stream);
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
    int begin(const int* gpuids, int ngpus) {
        TIGRE_CHECK(select_device(gpuids, ngpus));
        TIGRE_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        return TIGRE_OK;
    }
    int alloc(float** p, size_t n) {
        TIGRE_HIP_TRY(hipMallocAsync((void**)p, n * sizeof(float), stream));
        bufs.push_back(*p);
        return TIGRE_OK;
    }
    int finish() {
        TIGRE_HIP_TRY(hipStreamSynchronize(stream));
        return TIGRE_OK;
    }
};

size_t volume_elems(const TigreGeometry* g) {
    return (size_t)g->nVoxelX * (size_t)g->nVoxelY * (size_t)g->nVoxelZ;
}

int ax_host(const float* img, const TigreGeometry* geo, float** result, const float* angles, int nangles,
            int projection_type, int layout, const int* gpuids, int ngpus, bool parallel) {
    set_error("");
    TIGRE_REQUIRE(img != nullptr && result != nullptr, "null image or result pointer");
    HostGeometry g;
    TIGRE_CHECK(import_geometry(geo, angles, nangles, &g));
    g.parallel = parallel;
    HostCall call;
    TIGRE_CHECK(call.begin(gpuids, ngpus));
    const size_t nvox = volume_elems(geo);
    const size_t P = (size_t)geo->nDetecU * geo->nDetecV;
    float *d_img = nullptr, *d_proj = nullptr;
    TIGRE_CHECK(call.alloc(&d_img, nvox));
    TIGRE_CHECK(call.alloc(&d_proj, P * (size_t)nangles));
    TIGRE_HIP_TRY(hipMemcpyAsync(d_img, img, nvox * sizeof(float), hipMemcpyHostToDevice, call.stream));
    TIGRE_CHECK(ax_device(d_img, g, d_proj, projection_type, layout, TIGRE_DTYPE_F32, call.stream));
    for (int i = 0; i < nangles; ++i)
        TIGRE_HIP_TRY(hipMemcpyAsync(result[i], d_proj + P * (size_t)i, P * sizeof(float), hipMemcpyDeviceToHost,
                                     call.stream));
    return call.finish();
}

int atb_host(const float* projections, const TigreGeometry* geo, float* result, const float* angles, int nangles,
             int backprojection_type, int layout, const int* gpuids, int ngpus, bool parallel) {
    set_error("");
    TIGRE_REQUIRE(projections != nullptr && result != nullptr, "null projection or result pointer");
    HostGeometry g;
    TIGRE_CHECK(import_geometry(geo, angles, nangles, &g));
    g.parallel = parallel;
    HostCall call;
    TIGRE_CHECK(call.begin(gpuids, ngpus));
    const size_t nvox = volume_elems(geo);
    const size_t np = (size_t)geo->nDetecU * geo->nDetecV * (size_t)nangles;
    float *d_proj = nullptr, *d_vol = nullptr;
    TIGRE_CHECK(call.alloc(&d_proj, np));
    TIGRE_CHECK(call.alloc(&d_vol, nvox));
    TIGRE_HIP_TRY(hipMemcpyAsync(d_proj, projections, np * sizeof(float), hipMemcpyHostToDevice, call.stream));
    TIGRE_CHECK(atb_device(d_proj, g, d_vol, backprojection_type, layout, TIGRE_DTYPE_F32, call.stream));
    TIGRE_HIP_TRY(hipMemcpyAsync(result, d_vol, nvox * sizeof(float), hipMemcpyDeviceToHost, call.stream));
    return call.finish();
}

}  // namespace

extern "C" {

int tigre_ax(const float* img, const TigreGeometry* geo, float** result, const float* angles, int nangles,
             int projection_type, int layout, const int* gpuids, int ngpus) {
    return ax_host(img, geo, result, angles, nangles, projection_type, layout, gpuids, ngpus, false);
}

int tigre_ax_parallel(const float* img, const TigreGeometry* geo, float** result, const float* angles, int nangles,
                      int projection_type, int layout, const int* gpuids, int ngpus) {
    return ax_host(img, geo, result, angles, nangles, projection_type, layout, gpuids, ngpus, true);
}

int tigre_atb(const float* projections, const TigreGeometry* geo, float* result, const float* angles, int nangles,
              int backprojection_type, int layout, const int* gpuids, int ngpus) {
    return atb_host(projections, geo, result, angles, nangles, backprojection_type, layout, gpuids, ngpus, false);
}

int tigre_atb_parallel(const float* projections, const TigreGeometry* geo, float* result, const float* angles,
                       int nangles, int layout, const int* gpuids, int ngpus) {
    return atb_host(projections, geo, result, angles, nangles, TIGRE_BACKPROJ_FDK, layout, gpuids, ngpus, true);
}

int tigre_pocs_tv(const float* img, float* dst, float alpha, const long* image_size, int maxIter, const int* gpuids,
                  int ngpus) {
    set_error("");
    TIGRE_REQUIRE(img && dst && image_size, "null argument");
    TIGRE_REQUIRE(image_size[0] > 0 && image_size[1] > 0 && image_size[2] > 0, "image size must be positive");
    TIGRE_REQUIRE(maxIter >= 0, "maxIter must be non-negative");
    HostCall call;
    TIGRE_CHECK(call.begin(gpuids, ngpus));
    const size_t n = (size_t)image_size[0] * image_size[1] * image_size[2];
    float *d_a = nullptr, *d_b = nullptr;
    TIGRE_CHECK(call.alloc(&d_a, n));
    TIGRE_CHECK(call.alloc(&d_b, n));
    TIGRE_HIP_TRY(hipMemcpyAsync(d_a, img, n * sizeof(float), hipMemcpyHostToDevice, call.stream));
    TIGRE_CHECK(run_pocs_tv(d_a, d_b, alpha, image_size[0], image_size[1], image_size[2], maxIter, call.stream));
    TIGRE_HIP_TRY(hipMemcpyAsync(dst, d_b, n * sizeof(float), hipMemcpyDeviceToHost, call.stream));
    return call.finish();
}

int tigre_aw_pocs_tv(const float* img, float* dst, float alpha, const long* image_size, int maxIter, float delta,
                     const int* gpuids, int ngpus) {
    set_error("");
    TIGRE_REQUIRE(img && dst && image_size, "null argument");
    TIGRE_REQUIRE(image_size[0] > 0 && image_size[1] > 0 && image_size[2] > 0, "image size must be positive");
    TIGRE_REQUIRE(maxIter >= 0, "maxIter must be non-negative");
    HostCall call;
    TIGRE_CHECK(call.begin(gpuids, ngpus));
    const size_t n = (size_t)image_size[0] * image_size[1] * image_size[2];
    float *d_a = nullptr, *d_b = nullptr;
    TIGRE_CHECK(call.alloc(&d_a, n));
    TIGRE_CHECK(call.alloc(&d_b, n));
    TIGRE_HIP_TRY(hipMemcpyAsync(d_a, img, n * sizeof(float), hipMemcpyHostToDevice, call.stream));
    TIGRE_CHECK(run_aw_pocs_tv(d_a, d_b, alpha, delta, image_size[0], image_size[1], image_size[2], maxIter,
                               call.stream));
    TIGRE_HIP_TRY(hipMemcpyAsync(dst, d_b, n * sizeof(float), hipMemcpyDeviceToHost, call.stream));
    return call.finish();
}

int tigre_piccs_tv(const float* img, const float* prior, float* dst, float alpha, float ratio, const long* image_size,
                   int maxIter, const int* gpuids, int ngpus) {
    set_error("");
    TIGRE_REQUIRE(img && prior && dst && image_size, "null argument");
    TIGRE_REQUIRE(image_size[0] > 0 && image_size[1] > 0 && image_size[2] > 0, "image size must be positive");
    TIGRE_REQUIRE(maxIter >= 0, "maxIter must be non-negative");
    HostCall call;
    TIGRE_CHECK(call.begin(gpuids, ngpus));
    const size_t n = (size_t)image_size[0] * image_size[1] * image_size[2];
    float *d_a = nullptr, *d_p = nullptr, *d_b = nullptr;
    TIGRE_CHECK(call.alloc(&d_a, n));
    TIGRE_CHECK(call.alloc(&d_p, n));
    TIGRE_CHECK(call.alloc(&d_b, n));
    TIGRE_HIP_TRY(hipMemcpyAsync(d_a, img, n * sizeof(float), hipMemcpyHostToDevice, call.stream));
    TIGRE_HIP_TRY(hipMemcpyAsync(d_p, prior, n * sizeof(float), hipMemcpyHostToDevice, call.stream));
    TIGRE_CHECK(run_piccs_tv(d_a, d_p, d_b, alpha, ratio, image_size[0], image_size[1], image_size[2], maxIter,
                             call.stream));
    TIGRE_HIP_TRY(hipMemcpyAsync(dst, d_b, n * sizeof(float), hipMemcpyDeviceToHost, call.stream));
    return call.finish();
}

int tigre_tvdenoising(const float* src, float* dst, float lambda, const float* spacing, const long* image_size,
                      int maxIter, const int* gpuids, int ngpus) {
    set_error("");
    TIGRE_REQUIRE(src && dst && image_size && spacing, "null argument");
    TIGRE_REQUIRE(image_size[0] > 0 && image_size[1] > 0 && image_size[2] > 0, "image size must be positive");
    TIGRE_REQUIRE(maxIter >= 0, "maxIter must be non-negative");
    HostCall call;
    TIGRE_CHECK(call.begin(gpuids, ngpus));
    const size_t n = (size_t)image_size[0] * image_size[1] * image_size[2];
    float *d_a = nullptr, *d_b = nullptr;
    TIGRE_CHECK(call.alloc(&d_a, n));
    TIGRE_CHECK(call.alloc(&d_b, n));
    TIGRE_HIP_TRY(hipMemcpyAsync(d_a, src, n * sizeof(float), hipMemcpyHostToDevice, call.stream));
    TIGRE_CHECK(run_tvdenoise(d_a, d_b, lambda, spacing, image_size[0], image_size[1], image_size[2], maxIter,
                              call.stream));
    TIGRE_HIP_TRY(hipMemcpyAsync(dst, d_b, n * sizeof(float), hipMemcpyDeviceToHost, call.stream));
    return call.finish();
}

int tigre_ax_device(const float* d_img, const TigreGeometry* geo, float* d_proj, const float* angles, int nangles,
                    int projection_type, int layout, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_img != nullptr && d_proj != nullptr, "null device pointer");
    HostGeometry g;
    TIGRE_CHECK(import_geometry(geo, angles, nangles, &g));
    return ax_device(d_img, g, d_proj, projection_type, layout, TIGRE_DTYPE_F32, (hipStream_t)stream);
}

int tigre_ax_device_typed(const void* d_img, const TigreGeometry* geo, void* d_proj, const float* angles,
                          int nangles, int projection_type, int layout, int dtype, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_img != nullptr && d_proj != nullptr, "null device pointer");
    HostGeometry g;
    TIGRE_CHECK(import_geometry(geo, angles, nangles, &g));
    return ax_device(d_img, g, d_proj, projection_type, layout, dtype, (hipStream_t)stream);
}

int tigre_atb_device(const float* d_proj, const TigreGeometry* geo, float* d_vol, const float* angles, int nangles,
                     int backprojection_type, int layout, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_proj != nullptr && d_vol != nullptr, "null device pointer");
    HostGeometry g;
    TIGRE_CHECK(import_geometry(geo, angles, nangles, &g));
    return atb_device(d_proj, g, d_vol, backprojection_type, layout, TIGRE_DTYPE_F32, (hipStream_t)stream);
}

int tigre_atb_device_typed(const void* d_proj, const TigreGeometry* geo, void* d_vol, const float* angles,
                           int nangles, int backprojection_type, int layout, int dtype, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_proj != nullptr && d_vol != nullptr, "null device pointer");
    HostGeometry g;
    TIGRE_CHECK(import_geometry(geo, angles, nangles, &g));
    return atb_device(d_proj, g, d_vol, backprojection_type, layout, dtype, (hipStream_t)stream);
}

int tigre_ax_parallel_device(const void* d_img, const TigreGeometry* geo, void* d_proj, const float* angles,
                             int nangles, int projection_type, int layout, int dtype, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_img != nullptr && d_proj != nullptr, "null device pointer");
    HostGeometry g;
    TIGRE_CHECK(import_geometry(geo, angles, nangles, &g));
    g.parallel = true;
    return ax_device(d_img, g, d_proj, projection_type, layout, dtype, (hipStream_t)stream);
}

int tigre_atb_parallel_device(const void* d_proj, const TigreGeometry* geo, void* d_vol, const float* angles,
                              int nangles, int layout, int dtype, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_proj != nullptr && d_vol != nullptr, "null device pointer");
    HostGeometry g;
    TIGRE_CHECK(import_geometry(geo, angles, nangles, &g));
    g.parallel = true;
    return atb_device(d_proj, g, d_vol, TIGRE_BACKPROJ_FDK, layout, dtype, (hipStream_t)stream);
}

int tigre_pocs_tv_device(const float* d_img, float* d_dst, float alpha, const long* image_size, int maxIter,
                         void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_img && d_dst && image_size, "null argument");
    TIGRE_REQUIRE(image_size[0] > 0 && image_size[1] > 0 && image_size[2] > 0, "image size must be positive");
    TIGRE_REQUIRE(maxIter >= 0, "maxIter must be non-negative");
    return run_pocs_tv(d_img, d_dst, alpha, image_size[0], image_size[1], image_size[2], maxIter,
                       (hipStream_t)stream);
}

int tigre_aw_pocs_tv_device(const float* d_img, float* d_dst, float alpha, const long* image_size, int maxIter,
                            float delta, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_img && d_dst && image_size, "null argument");
    TIGRE_REQUIRE(image_size[0] > 0 && image_size[1] > 0 && image_size[2] > 0, "image size must be positive");
    TIGRE_REQUIRE(maxIter >= 0, "maxIter must be non-negative");
    return run_aw_pocs_tv(d_img, d_dst, alpha, delta, image_size[0], image_size[1], image_size[2], maxIter,
                          (hipStream_t)stream);
}

int tigre_piccs_tv_device(const float* d_img, const float* d_prior, float* d_dst, float alpha, float ratio,
                          const long* image_size, int maxIter, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_img && d_prior && d_dst && image_size, "null argument");
    TIGRE_REQUIRE(image_size[0] > 0 && image_size[1] > 0 && image_size[2] > 0, "image size must be positive");
    TIGRE_REQUIRE(maxIter >= 0, "maxIter must be non-negative");
    return run_piccs_tv(d_img, d_prior, d_dst, alpha, ratio, image_size[0], image_size[1], image_size[2], maxIter,
                        (hipStream_t)stream);
}

int tigre_tvdenoising_device(const float* d_src, float* d_dst, float lambda, const float* spacing,
                             const long* image_size, int maxIter, void* stream) {
    set_error("");
    TIGRE_REQUIRE(d_src && d_dst && image_size && spacing, "null argument");
    TIGRE_REQUIRE(image_size[0] > 0 && image_size[1] > 0 && image_size[2] > 0, "image size must be positive");
    TIGRE_REQUIRE(maxIter >= 0, "maxIter must be non-negative");
    return run_tvdenoise(d_src, d_dst, lambda, spacing, image_size[0], image_size[1], image_size[2], maxIter,
                         (hipStream_t)stream);
}

int tigre_get_gpu_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int tigre_get_gpu_name(int device, char* name, int len) {
    TIGRE_REQUIRE(name != nullptr && len > 0, "null name buffer");
    hipDeviceProp_t prop;
    TIGRE_HIP_TRY(hipGetDeviceProperties(&prop, device));
    std::strncpy(name, prop.name, (size_t)len - 1);
    name[len - 1] = '\0';
    return TIGRE_OK;
}

int tigre_kernel_timing(int enable) { return kernel_timing_enable(enable); }

int tigre_kernel_time(int kernel, double* total_ms, int* launches) {
    return kernel_time_collect(kernel, total_ms, launches);
}

const char* tigre_last_error(void) { return last_error(); }

const char* tigre_version(void) { return "tigre_amd 0.1.0 (gfx950)"; }

}  // extern "C"
