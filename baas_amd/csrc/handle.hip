// handle.hip -- library handles, status words and the error reporting path.
#include <cstring>

#include "blas.h"
#include "internal.h"

namespace baas {

size_t lds_limit(baas_handle *h) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) return 64 * 1024;
    size_t lim = prop.sharedMemPerBlock;
    return lim > 1024 ? lim - 1024 : lim;  // room for the static __shared__ words
}

}  // namespace baas

// Ready flags start at 0, so the first epoch is 1 (handle->counters[2]).
static bool set_first_epoch(baas_handle *h) {
    int one = 1;
    return hipMemcpy(h->counters + 2, &one, sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
}

extern "C" int baas_create(baas_handle **out, void *stream) {
    if (!out) return BAAS_INVALID_ARG;
    *out = nullptr;
    baas_handle *h = new baas_handle();
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return BAAS_HIP_ERROR;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) {
        delete h;
        return BAAS_HIP_ERROR;
    }
    h->num_cus = prop.multiProcessorCount;
    h->stream = (hipStream_t)stream;
    // this device's address of the allocation-free entry point's scratch
    int rc = baas::noalloc_scratch_init(h->device);
    if (rc != BAAS_OK) {
        delete h;
        return rc;
    }
    if (hipMalloc(&h->status, sizeof(baas::DevStatus)) != hipSuccess ||
        hipMalloc(&h->counters, BAAS_COUNTER_INTS * sizeof(int)) != hipSuccess) {
        if (h->status) (void)hipFree(h->status);
        delete h;
        return BAAS_ALLOC;
    }
    if (hipMemsetAsync(h->status, 0, sizeof(baas::DevStatus), h->stream) != hipSuccess ||
        hipMemsetAsync(h->counters, 0, BAAS_COUNTER_INTS * sizeof(int), h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess || !set_first_epoch(h)) {
        (void)hipFree(h->status);
        (void)hipFree(h->counters);
        delete h;
        return BAAS_HIP_ERROR;
    }
    *out = h;
    return BAAS_OK;
}

extern "C" int baas_destroy(baas_handle *h) {
    if (!h) return BAAS_INVALID_ARG;
    (void)hipStreamSynchronize(h->stream);
    if (h->status) (void)hipFree(h->status);
    if (h->counters) (void)hipFree(h->counters);
    if (h->dscratch) (void)hipFree(h->dscratch);
    delete h;
    return BAAS_OK;
}

extern "C" int baas_set_stream(baas_handle *h, void *stream) {
    if (!h) return BAAS_INVALID_ARG;
    h->stream = (hipStream_t)stream;
    return BAAS_OK;
}

extern "C" int baas_set_spin_limit(baas_handle *h, long long polls) {
    if (!h || polls < 1) return BAAS_INVALID_ARG;
    h->spin_limit = polls;
    return BAAS_OK;
}

extern "C" int baas_check(baas_handle *h) {
    if (!h) return BAAS_INVALID_ARG;
    BAAS_HIP_TRY(hipStreamSynchronize(h->stream));
    baas::DevStatus st{0, 0};
    BAAS_HIP_TRY(hipMemcpy(&st, h->status, sizeof(st), hipMemcpyDeviceToHost));
    if (st.code != 0) {
        baas::DevStatus z{0, 0};
        BAAS_HIP_TRY(hipMemcpy(h->status, &z, sizeof(z), hipMemcpyHostToDevice));
    }
    return st.code;
}

extern "C" const char *baas_status_string(int s) {
    switch (s) {
        case BAAS_OK: return "ok";
        case BAAS_INVALID_ARG: return "invalid argument";
        case BAAS_ZERO_PIVOT: return "zero pivot";
        case BAAS_NONPOSITIVE_PIVOT: return "non-positive pivot";
        case BAAS_HIP_ERROR: return "HIP error";
        case BAAS_TIMEOUT: return "ready-flag wait timed out";
        case BAAS_NOT_CONVERGED: return "not converged";
        case BAAS_ALLOC: return "allocation failed";
        case BAAS_UNSUPPORTED: return "unsupported";
        case BAAS_TOO_LARGE: return "input exceeds the fixed scratch of the allocation-free entry point";
        default: return "unknown status";
    }
}
