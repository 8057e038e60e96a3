// Device math on host arrays, for tests: the FP32 erfc fit and the FP32 / FP64 pair kernels exactly as the
// pair passes evaluate them, so they can be checked against the mathematics and against each other.
#include <string>

#include "../../include/pgm_hip.h"
#include "kernels.hpp"

namespace pgmhip {

void setGlobalError(const char* msg);

namespace {

__global__ void kEvalErfc(int n, const float* x, float* y) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        y[i] = erfcFast(x[i]);
}

// out[i][0..3] = G0..G3 of the pair kernels; out[i][4..5] = G0, G3 as the FP32 final and tangent passes rebuild
// them from G1, G2 (FP64: G0, G3 again).
__global__ void kEvalPair(int n, int precision, const double* rgSum, const double* r2, double beta, double* out) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= n)
        return;
    double* o = out+6*(size_t) i;
    if (precision == PGM_MIXED) {
        float G[4];
        const float rs = (float) rgSum[i], q2 = (float) r2[i], b = (float) beta;
        pairKernelsF<4>(rs, b, q2, G);
        float G0, G3;
        g03FromG12(G[1], G[2], q2, rs, b, G0, G3);
        for (int k = 0; k < 4; k++)
            o[k] = G[k];
        o[4] = G0;
        o[5] = G3;
    }
    else {
        double G[4];
        pairKernels<double, 4>(rsqrt(rgSum[i]), beta, sqrt(r2[i]), G);
        for (int k = 0; k < 4; k++)
            o[k] = G[k];
        o[4] = G[0];
        o[5] = G[3];
    }
}

// Copies in, one launch, copies out, on the current device; frees what it allocated on every path.
template<class Launch>
int onDevice(size_t inBytes0, const void* in0, size_t inBytes1, const void* in1, size_t outBytes, void* out,
             Launch launch) {
    void* d0 = nullptr;
    void* d1 = nullptr;
    void* dOut = nullptr;
    hipError_t e = hipMalloc(&d0, inBytes0 ? inBytes0 : 1);
    if (e == hipSuccess && inBytes1)
        e = hipMalloc(&d1, inBytes1);
    if (e == hipSuccess)
        e = hipMalloc(&dOut, outBytes ? outBytes : 1);
    if (e == hipSuccess && inBytes0)
        e = hipMemcpy(d0, in0, inBytes0, hipMemcpyHostToDevice);
    if (e == hipSuccess && inBytes1)
        e = hipMemcpy(d1, in1, inBytes1, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch(d0, d1, dOut);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    if (e == hipSuccess && outBytes)
        e = hipMemcpy(out, dOut, outBytes, hipMemcpyDeviceToHost);
    (void) hipFree(d0);
    (void) hipFree(d1);
    (void) hipFree(dOut);
    if (e != hipSuccess) {
        setGlobalError((std::string("pgm_hip: ")+hipGetErrorString(e)).c_str());
        return -1;
    }
    return 0;
}

}   // namespace
}   // namespace pgmhip

extern "C" {

int pgm_eval_erfc_f32(int n, const float* x, float* y) {
    if (n < 0 || (n > 0 && (!x || !y))) {
        pgmhip::setGlobalError("pgm_hip: pgm_eval_erfc_f32 needs n >= 0 and both arrays");
        return -1;
    }
    if (n == 0)
        return 0;
    const size_t bytes = sizeof(float)*(size_t) n;
    return pgmhip::onDevice(bytes, x, 0, nullptr, bytes, y, [&](void* dx, void*, void* dy) {
        pgmhip::kEvalErfc<<<(n+255)/256, 256>>>(n, static_cast<const float*>(dx), static_cast<float*>(dy));
    });
}

int pgm_eval_pair_kernels(int precision, int n, const double* rg_sum, const double* r2, double beta, double* out) {
    if (n < 0 || (n > 0 && (!rg_sum || !r2 || !out)) || (precision != PGM_DOUBLE && precision != PGM_MIXED)) {
        pgmhip::setGlobalError("pgm_hip: pgm_eval_pair_kernels needs a precision, n >= 0 and all arrays");
        return -1;
    }
    if (n == 0)
        return 0;
    const size_t bytes = sizeof(double)*(size_t) n;
    return pgmhip::onDevice(bytes, rg_sum, bytes, r2, 6*bytes, out, [&](void* dr, void* dq, void* dOut) {
        pgmhip::kEvalPair<<<(n+255)/256, 256>>>(n, precision, static_cast<const double*>(dr),
                                                 static_cast<const double*>(dq), beta, static_cast<double*>(dOut));
    });
}

}   // extern "C"
