// valu_rates.hip -- issue rate and accuracy of the VALU operations that the
// pair kernels are built from, on the device this runs on.
//
//   pn_amd/valu_rates            (prints one line per measurement)
//
// Rates: 4096 workgroups x 256 lanes, 2000 iterations of 8 independent
// chains per lane; lane-ops/s = lanes x iterations x 8 / kernel time.
// Accuracy: v_rsq_f64 raw and with the library's third-order correction,
// and v_rsq_f32, against 1/sqrt in long double on 2^20 log-uniform inputs
// over [1e-8, 1e4].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#define CHECK(x)                                                                     \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));             \
            return 1;                                                                \
        }                                                                            \
    } while (0)

__global__ void acc64(const double* x, double* raw, double* cor, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double y = __builtin_amdgcn_rsq(x[i]);
    raw[i] = y;
    const double h = x[i] * y;
    const double r = __builtin_fma(-h, y, 1.0);
    const double t = __builtin_fma(0.375, r, 0.5);
    cor[i] = __builtin_fma(y * r, t, y);
}

__global__ void acc32(const float* x, float* y, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = __builtin_amdgcn_rsqf(x[i]);
}

template <typename T, int OP>
__global__ void rate(T* out, int iters)
{
    T a[8];
    for (int k = 0; k < 8; ++k) a[k] = T(1) + T(1e-3) * T(threadIdx.x + k);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if constexpr (OP == 0) {
                if constexpr (sizeof(T) == 8) a[k] = __builtin_fma(a[k], T(0.999999), T(1e-7));
                else a[k] = __builtin_fmaf(a[k], T(0.999999), T(1e-7));
            }
            else if constexpr (OP == 1) a[k] = a[k] * T(0.999999);
            else {
                if constexpr (sizeof(T) == 8) a[k] = __builtin_amdgcn_rsq(a[k]);
                else a[k] = __builtin_amdgcn_rsqf(a[k]);
            }
        }
    }
    T s = 0;
    for (int k = 0; k < 8; ++k) s += a[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <typename T, int OP>
int time_rate(const char* name, int cus)
{
    const int blocks = 4096, th = 256, iters = 2000;
    T* out;
    CHECK(hipMalloc(&out, sizeof(T) * blocks * th));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    float ms = 0;
    for (int rep = 0; rep < 2; ++rep) {
        CHECK(hipEventRecord(e0));
        hipLaunchKernelGGL((rate<T, OP>), dim3(blocks), dim3(th), 0, 0, out, iters);
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        CHECK(hipEventElapsedTime(&ms, e0, e1));
    }
    const double ops = (double)blocks * th * iters * 8;
    std::printf("%-10s %8.3f ms  %.3e lane-ops/s  %6.1f lane-ops per CU-cycle at 2.4 GHz\n", name,
                ms, ops / (ms * 1e-3), ops / (ms * 1e-3) / (cus * 2.4e9));
    CHECK(hipFree(out));
    return 0;
}

int main()
{
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    std::printf("device %s, %d CUs\n", prop.gcnArchName, prop.multiProcessorCount);
    const int n = 1 << 20;
    std::vector<double> hx(n);
    std::mt19937_64 g(1);
    std::uniform_real_distribution<double> u(-8.0, 4.0);
    for (int i = 0; i < n; ++i) hx[i] = std::pow(10.0, u(g));
    std::vector<float> hxf(n);
    for (int i = 0; i < n; ++i) hxf[i] = (float)hx[i];
    double *dx, *draw, *dcor;
    float *dxf, *dyf;
    CHECK(hipMalloc(&dx, n * 8));
    CHECK(hipMalloc(&draw, n * 8));
    CHECK(hipMalloc(&dcor, n * 8));
    CHECK(hipMalloc(&dxf, n * 4));
    CHECK(hipMalloc(&dyf, n * 4));
    CHECK(hipMemcpy(dx, hx.data(), n * 8, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dxf, hxf.data(), n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(acc64, dim3(n / 256), dim3(256), 0, 0, dx, draw, dcor, n);
    hipLaunchKernelGGL(acc32, dim3(n / 256), dim3(256), 0, 0, dxf, dyf, n);
    std::vector<double> raw(n), cor(n);
    std::vector<float> yf(n);
    CHECK(hipMemcpy(raw.data(), draw, n * 8, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(cor.data(), dcor, n * 8, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(yf.data(), dyf, n * 4, hipMemcpyDeviceToHost));
    double mraw = 0, mcor = 0, mf = 0;
    for (int i = 0; i < n; ++i) {
        const long double ex = 1.0L / sqrtl((long double)hx[i]);
        const long double exf = 1.0L / sqrtl((long double)hxf[i]);
        mraw = std::fmax(mraw, (double)fabsl((raw[i] - ex) / ex) / 0x1p-53);
        mcor = std::fmax(mcor, (double)fabsl((cor[i] - ex) / ex) / 0x1p-53);
        mf = std::fmax(mf, (double)fabsl((yf[i] - exf) / exf) / 0x1p-24);
    }
    std::printf("v_rsq_f64 max rel err %.3e (%.1f half-ulp); with 3rd-order correction %.2f half-ulp\n",
                mraw * 0x1p-53, mraw, mcor);
    std::printf("v_rsq_f32 max rel err %.3e (%.2f half-ulp)\n", mf * 0x1p-24, mf);
    const int cus = prop.multiProcessorCount;
    if (time_rate<double, 0>("fma_f64", cus) || time_rate<double, 1>("mul_f64", cus) ||
        time_rate<double, 2>("rsq_f64", cus) || time_rate<float, 0>("fma_f32", cus) ||
        time_rate<float, 1>("mul_f32", cus) || time_rate<float, 2>("rsq_f32", cus))
        return 1;
    return 0;
}
