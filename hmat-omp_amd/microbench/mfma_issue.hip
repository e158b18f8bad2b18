// Issue-rate probe behind the sketch kernels' roofline ceilings.
//
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -mllvm -amdgpu-mfma-vgpr-form mfma_issue.hip -o mfma_issue && ./mfma_issue
//
// 2048 workgroups of 4 waves (8 waves per SIMD), 16 independent accumulators
// per wave, no memory traffic in the timed loops.  Reports the sustained
// FP64 (v_mfma_f64_16x16x4f64) and BF16 (v_mfma_f32_16x16x32_bf16) MFMA
// rates, and the SIMD cycles that 16 FP64 FMAs or 16 v_rcp_f64 add to 16
// FP64 MFMAs when interleaved with them (the clock is taken from the
// MFMA-only run at 64 cycles per 16x16x4 FP64 MFMA).
#include <hip/hip_runtime.h>

#include <cstdio>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

#define CHECK(x)                                                            \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
      return 1;                                                             \
    }                                                                       \
  } while (0)

// MODE 0: MFMA only; 1: + 16 FMA; 2: + 16 v_rcp_f64; 3: 16 FMA only; 4: 16 rcp only
template <int MODE>
__global__ __launch_bounds__(256) void k_f64(double* out, int iters) {
  v4d acc[16];
  double x[16];
  for (int i = 0; i < 16; ++i) {
    acc[i] = v4d{0, 0, 0, 0};
    x[i] = 1.0 + threadIdx.x * 1e-3 + i;
  }
  const double a = threadIdx.x * 1e-3, b = 1.0 + blockIdx.x * 1e-6;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (MODE <= 2) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
      if (MODE == 1 || MODE == 3) x[i] = fma(x[i], 0.999999, 1e-7);
      if (MODE == 2 || MODE == 4) x[i] = __builtin_amdgcn_rcp(x[i]);
    }
  }
  double s = 0;
  for (int i = 0; i < 16; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3] + x[i];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_bf16(double* out, int iters) {
  v4f acc[16];
  for (int i = 0; i < 16; ++i) acc[i] = v4f{0, 0, 0, 0};
  v8bf a, b;
  for (int j = 0; j < 8; ++j) {
    a[j] = (__bf16)(threadIdx.x * 1e-3f + j);
    b[j] = (__bf16)(1.0f + j);
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
  }
  double s = 0;
  for (int i = 0; i < 16; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

template <class K>
static int timed(K launch, float* ms) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  CHECK(hipEventRecord(e0));
  launch();
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(e1));
  CHECK(hipEventSynchronize(e1));
  CHECK(hipEventElapsedTime(ms, e0, e1));
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return 0;
}

int main() {
  const int nwg = 256 * 8, iters = 10000;
  const double waves_per_simd = nwg * 4.0 / 1024.0;
  double* out = nullptr;
  CHECK(hipMalloc(&out, (size_t)nwg * 256 * sizeof(double)));
  float t[5] = {0, 0, 0, 0, 0}, tb = 0;
  for (int rep = 0; rep < 2; ++rep) {  // the first round warms up
    if (timed([&] { k_f64<0><<<nwg, 256>>>(out, iters); }, &t[0]) ||
        timed([&] { k_f64<1><<<nwg, 256>>>(out, iters); }, &t[1]) ||
        timed([&] { k_f64<2><<<nwg, 256>>>(out, iters); }, &t[2]) ||
        timed([&] { k_f64<3><<<nwg, 256>>>(out, iters); }, &t[3]) ||
        timed([&] { k_f64<4><<<nwg, 256>>>(out, iters); }, &t[4]) ||
        timed([&] { k_bf16<<<nwg, 256>>>(out, 5 * iters); }, &tb))
      return 1;
  }
  const double f64 = (double)nwg * 4 * 16 * 2048.0 * iters, bf = (double)nwg * 4 * 16 * 16384.0 * 5 * iters;
  std::printf("FP64 MFMA 16x16x4 only : %7.2f TFLOP/s\n", f64 / t[0] / 1e9);
  std::printf("BF16 MFMA 16x16x32 only: %7.1f TFLOP/s\n", bf / tb / 1e9);
  const double cyc_per_ms = 16.0 * 64.0 * iters * waves_per_simd / t[0];
  const char* nm[5] = {"16 FP64 MFMA", "16 MFMA + 16 FMA", "16 MFMA + 16 v_rcp_f64", "16 FMA", "16 v_rcp_f64"};
  for (int m = 0; m < 5; ++m)
    std::printf("%-24s %7.1f SIMD cycles per wave-iteration\n", nm[m], t[m] * cyc_per_ms / (iters * waves_per_simd));
  CHECK(hipFree(out));
  return 0;
}
