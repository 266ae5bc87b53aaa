// valu_rate -- the VALU issue ceiling of this GPU for the instruction mix of
// the face kernels: whole-chip wave-instructions per second of v_mul_f32,
// v_rcp_f32 and v_sqrt_f32 (non-packed FP32), each as 8 independent chains
// per thread on 8 waves per SIMD, timed with HIP events.  The face kernel's
// SQ_INSTS_VALU over its duration is compared with these rates to place it
// against the VALU roofline it is bound by.
//
// Usage: valu_rate   (prints one line per instruction)
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

template <int Op>
__global__ void __launch_bounds__(256) chains(float *out, int iters, float s) {
  float a[8];
  for (int i = 0; i < 8; ++i) a[i] = 1.0f + 0.001f * (float)(threadIdx.x + i);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (Op == 0)
        a[i] = a[i] * s;
      else if (Op == 1)
        a[i] = __builtin_amdgcn_rcpf(a[i]);
      else
        a[i] = __builtin_amdgcn_sqrtf(a[i]);
    }
  }
  float r = 0.0f;
  for (int i = 0; i < 8; ++i) r += a[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

bool ok(hipError_t e, const char *what) {
  if (e != hipSuccess) std::fprintf(stderr, "valu_rate: %s: %s\n", what, hipGetErrorString(e));
  return e == hipSuccess;
}

}  // namespace

int main() {
  hipDeviceProp_t prop;
  if (!ok(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties")) return 1;
  const int cus = prop.multiProcessorCount, threads = 256, iters = 4096;
  const int blocks = cus * 8;  // 8 workgroups of 4 waves per CU: 8 waves per SIMD
  float *out = nullptr;
  if (!ok(hipMalloc(&out, sizeof(float) * blocks * threads), "hipMalloc")) return 1;
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate"))
    return 1;
  const char *names[3] = {"v_mul_f32", "v_rcp_f32", "v_sqrt_f32"};
  const double winst = (double)blocks * (threads / 64) * iters * 8;
  for (int op = 0; op < 3; ++op) {
    float ms = 0.0f;
    for (int rep = 0; rep < 2; ++rep) {  // the first launch warms up
      if (!ok(hipEventRecord(e0, 0), "hipEventRecord")) return 1;
      if (op == 0) hipLaunchKernelGGL(chains<0>, dim3(blocks), dim3(threads), 0, 0, out, iters, 1.0000001f);
      if (op == 1) hipLaunchKernelGGL(chains<1>, dim3(blocks), dim3(threads), 0, 0, out, iters, 1.0f);
      if (op == 2) hipLaunchKernelGGL(chains<2>, dim3(blocks), dim3(threads), 0, 0, out, iters, 1.0f);
      if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e1, 0), "hipEventRecord") ||
          !ok(hipEventSynchronize(e1), "hipEventSynchronize") ||
          !ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime"))
        return 1;
    }
    std::printf("%-11s %8.3f ms  %.3e wave-instructions/s on %d CUs\n", names[op], ms,
                winst / (ms * 1e-3), cus);
  }
  (void)hipFree(out);
  return 0;
}
