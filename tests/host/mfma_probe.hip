// Lane-layout probe of the 16x16x4 MFMA builtins the assembly kernels use
// (one wave, every access in bounds).  With A[i][k] = (k==0 ? i : k==1 ? 1 : 0)
// and B[k][j] = (k==0 ? 1000 : k==1 ? j : 0), D[i][j] = 1000 i + j, so each
// output value names its (i, j).  Lane l supplies A from (i=l%16, k=l/16)
// and B from (k=l/16, j=l%16); the test reads back which D entries each lane
// holds.  `chain` repeats the accumulation `chain` times through a loop, as
// the kernels do, to check dependent MFMA sequences.
#include <hip/hip_runtime.h>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64) void probe_f64(double *out, int chain) {
  const int l = threadIdx.x;
  const int i = l & 15, k = l >> 4;
  const double a = k == 0 ? (double)i : (k == 1 ? 1.0 : 0.0);
  const double b = k == 0 ? 1000.0 : (k == 1 ? (double)(l & 15) : 0.0);
  v4d c = {0, 0, 0, 0};
  for (int t = 0; t < chain; t++) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  for (int r = 0; r < 4; r++) out[l * 4 + r] = c[r];
}

__global__ __launch_bounds__(64) void probe_f32(float *out, int chain) {
  const int l = threadIdx.x;
  const int i = l & 15, k = l >> 4;
  const float a = k == 0 ? (float)i : (k == 1 ? 1.0f : 0.0f);
  const float b = k == 0 ? 1000.0f : (k == 1 ? (float)(l & 15) : 0.0f);
  v4f c = {0, 0, 0, 0};
  for (int t = 0; t < chain; t++) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  for (int r = 0; r < 4; r++) out[l * 4 + r] = c[r];
}

// v_mfma_f32_16x16x32_f16: lane l supplies A[l%16][8*(l/16) + i] and
// B[8*(l/16) + i][l%16], i = 0..7 (the layout this probe checks); the same
// A and B values as above, so D[i][j] = 1000 i + j again.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));

__global__ __launch_bounds__(64) void probe_f16x32(float *out, int chain) {
  const int l = threadIdx.x;
  const int m = l & 15, g = l >> 4;
  half8 a, b;
  for (int i = 0; i < 8; i++) {
    const int k = 8 * g + i;
    a[i] = (_Float16)(k == 0 ? (float)m : (k == 1 ? 1.0f : 0.0f));
    b[i] = (_Float16)(k == 0 ? 1000.0f : (k == 1 ? (float)m : 0.0f));
  }
  v4f c = {0, 0, 0, 0};
  for (int t = 0; t < chain; t++) c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  for (int r = 0; r < 4; r++) out[l * 4 + r] = c[r];
}

extern "C" int mfma_probe_f16x32(int chain, double *host_out) {
  void *d = nullptr;
  if (hipMalloc(&d, 256 * sizeof(float)) != hipSuccess) return -1;
  float tmp[256];
  probe_f16x32<<<1, 64>>>((float *)d, chain);
  if (hipMemcpy(tmp, d, 256 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return -2;
  for (int i = 0; i < 256; i++) host_out[i] = tmp[i];
  if (hipDeviceSynchronize() != hipSuccess) return -3;
  (void)hipFree(d);
  return 0;
}

extern "C" int mfma_probe(int f64, int chain, double *host_out) {
  void *d = nullptr;
  if (hipMalloc(&d, 256 * sizeof(double)) != hipSuccess) return -1;
  if (f64) {
    probe_f64<<<1, 64>>>((double *)d, chain);
    if (hipMemcpy(host_out, d, 256 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -2;
  } else {
    float tmp[256];
    probe_f32<<<1, 64>>>((float *)d, chain);
    if (hipMemcpy(tmp, d, 256 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return -2;
    for (int i = 0; i < 256; i++) host_out[i] = tmp[i];
  }
  if (hipDeviceSynchronize() != hipSuccess) return -3;
  (void)hipFree(d);
  return 0;
}
