// The memory traffic of one E curl half-step (k_update_e) with the curl
// arithmetic, CPML, TFSF and coefficient tables taken out: what the access
// pattern alone costs on this GPU, the practical ceiling of the split curl
// kernels.
//
// Same fields and layout as the engine: (n+1)^2 rows of pz = (n+1) rounded
// up to 32 elements, z fastest; the same tiles (64 k x 4 j threads marching
// 8 x planes, tile ids in z, y, x order); per cell the same loads (Ex, Ey,
// Ez, Hx, Hy, Hz, the material byte, Hy and Hx at k-1, Hz and Hx at j-1; Hz
// and Hy at i-1 carried from the previous plane) and the same in-place
// stores of Ex, Ey, Ez.  Prints one JSON line per precision.
//
//   pattern_model [n = 400] [reps = 20]
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

namespace {

constexpr int TZ = 64, TY = 4, TX = 8;

template <typename T>
__global__ void __launch_bounds__(TZ * TY) k_pattern(T *__restrict__ ex, T *__restrict__ ey, T *__restrict__ ez,
                                                     const T *__restrict__ hx, const T *__restrict__ hy,
                                                     const T *__restrict__ hz, const unsigned char *__restrict__ mat,
                                                     int n, int pz, int ntz, int nty) {
  const int b = blockIdx.x;
  const int bz = b % ntz, by = (b / ntz) % nty, bx = b / (ntz * nty);
  const int k = bz * TZ + (int)(threadIdx.x % TZ), j = by * TY + (int)(threadIdx.x / TZ);
  if (k > n || j > n) return;
  const long long sj = pz, si = (long long)(n + 1) * pz;
  const int i0 = bx * TX, i1 = i0 + TX < n + 1 ? i0 + TX : n + 1;
  const long long dk = k > 0 ? 1 : 0, dj = j > 0 ? sj : 0;
  long long c = (long long)i0 * si + (long long)j * sj + k;
  T hz_prev = T(0), hy_prev = T(0);
  if (i0 > 0) {
    hz_prev = hz[c - si];
    hy_prev = hy[c - si];
  }
  for (int i = i0; i < i1; ++i, c += si) {
    const T m = (T)mat[c];
    const T e0 = ex[c], e1 = ey[c], e2 = ez[c];
    const T h0 = hx[c], h1 = hy[c], h2 = hz[c];
    const T h1k = hy[c - dk], h0k = hx[c - dk], h2j = hz[c - dj], h0j = hx[c - dj];
    ex[c] = m * e0 + (h1 - h1k) - (h2 - h2j);
    ey[c] = m * e1 + (h0 - h0k) - (h2 - hz_prev);
    ez[c] = m * e2 + (h0 - h0j) - (h1 - hy_prev);
    hz_prev = h2;
    hy_prev = h1;
  }
}

#define CHECK(x)                                                                     \
  do {                                                                               \
    hipError_t err_ = (x);                                                           \
    if (err_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "pattern_model: %s at line %d\n", hipGetErrorString(err_), \
                   __LINE__);                                                        \
      return 1;                                                                      \
    }                                                                                \
  } while (0)

template <typename T>
int measure(int n, int reps, const char *name) {
  const int pz = ((n + 1) + 31) / 32 * 32;
  const size_t cells = (size_t)(n + 1) * (n + 1) * pz;
  T *f[6] = {};
  unsigned char *mat = nullptr;
  for (int q = 0; q < 6; ++q) {
    CHECK(hipMalloc(&f[q], cells * sizeof(T)));
    CHECK(hipMemset(f[q], 0, cells * sizeof(T)));
  }
  CHECK(hipMalloc(&mat, cells));
  CHECK(hipMemset(mat, 1, cells));
  const int ntz = (pz + TZ - 1) / TZ, nty = (n + 1 + TY - 1) / TY, ntx = (n + 1 + TX - 1) / TX;
  const dim3 grid((unsigned)ntz * nty * ntx), block(TZ * TY);
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  for (int w = 0; w < 3; ++w)
    hipLaunchKernelGGL(k_pattern<T>, grid, block, 0, 0, f[0], f[1], f[2], f[3], f[4], f[5], mat, n, pz, ntz, nty);
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_pattern<T>, grid, block, 0, 0, f[0], f[1], f[2], f[3], f[4], f[5], mat, n, pz, ntz, nty);
  CHECK(hipEventRecord(e1, 0));
  CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  ms /= (float)reps;
  // the curl kernels' algorithmic bytes: n^3 cells x (9 reals + 1 byte)
  const double bytes = (double)n * n * n * (9.0 * sizeof(T) + 1.0);
  std::printf("{\"precision\": \"%s\", \"grid\": %d, \"ms\": %.4f, \"bytes_per_launch\": %.0f, "
              "\"algorithmic_gbs\": %.1f}\n",
              name, n, ms, bytes, bytes / (ms * 1e-3) / 1e9);
  std::fflush(stdout);
  for (int q = 0; q < 6; ++q) CHECK(hipFree(f[q]));
  CHECK(hipFree(mat));
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 400;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
  if (n < 2 || n > 1000 || reps < 1) {
    std::fprintf(stderr, "pattern_model: grid edge in [2, 1000] and reps >= 1\n");
    return 2;
  }
  if (measure<float>(n, reps, "f32")) return 1;
  if (measure<double>(n, reps, "f64")) return 1;
  return 0;
}
