// Daylight-coefficient x sky-matrix product (dctimestep / rmtxop multiply,
// util/cmatrix.c:421-474, util/rmatrix.c:806-836):
//   C[m][n][c] = sum_k A[m][k][c] * B[k][n][c]   for c = R, G, B.
//
// mode 1 (default): BF16 MFMA with split operands.  Each FP32 value x is
// carried as hi = bf16(x), lo = bf16(x - hi); the product is accumulated in
// FP32 as hi*hi + hi*lo + lo*hi, which leaves a relative error near 2^-16
// per term instead of bf16's 2^-8.  v_mfma_f32_16x16x32_bf16 (gfx950) does
// the products; both operands are first re-laid out K-contiguous per channel
// so every lane fetches its 8 bf16 values with one 16-byte load.
// mode 0: plain FP32 FMA, tiled through LDS (accuracy baseline).
#include <hip/hip_runtime.h>
#include "rg_render.hpp"

namespace rg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short f2bf(float f) {   // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

// A[M][K][3] -> hi/lo planes [3][M][Kp]
__global__ void k_split_a(const float* A, int M, int K, int Kp, unsigned short* hi, unsigned short* lo) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t tot = (size_t)3 * M * Kp;
  if (i >= tot) return;
  int k = (int)(i % Kp);
  size_t r = i / Kp;
  int m = (int)(r % M);
  int c = (int)(r / M);
  float x = k < K ? A[((size_t)m * K + k) * 3 + c] : 0.0f;
  unsigned short h = f2bf(x);
  hi[i] = h;
  lo[i] = f2bf(x - bf2f(h));
}

// B[K][N][3] -> hi/lo planes [3][N][Kp] (transposed through LDS)
__global__ void k_split_b(const float* B, int K, int N, int Kp, unsigned short* hi, unsigned short* lo) {
  __shared__ float tile[32][33];
  int c = blockIdx.z;
  int n0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
  int tx = threadIdx.x, ty = threadIdx.y;   // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    int k = k0 + j, n = n0 + tx;
    tile[j][tx] = (k < K && n < N) ? B[((size_t)k * N + n) * 3 + c] : 0.0f;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    int n = n0 + j, k = k0 + tx;
    if (n < N && k < Kp) {
      float x = tile[tx][j];
      unsigned short h = f2bf(x);
      size_t o = ((size_t)c * N + n) * Kp + k;
      hi[o] = h;
      lo[o] = f2bf(x - bf2f(h));
    }
  }
}

constexpr int DB = 256;        // 4 wavefronts, 2 x 2
constexpr int BM = 64, BN = 64;

__device__ __forceinline__ bf16x8 ld8(const unsigned short* p) {
  u16x8 v = *(const u16x8*)p;
  return __builtin_bit_cast(bf16x8, v);
}

__global__ void __launch_bounds__(DB) k_dcsky_mfma(const unsigned short* Ah, const unsigned short* Al,
                                                   const unsigned short* Bh, const unsigned short* Bl, float* C,
                                                   int M, int N, int Kp) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * BM + wm * 32, n0 = blockIdx.x * BN + wn * 32;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  f32x4 acc[3][2][2];
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) acc[c][a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  // clamp rows/cols so loads stay in bounds; their results are not stored
  int mr[2], nr[2];
  for (int a = 0; a < 2; ++a) {
    mr[a] = min(m0 + 16 * a + r16, M - 1);
    nr[a] = min(n0 + 16 * a + r16, N - 1);
  }
  for (int c = 0; c < 3; ++c) {
    const unsigned short* ah = Ah + (size_t)c * M * Kp;
    const unsigned short* al = Al + (size_t)c * M * Kp;
    const unsigned short* bh = Bh + (size_t)c * N * Kp;
    const unsigned short* bl = Bl + (size_t)c * N * Kp;
    for (int k = 0; k < Kp; k += 32) {
      bf16x8 fah[2], fal[2], fbh[2], fbl[2];
      for (int a = 0; a < 2; ++a) {
        fah[a] = ld8(ah + (size_t)mr[a] * Kp + k + kq);
        fal[a] = ld8(al + (size_t)mr[a] * Kp + k + kq);
        fbh[a] = ld8(bh + (size_t)nr[a] * Kp + k + kq);
        fbl[a] = ld8(bl + (size_t)nr[a] * Kp + k + kq);
      }
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          acc[c][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fah[a], fbh[b], acc[c][a][b], 0, 0, 0);
          acc[c][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fah[a], fbl[b], acc[c][a][b], 0, 0, 0);
          acc[c][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fal[a], fbh[b], acc[c][a][b], 0, 0, 0);
        }
    }
  }
  // D layout: lane holds column (lane & 15), rows 4*(lane>>4) + i.  The
  // block's 64 x 64 x 3 tile is staged in LDS in C's [m][n][c] order, then
  // written row by row with coalesced stores.
  __shared__ float ct[BM][BN * 3 + 4];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      int nl = wn * 32 + 16 * b + r16;
      for (int i = 0; i < 4; ++i) {
        int ml = wm * 32 + 16 * a + (lane >> 4) * 4 + i;
        ct[ml][3 * nl + 0] = acc[0][a][b][i];
        ct[ml][3 * nl + 1] = acc[1][a][b][i];
        ct[ml][3 * nl + 2] = acc[2][a][b][i];
      }
    }
  __syncthreads();
  const int mb = blockIdx.y * BM, nb = blockIdx.x * BN;
  const int ncols = min(BN, N - nb) * 3;     // valid floats per tile row
  const bool vec = (ncols == BN * 3) && (((size_t)N * 3) % 4 == 0);
  for (int ml = wave; ml < BM; ml += DB / 64) {
    int m = mb + ml;
    if (m >= M) break;
    float* row = C + ((size_t)m * N + nb) * 3;
    if (vec) {
      for (int e = lane; e < BN * 3 / 4; e += 64) {
        float4 v = make_float4(ct[ml][4 * e], ct[ml][4 * e + 1], ct[ml][4 * e + 2], ct[ml][4 * e + 3]);
        *(float4*)(row + 4 * e) = v;
      }
    } else {
      for (int e = lane; e < ncols; e += 64) row[e] = ct[ml][e];
    }
  }
}

// FP32 reference path: 16x16 threads, 64x64 tile, LDS staged.
__global__ void __launch_bounds__(256) k_dcsky_f32(const float* A, const float* B, float* C, int M, int K, int N) {
  __shared__ float As[3][64][17];
  __shared__ float Bs[3][16][65];
  int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  float acc[3][4][4] = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
    for (int t = threadIdx.x; t < 3 * 64 * 16; t += 256) {
      int c = t % 3, kk = (t / 3) % 16, mm = t / 48;
      int m = m0 + mm, k = k0 + kk;
      As[c][mm][kk] = (m < M && k < K) ? A[((size_t)m * K + k) * 3 + c] : 0.f;
    }
    for (int t = threadIdx.x; t < 3 * 64 * 16; t += 256) {
      int c = t % 3, nn = (t / 3) % 64, kk = t / 192;
      int n = n0 + nn, k = k0 + kk;
      Bs[c][kk][nn] = (n < N && k < K) ? B[((size_t)k * N + n) * 3 + c] : 0.f;
    }
    __syncthreads();
    for (int kk = 0; kk < 16; ++kk)
      for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 4; ++i) {
          float a = As[c][ty * 4 + i][kk];
          for (int j = 0; j < 4; ++j) acc[c][i][j] = fmaf(a, Bs[c][kk][tx * 4 + j], acc[c][i][j]);
        }
    __syncthreads();
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      int m = m0 + ty * 4 + i, n = n0 + tx * 4 + j;
      if (m < M && n < N)
        for (int c = 0; c < 3; ++c) C[((size_t)m * N + n) * 3 + c] = acc[c][i][j];
    }
}

static void dck(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}

void dcsky(const float* A, const float* B, float* C, int M, int K, int N, int mode, hipStream_t st) {
  if (mode == 0) {
    hipLaunchKernelGGL(k_dcsky_f32, dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, st, A, B, C, M, K, N);
    return;
  }
  int Kp = (K + 31) / 32 * 32;
  size_t na = (size_t)3 * M * Kp, nb = (size_t)3 * N * Kp;
  unsigned short* buf = nullptr;
  dck(hipMallocAsync((void**)&buf, 2 * (na + nb) * sizeof(unsigned short), st), "dcsky scratch");
  unsigned short *Ah = buf, *Al = buf + na, *Bh = buf + 2 * na, *Bl = buf + 2 * na + nb;
  hipLaunchKernelGGL(k_split_a, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, st, A, M, K, Kp, Ah, Al);
  hipLaunchKernelGGL(k_split_b, dim3((N + 31) / 32, Kp / 32, 3), dim3(32, 8), 0, st, B, K, N, Kp, Bh, Bl);
  hipLaunchKernelGGL(k_dcsky_mfma, dim3((N + BN - 1) / BN, (M + BM - 1) / BM), dim3(DB), 0, st, Ah, Al, Bh, Bl, C, M,
                     N, Kp);
  dck(hipFreeAsync(buf, st), "dcsky scratch free");
}

}  // namespace rg
