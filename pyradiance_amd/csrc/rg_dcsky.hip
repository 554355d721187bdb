// Daylight-coefficient x sky-matrix product (dctimestep / rmtxop multiply,
// util/cmatrix.c:421-474, util/rmatrix.c:806-836):
//   C[m][n][c] = sum_k A[m][k][c] * B[k][n][c]   for c = R, G, B.
//
// mode 1 (default): BF16 MFMA with split operands.  Each FP32 value x is
// carried as hi = bf16(x), lo = bf16(x - hi); the product is accumulated in
// FP32 as hi*hi + hi*lo + lo*hi, which leaves a relative error near 2^-16
// per term instead of bf16's 2^-8.  v_mfma_f32_16x16x32_bf16 (gfx950) does
// the products; both operands are first re-laid out K-contiguous per channel
// so every lane fetches its 8 bf16 values with one 16-byte load, and the next
// K step is fetched while the current one is on the matrix cores.
// mode 0: plain FP32 FMA, tiled through LDS (accuracy baseline).
#include <hip/hip_runtime.h>
#include <mutex>
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

// Both operands to hi/lo planes in one launch: blocks [0, nba) split
// A[M][K][3] into [3][M][Kp] (one element per thread), the others take a
// 32 (k) x 32 (n) tile of B[K][N][3] with its three channels (96 contiguous
// floats per k row) through LDS into [3][N][Kp] (K-contiguous per channel).
// K is zero-padded to Kp.
__global__ void __launch_bounds__(256) k_split(const float* A, const float* B, int M, int K, int N, int Kp,
                                               unsigned nba, unsigned short* Ah, unsigned short* Al,
                                               unsigned short* Bh, unsigned short* Bl) {
  __shared__ float tile[3][32][33];
  if (blockIdx.x < nba) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)3 * M * Kp) return;
    const int k = (int)(i % Kp);
    const size_t r = i / Kp;
    const int m = (int)(r % M), c = (int)(r / M);
    const float x = k < K ? A[((size_t)m * K + k) * 3 + c] : 0.0f;
    const unsigned short h = f2bf(x);
    Ah[i] = h;
    Al[i] = f2bf(x - bf2f(h));
    return;
  }
  const unsigned bt = blockIdx.x - nba;
  const int ntn = (N + 31) / 32;
  const int n0 = (int)(bt % ntn) * 32, k0 = (int)(bt / ntn) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int k = k0 + j;
#pragma unroll
    for (int e = tx; e < 96; e += 32) {
      const int n = n0 + e / 3;
      tile[e % 3][j][e / 3] = (k < K && n < N) ? B[((size_t)k * N + n) * 3 + e % 3] : 0.0f;
    }
  }
  __syncthreads();
  for (int j = ty; j < 96; j += 8) {
    const int c = j / 32, n = n0 + j % 32, k = k0 + tx;
    if (n < N && k < Kp) {
      const float x = tile[c][tx][j % 32];
      const unsigned short h = f2bf(x);
      const size_t o = ((size_t)c * N + n) * Kp + k;
      Bh[o] = h;
      Bl[o] = f2bf(x - bf2f(h));
    }
  }
}

constexpr int DB = 256;        // 4 wavefronts, 2 (m) x 2 (n)
#ifndef RG_DCSKY_WB
#define RG_DCSKY_WB 4          // 16-column tiles per wavefront along n
#endif
constexpr int WA = 2;          // 16-row tiles per wavefront along m
constexpr int WB = RG_DCSKY_WB;
constexpr int BM = 32 * WA, BN = 32 * WB;   // block tile
constexpr int GROUP_M = 8;     // m tiles walked together (L2 reuse of B)

__device__ __forceinline__ bf16x8 ld8(const unsigned short* p) {
  u16x8 v = *(const u16x8*)p;
  return __builtin_bit_cast(bf16x8, v);
}

// Block -> tile.  Workgroups are dispatched round-robin over the 8 XCDs, so
// block b runs on XCD b % 8; each XCD is given a contiguous range of tiles,
// walked in groups of GROUP_M row tiles, so the operand rows it re-reads stay
// in its own L2.
__device__ __forceinline__ bool tile_of(int Mt, int Nt, int& tm, int& tn) {
  const int T = Mt * Nt;
  const int per = (T + 7) / 8;
  const int b = blockIdx.x;
  const int slot = b >> 3;
  const int tile = (b & 7) * per + slot;
  if (slot >= per || tile >= T) return false;
  const int group = tile / (GROUP_M * Nt);
  const int first = group * GROUP_M;
  const int gsz = min(Mt - first, GROUP_M);
  const int in = tile - group * GROUP_M * Nt;
  tm = first + in % gsz;
  tn = in / gsz;
  return true;
}

// The MFMA computes D = X * Y with X the B fragment (rows = sky columns n)
// and Y the A fragment (columns = sensors m), so each lane ends up holding 4
// consecutive n of one m: with the three channels that is 12 consecutive
// floats of C, written as three 16-byte stores.
__global__ void __launch_bounds__(DB) k_dcsky_mfma(const unsigned short* Ah, const unsigned short* Al,
                                                   const unsigned short* Bh, const unsigned short* Bl, float* C,
                                                   int M, int N, int Kp, int Mt, int Nt) {
  int tm, tn;
  if (!tile_of(Mt, Nt, tm, tn)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = tm * BM + wm * 16 * WA, n0 = tn * BN + wn * 16 * WB;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  f32x4 acc[3][WA][WB];
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < WA; ++a)
      for (int b = 0; b < WB; ++b) acc[c][a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  // clamp rows/cols so loads stay in bounds; their results are not stored
  size_t ao[WA], bo[WB];
  for (int a = 0; a < WA; ++a) ao[a] = (size_t)min(m0 + 16 * a + r16, M - 1) * Kp + kq;
  for (int b = 0; b < WB; ++b) bo[b] = (size_t)min(n0 + 16 * b + r16, N - 1) * Kp + kq;
  for (int c = 0; c < 3; ++c) {
    const unsigned short* ah = Ah + (size_t)c * M * Kp;
    const unsigned short* al = Al + (size_t)c * M * Kp;
    const unsigned short* bh = Bh + (size_t)c * N * Kp;
    const unsigned short* bl = Bl + (size_t)c * N * Kp;
    bf16x8 fah[WA], fal[WA], fbh[WB], fbl[WB];
    for (int a = 0; a < WA; ++a) { fah[a] = ld8(ah + ao[a]); fal[a] = ld8(al + ao[a]); }
    for (int b = 0; b < WB; ++b) { fbh[b] = ld8(bh + bo[b]); fbl[b] = ld8(bl + bo[b]); }
    for (int k = 0; k < Kp; k += 32) {
      // prefetch the next K step while this one is on the matrix cores
      bf16x8 nah[WA], nal[WA], nbh[WB], nbl[WB];
      const int kn = k + 32 < Kp ? k + 32 : k;
      for (int a = 0; a < WA; ++a) { nah[a] = ld8(ah + ao[a] + kn); nal[a] = ld8(al + ao[a] + kn); }
      for (int b = 0; b < WB; ++b) { nbh[b] = ld8(bh + bo[b] + kn); nbl[b] = ld8(bl + bo[b] + kn); }
      for (int a = 0; a < WA; ++a)
        for (int b = 0; b < WB; ++b) {
          acc[c][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fbh[b], fah[a], acc[c][a][b], 0, 0, 0);
          acc[c][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fbl[b], fah[a], acc[c][a][b], 0, 0, 0);
          acc[c][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fbh[b], fal[a], acc[c][a][b], 0, 0, 0);
        }
      for (int a = 0; a < WA; ++a) { fah[a] = nah[a]; fal[a] = nal[a]; }
      for (int b = 0; b < WB; ++b) { fbh[b] = nbh[b]; fbl[b] = nbl[b]; }
    }
  }
  // D layout: lane holds D[n = 4 * (lane >> 4) + i][m = lane & 15]
  const bool vec = (N & 3) == 0;
  for (int a = 0; a < WA; ++a) {
    const int m = m0 + 16 * a + r16;
    if (m >= M) continue;
    for (int b = 0; b < WB; ++b) {
      const int nb = n0 + 16 * b + (lane >> 4) * 4;
      if (nb >= N) continue;
      float* dst = C + ((size_t)m * N + nb) * 3;
      const f32x4 r = acc[0][a][b], g = acc[1][a][b], u = acc[2][a][b];
      if (vec) {   // nb + 3 < N as N is a multiple of 4; 16-byte aligned
        float4* d4 = (float4*)dst;
        d4[0] = make_float4(r[0], g[0], u[0], r[1]);
        d4[1] = make_float4(g[1], u[1], r[2], g[2]);
        d4[2] = make_float4(u[2], r[3], g[3], u[3]);
      } else {
        for (int i = 0; i < 4 && nb + i < N; ++i) {
          dst[3 * i + 0] = r[i];
          dst[3 * i + 1] = g[i];
          dst[3 * i + 2] = u[i];
        }
      }
    }
  }
}

// LDS-staged form.  A block of 16 wavefronts (4 x 4) computes a
// 128 x 128 tile of all three channels; each wavefront holds a 32 x 32 tile
// (2 x 2 MFMA tiles x 3 channels of accumulators).  For each K step of 32
// and channel, the block's A and B hi/lo tiles (128 rows x 32 K each, 32 KB)
// are loaded once into LDS and read by all 16 wavefronts; the next stage's
// global loads are issued before the current stage's MFMAs, into the second
// LDS buffer.  Rows are padded to 40 bf16 so the 16 lanes that read one
// fragment column hit different banks.
constexpr int LBM = 128, LBN = 128;
#ifndef RG_DCSKY_LKS
#define RG_DCSKY_LKS 64
#endif
constexpr int LKS = RG_DCSKY_LKS;             // K per stage (32 or 64; 64 measured 22-28 % faster at K = 2306)
constexpr int LROW = LKS + 8;                 // padded row length (bf16)
constexpr int LTILE = 128 * LROW;             // one 128 x LKS plane tile in LDS (bf16)
constexpr int LCPR = LKS / 8;                 // 16-byte chunks per tile row

// WM x WN wavefronts, each holding TA x TB 16 x 16 MFMA tiles of all three
// channels (16 * TA * WM = 16 * TB * WN = 128).
template <int WM, int WN, int TA, int TB>
__global__ void __launch_bounds__(64 * WM * WN) k_dcsky_lds(const unsigned short* Ah, const unsigned short* Al,
                                                            const unsigned short* Bh, const unsigned short* Bl,
                                                            float* C, int M, int N, int Kp, int Mt, int Nt) {
  static_assert(16 * TA * WM == LBM && 16 * TB * WN == LBN, "block tile");
  constexpr int NT = 64 * WM * WN;
  constexpr int NCH = 4 * 128 * LCPR / NT;    // 16-byte chunks each thread stages per stage
  __shared__ __attribute__((aligned(16))) unsigned short lds[2][4][LTILE];   // [buffer][Ah, Al, Bh, Bl]
  int tm, tn;
  if (!tile_of(Mt, Nt, tm, tn)) return;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int mb = tm * LBM, nb = tn * LBN;
  // staging: chunk q (16 B = 8 bf16) of the stage: plane q / (128 LCPR),
  // row (q mod 128 LCPR) / LCPR, part q mod LCPR
  int soff[NCH], splane[NCH];
  size_t sg[NCH];
  const size_t csz[4] = {(size_t)M * Kp, (size_t)M * Kp, (size_t)N * Kp, (size_t)N * Kp};
  for (int j = 0; j < NCH; ++j) {
    const int q = t + j * NT;
    const int pl = q / (128 * LCPR), row = (q % (128 * LCPR)) / LCPR, part = q % LCPR;
    splane[j] = pl;
    soff[j] = row * LROW + part * 8;
    sg[j] = (pl < 2 ? (size_t)min(mb + row, M - 1) : (size_t)min(nb + row, N - 1)) * Kp + part * 8;
  }
  const unsigned short* planes[4] = {Ah, Al, Bh, Bl};
  auto gload = [&](int stage, u16x8* reg) {
    const int k0 = (stage / 3) * LKS, c = stage % 3;
    for (int j = 0; j < NCH; ++j) reg[j] = *(const u16x8*)(planes[splane[j]] + c * csz[splane[j]] + sg[j] + k0);
  };
  auto lstore = [&](int buf, const u16x8* reg) {
    for (int j = 0; j < NCH; ++j) *(u16x8*)&lds[buf][splane[j]][soff[j]] = reg[j];
  };
  f32x4 acc[3][TA][TB];
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < TA; ++a)
      for (int b = 0; b < TB; ++b) acc[c][a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nstage = 3 * (Kp / LKS);
  u16x8 reg[NCH];
  gload(0, reg);
  lstore(0, reg);
  __syncthreads();
  for (int s = 0; s < nstage; ++s) {
    const int buf = s & 1;
    if (s + 1 < nstage) gload(s + 1, reg);
    const int c = s % 3;
#pragma unroll
    for (int kk = 0; kk < LKS; kk += 32) {
      bf16x8 fah[TA], fal[TA], fbh[TB], fbl[TB];
      for (int a = 0; a < TA; ++a) {
        const int row = wm * 16 * TA + 16 * a + r16;
        fah[a] = __builtin_bit_cast(bf16x8, *(const u16x8*)&lds[buf][0][row * LROW + kk + kq]);
        fal[a] = __builtin_bit_cast(bf16x8, *(const u16x8*)&lds[buf][1][row * LROW + kk + kq]);
      }
      for (int b = 0; b < TB; ++b) {
        const int row = wn * 16 * TB + 16 * b + r16;
        fbh[b] = __builtin_bit_cast(bf16x8, *(const u16x8*)&lds[buf][2][row * LROW + kk + kq]);
        fbl[b] = __builtin_bit_cast(bf16x8, *(const u16x8*)&lds[buf][3][row * LROW + kk + kq]);
      }
      for (int a = 0; a < TA; ++a)
        for (int b = 0; b < TB; ++b) {
          f32x4 v = c == 0 ? acc[0][a][b] : (c == 1 ? acc[1][a][b] : acc[2][a][b]);
          v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fbh[b], fah[a], v, 0, 0, 0);
          v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fbl[b], fah[a], v, 0, 0, 0);
          v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fbh[b], fal[a], v, 0, 0, 0);
          if (c == 0) acc[0][a][b] = v;
          else if (c == 1) acc[1][a][b] = v;
          else acc[2][a][b] = v;
        }
    }
    if (s + 1 < nstage) lstore(buf ^ 1, reg);
    __syncthreads();
  }
  // D layout: lane holds D[n = 4 * (lane >> 4) + i][m = lane & 15]
  const bool vec = (N & 3) == 0;
  for (int a = 0; a < TA; ++a) {
    const int m = mb + wm * 16 * TA + 16 * a + r16;
    if (m >= M) continue;
    for (int b = 0; b < TB; ++b) {
      const int nq = nb + wn * 16 * TB + 16 * b + (lane >> 4) * 4;
      if (nq >= N) continue;
      float* dst = C + ((size_t)m * N + nq) * 3;
      const f32x4 r = acc[0][a][b], g = acc[1][a][b], u = acc[2][a][b];
      if (vec) {
        float4* d4 = (float4*)dst;
        d4[0] = make_float4(r[0], g[0], u[0], r[1]);
        d4[1] = make_float4(g[1], u[1], r[2], g[2]);
        d4[2] = make_float4(u[2], r[3], g[3], u[3]);
      } else {
        for (int i = 0; i < 4 && nq + i < N; ++i) {
          dst[3 * i + 0] = r[i];
          dst[3 * i + 1] = g[i];
          dst[3 * i + 2] = u[i];
        }
      }
    }
  }
}

// Short-K form (K <= 160, e.g. 146 Klems patches): the product is bound by
// the output write, and the operands are small enough that a wavefront keeps
// its 16 rows of A (hi and lo, all K, three channels) in registers for the
// whole launch.  A block of 8 wavefronts owns a slab of 128 rows of A and
// walks a range of sky columns in chunks of 32: each chunk's B planes (hi/lo,
// three channels, all K) are staged into LDS once, double-buffered (the next
// chunk's global loads are in flight during this chunk's MFMAs), and every
// wavefront multiplies its rows by them and stores its 16 x 32 x 3 outputs
// before the next chunk.  No K loop over global memory, one barrier per chunk.
// Every lane issues the same loads and stores in every chunk (pieces past
// the chunk load a valid piece again; outputs outside C go to `sink`), so the
// wait for the next chunk's loads never waits for this chunk's stores.
constexpr int SW = 8;                  // wavefronts per block
constexpr int SROWS = 16 * SW;         // rows of A per block (a slab)
constexpr int SCOLS = 32;              // sky columns per staged chunk
constexpr int SHORT_KS = 5;            // K steps of 32 the short form covers (Kp <= 160)
constexpr int SINK_FLOATS = 64 * 12;   // one wavefront's worth of redirected outputs

template <int KS, bool VEC>
__global__ void __launch_bounds__(64 * SW) k_dcsky_short(const unsigned short* Ah, const unsigned short* Al,
                                                         const unsigned short* Bh, const unsigned short* Bl, float* C,
                                                         float* sink, int M, int N, int nslab, int nrange) {
  constexpr int Kp = 32 * KS;
  constexpr int ROW = Kp + 8;                  // padded LDS row (bf16)
  constexpr int PL = SCOLS * ROW;              // one (channel, hi|lo) tile
  constexpr int NT = 64 * SW;
  constexpr int CPR = Kp / 8;                  // 16-byte pieces per row
  constexpr int NPC = 6 * SCOLS * CPR;         // pieces per chunk
  constexpr int NLD = (NPC + NT - 1) / NT;     // pieces per thread
  __shared__ __attribute__((aligned(16))) unsigned short lds[2][6][PL];   // [buffer][2 c + lo]
  // block -> (slab, column range); blocks run on XCD b % 8, and the blocks
  // of one XCD take consecutive q: all slabs of one column range, which
  // share its B chunks in that XCD's L2
  const int per = (int)((gridDim.x + 7) / 8);
  const int q = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (q >= nslab * nrange) return;
  const int slab = q % nslab, range = q / nslab;
  const int nchunk = (N + SCOLS - 1) / SCOLS;
  const int cb = (int)((long long)range * nchunk / nrange), ce = (int)((long long)(range + 1) * nchunk / nrange);
  if (cb >= ce) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int m = slab * SROWS + wave * 16 + r16;
  const int mc = min(m, M - 1);                // clamped for loads; rows >= M are not stored
  bf16x8 ah[3][KS], al[3][KS];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const size_t o = ((size_t)c * M + mc) * Kp + s * 32 + kq;
      ah[c][s] = ld8(Ah + o);
      al[c][s] = ld8(Al + o);
    }
  // staging piece j of this thread: plane pl = 2 c + lo, row, 16-byte part
  int soff[NLD], spl[NLD], srow[NLD], spart[NLD];
#pragma unroll
  for (int j = 0; j < NLD; ++j) {
    const int p = min(t + j * NT, NPC - 1);
    const int rem = p % (SCOLS * CPR);
    spl[j] = p / (SCOLS * CPR);
    srow[j] = rem / CPR;
    spart[j] = rem % CPR;
    soff[j] = srow[j] * ROW + spart[j] * 8;
  }
  auto gload = [&](int chunk, u16x8* reg) {
    const int n0 = chunk * SCOLS;
#pragma unroll
    for (int j = 0; j < NLD; ++j) {
      const unsigned short* src = (spl[j] & 1) ? Bl : Bh;
      reg[j] = *(const u16x8*)(src + ((size_t)(spl[j] >> 1) * N + min(n0 + srow[j], N - 1)) * Kp + spart[j] * 8);
    }
  };
  auto lstore = [&](int buf, const u16x8* reg) {
#pragma unroll
    for (int j = 0; j < NLD; ++j) *(u16x8*)&lds[buf][spl[j]][soff[j]] = reg[j];
  };
  u16x8 reg[NLD];
  gload(cb, reg);
  lstore(0, reg);
  // every load so far (A and the first chunk) complete here, so that no wait
  // inside the loop, where this chunk's stores are also outstanding, is
  // counted against them
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  float* const mysink = sink + lane * 12;
  for (int ch = cb; ch < ce; ++ch) {
    const int buf = (ch - cb) & 1;
    gload(min(ch + 1, ce - 1), reg);
    __builtin_amdgcn_sched_barrier(0);     // the next chunk's loads go out before the MFMAs
    f32x4 acc[3][2];
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c][0] = acc[c][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int row = 16 * b + r16;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const bf16x8 bh = __builtin_bit_cast(bf16x8, *(const u16x8*)&lds[buf][2 * c][row * ROW + s * 32 + kq]);
          const bf16x8 bl = __builtin_bit_cast(bf16x8, *(const u16x8*)&lds[buf][2 * c + 1][row * ROW + s * 32 + kq]);
          acc[c][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, ah[c][s], acc[c][b], 0, 0, 0);
          acc[c][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, ah[c][s], acc[c][b], 0, 0, 0);
          acc[c][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, al[c][s], acc[c][b], 0, 0, 0);
        }
      }
    // D layout: lane holds D[n = 4 * (lane >> 4) + i][m = lane & 15], 12
    // consecutive floats of row m
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int nq = ch * SCOLS + 16 * b + (lane >> 4) * 4;
      const f32x4 r = acc[0][b], g = acc[1][b], u = acc[2][b];
      if (VEC) {   // N a multiple of 4: nq < N implies nq + 3 < N, 16-byte aligned
        float* dst = (m < M && nq < N) ? C + ((size_t)m * N + nq) * 3 : mysink;
        float4* d4 = (float4*)dst;
        d4[0] = make_float4(r[0], g[0], u[0], r[1]);
        d4[1] = make_float4(g[1], u[1], r[2], g[2]);
        d4[2] = make_float4(u[2], r[3], g[3], u[3]);
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float* dst = (m < M && nq + i < N) ? C + ((size_t)m * N + nq + i) * 3 : mysink + 3 * i;
          dst[0] = r[i];
          dst[1] = g[i];
          dst[2] = u[i];
        }
      }
    }
    lstore(buf ^ 1, reg);    // the last chunk stages a copy of itself into the idle buffer
    __syncthreads();
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

// Split planes of a call, kept per (device, stream) and grown as needed (a
// stream-ordered allocation and free per call cost about as much as the whole
// product at K = 146).  The lock is held while a call queues its kernels, so a
// buffer is only replaced after the work queued on it has finished.
struct DcScratch {
  void* p = nullptr;
  size_t bytes = 0;
};
static std::mutex g_dc_mu;
static std::map<std::pair<int, hipStream_t>, DcScratch> g_dc_scratch;

static unsigned short* dc_scratch(size_t bytes, hipStream_t st) {
  int dev = 0;
  dck(hipGetDevice(&dev), "dcsky device");
  DcScratch& s = g_dc_scratch[{dev, st}];
  if (s.bytes < bytes) {
    if (s.p) {
      dck(hipStreamSynchronize(st), "dcsky scratch sync");
      dck(hipFree(s.p), "dcsky scratch free");
      s.p = nullptr;
      s.bytes = 0;
    }
    dck(hipMalloc(&s.p, bytes), "dcsky scratch");
    s.bytes = bytes;
  }
  return (unsigned short*)s.p;
}

// Column ranges per slab for the short-K form: one block per CU at a time
// (its LDS), so the fewest ranges that keep at least 90 % of the CUs busy
// over whole rounds of blocks.
static int short_ranges(int nslab, int nchunk) {
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  ncu = std::max(ncu, 1);
  int best = 1;
  double beff = 0.0;
  for (int r = 1; r <= std::min(nchunk, 64); ++r) {
    const long long G = (long long)nslab * r;
    const long long rounds = (G + ncu - 1) / ncu;
    const double eff = (double)G / (double)(rounds * ncu);
    if (eff > beff + 1e-9) { beff = eff; best = r; }
    if (eff >= 0.9) break;
  }
  return best;
}

template <int KS>
static void launch_short(const unsigned short* Ah, const unsigned short* Al, const unsigned short* Bh,
                         const unsigned short* Bl, float* C, float* sink, int M, int N, hipStream_t st) {
  const int nslab = (M + SROWS - 1) / SROWS, nchunk = (N + SCOLS - 1) / SCOLS;
  const int nrange = short_ranges(nslab, nchunk);
  const long long G = (long long)nslab * nrange;
  if (G > 0x7fffffffLL - 8) throw std::runtime_error("dcsky: matrix too large");
  const unsigned nblk = (unsigned)((G + 7) / 8 * 8);
  if (N % 4 == 0)
    hipLaunchKernelGGL((k_dcsky_short<KS, true>), dim3(nblk), dim3(64 * SW), 0, st, Ah, Al, Bh, Bl, C, sink, M, N,
                       nslab, nrange);
  else
    hipLaunchKernelGGL((k_dcsky_short<KS, false>), dim3(nblk), dim3(64 * SW), 0, st, Ah, Al, Bh, Bl, C, sink, M, N,
                       nslab, nrange);
}

void dcsky(const float* A, const float* B, float* C, int M, int K, int N, int mode, hipStream_t st) {
  if (mode == 0) {
    hipLaunchKernelGGL(k_dcsky_f32, dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, st, A, B, C, M, K, N);
    return;
  }
  // Short K (a few K steps) is bound by the output write: K <= 160 keeps A in
  // registers (mode 4 forces that form), up to 256 the per-wavefront kernel;
  // long K is MFMA work and needs the LDS-staged tiles.  mode 2 forces the
  // per-wavefront kernel, mode 3 the staged one.
  const int k32 = (K + 31) / 32 * 32;
  if (mode == 4 && k32 > 32 * SHORT_KS) throw std::runtime_error("dcsky: mode 4 needs K <= 160");
  const bool shortk = mode == 4 || (mode == 1 && k32 <= 32 * SHORT_KS);
  const bool staged = !shortk && (mode == 3 || (mode == 1 && K > 256));
  const int kstep = staged ? LKS : 32;
  int Kp = (K + kstep - 1) / kstep * kstep;
  size_t na = (size_t)3 * M * Kp, nb = (size_t)3 * N * Kp;
  std::lock_guard<std::mutex> lk(g_dc_mu);
  // split planes, then (16-byte aligned) the short form's output sink
  const size_t plane_bytes = (2 * (na + nb) * sizeof(unsigned short) + 255) / 256 * 256;
  unsigned short* buf = dc_scratch(plane_bytes + SINK_FLOATS * sizeof(float), st);
  unsigned short *Ah = buf, *Al = buf + na, *Bh = buf + 2 * na, *Bl = buf + 2 * na + nb;
  float* sink = (float*)((char*)buf + plane_bytes);
  const size_t nba = (na + 255) / 256, nbb = (size_t)((N + 31) / 32) * (Kp / 32);
  if (nba + nbb > 0x7fffffffu) throw std::runtime_error("dcsky: matrix too large");
  hipLaunchKernelGGL(k_split, dim3((unsigned)(nba + nbb)), dim3(256), 0, st, A, B, M, K, N, Kp, (unsigned)nba, Ah, Al,
                     Bh, Bl);
  if (shortk) {
    switch (Kp / 32) {
      case 1: launch_short<1>(Ah, Al, Bh, Bl, C, sink, M, N, st); break;
      case 2: launch_short<2>(Ah, Al, Bh, Bl, C, sink, M, N, st); break;
      case 3: launch_short<3>(Ah, Al, Bh, Bl, C, sink, M, N, st); break;
      case 4: launch_short<4>(Ah, Al, Bh, Bl, C, sink, M, N, st); break;
      default: launch_short<5>(Ah, Al, Bh, Bl, C, sink, M, N, st); break;
    }
  } else if (!staged) {
    const int Mt = (M + BM - 1) / BM, Nt = (N + BN - 1) / BN;
    const long long tiles = (long long)Mt * Nt;
    if (tiles > 0x7fffffffLL - 8) throw std::runtime_error("dcsky: matrix too large");
    const unsigned nblk = (unsigned)((tiles + 7) / 8 * 8);
    hipLaunchKernelGGL(k_dcsky_mfma, dim3(nblk), dim3(DB), 0, st, Ah, Al, Bh, Bl, C, M, N, Kp, Mt, Nt);
  } else {
    const int Mt = (M + LBM - 1) / LBM, Nt = (N + LBN - 1) / LBN;
    const long long tiles = (long long)Mt * Nt;
    if (tiles > 0x7fffffffLL - 8) throw std::runtime_error("dcsky: matrix too large");
    const unsigned nblk = (unsigned)((tiles + 7) / 8 * 8);
    hipLaunchKernelGGL((k_dcsky_lds<4, 4, 2, 2>), dim3(nblk), dim3(1024), 0, st, Ah, Al, Bh, Bl, C, M, N, Kp, Mt,
                       Nt);
  }
}

}  // namespace rg
