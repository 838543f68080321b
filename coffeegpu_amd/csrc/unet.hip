// U-Net building blocks for gfx950 (include/coffee_unet.h): BF16 activations
// and weights, FP32 accumulation on the matrix cores.
//
// The 3x3 convolution is an implicit GEMM.  With Cin % 32 == 0 it runs on
// v_mfma_f32_16x16x32_bf16 with the weights staged in LDS (conv3x3_k32
// below); otherwise (the 16-channel input layer) on v_mfma_f32_16x16x16_bf16: the
// M dimension is 16 pixels per wavefront (4 wavefronts = 64 pixels per
// workgroup), N is 16*NB output channels, K runs over the 9 taps times the
// input channels in steps of 16.  NHWC with channels padded to 16 makes every
// A and B fragment one 8-byte load per lane (4 consecutive BF16 channels).
// Operand layout of the 16x16x16 instruction (wave64):
//   A[i][k] from lane i + 16*(k/4), element k%4;
//   B[k][j] from lane j + 16*(k/4), element k%4;
//   D[i][j] in lane j + 16*(i/4), element i%4.
// The backward data pass is the same kernel with the taps mirrored and the
// transposed weights.  The weight gradient reduces over pixels (K = pixels),
// split into ranges whose FP32 partial tiles are summed in a second pass, so
// results do not depend on scheduling.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "coffee_hip.h"
#include "coffee_unet.h"
#include "ctx.hpp"

namespace coffee {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

__device__ __forceinline__ float bf2f(bf16 v) { return __uint_as_float((uint32_t)v << 16); }

__device__ __forceinline__ bf16 f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (bf16)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16)(u >> 16);
}

int launch_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(nullptr, COFFEE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return COFFEE_OK;
}

struct ConvArgs {
  const bf16* in;
  const bf16* w;
  const float* bias;
  bf16* out;
  float* out32;
  int N, H, W, Cin, Cout, flip, relu;
};

template <int NB>
__global__ void __launch_bounds__(256) conv3x3_mfma(const ConvArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const long P = (long)a.N * a.H * a.W;
  const long pix0 = ((long)blockIdx.x * 4 + wave) * 16;
  const int co0 = blockIdx.y * 16 * NB;
  const long pa = pix0 + li;
  int n = 0, y = 0, x = 0;
  const bool pvalid = pa < P;
  if (pvalid) {
    x = (int)(pa % a.W);
    const long t = pa / a.W;
    y = (int)(t % a.H);
    n = (int)(t / a.H);
  }
  f4 acc[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
  bool cvalid[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) cvalid[b] = co0 + b * 16 + li < a.Cout;
  for (int tap = 0; tap < 9; tap++) {
    int dy = tap / 3 - 1, dx = tap % 3 - 1;
    if (a.flip) {
      dy = -dy;
      dx = -dx;
    }
    const int yy = y + dy, xx = x + dx;
    const bool ok = pvalid && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
    const bf16* src = a.in + ((long)(n * a.H + (ok ? yy : 0)) * a.W + (ok ? xx : 0)) * a.Cin + 4 * lk;
    const bf16* wt = a.w + (long)tap * a.Cout * a.Cin + 4 * lk;
    for (int c0 = 0; c0 < a.Cin; c0 += 16) {
      s4 av = s4{0, 0, 0, 0};
      if (ok) av = *(const s4*)(src + c0);
#pragma unroll
      for (int b = 0; b < NB; b++) {
        s4 bv = s4{0, 0, 0, 0};
        if (cvalid[b]) bv = *(const s4*)(wt + (long)(co0 + b * 16 + li) * a.Cin + c0);
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[b], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const int co = co0 + b * 16 + li;
    if (!cvalid[b]) continue;
    const float bb = a.bias ? a.bias[co] : 0.f;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const long p = pix0 + 4 * lk + v;
      if (p >= P) continue;
      float val = acc[b][v] + bb;
      if (a.relu) val = val > 0.f ? val : 0.f;
      if (a.out) a.out[p * a.Cout + co] = f2bf(val);
      if (a.out32) a.out32[p * a.Cout + co] = val;
    }
  }
}

typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// The same convolution on v_mfma_f32_16x16x32_bf16 (twice the K of the
// 16x16x16 instruction per issue) for Cin % 32 == 0.  A workgroup of 4
// wavefronts computes 64*MB pixels x 16*NB output channels; K runs over the 9
// taps x Cin in steps of 32.  The weight slice of a step (16*NB channels x 32
// input channels) is staged in LDS, double-buffered and shared by the 4
// wavefronts; each wavefront's input fragments (16 B per lane) and the next
// weight slice are requested one step ahead.  Operand layout (wave64):
//   A[i][k] from lane i + 16*(k/8), element k%8;
//   B[k][j] from lane j + 16*(k/8), element k%8;
//   D[i][j] in lane j + 16*(i/4), element i%4.
template <int NB, int MB>
__global__ void __launch_bounds__(256) conv3x3_k32(const ConvArgs a) {
  constexpr int LDW = 40;  // LDS row: 32 input channels + 8 of padding (bank spread)
  __shared__ __attribute__((aligned(16))) bf16 wl[2][16 * NB * LDW];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const long P = (long)a.N * a.H * a.W;
  const long pixw = ((long)blockIdx.x * 4 + wave) * (16 * MB);
  const int co0 = blockIdx.y * 16 * NB;
  int pn[MB], py[MB], px[MB];
  bool pv[MB];
#pragma unroll
  for (int m = 0; m < MB; m++) {
    const long pa = pixw + m * 16 + li;
    pv[m] = pa < P;
    pn[m] = py[m] = px[m] = 0;
    if (pv[m]) {
      px[m] = (int)(pa % a.W);
      const long t = pa / a.W;
      py[m] = (int)(t % a.H);
      pn[m] = (int)(t / a.H);
    }
  }
  // weight staging: thread t loads 8 input channels (16 B) of row t/4
  const int wrow = threadIdx.x >> 2, wpart = threadIdx.x & 3;
  const bool wload = wrow < 16 * NB;
  const bool wvalid = wload && co0 + wrow < a.Cout;
  const int nci = a.Cin / 32, nk = 9 * nci;
  auto load_w = [&](int s) {
    s8 v = {0, 0, 0, 0, 0, 0, 0, 0};
    const int tap = s / nci, c0 = (s - tap * nci) * 32;
    if (wvalid) v = *(const s8*)(a.w + ((long)tap * a.Cout + co0 + wrow) * a.Cin + c0 + 8 * wpart);
    return v;
  };
  auto load_a = [&](int s, s8 (&av)[MB]) {
    const int tap = s / nci, c0 = (s - tap * nci) * 32;
    int dy = tap / 3 - 1, dx = tap % 3 - 1;
    if (a.flip) {
      dy = -dy;
      dx = -dx;
    }
#pragma unroll
    for (int m = 0; m < MB; m++) {
      const int yy = py[m] + dy, xx = px[m] + dx;
      av[m] = s8{0, 0, 0, 0, 0, 0, 0, 0};
      if (pv[m] && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
        av[m] = *(const s8*)(a.in + ((long)(pn[m] * a.H + yy) * a.W + xx) * a.Cin + c0 + 8 * lk);
    }
  };
  f4 acc[MB][NB];
#pragma unroll
  for (int m = 0; m < MB; m++)
#pragma unroll
    for (int b = 0; b < NB; b++) acc[m][b] = f4{0.f, 0.f, 0.f, 0.f};
  s8 acur[MB], anext[MB];
  s8 wnext = load_w(0);
  load_a(0, acur);
  if (wload) *(s8*)&wl[0][wrow * LDW + 8 * wpart] = wnext;
  __syncthreads();
  for (int s = 0; s < nk; s++) {
    if (s + 1 < nk) {
      wnext = load_w(s + 1);
      load_a(s + 1, anext);
    }
    const bf16* wb = wl[s & 1];
    bf8 bfr[NB];
#pragma unroll
    for (int b = 0; b < NB; b++)
      bfr[b] = __builtin_bit_cast(bf8, *(const s8*)&wb[(b * 16 + li) * LDW + 8 * lk]);
#pragma unroll
    for (int m = 0; m < MB; m++) {
      const bf8 af = __builtin_bit_cast(bf8, acur[m]);
#pragma unroll
      for (int b = 0; b < NB; b++)
        acc[m][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr[b], acc[m][b], 0, 0, 0);
    }
    if (s + 1 < nk) {
      // buffer (s + 1) & 1 was last read in step s - 1, before its barrier
      if (wload) *(s8*)&wl[(s + 1) & 1][wrow * LDW + 8 * wpart] = wnext;
#pragma unroll
      for (int m = 0; m < MB; m++) acur[m] = anext[m];
    }
    __syncthreads();
  }
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const int co = co0 + b * 16 + li;
    if (co >= a.Cout) continue;
    const float bb = a.bias ? a.bias[co] : 0.f;
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const long p = pixw + m * 16 + 4 * lk + v;
        if (p >= P) continue;
        float val = acc[m][b][v] + bb;
        if (a.relu) val = val > 0.f ? val : 0.f;
        if (a.out) a.out[p * a.Cout + co] = f2bf(val);
        if (a.out32) a.out32[p * a.Cout + co] = val;
      }
  }
}

// 16-byte chunk swizzle of the 64-byte (32-channel) LDS rows of conv3x3_tile:
// chunk c of row k lives at c ^ csw(k).  On gfx950 a ds_read_b128 whose 64
// lanes read chunk lane / 16 of rows base + lane % 16 then runs as fast as a
// contiguous 1 KiB read, for every base (measured); unswizzled 64-byte rows,
// rows padded to 80 bytes and the swizzles (k >> 2) & 3 or k & 3 took 1.35x
// as long, and the forward convolution counted a third of its LDS cycles as
// bank conflicts with the padded rows.
__device__ __forceinline__ int csw(int k) { return (k >> 1) & 3; }

// The same convolution with the input tile in LDS, for Cin % 32 == 0,
// Cout % 64 == 0 and W % 16 == 0.  A workgroup (4 wavefronts) computes a tile
// of TR image rows x 16 columns of one image x 64 output channels.  For each
// chunk of 32 input channels it stages the (TR + 2) x 18 input pixels (zero
// outside the image) and the 9 x 64 x 32 weights in LDS once, and all nine taps
// read their A fragments (one 16-pixel row segment, shifted by the tap) and B
// fragments from there, so every input value crosses the memory hierarchy
// once per chunk instead of once per tap.  The next chunk's global loads are
// issued before the current chunk's MFMAs, into registers; NBUF = 2
// double-buffers the LDS, NBUF = 1 stores after a barrier (half the LDS, so two
// workgroups share a CU).  LDS rows
// are the 32 channels (64 bytes) with their 16-byte chunks swizzled by csw,
// so the 16-byte fragment reads of consecutive lanes fall in distinct banks.
// Wavefront w owns rows
// w*TR/4 .. (w+1)*TR/4 - 1: TR/4 M fragments x 4 N fragments of
// v_mfma_f32_16x16x32_bf16.  The epilogue goes through LDS, so every output
// pixel's 64 channels are written as 16-byte stores.
template <int TR, int NBUF>
__global__ void __launch_bounds__(256) conv3x3_tile(const ConvArgs a) {
  constexpr int TW = 16, CP = 32, IH = TR + 2, IW = TW + 2, NIN = IH * IW;
  constexpr int MB = TR / 4, NB = 4;
  constexpr int IN_E = NIN * CP, W_E = 9 * 64 * CP, BUF_E = IN_E + W_E;
  constexpr int IN_LD = (NIN * 4 + 255) / 256;  // 16-byte input loads per thread
  constexpr int W_LD = 9;                       // 9 x 64 x 4 16-byte weight loads / 256
  constexpr int EP = 68;                        // epilogue row pitch (floats)
  constexpr int LDS_E = NBUF * BUF_E > TR * TW * EP * 2 ? NBUF * BUF_E : TR * TW * EP * 2;
  static_assert(TR % 4 == 0, "rows per wavefront");
  __shared__ __attribute__((aligned(16))) bf16 lds[LDS_E];

  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int tiles_x = a.W / TW, tiles_y = (a.H + TR - 1) / TR;
  const int tile = blockIdx.x;
  const int img = tile / (tiles_x * tiles_y);
  const int rest = tile - img * tiles_x * tiles_y;
  const int y0 = (rest / tiles_x) * TR, x0 = (rest % tiles_x) * TW;
  const int co0 = blockIdx.y * 64;
  const int nch = a.Cin / 32;

  s8 rin[IN_LD], rw[W_LD];
  auto load_chunk = [&](int ch) __attribute__((always_inline)) {
    const int c0 = ch * 32;
#pragma unroll
    for (int u = 0; u < IN_LD; u++) {
      const int idx = t + 256 * u;
      rin[u] = s8{0, 0, 0, 0, 0, 0, 0, 0};
      if (idx < NIN * 4) {
        const int pix = idx >> 2, part = idx & 3;
        const int yy = y0 - 1 + pix / IW, xx = x0 - 1 + pix % IW;
        if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
          rin[u] = *(const s8*)(a.in + ((long)(img * a.H + yy) * a.W + xx) * a.Cin + c0 + 8 * part);
      }
    }
#pragma unroll
    for (int u = 0; u < W_LD; u++) {
      const int idx = t + 256 * u;  // < 9 * 64 * 4
      const int row = idx >> 2, part = idx & 3;  // row = tap * 64 + co
      const int tap = row >> 6, co = row & 63;
      rw[u] = *(const s8*)(a.w + ((long)tap * a.Cout + co0 + co) * a.Cin + c0 + 8 * part);
    }
  };
  auto store_chunk = [&](bf16* buf) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < IN_LD; u++) {
      const int idx = t + 256 * u;
      if (idx < NIN * 4) *(s8*)&buf[(idx >> 2) * CP + 8 * ((idx & 3) ^ csw(idx >> 2))] = rin[u];
    }
#pragma unroll
    for (int u = 0; u < W_LD; u++) {
      const int idx = t + 256 * u;
      *(s8*)&buf[IN_E + (idx >> 2) * CP + 8 * ((idx & 3) ^ csw(idx >> 2))] = rw[u];
    }
  };

  f4 acc[MB][NB];
#pragma unroll
  for (int m = 0; m < MB; m++)
#pragma unroll
    for (int b = 0; b < NB; b++) acc[m][b] = f4{0.f, 0.f, 0.f, 0.f};

  load_chunk(0);
  store_chunk(lds);
  __syncthreads();
  for (int ch = 0; ch < nch; ch++) {
    const bf16* buf = lds + (NBUF == 2 ? (ch & 1) * BUF_E : 0);
    if (ch + 1 < nch) load_chunk(ch + 1);
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
      int dy = tap / 3 - 1, dx = tap % 3 - 1;
      if (a.flip) {
        dy = -dy;
        dx = -dx;
      }
      bf8 bfr[NB];
#pragma unroll
      for (int b = 0; b < NB; b++)  // csw(tap * 64 + b * 16 + li) = csw(li)
        bfr[b] = __builtin_bit_cast(
            bf8, *(const s8*)&buf[IN_E + (tap * 64 + b * 16 + li) * CP + 8 * (lk ^ csw(li))]);
#pragma unroll
      for (int m = 0; m < MB; m++) {
        const int r = wave * MB + m;
        const int row = (r + 1 + dy) * IW + (li + 1 + dx);
        const bf8 af = __builtin_bit_cast(bf8, *(const s8*)&buf[row * CP + 8 * (lk ^ csw(row))]);
#pragma unroll
        for (int b = 0; b < NB; b++)
          acc[m][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr[b], acc[m][b], 0, 0, 0);
      }
    }
    if (NBUF == 2) {
      // the other buffer was last read in the previous chunk, before its barrier
      if (ch + 1 < nch) store_chunk(lds + ((ch + 1) & 1) * BUF_E);
    } else if (ch + 1 < nch) {
      __syncthreads();  // every wavefront is done with this chunk
      store_chunk(lds);
    }
    __syncthreads();
  }

  // epilogue: accumulators -> LDS (pixel-major, 64 channels) -> 16-byte stores
  float* ep = (float*)lds;
#pragma unroll
  for (int m = 0; m < MB; m++)
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int v = 0; v < 4; v++)
        ep[((wave * MB + m) * TW + 4 * lk + v) * EP + b * 16 + li] = acc[m][b][v];
  __syncthreads();
  for (int idx = t; idx < TR * TW * 8; idx += 256) {
    const int px = idx >> 3, c8 = idx & 7;
    const int yy = y0 + px / TW, xx = x0 + px % TW;
    if (yy >= a.H) continue;
    const long p = (long)(img * a.H + yy) * a.W + xx;
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      float val = ep[px * EP + 8 * c8 + e] + (a.bias ? a.bias[co0 + 8 * c8 + e] : 0.f);
      if (a.relu) val = val > 0.f ? val : 0.f;
      v[e] = val;
    }
    if (a.out) {
      s8 o;
#pragma unroll
      for (int e = 0; e < 8; e++) o[e] = (short)f2bf(v[e]);
      *(s8*)(a.out + p * a.Cout + co0 + 8 * c8) = o;
    }
    if (a.out32) {
      f4* dst = (f4*)(a.out32 + p * a.Cout + co0 + 8 * c8);
      dst[0] = f4{v[0], v[1], v[2], v[3]};
      dst[1] = f4{v[4], v[5], v[6], v[7]};
    }
  }
}

// One wavefront per (pixel range, tap, 16x16 tile of (co, ci)).
__global__ void __launch_bounds__(64)
    wgrad_mfma(const bf16* __restrict__ x, const bf16* __restrict__ dz, float* __restrict__ part,
               int N, int H, int W, int Cin, int Cout, long chunk) {
  const int lane = threadIdx.x;
  const int li = lane & 15, lk = lane >> 4;
  const int split = blockIdx.x, tap = blockIdx.y;
  const int ncib = Cin / 16;
  const int cob = blockIdx.z / ncib, cib = blockIdx.z % ncib;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const long P = (long)N * H * W;
  const long p_begin = (long)split * chunk;
  const long p_end = p_begin + chunk < P ? p_begin + chunk : P;
  const int co = cob * 16 + li, ci = cib * 16 + li;
  f4 acc = f4{0.f, 0.f, 0.f, 0.f};
  for (long p0 = p_begin; p0 < p_end; p0 += 16) {
    short av[4], bv[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const long p = p0 + 4 * lk + v;
      av[v] = 0;
      bv[v] = 0;
      if (p < p_end) {
        av[v] = (short)dz[p * Cout + co];
        const int px = (int)(p % W);
        const long t = p / W;
        const int py = (int)(t % H);
        const int pn = (int)(t / H);
        const int yy = py + dy, xx = px + dx;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W)
          bv[v] = (short)x[((long)(pn * H + yy) * W + xx) * Cin + ci];
      }
    }
    const s4 a4 = s4{av[0], av[1], av[2], av[3]};
    const s4 b4 = s4{bv[0], bv[1], bv[2], bv[3]};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a4, b4, acc, 0, 0, 0);
  }
  // D[i = co][j = ci]: lane holds co = cob*16 + 4*lk + v, ci = cib*16 + li
  float* dst = part + ((long)split * 9 + tap) * Cout * Cin;
#pragma unroll
  for (int v = 0; v < 4; v++) dst[(long)(cob * 16 + 4 * lk + v) * Cin + cib * 16 + li] = acc[v];
}

// Weight gradient on v_mfma_f32_16x16x32_bf16: for one tap and one pixel
// range a workgroup (4 wavefronts) computes a 64 (co) x 64 (ci) tile of
// dW[tap][co][ci] = sum_p dz[p][co] x[p + tap][ci] (GEMM with K = pixels).
// Each step stages 32 pixels of dz and of the shifted input in LDS, transposed
// to channel-major rows so that a lane's 8 consecutive pixels (one MFMA
// fragment) are one 16-byte read; the global loads (16 B = 8 channels of one
// pixel per thread) for the next step are issued before the current step's
// MFMAs.  Wavefront w computes channels co0 + 16 w .. + 15 against the 64
// input channels (4 MFMAs per step).  Channels beyond Cout / Cin and pixels
// beyond the range or the image are zero.
__global__ void __launch_bounds__(256)
    wgrad_k32(const bf16* __restrict__ x, const bf16* __restrict__ dz, float* __restrict__ part,
              int N, int H, int W, int Cin, int Cout, long chunk) {
  constexpr int LDP = 40;  // LDS row: 32 pixels + 8 of padding (bank spread)
  __shared__ __attribute__((aligned(16))) bf16 dzt[2][64 * LDP];
  __shared__ __attribute__((aligned(16))) bf16 xt[2][64 * LDP];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int split = blockIdx.x, tap = blockIdx.y;
  const int ncib = (Cin + 63) / 64;
  const int co0 = (blockIdx.z / ncib) * 64, ci0 = (blockIdx.z % ncib) * 64;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const long P = (long)N * H * W;
  const long p_begin = (long)split * chunk;
  const long p_end = p_begin + chunk < P ? p_begin + chunk : P;
  // staging: thread t moves 8 channels (16 B) of pixel t % 32 of the step,
  // channels 8 (t / 32) ..; consecutive lanes write consecutive LDS columns
  const int sp = threadIdx.x & 31, sc = (threadIdx.x >> 5) * 8;
  const bool dz_ok = co0 + sc < Cout, x_ok = ci0 + sc < Cin;
  auto fetch = [&](long k0, s8& vdz, s8& vx) {
    vdz = s8{0, 0, 0, 0, 0, 0, 0, 0};
    vx = s8{0, 0, 0, 0, 0, 0, 0, 0};
    const long p = k0 + sp;
    if (p >= p_end) return;
    if (dz_ok) vdz = *(const s8*)(dz + p * Cout + co0 + sc);
    const int px = (int)(p % W);
    const long t = p / W;
    const int py = (int)(t % H), pn = (int)(t / H);
    const int yy = py + dy, xx = px + dx;
    if (x_ok && yy >= 0 && yy < H && xx >= 0 && xx < W)
      vx = *(const s8*)(x + ((long)(pn * H + yy) * W + xx) * Cin + ci0 + sc);
  };
  auto stage = [&](int buf, const s8& vdz, const s8& vx) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
      dzt[buf][(sc + e) * LDP + sp] = (bf16)vdz[e];
      xt[buf][(sc + e) * LDP + sp] = (bf16)vx[e];
    }
  };
  f4 acc[4];
#pragma unroll
  for (int b = 0; b < 4; b++) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
  s8 ndz, nx;
  fetch(p_begin, ndz, nx);
  stage(0, ndz, nx);
  __syncthreads();
  int buf = 0;
  for (long k0 = p_begin; k0 < p_end; k0 += 32) {
    const bool more = k0 + 32 < p_end;
    if (more) fetch(k0 + 32, ndz, nx);
    const bf8 af = __builtin_bit_cast(bf8, *(const s8*)&dzt[buf][(wave * 16 + li) * LDP + 8 * lk]);
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const bf8 bfr = __builtin_bit_cast(bf8, *(const s8*)&xt[buf][(b * 16 + li) * LDP + 8 * lk]);
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr, acc[b], 0, 0, 0);
    }
    // buffer buf ^ 1 was last read in the previous step, before its barrier
    if (more) stage(buf ^ 1, ndz, nx);
    __syncthreads();
    buf ^= 1;
  }
  // D[i = co][j = ci]: lane holds co = co0 + 16 wave + 4 lk + v, ci = ci0 + 16 b + li
  float* dst = part + ((long)split * 9 + tap) * Cout * Cin;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const int ci = ci0 + b * 16 + li;
    if (ci >= Cin) continue;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int co = co0 + wave * 16 + 4 * lk + v;
      if (co < Cout) dst[(long)co * Cin + ci] = acc[b][v];
    }
  }
}

// Transposing 64-bit LDS read of gfx950 (ds_read_b64_tr_b16).  Within each
// group of 16 lanes, lane 4 r + q names 4 consecutive 16-bit values (columns
// 4 q .. 4 q + 3 of "row" r); lane i of the group receives column i of rows
// 0 .. 3.  Two such reads give the 8 K values of a 16x16x32 MFMA operand
// from a matrix stored K-major (pixels as rows, channels contiguous).
typedef __attribute__((address_space(3))) s4 lds_s4;
__device__ __forceinline__ s4 lds_read_tr(const bf16* p) {
  return __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s4*)p);
}

// 16-byte chunk swizzle of the dz tile of wgrad_tr (pixel rows of 64
// channels = 8 chunks, 128 bytes): chunk c of pixel k lives at c ^ wswz(k),
// which puts the 32 lanes of every transposing read (pixel rows 8 g + 4 h + r
// of a 32-pixel step) on distinct banks with no padding.  The bits it uses
// (1 and 3) come only from the lane's own part of the pixel index.
__device__ __forceinline__ int wswz(int k) { return (k & 2) | ((k >> 1) & 4); }

// Weight gradient with spatial tiles in LDS, for Cin % 64 == 0, Cout % 64 == 0
// and W % 16 == 0: a workgroup computes the 9 x 64 (co) x 64 (ci) block of
// dW[tap][co][ci] = sum_p dz[p][co] x[p + tap][ci] over a range of tiles of
// TR image rows x 16 columns.  dz (TR x 16 pixels x 64 co) and the input
// with its one-pixel halo are staged in LDS as they lie in memory (pixel
// rows, channels contiguous; 16-byte stores), and both MFMA operands are read
// with the transposing LDS read, so a tap's shift is only another row
// address (a constant offset from the lane's base) and every staged value
// serves all nine taps.  Input rows are padded to XS = 72 values.  Wavefront
// w owns input channels ci0 + 16 w .. + 15 against the 64 output channels:
// per 32-pixel step it reads 4 A fragments (shared by the taps) and 9 B
// fragments for 36 MFMAs (v_mfma_f32_16x16x32_bf16).  The next tile's
// global loads are in flight during the current tile's MFMAs.  Pixels beyond
// the image are zero.  Partial blocks per range go to
// part[range][tap][co][ci] (summed by wgrad_reduce in range order); when
// bpart is given, the workgroups of the first input-channel block also sum dz
// over their range, per output channel in a fixed order, into
// bpart[range][co].
template <int TR>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
    wgrad_tr(const bf16* __restrict__ x, const bf16* __restrict__ dz, float* __restrict__ part,
             float* __restrict__ bpart, int N, int H, int W, int Cin, int Cout,
             int tiles_per_range, int ntiles) {
  constexpr int TW = 16, TP = TR * TW, XW = TW + 2, XP = (TR + 2) * XW, XS = 72;
  constexpr int NDZ = TP * 8 / 256;            // 16-byte dz loads per thread
  constexpr int NX = (XP * 8 + 255) / 256;     // 16-byte input loads per thread
  static_assert(TP % 32 == 0, "whole 32-pixel steps per tile");
  static_assert(32 * 64 <= XP * XS, "bias reduction fits in the input tile");
  __shared__ __attribute__((aligned(16))) bf16 dzs[TP * 64];
  __shared__ __attribute__((aligned(16))) bf16 xs[XP * XS];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int g = lane >> 4, r = (lane >> 2) & 3, q = lane & 3;
  const int range = blockIdx.x;
  const int ncib = Cin / 64;
  const int co0 = (blockIdx.y / ncib) * 64, ci0 = (blockIdx.y % ncib) * 64;
  const bool do_bias = bpart && blockIdx.y % ncib == 0;
  const int tiles_x = W / TW, tiles_y = (H + TR - 1) / TR;
  const int t_begin = range * tiles_per_range;
  const int t_end = min(t_begin + tiles_per_range, ntiles);

  s8 rdz[NDZ], rx[NX];
  auto load_tile = [&](int tile) __attribute__((always_inline)) {
    const int img = tile / (tiles_x * tiles_y);
    const int rest = tile - img * tiles_x * tiles_y;
    const int y0 = (rest / tiles_x) * TR, x0 = (rest % tiles_x) * TW;
#pragma unroll
    for (int u = 0; u < NDZ; u++) {
      const int idx = t + 256 * u, px = idx >> 3, c8 = idx & 7;
      const int yy = y0 + px / TW, xx = x0 + px % TW;
      rdz[u] = s8{0, 0, 0, 0, 0, 0, 0, 0};
      if (yy < H) rdz[u] = *(const s8*)(dz + ((long)(img * H + yy) * W + xx) * Cout + co0 + 8 * c8);
    }
#pragma unroll
    for (int u = 0; u < NX; u++) {
      const int idx = t + 256 * u;
      rx[u] = s8{0, 0, 0, 0, 0, 0, 0, 0};
      if (idx < XP * 8) {
        const int pix = idx >> 3, c8 = idx & 7;
        const int yy = y0 - 1 + pix / XW, xx = x0 - 1 + pix % XW;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W)
          rx[u] = *(const s8*)(x + ((long)(img * H + yy) * W + xx) * Cin + ci0 + 8 * c8);
      }
    }
  };
  auto store_tile = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < NDZ; u++) {
      const int idx = t + 256 * u, px = idx >> 3, c8 = idx & 7;
      *(s8*)&dzs[px * 64 + 8 * (c8 ^ wswz(px))] = rdz[u];
    }
#pragma unroll
    for (int u = 0; u < NX; u++) {
      const int idx = t + 256 * u;
      if (idx < XP * 8) *(s8*)&xs[(idx >> 3) * XS + 8 * (idx & 7)] = rx[u];
    }
  };
  // A: K index 8 g + 4 h + r of step st is pixel 32 st + 16 (g / 2) + 8 (g % 2)
  // + 4 h + r; the lane's part ka fixes the swizzle, st and h add 64-value rows.
  const int ka = 16 * (g >> 1) + 8 * (g & 1) + r;
  const bf16* abase[4];
#pragma unroll
  for (int m = 0; m < 4; m++) abase[m] = &dzs[ka * 64 + 8 * ((2 * m + (q >> 1)) ^ wswz(ka)) + 4 * (q & 1)];
  // B: the same pixel shifted by the tap, in the haloed input tile
  const bf16* xbase = &xs[(((g >> 1) + 1) * XW + 8 * (g & 1) + r + 1) * XS + 8 * (2 * wave + (q >> 1)) +
                          4 * (q & 1)];

  f4 acc[9][4];
#pragma unroll
  for (int k = 0; k < 9; k++)
#pragma unroll
    for (int m = 0; m < 4; m++) acc[k][m] = f4{0.f, 0.f, 0.f, 0.f};
  float bsum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (t_begin < t_end) load_tile(t_begin);
  for (int tile = t_begin; tile < t_end; tile++) {
    __syncthreads();  // every wavefront is done with the previous tile
    store_tile();
    __syncthreads();
    if (tile + 1 < t_end) load_tile(tile + 1);
    if (do_bias) {
      // thread (pixel group t / 8, chunk t % 8): pixels t / 8, t / 8 + 32, ...
#pragma unroll
      for (int i = 0; i < TP / 32; i++) {
        const int px = (t >> 3) + 32 * i, c8 = t & 7;
        const s8 v = *(const s8*)&dzs[px * 64 + 8 * (c8 ^ wswz(px))];
#pragma unroll
        for (int e = 0; e < 8; e++) bsum[e] += bf2f((bf16)v[e]);
      }
    }
#pragma unroll 1
    for (int st = 0; st < TP / 32; st++) {
      bf8 af[4];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const s4 lo = lds_read_tr(abase[m] + (32 * st) * 64);
        const s4 hi = lds_read_tr(abase[m] + (32 * st + 4) * 64);
        af[m] = __builtin_bit_cast(bf8, s8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]});
      }
#pragma unroll
      for (int tap = 0; tap < 9; tap++) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const int off = ((2 * st + dy) * XW + dx) * XS;
        const s4 lo = lds_read_tr(xbase + off);
        const s4 hi = lds_read_tr(xbase + off + 4 * XS);
        const bf8 bfr = __builtin_bit_cast(bf8, s8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]});
#pragma unroll
        for (int m = 0; m < 4; m++)
          acc[tap][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], bfr, acc[tap][m], 0, 0, 0);
      }
    }
  }
  // D[i = co][j = ci]: lane holds co = co0 + 16 m + 4 lk + v, ci = ci0 + 16 wave + li
  const int ci = ci0 + wave * 16 + li;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    float* dst = part + ((long)range * 9 + k) * Cout * Cin;
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int v = 0; v < 4; v++) dst[(long)(co0 + m * 16 + 4 * lk + v) * Cin + ci] = acc[k][m][v];
  }
  if (do_bias) {
    // the 32 pixel groups' sums of each channel, added in group order
    __syncthreads();  // every wavefront is done with the last tile
    float* red = (float*)xs;  // 32 x 64 floats, inside the input tile
#pragma unroll
    for (int e = 0; e < 8; e++) red[(t >> 3) * 64 + 8 * (t & 7) + e] = bsum[e];
    __syncthreads();
    if (t < 64) {
      float b = 0.f;
      for (int k = 0; k < 32; k++) b += red[k * 64 + t];
      bpart[(long)range * Cout + co0 + t] = b;
    }
  }
}

// Sum of m values v(k), k = 0 .. m-1, as 8 interleaved partial sums (k mod 8)
// combined pairwise: a fixed order (deterministic results) whose loads do
// not wait on each other.
template <typename F>
__device__ __forceinline__ float sum8(int m, F&& v) {
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int k = 0;
  for (; k + 8 <= m; k += 8)
#pragma unroll
    for (int u = 0; u < 8; u++) s[u] += v(k + u);
  for (int u = 0; k + u < m; u++) s[u] += v(k + u);
  return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// dW[i] = sum over ranges of part[range][i]; with bpart, also db[c] = sum
// over ranges of bpart[range][c] (threads n .. n + nb - 1), in range order.
__global__ void wgrad_reduce(const float* __restrict__ part, float* __restrict__ dW, long n,
                             int nsplit, const float* __restrict__ bpart, float* __restrict__ db,
                             int nb) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n + nb; i += (long)gridDim.x * blockDim.x) {
    if (i < n) dW[i] = sum8(nsplit, [&](int k) { return part[(long)k * n + i]; });
    else db[i - n] = sum8(nsplit, [&](int k) { return bpart[(long)k * nb + (i - n)]; });
  }
}

// Bias gradient db[c] = sum_p dz[p][c] in two deterministic passes.  Pass 1:
// block b sums the pixels [b chunk, (b + 1) chunk) of every channel, threads
// laid out channel-fastest so that a pixel's channels are one coalesced row
// (channel groups of up to 256; 256 / group pixel rows in parallel, summed in
// a fixed order).  Pass 2 adds the block partials in block order.
__global__ void __launch_bounds__(256) bias_grad_part(const bf16* __restrict__ dz,
                                                      float* __restrict__ part, long P, int C,
                                                      long chunk) {
  __shared__ float red[256];
  const long p0 = (long)blockIdx.x * chunk;
  const long p1 = p0 + chunk < P ? p0 + chunk : P;
  for (int cb = 0; cb < C; cb += 256) {
    const int cw = C - cb < 256 ? C - cb : 256;
    const int rows = 256 / cw;
    const int c = (int)threadIdx.x % cw, r = (int)threadIdx.x / cw;
    float s = 0.f;
    if (r < rows)
      for (long p = p0 + r; p < p1; p += rows) s += bf2f(dz[p * C + cb + c]);
    red[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < cw) {
      float t = 0.f;
      for (int k = 0; k < rows; k++) t += red[k * cw + threadIdx.x];
      part[(long)blockIdx.x * C + cb + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

// Pass 2: 4 partial rows x 64 channels per block, each thread adding every
// 4th block partial (sum8), then the 4 row sums in row order.
__global__ void __launch_bounds__(256) bias_grad_sum(const float* __restrict__ part,
                                                     float* __restrict__ db, int nb, int C) {
  __shared__ float red[256];
  const int cl = (int)threadIdx.x & 63, r = (int)threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + cl;
  float s = 0.f;
  if (c < C) {
    const int m = nb > r ? (nb - r + 3) / 4 : 0;  // partials r, r + 4, ...
    s = sum8(m, [&](int k) { return part[(long)(r + 4 * k) * C + c]; });
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (r == 0 && c < C) db[c] = ((red[cl] + red[64 + cl]) + red[128 + cl]) + red[192 + cl];
}

__global__ void combine_kernel(bf16* __restrict__ out, int ost, int ooff, const bf16* __restrict__ a,
                               int ast, int aoff, const bf16* __restrict__ b, int bst, int boff,
                               const bf16* __restrict__ mask, int mst, int moff, int nc, long npix) {
  const long total = npix * nc;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long p = i / nc;
    const int c = (int)(i % nc);
    float v = bf2f(a[p * ast + aoff + c]);
    if (b) v += bf2f(b[p * bst + boff + c]);
    if (mask && !(bf2f(mask[p * mst + moff + c]) > 0.f)) v = 0.f;
    out[p * ost + ooff + c] = f2bf(v);
  }
}

// h, w: sizes of the fine grid.
__global__ void pool2_kernel(const bf16* __restrict__ in, bf16* __restrict__ out, int N, int h, int w,
                             int C) {
  const int h2 = h / 2, w2 = w / 2;
  const long total = (long)N * h2 * w2 * C;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long t = i / C;
    const int x = (int)(t % w2);
    t /= w2;
    const int y = (int)(t % h2);
    const int n = (int)(t / h2);
    const long b = ((long)(n * h + 2 * y) * w + 2 * x) * C + c;
    const float s = (bf2f(in[b]) + bf2f(in[b + C])) + (bf2f(in[b + (long)w * C]) + bf2f(in[b + (long)w * C + C]));
    out[i] = f2bf(0.25f * s);
  }
}

__global__ void pool2_bwd_kernel(const bf16* __restrict__ dy, bf16* __restrict__ dx, int N, int h,
                                 int w, int C) {
  const long total = (long)N * h * w * C;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long t = i / C;
    const int x = (int)(t % w);
    t /= w;
    const int y = (int)(t % h);
    const int n = (int)(t / h);
    dx[i] = f2bf(0.25f * bf2f(dy[((long)(n * (h / 2) + y / 2) * (w / 2) + x / 2) * C + c]));
  }
}

__global__ void up2_kernel(const bf16* __restrict__ in, bf16* __restrict__ out, int N, int h, int w,
                           int C) {
  const long total = (long)N * h * w * C;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long t = i / C;
    const int x = (int)(t % w);
    t /= w;
    const int y = (int)(t % h);
    const int n = (int)(t / h);
    out[i] = in[((long)(n * (h / 2) + y / 2) * (w / 2) + x / 2) * C + c];
  }
}

__global__ void up2_bwd_kernel(const bf16* __restrict__ dy, bf16* __restrict__ dx, int N, int h,
                               int w, int C) {
  const int h2 = h / 2, w2 = w / 2;
  const long total = (long)N * h2 * w2 * C;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long t = i / C;
    const int x = (int)(t % w2);
    t /= w2;
    const int y = (int)(t % h2);
    const int n = (int)(t / h2);
    const long b = ((long)(n * h + 2 * y) * w + 2 * x) * C + c;
    const float s = (bf2f(dy[b]) + bf2f(dy[b + C])) + (bf2f(dy[b + (long)w * C]) + bf2f(dy[b + (long)w * C + C]));
    dx[i] = f2bf(s);
  }
}

__global__ void __launch_bounds__(256)
    mse_kernel(const float* __restrict__ out32, int cst, const float* __restrict__ target, int nc,
               long npix, bf16* __restrict__ dout, double* __restrict__ part) {
  __shared__ double red[256];
  const double scale = 2.0 / ((double)npix * nc);
  double s = 0.0;
  const long total = npix * cst;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long p = i / cst;
    const int c = (int)(i % cst);
    float g = 0.f;
    if (c < nc) {
      const double d = (double)out32[i] - (double)target[p * nc + c];
      s += d * d;
      g = (float)(scale * d);
    }
    dout[i] = f2bf(g);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int m = 128; m > 0; m >>= 1) {
    if ((int)threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

__global__ void weights_bf16_kernel(const float* __restrict__ w32, bf16* __restrict__ wb,
                                    bf16* __restrict__ wbt, int Cout, int Cin) {
  const long total = 9L * Cout * Cin;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int ci = (int)(i % Cin);
    const long t = i / Cin;
    const int co = (int)(t % Cout);
    const int tap = (int)(t / Cout);
    const bf16 v = f2bf(w32[i]);
    wb[i] = v;
    wbt[((long)tap * Cin + ci) * Cout + co] = v;
  }
}

__global__ void to_bf16_kernel(const float* __restrict__ in, bf16* __restrict__ out, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = f2bf(in[i]);
}

__global__ void from_bf16_kernel(const bf16* __restrict__ in, float* __restrict__ out, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = bf2f(in[i]);
}

template <typename T>
__global__ void pack_kernel(const T* f0, const T* f1, const T* f2, const T* f3, const T* f4_,
                            const T* f5, int dims0, int g0, int g1, int n0, int n1, double x0,
                            double dx, double dy, double rpow, double extra, void* out, int out32,
                            int cst, long base) {
  const long total = (long)n0 * n1 * cst;
  const T* F[6] = {f0, f1, f2, f3, f4_, f5};
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int c = (int)(t % cst);
    const long pix = t / cst;
    const int i = (int)(pix % n0);
    const int j = (int)(pix / n0);
    double v = 0.0;
    if (c < 6) {
      const double r = exp(x0 + dx * i);
      const double th = dy * j;
      const double h = (c % 3 == 2) ? r * sin(th) : r;
      v = (double)F[c][(long)(i + g0) + (long)(j + g1) * dims0] * h * pow(r, rpow);
    } else if (c == 6) {
      v = extra;
    }
    const long o = base + t;
    if (out32) ((float*)out)[o] = (float)v;
    else ((bf16*)out)[o] = f2bf((float)v);
  }
}

template <typename T>
__global__ void unpack_kernel(const float* in, int cst, long base, T* f0, T* f1, T* f2, T* f3,
                              T* f4_, T* f5, int dims0, int g0, int g1, int n0, int n1, double x0,
                              double dx, double dy, double rpow) {
  const long total = (long)n0 * n1 * 6;
  T* F[6] = {f0, f1, f2, f3, f4_, f5};
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int c = (int)(t % 6);
    const long pix = t / 6;
    const int i = (int)(pix % n0);
    const int j = (int)(pix / n0);
    const double r = exp(x0 + dx * i);
    const double th = dy * j;
    const double h = (c % 3 == 2) ? r * sin(th) : r;
    const double den = h * pow(r, rpow);
    const double v = fabs(den) > 1e-12 ? (double)in[base + pix * cst + c] / den : 0.0;
    F[c][(long)(i + g0) + (long)(j + g1) * dims0] = (T)v;
  }
}

// Corotation: E = -(Omega x r) x B with v = Omega r sin(theta) (phi direction),
// in upper components of x = ln r, theta: E^x = v B^theta, E^theta = -v B^x,
// E^phi = 0 (boundary_pulsar_sph applies the same relation inside the star).
template <typename T>
__global__ void corotate_kernel(T* ex, T* ey, T* ez, const T* bx, const T* by, int dims0, int g0,
                                int g1, int n0, int n1, double x0, double dx, double dy,
                                double omega) {
  const long total = (long)n0 * n1;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % n0), j = (int)(t / n0);
    const long c = (long)(i + g0) + (long)(j + g1) * dims0;
    const double v = omega * exp(x0 + dx * i) * sin(dy * j);
    ex[c] = (T)(v * (double)by[c]);
    ey[c] = (T)(-v * (double)bx[c]);
    ez[c] = T(0);
  }
}

int grid_for(long total, int block = 256) {
  long b = (total + block - 1) / block;
  return (int)(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

}  // namespace
}  // namespace coffee

using namespace coffee;

extern "C" {

int coffee_unet_conv3x3(const void* in, const void* w, const float* bias, void* out, float* out32,
                        int n, int h, int wd, int cin, int cout, int flip, int relu, void* stream) {
  if (!in || !w || (!out && !out32) || n < 1 || h < 1 || wd < 1 || cin < 16 || cout < 16 ||
      cin % 16 || cout % 16)
    return fail(nullptr, COFFEE_ERR_INVALID, "conv3x3: bad arguments (channels must be multiples of 16)");
  ConvArgs a{(const bf16*)in, (const bf16*)w, bias, (bf16*)out, out32, n, h, wd, cin, cout, flip, relu};
  hipStream_t s = (hipStream_t)stream;
  const long P = (long)n * h * wd;
  const long bx = (P + 63) / 64;
  if (bx > 2147483647L) return fail(nullptr, COFFEE_ERR_SIZE, "conv3x3: too many pixels");
  const bool aligned16 = ((uintptr_t)in % 16 == 0) && ((uintptr_t)w % 16 == 0);
  const bool out_aligned = ((uintptr_t)out % 16 == 0) && ((uintptr_t)out32 % 16 == 0);
  if (cin % 32 == 0 && cout % 64 == 0 && wd % 16 == 0 && aligned16 && out_aligned) {
    // LDS-tiled kernel: 16-row tiles where the image has them, else 8-row
    // single-buffered LDS (72 / 60 KB), two workgroups per CU; 16-row tiles
    // where they still give two workgroups per CU, else 8-row tiles (measured
    // at the U-Net shapes: 8x256x128 64->64 44 us, 8x128x64 128->128 32 us,
    // 8x64x32 256->256 31 us; the double-buffered variant, one workgroup per
    // CU, was 25-35 % slower on the first two)
    const long cb = cout / 64;
    int tr = 16;
    if (h < 16 || (long)n * (wd / 16) * ((h + 15) / 16) * cb < 512) tr = 8;
    const long tiles = (long)n * (wd / 16) * ((h + tr - 1) / tr);
    if (tiles > 2147483647L) return fail(nullptr, COFFEE_ERR_SIZE, "conv3x3: too many tiles");
    const dim3 g((unsigned)tiles, (unsigned)cb);
    if (tr == 16) conv3x3_tile<16, 1><<<g, 256, 0, s>>>(a);
    else conv3x3_tile<8, 1><<<g, 256, 0, s>>>(a);
  } else if (cin % 32 == 0 && aligned16) {
    // K = 32 kernel; two 16-pixel row blocks per wavefront when that still
    // leaves two workgroups per CU
    const int nb = cout % 64 == 0 ? 4 : (cout % 32 == 0 ? 2 : 1);
    const long b2 = (P + 127) / 128, cb = cout / (16 * nb);
    const bool mb2 = b2 * cb >= 512;
    const unsigned gx = (unsigned)(mb2 ? b2 : bx);
    const dim3 g(gx, (unsigned)cb);
#define COFFEE_CONV32(NBV)                                              \
  do {                                                                  \
    if (mb2) conv3x3_k32<NBV, 2><<<g, 256, 0, s>>>(a);                 \
    else conv3x3_k32<NBV, 1><<<g, 256, 0, s>>>(a);                     \
  } while (0)
    if (nb == 4) COFFEE_CONV32(4);
    else if (nb == 2) COFFEE_CONV32(2);
    else COFFEE_CONV32(1);
#undef COFFEE_CONV32
  } else if (cout % 64 == 0) {
    conv3x3_mfma<4><<<dim3((unsigned)bx, cout / 64), 256, 0, s>>>(a);
  } else if (cout % 32 == 0) {
    conv3x3_mfma<2><<<dim3((unsigned)bx, cout / 32), 256, 0, s>>>(a);
  } else {
    conv3x3_mfma<1><<<dim3((unsigned)bx, cout / 16), 256, 0, s>>>(a);
  }
  return launch_check("conv3x3");
}

int coffee_unet_conv3x3_wgrad(const void* x, const void* dz, float* partial, int nsplit, float* dW,
                              float* db, int n, int h, int wd, int cin, int cout, void* stream) {
  if (!x || !dz || !partial || !dW || nsplit < 1 || cin % 16 || cout % 16 || cin < 16 || cout < 16)
    return fail(nullptr, COFFEE_ERR_INVALID, "conv3x3_wgrad: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const long P = (long)n * h * wd;
  long chunk = (P + nsplit - 1) / nsplit;
  chunk = (chunk + 15) / 16 * 16;
  int used = (int)((P + chunk - 1) / chunk);
  const bool al16 = (uintptr_t)x % 16 == 0 && (uintptr_t)dz % 16 == 0;
  bool fused_bias = false;
  if (al16 && cin % 64 == 0 && cout % 64 == 0 && wd % 16 == 0) {
    // spatial tiles of 8 rows (4 where the image is shorter), split over at
    // most nsplit ranges of whole tiles
    const int tr = h >= 8 ? 8 : 4;
    const long ntiles = (long)n * (wd / 16) * ((h + tr - 1) / tr);
    if (ntiles > 2147483647L) return fail(nullptr, COFFEE_ERR_SIZE, "conv3x3_wgrad: too many tiles");
    // with a bias gradient, the ranges leave room after their weight partials
    // for one row of cout bias partials each: at most nsplit - 1 ranges
    fused_bias = db && nsplit >= 2 && nsplit - 1 <= 9 * cin;
    const long nr = fused_bias ? nsplit - 1 : nsplit;
    const int per = (int)((ntiles + nr - 1) / nr);
    used = (int)((ntiles + per - 1) / per);
    const dim3 g((unsigned)used, (unsigned)((cout / 64) * (cin / 64)));
    float* bpart = fused_bias ? partial + (long)used * 9 * cout * cin : nullptr;
    if (tr == 8)
      wgrad_tr<8><<<g, 256, 0, s>>>((const bf16*)x, (const bf16*)dz, partial, bpart, n, h, wd, cin,
                                     cout, per, (int)ntiles);
    else
      wgrad_tr<4><<<g, 256, 0, s>>>((const bf16*)x, (const bf16*)dz, partial, bpart, n, h, wd, cin,
                                     cout, per, (int)ntiles);
  } else if (al16)
    wgrad_k32<<<dim3(used, 9, ((cout + 63) / 64) * ((cin + 63) / 64)), 256, 0, s>>>(
        (const bf16*)x, (const bf16*)dz, partial, n, h, wd, cin, cout, chunk);
  else
    wgrad_mfma<<<dim3(used, 9, (cout / 16) * (cin / 16)), 64, 0, s>>>(
        (const bf16*)x, (const bf16*)dz, partial, n, h, wd, cin, cout, chunk);
  int rc = launch_check("conv3x3_wgrad");
  if (rc) return rc;
  const long nw = 9L * cout * cin;
  // with the bias partials of wgrad_tr, one pass reduces both
  const int nb = fused_bias ? cout : 0;
  wgrad_reduce<<<grid_for(nw + nb), 256, 0, s>>>(partial, dW, nw, used, partial + nw * used, db, nb);
  rc = launch_check("wgrad_reduce");
  if (rc || !db || fused_bias) return rc;
  // the weight partials have been reduced: their buffer holds the bias
  // partials, one row of cout per block (at most nsplit x 9 x cin rows); a
  // block covers 64 pixels per thread row
  const long rows = 256 / (cout < 256 ? cout : 256);
  const long cap = (long)nsplit * 9 * cin;
  long bchunk = rows * 64;
  long nbb = (P + bchunk - 1) / bchunk;
  if (nbb > cap) {
    bchunk = (P + cap - 1) / cap;
    nbb = (P + bchunk - 1) / bchunk;
  }
  bias_grad_part<<<(unsigned)nbb, 256, 0, s>>>((const bf16*)dz, partial, P, cout, bchunk);
  rc = launch_check("bias_grad_part");
  if (rc) return rc;
  bias_grad_sum<<<(cout + 63) / 64, 256, 0, s>>>(partial, db, (int)nbb, cout);
  return launch_check("bias_grad_sum");
}

int coffee_unet_combine(void* out, int ost, int ooff, const void* a, int ast, int aoff,
                        const void* b, int bst, int boff, const void* mask, int mst, int moff,
                        int nc, long npix, void* stream) {
  if (!out || !a || nc < 1 || npix < 1 || ooff + nc > ost || aoff + nc > ast ||
      (b && boff + nc > bst) || (mask && moff + nc > mst))
    return fail(nullptr, COFFEE_ERR_INVALID, "combine: bad arguments");
  combine_kernel<<<grid_for(npix * nc), 256, 0, (hipStream_t)stream>>>(
      (bf16*)out, ost, ooff, (const bf16*)a, ast, aoff, (const bf16*)b, bst, boff,
      (const bf16*)mask, mst, moff, nc, npix);
  return launch_check("combine");
}

int coffee_unet_pool2(const void* in, void* out, int n, int h, int wd, int c, void* stream) {
  if (!in || !out || h % 2 || wd % 2) return fail(nullptr, COFFEE_ERR_INVALID, "pool2: even sizes required");
  pool2_kernel<<<grid_for((long)n * h * wd * c / 4), 256, 0, (hipStream_t)stream>>>(
      (const bf16*)in, (bf16*)out, n, h, wd, c);
  return launch_check("pool2");
}

int coffee_unet_pool2_bwd(const void* dy, void* dx, int n, int h, int wd, int c, void* stream) {
  if (!dy || !dx || h % 2 || wd % 2) return fail(nullptr, COFFEE_ERR_INVALID, "pool2_bwd: even sizes required");
  pool2_bwd_kernel<<<grid_for((long)n * h * wd * c), 256, 0, (hipStream_t)stream>>>(
      (const bf16*)dy, (bf16*)dx, n, h, wd, c);
  return launch_check("pool2_bwd");
}

int coffee_unet_up2(const void* in, void* out, int n, int h, int wd, int c, void* stream) {
  if (!in || !out || h % 2 || wd % 2) return fail(nullptr, COFFEE_ERR_INVALID, "up2: even sizes required");
  up2_kernel<<<grid_for((long)n * h * wd * c), 256, 0, (hipStream_t)stream>>>(
      (const bf16*)in, (bf16*)out, n, h, wd, c);
  return launch_check("up2");
}

int coffee_unet_up2_bwd(const void* dy, void* dx, int n, int h, int wd, int c, void* stream) {
  if (!dy || !dx || h % 2 || wd % 2) return fail(nullptr, COFFEE_ERR_INVALID, "up2_bwd: even sizes required");
  up2_bwd_kernel<<<grid_for((long)n * h * wd * c / 4), 256, 0, (hipStream_t)stream>>>(
      (const bf16*)dy, (bf16*)dx, n, h, wd, c);
  return launch_check("up2_bwd");
}

int coffee_unet_mse(const float* out32, int cst, const float* target, int nc, long npix,
                    void* dout, double* loss, void* stream) {
  if (!out32 || !target || !dout || !loss || nc < 1 || nc > cst)
    return fail(nullptr, COFFEE_ERR_INVALID, "mse: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const int blocks = 1024;
  double* part = nullptr;
  if (hipMalloc(&part, sizeof(double) * blocks) != hipSuccess)
    return fail(nullptr, COFFEE_ERR_HIP, "mse: hipMalloc failed");
  mse_kernel<<<blocks, 256, 0, s>>>(out32, cst, target, nc, npix, (bf16*)dout, part);
  hipError_t e = hipGetLastError();
  double h[1024];
  if (e == hipSuccess) e = hipMemcpyAsync(h, part, sizeof(double) * blocks, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(part);
  if (e != hipSuccess) return fail(nullptr, COFFEE_ERR_HIP, std::string("mse: ") + hipGetErrorString(e));
  double tot = 0.0;
  for (int b = 0; b < blocks; b++) tot += h[b];
  *loss = tot / ((double)npix * nc);
  return COFFEE_OK;
}

int coffee_unet_adam(float* p, const float* g, float* m, float* v, long n, float lr, float beta1,
                     float beta2, float eps, int t, void* wb, void* wbt, int cout, int cin,
                     void* stream) {
  if (!p || !g || !m || !v || n < 1 || t < 1) return fail(nullptr, COFFEE_ERR_INVALID, "adam: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
  adam_kernel<<<grid_for(n), 256, 0, s>>>(p, g, m, v, n, lr, beta1, beta2, eps, c1, c2);
  int rc = launch_check("adam");
  if (rc || !wb) return rc;
  if (!wbt || 9L * cout * cin != n) return fail(nullptr, COFFEE_ERR_INVALID, "adam: weight shape mismatch");
  weights_bf16_kernel<<<grid_for(n), 256, 0, s>>>(p, (bf16*)wb, (bf16*)wbt, cout, cin);
  return launch_check("weights_bf16");
}

int coffee_unet_weights_bf16(const float* w32, void* wb, void* wbt, int cout, int cin, void* stream) {
  if (!w32 || !wb || !wbt) return fail(nullptr, COFFEE_ERR_INVALID, "weights_bf16: null argument");
  weights_bf16_kernel<<<grid_for(9L * cout * cin), 256, 0, (hipStream_t)stream>>>(
      w32, (bf16*)wb, (bf16*)wbt, cout, cin);
  return launch_check("weights_bf16");
}

int coffee_unet_to_bf16(const float* in, void* out, long n, void* stream) {
  if (!in || !out) return fail(nullptr, COFFEE_ERR_INVALID, "to_bf16: null argument");
  to_bf16_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(in, (bf16*)out, n);
  return launch_check("to_bf16");
}

int coffee_unet_from_bf16(const void* in, float* out, long n, void* stream) {
  if (!in || !out) return fail(nullptr, COFFEE_ERR_INVALID, "from_bf16: null argument");
  from_bf16_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>((const bf16*)in, out, n);
  return launch_check("from_bf16");
}

int coffee_unet_pack(void* const fields[6], int scalar_bytes, int dims0, int g0, int g1, int n0,
                     int n1, double x0, double dx, double dy, double rpow, double extra_channel,
                     void* out, int out_fp32, int cst, int image, void* stream) {
  if (!fields || !out || cst < 6 || n0 < 1 || n1 < 1 || image < 0)
    return fail(nullptr, COFFEE_ERR_INVALID, "pack: bad arguments");
  for (int q = 0; q < 6; q++)
    if (!fields[q]) return fail(nullptr, COFFEE_ERR_INVALID, "pack: null field");
  const long base = (long)image * n0 * n1 * cst;
  const int blocks = grid_for((long)n0 * n1 * cst);
  hipStream_t s = (hipStream_t)stream;
  if (scalar_bytes == 4) {
    float* const* f = (float* const*)fields;
    pack_kernel<float><<<blocks, 256, 0, s>>>(f[0], f[1], f[2], f[3], f[4], f[5], dims0, g0, g1, n0,
                                              n1, x0, dx, dy, rpow, extra_channel, out, out_fp32,
                                              cst, base);
  } else {
    double* const* f = (double* const*)fields;
    pack_kernel<double><<<blocks, 256, 0, s>>>(f[0], f[1], f[2], f[3], f[4], f[5], dims0, g0, g1,
                                               n0, n1, x0, dx, dy, rpow, extra_channel, out,
                                               out_fp32, cst, base);
  }
  return launch_check("pack");
}

int coffee_unet_unpack(const float* in, int cst, int image, void* const fields[6], int scalar_bytes,
                       int dims0, int g0, int g1, int n0, int n1, double x0, double dx, double dy,
                       double rpow, void* stream) {
  if (!in || !fields || cst < 6 || n0 < 1 || n1 < 1 || image < 0)
    return fail(nullptr, COFFEE_ERR_INVALID, "unpack: bad arguments");
  for (int q = 0; q < 6; q++)
    if (!fields[q]) return fail(nullptr, COFFEE_ERR_INVALID, "unpack: null field");
  const long base = (long)image * n0 * n1 * cst;
  const int blocks = grid_for((long)n0 * n1 * 6);
  hipStream_t s = (hipStream_t)stream;
  if (scalar_bytes == 4) {
    float* const* f = (float* const*)fields;
    unpack_kernel<float><<<blocks, 256, 0, s>>>(in, cst, base, f[0], f[1], f[2], f[3], f[4], f[5],
                                                dims0, g0, g1, n0, n1, x0, dx, dy, rpow);
  } else {
    double* const* f = (double* const*)fields;
    unpack_kernel<double><<<blocks, 256, 0, s>>>(in, cst, base, f[0], f[1], f[2], f[3], f[4], f[5],
                                                 dims0, g0, g1, n0, n1, x0, dx, dy, rpow);
  }
  return launch_check("unpack");
}

int coffee_unet_corotate(void* const fields[6], int scalar_bytes, int dims0, int g0, int g1,
                         int n0, int n1, double x0, double dx, double dy, double omega,
                         void* stream) {
  if (!fields || n0 < 1 || n1 < 1 || dims0 < n0 + 2 * g0 || g0 < 0 || g1 < 0)
    return fail(nullptr, COFFEE_ERR_INVALID, "corotate: bad arguments");
  for (int q = 0; q < 5; q++)
    if (!fields[q]) return fail(nullptr, COFFEE_ERR_INVALID, "corotate: null field");
  const int blocks = grid_for((long)n0 * n1);
  hipStream_t s = (hipStream_t)stream;
  if (scalar_bytes == 4) {
    float* const* f = (float* const*)fields;
    corotate_kernel<float><<<blocks, 256, 0, s>>>(f[0], f[1], f[2], f[3], f[4], dims0, g0, g1, n0,
                                                  n1, x0, dx, dy, omega);
  } else if (scalar_bytes == 8) {
    double* const* f = (double* const*)fields;
    corotate_kernel<double><<<blocks, 256, 0, s>>>(f[0], f[1], f[2], f[3], f[4], dims0, g0, g1,
                                                   n0, n1, x0, dx, dy, omega);
  } else {
    return fail(nullptr, COFFEE_ERR_INVALID, "corotate: scalar_bytes must be 4 or 8");
  }
  return launch_check("corotate");
}

}  // extern "C"
