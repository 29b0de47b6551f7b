// xt_unet.hip -- the learned warm start: a 3D U-Net (4 levels, channels
// 16/32/64/128, 3x3x3 convs) trained and run by this library's own kernels.
//
// Activations are channels-last bf16 [V][C] over V = d0*d1*d2 voxels (d2
// fastest), so the 8 consecutive channels one MFMA lane feeds are one 16-byte
// load.  Every 3x3x3 conv is an implicit GEMM on v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation:
//   forward        out[v][co] = sum_{tap,ci} in[v+off(tap)][ci] * W[co][tap][ci]
//   backward-data  gin[v][ci] = sum_{tap,co} gout[v+off(tap)][co] * W[co][26-tap][ci]
//                  (the same kernel with the flipped, transposed weights)
//   weight grad    dW[co][tap*Ci+ci] = sum_v gout[v][co] * in[v+off(tap)][ci]
//                  (split over voxel chunks, FP32 atomics)
// FP32 master weights with Adam; bf16 copies of both weight layouts are
// regenerated after every update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "xt_unet_internal.h"

typedef uint16_t bf16_t;  // raw bfloat16 bits
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

#define UN_HIP(call)                                                                         \
  do {                                                                                       \
    hipError_t _e = (call);                                                                  \
    if (_e != hipSuccess)                                                                    \
      return xt_report_error(XT_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                            \
  } while (0)
#define UN_TRY(expr)               \
  do {                             \
    int _rc = (expr);              \
    if (_rc != XT_OK) return _rc;  \
  } while (0)

// float -> bf16, round to nearest even (finite inputs)
__device__ __host__ inline bf16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ inline float bf2f(bf16_t b) { return __uint_as_float((uint32_t)b << 16); }
// ReLU mask of a stored activation: strictly positive
__device__ inline bool bf_pos(bf16_t b) { return b != 0 && !(b & 0x8000u); }

__device__ inline bf16x8 as_bf16x8(uint4 q) { return __builtin_bit_cast(bf16x8, q); }

// ---------------------------------------------------------------------------
// 3x3x3 conv, forward (also backward-data).  Workgroup = 4 waves; a wave
// computes 16*MB voxels x 16*NB output channels.  K runs over kk = tap*Ci + ci
// padded to Kpad (multiple of 32) with zero weights.  VEC: Ci % 8 == 0 (one
// 16-byte load per lane and fragment), else element gathers (the 1-channel
// stem).  Optional epilogue: + bias, ReLU, x (mask > 0).
template <bool VEC, int MB, int NB>
__global__ void __launch_bounds__(256) k_conv3(const bf16_t *__restrict__ in, int Ci,
                                               const bf16_t *__restrict__ W, int Kpad,
                                               const float *__restrict__ bias, int Co, int d0,
                                               int d1, int d2, bf16_t *__restrict__ out, int relu,
                                               const bf16_t *__restrict__ mask) {
  const int V = d0 * d1 * d2, P = d1 * d2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int m0 = (blockIdx.x * 4 + wave) * (16 * MB);
  const int n0 = blockIdx.y * (16 * NB);
  if (m0 >= V) return;
  int c0[MB], c1[MB], c2[MB];
#pragma unroll
  for (int mb = 0; mb < MB; mb++) {
    const int v = m0 + mb * 16 + r;
    c0[mb] = v < V ? v / P : -100;  // out-of-range rows never pass the bounds test
    c1[mb] = (v / d2) % d1;
    c2[mb] = v % d2;
  }
  f32x4 acc[MB][NB];
#pragma unroll
  for (int mb = 0; mb < MB; mb++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int Ktot = 27 * Ci;
  const uint4 zero4 = {0u, 0u, 0u, 0u};
  int tap = (g * 8) / Ci, ci = g * 8 - tap * Ci;  // (tap, ci) of kk = k0 + g*8
  for (int k0 = 0; k0 < Kpad; k0 += 32) {
    bf16x8 a[MB], b[NB];
    if constexpr (VEC) {
      const int a0 = tap / 9 - 1, a1 = (tap / 3) % 3 - 1, a2 = tap % 3 - 1;
      const bool tok = tap < 27;
      const int off = a0 * P + a1 * d2 + a2;
#pragma unroll
      for (int mb = 0; mb < MB; mb++) {
        const int p0 = c0[mb] + a0, p1 = c1[mb] + a1, p2 = c2[mb] + a2;
        const bool ok = tok && (unsigned)p0 < (unsigned)d0 && (unsigned)p1 < (unsigned)d1 &&
                        (unsigned)p2 < (unsigned)d2;
        const int v = m0 + mb * 16 + r;
        a[mb] = as_bf16x8(ok ? *(const uint4 *)(in + (size_t)(v + off) * Ci + ci) : zero4);
      }
      ci += 32;
      while (ci >= Ci) {
        ci -= Ci;
        tap++;
      }
    } else {
      const int kk = k0 + g * 8;
#pragma unroll
      for (int mb = 0; mb < MB; mb++) {
        const int v = m0 + mb * 16 + r;
        uint16_t e[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const int kj = kk + j, t = kj / Ci, cj = kj - t * Ci;
          const int a0 = t / 9 - 1, a1 = (t / 3) % 3 - 1, a2 = t % 3 - 1;
          const int p0 = c0[mb] + a0, p1 = c1[mb] + a1, p2 = c2[mb] + a2;
          const bool ok = kj < Ktot && (unsigned)p0 < (unsigned)d0 && (unsigned)p1 < (unsigned)d1 &&
                          (unsigned)p2 < (unsigned)d2;
          e[j] = ok ? in[(size_t)(v + a0 * P + a1 * d2 + a2) * Ci + cj] : (uint16_t)0;
        }
        uint4 q;
        q.x = e[0] | ((uint32_t)e[1] << 16);
        q.y = e[2] | ((uint32_t)e[3] << 16);
        q.z = e[4] | ((uint32_t)e[5] << 16);
        q.w = e[6] | ((uint32_t)e[7] << 16);
        a[mb] = as_bf16x8(q);
      }
    }
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      const int co = n0 + nb * 16 + r;
      b[nb] = as_bf16x8(co < Co ? *(const uint4 *)(W + (size_t)co * Kpad + k0 + g * 8) : zero4);
    }
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
      for (int nb = 0; nb < NB; nb++)
        acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb], b[nb], acc[mb][nb], 0, 0, 0);
  }
  // accumulator lane layout: column (output channel) r, rows (voxels) g*4 + i
#pragma unroll
  for (int nb = 0; nb < NB; nb++) {
    const int co = n0 + nb * 16 + r;
    if (co >= Co) continue;
    const float bv = bias ? bias[co] : 0.f;
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int v = m0 + mb * 16 + g * 4 + i;
        if (v >= V) continue;
        float val = acc[mb][nb][i] + bv;
        if (relu) val = fmaxf(val, 0.f);
        if (mask && !bf_pos(mask[(size_t)v * Co + co])) val = 0.f;
        out[(size_t)v * Co + co] = f2bf(val);
      }
  }
}

// LDS-tiled variant for Ci % 16 == 0: a workgroup's 64*MB voxels form a TY x TX
// tile of one d0-plane (TX along d2); per chunk of CC input channels the tile
// and its one-voxel halo in d0, d1 and d2 are staged in LDS once and every
// tap's A fragment is read from there (the direct kernel reads each input 27
// times from L2).  K order: channel chunks outer, taps inner (CC = 32: one tap
// per MFMA K-step; CC = 16: two).  TY | d1 and TX | d2.
template <int MB, int NB, int CC>
__global__ void __launch_bounds__(256) k_conv3_lds(const bf16_t *__restrict__ in, int Ci,
                                                   const bf16_t *__restrict__ W, int Kpad,
                                                   const float *__restrict__ bias, int Co, int d0,
                                                   int d1, int d2, int TY, int TX,
                                                   bf16_t *__restrict__ out, int relu,
                                                   const bf16_t *__restrict__ mask) {
  constexpr int CP = CC + 8;  // padded channel stride of a staged voxel (bf16)
  constexpr int TPS = 32 / CC;
  extern __shared__ uint4 lds4[];
  bf16_t *tile = (bf16_t *)lds4;  // [3][TY+2][TX+2][CP]
  const int P = d1 * d2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int n0 = blockIdx.y * (16 * NB);
  const int ntx = d2 / TX, nty = d1 / TY;
  int bi = blockIdx.x;
  const int bx = bi % ntx;
  bi /= ntx;
  const int by = bi % nty;
  const int p0 = bi / nty;
  const int y0 = by * TY, x0 = bx * TX;
  const int HY = TY + 2, HX = TX + 2, HV = 3 * HY * HX;
  int ty[MB], tx[MB];
  bool rok[MB];
#pragma unroll
  for (int mb = 0; mb < MB; mb++) {
    const int m = (wave * MB + mb) * 16 + r;
    ty[mb] = m / TX;
    tx[mb] = m % TX;
    rok[mb] = ty[mb] < TY;
  }
  f32x4 acc[MB][NB];
#pragma unroll
  for (int mb = 0; mb < MB; mb++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint4 zero4 = {0u, 0u, 0u, 0u};
  for (int c0 = 0; c0 < Ci; c0 += CC) {
    __syncthreads();  // the previous chunk's fragments are read
    for (int q = threadIdx.x; q < HV * (CC / 8); q += 256) {
      const int hv = q / (CC / 8), c8 = q % (CC / 8);
      const int hz = hv / (HY * HX), rem = hv % (HY * HX);
      const int pp = p0 + hz - 1, yy = y0 + rem / HX - 1, xx = x0 + rem % HX - 1;
      uint4 v = zero4;
      if ((unsigned)pp < (unsigned)d0 && (unsigned)yy < (unsigned)d1 && (unsigned)xx < (unsigned)d2)
        v = *(const uint4 *)(in + ((size_t)pp * P + (size_t)yy * d2 + xx) * Ci + c0 + 8 * c8);
      *(uint4 *)(tile + (size_t)hv * CP + 8 * c8) = v;
    }
    __syncthreads();
    for (int t0 = 0; t0 < 27; t0 += TPS) {
      const int tap = t0 + (TPS == 2 ? (g >> 1) : 0);
      const int cg = (TPS == 2 ? (g & 1) : g) * 8;
      const bool tok = tap < 27;
      const int a0 = tap / 9, a1 = (tap / 3) % 3, a2 = tap % 3;  // halo offsets 0..2
      bf16x8 a[MB], b[NB];
#pragma unroll
      for (int mb = 0; mb < MB; mb++) {
        const int hv = (a0 * HY + ty[mb] + a1) * HX + tx[mb] + a2;
        a[mb] = as_bf16x8((tok && rok[mb]) ? *(const uint4 *)(tile + (size_t)hv * CP + cg) : zero4);
      }
#pragma unroll
      for (int nb = 0; nb < NB; nb++) {
        const int co = n0 + nb * 16 + r;
        b[nb] = as_bf16x8((tok && co < Co)
                              ? *(const uint4 *)(W + (size_t)co * Kpad + (size_t)tap * Ci + c0 + cg)
                              : zero4);
      }
#pragma unroll
      for (int mb = 0; mb < MB; mb++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb], b[nb], acc[mb][nb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int nb = 0; nb < NB; nb++) {
    const int co = n0 + nb * 16 + r;
    if (co >= Co) continue;
    const float bv = bias ? bias[co] : 0.f;
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = (wave * MB + mb) * 16 + g * 4 + i;
        const int yy = m / TX, xx = m % TX;
        if (yy >= TY) continue;
        const size_t v = ((size_t)p0 * d1 + y0 + yy) * d2 + x0 + xx;
        float val = acc[mb][nb][i] + bv;
        if (relu) val = fmaxf(val, 0.f);
        if (mask && !bf_pos(mask[v * Co + co])) val = 0.f;
        out[v * Co + co] = f2bf(val);
      }
  }
}

// Weight gradient: M = output channels (16*MBo per wave), N = kk columns (16*NBk
// per wave), K = voxels in steps of 32.  Each workgroup reduces one chunk of
// voxels (split over its 4 waves) and adds its partial tile to dW with FP32
// atomics.
template <int MBo, int NBk>
__global__ void __launch_bounds__(256) k_conv3_wgrad(const bf16_t *__restrict__ in, int Ci,
                                                     const bf16_t *__restrict__ gout, int Co, int d0,
                                                     int d1, int d2, int vchunk,
                                                     float *__restrict__ dW) {
  const int V = d0 * d1 * d2, P = d1 * d2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int co0 = blockIdx.y * (16 * MBo), kk0 = blockIdx.z * (16 * NBk);
  const int Ktot = 27 * Ci;
  const int wchunk = vchunk / 4;
  const int vbeg = blockIdx.x * vchunk + wave * wchunk;
  const int vend = min(vbeg + wchunk, V);
  if (vbeg >= vend) return;
  int ka0[NBk], ka1[NBk], ka2[NBk], kci[NBk];
  bool kok[NBk];
#pragma unroll
  for (int nb = 0; nb < NBk; nb++) {
    const int kk = kk0 + nb * 16 + r;
    kok[nb] = kk < Ktot;
    const int t = kok[nb] ? kk / Ci : 13;
    kci[nb] = kok[nb] ? kk - t * Ci : 0;
    ka0[nb] = t / 9 - 1;
    ka1[nb] = (t / 3) % 3 - 1;
    ka2[nb] = t % 3 - 1;
  }
  f32x4 acc[MBo][NBk];
#pragma unroll
  for (int mb = 0; mb < MBo; mb++)
#pragma unroll
    for (int nb = 0; nb < NBk; nb++) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int v0 = vbeg; v0 < vend; v0 += 32) {
    // this lane's 8 voxels: v0 + g*8 + j
    int v = v0 + g * 8;
    int p0 = v / P, p1 = (v / d2) % d1, p2 = v % d2;
    uint16_t ea[MBo][8], eb[NBk][8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const bool vok = v < vend;
#pragma unroll
      for (int mb = 0; mb < MBo; mb++) {
        const int co = co0 + mb * 16 + r;
        ea[mb][j] = (vok && co < Co) ? gout[(size_t)v * Co + co] : (uint16_t)0;
      }
#pragma unroll
      for (int nb = 0; nb < NBk; nb++) {
        const int q0 = p0 + ka0[nb], q1 = p1 + ka1[nb], q2 = p2 + ka2[nb];
        const bool ok = vok && kok[nb] && (unsigned)q0 < (unsigned)d0 &&
                        (unsigned)q1 < (unsigned)d1 && (unsigned)q2 < (unsigned)d2;
        eb[nb][j] = ok ? in[(size_t)(v + ka0[nb] * P + ka1[nb] * d2 + ka2[nb]) * Ci + kci[nb]]
                       : (uint16_t)0;
      }
      v++;
      if (++p2 == d2) {
        p2 = 0;
        if (++p1 == d1) {
          p1 = 0;
          p0++;
        }
      }
    }
    bf16x8 a[MBo], b[NBk];
#pragma unroll
    for (int mb = 0; mb < MBo; mb++) {
      uint4 q;
      q.x = ea[mb][0] | ((uint32_t)ea[mb][1] << 16);
      q.y = ea[mb][2] | ((uint32_t)ea[mb][3] << 16);
      q.z = ea[mb][4] | ((uint32_t)ea[mb][5] << 16);
      q.w = ea[mb][6] | ((uint32_t)ea[mb][7] << 16);
      a[mb] = as_bf16x8(q);
    }
#pragma unroll
    for (int nb = 0; nb < NBk; nb++) {
      uint4 q;
      q.x = eb[nb][0] | ((uint32_t)eb[nb][1] << 16);
      q.y = eb[nb][2] | ((uint32_t)eb[nb][3] << 16);
      q.z = eb[nb][4] | ((uint32_t)eb[nb][5] << 16);
      q.w = eb[nb][6] | ((uint32_t)eb[nb][7] << 16);
      b[nb] = as_bf16x8(q);
    }
#pragma unroll
    for (int mb = 0; mb < MBo; mb++)
#pragma unroll
      for (int nb = 0; nb < NBk; nb++)
        acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb], b[nb], acc[mb][nb], 0, 0, 0);
  }
  // lane holds column kk0+nb*16+r and rows co0+mb*16+g*4+i
#pragma unroll
  for (int nb = 0; nb < NBk; nb++) {
    if (!kok[nb]) continue;
    const int kk = kk0 + nb * 16 + r;
#pragma unroll
    for (int mb = 0; mb < MBo; mb++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int co = co0 + mb * 16 + g * 4 + i;
        if (co < Co) atomicAdd(dW + (size_t)co * Ktot + kk, acc[mb][nb][i]);
      }
  }
}

// db[c] += sum_v gout[v][c]; 256 % C == 0.
__global__ void __launch_bounds__(256) k_bias_grad(const bf16_t *__restrict__ gout, int C, int V,
                                                   int vchunk, float *__restrict__ db) {
  __shared__ float red[256];
  const int t = threadIdx.x, c = t % C, grp = t / C, ng = 256 / C;
  const int vbeg = blockIdx.x * vchunk, vend = min(vbeg + vchunk, V);
  float s = 0.f;
  for (int v = vbeg + grp; v < vend; v += ng) s += bf2f(gout[(size_t)v * C + c]);
  red[t] = s;
  __syncthreads();
  if (grp == 0) {
    for (int k = 1; k < ng; k++) s += red[k * C + c];
    atomicAdd(db + c, s);
  }
}

// 2x2x2 average pool: fine dims (d0, d1, d2) -> coarse (d0/2, d1/2, d2/2); one
// thread per coarse voxel and 8 channels (C % 8 == 0), 16-byte loads and stores
__global__ void k_pool(const bf16_t *__restrict__ in, int C, int d0, int d1, int d2,
                       bf16_t *__restrict__ out) {
  const int e1 = d1 / 2, e2 = d2 / 2, C8 = C / 8;
  const size_t n = (size_t)(d0 / 2) * e1 * e2 * C8;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c8 = (int)(i % C8);
    const int w = (int)(i / C8);  // coarse voxel (V < 2^31)
    const int q2 = w % e2, q1 = (w / e2) % e1, q0 = w / (e1 * e2);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const size_t v = ((size_t)(2 * q0 + (k >> 2)) * d1 + (2 * q1 + ((k >> 1) & 1))) * d2 +
                       (2 * q2 + (k & 1));
      const uint4 q = *(const uint4 *)(in + v * C + 8 * c8);
      const uint32_t u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        s[2 * j] += bf2f((bf16_t)(u[j] & 0xFFFFu));
        s[2 * j + 1] += bf2f((bf16_t)(u[j] >> 16));
      }
    }
    uint4 o;
    uint32_t *ou = (uint32_t *)&o;
#pragma unroll
    for (int j = 0; j < 4; j++)
      ou[j] = (uint32_t)f2bf(s[2 * j] * 0.125f) | ((uint32_t)f2bf(s[2 * j + 1] * 0.125f) << 16);
    *(uint4 *)(out + (size_t)w * C + 8 * c8) = o;
  }
}

// out[v] = [ low[parent(v)] (Ca channels) | skip[v] (Cb channels) ] on fine dims;
// one thread per 8 channels (Ca, Cb % 8 == 0), 16-byte copies
__global__ void k_upcat(const bf16_t *__restrict__ low, int Ca, const bf16_t *__restrict__ skip,
                        int Cb, int d0, int d1, int d2, bf16_t *__restrict__ out) {
  const int Ct8 = (Ca + Cb) / 8, Ca8 = Ca / 8;
  const size_t n = (size_t)d0 * d1 * d2 * Ct8;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c8 = (int)(i % Ct8);
    const int v = (int)(i / Ct8);  // V < 2^31
    uint4 q;
    if (c8 < Ca8) {
      const int p2 = v % d2, p1 = (v / d2) % d1, p0 = v / (d1 * d2);
      const size_t u = ((size_t)(p0 / 2) * (d1 / 2) + p1 / 2) * (d2 / 2) + p2 / 2;
      q = *(const uint4 *)(low + u * Ca + 8 * c8);
    } else {
      q = *(const uint4 *)(skip + (size_t)v * Cb + 8 * (c8 - Ca8));
    }
    *(uint4 *)(out + i * 8) = q;
  }
}

// glow[u][c] = sum over the 8 children v of u of gcat[v][c], c < Ca (coarse dims)
__global__ void k_up_bwd(const bf16_t *__restrict__ gcat, int Ca, int Ct, int e0, int e1, int e2,
                         bf16_t *__restrict__ glow) {
  const size_t n = (size_t)e0 * e1 * e2 * Ca;
  const int d1 = 2 * e1, d2 = 2 * e2;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % Ca);
    const size_t w = i / Ca;
    const int q2 = (int)(w % e2), q1 = (int)((w / e2) % e1), q0 = (int)(w / ((size_t)e1 * e2));
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const size_t v = ((size_t)(2 * q0 + (k >> 2)) * d1 + (2 * q1 + ((k >> 1) & 1))) * d2 +
                       (2 * q2 + (k & 1));
      s += bf2f(gcat[v * Ct + c]);
    }
    glow[i] = f2bf(s);
  }
}

// ge[v][c] = gcat[v][Ca+c] + [e[v][c] > 0] * gpool[parent(v)][c] / 8 (fine dims);
// gcat's skip part is already masked by the concatenated activation.
__global__ void k_skip_bwd(const bf16_t *__restrict__ gcat, int Ca, int Ct,
                           const bf16_t *__restrict__ gpool, const bf16_t *__restrict__ e, int Cb,
                           int d0, int d1, int d2, bf16_t *__restrict__ ge) {
  const size_t n = (size_t)d0 * d1 * d2 * Cb;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % Cb);
    const size_t v = i / Cb;
    const int p2 = (int)(v % d2), p1 = (int)((v / d2) % d1), p0 = (int)(v / ((size_t)d1 * d2));
    const size_t u = ((size_t)(p0 / 2) * (d1 / 2) + p1 / 2) * (d2 / 2) + p2 / 2;
    float s = bf2f(gcat[v * Ct + Ca + c]);
    if (bf_pos(e[i])) s += bf2f(gpool[u * Cb + c]) * 0.125f;
    ge[i] = f2bf(s);
  }
}

// network input: t[v] = bf16(x[v] * scale)
__global__ void k_prep(const float *__restrict__ x, float scale, size_t n, bf16_t *__restrict__ t) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    t[i] = f2bf(x[i] * scale);
}

// head: out[v] = (x[v]*scale + sum_c d[v][c]*wh[c] + bh) / scale
__global__ void k_head_fwd(const bf16_t *__restrict__ d, int C, const float *__restrict__ wh,
                           const float *__restrict__ bh, const float *x, float scale, size_t n,
                           float *out) {
  // C == 16: two 16-byte loads per voxel
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (size_t)gridDim.x * blockDim.x) {
    const uint4 q0 = *(const uint4 *)(d + v * C), q1 = *(const uint4 *)(d + v * C + 8);
    const uint32_t u[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
    float h = bh[0];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      h += bf2f((bf16_t)(u[j] & 0xFFFFu)) * wh[2 * j];
      h += bf2f((bf16_t)(u[j] >> 16)) * wh[2 * j + 1];
    }
    out[v] = (x[v] * scale + h) / scale;
  }
}

// head backward and the loss mean((x*s + h - t*s)^2):
// g = 2 res / n; loss += res^2 / n; dwh[c] += g d[v][c]; dbh += g;
// gd[v][c] = bf16(g wh[c]) [d[v][c] > 0].  Block size 256, C <= 16.
__global__ void __launch_bounds__(256) k_head_bwd(const bf16_t *__restrict__ d, int C,
                                                  const float *__restrict__ wh,
                                                  const float *__restrict__ bh,
                                                  const float *__restrict__ x,
                                                  const float *__restrict__ tgt, float scale,
                                                  size_t n, double *loss, float *dwh, float *dbh,
                                                  bf16_t *__restrict__ gd) {
  __shared__ float red[17][256];
  __shared__ double redl[256];
  const int t = threadIdx.x;
  float sw[16] = {0.f}, sb = 0.f;
  double sl = 0.0;
  const float inv_n = 1.f / (float)n;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + t; v < n; v += (size_t)gridDim.x * blockDim.x) {
    float h = bh[0];
    float dv[16];
    for (int c = 0; c < C; c++) {
      dv[c] = bf2f(d[v * C + c]);
      h += dv[c] * wh[c];
    }
    const float res = x[v] * scale + h - tgt[v] * scale;
    const float gv = 2.f * res * inv_n;
    sl += (double)res * res;
    sb += gv;
    for (int c = 0; c < C; c++) {
      sw[c] += gv * dv[c];
      gd[v * C + c] = dv[c] > 0.f ? f2bf(gv * wh[c]) : (bf16_t)0;
    }
  }
  for (int c = 0; c < C; c++) red[c][t] = sw[c];
  red[16][t] = sb;
  redl[t] = sl;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      for (int c = 0; c < C; c++) red[c][t] += red[c][t + s];
      red[16][t] += red[16][t + s];
      redl[t] += redl[t + s];
    }
    __syncthreads();
  }
  if (t < C) atomicAdd(dwh + t, red[t][0]);
  if (t == 0) {
    atomicAdd(dbh, red[16][0]);
    atomicAdd(loss, redl[0] / (double)n);
  }
}

// bf16 weight layouts from the FP32 master [Co][27*Ci]:
//   forward  Wf[co][kk], kk = tap*Ci + ci < Kpad (zero beyond 27*Ci)
//   backward Wb[ci][kk'], kk' = tap'*Co + co < Kpadb, Wb = W[co][(26-tap')*Ci + ci]
__global__ void k_wprep(const float *__restrict__ w, int Ci, int Co, int Kpad, int Kpadb,
                        bf16_t *__restrict__ Wf, bf16_t *__restrict__ Wb) {
  const size_t nf = (size_t)Co * Kpad, nb = (size_t)Ci * Kpadb;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf + nb;
       i += (size_t)gridDim.x * blockDim.x) {
    if (i < nf) {
      const int co = (int)(i / Kpad), kk = (int)(i % Kpad);
      Wf[i] = kk < 27 * Ci ? f2bf(w[(size_t)co * 27 * Ci + kk]) : (bf16_t)0;
    } else {
      const size_t j = i - nf;
      const int ci = (int)(j / Kpadb), kk = (int)(j % Kpadb);
      bf16_t val = 0;
      if (kk < 27 * Co) {
        const int tp = kk / Co, co = kk % Co;
        val = f2bf(w[(size_t)co * 27 * Ci + (26 - tp) * Ci + ci]);
      }
      Wb[j] = val;
    }
  }
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  }
}

__global__ void k_fill(float *__restrict__ p, size_t n, float val) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = val;
}

__global__ void k_fill_d(double *p, double val) { *p = val; }

// ---------------------------------------------------------------------------
// host side
namespace {

constexpr int NCONV = 14;
// (level, C_in, C_out) of each 3x3x3 conv with base channels 16
const int kConvSpec[NCONV][3] = {{0, 1, 16},   {0, 16, 16},  {1, 16, 32}, {1, 32, 32},
                                 {2, 32, 64},  {2, 64, 64},  {3, 64, 128}, {3, 128, 128},
                                 {2, 192, 64}, {2, 64, 64},  {1, 96, 32}, {1, 32, 32},
                                 {0, 48, 16},  {0, 16, 16}};
constexpr int HEAD_C = 16;

struct Conv {
  int lvl, Ci, Co, Kpad, Kpadb;
  size_t w_off, b_off;  // into the parameter vector
  bf16_t *Wf = nullptr, *Wb = nullptr;
};

int grid1(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 8192); }

}  // namespace

struct xt_unet {
  int dev = 0;
  hipStream_t st = nullptr;  // the stream of the synchronous entry points
  int dims[4][3];
  size_t V[4];
  float scale = 1.f;
  bool train = false;
  Conv conv[NCONV];
  size_t head_w = 0, head_b = 0;
  int64_t nparam = 0;
  float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  int64_t adam_t = 0;
  bool wdirty = true;
  bf16_t *wbuf = nullptr;
  // activations
  bf16_t *t_in, *e1a, *e1, *p1, *e2a, *e2, *p2, *e3a, *e3, *p3, *ba, *bb, *u3, *d3a, *d3, *u2,
      *d2a, *d2, *u1, *d1a, *d1;
  // activation gradients (train)
  bf16_t *g_d1, *g_d1a, *g_u1, *g_d2, *g_d2a, *g_u2, *g_d3, *g_d3a, *g_u3, *g_bb, *g_ba, *g_p3,
      *g_e3, *g_e3a, *g_p2, *g_e2, *g_e2a, *g_p1, *g_e1, *g_e1a;
  bf16_t *abuf = nullptr, *gbuf = nullptr;
  double *loss_dev = nullptr;
  float *stage = nullptr;  // [2][V0] host-call staging (input, target/output)
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double fwd_ms = 0, bwd_ms = 0;
  bool lds_conv = true;  // LDS-tiled convolutions where Ci % 16 == 0 and Ci >= 32
};

namespace {

// carve `count` bf16 elements (16-byte aligned) from a running offset
struct Carver {
  size_t off = 0;
  std::vector<std::pair<bf16_t **, size_t>> items;
  void add(bf16_t **p, size_t count) {
    items.push_back({p, off});
    off += (count + 127) / 128 * 128;
  }
  void assign(bf16_t *base) {
    for (auto &it : items) *it.first = base + it.second;
  }
};

int pow2_divisor(int n, int cap) {  // largest power of two dividing n, at most cap
  int p = 1;
  while (p * 2 <= cap && n % (p * 2) == 0) p *= 2;
  return p;
}

void launch_conv(xt_unet *u, hipStream_t st, const bf16_t *in, int Ci, const bf16_t *W, int Kpad,
                 const float *bias, int Co, const int *d, bf16_t *out, int relu,
                 const bf16_t *mask) {
  const int V = d[0] * d[1] * d[2];
  const int NB = (Co % 64 == 0) ? 4 : (Co % 32 == 0) ? 2 : 1;
  // LDS tiles pay from 32 input channels on: with 16 the 4x halo staging costs
  // more than the L2 re-reads it saves (16->16 at 512x512x32: 785 against 547 us)
  if (Ci % 16 == 0 && Ci >= 32 && u->lds_conv) {
    const int CC = (Ci % 32 == 0) ? 32 : 16;
    const int MBl = (NB == 1 && CC == 16) ? 4 : 2;
    const int TX = pow2_divisor(d[2], std::min(32, 64 * MBl));
    const int TY = pow2_divisor(d[1], 64 * MBl / TX);
    const size_t lds = (size_t)3 * (TY + 2) * (TX + 2) * (CC + 8) * sizeof(bf16_t);
    dim3 grid(d[0] * (d[1] / TY) * (d[2] / TX), Co / (16 * NB));
#define UN_LDS(MB_, NB_, CC_)                                                                     \
  hipLaunchKernelGGL((k_conv3_lds<MB_, NB_, CC_>), grid, dim3(256), lds, st, in, Ci, W, Kpad, bias, \
                     Co, d[0], d[1], d[2], TY, TX, out, relu, mask)
    if (NB == 1 && CC == 16) UN_LDS(4, 1, 16);
    else if (NB == 1) UN_LDS(2, 1, 32);
    else if (NB == 2 && CC == 16) UN_LDS(2, 2, 16);
    else if (NB == 2) UN_LDS(2, 2, 32);
    else if (CC == 16) UN_LDS(2, 4, 16);
    else UN_LDS(2, 4, 32);
#undef UN_LDS
    return;
  }
  const int MB = 8 / NB;
  dim3 grid((V + 64 * MB - 1) / (64 * MB), Co / (16 * NB));
  if (Ci % 8 != 0) {
    hipLaunchKernelGGL((k_conv3<false, 8, 1>), grid, dim3(256), 0, st, in, Ci, W, Kpad, bias, Co,
                       d[0], d[1], d[2], out, relu, mask);
  } else if (NB == 1) {
    hipLaunchKernelGGL((k_conv3<true, 8, 1>), grid, dim3(256), 0, st, in, Ci, W, Kpad, bias, Co,
                       d[0], d[1], d[2], out, relu, mask);
  } else if (NB == 2) {
    hipLaunchKernelGGL((k_conv3<true, 4, 2>), grid, dim3(256), 0, st, in, Ci, W, Kpad, bias, Co,
                       d[0], d[1], d[2], out, relu, mask);
  } else {
    hipLaunchKernelGGL((k_conv3<true, 2, 4>), grid, dim3(256), 0, st, in, Ci, W, Kpad, bias, Co,
                       d[0], d[1], d[2], out, relu, mask);
  }
}

void launch_wgrad(hipStream_t st, const bf16_t *in, int Ci, const bf16_t *gout, int Co,
                  const int *d, float *dW) {
  const int V = d[0] * d[1] * d[2];
  const int MBo = std::min(Co / 16, 4);
  const int gy = Co / (16 * MBo), gz = (27 * Ci + 63) / 64;
  int nx = std::max(1, 2048 / (gy * gz));
  int vchunk = (V + nx - 1) / nx;
  vchunk = std::max(128, (vchunk + 127) / 128 * 128);
  nx = (V + vchunk - 1) / vchunk;
  dim3 grid(nx, gy, gz);
  if (MBo == 1)
    hipLaunchKernelGGL((k_conv3_wgrad<1, 4>), grid, dim3(256), 0, st, in, Ci, gout, Co, d[0], d[1],
                       d[2], vchunk, dW);
  else if (MBo == 2)
    hipLaunchKernelGGL((k_conv3_wgrad<2, 4>), grid, dim3(256), 0, st, in, Ci, gout, Co, d[0], d[1],
                       d[2], vchunk, dW);
  else
    hipLaunchKernelGGL((k_conv3_wgrad<4, 4>), grid, dim3(256), 0, st, in, Ci, gout, Co, d[0], d[1],
                       d[2], vchunk, dW);
}

void launch_bias_grad(hipStream_t st, const bf16_t *gout, int C, int V, float *db) {
  const int vchunk = std::max(256, (V + 1023) / 1024);
  hipLaunchKernelGGL(k_bias_grad, dim3((V + vchunk - 1) / vchunk), dim3(256), 0, st, gout, C, V,
                     vchunk, db);
}

// forward of every layer on stream st; x is the float32 input volume
void enqueue_forward(xt_unet *u, const float *x, float *out, hipStream_t st) {
  if (u->wdirty) {
    for (int l = 0; l < NCONV; l++) {
      const Conv &c = u->conv[l];
      const size_t n = (size_t)c.Co * c.Kpad + (size_t)c.Ci * c.Kpadb;
      hipLaunchKernelGGL(k_wprep, dim3(grid1(n)), dim3(256), 0, st, u->params + c.w_off, c.Ci,
                         c.Co, c.Kpad, c.Kpadb, c.Wf, c.Wb);
    }
    u->wdirty = false;
  }
  const size_t V0 = u->V[0];
  hipLaunchKernelGGL(k_prep, dim3(grid1(V0)), dim3(256), 0, st, x, u->scale, V0, u->t_in);
  auto conv = [&](int l, const bf16_t *in, bf16_t *o) {
    const Conv &c = u->conv[l];
    launch_conv(u, st, in, c.Ci, c.Wf, c.Kpad, u->params + c.b_off, c.Co, u->dims[c.lvl], o, 1,
                nullptr);
  };
  auto pool = [&](int lvl, const bf16_t *in, int C, bf16_t *o) {
    const int *d = u->dims[lvl];
    hipLaunchKernelGGL(k_pool, dim3(grid1(u->V[lvl + 1] * C / 8)), dim3(256), 0, st, in, C, d[0],
                       d[1], d[2], o);
  };
  auto upcat = [&](int lvl, const bf16_t *low, int Ca, const bf16_t *skip, int Cb, bf16_t *o) {
    const int *d = u->dims[lvl];
    hipLaunchKernelGGL(k_upcat, dim3(grid1(u->V[lvl] * (Ca + Cb) / 8)), dim3(256), 0, st, low, Ca,
                       skip, Cb, d[0], d[1], d[2], o);
  };
  conv(0, u->t_in, u->e1a);
  conv(1, u->e1a, u->e1);
  pool(0, u->e1, 16, u->p1);
  conv(2, u->p1, u->e2a);
  conv(3, u->e2a, u->e2);
  pool(1, u->e2, 32, u->p2);
  conv(4, u->p2, u->e3a);
  conv(5, u->e3a, u->e3);
  pool(2, u->e3, 64, u->p3);
  conv(6, u->p3, u->ba);
  conv(7, u->ba, u->bb);
  upcat(2, u->bb, 128, u->e3, 64, u->u3);
  conv(8, u->u3, u->d3a);
  conv(9, u->d3a, u->d3);
  upcat(1, u->d3, 64, u->e2, 32, u->u2);
  conv(10, u->u2, u->d2a);
  conv(11, u->d2a, u->d2);
  upcat(0, u->d2, 32, u->e1, 16, u->u1);
  conv(12, u->u1, u->d1a);
  conv(13, u->d1a, u->d1);
  if (out)  // training recomputes the head inside its backward
    hipLaunchKernelGGL(k_head_fwd, dim3(grid1(V0)), dim3(256), 0, st, u->d1, HEAD_C,
                       u->params + u->head_w, u->params + u->head_b, x, u->scale, V0, out);
}

// backward after enqueue_forward(x): gradient of mean((out - tgt)^2 * scale^2)
void enqueue_backward(xt_unet *u, const float *x, const float *tgt, hipStream_t st) {
  const size_t V0 = u->V[0];
  hipLaunchKernelGGL(k_fill, dim3(grid1(u->nparam)), dim3(256), 0, st, u->grads,
                     (size_t)u->nparam, 0.f);
  hipLaunchKernelGGL(k_fill_d, dim3(1), dim3(1), 0, st, u->loss_dev, 0.0);
  hipLaunchKernelGGL(k_head_bwd, dim3(std::min<size_t>((V0 + 255) / 256, 1024)), dim3(256), 0, st,
                     u->d1, HEAD_C, u->params + u->head_w, u->params + u->head_b, x, tgt, u->scale,
                     V0, u->loss_dev, u->grads + u->head_w, u->grads + u->head_b, u->g_d1);
  // conv l: input `in`, pre-activation gradient `g`; input gradient to `gin`
  // (masked by `in` when it is a ReLU output), skipped for the stem
  auto back = [&](int l, const bf16_t *in, const bf16_t *g, bf16_t *gin, bool in_relu) {
    const Conv &c = u->conv[l];
    const int *d = u->dims[c.lvl];
    launch_wgrad(st, in, c.Ci, g, c.Co, d, u->grads + c.w_off);
    launch_bias_grad(st, g, c.Co, (int)u->V[c.lvl], u->grads + c.b_off);
    if (gin)
      launch_conv(u, st, g, c.Co, c.Wb, c.Kpadb, nullptr, c.Ci, d, gin, 0, in_relu ? in : nullptr);
  };
  auto up_bwd = [&](int lvl_coarse, const bf16_t *gcat, int Ca, int Ct, bf16_t *glow) {
    const int *e = u->dims[lvl_coarse];
    hipLaunchKernelGGL(k_up_bwd, dim3(grid1(u->V[lvl_coarse] * Ca)), dim3(256), 0, st, gcat, Ca,
                       Ct, e[0], e[1], e[2], glow);
  };
  auto skip_bwd = [&](int lvl, const bf16_t *gcat, int Ca, int Ct, const bf16_t *gpool,
                      const bf16_t *e, int Cb, bf16_t *ge) {
    const int *d = u->dims[lvl];
    hipLaunchKernelGGL(k_skip_bwd, dim3(grid1(u->V[lvl] * Cb)), dim3(256), 0, st, gcat, Ca, Ct,
                       gpool, e, Cb, d[0], d[1], d[2], ge);
  };
  back(13, u->d1a, u->g_d1, u->g_d1a, true);
  back(12, u->u1, u->g_d1a, u->g_u1, true);
  up_bwd(1, u->g_u1, 32, 48, u->g_d2);
  back(11, u->d2a, u->g_d2, u->g_d2a, true);
  back(10, u->u2, u->g_d2a, u->g_u2, true);
  up_bwd(2, u->g_u2, 64, 96, u->g_d3);
  back(9, u->d3a, u->g_d3, u->g_d3a, true);
  back(8, u->u3, u->g_d3a, u->g_u3, true);
  up_bwd(3, u->g_u3, 128, 192, u->g_bb);
  back(7, u->ba, u->g_bb, u->g_ba, true);
  back(6, u->p3, u->g_ba, u->g_p3, false);
  skip_bwd(2, u->g_u3, 128, 192, u->g_p3, u->e3, 64, u->g_e3);
  back(5, u->e3a, u->g_e3, u->g_e3a, true);
  back(4, u->p2, u->g_e3a, u->g_p2, false);
  skip_bwd(1, u->g_u2, 64, 96, u->g_p2, u->e2, 32, u->g_e2);
  back(3, u->e2a, u->g_e2, u->g_e2a, true);
  back(2, u->p1, u->g_e2a, u->g_p1, false);
  skip_bwd(0, u->g_u1, 32, 48, u->g_p1, u->e1, 16, u->g_e1);
  back(1, u->e1a, u->g_e1, u->g_e1a, true);
  back(0, u->t_in, u->g_e1a, nullptr, false);
}

void free_net(xt_unet *u) {
  if (!u) return;
  void *bufs[] = {u->params, u->grads, u->adam_m, u->adam_v, u->wbuf,
                  u->abuf,   u->gbuf,  u->loss_dev, u->stage};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : u->ev)
    if (e) (void)hipEventDestroy(e);
  if (u->st) (void)hipStreamDestroy(u->st);
  delete u;
}

}  // namespace

void xt_unet_shape(const xt_unet *u, int *d0, int *d1, int *d2) {
  *d0 = u->dims[0][0];
  *d1 = u->dims[0][1];
  *d2 = u->dims[0][2];
}

int xt_unet_forward_on(xt_unet *u, const float *in, float *out, hipStream_t st) {
  enqueue_forward(u, in, out, st);
  UN_HIP(hipGetLastError());
  return XT_OK;
}

extern "C" int xt_unet_create(int32_t d0, int32_t d1, int32_t d2, float scale, uint64_t seed,
                              int32_t train, xt_unet **net) {
  if (!net) return xt_report_error(XT_E_ARG, "null net pointer");
  *net = nullptr;
  if (d0 <= 0 || d1 <= 0 || d2 <= 0 || d0 % 8 || d1 % 8 || d2 % 8)
    return xt_report_error(XT_E_ARG, "U-Net volume %dx%dx%d: each size must be a positive "
                           "multiple of 8", d0, d1, d2);
  if ((int64_t)d0 * d1 * d2 * 192 >= (int64_t)1 << 31)
    return xt_report_error(XT_E_ARG, "U-Net volume %dx%dx%d too large (V*192 must be < 2^31)",
                           d0, d1, d2);
  if (!(scale > 0.f) || !std::isfinite(scale))
    return xt_report_error(XT_E_ARG, "U-Net input scale must be positive");
  xt_unet *u = new xt_unet();
  u->scale = scale;
  u->train = train != 0;
  (void)hipGetDevice(&u->dev);
  for (int l = 0; l < 4; l++) {
    u->dims[l][0] = d0 >> l;
    u->dims[l][1] = d1 >> l;
    u->dims[l][2] = d2 >> l;
    u->V[l] = (size_t)u->dims[l][0] * u->dims[l][1] * u->dims[l][2];
  }
  // parameters
  size_t off = 0, wcount = 0;
  for (int l = 0; l < NCONV; l++) {
    Conv &c = u->conv[l];
    c.lvl = kConvSpec[l][0];
    c.Ci = kConvSpec[l][1];
    c.Co = kConvSpec[l][2];
    c.Kpad = (27 * c.Ci + 31) / 32 * 32;
    c.Kpadb = (27 * c.Co + 31) / 32 * 32;
    c.w_off = off;
    off += (size_t)c.Co * 27 * c.Ci;
    c.b_off = off;
    off += c.Co;
    wcount += (size_t)c.Co * c.Kpad + (size_t)c.Ci * c.Kpadb;
  }
  u->head_w = off;
  off += HEAD_C;
  u->head_b = off;
  off += 1;
  u->nparam = (int64_t)off;
  auto fail = [&](int rc) {
    free_net(u);
    return rc;
  };
#define UN_ALLOC(ptr, bytes)                                                               \
  do {                                                                                     \
    hipError_t _e = hipMalloc((void **)&(ptr), (bytes));                                   \
    if (_e != hipSuccess)                                                                  \
      return fail(xt_report_error(XT_E_HIP, "U-Net allocation of %zu bytes: %s",           \
                                  (size_t)(bytes), hipGetErrorString(_e)));                \
  } while (0)
  UN_ALLOC(u->params, off * sizeof(float));
  UN_ALLOC(u->wbuf, wcount * sizeof(bf16_t) + 256);
  {
    size_t w = 0;
    for (int l = 0; l < NCONV; l++) {
      Conv &c = u->conv[l];
      c.Wf = u->wbuf + w;
      w += (size_t)c.Co * c.Kpad;
      c.Wb = u->wbuf + w;
      w += (size_t)c.Ci * c.Kpadb;
    }
  }
  const size_t *V = u->V;
  Carver a;
  a.add(&u->t_in, V[0]);
  a.add(&u->e1a, V[0] * 16);
  a.add(&u->e1, V[0] * 16);
  a.add(&u->p1, V[1] * 16);
  a.add(&u->e2a, V[1] * 32);
  a.add(&u->e2, V[1] * 32);
  a.add(&u->p2, V[2] * 32);
  a.add(&u->e3a, V[2] * 64);
  a.add(&u->e3, V[2] * 64);
  a.add(&u->p3, V[3] * 64);
  a.add(&u->ba, V[3] * 128);
  a.add(&u->bb, V[3] * 128);
  a.add(&u->u3, V[2] * 192);
  a.add(&u->d3a, V[2] * 64);
  a.add(&u->d3, V[2] * 64);
  a.add(&u->u2, V[1] * 96);
  a.add(&u->d2a, V[1] * 32);
  a.add(&u->d2, V[1] * 32);
  a.add(&u->u1, V[0] * 48);
  a.add(&u->d1a, V[0] * 16);
  a.add(&u->d1, V[0] * 16);
  UN_ALLOC(u->abuf, a.off * sizeof(bf16_t));
  a.assign(u->abuf);
  if (u->train) {
    Carver gc;
    gc.add(&u->g_d1, V[0] * 16);
    gc.add(&u->g_d1a, V[0] * 16);
    gc.add(&u->g_u1, V[0] * 48);
    gc.add(&u->g_d2, V[1] * 32);
    gc.add(&u->g_d2a, V[1] * 32);
    gc.add(&u->g_u2, V[1] * 96);
    gc.add(&u->g_d3, V[2] * 64);
    gc.add(&u->g_d3a, V[2] * 64);
    gc.add(&u->g_u3, V[2] * 192);
    gc.add(&u->g_bb, V[3] * 128);
    gc.add(&u->g_ba, V[3] * 128);
    gc.add(&u->g_p3, V[3] * 64);
    gc.add(&u->g_e3, V[2] * 64);
    gc.add(&u->g_e3a, V[2] * 64);
    gc.add(&u->g_p2, V[2] * 32);
    gc.add(&u->g_e2, V[1] * 32);
    gc.add(&u->g_e2a, V[1] * 32);
    gc.add(&u->g_p1, V[1] * 16);
    gc.add(&u->g_e1, V[0] * 16);
    gc.add(&u->g_e1a, V[0] * 16);
    UN_ALLOC(u->gbuf, gc.off * sizeof(bf16_t));
    gc.assign(u->gbuf);
    UN_ALLOC(u->grads, off * sizeof(float));
    UN_ALLOC(u->adam_m, off * sizeof(float));
    UN_ALLOC(u->adam_v, off * sizeof(float));
    UN_ALLOC(u->loss_dev, sizeof(double));
  }
#undef UN_ALLOC
  if (hipStreamCreate(&u->st) != hipSuccess) return fail(xt_report_error(XT_E_HIP, "stream"));
  for (auto &e : u->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(xt_report_error(XT_E_HIP, "events"));
  // He-normal weights, zero biases, zero head
  std::vector<float> h(off, 0.f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < NCONV; l++) {
    const Conv &c = u->conv[l];
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (27.f * c.Ci)));
    for (size_t i = 0; i < (size_t)c.Co * 27 * c.Ci; i++) h[c.w_off + i] = nd(rng);
  }
  int rc = xt_unet_set_params(u, h.data(), (int64_t)off);
  if (rc != XT_OK) return fail(rc);
  *net = u;
  return XT_OK;
}

extern "C" void xt_unet_destroy(xt_unet *u) { free_net(u); }

extern "C" int64_t xt_unet_param_count(const xt_unet *u) { return u ? u->nparam : 0; }

extern "C" int xt_unet_get_scale(const xt_unet *u, float *scale) {
  if (!u || !scale) return xt_report_error(XT_E_ARG, "null argument");
  *scale = u->scale;
  return XT_OK;
}

extern "C" int xt_unet_get_params(xt_unet *u, float *host, int64_t n) {
  if (!u || !host || n != u->nparam)
    return xt_report_error(XT_E_ARG, "get_params: need %lld floats", (long long)(u ? u->nparam : 0));
  UN_HIP(hipStreamSynchronize(u->st));
  UN_HIP(hipMemcpy(host, u->params, n * sizeof(float), hipMemcpyDeviceToHost));
  return XT_OK;
}

extern "C" int xt_unet_set_params(xt_unet *u, const float *host, int64_t n) {
  if (!u || !host || n != u->nparam)
    return xt_report_error(XT_E_ARG, "set_params: need %lld floats", (long long)(u ? u->nparam : 0));
  UN_HIP(hipStreamSynchronize(u->st));
  UN_HIP(hipMemcpy(u->params, host, n * sizeof(float), hipMemcpyHostToDevice));
  if (u->train) {
    hipLaunchKernelGGL(k_fill, dim3(grid1(n)), dim3(256), 0, u->st, u->adam_m, (size_t)n, 0.f);
    hipLaunchKernelGGL(k_fill, dim3(grid1(n)), dim3(256), 0, u->st, u->adam_v, (size_t)n, 0.f);
    hipLaunchKernelGGL(k_fill, dim3(grid1(n)), dim3(256), 0, u->st, u->grads, (size_t)n, 0.f);
    u->adam_t = 0;
  }
  u->wdirty = true;
  UN_HIP(hipStreamSynchronize(u->st));
  return XT_OK;
}

extern "C" int xt_unet_forward(xt_unet *u, const float *in, float *out) {
  if (!u || !in || !out) return xt_report_error(XT_E_ARG, "null argument");
  UN_HIP(hipEventRecord(u->ev[0], u->st));
  UN_TRY(xt_unet_forward_on(u, in, out, u->st));
  UN_HIP(hipEventRecord(u->ev[1], u->st));
  UN_HIP(hipEventSynchronize(u->ev[1]));
  float ms = 0.f;
  UN_HIP(hipEventElapsedTime(&ms, u->ev[0], u->ev[1]));
  u->fwd_ms = ms;
  return XT_OK;
}

extern "C" int xt_unet_loss_grad(xt_unet *u, const float *in, const float *tgt, double *loss,
                                 float *grad_host) {
  if (!u || !in || !tgt) return xt_report_error(XT_E_ARG, "null argument");
  if (!u->train) return xt_report_error(XT_E_ARG, "network was created without training buffers");
  UN_HIP(hipEventRecord(u->ev[0], u->st));
  enqueue_forward(u, in, nullptr, u->st);
  UN_HIP(hipEventRecord(u->ev[1], u->st));
  enqueue_backward(u, in, tgt, u->st);
  UN_HIP(hipGetLastError());
  UN_HIP(hipEventRecord(u->ev[2], u->st));
  UN_HIP(hipEventSynchronize(u->ev[2]));
  float f = 0.f, b = 0.f;
  UN_HIP(hipEventElapsedTime(&f, u->ev[0], u->ev[1]));
  UN_HIP(hipEventElapsedTime(&b, u->ev[1], u->ev[2]));
  u->fwd_ms = f;
  u->bwd_ms = b;
  if (loss) UN_HIP(hipMemcpy(loss, u->loss_dev, sizeof(double), hipMemcpyDeviceToHost));
  if (grad_host)
    UN_HIP(hipMemcpy(grad_host, u->grads, u->nparam * sizeof(float), hipMemcpyDeviceToHost));
  return XT_OK;
}

extern "C" int xt_unet_adam_step(xt_unet *u, float lr, float b1, float b2, float eps) {
  if (!u || !u->train) return xt_report_error(XT_E_ARG, "network has no training buffers");
  u->adam_t++;
  const float bc1 = 1.f - std::pow(b1, (float)u->adam_t), bc2 = 1.f - std::pow(b2, (float)u->adam_t);
  hipLaunchKernelGGL(k_adam, dim3(grid1(u->nparam)), dim3(256), 0, u->st, u->params, u->grads,
                     u->adam_m, u->adam_v, (size_t)u->nparam, lr, b1, b2, eps, bc1, bc2);
  UN_HIP(hipGetLastError());
  u->wdirty = true;
  UN_HIP(hipStreamSynchronize(u->st));
  return XT_OK;
}

extern "C" int xt_unet_train_step(xt_unet *u, const float *in, const float *tgt, float lr,
                                  double *loss) {
  UN_TRY(xt_unet_loss_grad(u, in, tgt, loss, nullptr));
  return xt_unet_adam_step(u, lr, 0.9f, 0.999f, 1e-8f);
}

static int stage_in(xt_unet *u, const float *a, const float *b) {
  if (!u->stage) UN_HIP(hipMalloc((void **)&u->stage, 2 * u->V[0] * sizeof(float)));
  UN_HIP(hipMemcpy(u->stage, a, u->V[0] * sizeof(float), hipMemcpyHostToDevice));
  if (b) UN_HIP(hipMemcpy(u->stage + u->V[0], b, u->V[0] * sizeof(float), hipMemcpyHostToDevice));
  return XT_OK;
}

extern "C" int xt_unet_forward_host(xt_unet *u, const float *in, float *out) {
  if (!u || !in || !out) return xt_report_error(XT_E_ARG, "null argument");
  UN_TRY(stage_in(u, in, nullptr));
  UN_TRY(xt_unet_forward(u, u->stage, u->stage + u->V[0]));
  UN_HIP(hipMemcpy(out, u->stage + u->V[0], u->V[0] * sizeof(float), hipMemcpyDeviceToHost));
  return XT_OK;
}

extern "C" int xt_unet_loss_grad_host(xt_unet *u, const float *in, const float *tgt, double *loss,
                                      float *grad_host) {
  if (!u || !in || !tgt) return xt_report_error(XT_E_ARG, "null argument");
  UN_TRY(stage_in(u, in, tgt));
  return xt_unet_loss_grad(u, u->stage, u->stage + u->V[0], loss, grad_host);
}

extern "C" int xt_unet_train_step_host(xt_unet *u, const float *in, const float *tgt, float lr,
                                       double *loss) {
  if (!u || !in || !tgt) return xt_report_error(XT_E_ARG, "null argument");
  UN_TRY(stage_in(u, in, tgt));
  return xt_unet_train_step(u, u->stage, u->stage + u->V[0], lr, loss);
}

extern "C" int xt_unet_last_ms(const xt_unet *u, double *fwd_ms, double *bwd_ms) {
  if (!u) return xt_report_error(XT_E_ARG, "null net");
  if (fwd_ms) *fwd_ms = u->fwd_ms;
  if (bwd_ms) *bwd_ms = u->bwd_ms;
  return XT_OK;
}
