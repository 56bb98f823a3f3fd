// cnn.hip -- half-width proposal network: patch features, forward, backward,
// Adam, and the proposal pass of AVISM_GROW_LEARNED.
//
// The network has no counterpart in the reference.  It sees a 9^3 patch of
// (delta, div/div_rms, |grad delta|/GRAD_THRE) around a candidate centre and
// predicts the half-width that VOIDFIND's growth loop (voidfinding.f90:500-654)
// reaches; the exact criteria re-verify every proposal (voids.hip), so the
// catalogue never depends on the prediction.
//
//   conv1 3->16, conv2 16->16, conv3 16->16 (3^3, valid, ReLU), head 432->8.
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x16_bf16 (bf16 operands,
// fp32 accumulation).  Output positions are the M rows, output channels the 16
// N columns and K = (3^3 offset, input channel) with the channel fastest, so a
// lane's four K values are four adjacent channels of one LDS row (one 8-byte
// LDS read).  conv1's 3 input channels are padded to 4.  One wave per patch;
// activations are stored as bf16 in LDS between layers.
//
// Operand layout of v_mfma_f32_16x16x16_bf16 (64 lanes, l = lane):
//   A: lane l holds A[l % 16][4*(l/16) + v], v = 0..3
//   B: lane l holds B[4*(l/16) + v][l % 16]
//   D: lane l holds D[4*(l/16) + v][l % 16]
//
// Parameter vector (fp32): W1[108][16] b1[16] W2[432][16] b2[16] W3[432][16]
// b3[16] WH[432][8] bh[8]; weight row index = offset*CIN + channel, the head's
// = position*16 + channel.  Weights are rounded to bf16 when used.
#include <cstring>

#include "avism_internal.h"

namespace avism {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int OFF_W1 = 0, OFF_B1 = 1728, OFF_W2 = 1744, OFF_B2 = 8656, OFF_W3 = 8672, OFF_B3 = 15584,
              OFF_WH = 15600, OFF_BH = 19056;
constexpr int NCLS = 8;
constexpr int CACHE_PER = 343 * 16 + 125 * 16 + 27 * 16;  // bf16 activations per patch
static_assert(CACHE_PER == AVISM_CNN_CACHE, "activation cache size differs from the header");
static_assert(OFF_BH + NCLS == AVISM_CNN_NPARAM, "parameter count differs from the header");

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float bfr(float f) { return bf2f(f2bf(f)); }

__device__ __forceinline__ v4s pack4(float a, float b, float c, float d) {
  v4s r;
  r[0] = (short)f2bf(a);
  r[1] = (short)f2bf(b);
  r[2] = (short)f2bf(c);
  r[3] = (short)f2bf(d);
  return r;
}

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// input-map position of output position p (O^3 map, x fastest) at 3^3 offset off
template <int O>
__device__ __forceinline__ int inpos(int p, int off) {
  constexpr int S = O + 2;
  int ox = p % O, oy = (p / O) % O, oz = p / (O * O);
  int dx = off % 3, dy = (off / 3) % 3, dz = off / 9;
  return ((oz + dz) * S + (oy + dy)) * S + (ox + dx);
}

struct Feat {
  const float* u1co;
  const float* div;
  const int32_t* lin;
  const int32_t* order;
  int n, low1, ne, periodic;
  float inv_rms, inv_two_dx, inv_grad;
  const double* div_sq;  // sum of DIVERCO^2 over the inner cells (device)
  double ncell;
  // 1 / RMS(DIVERCO) from the device sum (0 when the field is zero)
  __device__ __forceinline__ float rms_inv() const {
    const double rms = sqrt(*div_sq / ncell);
    return rms > 0 ? (float)(1.0 / rms) : 0.0f;
  }
  __device__ __forceinline__ int wrapi(int i) const {
    int m = (i - 1) % n;
    if (m < 0) m += n;
    return m + 1;
  }
  __device__ __forceinline__ bool inside(int i, int j, int k) const {
    return periodic || (i >= 1 && i <= n && j >= 1 && j <= n && k >= 1 && k <= n);
  }
  // inner linear index; periodic wrap, else clamped to the box
  __device__ __forceinline__ size_t at(int i, int j, int k) const {
    if (periodic) {
      i = wrapi(i);
      j = wrapi(j);
      k = wrapi(k);
    }
    i = max(1, min(n, i));
    j = max(1, min(n, j));
    k = max(1, min(n, k));
    return (size_t)(i - 1) + (size_t)n * ((size_t)(j - 1) + (size_t)n * (k - 1));
  }
  __device__ __forceinline__ void values(int i, int j, int k, float& f0, float& f1, float& f2) const {
    if (!inside(i, j, k)) {
      f0 = f1 = f2 = 0.0f;
      return;
    }
    size_t c = at(i, j, k);
    f0 = u1co[c] - 1.0f;
    f1 = div[c] * inv_rms;
    float gx = (u1co[at(i + 1, j, k)] - u1co[at(i - 1, j, k)]) * inv_two_dx;
    float gy = (u1co[at(i, j + 1, k)] - u1co[at(i, j - 1, k)]) * inv_two_dx;
    float gz = (u1co[at(i, j, k + 1)] - u1co[at(i, j, k - 1)]) * inv_two_dx;
    f2 = sqrtf(gx * gx + gy * gy + gz * gz) * inv_grad;
  }
  // centre (ext reference indices) of the candidate processed at rank t
  __device__ __forceinline__ void centre(int64_t t, int& I, int& J, int& K) const {
    int e = lin[order[t]];
    I = e % ne + low1;
    J = (e / ne) % ne + low1;
    K = e / (ne * ne) + low1;
  }
};

// features of candidate rank t into X0 (LDS [729][4] bf16, 4th channel 0)
__device__ void build_patch(const Feat& F, int64_t t, uint16_t* X0, int lane) {
  int I, J, K;
  F.centre(t, I, J, K);
  for (int q = lane; q < 729; q += 64) {
    int dx = q % 9 - 4, dy = (q / 9) % 9 - 4, dz = q / 81 - 4;
    float f0, f1, f2;
    F.values(I + dx, J + dy, K + dz, f0, f1, f2);
    X0[q * 4 + 0] = f2bf(f0);
    X0[q * 4 + 1] = f2bf(f1);
    X0[q * 4 + 2] = f2bf(f2);
    X0[q * 4 + 3] = 0;
  }
}

__device__ void load_patch(const uint16_t* __restrict__ patches, int64_t p, uint16_t* X0, int lane) {
  for (int q = lane; q < 729; q += 64) {
    X0[q * 4 + 0] = patches[(p * 3 + 0) * 729 + q];
    X0[q * 4 + 1] = patches[(p * 3 + 1) * 729 + q];
    X0[q * 4 + 2] = patches[(p * 3 + 2) * 729 + q];
    X0[q * 4 + 3] = 0;
  }
}

// B fragments of the three convolutions, held in registers for a whole launch
struct Frags {
  v4s b1[7], b2[27], b3[27];
  float bias1, bias2, bias3;
};

__device__ void load_frags(const float* __restrict__ P, Frags& F, int lane) {
  const int n = lane & 15, kb = 4 * (lane >> 4);
  for (int ks = 0; ks < 7; ++ks) {
    float w[4];
    for (int v = 0; v < 4; ++v) {
      int k = ks * 16 + kb + v;
      w[v] = k < 108 ? P[OFF_W1 + k * 16 + n] : 0.0f;
    }
    F.b1[ks] = pack4(w[0], w[1], w[2], w[3]);
  }
  for (int ks = 0; ks < 27; ++ks) {
    int k = ks * 16 + kb;
    F.b2[ks] = pack4(P[OFF_W2 + k * 16 + n], P[OFF_W2 + (k + 1) * 16 + n], P[OFF_W2 + (k + 2) * 16 + n],
                     P[OFF_W2 + (k + 3) * 16 + n]);
    F.b3[ks] = pack4(P[OFF_W3 + k * 16 + n], P[OFF_W3 + (k + 1) * 16 + n], P[OFF_W3 + (k + 2) * 16 + n],
                     P[OFF_W3 + (k + 3) * 16 + n]);
  }
  F.bias1 = P[OFF_B1 + n];
  F.bias2 = P[OFF_B2 + n];
  F.bias3 = P[OFF_B3 + n];
}

// forward of the patch in X0.  A2 may alias X0 and A3 may alias A1 (conv2
// reads only A1, conv3 only A2).  Returns the logits in lanes 0..7.
__device__ float forward_patch(const float* __restrict__ P, const Frags& F, const uint16_t* X0, uint16_t* A1,
                               uint16_t* A2, uint16_t* A3, int lane) {
  const int n = lane & 15, kb = 4 * (lane >> 4);
  // conv1: 343 outputs (7^3) in 22 row tiles, K = 27*4 = 108 (7 steps)
  for (int mt = 0; mt < 22; ++mt) {
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int pos = mt * 16 + n;
    for (int ks = 0; ks < 7; ++ks) {
      const int off = ks * 4 + (lane >> 4);
      v4s a = {0, 0, 0, 0};
      if (pos < 343 && off < 27) a = *(const v4s*)&X0[inpos<7>(pos, off) * 4];
      acc = mfma(a, F.b1[ks], acc);
    }
    for (int v = 0; v < 4; ++v) {
      const int row = mt * 16 + kb + v;
      if (row < 343) A1[row * 16 + n] = f2bf(fmaxf(acc[v] + F.bias1, 0.0f));
    }
  }
  __syncthreads();
  // conv2: 125 outputs (5^3), K = 27*16
  for (int mt = 0; mt < 8; ++mt) {
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int pos = mt * 16 + n;
    for (int ks = 0; ks < 27; ++ks) {
      v4s a = {0, 0, 0, 0};
      if (pos < 125) a = *(const v4s*)&A1[inpos<5>(pos, ks) * 16 + kb];
      acc = mfma(a, F.b2[ks], acc);
    }
    for (int v = 0; v < 4; ++v) {
      const int row = mt * 16 + kb + v;
      if (row < 125) A2[row * 16 + n] = f2bf(fmaxf(acc[v] + F.bias2, 0.0f));
    }
  }
  __syncthreads();
  // conv3: 27 outputs (3^3)
  for (int mt = 0; mt < 2; ++mt) {
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int pos = mt * 16 + n;
    for (int ks = 0; ks < 27; ++ks) {
      v4s a = {0, 0, 0, 0};
      if (pos < 27) a = *(const v4s*)&A2[inpos<3>(pos, ks) * 16 + kb];
      acc = mfma(a, F.b3[ks], acc);
    }
    for (int v = 0; v < 4; ++v) {
      const int row = mt * 16 + kb + v;
      if (row < 27) A3[row * 16 + n] = f2bf(fmaxf(acc[v] + F.bias3, 0.0f));
    }
  }
  __syncthreads();
  // head: lane (j = lane & 7, chunk = lane >> 3) sums 54 of the 432 features
  const int j = lane & 7, ch = lane >> 3;
  float s = 0.0f;
  for (int k = ch * 54; k < ch * 54 + 54; ++k) s += bf2f(A3[k]) * bfr(P[OFF_WH + k * NCLS + j]);
  s += __shfl_xor(s, 8);
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  return s + P[OFF_BH + j];
}

// ---------------------------------------------------------------- kernels
// sum of DIVERCO^2 in FP64, in a fixed order: kDivSqBlocks partial sums
// (float4 loads, four accumulators per thread), then one block adds them
constexpr int kDivSqBlocks = 2048;
__global__ void __launch_bounds__(256) k_div_sq(int64_t n, const float* __restrict__ div, double* part) {
  __shared__ double red[256];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const int64_t n4 = n / 4;
  const float4* d4 = (const float4*)div;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (int64_t)kDivSqBlocks * 256) {
    const float4 v = d4[i];
    s0 += (double)v.x * v.x;
    s1 += (double)v.y * v.y;
    s2 += (double)v.z * v.z;
    s3 += (double)v.w * v.w;
  }
  if (blockIdx.x == 0 && threadIdx.x < n - 4 * n4) {
    const double v = div[4 * n4 + threadIdx.x];
    s0 += v * v;
  }
  red[threadIdx.x] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_div_sq_sum(const double* __restrict__ part, double* out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < kDivSqBlocks; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// patches [n][3][729] bf16 (channel, z, y, x) of the candidates at ranks idx[]
__global__ void __launch_bounds__(64) k_patches(Feat F, int64_t n, const int64_t* __restrict__ idx,
                                                uint16_t* out) {
  __shared__ __attribute__((aligned(16))) uint16_t X0[729 * 4];
  const int lane = threadIdx.x;
  F.inv_rms = F.rms_inv();
  for (int64_t p = blockIdx.x; p < n; p += gridDim.x) {
    build_patch(F, idx[p], X0, lane);
    __syncthreads();
    for (int q = lane; q < 729; q += 64)
      for (int c = 0; c < 3; ++c) out[(p * 3 + c) * 729 + q] = X0[q * 4 + c];
    __syncthreads();
  }
}

// forward for training: logits [n][8] and, if cache != null, the bf16
// activations A1 [343][16], A2 [125][16], A3 [27][16] of every patch
__global__ void __launch_bounds__(64) k_forward(const float* __restrict__ P, int64_t n,
                                                const uint16_t* __restrict__ patches, float* logits,
                                                uint16_t* cache) {
  __shared__ __attribute__((aligned(16))) uint16_t X0[729 * 4];
  __shared__ __attribute__((aligned(16))) uint16_t A1[352 * 16];
  __shared__ __attribute__((aligned(16))) uint16_t A2[128 * 16];
  __shared__ __attribute__((aligned(16))) uint16_t A3[32 * 16];
  const int lane = threadIdx.x;
  Frags F;
  load_frags(P, F, lane);
  for (int64_t p = blockIdx.x; p < n; p += gridDim.x) {
    load_patch(patches, p, X0, lane);
    __syncthreads();
    float lg = forward_patch(P, F, X0, A1, A2, A3, lane);
    if (lane < NCLS) logits[p * NCLS + lane] = lg;
    if (cache) {
      uint16_t* cp = cache + p * CACHE_PER;
      for (int q = lane; q < 343 * 16; q += 64) cp[q] = A1[q];
      for (int q = lane; q < 125 * 16; q += 64) cp[343 * 16 + q] = A2[q];
      for (int q = lane; q < 27 * 16; q += 64) cp[343 * 16 + 125 * 16 + q] = A3[q];
    }
    __syncthreads();
  }
}

// backward of the mean softmax cross-entropy over n patches; weight and bias
// gradients are accumulated into g (zeroed by the host) with fp32 atomics.
// Activations and weights enter the MFMAs as the bf16 values the forward used;
// the fp32 gradients dZ enter as a bf16 pair hi + lo (hi = bf16(dZ),
// lo = bf16(dZ - hi)), two MFMAs per tile, so the weight gradients keep
// about 16 bits of dZ instead of 8.
//
// LDS (bytes): X0 [0,5840) | A1 [5840,17104) | DZ2 fp32 [17104,25296) |
// region R [25296,47248): A2, A3, DZ3, DA2 while layers 3 and 2 are
// differentiated, then DA1 (dA1, masked in place into dZ1) over all of them.
constexpr int R_BASE = 25296;
constexpr int BWD_LDS = R_BASE + 343 * 16 * 4;

__device__ __forceinline__ void split4(const float* x, v4s& hi, v4s& lo) {
  float h[4], l[4];
  for (int v = 0; v < 4; ++v) {
    h[v] = bfr(x[v]);
    l[v] = x[v] - h[v];
  }
  hi = pack4(h[0], h[1], h[2], h[3]);
  lo = pack4(l[0], l[1], l[2], l[3]);
}

__global__ void __launch_bounds__(64) k_backward(const float* __restrict__ P, int64_t n,
                                                 const uint16_t* __restrict__ patches,
                                                 const uint16_t* __restrict__ cache,
                                                 const float* __restrict__ logits,
                                                 const int32_t* __restrict__ labels, float inv_b, float* g,
                                                 float* loss) {
  __shared__ __attribute__((aligned(16))) unsigned char sm[BWD_LDS];
  uint16_t* X0 = (uint16_t*)sm;
  uint16_t* A1 = (uint16_t*)(sm + 5840);
  float* DZ2 = (float*)(sm + 17104);
  uint16_t* A2 = (uint16_t*)(sm + R_BASE);
  uint16_t* A3 = (uint16_t*)(sm + R_BASE + 4096);
  float* DZ3 = (float*)(sm + R_BASE + 5120);
  float* DA2 = (float*)(sm + R_BASE + 7168);
  float* DA1 = (float*)(sm + R_BASE);
  const int lane = threadIdx.x;
  const int nn = lane & 15, kb = 4 * (lane >> 4);
  for (int64_t p = blockIdx.x; p < n; p += gridDim.x) {
    const uint16_t* cp = cache + p * CACHE_PER;
    load_patch(patches, p, X0, lane);
    for (int q = lane; q < 352 * 16; q += 64) A1[q] = q < 343 * 16 ? cp[q] : (uint16_t)0;
    for (int q = lane; q < 128 * 16; q += 64) A2[q] = q < 125 * 16 ? cp[343 * 16 + q] : (uint16_t)0;
    for (int q = lane; q < 32 * 16; q += 64) A3[q] = q < 27 * 16 ? cp[343 * 16 + 125 * 16 + q] : (uint16_t)0;
    // softmax cross-entropy (every lane computes the 8 class gradients)
    float lg[NCLS], dl[NCLS];
    float mx = -INFINITY;
    for (int j = 0; j < NCLS; ++j) {
      lg[j] = logits[p * NCLS + j];
      mx = fmaxf(mx, lg[j]);
    }
    float se = 0.0f;
    for (int j = 0; j < NCLS; ++j) se += expf(lg[j] - mx);
    const int lab = min(max(labels[p], 0), NCLS - 1);
    for (int j = 0; j < NCLS; ++j) dl[j] = (expf(lg[j] - mx) / se - (j == lab ? 1.0f : 0.0f)) * inv_b;
    if (lane == 0) atomicAdd(loss, (logf(se) + mx - lg[lab]) * inv_b);
    __syncthreads();
    // head: dWH, dbh; dfeat -> dz3 = dfeat * (a3 > 0)
    for (int k = lane; k < 432; k += 64) {
      const float f = bf2f(A3[k]);
      float df = 0.0f;
      for (int j = 0; j < NCLS; ++j) {
        atomicAdd(&g[OFF_WH + k * NCLS + j], f * dl[j]);
        df += dl[j] * bfr(P[OFF_WH + k * NCLS + j]);
      }
      DZ3[k] = f > 0.0f ? df : 0.0f;
    }
    for (int k = 432 + lane; k < 512; k += 64) DZ3[k] = 0.0f;
    if (lane < NCLS) atomicAdd(&g[OFF_BH + lane], dl[lane]);
    for (int q = lane; q < 125 * 16; q += 64) DA2[q] = 0.0f;
    __syncthreads();
    if (lane < 16) {
      float s = 0.0f;
      for (int q = 0; q < 27; ++q) s += DZ3[q * 16 + lane];
      atomicAdd(&g[OFF_B3 + lane], s);
    }
    // dW3[(off,c)][n] = sum_o A2[inpos(o,off)][c] dZ3[o][n]: M = c, N = n, K = o
    for (int off = 0; off < 27; ++off) {
      v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
      for (int ks = 0; ks < 2; ++ks) {
        float av[4], bv[4];
        for (int v = 0; v < 4; ++v) {
          const int o = ks * 16 + kb + v;
          av[v] = o < 27 ? bf2f(A2[inpos<3>(o, off) * 16 + nn]) : 0.0f;
          bv[v] = DZ3[o * 16 + nn];
        }
        v4s bh, bl;
        split4(bv, bh, bl);
        const v4s a = pack4(av[0], av[1], av[2], av[3]);
        acc = mfma(a, bh, acc);
        acc = mfma(a, bl, acc);
      }
      for (int v = 0; v < 4; ++v) atomicAdd(&g[OFF_W3 + (off * 16 + kb + v) * 16 + nn], acc[v]);
    }
    // dA2 = col2im(dZ3 W3^T): per offset, D[o][c] = sum_n dZ3[o][n] W3[(off,c)][n]
    for (int mt = 0; mt < 2; ++mt) {
      v4s ah, al;
      split4(&DZ3[(mt * 16 + nn) * 16 + kb], ah, al);
      for (int off = 0; off < 27; ++off) {
        const float* w = P + OFF_W3 + (off * 16 + nn) * 16 + kb;
        const v4s b = pack4(w[0], w[1], w[2], w[3]);
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mfma(ah, b, acc);
        acc = mfma(al, b, acc);
        for (int v = 0; v < 4; ++v) {
          const int o = mt * 16 + kb + v;
          if (o < 27) DA2[inpos<3>(o, off) * 16 + nn] += acc[v];
        }
        __syncthreads();
      }
    }
    // dz2 = dA2 * (a2 > 0), rows 125..127 zero
    for (int q = lane; q < 128 * 16; q += 64)
      DZ2[q] = (q < 125 * 16 && bf2f(A2[q]) > 0.0f) ? DA2[q] : 0.0f;
    __syncthreads();
    if (lane < 16) {
      float s = 0.0f;
      for (int q = 0; q < 125; ++q) s += DZ2[q * 16 + lane];
      atomicAdd(&g[OFF_B2 + lane], s);
    }
    // dW2: K = 125 outputs (8 steps)
    for (int off = 0; off < 27; ++off) {
      v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
      for (int ks = 0; ks < 8; ++ks) {
        float av[4], bv[4];
        for (int v = 0; v < 4; ++v) {
          const int o = ks * 16 + kb + v;
          av[v] = o < 125 ? bf2f(A1[inpos<5>(o, off) * 16 + nn]) : 0.0f;
          bv[v] = DZ2[o * 16 + nn];
        }
        v4s bh, bl;
        split4(bv, bh, bl);
        const v4s a = pack4(av[0], av[1], av[2], av[3]);
        acc = mfma(a, bh, acc);
        acc = mfma(a, bl, acc);
      }
      for (int v = 0; v < 4; ++v) atomicAdd(&g[OFF_W2 + (off * 16 + kb + v) * 16 + nn], acc[v]);
    }
    __syncthreads();  // region R is reused for DA1 from here on
    for (int q = lane; q < 343 * 16; q += 64) DA1[q] = 0.0f;
    __syncthreads();
    // dA1 = col2im(dZ2 W2^T)
    for (int mt = 0; mt < 8; ++mt) {
      v4s ah, al;
      split4(&DZ2[(mt * 16 + nn) * 16 + kb], ah, al);
      for (int off = 0; off < 27; ++off) {
        const float* w = P + OFF_W2 + (off * 16 + nn) * 16 + kb;
        const v4s b = pack4(w[0], w[1], w[2], w[3]);
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mfma(ah, b, acc);
        acc = mfma(al, b, acc);
        for (int v = 0; v < 4; ++v) {
          const int o = mt * 16 + kb + v;
          if (o < 125) DA1[inpos<5>(o, off) * 16 + nn] += acc[v];
        }
        __syncthreads();
      }
    }
    // dz1 = dA1 * (a1 > 0), in place
    for (int q = lane; q < 343 * 16; q += 64)
      if (!(bf2f(A1[q]) > 0.0f)) DA1[q] = 0.0f;
    __syncthreads();
    if (lane < 16) {
      float s = 0.0f;
      for (int q = 0; q < 343; ++q) s += DA1[q * 16 + lane];
      atomicAdd(&g[OFF_B1 + lane], s);
    }
    // dW1[k = off*4 + c][n]: M = k (7 tiles of 16, 108 used), K = 343 outputs (22 steps)
    for (int mt = 0; mt < 7; ++mt) {
      v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
      const int k = mt * 16 + nn;
      const int off = k >> 2, c = k & 3;
      for (int ks = 0; ks < 22; ++ks) {
        float av[4], bv[4];
        for (int v = 0; v < 4; ++v) {
          const int o = ks * 16 + kb + v;
          av[v] = (o < 343 && k < 108) ? bf2f(X0[inpos<7>(o, off) * 4 + c]) : 0.0f;
          bv[v] = o < 343 ? DA1[o * 16 + nn] : 0.0f;
        }
        v4s bh, bl;
        split4(bv, bh, bl);
        const v4s a = pack4(av[0], av[1], av[2], av[3]);
        acc = mfma(a, bh, acc);
        acc = mfma(a, bl, acc);
      }
      for (int v = 0; v < 4; ++v) {
        const int kk = mt * 16 + kb + v;
        if (kk < 108) atomicAdd(&g[OFF_W1 + kk * 16 + nn], acc[v]);
      }
    }
    __syncthreads();
  }
}

__global__ void k_adam(int n, float* p, const float* __restrict__ g, float* m, float* v, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------- proposal pass
// k_propose: features, forward and argmax (first maximum wins) of every
// candidate, on gfx950's v_mfma_f32_16x16x32_bf16:
//   A: lane l holds A[l % 16][8*(l/16) + v], v = 0..7
//   B: lane l holds B[8*(l/16) + v][l % 16]
//   D: lane l holds D[4*(l/16) + v][l % 16], v = 0..3
// A K step of 32 covers 8 offsets of conv1 (4 channels each) or 2 offsets of
// conv2/conv3 (16 channels each); K steps: conv1 4 (108 padded to 128), conv2
// and conv3 14 (432 padded to 448).
//
// One wave per candidate; the waves of a block share only the weights, so
// each runs through its own candidates with wave-level ordering.  Every
// B fragment (bf16, 32 K steps x 16 bytes per lane) and the head weights are
// in LDS, written once per block.  The A operands are read from LDS one K
// step (conv2, conv3) or one pair of row tiles (conv1) ahead of the MFMAs that
// use them.  Rows past a layer's outputs read the last real row and write a
// sink row (MFMA rows are independent, so they never reach a real output);
// padded K slots read a real offset against B rows that are zero.  A
// candidate's inputs (U1CO on the 11^3 cells around it, DIVERCO on the 9^3
// patch) are loaded into registers while the previous candidate's layers
// run, then staged in LDS, where the features are computed.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef short v8s __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v4f mfma32(v8s a, v8s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                  0);
}
// round to nearest even (the hardware conversion; equal to f2bf for finite values)
__device__ __forceinline__ uint32_t bfpack(float lo, float hi) {
  v2bf b = {(__bf16)lo, (__bf16)hi};
  return __builtin_bit_cast(uint32_t, b);
}
__device__ __forceinline__ uint16_t bf1(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

constexpr int KS1 = 4, KS2 = 14, KS3 = 14, KSALL = KS1 + KS2 + KS3;
constexpr int PR_WF = KSALL * 64 * 16;        // B fragments [K step][lane]
constexpr int PR_WH = NCLS * 432 * 2;         // head weights [class][feature], bf16
constexpr int PR_BIAS = 64 * 4;               // b1, b2, b3 [16] and bh [8], fp32
constexpr int PR_X0 = 5840;                   // X0 [729][4] bf16; A2 [126][16] later (row 125: sink)
constexpr int PR_A1 = 11008;                  // inputs U1CO [1344] + DIVERCO [768] (f32); A1 [344][16]
                                              // (row 343: sink); A3 [28][16] (row 27: sink)
constexpr int PR_HU = 21 * 64;                // DIVERCO's offset in the inputs (floats)
constexpr int PR_AX = 272;                    // two sets of 33 axis indices
constexpr int PR_WAVE = PR_X0 + PR_A1 + PR_AX;
constexpr int PW = 4;                         // waves per block
constexpr int PR_LDS = PR_WF + PR_WH + PR_BIAS + PW * PR_WAVE;
static_assert(PR_LDS <= 160 * 1024, "proposal block exceeds the LDS of a CU");

// offset of 3^3 offset o in an S^3 map
template <int S>
__device__ __forceinline__ int odelta(int o) {
  return (o / 9) * S * S + ((o / 3) % 3) * S + o % 3;
}
// first input position of output position p of an O^3 map in its (O+2)^3 input
template <int O>
__device__ __forceinline__ int obase(int p) {
  return ((p / (O * O)) * (O + 2) + (p / O) % O) * (O + 2) + p % O;
}

// wrapped (periodic) or clamped inner index 0..n-1 of the 33 coordinates
// C - 5 .. C + 5 of the centre on each axis; bit 30 marks a coordinate outside
// a non-periodic box.  Lanes 0..32.
__device__ __forceinline__ void set_axes(const Feat& F, int e, int* ax, int lane) {
  if (lane >= 33) return;
  const int a = lane / 11, t = lane % 11;
  const int c = a == 0 ? e % F.ne : a == 1 ? (e / F.ne) % F.ne : e / (F.ne * F.ne);
  const int i = c + F.low1 - 5 + t;
  int m, out = 0;
  if (F.periodic) {
    m = (i - 1) % F.n;
    if (m < 0) m += F.n;
  } else {
    out = (i < 1 || i > F.n) ? 1 : 0;
    m = min(max(i, 1), F.n) - 1;
  }
  ax[lane] = m | (out << 30);
}

__device__ __forceinline__ size_t ax_index(const Feat& F, const int* ax, int hx, int hy, int hz) {
  const int m = (1 << 30) - 1;
  return (size_t)(ax[hx] & m) + (size_t)F.n * ((size_t)(ax[11 + hy] & m) + (size_t)F.n * (size_t)(ax[22 + hz] & m));
}

// the 1331 U1CO values around the centre (11^3, x fastest) and the 729
// DIVERCO values of the patch, lane-strided (lane l of load i: element
// 64 i + l), into registers; lanes past the end load the last element again
__device__ __forceinline__ void load_halo(const Feat& F, const int* ax, int lane, float (&pu)[21], float (&pd)[12]) {
  // every index first (they read the axes in LDS), then the loads
  const int m = (1 << 30) - 1;
  uint32_t idx[33];
#pragma unroll
  for (int i = 0; i < 21; ++i) {
    const int h = min(lane + 64 * i, 1330);
    idx[i] = (uint32_t)(ax[h % 11] & m) +
             (uint32_t)F.n * ((uint32_t)(ax[11 + (h / 11) % 11] & m) + (uint32_t)F.n * (uint32_t)(ax[22 + h / 121] & m));
  }
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    const int q = min(lane + 64 * i, 728);
    idx[21 + i] = (uint32_t)(ax[q % 9 + 1] & m) +
                  (uint32_t)F.n * ((uint32_t)(ax[12 + (q / 9) % 9] & m) + (uint32_t)F.n * (uint32_t)(ax[23 + q / 81] & m));
  }
  asm volatile("" ::: "memory");
#pragma unroll
  for (int i = 0; i < 21; ++i) pu[i] = F.u1co[idx[i]];
#pragma unroll
  for (int i = 0; i < 12; ++i) pd[i] = F.div[idx[21 + i]];
}

// B fragment of K step s (0..31 over conv1, conv2, conv3) for lane l: rows
// 32 k + 8 (l / 16) + v of the layer's W[rows][16], column l % 16; rows past
// the layer's K are zero
__device__ __forceinline__ v8s b_frag(const float* __restrict__ P, int s, int l) {
  int off, rows, ks;
  if (s < KS1) {
    off = OFF_W1, rows = 108, ks = s;
  } else if (s < KS1 + KS2) {
    off = OFF_W2, rows = 432, ks = s - KS1;
  } else {
    off = OFF_W3, rows = 432, ks = s - KS1 - KS2;
  }
  v8s f;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int r = 32 * ks + 8 * (l >> 4) + v;
    f[v] = (short)f2bf(r < rows ? P[off + r * 16 + (l & 15)] : 0.0f);
  }
  return f;
}

__device__ __forceinline__ v8s cat2(uint2 lo, uint2 hi) {
  uint4 u = {lo.x, lo.y, hi.x, hi.y};
  return __builtin_bit_cast(v8s, u);
}

// orders this wave's LDS accesses (the waves never read each other's regions)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int W>
__global__ void __launch_bounds__(64 * W) k_propose(const float* __restrict__ P, Feat F, int64_t ncand,
                                                    int8_t* prop) {
  F.inv_rms = F.rms_inv();
  __shared__ __attribute__((aligned(16))) unsigned char sm[PR_WF + PR_WH + PR_BIAS + W * PR_WAVE];
  const v8s* WF = (const v8s*)sm;
  uint16_t* WHT = (uint16_t*)(sm + PR_WF);
  float* BIAS = (float*)(sm + PR_WF + PR_WH);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  constexpr int WAVE0 = PR_WF + PR_WH + PR_BIAS;
  unsigned char* wb = sm + WAVE0 + wave * PR_WAVE;
  uint16_t* X0 = (uint16_t*)wb;
  uint16_t* A1 = (uint16_t*)(wb + PR_X0);
  float* U = (float*)(wb + PR_X0);
  float* D = U + PR_HU;
  int* AX = (int*)(wb + PR_X0 + PR_A1);

  for (int i = tid; i < KSALL * 64; i += 64 * W) ((v8s*)sm)[i] = b_frag(P, i >> 6, i & 63);
  for (int i = tid; i < NCLS * 432; i += 64 * W) WHT[i] = f2bf(P[OFF_WH + (i % 432) * NCLS + i / 432]);
  if (tid < 56)
    BIAS[tid] = tid < 16 ? P[OFF_B1 + tid] : tid < 32 ? P[OFF_B2 + tid - 16] : tid < 48 ? P[OFF_B3 + tid - 32]
                                                                                      : P[OFF_BH + tid - 48];
  const int n16 = lane & 15, q4 = lane >> 4;
  const int j8 = lane & 7, c8 = lane >> 3;
  // conv1's per-lane A offsets (bytes) of K step s, half h: offset
  // 8 s + 2 (l / 16) + h, clamped to a real one (its B rows are zero)
  int d9[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) d9[k] = odelta<9>(min(8 * (k >> 1) + 2 * q4 + (k & 1), 26)) * 8;
  // the patch cells of this lane: halo cell, patch axes (packed)
  int cell[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    const int qq = min(lane + 64 * i, 728);
    const int dx = qq % 9, dy = (qq / 9) % 9, dz = qq / 81;
    cell[i] = ((dx + 1) + 11 * (dy + 1) + 121 * (dz + 1)) | (dx << 11) | (dy << 15) | (dz << 19);
  }

  const int64_t stride = (int64_t)gridDim.x * W;
  int64_t t = (int64_t)blockIdx.x * W + wave;
  float pu[21], pd[12];
  // prologue: axes and inputs of the first candidate, the centre of the
  // second and the sorted index of the third.  In the loop these two are
  // loaded before the next candidate's inputs, so that no wait for one of
  // them waits for the inputs as well (loads complete in order).
  int e1 = 0, o2 = 0;
  if (t < ncand) set_axes(F, F.lin[F.order[t]], AX, lane);
  __syncthreads();  // the weights, and this wave's axes
  if (t + stride < ncand) e1 = F.lin[F.order[t + stride]];
  if (t + 2 * stride < ncand) o2 = F.order[t + 2 * stride];
  if (t < ncand) load_halo(F, AX, lane, pu, pd);

  for (int it = 0; t < ncand; ++it, t += stride) {
    const bool valid = true;
    const int* axc = AX + (it & 1) * 34;
    int* axn = AX + ((it + 1) & 1) * 34;
    const int64_t tn = t + stride;
    // this candidate's inputs to LDS; axes of the next one
#pragma unroll
    for (int i = 0; i < 21; ++i) U[lane + 64 * i] = pu[i];
#pragma unroll
    for (int i = 0; i < 12; ++i) D[lane + 64 * i] = pd[i];
    if (tn < ncand) set_axes(F, e1, axn, lane);
    wave_sync();
    // features of this candidate; the next one's inputs are loaded meanwhile
    if (valid) {
      const bool per = F.periodic != 0;
#pragma unroll 4
      for (int i = 0; i < 12; ++i) {
        const int qq = lane + 64 * i;
        if (qq < 729) {
          const int c = cell[i] & 2047;
          bool in = true;
          if (!per) {
            const int dx = (cell[i] >> 11) & 15, dy = (cell[i] >> 15) & 15, dz = cell[i] >> 19;
            in = !((axc[dx + 1] | axc[12 + dy] | axc[23 + dz]) >> 30);
          }
          uint2 w = {0u, 0u};
          if (in) {
            const float f0 = U[c] - 1.0f;
            const float f1 = D[qq] * F.inv_rms;
            const float gx = (U[c + 1] - U[c - 1]) * F.inv_two_dx;
            const float gy = (U[c + 11] - U[c - 11]) * F.inv_two_dx;
            const float gz = (U[c + 121] - U[c - 121]) * F.inv_two_dx;
            const float f2 = sqrtf(gx * gx + gy * gy + gz * gz) * F.inv_grad;
            w.x = bfpack(f0, f1);
            w.y = bfpack(f2, 0.0f);
          }
          *(uint2*)&X0[qq * 4] = w;
        }
      }
    }
    if (tn + stride < ncand) e1 = F.lin[o2];
    if (tn + 2 * stride < ncand) o2 = F.order[tn + 2 * stride];
    if (tn < ncand) load_halo(F, axn, lane, pu, pd);
    wave_sync();
    if (valid) {
      // conv1: 343 outputs (7^3) in row tiles 0..23, in pairs; two pairs per
      // iteration so the read buffers alternate
      const int wo = wave * PR_WAVE;
      const unsigned char* x0b = sm + WAVE0 + wo;
      uint16_t* a1w = (uint16_t*)(sm + WAVE0 + wo + PR_X0);
      v8s b1[KS1];
#pragma unroll
      for (int s = 0; s < KS1; ++s) b1[s] = WF[s * 64 + lane];
      const float bias1 = BIAS[n16];
      uint2 ra[16], rb[16];
      auto load1 = [&](uint2 (&r)[16], int pr) {
        const unsigned char* r0 = x0b + obase<7>(min(32 * pr + n16, 342)) * 8;
        const unsigned char* r1 = x0b + obase<7>(min(32 * pr + 16 + n16, 342)) * 8;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          r[k] = *(const uint2*)(r0 + d9[k]);
          r[8 + k] = *(const uint2*)(r1 + d9[k]);
        }
      };
      auto comp1 = [&](const uint2 (&r)[16], int pr) {
        v4f acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < KS1; ++s) {
          acc0 = mfma32(cat2(r[2 * s], r[2 * s + 1]), b1[s], acc0);
          acc1 = mfma32(cat2(r[8 + 2 * s], r[8 + 2 * s + 1]), b1[s], acc1);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row0 = 32 * pr + 4 * q4 + v;
          a1w[min(row0, 343) * 16 + n16] = bf1(fmaxf(acc0[v] + bias1, 0.0f));
          a1w[min(row0 + 16, 343) * 16 + n16] = bf1(fmaxf(acc1[v] + bias1, 0.0f));
        }
      };
      load1(ra, 0);
#pragma unroll 1
      for (int pr = 0; pr < 12; pr += 2) {
        load1(rb, pr + 1);
        comp1(ra, pr);
        load1(ra, pr + 2 < 12 ? pr + 2 : 11);
        comp1(rb, pr + 1);
      }
    }
    wave_sync();
    if (valid) {
      // conv2: 125 outputs (5^3) in 8 row tiles, 14 K steps, two per
      // iteration so the read buffers alternate
      const int wo = wave * PR_WAVE;
      const unsigned char* a1b = sm + WAVE0 + wo + PR_X0 + 16 * (q4 & 1);
      uint16_t* a2w = (uint16_t*)(sm + WAVE0 + wo);
      const float bias2 = BIAS[16 + n16];
      int rows[8];
#pragma unroll
      for (int mt = 0; mt < 8; ++mt) rows[mt] = obase<5>(min(mt * 16 + n16, 124)) * 32;
      v4f acc[8];
      v8s ca[8], cb[8], ba, bb;
      auto load2 = [&](v8s (&r)[8], v8s& b, int s) {
        b = WF[(KS1 + s) * 64 + lane];
        const unsigned char* p = a1b + odelta<7>(min(2 * s + (q4 >> 1), 26)) * 32;
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) r[mt] = *(const v8s*)(p + rows[mt]);
      };
#pragma unroll
      for (int mt = 0; mt < 8; ++mt) acc[mt] = {0.0f, 0.0f, 0.0f, 0.0f};
      load2(ca, ba, 0);
#pragma unroll 1
      for (int s = 0; s < KS2; s += 2) {
        load2(cb, bb, s + 1);
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) acc[mt] = mfma32(ca[mt], ba, acc[mt]);
        load2(ca, ba, s + 2 < KS2 ? s + 2 : KS2 - 1);
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) acc[mt] = mfma32(cb[mt], bb, acc[mt]);
      }
#pragma unroll
      for (int mt = 0; mt < 8; ++mt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = min(mt * 16 + 4 * q4 + v, 125);
          a2w[row * 16 + n16] = bf1(fmaxf(acc[mt][v] + bias2, 0.0f));
        }
    }
    wave_sync();
    if (valid) {
      // conv3: 27 outputs (3^3) in 2 row tiles, fully unrolled; K steps in
      // pairs into two partial sums per tile, the next pair's A read ahead
      const int wo = wave * PR_WAVE;
      const unsigned char* a2b = sm + WAVE0 + wo + 16 * (q4 & 1);
      uint16_t* a3w = (uint16_t*)(sm + WAVE0 + wo + PR_X0);
      const float bias3 = BIAS[32 + n16];
      const int r0 = obase<3>(min(n16, 26)) * 32, r1 = obase<3>(min(16 + n16, 26)) * 32;
      auto src3 = [&](int mt, int s) -> const v8s* {
        return (const v8s*)(a2b + (mt ? r1 : r0) + odelta<5>(min(2 * s + (q4 >> 1), 26)) * 32);
      };
      v4f acc[4];
      v8s cur[4], nxt[4], bc[2], bn[2];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        acc[k] = {0.0f, 0.0f, 0.0f, 0.0f};
        cur[k] = *src3(k >> 1, k & 1);
      }
      bc[0] = WF[(KS1 + KS2) * 64 + lane];
      bc[1] = WF[(KS1 + KS2 + 1) * 64 + lane];
#pragma unroll
      for (int s = 0; s < KS3; s += 2) {
        if (s + 2 < KS3) {
          bn[0] = WF[(KS1 + KS2 + s + 2) * 64 + lane];
          bn[1] = WF[(KS1 + KS2 + s + 3) * 64 + lane];
#pragma unroll
          for (int k = 0; k < 4; ++k) nxt[k] = *src3(k >> 1, s + 2 + (k & 1));
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] = mfma32(cur[k], bc[k & 1], acc[k]);
        if (s + 2 < KS3) {
          bc[0] = bn[0];
          bc[1] = bn[1];
#pragma unroll
          for (int k = 0; k < 4; ++k) cur[k] = nxt[k];
        }
      }
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = min(mt * 16 + 4 * q4 + v, 27);
          a3w[row * 16 + n16] = bf1(fmaxf(acc[2 * mt][v] + acc[2 * mt + 1][v] + bias3, 0.0f));
        }
    }
    wave_sync();
    if (valid) {
      // head: lane (class j8, chunk c8) sums 54 of the 432 features in order
      float sum = 0.0f;
      const uint32_t* a3 = (const uint32_t*)(A1 + c8 * 54);
      const uint32_t* wh = (const uint32_t*)(WHT + j8 * 432 + c8 * 54);
#pragma unroll 9
      for (int k = 0; k < 27; ++k) {
        const uint32_t av = a3[k], wv = wh[k];
        sum += bf2f((uint16_t)(av & 0xffff)) * bf2f((uint16_t)(wv & 0xffff));
        sum += bf2f((uint16_t)(av >> 16)) * bf2f((uint16_t)(wv >> 16));
      }
      sum += __shfl_xor(sum, 8);
      sum += __shfl_xor(sum, 16);
      sum += __shfl_xor(sum, 32);
      float best = lane < NCLS ? sum + BIAS[48 + j8] : -INFINITY;
      int bi = lane < NCLS ? lane : NCLS;
      for (int o = 1; o < 8; o <<= 1) {
        const float ob = __shfl_xor(best, o);
        const int oi = __shfl_xor(bi, o);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (lane == 0) prop[t] = (int8_t)bi;
    }
    wave_sync();
  }
}

inline unsigned nbk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

int make_feat(avism_ctx* c, Feat& F) {
  if (!c->fields_ready) AV_FAIL(c, AVISM_ERR_STATE, "fields not computed");
  const GridGeom& g = c->g;
  const int64_t ncell = (int64_t)g.n * g.n * g.n;
  // the sum stays on the device: the kernels take 1/RMS from it
  if (!c->cnn_work.ensure((size_t)(kDivSqBlocks + 1) * 8)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory");
  double* part = c->cnn_work.as<double>();
  k_div_sq<<<kDivSqBlocks, 256, 0, c->stream>>>(ncell, c->div.as<float>(), part);
  k_div_sq_sum<<<1, 256, 0, c->stream>>>(part, part + kDivSqBlocks);
  AV_CHECK_HIP(c, hipGetLastError());
  F.div_sq = part + kDivSqBlocks;
  F.ncell = (double)ncell;
  F.u1co = c->u1co.as<float>();
  F.div = c->div.as<float>();
  F.lin = c->cand_lin.as<int32_t>();
  F.order = c->misc_i32.as<int32_t>();
  F.n = g.n;
  F.low1 = g.low1;
  F.ne = g.next;
  F.periodic = g.periodic;
  F.inv_rms = 0.0f;  // set on the device from div_sq
  F.inv_two_dx = 1.0f / (2.0f * g.dx);
  F.inv_grad = c->vp.grad_thre > 0 ? 1.0f / c->vp.grad_thre : 1.0f;
  return AVISM_OK;
}

}  // namespace

int cnn_patches(avism_ctx* c, const int64_t* ranks, int64_t n, uint16_t* out) {
  if (n <= 0) return AVISM_OK;
  if (c->ncand <= 0) AV_FAIL(c, AVISM_ERR_STATE, "no candidates: run avism_find_voids first");
  Feat F;
  int rc = make_feat(c, F);
  if (rc) return rc;
  k_patches<<<(unsigned)std::min<int64_t>(n, 1 << 20), 64, 0, c->stream>>>(F, n, ranks, out);
  AV_CHECK_HIP(c, hipGetLastError());
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return AVISM_OK;
}

int cnn_forward(avism_ctx* c, const float* params, int64_t n, const uint16_t* patches, float* logits,
                uint16_t* cache) {
  if (n <= 0) return AVISM_OK;
  k_forward<<<(unsigned)std::min<int64_t>(n, 1 << 20), 64, 0, c->stream>>>(params, n, patches, logits, cache);
  AV_CHECK_HIP(c, hipGetLastError());
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return AVISM_OK;
}

int cnn_backward(avism_ctx* c, const float* params, int64_t n, const uint16_t* patches, const uint16_t* cache,
                 const float* logits, const int32_t* labels, float* grads, float* loss) {
  AV_CHECK_HIP(c, hipMemsetAsync(grads, 0, (size_t)AVISM_CNN_NPARAM * 4, c->stream));
  AV_CHECK_HIP(c, hipMemsetAsync(loss, 0, 4, c->stream));
  if (n > 0) {
    k_backward<<<(unsigned)std::min<int64_t>(n, 1 << 20), 64, 0, c->stream>>>(
        params, n, patches, cache, logits, labels, 1.0f / (float)n, grads, loss);
    AV_CHECK_HIP(c, hipGetLastError());
  }
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return AVISM_OK;
}

int cnn_adam(avism_ctx* c, float* params, const float* grads, float* m, float* v, float lr, int32_t step) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  k_adam<<<nbk(AVISM_CNN_NPARAM), 256, 0, c->stream>>>(AVISM_CNN_NPARAM, params, grads, m, v, lr, b1, b2,
                                                       eps, c1, c2);
  AV_CHECK_HIP(c, hipGetLastError());
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  return AVISM_OK;
}

// proposals for the current candidates, in processing order, into c->proposals
int cnn_propose(avism_ctx* c) {
  if (!c->cnn_ready) AV_FAIL(c, AVISM_ERR_STATE, "no network parameters: call avism_cnn_set_params");
  const int64_t nc = c->ncand;
  if (!c->proposals.ensure((size_t)nc + 1)) AV_FAIL(c, AVISM_ERR_HIP, "out of device memory (proposals)");
  if (nc == 0) return AVISM_OK;
  StageTimer t(c, "cnn_propose");
  Feat F;
  int rc = make_feat(c, F);
  if (rc) return rc;
  // one block per CU: the block's LDS holds the weights once for PW waves
  const unsigned grid = (unsigned)std::min<int64_t>((nc + PW - 1) / PW, 256);
  k_propose<PW><<<grid, 64 * PW, 0, c->stream>>>(c->cnn_params.as<float>(), F, nc, c->proposals.as<int8_t>());
  AV_CHECK_HIP(c, hipGetLastError());
  return AVISM_OK;
}

}  // namespace avism
