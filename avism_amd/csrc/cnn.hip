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
__global__ void k_div_sq(int64_t n, const float* __restrict__ div, double* out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double v = div[i];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

// patches [n][3][729] bf16 (channel, z, y, x) of the candidates at ranks idx[]
__global__ void __launch_bounds__(64) k_patches(Feat F, int64_t n, const int64_t* __restrict__ idx,
                                                uint16_t* out) {
  __shared__ __attribute__((aligned(16))) uint16_t X0[729 * 4];
  const int lane = threadIdx.x;
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

// proposal of every candidate: features, forward, argmax (first maximum wins).
// Two waves per SIMD: the register cap spills 160 bytes per lane of weight
// fragments, which costs less than the latency one wave per SIMD leaves
// exposed (5.6 -> 4.0 ms at 123,455 candidates; three waves, with 540 bytes
// spilled, took 5.0 ms)
constexpr int FWD_LDS = 5840 + 352 * 16 * 2;

__global__ void __launch_bounds__(64, 2) k_propose(const float* __restrict__ P, Feat F, int64_t ncand,
                                                int8_t* prop) {
  __shared__ __attribute__((aligned(16))) unsigned char sm[FWD_LDS];
  uint16_t* X0 = (uint16_t*)sm;
  uint16_t* A1 = (uint16_t*)(sm + 5840);
  const int lane = threadIdx.x;
  Frags Fr;
  load_frags(P, Fr, lane);
  for (int64_t t = blockIdx.x; t < ncand; t += gridDim.x) {
    build_patch(F, t, X0, lane);
    __syncthreads();
    const float lg = forward_patch(P, Fr, X0, A1, X0, A1, lane);
    float best = lane < NCLS ? lg : -INFINITY;
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
    __syncthreads();
  }
}

inline unsigned nbk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

int make_feat(avism_ctx* c, Feat& F) {
  if (!c->fields_ready) AV_FAIL(c, AVISM_ERR_STATE, "fields not computed");
  const GridGeom& g = c->g;
  const int64_t ncell = (int64_t)g.n * g.n * g.n;
  double* d = (double*)((char*)c->scalars.p + 3072);
  AV_CHECK_HIP(c, hipMemsetAsync(d, 0, 8, c->stream));
  k_div_sq<<<1024, 256, 0, c->stream>>>(ncell, c->div.as<float>(), d);
  AV_CHECK_HIP(c, hipGetLastError());
  double s = 0.0;
  AV_CHECK_HIP(c, hipMemcpyAsync(&s, d, 8, hipMemcpyDeviceToHost, c->stream));
  AV_CHECK_HIP(c, hipStreamSynchronize(c->stream));
  const double rms = sqrt(s / (double)ncell);
  F.u1co = c->u1co.as<float>();
  F.div = c->div.as<float>();
  F.lin = c->cand_lin.as<int32_t>();
  F.order = c->misc_i32.as<int32_t>();
  F.n = g.n;
  F.low1 = g.low1;
  F.ne = g.next;
  F.periodic = g.periodic;
  F.inv_rms = rms > 0 ? (float)(1.0 / rms) : 0.0f;
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
  const unsigned grid = (unsigned)std::min<int64_t>(nc, 16384);
  k_propose<<<grid, 64, 0, c->stream>>>(c->cnn_params.as<float>(), F, nc, c->proposals.as<int8_t>());
  AV_CHECK_HIP(c, hipGetLastError());
  return AVISM_OK;
}

}  // namespace avism
