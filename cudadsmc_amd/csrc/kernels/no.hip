// Neural-operator kernels for gfx950: BF16 MFMA (v_mfma_f32_32x32x16_bf16)
// with FP32 accumulation.
//
// The operator that warm-starts particle loading is a stack of 3-D
// convolutions on the 64x32x32 simulation grid. Activations are BF16,
// channels-last ([B][Z][Y][X][C], C a multiple of 16); weights are BF16
// [taps][Cout][Cin] (K-contiguous per output channel) with FP32 master
// copies; gradients are FP32.
//
//  * conv_fwd: implicit GEMM, M = voxels, N = Cout, K = taps*Cin. Each wave
//    owns 32 voxels and all Cout (Cout/32 MFMA tiles); A fragments are
//    16-byte loads of 8 channels of the tap's neighbour voxel (zero outside
//    the grid), B fragments 16-byte LDS reads of the K-contiguous weights,
//    which the workgroup stages one tap ahead.
//    Epilogue: bias, optional ReLU, optional residual add, BF16 or FP32 out.
//  * the data gradient is conv_fwd on mirrored, transposed weights
//    (weight_flip), applied to the ReLU-masked output gradient.
//  * conv_bwd_w: per tap a GEMM M = Cout, N = Cin, K = voxels. A chunk of
//    voxels is staged transposed in LDS ([C][k], k contiguous) so both MFMA
//    operands are 16-byte LDS reads; each chunk writes an FP32 partial that
//    reduce_partials sums in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <cstdint>

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

namespace {

constexpr int kWaves = 4;
constexpr int kMaxNTiles = 4;   // Cout <= 128
constexpr int kChunk = 128;     // voxels staged per LDS round in conv_bwd_w (two per lane)
constexpr int kLdk = kChunk + 8;  // LDS row length of the staged operands
constexpr int kMaxC = 128;
constexpr int kWRegs = kMaxC * kMaxC / 8 / 256;  // 16-byte weight pieces per thread when staging a tap

__device__ __forceinline__ void tap_offset(int t, int taps, int dil, int& dz, int& dy, int& dx) {
  if (taps == 1) {
    dz = dy = dx = 0;
  } else {
    dz = (t / 9 - 1) * dil;
    dy = ((t / 3) % 3 - 1) * dil;
    dx = (t % 3 - 1) * dil;
  }
}

__device__ __forceinline__ float bf2f(bf16 v) { return (float)v; }

}  // namespace

struct NoConvArgs {
  const bf16* x;
  const bf16* w;
  const float* bias;
  const bf16* res;
  bf16* y;
  float* y32;
  bf16* h;
  int B, Z, Y, X, Cin, Cout, taps, dil, relu;
};

__global__ void __launch_bounds__(256) k_no_conv_fwd(NoConvArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int V = a.Z * a.Y * a.X;
  const long M = (long)a.B * V;
  const long m = ((long)blockIdx.x * kWaves + wave) * 32 + (lane & 31);
  const int kg = lane >> 5;  // which 8 of the 16 K values this lane supplies
  const int ntiles = a.Cout >> 5;
  f32x16 acc[kMaxNTiles];
#pragma unroll
  for (int j = 0; j < kMaxNTiles; ++j)
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[j][i] = 0.0f;
  int b = 0, z = 0, y = 0, x = 0;
  const bool valid = m < M;
  if (valid) {
    b = (int)(m / V);
    int r = (int)(m - (long)b * V);
    z = r / (a.Y * a.X);
    r -= z * a.Y * a.X;
    y = r / a.X;
    x = r - y * a.X;
  }
  // The workgroup's four waves share each tap's weights: they are staged in
  // LDS ([Cout][Cin + 8], two buffers), tap t + 1's copied in while tap t is
  // computed, so the B fragments are LDS reads instead of a per-wave reload
  // through L1.
  extern __shared__ __align__(16) bf16 sw[];
  const int ldw = a.Cin + 8;
  const int wchunks = a.Cout * (a.Cin >> 3);  // 16-byte pieces of one tap's weights
  bf16x8 wreg[kWRegs];
  auto fetch_w = [&](int t) {
    const bf16* src = a.w + (long)t * a.Cout * a.Cin;
#pragma unroll
    for (int r = 0; r < kWRegs; ++r) {
      const int q = threadIdx.x + r * 256;
      if (q < wchunks) wreg[r] = *reinterpret_cast<const bf16x8*>(src + (long)q * 8);
    }
  };
  auto store_w = [&](int buf) {
    bf16* dst = sw + (long)buf * a.Cout * ldw;
#pragma unroll
    for (int r = 0; r < kWRegs; ++r) {
      const int q = threadIdx.x + r * 256;
      if (q < wchunks) {
        const int co = q / (a.Cin >> 3), c8 = q - co * (a.Cin >> 3);
        *reinterpret_cast<bf16x8*>(dst + co * ldw + c8 * 8) = wreg[r];
      }
    }
  };
  fetch_w(0);
  store_w(0);
  __syncthreads();
  for (int t = 0; t < a.taps; ++t) {
    if (t + 1 < a.taps) fetch_w(t + 1);
    int dz, dy, dx;
    tap_offset(t, a.taps, a.dil, dz, dy, dx);
    const int zz = z + dz, yy = y + dy, xx = x + dx;
    const bool in = valid && zz >= 0 && zz < a.Z && yy >= 0 && yy < a.Y && xx >= 0 && xx < a.X;
    const bf16* xp = a.x + ((((long)b * a.Z + zz) * a.Y + yy) * a.X + xx) * a.Cin;
    const bf16* wl = sw + (long)(t & 1) * a.Cout * ldw;
    for (int c0 = 0; c0 < a.Cin; c0 += 16) {
      bf16x8 av;
      if (in) {
        av = *reinterpret_cast<const bf16x8*>(xp + c0 + kg * 8);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) av[e] = (bf16)0.0f;
      }
#pragma unroll
      for (int j = 0; j < kMaxNTiles; ++j) {
        if (j < ntiles) {
          const int co = j * 32 + (lane & 31);
          const bf16x8 bv = *reinterpret_cast<const bf16x8*>(wl + co * ldw + c0 + kg * 8);
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[j], 0, 0, 0);
        }
      }
    }
    if (t + 1 < a.taps) store_w((t + 1) & 1);  // that buffer was last read in tap t - 1
    __syncthreads();
  }
  // epilogue: lane holds D[row][col], col = lane % 32,
  // row = 8 * (i / 4) + 4 * (lane / 32) + i % 4
  const long mbase = ((long)blockIdx.x * kWaves + wave) * 32;
#pragma unroll
  for (int j = 0; j < kMaxNTiles; ++j) {
    if (j >= ntiles) continue;
    const int co = j * 32 + (lane & 31);
    const float bias = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const long mr = mbase + 8 * (i / 4) + 4 * kg + (i % 4);
      if (mr >= M) continue;
      float v = acc[j][i] + bias;
      if (a.relu) v = v > 0.0f ? v : 0.0f;
      const long o = mr * a.Cout + co;
      if (a.h) a.h[o] = (bf16)v;
      if (a.res) v += bf2f(a.res[o]);
      if (a.y32) a.y32[o] = v;
      if (a.y) a.y[o] = (bf16)v;
    }
  }
}

// wb[t'][ci][co] = w[t][co][ci] with t' the mirrored tap (offset -d): the
// data gradient of conv(x, w) is conv(dy, wb).
__global__ void k_no_weight_flip(const bf16* w, bf16* wb, int taps, int Cout, int Cin) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long n = (long)taps * Cout * Cin;
  if (i >= n) return;
  int ci = (int)(i % Cin);
  int co = (int)((i / Cin) % Cout);
  int t = (int)(i / ((long)Cin * Cout));
  int tm = taps - 1 - t;  // (dz,dy,dx) -> (-dz,-dy,-dx)
  wb[((long)tm * Cin + ci) * Cout + co] = w[i];
}

struct NoBwdWArgs {
  const bf16* x;   // conv input [M][Cin]
  const bf16* dy;  // output gradient [M][Cout]
  float* part;     // [n_chunks][taps][Cout][Cin]
  int B, Z, Y, X, Cin, Cout, taps, dil, chunk;  // voxels per block
};

__global__ void __launch_bounds__(256) k_no_conv_bwd_w(NoBwdWArgs a) {
  // dy and the shifted x of kChunk voxels, transposed: [Cout][kLdk] then
  // [Cin][kLdk] (dynamic LDS sized by the call's channel counts)
  extern __shared__ __align__(16) bf16 sbw[];
  bf16* const sdy = sbw;
  bf16* const sx = sbw + (long)a.Cout * kLdk;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kg = lane >> 5;
  const int t = blockIdx.y;
  const int V = a.Z * a.Y * a.X;
  const long M = (long)a.B * V;
  const long m0 = (long)blockIdx.x * a.chunk;
  int dz, dyo, dx;
  tap_offset(t, a.taps, a.dil, dz, dyo, dx);
  const int tm = a.Cout >> 5, tn = a.Cin >> 5;
  const int ntile = tm * tn;
  f32x16 acc[kMaxNTiles];
#pragma unroll
  for (int j = 0; j < kMaxNTiles; ++j)
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[j][i] = 0.0f;
  const int cgy = a.Cout / 8, cgx = a.Cin / 8;
  const long m_end = m0 + a.chunk < M ? m0 + a.chunk : M;  // this workgroup's voxels: [m0, m_end)
  for (long k0 = m0; k0 < m_end; k0 += kChunk) {
    __syncthreads();
    // stage dy and the shifted x of kChunk voxels, transposed to [c][k]:
    // lane l takes voxels 2l and 2l + 1, so a channel's two values form one
    // 32-bit LDS word and consecutive lanes write consecutive words (no bank
    // conflicts); each lane works out its voxels' sources once, and the
    // waves share out the 8-channel groups
    {
      static_assert(kChunk == 128, "two staged voxels per lane of a wavefront");
      long mm[2], src[2];  // the voxels, and the voxels x is read from for this tap (-1: outside)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        mm[h] = k0 + 2 * lane + h;
        src[h] = -1;
        if (mm[h] < m_end) {
          const int b = (int)(mm[h] / V);
          int r = (int)(mm[h] - (long)b * V);
          const int z = r / (a.Y * a.X);
          r -= z * a.Y * a.X;
          const int y = r / a.X;
          const int x = r - y * a.X;
          const int zz = z + dz, yy = y + dyo, xx = x + dx;
          if (zz >= 0 && zz < a.Z && yy >= 0 && yy < a.Y && xx >= 0 && xx < a.X)
            src[h] = (((long)b * a.Z + zz) * a.Y + yy) * a.X + xx;
        }
      }
      for (int g = wave; g < cgy + cgx; g += kWaves) {
        bf16x8 v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
          for (int e = 0; e < 8; ++e) v[h][e] = (bf16)0.0f;
          if (g < cgy) {
            if (mm[h] < m_end) v[h] = *reinterpret_cast<const bf16x8*>(a.dy + mm[h] * a.Cout + g * 8);
          } else if (src[h] >= 0) {
            v[h] = *reinterpret_cast<const bf16x8*>(a.x + src[h] * a.Cin + (g - cgy) * 8);
          }
        }
        bf16* row = g < cgy ? sdy + (long)(g * 8) * kLdk : sx + (long)((g - cgy) * 8) * kLdk;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          bf16x2 pr;
          pr[0] = v[0][e];
          pr[1] = v[1][e];
          *reinterpret_cast<bf16x2*>(row + (long)e * kLdk + 2 * lane) = pr;
        }
      }
    }
    __syncthreads();
    // tiles assigned round-robin to the waves
#pragma unroll
    for (int j = 0; j < kMaxNTiles; ++j) {
      const int tile = wave + j * kWaves;
      if (tile >= ntile) continue;
      const int co = (tile / tn) * 32 + (lane & 31);
      const int ci = (tile % tn) * 32 + (lane & 31);
#pragma unroll
      for (int kc = 0; kc < kChunk; kc += 16) {
        bf16x8 av = *reinterpret_cast<const bf16x8*>(sdy + (long)co * kLdk + kc + kg * 8);
        bf16x8 bv = *reinterpret_cast<const bf16x8*>(sx + (long)ci * kLdk + kc + kg * 8);
        acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[j], 0, 0, 0);
      }
    }
  }
  float* out = a.part + ((long)blockIdx.x * a.taps + t) * a.Cout * a.Cin;
#pragma unroll
  for (int j = 0; j < kMaxNTiles; ++j) {
    const int tile = wave + j * kWaves;
    if (tile >= ntile) continue;
    const int ci = (tile % tn) * 32 + (lane & 31);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int co = (tile / tn) * 32 + 8 * (i / 4) + 4 * kg + (i % 4);
      out[(long)co * a.Cin + ci] = acc[j][i];
    }
  }
}

// out[i] (+)= sum over chunks of part[c][i], fixed order.
__global__ void k_no_reduce_partials(const float* part, long n, int chunks, float* out, int accumulate) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int c = 0; c < chunks; ++c) s += part[(long)c * n + i];
  out[i] = accumulate ? out[i] + s : s;
}

// db[co] partial sums: workgroup b sums voxels [b * kBiasChunk, (b + 1) *
// kBiasChunk) of every channel into part[b][co]; each thread takes channel
// co = 32 j + (t mod 32) and every eighth voxel (coalesced 64-byte rows),
// and the eight voxel lanes are added in a fixed order.
constexpr int kBiasChunk = 512;

__global__ void __launch_bounds__(256) k_no_bias_grad(const bf16* dy, long M, int Cout, float* part) {
  __shared__ float red[8][kMaxC];
  const int t = threadIdx.x, cl = t & 31, vl = t >> 5;
  const long m0 = (long)blockIdx.x * kBiasChunk;
  for (int j = 0; j < Cout / 32; ++j) {
    const int co = j * 32 + cl;
    float s = 0.0f;
#pragma unroll 8
    for (int i = vl; i < kBiasChunk; i += 8) {
      const long m = m0 + i;
      if (m < M) s += bf2f(dy[m * Cout + co]);
    }
    red[vl][co] = s;
  }
  __syncthreads();
  if (t < Cout) {
    float s = 0.0f;
    for (int v = 0; v < 8; ++v) s += red[v][t];
    part[(long)blockIdx.x * Cout + t] = s;
  }
}

// out[i] (+)= sum over c of part[c][i] for many partial rows: a workgroup
// per output, the rows split over its threads and added in a fixed order.
__global__ void __launch_bounds__(256) k_no_reduce_rows(const float* part, int n, int rows, float* out,
                                                        int accumulate) {
  __shared__ float red[256];
  const int i = blockIdx.x, t = threadIdx.x;
  float s = 0.0f;
  for (int c = t; c < rows; c += 256) s += part[(long)c * n + i];
  red[t] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) out[i] = accumulate ? out[i] + red[0] : red[0];
}

// dz = dy * (h > 0)  (ReLU backward through the stored activation)
__global__ void k_no_relu_bwd(const bf16* dy, const bf16* h, bf16* dz, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dz[i] = bf2f(h[i]) > 0.0f ? dy[i] : (bf16)0.0f;
}

// c = a + b
__global__ void k_no_add(const bf16* a, const bf16* b, bf16* c, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  c[i] = (bf16)(bf2f(a[i]) + bf2f(b[i]));
}

// Masked mean-squared error of pred [M][C] (FP32, only the first Cf channels
// are fields) against target [M][Cf]; dpred = 2 (pred - target) mask / count,
// written as BF16 [M][C] (zero in the padding channels); per-block loss sums.
__global__ void k_no_mse(const float* pred, const float* target, const float* mask, long M, int C, int Cf,
                         float inv_count, bf16* dpred, float* loss_part) {
  __shared__ float red[256];
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < M * C) {
    long m = i / C;
    int c = (int)(i - m * C);
    float g = 0.0f;
    if (c < Cf) {
      float w = mask ? mask[m] : 1.0f;
      float d = (pred[i] - target[m * Cf + c]) * w;
      l = d * d * inv_count;
      g = 2.0f * d * inv_count;
    }
    dpred[i] = (bf16)g;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// Adam on FP32 master weights, BF16 copy refreshed.
__global__ void k_no_adam(float* w, const float* g, float* m1, float* m2, bf16* wbf, long n, float lr, float b1,
                          float b2, float eps, float c1, float c2, float wd) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i] + wd * w[i];
  float a = b1 * m1[i] + (1.0f - b1) * gi;
  float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  float upd = (a / c1) / (sqrtf(b / c2) + eps);
  float nw = w[i] - lr * upd;
  w[i] = nw;
  if (wbf) wbf[i] = (bf16)nw;
}

__global__ void k_no_f32_to_bf16(const float* a, bf16* b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (bf16)a[i];
}

// Per-voxel input features of the operator: x, y, z in [-1, 1], signed
// distance to the sphere surface in units of the box height, and the five
// normalised freestream/body parameters broadcast (p[b][5]); channels
// 9..C-1 are zero. sphere[b] = {cx, cy, cz, r} in metres, box = {xmin, ymin,
// zmin, xmax, ymax, zmax}.
__global__ void k_no_inputs(const float* p, const float* sphere, const float* box, int B, int Z, int Y, int X, int C,
                            bf16* out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long V = (long)Z * Y * X;
  if (i >= (long)B * V) return;
  int b = (int)(i / V);
  int r = (int)(i - (long)b * V);
  int z = r / (Y * X);
  r -= z * Y * X;
  int y = r / X;
  int x = r - y * X;
  float fx = (x + 0.5f) / X, fy = (y + 0.5f) / Y, fz = (z + 0.5f) / Z;
  float px = box[0] + fx * (box[3] - box[0]);
  float py = box[1] + fy * (box[4] - box[1]);
  float pz = box[2] + fz * (box[5] - box[2]);
  const float* s = sphere + 4 * b;
  float dx = px - s[0], dy = py - s[1], dz = pz - s[2];
  float sd = (sqrtf(dx * dx + dy * dy + dz * dz) - s[3]) / (box[4] - box[1]);
  bf16* o = out + i * C;
  float f[9] = {2.0f * fx - 1.0f, 2.0f * fy - 1.0f, 2.0f * fz - 1.0f, sd, p[5 * b], p[5 * b + 1], p[5 * b + 2],
                p[5 * b + 3], p[5 * b + 4]};
  for (int c = 0; c < C; ++c) o[c] = (bf16)(c < 9 ? f[c] : 0.0f);
}

// ------------------------------------------------------------------ C ABI
static inline unsigned nb(long n, int t) { return (unsigned)((n + t - 1) / t); }
#define NO_EMPTY(n) \
  if ((n) <= 0) return 0

extern "C" {

int dsmc_no_conv_fwd(const void* x, const void* w, const float* bias, const void* res, void* y, float* y32,
                     void* h, int B, int Z, int Y, int X, int Cin, int Cout, int taps, int dil, int relu,
                     void* stream) {
  if (Cin % 16 || Cin > kMaxC || Cout % 32 || Cout > 32 * kMaxNTiles || (taps != 1 && taps != 27) || B <= 0)
    return -1;
  NoConvArgs a{(const bf16*)x, (const bf16*)w, bias, (const bf16*)res, (bf16*)y, y32, (bf16*)h,
               B, Z, Y, X, Cin, Cout, taps, dil, relu};
  long M = (long)B * Z * Y * X;
  NO_EMPTY(M);
  const size_t lds = 2ull * Cout * (Cin + 8) * sizeof(bf16);  // two taps of weights
  hipLaunchKernelGGL(k_no_conv_fwd, dim3(nb(M, 32 * kWaves)), dim3(256), lds, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int dsmc_no_weight_flip(const void* w, void* wb, int taps, int Cout, int Cin, void* stream) {
  long n = (long)taps * Cout * Cin;
  NO_EMPTY(n);
  hipLaunchKernelGGL(k_no_weight_flip, dim3(nb(n, 256)), dim3(256), 0, (hipStream_t)stream, (const bf16*)w,
                     (bf16*)wb, taps, Cout, Cin);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int dsmc_no_conv_bwd_w(const void* x, const void* dy, float* dw, float* part, long part_floats, int B, int Z,
                       int Y, int X, int Cin, int Cout, int taps, int dil, int chunk, int accumulate, void* stream) {
  if (Cin % 32 || Cout % 32 || Cin > kMaxC || Cout > kMaxC || chunk <= 0 || chunk % 64 || (taps != 1 && taps != 27))
    return -1;
  if ((Cout / 32) * (Cin / 32) > kWaves * kMaxNTiles) return -1;
  long M = (long)B * Z * Y * X;
  NO_EMPTY(M);
  int chunks = (int)nb(M, chunk);
  long n = (long)taps * Cout * Cin;
  if ((long)chunks * n > part_floats) return -5;
  NoBwdWArgs a{(const bf16*)x, (const bf16*)dy, part, B, Z, Y, X, Cin, Cout, taps, dil, chunk};
  const size_t lds = (size_t)(Cout + Cin) * kLdk * sizeof(bf16);
  hipLaunchKernelGGL(k_no_conv_bwd_w, dim3(chunks, taps), dim3(256), lds, (hipStream_t)stream, a);
  hipLaunchKernelGGL(k_no_reduce_partials, dim3(nb(n, 256)), dim3(256), 0, (hipStream_t)stream, part, n, chunks,
                     dw, accumulate);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int dsmc_no_bias_grad(const void* dy, long M, int Cout, float* db, float* part, long part_floats, int accumulate,
                      void* stream) {
  NO_EMPTY(M);
  if (Cout % 32 || Cout > kMaxC) return -1;
  int chunks = (int)nb(M, kBiasChunk);
  if ((long)chunks * Cout > part_floats) return -5;
  hipLaunchKernelGGL(k_no_bias_grad, dim3(chunks), dim3(256), 0, (hipStream_t)stream, (const bf16*)dy, M, Cout,
                     part);
  hipLaunchKernelGGL(k_no_reduce_rows, dim3(Cout), dim3(256), 0, (hipStream_t)stream, part, Cout, chunks, db,
                     accumulate);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int dsmc_no_relu_bwd(const void* dy, const void* h, void* dz, long n, void* stream) {
  NO_EMPTY(n);
  hipLaunchKernelGGL(k_no_relu_bwd, dim3(nb(n, 256)), dim3(256), 0, (hipStream_t)stream, (const bf16*)dy,
                     (const bf16*)h, (bf16*)dz, n);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int dsmc_no_add(const void* a, const void* b, void* c, long n, void* stream) {
  NO_EMPTY(n);
  hipLaunchKernelGGL(k_no_add, dim3(nb(n, 256)), dim3(256), 0, (hipStream_t)stream, (const bf16*)a,
                     (const bf16*)b, (bf16*)c, n);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int dsmc_no_mse(const float* pred, const float* target, const float* mask, long M, int C, int Cf, float inv_count,
                void* dpred, float* loss_part, void* stream) {
  NO_EMPTY(M * C);
  hipLaunchKernelGGL(k_no_mse, dim3(nb(M * C, 256)), dim3(256), 0, (hipStream_t)stream, pred, target, mask, M, C,
                     Cf, inv_count, (bf16*)dpred, loss_part);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int dsmc_no_adam(float* w, const float* g, float* m1, float* m2, void* wbf, long n, float lr, float b1, float b2,
                 float eps, float c1, float c2, float wd, void* stream) {
  NO_EMPTY(n);
  hipLaunchKernelGGL(k_no_adam, dim3(nb(n, 256)), dim3(256), 0, (hipStream_t)stream, w, g, m1, m2, (bf16*)wbf, n,
                     lr, b1, b2, eps, c1, c2, wd);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int dsmc_no_f32_to_bf16(const float* a, void* b, long n, void* stream) {
  NO_EMPTY(n);
  hipLaunchKernelGGL(k_no_f32_to_bf16, dim3(nb(n, 256)), dim3(256), 0, (hipStream_t)stream, a, (bf16*)b, n);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int dsmc_no_inputs(const float* p, const float* sphere, const float* box, int B, int Z, int Y, int X, int C,
                   void* out, void* stream) {
  if (C < 9) return -1;
  long n = (long)B * Z * Y * X;
  NO_EMPTY(n);
  hipLaunchKernelGGL(k_no_inputs, dim3(nb(n, 256)), dim3(256), 0, (hipStream_t)stream, p, sphere, box, B, Z, Y, X,
                     C, (bf16*)out);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // extern "C"
