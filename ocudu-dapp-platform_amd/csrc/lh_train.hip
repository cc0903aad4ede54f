// FP32 forward and backward passes of the learned (neural) min-sum decoder, for gfx950.
//
// Model: the layered normalised-offset min-sum decoder unrolled over T iterations, with
// one scale alpha[t][l] and one offset beta[t][l] per iteration t and layer l.  For check
// row r of layer l (edges e with variable-node soft values s at the cyclically shifted
// positions of the lifted graph):
//   v_e   = s_e - c_e(t-1)                               (c(-1) = 0)
//   c_e(t) = sgn_e * relu(alpha[t][l] * mag_e - beta[t][l]),
//            mag_e = min over the other edges of |v|, sgn_e = product of their signs
//   s_e  <- v_e + c_e(t)
// Soft values start at the clamped channel values ([-64, 64], two punctured nodes at 0) as
// in the int8 decoder, and are not saturated.  The loss is the mean binary cross-entropy of
// the message bits after every iteration, with logits tau * s and weight w[t].
//
// One workgroup per codeblock, one thread per check row; the soft values (forward) and
// their gradients (backward) live in LDS.  The forward pass stores every v_e in a
// per-codeblock workspace; the backward pass walks the layers in reverse and recomputes
// each row's minima from them.  Workspace of a codeblock (floats):
//   hist[T][E_L][Z]   v_e per iteration, edge (layer-ordered) and row
//   gl[T][K_b Z]      d loss_t / d s of the message bits
//   gnext[E_L][Z]     gradient of the next iteration's check messages
// with E_L the number of base-graph edges of the first L layers.
#include <hip/hip_runtime.h>

#include "bg_tables.inc"
#include "lh_internal.h"

#pragma clang fp contract(off)

namespace lh {

namespace {

template <int BG>
struct tbg;
template <>
struct tbg<1> {
  static constexpr int K = 22, M = 46, DMAX = 19;
  __device__ static int rs(int i) { return lh_bg1_row_start[i]; }
};
template <>
struct tbg<2> {
  static constexpr int K = 10, M = 42, DMAX = 10;
  __device__ static int rs(int i) { return lh_bg2_row_start[i]; }
};

// Two smallest magnitudes of a row (first occurrence wins ties, as in the int8 decoder),
// their edges and signs, and the parity of the negative messages.
struct row_stats {
  float m1, m2;
  int a1, a2;
  float s1, s2; // signs (+-1) of v at a1, a2
  int neg;      // parity of negative v
};

template <int DMAX>
__device__ __forceinline__ row_stats stats_of(const float (&v)[DMAX], int deg)
{
  row_stats st{3.0e38f, 3.0e38f, 0, 0, 1.0f, 1.0f, 0};
#pragma unroll
  for (int e = 0; e < DMAX; ++e) {
    if (e < deg) {
      const float a = fabsf(v[e]);
      const float sg = v[e] < 0.0f ? -1.0f : 1.0f;
      if (a < st.m1) {
        st.m2 = st.m1;
        st.a2 = st.a1;
        st.s2 = st.s1;
        st.m1 = a;
        st.a1 = e;
        st.s1 = sg;
      } else if (a < st.m2) {
        st.m2 = a;
        st.a2 = e;
        st.s2 = sg;
      }
      st.neg ^= v[e] < 0.0f ? 1 : 0;
    }
  }
  return st;
}

// Sign of the check message on edge e: product of the other edges' signs.
__device__ __forceinline__ float msg_sign(const row_stats& st, float v)
{
  const int n = st.neg ^ (v < 0.0f ? 1 : 0);
  return n ? -1.0f : 1.0f;
}

__device__ __forceinline__ float wave_sum(float x)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    x += __shfl_xor(x, o);
  }
  return x;
}

} // namespace

template <int BG>
__global__ void __launch_bounds__(kMaxThreads) k_nms_train(const train_job* __restrict__ jobs, int T,
                                                            const float* __restrict__ alpha,
                                                            const float* __restrict__ beta,
                                                            const float* __restrict__ wt, float tau,
                                                            const int8_t* __restrict__ in_base,
                                                            const uint8_t* __restrict__ tgt_base,
                                                            float* __restrict__ ws_base, float* __restrict__ loss_out,
                                                            float* __restrict__ galpha, float* __restrict__ gbeta,
                                                            const uint32_t* __restrict__ gtab, int backward)
{
  using B = tbg<BG>;
  constexpr int K = B::K, M = B::M, DMAX = B::DMAX;
  extern __shared__ __align__(16) float buf[]; // (K + L) * Z: soft values, then their gradients
  __shared__ float gacc[2 * kLearnedIters * M];
  __shared__ float red[8];
  const train_job jb = jobs[blockIdx.x];
  const int Z = jb.Z, L = jb.L, t0 = (int)threadIdx.x, NT = (int)blockDim.x;
  const bool row = t0 < Z;
  const int r = t0;
  const int EL = B::rs(L);
  const int KZ = K * Z;
  const int nbits = KZ - jb.F;
  const uint32_t* gt = gtab + jb.gt_off;
  float* hist = ws_base + jb.ws_off;
  float* gl = hist + (size_t)T * EL * Z;
  float* gnext = gl + (size_t)T * KZ;
  const int8_t* in = in_base + jb.in_off;
  const uint8_t* tgt = tgt_base + jb.tgt_off;

  for (int i = t0; i < 2 * kLearnedIters * M; i += NT) {
    gacc[i] = 0.0f;
  }
  // Soft values: two punctured nodes at zero, the rest clamped to [-64, 64].
  const int nsoft = (K + L) * Z;
  for (int i = t0; i < nsoft; i += NT) {
    float x = 0.0f;
    if (i >= 2 * Z && i - 2 * Z < (int)jb.n_in) {
      x = (float)min(max((int)in[i - 2 * Z], -64), 64);
    }
    buf[i] = x;
  }
  __syncthreads();

  // ---- forward ----
  float loss = 0.0f;
  for (int t = 0; t < T; ++t) {
    for (int l = 0; l < L; ++l) {
      if (row) {
        const int e0 = B::rs(l), deg = B::rs(l + 1) - e0;
        int addr[DMAX];
        float v[DMAX];
#pragma unroll
        for (int e = 0; e < DMAX; ++e) {
          if (e < deg) {
            const uint32_t g = gt[e0 + e];
            int j = r + (int)(g >> 16);
            j = j >= Z ? j - Z : j;
            addr[e] = (int)(g & 0xffffu) + j;
          }
        }
        if (t > 0) {
          const float* hp = hist + ((size_t)(t - 1) * EL + e0) * Z + r;
          float pv[DMAX];
#pragma unroll
          for (int e = 0; e < DMAX; ++e) {
            pv[e] = e < deg ? hp[(size_t)e * Z] : 0.0f;
          }
          const row_stats ps = stats_of<DMAX>(pv, deg);
          const float a = alpha[(t - 1) * M + l], b = beta[(t - 1) * M + l];
#pragma unroll
          for (int e = 0; e < DMAX; ++e) {
            if (e < deg) {
              const float mag = e == ps.a1 ? ps.m2 : ps.m1;
              const float c = msg_sign(ps, pv[e]) * fmaxf(a * mag - b, 0.0f);
              v[e] = buf[addr[e]] - c;
            }
          }
        } else {
#pragma unroll
          for (int e = 0; e < DMAX; ++e) {
            if (e < deg) {
              v[e] = buf[addr[e]];
            }
          }
        }
        float* hq = hist + ((size_t)t * EL + e0) * Z + r;
#pragma unroll
        for (int e = 0; e < DMAX; ++e) {
          if (e < deg) {
            hq[(size_t)e * Z] = v[e];
          }
        }
        const row_stats st = stats_of<DMAX>(v, deg);
        const float a = alpha[t * M + l], b = beta[t * M + l];
#pragma unroll
        for (int e = 0; e < DMAX; ++e) {
          if (e < deg) {
            const float mag = e == st.a1 ? st.m2 : st.m1;
            const float c = msg_sign(st, v[e]) * fmaxf(a * mag - b, 0.0f);
            buf[addr[e]] = v[e] + c;
          }
        }
      }
      __syncthreads();
    }
    // Loss of iteration t and its gradient with respect to the message soft values.
    const float w = wt[t] / (float)nbits;
    for (int j = t0; j < nbits; j += NT) {
      const float y = (float)((tgt[j >> 3] >> (7 - (j & 7))) & 1u);
      const float sgn = 1.0f - 2.0f * y;
      const float z = sgn * buf[j] * tau; // > 0: correct decision
      const float sp = fmaxf(-z, 0.0f) + log1pf(expf(-fabsf(z)));
      loss += w * sp;
      gl[(size_t)t * KZ + j] = -w * sgn * tau / (1.0f + expf(z));
    }
    __syncthreads(); // the next iteration updates the soft values read here
  }
  float ls = wave_sum(loss);
  if ((t0 & 63) == 0) {
    red[t0 >> 6] = ls;
  }
  __syncthreads();
  if (t0 == 0) {
    float s = 0.0f;
    for (int i = 0; i < NT / 64; ++i) {
      s += red[i];
    }
    loss_out[blockIdx.x] = s;
  }
  if (!backward) {
    return;
  }

  // ---- backward ----
  for (int i = t0; i < nsoft; i += NT) {
    buf[i] = 0.0f;
  }
  __syncthreads();
  for (int t = T - 1; t >= 0; --t) {
    for (int j = t0; j < nbits; j += NT) {
      buf[j] += gl[(size_t)t * KZ + j];
    }
    __syncthreads();
    for (int l = L - 1; l >= 0; --l) {
      float ga = 0.0f, gb = 0.0f;
      if (row) {
        const int e0 = B::rs(l), deg = B::rs(l + 1) - e0;
        int addr[DMAX];
        float v[DMAX], gv[DMAX];
        const float* hq = hist + ((size_t)t * EL + e0) * Z + r;
        float* gn = gnext + (size_t)e0 * Z + r;
#pragma unroll
        for (int e = 0; e < DMAX; ++e) {
          if (e < deg) {
            const uint32_t g = gt[e0 + e];
            int j = r + (int)(g >> 16);
            j = j >= Z ? j - Z : j;
            addr[e] = (int)(g & 0xffffu) + j;
            v[e] = hq[(size_t)e * Z];
          }
        }
        const row_stats st = stats_of<DMAX>(v, deg);
        const float a = alpha[t * M + l], b = beta[t * M + l];
        float g1 = 0.0f, g2 = 0.0f; // gradients reaching |v| at a1 and a2
#pragma unroll
        for (int e = 0; e < DMAX; ++e) {
          if (e < deg) {
            const float gs = buf[addr[e]];
            const float gc = gs + (t < T - 1 ? gn[(size_t)e * Z] : 0.0f);
            gv[e] = gs;
            const bool am = e == st.a1;
            const float mag = am ? st.m2 : st.m1;
            const float pre = a * mag - b;
            if (pre > 0.0f) {
              const float sg = msg_sign(st, v[e]);
              ga += gc * sg * mag;
              gb -= gc * sg;
              const float gm = gc * sg * a;
              if (am) {
                g2 += gm * st.s2;
              } else {
                g1 += gm * st.s1;
              }
            }
          }
        }
#pragma unroll
        for (int e = 0; e < DMAX; ++e) {
          if (e < deg) {
            const float g = gv[e] + (e == st.a1 ? g1 : 0.0f) + (e == st.a2 ? g2 : 0.0f);
            buf[addr[e]] = g;
            gn[(size_t)e * Z] = -g;
          }
        }
      }
      ga = wave_sum(ga);
      gb = wave_sum(gb);
      if ((t0 & 63) == 0) {
        atomicAdd(&gacc[2 * (t * M + l)], ga);
        atomicAdd(&gacc[2 * (t * M + l) + 1], gb);
      }
      __syncthreads();
    }
  }
  for (int i = t0; i < T * M; i += NT) {
    const float x = gacc[2 * i], y = gacc[2 * i + 1];
    if (x != 0.0f) {
      atomicAdd(&galpha[i], x);
    }
    if (y != 0.0f) {
      atomicAdd(&gbeta[i], y);
    }
  }
}

// Adam on the (alpha, beta) tables, then alpha clamped to [0, 511/256] and beta to
// [0, 64].  A non-negative offset keeps the model continuous: a message whose other
// edges include a zero v has magnitude max(0, -beta) = 0, so the sign of that zero
// (which rounding decides) does not matter.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, int n_alpha, float lr, float b1, float b2, float eps, float c1,
                       float c2)
{
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) {
    return;
  }
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float x = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  x = i < n_alpha ? fminf(fmaxf(x, 0.0f), 511.0f / 256.0f) : fminf(fmaxf(x, 0.0f), 64.0f);
  p[i] = x;
}

size_t nms_train_lds_bytes(int bg, int Z, int L)
{
  return (size_t)((bg == 1 ? 22 : 10) + L) * Z * 4;
}

hipError_t launch_nms_train(int bg, const train_job* d_jobs, unsigned n, int nt, size_t lds, int T,
                            const float* alpha, const float* beta, const float* wt, float tau, const int8_t* in,
                            const uint8_t* tgt, float* ws, float* loss, float* galpha, float* gbeta,
                            const uint32_t* gtab, int backward, hipStream_t s)
{
  const void* fn = bg == 1 ? (const void*)k_nms_train<1> : (const void*)k_nms_train<2>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) {
    return e;
  }
  if (bg == 1) {
    hipLaunchKernelGGL(k_nms_train<1>, dim3(n), dim3(nt), lds, s, d_jobs, T, alpha, beta, wt, tau, in, tgt, ws, loss,
                       galpha, gbeta, gtab, backward);
  } else {
    hipLaunchKernelGGL(k_nms_train<2>, dim3(n), dim3(nt), lds, s, d_jobs, T, alpha, beta, wt, tau, in, tgt, ws, loss,
                       galpha, gbeta, gtab, backward);
  }
  return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, int n_alpha, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t s)
{
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, p, g, m, v, n, n_alpha, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

} // namespace lh
