// HIP kernels of libtcm for CDNA4 (gfx950): gate application (one kernel per
// gate, and LDS-tiled fused passes), adjoint reverse sweeps, Pauli-sum
// energies / H|psi>, and the fp64 reductions behind them.
//
// State layout: [batch][2^n] complex (float2 or double2), qubit 0 = the most
// significant index bit (circuit.py:803-810).  A gate on qubit q acts on index
// bit n-1-q.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "kernels.h"

namespace tcm {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
template <int V>
using IC = std::integral_constant<int, V>;

template <typename T>
__device__ __forceinline__ V2<T> mk(T a, T b) {
  V2<T> r;
  r.x = a;
  r.y = b;
  return r;
}
// Im(conj(a) b), Re(conj(a) b)
template <typename T>
__device__ __forceinline__ T imcj(V2<T> a, V2<T> b) { return a.x * b.y - a.y * b.x; }
template <typename T>
__device__ __forceinline__ T recj(V2<T> a, V2<T> b) { return a.x * b.x + a.y * b.y; }

__device__ __forceinline__ uint32_t swz(uint32_t t) {
  return t ^ (((t >> 4) ^ (t >> 8) ^ (t >> 12) ^ (t >> 16)) & 0xFu);
}

// Offset (in amplitudes, from the block base) of tile element x: the tile's
// L low bits are contiguous in the state and the run-offset table `ro` maps
// the tile bits above them to their state bit positions.  The map is linear
// over disjoint bit sets, so element it * nthr + tid sits at
// tile_dep(it * nthr) (block-uniform, scalar loads) + tile_dep(tid).
__device__ __forceinline__ uint32_t tile_dep(const uint32_t* ro, int L, uint32_t x) {
  return ro[x >> L] + (x & ((1u << L) - 1));
}

// a pointer known to be the same in every lane, moved to scalar registers
template <typename P>
__device__ __forceinline__ P* uniform_ptr(P* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<P*>(((uint64_t)hi << 32) | lo);
}

// A block-uniform zero the optimiser cannot see through: addresses rebuilt
// from it are recomputed (cheap scalar work) instead of held live in
// registers across the group loop.
__device__ __forceinline__ uint32_t opaque_zero() {
  uint32_t z = 0;
  asm volatile("" : "+v"(z));
  return __builtin_amdgcn_readfirstlane(z);
}
__device__ __forceinline__ void rematerialize(uint32_t& dlo, uint32_t& slo, const uint32_t*& ro) {
  asm volatile("" : "+v"(dlo), "+v"(slo));
  ro += opaque_zero();
}

// tile position of a group's thread bits: the thread index bits moved as
// runs of consecutive bits to their tile positions
__device__ __forceinline__ uint32_t group_base(const DGroup* G, uint32_t tid) {
  uint32_t tb = 0;
  for (int r = 0; r < G->nrun; ++r) {
    const uint32_t w = (uint32_t)G->run[r];
    tb |= ((tid >> (w & 0xFFu)) & ((1u << ((w >> 8) & 0xFFu)) - 1u)) << (w >> 16);
  }
  return tb;
}

// insert a zero bit at position b
__device__ __forceinline__ uint64_t ins0(uint64_t x, int b) {
  uint64_t lo = x & ((1ull << b) - 1);
  return ((x >> b) << (b + 1)) | lo;
}

template <typename T>
__device__ __forceinline__ const T* mptr(int moff, const T* mt, const T* ct) {
  return moff >= 0 ? mt + moff : ct + (-moff - 1);
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// one DPP row-shift of a double (lanes without a source read 0)
template <int CTRL>
__device__ __forceinline__ double dpp_shr(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double read_lane(double x, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                          __builtin_amdgcn_readlane(__double2loint(x), lane));
}
// fp64 sum over a full 64-lane wave, uniform result: inclusive row prefix
// sums by DPP row_shr 1/2/4/8 leave each 16-lane row's total in its last
// lane; the four row totals are read back as scalars.
__device__ __forceinline__ double wave_sum_dpp(double x) {
  x += dpp_shr<0x111>(x);
  x += dpp_shr<0x112>(x);
  x += dpp_shr<0x114>(x);
  x += dpp_shr<0x118>(x);
  return (read_lane(x, 15) + read_lane(x, 31)) + (read_lane(x, 47) + read_lane(x, 63));
}

// block-wide sum of (a, b) in fp64; result valid in thread 0.  blockDim.x
// must be a multiple of 64 and <= 1024.
__device__ __forceinline__ void block_sum2(double& a, double& b) {
  __shared__ double red[2][16];
  a = wave_sum_dpp(a);
  b = wave_sum_dpp(b);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][w] = a;
    red[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    double sa = 0, sb = 0;
    for (int i = 0; i < nw; ++i) {
      sa += red[0][i];
      sb += red[1][i];
    }
    a = sa;
    b = sb;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// slot dispatch: runtime register slot -> compile-time template argument.
// Plain switches (no lambdas capturing the register array) so that the
// per-thread amplitude array stays in VGPRs.
// ---------------------------------------------------------------------------
#define TCM_SW1(R, s, ...)                                          \
  switch (s) {                                                      \
    case 0: { constexpr int S_ = 0; __VA_ARGS__; } break;           \
    case 1: if constexpr (R > 1) { constexpr int S_ = 1; __VA_ARGS__; } break; \
    case 2: if constexpr (R > 2) { constexpr int S_ = 2; __VA_ARGS__; } break; \
    case 3: if constexpr (R > 3) { constexpr int S_ = 3; __VA_ARGS__; } break; \
    case 4: if constexpr (R > 4) { constexpr int S_ = 4; __VA_ARGS__; } break; \
    default: break;                                                 \
  }

#define TCM_SW2B(R, A_, s1, ...)                                                         \
  switch (s1) {                                                                          \
    case 0: if constexpr (A_ != 0) { constexpr int B_ = 0; __VA_ARGS__; } break;         \
    case 1: if constexpr (R > 1 && A_ != 1) { constexpr int B_ = 1; __VA_ARGS__; } break; \
    case 2: if constexpr (R > 2 && A_ != 2) { constexpr int B_ = 2; __VA_ARGS__; } break; \
    case 3: if constexpr (R > 3 && A_ != 3) { constexpr int B_ = 3; __VA_ARGS__; } break; \
    case 4: if constexpr (R > 4 && A_ != 4) { constexpr int B_ = 4; __VA_ARGS__; } break; \
    default: break;                                                                      \
  }

#define TCM_SW2(R, s0, s1, ...)                                                           \
  switch (s0) {                                                                           \
    case 0: { constexpr int A_ = 0; TCM_SW2B(R, A_, s1, __VA_ARGS__) } break;             \
    case 1: if constexpr (R > 1) { constexpr int A_ = 1; TCM_SW2B(R, A_, s1, __VA_ARGS__) } break; \
    case 2: if constexpr (R > 2) { constexpr int A_ = 2; TCM_SW2B(R, A_, s1, __VA_ARGS__) } break; \
    case 3: if constexpr (R > 3) { constexpr int A_ = 3; TCM_SW2B(R, A_, s1, __VA_ARGS__) } break; \
    case 4: if constexpr (R > 4) { constexpr int A_ = 4; TCM_SW2B(R, A_, s1, __VA_ARGS__) } break; \
    default: break;                                                                       \
  }

// ---------------------------------------------------------------------------
// in-register gate application on v[0 .. 2^R): register index bit S <-> the
// gate qubit.  DAG applies U^dagger.
// ---------------------------------------------------------------------------
template <typename T, int R, int S, bool DAG>
__device__ __forceinline__ void ap_u1(V2<T>* v, const T* m) {
  T ar = m[0], ai = m[1], br = m[2], bi = m[3], cr = m[4], ci = m[5], dr = m[6], di = m[7];
  if constexpr (DAG) {
    T tr = br, ti = bi;
    br = cr; bi = -ci;
    cr = tr; ci = -ti;
    ai = -ai; di = -di;
  }
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    if (j & (1 << S)) continue;
    V2<T> x = v[j], y = v[j | (1 << S)];
    v[j] = mk<T>(ar * x.x - ai * x.y + br * y.x - bi * y.y, ar * x.y + ai * x.x + br * y.y + bi * y.x);
    v[j | (1 << S)] = mk<T>(cr * x.x - ci * x.y + dr * y.x - di * y.y, cr * x.y + ci * x.x + dr * y.y + di * y.x);
  }
}

template <typename T, int R, int S, bool DAG>
__device__ __forceinline__ void ap_r1(V2<T>* v, const T* m) {
  T a = m[0], b = m[1], c = m[2], d = m[3];
  if constexpr (DAG) {
    T t = b; b = c; c = t;
  }
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    if (j & (1 << S)) continue;
    V2<T> x = v[j], y = v[j | (1 << S)];
    v[j] = mk<T>(a * x.x + b * y.x, a * x.y + b * y.y);
    v[j | (1 << S)] = mk<T>(c * x.x + d * y.x, c * x.y + d * y.y);
  }
}

template <typename T, int R, int S, bool DAG>
__device__ __forceinline__ void ap_d1(V2<T>* v, const T* m) {
  T d0r = m[0], d0i = m[1], d1r = m[2], d1i = m[3];
  if constexpr (DAG) { d0i = -d0i; d1i = -d1i; }
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    V2<T> x = v[j];
    if (j & (1 << S)) v[j] = mk<T>(d1r * x.x - d1i * x.y, d1r * x.y + d1i * x.x);
    else v[j] = mk<T>(d0r * x.x - d0i * x.y, d0r * x.y + d0i * x.x);
  }
}

template <typename T, int R, int S>
__device__ __forceinline__ void ap_x1(V2<T>* v) {
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    if (j & (1 << S)) continue;
    V2<T> t = v[j];
    v[j] = v[j | (1 << S)];
    v[j | (1 << S)] = t;
  }
}

// 2-qubit: matrix index r = (bit S0) << 1 | (bit S1)
template <typename T, int R, int S0, int S1, bool DAG>
__device__ __forceinline__ void ap_u2(V2<T>* v, const T* m) {
  T mr[16], mi[16];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if constexpr (DAG) {
        mr[r * 4 + c] = m[2 * (c * 4 + r)];
        mi[r * 4 + c] = -m[2 * (c * 4 + r) + 1];
      } else {
        mr[r * 4 + c] = m[2 * (r * 4 + c)];
        mi[r * 4 + c] = m[2 * (r * 4 + c) + 1];
      }
    }
  constexpr int m0 = 1 << S0, m1 = 1 << S1;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    if (j & (m0 | m1)) continue;
    const int idx[4] = {j, j | m1, j | m0, j | m0 | m1};
    V2<T> x[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = v[idx[c]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      T sr = 0, si = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        sr += mr[r * 4 + c] * x[c].x - mi[r * 4 + c] * x[c].y;
        si += mr[r * 4 + c] * x[c].y + mi[r * 4 + c] * x[c].x;
      }
      v[idx[r]] = mk<T>(sr, si);
    }
  }
}

template <typename T, int R, int S0, int S1, bool DAG>
__device__ __forceinline__ void ap_d2(V2<T>* v, const T* m) {
  T dr[4], di[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    dr[r] = m[2 * r];
    di[r] = DAG ? -m[2 * r + 1] : m[2 * r + 1];
  }
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    const int r = (((j >> S0) & 1) << 1) | ((j >> S1) & 1);
    V2<T> x = v[j];
    v[j] = mk<T>(dr[r] * x.x - di[r] * x.y, dr[r] * x.y + di[r] * x.x);
  }
}

template <typename T, int R, int S0, int S1>
__device__ __forceinline__ void ap_cx(V2<T>* v) {
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    if (!(j & (1 << S0)) || (j & (1 << S1))) continue;
    V2<T> t = v[j];
    v[j] = v[j | (1 << S1)];
    v[j | (1 << S1)] = t;
  }
}

template <typename T, int R, int S0, int S1>
__device__ __forceinline__ void ap_cz(V2<T>* v) {
#pragma unroll
  for (int j = 0; j < (1 << R); ++j)
    if ((j & (1 << S0)) && (j & (1 << S1))) v[j] = mk<T>(-v[j].x, -v[j].y);
}

template <typename T, int R, int S0, int S1>
__device__ __forceinline__ void ap_swap(V2<T>* v) {
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    if (!(j & (1 << S0)) || (j & (1 << S1))) continue;
    const int jj = (j ^ (1 << S0)) | (1 << S1);
    V2<T> t = v[j];
    v[j] = v[jj];
    v[jj] = t;
  }
}

template <typename T, int R, bool DAG>
__device__ __forceinline__ void ap_dg(V2<T>* v, const T* d) {
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    T dr = d[2 * j], di = DAG ? -d[2 * j + 1] : d[2 * j + 1];
    V2<T> x = v[j];
    v[j] = mk<T>(dr * x.x - di * x.y, dr * x.y + di * x.x);
  }
}

// apply one op (kinds with a register form) on v.  FULL = false drops the
// dense 4x4 kind, whose 32-value matrix dominates register pressure; passes
// without it run a leaner kernel with higher occupancy.
template <typename T, int R, bool DAG, bool FULL = true>
__device__ __forceinline__ void apply_op(V2<T>* v, const DOp& op, const T* mt, const T* ct) {
  switch (op.kind) {
    case OK_U1: {
      const T* m = mptr<T>(op.moff, mt, ct);
      TCM_SW1(R, op.s0, ap_u1<T, R, S_, DAG>(v, m))
    } break;
    case OK_R1: {
      const T* m = mptr<T>(op.moff, mt, ct);
      TCM_SW1(R, op.s0, ap_r1<T, R, S_, DAG>(v, m))
    } break;
    case OK_D1: {
      const T* m = mptr<T>(op.moff, mt, ct);
      TCM_SW1(R, op.s0, ap_d1<T, R, S_, DAG>(v, m))
    } break;
    case OK_X1:
      TCM_SW1(R, op.s0, ap_x1<T, R, S_>(v))
      break;
    case OK_U2: {
      if constexpr (FULL) {
        const T* m = mptr<T>(op.moff, mt, ct);
        TCM_SW2(R, op.s0, op.s1, ap_u2<T, R, A_, B_, DAG>(v, m))
      }
    } break;
    case OK_D2: {
      const T* m = mptr<T>(op.moff, mt, ct);
      TCM_SW2(R, op.s0, op.s1, ap_d2<T, R, A_, B_, DAG>(v, m))
    } break;
    case OK_CX:
      TCM_SW2(R, op.s0, op.s1, ap_cx<T, R, A_, B_>(v))
      break;
    case OK_CZ:
      TCM_SW2(R, op.s0, op.s1, ap_cz<T, R, A_, B_>(v))
      break;
    case OK_SWAP:
      TCM_SW2(R, op.s0, op.s1, ap_swap<T, R, A_, B_>(v))
      break;
    case OK_DG:
      ap_dg<T, R, DAG>(v, mt + op.moff);
      break;
    default:
      break;
  }
}

// ---------------------------------------------------------------------------
// gradient kernels on registers: Im <l | G | p> summed over the thread's
// amplitudes (p = psi after the gate, l = lambda after the gate)
// ---------------------------------------------------------------------------
template <typename T, int R, int S>
__device__ __forceinline__ T g_x(const V2<T>* p, const V2<T>* l) {
  T a = 0;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    if (j & (1 << S)) continue;
    const int jj = j | (1 << S);
    a += imcj<T>(l[j], p[jj]) + imcj<T>(l[jj], p[j]);
  }
  return a;
}
template <typename T, int R, int S>
__device__ __forceinline__ T g_y(const V2<T>* p, const V2<T>* l) {
  T a = 0;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    if (j & (1 << S)) continue;
    const int jj = j | (1 << S);
    a += recj<T>(l[jj], p[j]) - recj<T>(l[j], p[jj]);
  }
  return a;
}
template <typename T, int R, int S>
__device__ __forceinline__ T g_z(const V2<T>* p, const V2<T>* l) {
  T a = 0;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    T t = imcj<T>(l[j], p[j]);
    a += (j & (1 << S)) ? -t : t;
  }
  return a;
}
template <typename T, int R, int S>
__device__ __forceinline__ T g_p1(const V2<T>* p, const V2<T>* l) {
  T a = 0;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j)
    if (j & (1 << S)) a += imcj<T>(l[j], p[j]);
  return a;
}
template <typename T, int R, int S0, int S1>
__device__ __forceinline__ T g_xx(const V2<T>* p, const V2<T>* l) {
  T a = 0;
  constexpr int m = (1 << S0) | (1 << S1);
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) a += imcj<T>(l[j], p[j ^ m]);
  return a;
}
template <typename T, int R, int S0, int S1>
__device__ __forceinline__ T g_yy(const V2<T>* p, const V2<T>* l) {
  T a = 0;
  constexpr int m = (1 << S0) | (1 << S1);
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    T t = imcj<T>(l[j], p[j ^ m]);
    const int par = ((j >> S0) ^ (j >> S1)) & 1;
    a += par ? t : -t;
  }
  return a;
}
template <typename T, int R, int S0, int S1>
__device__ __forceinline__ T g_zz(const V2<T>* p, const V2<T>* l) {
  T a = 0;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    T t = imcj<T>(l[j], p[j]);
    const int par = ((j >> S0) ^ (j >> S1)) & 1;
    a += par ? -t : t;
  }
  return a;
}
// general Hermitian generator, 2x2 (U1 layout) on slot S
template <typename T, int R, int S>
__device__ __forceinline__ T g_m1(const V2<T>* p, const V2<T>* l, const T* g) {
  T a = 0;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    if (j & (1 << S)) continue;
    const int jj = j | (1 << S);
    V2<T> x = p[j], y = p[jj];
    V2<T> u = mk<T>(g[0] * x.x - g[1] * x.y + g[2] * y.x - g[3] * y.y, g[0] * x.y + g[1] * x.x + g[2] * y.y + g[3] * y.x);
    V2<T> w = mk<T>(g[4] * x.x - g[5] * x.y + g[6] * y.x - g[7] * y.y, g[4] * x.y + g[5] * x.x + g[6] * y.y + g[7] * y.x);
    a += imcj<T>(l[j], u) + imcj<T>(l[jj], w);
  }
  return a;
}
template <typename T, int R, int S0, int S1>
__device__ __forceinline__ T g_m2(const V2<T>* p, const V2<T>* l, const T* g) {
  T a = 0;
  constexpr int m0 = 1 << S0, m1 = 1 << S1;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) {
    if (j & (m0 | m1)) continue;
    const int idx[4] = {j, j | m1, j | m0, j | m0 | m1};
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      T sr = 0, si = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const T gr = g[2 * (r * 4 + c)], gi = g[2 * (r * 4 + c) + 1];
        sr += gr * p[idx[c]].x - gi * p[idx[c]].y;
        si += gr * p[idx[c]].y + gi * p[idx[c]].x;
      }
      a += imcj<T>(l[idx[r]], mk<T>(sr, si));
    }
  }
  return a;
}

template <typename T, int R, bool FULL = true>
__device__ __forceinline__ T grad_op(const V2<T>* p, const V2<T>* l, int gen, int s0, int s1, const T* ct, int goff) {
  T a = 0;
  switch (gen) {
    case GN_X: TCM_SW1(R, s0, a = (g_x<T, R, S_>(p, l))) break;
    case GN_Y: TCM_SW1(R, s0, a = (g_y<T, R, S_>(p, l))) break;
    case GN_Z: TCM_SW1(R, s0, a = (g_z<T, R, S_>(p, l))) break;
    case GN_P1: TCM_SW1(R, s0, a = (g_p1<T, R, S_>(p, l))) break;
    case GN_XX: TCM_SW2(R, s0, s1, a = (g_xx<T, R, A_, B_>(p, l))) break;
    case GN_YY: TCM_SW2(R, s0, s1, a = (g_yy<T, R, A_, B_>(p, l))) break;
    case GN_ZZ: TCM_SW2(R, s0, s1, a = (g_zz<T, R, A_, B_>(p, l))) break;
    case GN_M1: TCM_SW1(R, s0, a = (g_m1<T, R, S_>(p, l, ct + goff))) break;
    case GN_M2:
      if constexpr (FULL) { TCM_SW2(R, s0, s1, a = (g_m2<T, R, A_, B_>(p, l, ct + goff))) }
      break;
    default: break;
  }
  return a;
}

// gradient of a diagonal-generator member from precomputed im[j] = Im(conj(l_j) p_j)
template <typename T, int R, int S>
__device__ __forceinline__ T gm_z(const T* im) {
  T a = 0;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) a += ((j >> S) & 1) ? -im[j] : im[j];
  return a;
}
template <typename T, int R, int S>
__device__ __forceinline__ T gm_p1(const T* im) {
  T a = 0;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j)
    if ((j >> S) & 1) a += im[j];
  return a;
}
template <typename T, int R, int S0, int S1>
__device__ __forceinline__ T gm_zz(const T* im) {
  T a = 0;
#pragma unroll
  for (int j = 0; j < (1 << R); ++j) a += (((j >> S0) ^ (j >> S1)) & 1) ? -im[j] : im[j];
  return a;
}
template <typename T, int R>
__device__ __forceinline__ T grad_member(const T* im, int gen, int s0, int s1) {
  T a = 0;
  switch (gen) {
    case GN_Z: TCM_SW1(R, s0, a = (gm_z<T, R, S_>(im))) break;
    case GN_P1: TCM_SW1(R, s0, a = (gm_p1<T, R, S_>(im))) break;
    case GN_ZZ: TCM_SW2(R, s0, s1, a = (gm_zz<T, R, A_, B_>(im))) break;
    default: break;
  }
  return a;
}

// ---------------------------------------------------------------------------
// matrix preparation: per batch item, evaluate every recipe from params
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_prep(const DRecipe* __restrict__ rec, int nrec, int first_dg, int dg_log,
                       const DFactor* __restrict__ fac, const T* __restrict__ params, int64_t pstride,
                       const T* __restrict__ ct, T* __restrict__ mtab, int64_t mstride) {
  const int64_t b = blockIdx.x;
  T* mt = mtab + b * mstride;
  const T* pr = params ? params + b * pstride : nullptr;
  for (int i = threadIdx.x; i < first_dg; i += blockDim.x) {
    const DRecipe r = rec[i];
    if (r.kind < 0) continue;
    double a = r.param >= 0 ? r.scale * (double)pr[r.param] : r.theta;
    T* o = mt + r.out;
    switch (r.kind) {
      case TCM_G_RX: {
        double c, s;
        sincos(0.5 * a, &s, &c);
        o[0] = (T)c; o[1] = 0; o[2] = 0; o[3] = (T)-s; o[4] = 0; o[5] = (T)-s; o[6] = (T)c; o[7] = 0;
      } break;
      case TCM_G_RY: {
        double c, s;
        sincos(0.5 * a, &s, &c);
        o[0] = (T)c; o[1] = (T)-s; o[2] = (T)s; o[3] = (T)c;
        // three-shear form (Paeth): the rotation by phi' is x -= t y,
        // y += s' x, x -= t y with t = tan(phi' / 2), s' = sin phi'; phi' is
        // phi = a / 2 reduced into [-pi/2, pi/2] (|t| <= 1) by
        // R(phi) = -R(phi - pi), the sign kept in sigma
        double ph = remainder(0.5 * a, 2.0 * M_PI);
        double sg = 1.0;
        if (fabs(ph) > 0.5 * M_PI) {
          ph -= copysign(M_PI, ph);
          sg = -1.0;
        }
        o[4] = (T)tan(0.5 * ph); o[5] = (T)sin(ph); o[6] = (T)sg; o[7] = 0;
      } break;
      case TCM_G_RZ: {
        double c, s;
        sincos(0.5 * a, &s, &c);
        o[0] = (T)c; o[1] = (T)-s; o[2] = (T)c; o[3] = (T)s;
      } break;
      case TCM_G_RZZ: {
        double c, s;
        sincos(0.5 * a, &s, &c);
        o[0] = (T)c; o[1] = (T)-s; o[2] = (T)c; o[3] = (T)s;
        o[4] = (T)c; o[5] = (T)s; o[6] = (T)c; o[7] = (T)-s;
      } break;
      case TCM_G_RXX:
      case TCM_G_RYY: {
        double c, s;
        sincos(0.5 * a, &s, &c);
        for (int q = 0; q < 32; ++q) o[q] = 0;
        for (int d = 0; d < 4; ++d) o[2 * (d * 4 + d)] = (T)c;
        if (r.kind == TCM_G_RXX) {
          // -i s XX : anti-diagonal ones
          for (int d = 0; d < 4; ++d) o[2 * (d * 4 + (3 - d)) + 1] = (T)-s;
        } else {
          // -i s YY : YY[0][3] = -1, YY[1][2] = 1, YY[2][1] = 1, YY[3][0] = -1
          o[2 * (0 * 4 + 3) + 1] = (T)s;
          o[2 * (1 * 4 + 2) + 1] = (T)-s;
          o[2 * (2 * 4 + 1) + 1] = (T)-s;
          o[2 * (3 * 4 + 0) + 1] = (T)s;
        }
      } break;
      case TCM_G_PHASE: {
        double c, s;
        sincos(a, &s, &c);
        o[0] = 1; o[1] = 0; o[2] = (T)c; o[3] = (T)s;
      } break;
      case TCM_G_EXP1: {
        // cos(a) I - i sin(a) G, G from the constant table (full matrix)
        double c, s;
        sincos(a, &s, &c);
        const T* g = ct + r.goff;
        const int d = (r.layout == OK_U1 || r.layout == OK_D1) ? 2 : 4;
        if (r.layout == OK_D1 || r.layout == OK_D2) {
          for (int q = 0; q < d; ++q) {
            double gr = g[2 * (q * d + q)], gi = g[2 * (q * d + q) + 1];
            o[2 * q] = (T)(c + s * gi);
            o[2 * q + 1] = (T)(-s * gr);
          }
        } else {
          for (int rr = 0; rr < d; ++rr)
            for (int cc = 0; cc < d; ++cc) {
              double gr = g[2 * (rr * d + cc)], gi = g[2 * (rr * d + cc) + 1];
              o[2 * (rr * d + cc)] = (T)((rr == cc ? c : 0.0) + s * gi);
              o[2 * (rr * d + cc) + 1] = (T)(-s * gr);
            }
        }
      } break;
      case TCM_G_EXPG: {
        // (I - Q) + cos(a) Q - i sin(a) G; G then Q = G^2 in the constant table
        double c, s;
        sincos(a, &s, &c);
        const int d = (r.layout == OK_U1 || r.layout == OK_D1) ? 2 : 4;
        const T* g = ct + r.goff;
        const T* q = g + 2 * d * d;
        if (r.layout == OK_D1 || r.layout == OK_D2) {
          for (int k = 0; k < d; ++k) {
            const int e = 2 * (k * d + k);
            o[2 * k] = (T)(1.0 + (c - 1.0) * (double)q[e] + s * (double)g[e + 1]);
            o[2 * k + 1] = (T)((c - 1.0) * (double)q[e + 1] - s * (double)g[e]);
          }
        } else {
          for (int rr = 0; rr < d; ++rr)
            for (int cc = 0; cc < d; ++cc) {
              const int e = 2 * (rr * d + cc);
              o[e] = (T)((rr == cc ? 1.0 : 0.0) + (c - 1.0) * (double)q[e] + s * (double)g[e + 1]);
              o[e + 1] = (T)((c - 1.0) * (double)q[e + 1] - s * (double)g[e]);
            }
        }
      } break;
      default: break;
    }
  }
  __syncthreads();
  // fused diagonals (recipes first_dg..nrec-1): products of member diagonals
  // over the group index j; thread t takes recipe first_dg + (t >> dg_log)
  // and entry t & (2^dg_log - 1), so all recipes are built side by side
  const int ndg = nrec - first_dg;
  for (int t = threadIdx.x; t < (ndg << dg_log); t += blockDim.x) {
    const DRecipe r = rec[first_dg + (t >> dg_log)];
    const int j = t & ((1 << dg_log) - 1);
    if (r.kind >= 0 || j >= (1 << r.R)) continue;
    T* o = mt + r.out;
    {
      double pr_ = 1.0, pi_ = 0.0;
      for (int f = 0; f < r.nfac; ++f) {
        const DFactor F = fac[r.goff + f];
        double er, ei;
        if (F.layout == OK_CZ) {
          er = (((j >> F.s0) & 1) && ((j >> F.s1) & 1)) ? -1.0 : 1.0;
          ei = 0.0;
        } else {
          const T* src = F.src >= 0 ? mt + F.src : ct + (-F.src - 1);
          int idx = F.layout == OK_D1 ? ((j >> F.s0) & 1) : ((((j >> F.s0) & 1) << 1) | ((j >> F.s1) & 1));
          er = src[2 * idx];
          ei = src[2 * idx + 1];
        }
        double nr = pr_ * er - pi_ * ei, ni = pr_ * ei + pi_ * er;
        pr_ = nr;
        pi_ = ni;
      }
      o[2 * j] = (T)pr_;
      o[2 * j + 1] = (T)pi_;
    }
  }
}

// |0..0> on every state of the batch
template <typename T>
__global__ void k_init_zero(V2<T>* __restrict__ st, int64_t total, int64_t N) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
    st[i] = mk<T>((i % N) == 0 ? T(1) : T(0), T(0));
}

// ---------------------------------------------------------------------------
// plain kernels: one launch per gate, grid-stride over (batch, item)
// ---------------------------------------------------------------------------
template <typename T, int KIND, bool DAG>
__global__ void __launch_bounds__(256) k_plain1(DOp op, const T* __restrict__ mtab, int64_t mstride,
                                                const T* __restrict__ ct, V2<T>* __restrict__ st,
                                                int64_t sstride, int n, int64_t batch) {
  const int64_t items = 1ll << (n - 1);
  const int64_t total = items * batch;
  const int b0 = op.s0;
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < total; w += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = w >> (n - 1);
    const int64_t idx = w & (items - 1);
    const uint64_t i0 = ins0(idx, b0), i1 = i0 | (1ull << b0);
    V2<T>* s = st + b * sstride;
    V2<T> v[2] = {s[i0], s[i1]};
    const T* mt = mtab + b * mstride;
    DOp o = op;
    o.s0 = 0;
    apply_op<T, 1, DAG>(v, o, mt, ct);
    s[i0] = v[0];
    s[i1] = v[1];
  }
}

template <typename T, int KIND, bool DAG>
__global__ void __launch_bounds__(256) k_plain2(DOp op, const T* __restrict__ mtab, int64_t mstride,
                                                const T* __restrict__ ct, V2<T>* __restrict__ st,
                                                int64_t sstride, int n, int64_t batch) {
  const int64_t items = 1ll << (n - 2);
  const int64_t total = items * batch;
  const int ba = op.s0, bb = op.s1;  // bit of qubit 0 (matrix MSB), qubit 1
  const int lo = ba < bb ? ba : bb, hi = ba < bb ? bb : ba;
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < total; w += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = w >> (n - 2);
    const int64_t idx = w & (items - 1);
    const uint64_t i00 = ins0(ins0(idx, lo), hi);
    const uint64_t ma = 1ull << ba, mb = 1ull << bb;
    V2<T>* s = st + b * sstride;
    // register index: bit 1 <-> qubit 0 (matrix MSB), bit 0 <-> qubit 1
    V2<T> v[4] = {s[i00], s[i00 | mb], s[i00 | ma], s[i00 | ma | mb]};
    const T* mt = mtab + b * mstride;
    DOp o = op;
    o.s0 = 1;
    o.s1 = 0;
    apply_op<T, 2, DAG>(v, o, mt, ct);
    s[i00] = v[0];
    s[i00 | mb] = v[1];
    s[i00 | ma] = v[2];
    s[i00 | ma | mb] = v[3];
  }
}

// dense k-qubit gate (k = K), constant matrix; bits[0] = matrix MSB
template <typename T, int K, bool DAG>
__global__ void __launch_bounds__(256) k_plaink(DOp op, const int32_t* __restrict__ ukbits,
                                                const T* __restrict__ ct, V2<T>* __restrict__ st,
                                                int64_t sstride, int n, int64_t batch) {
  const int64_t items = 1ll << (n - K);
  const int64_t total = items * batch;
  int bits[K], sorted[K];
  for (int i = 0; i < K; ++i) sorted[i] = bits[i] = ukbits[op.mem0 + i];
  for (int i = 0; i < K; ++i)
    for (int j = i + 1; j < K; ++j)
      if (sorted[j] < sorted[i]) { int t = sorted[i]; sorted[i] = sorted[j]; sorted[j] = t; }
  const T* m = ct + (-op.moff - 1);
  constexpr int D = 1 << K;
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < total; w += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = w >> (n - K);
    uint64_t base = w & (items - 1);
    for (int i = 0; i < K; ++i) base = ins0(base, sorted[i]);
    V2<T>* s = st + b * sstride;
    V2<T> x[D];
    uint64_t ad[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      uint64_t a = base;
#pragma unroll
      for (int i = 0; i < K; ++i)
        if ((c >> (K - 1 - i)) & 1) a |= 1ull << bits[i];
      ad[c] = a;
      x[c] = s[a];
    }
    for (int r = 0; r < D; ++r) {
      T sr = 0, si = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        T mr, mi;
        if constexpr (DAG) { mr = m[2 * (c * D + r)]; mi = -m[2 * (c * D + r) + 1]; }
        else { mr = m[2 * (r * D + c)]; mi = m[2 * (r * D + c) + 1]; }
        sr += mr * x[c].x - mi * x[c].y;
        si += mr * x[c].y + mi * x[c].x;
      }
      s[ad[r]] = mk<T>(sr, si);
    }
  }
}

// plain adjoint step for a 1- or 2-qubit gate: gradient partial (if the gate
// is trainable), then psi <- U^dag psi, lambda <- U^dag lambda.
// grid = (slots, batch); partial[b * pstride + blockIdx.x]
template <typename T, int NQ>
__global__ void __launch_bounds__(256) k_plain_adj(DOp op, const T* __restrict__ mtab, int64_t mstride,
                                                   const T* __restrict__ ct, const V2<T>* psi_in, V2<T>* psi_out,
                                                   const V2<T>* lam_in, V2<T>* lam_out, int64_t sstride, int n,
                                                   double* __restrict__ part, int64_t pstride) {
  const int64_t items = 1ll << (n - NQ);
  const int64_t b = blockIdx.y;
  const T* mt = mtab + b * mstride;
  const V2<T>* pi = psi_in + b * sstride;
  const V2<T>* li = lam_in + b * sstride;
  V2<T>* po = psi_out ? psi_out + b * sstride : nullptr;
  V2<T>* lo_ = lam_out + b * sstride;
  double acc = 0.0;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < items; idx += (int64_t)gridDim.x * blockDim.x) {
    if constexpr (NQ == 1) {
      const int b0 = op.s0;
      const uint64_t i0 = ins0(idx, b0), i1 = i0 | (1ull << b0);
      V2<T> p[2] = {pi[i0], pi[i1]}, l[2] = {li[i0], li[i1]};
      if (op.gen != GN_NONE) acc += (double)grad_op<T, 1>(p, l, op.gen, 0, -1, ct, op.goff);
      DOp o = op;
      o.s0 = 0;
      apply_op<T, 1, true>(p, o, mt, ct);
      apply_op<T, 1, true>(l, o, mt, ct);
      if (po) { po[i0] = p[0]; po[i1] = p[1]; }
      lo_[i0] = l[0];
      lo_[i1] = l[1];
    } else {
      const int ba = op.s0, bb = op.s1;
      const int lo = ba < bb ? ba : bb, hi = ba < bb ? bb : ba;
      const uint64_t i00 = ins0(ins0(idx, lo), hi);
      const uint64_t ma = 1ull << ba, mb = 1ull << bb;
      const uint64_t ad[4] = {i00, i00 | mb, i00 | ma, i00 | ma | mb};
      V2<T> p[4], l[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) { p[c] = pi[ad[c]]; l[c] = li[ad[c]]; }
      if (op.gen != GN_NONE) acc += (double)grad_op<T, 2>(p, l, op.gen, 1, 0, ct, op.goff);
      DOp o = op;
      o.s0 = 1;
      o.s1 = 0;
      apply_op<T, 2, true>(p, o, mt, ct);
      apply_op<T, 2, true>(l, o, mt, ct);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (po) po[ad[c]] = p[c];
        lo_[ad[c]] = l[c];
      }
    }
  }
  if (op.gop >= 0) {
    double dummy = 0.0;
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) part[b * pstride + blockIdx.x] = acc;
  }
}

// ---------------------------------------------------------------------------
// fused passes: a tile of 2^k amplitudes (k tile bits) is staged in LDS,
// register groups of 2^R amplitudes per thread apply runs of gates, and the
// tile is written back -- one HBM read + write for all gates of the pass.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t block_base(const DPass* P, uint64_t blk) {
  uint64_t base = 0;
  for (int i = 0; i < P->nblk_log; ++i)
    if ((blk >> i) & 1) base |= 1ull << P->blk_bits[i];
  return base;
}

// Body of one fused forward pass over tile `blk` of batch item `b`; threads
// tid >= nthr (a smaller tile than the block) only take part in barriers.
template <typename T, int R, bool FULL>
__device__ __forceinline__ void fused_fwd_tile(const DPass* P, const DGroup* __restrict__ groups,
                                               const DOp* __restrict__ ops, const uint32_t* __restrict__ runoff,
                                               const T* mt, const T* __restrict__ ct, V2<T>* __restrict__ s,
                                               V2<T>* __restrict__ c, int nthr, V2<T>* tile) {
  const int k = P->k, L = P->L;
  const int tid = threadIdx.x;
  const bool act = tid < nthr;
  const uint32_t* ro = runoff + P->runoff0;
  uint32_t dlo = act ? tile_dep(ro, L, tid) : 0u;
  uint32_t slo = swz(tid);
  if (act) {
    // all loads in flight before the first LDS write
    V2<T> buf[1 << R];
#pragma unroll
    for (int it = 0; it < (1 << R); ++it) buf[it] = s[tile_dep(ro, L, it * nthr) + dlo];
#pragma unroll
    for (int it = 0; it < (1 << R); ++it) tile[swz(it * nthr) ^ slo] = buf[it];
  }
  __syncthreads();
  for (int g = 0; g < P->ngroups; ++g) {
    const DGroup* G = groups + P->group0 + g;
    if (act) {
      const uint32_t a0 = swz(group_base(G, tid));
      V2<T> v[1 << R];
#pragma unroll
      for (int j = 0; j < (1 << R); ++j) v[j] = tile[a0 ^ (uint32_t)G->S[j]];
      for (int o = 0; o < G->nops; ++o) apply_op<T, R, false, FULL>(v, ops[G->op0 + o], mt, ct);
      const DGroup* Gs = G + opaque_zero();
      uint32_t as = a0;
      asm volatile("" : "+v"(as));
#pragma unroll
      for (int j = 0; j < (1 << R); ++j) tile[as ^ (uint32_t)Gs->S[j]] = v[j];
    }
    __syncthreads();
  }
  rematerialize(dlo, slo, ro);
  if (act) {
#pragma unroll
    for (int it = 0; it < (1 << R); ++it) {
      const uint32_t ah = it * nthr;
      const V2<T> x = tile[swz(ah) ^ slo];
      const uint32_t off = tile_dep(ro, L, ah) + dlo;
      s[off] = x;
      if (c) c[off] = x;
    }
  }
}

template <typename T, int R, int TB, bool FULL>
__global__ void __launch_bounds__(TB) k_fused_fwd(const DPass* __restrict__ passes, int pi,
                                                  const DGroup* __restrict__ groups, const DOp* __restrict__ ops,
                                                  const uint32_t* __restrict__ runoff, const T* __restrict__ mtab,
                                                  int64_t mstride, const T* __restrict__ ct,
                                                  V2<T>* __restrict__ st, int64_t sstride,
                                                  V2<T>* __restrict__ ckout) {
  extern __shared__ __align__(16) unsigned char smem[];
  const DPass* P = passes + pi;
  const int nbl = P->nblk_log;
  const uint64_t blk = blockIdx.x & ((1ull << nbl) - 1);
  const uint64_t b = (uint64_t)blockIdx.x >> nbl;
  const uint64_t base = block_base(P, blk);
  fused_fwd_tile<T, R, FULL>(P, groups, ops, runoff, mtab + b * mstride, ct, st + b * sstride + base,
                             ckout ? ckout + b * sstride + base : nullptr, blockDim.x,
                             reinterpret_cast<V2<T>*>(smem));
}

// Several programs' passes in one launch (circuits of different qubit
// counts): block -> job by the prefix of the jobs' block counts.
template <typename T, int R, int TB>
__global__ void __launch_bounds__(TB) k_fused_fwd_multi(MultiJobs J) {
  extern __shared__ __align__(16) unsigned char smem[];
  int j = 0;
  while (j + 1 < J.njobs && (int64_t)blockIdx.x >= J.job[j + 1].block0) ++j;
  j = __builtin_amdgcn_readfirstlane(j);  // block-uniform
  const MultiJob& Q = J.job[j];
  // the job's fields are block-uniform: keep them in scalar registers
  const DPass* P = uniform_ptr((const DPass*)Q.dpass) + __builtin_amdgcn_readfirstlane(Q.pass);
  const DGroup* groups = uniform_ptr((const DGroup*)Q.dgroup);
  const DOp* ops = uniform_ptr((const DOp*)Q.dop);
  const uint32_t* runoff = uniform_ptr((const uint32_t*)Q.runoff);
  const T* ct = uniform_ptr((const T*)Q.ctab);
  const int nbl = P->nblk_log;
  const uint64_t lb = (uint64_t)blockIdx.x - Q.block0;
  const uint64_t blk = lb & ((1ull << nbl) - 1);
  const uint64_t b = lb >> nbl;
  const uint64_t base = block_base(P, blk);
  const T* mt = uniform_ptr((const T*)Q.mtab + b * Q.mstride);
  V2<T>* s = uniform_ptr((V2<T>*)Q.state + b * Q.sstride + base);
  if (P->full)
    fused_fwd_tile<T, R, true>(P, groups, ops, runoff, mt, ct, s, nullptr, 1 << P->tthreads_log,
                               reinterpret_cast<V2<T>*>(smem));
  else
    fused_fwd_tile<T, R, false>(P, groups, ops, runoff, mt, ct, s, nullptr, 1 << P->tthreads_log,
                                reinterpret_cast<V2<T>*>(smem));
}

// Reverse sweep over one fused pass.  psi/lambda tiles are read from
// psi_in/lam_in, gates are undone in reverse order, gradient partials of the
// pass's trainable gates are reduced per tile into part[b][gloc][blk], and
// the tiles are written to psi_out (skipped when NULL: a checkpoint replaces
// it) and lam_out.
template <typename T, int R, int TB, bool FULL>
__global__ void __launch_bounds__(TB) k_fused_adj(const DPass* __restrict__ passes, int pi,
                                                    const DGroup* __restrict__ groups, const DOp* __restrict__ ops,
                                                    const DMember* __restrict__ mems,
                                                    const uint32_t* __restrict__ runoff, const T* __restrict__ mtab,
                                                    int64_t mstride, const T* __restrict__ ct,
                                                    const V2<T>* psi_in, V2<T>* psi_out,
                                                    const V2<T>* lam_in, V2<T>* lam_out, int64_t sstride,
                                                    double* __restrict__ part, int64_t pstride) {
  extern __shared__ __align__(16) unsigned char smem[];
  const DPass* P = passes + pi;
  const int k = P->k, L = P->L, nbl = P->nblk_log;
  V2<T>* tp = reinterpret_cast<V2<T>*>(smem);
  V2<T>* tl = tp + (1 << k);
  // red[gloc][wave]: fp64 wave sums of each trainable gate of the pass
  double* red = reinterpret_cast<double*>(tl + (1 << k));
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lane = tid & 63, wv = tid >> 6, nw = nthr >> 6;
  const uint64_t blk = blockIdx.x & ((1ull << nbl) - 1);
  const uint64_t b = (uint64_t)blockIdx.x >> nbl;
  const uint64_t base = block_base(P, blk);
  const T* mt = mtab + b * mstride;
  const uint32_t* ro = runoff + P->runoff0;
  uint32_t dlo = tile_dep(ro, L, tid);
  uint32_t slo = swz(tid);
  {
    const V2<T>* sp = psi_in + b * sstride + base;
    const V2<T>* sl = lam_in + b * sstride + base;
    V2<T> bp[1 << R], bl[1 << R];
#pragma unroll
    for (int it = 0; it < (1 << R); ++it) {
      const uint32_t off = tile_dep(ro, L, it * nthr) + dlo;
      bp[it] = sp[off];
      bl[it] = sl[off];
    }
#pragma unroll
    for (int it = 0; it < (1 << R); ++it) {
      tp[swz(it * nthr) ^ slo] = bp[it];
      tl[swz(it * nthr) ^ slo] = bl[it];
    }
  }
  __syncthreads();
  for (int g = P->ngroups - 1; g >= 0; --g) {
    const DGroup* G = groups + P->group0 + g;
    const uint32_t a0 = swz(group_base(G, tid));
    V2<T> p[1 << R], l[1 << R];
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
      p[j] = tp[a0 ^ (uint32_t)G->S[j]];
      l[j] = tl[a0 ^ (uint32_t)G->S[j]];
    }
    for (int o = G->nops - 1; o >= 0; --o) {
      const DOp& op = ops[G->op0 + o];
      if (op.kind == OK_DG && op.nmem > 0) {
        T im[1 << R];
#pragma unroll
        for (int j = 0; j < (1 << R); ++j) im[j] = imcj<T>(l[j], p[j]);
        for (int q = 0; q < op.nmem; ++q) {
          const DMember& M = mems[op.mem0 + q];
          const double gs = wave_sum_dpp((double)grad_member<T, R>(im, M.gen, M.s0, M.s1));
          if (lane == 0) red[M.gloc * nw + wv] = gs;
        }
      } else if (op.gen != GN_NONE) {
        const double gs = wave_sum_dpp((double)grad_op<T, R, FULL>(p, l, op.gen, op.s0, op.s1, ct, op.goff));
        if (lane == 0) red[op.gloc * nw + wv] = gs;
      }
      apply_op<T, R, true, FULL>(p, op, mt, ct);
      apply_op<T, R, true, FULL>(l, op, mt, ct);
    }
    const DGroup* Gs = G + opaque_zero();
    uint32_t as = a0;
    asm volatile("" : "+v"(as));
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
      tp[as ^ (uint32_t)Gs->S[j]] = p[j];
      tl[as ^ (uint32_t)Gs->S[j]] = l[j];
    }
    __syncthreads();
  }
  // recompute the store addresses instead of holding the load addresses in
  // registers through the group loop
  rematerialize(dlo, slo, ro);
  {
    V2<T>* sp = psi_out ? psi_out + b * sstride + base : nullptr;
    V2<T>* sl = lam_out + b * sstride + base;
#pragma unroll
    for (int it = 0; it < (1 << R); ++it) {
      const uint32_t ah = it * nthr;
      const uint32_t off = tile_dep(ro, L, ah) + dlo;
      if (sp) sp[off] = tp[swz(ah) ^ slo];
      sl[off] = tl[swz(ah) ^ slo];
    }
  }
  const int ng = P->ngop;
  for (int q = tid; q < ng; q += nthr) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[q * nw + w];
    part[b * pstride + (int64_t)q * (1ll << nbl) + blk] = s;
  }
}

// ---------------------------------------------------------------------------
// dense 5-qubit gate on FP32 matrix cores.  The state is viewed as a
// 32 x (2^(n-5)) matrix X (rows = the 5 gate bits, row MSB = gate qubit 0,
// columns = all other bits) and Y = U X is computed per 32-column chunk by
// one wave with v_mfma_f32_32x32x2f32 (K = 2 per instruction), as three
// real products (Gauss's complex multiplication):
//   P1 = Ur Xr,  P2 = Ui Xi,  P3 = (Ur + Ui)(Xr + Xi),
//   Yr = P1 - P2,  Yi = P3 - P1 - P2.
// Operand layouts (CDNA3/4): lane l supplies A[l % 32][l / 32] and
// B[l / 32][l % 32]; accumulator register r of lane l holds
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
// ---------------------------------------------------------------------------
typedef float tcm_v16f __attribute__((ext_vector_type(16)));
// two complex64 amplitudes (one 16-byte access; a native vector, so copies
// stay in registers)
typedef float tcm_v4f __attribute__((ext_vector_type(4)));

template <bool DAG>
__global__ void __launch_bounds__(256) k_mfma_u5(DOp op, const int32_t* __restrict__ ukbits,
                                                 const float* __restrict__ ct, float2* __restrict__ st,
                                                 int64_t sstride, int n, int64_t batch) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const float* U = ct + (-op.moff - 1);
  int bits[5], sorted[5];
  for (int i = 0; i < 5; ++i) sorted[i] = bits[i] = ukbits[op.mem0 + i];
  for (int i = 0; i < 5; ++i)
    for (int j = i + 1; j < 5; ++j)
      if (sorted[j] < sorted[i]) { int t = sorted[i]; sorted[i] = sorted[j]; sorted[j] = t; }
  // A operands for the 16 K-steps: U[i][2s + l/32] (U^dag: conj(U[2s + l/32][i]))
  const int ai = lane & 31, ak = lane >> 5;
  float ur[16], ui[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int kk = 2 * s + ak;
    if (DAG) {
      ur[s] = U[2 * (kk * 32 + ai)];
      ui[s] = -U[2 * (kk * 32 + ai) + 1];
    } else {
      ur[s] = U[2 * (ai * 32 + kk)];
      ui[s] = U[2 * (ai * 32 + kk) + 1];
    }
  }
  // index offset of gate row r (5 bits, MSB = bits[0])
  auto rowoff = [&](int r) {
    uint64_t o = 0;
#pragma unroll
    for (int i = 0; i < 5; ++i)
      if ((r >> (4 - i)) & 1) o |= 1ull << bits[i];
    return o;
  };
  const int64_t ncol = 1ll << (n - 5);
  const int64_t chunks_per_state = ncol >> 5;
  const int64_t nchunks = chunks_per_state * batch;
  // chunk ch: 32 consecutive columns of one state; the next chunk's loads
  // are issued before this chunk's matrix-core work (software pipelining)
  auto chunk_ptr = [&](int64_t ch) {
    const int64_t b = ch / chunks_per_state;
    const uint64_t c = ((uint64_t)(ch % chunks_per_state) << 5) + (lane & 31);
    uint64_t base = c;
#pragma unroll
    for (int i = 0; i < 5; ++i) base = ins0(base, sorted[i]);
    return st + b * sstride + base;
  };
  if (wave >= nchunks) return;
  float2* s = chunk_ptr(wave);
  float2 x[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) x[q] = s[rowoff(2 * q + ak)];
  for (int64_t ch = wave; ch < nchunks; ch += nwaves) {
    const bool more = ch + nwaves < nchunks;
    float2* sn = more ? chunk_ptr(ch + nwaves) : s;
    float2 xn[16];
    if (more) {
#pragma unroll
      for (int q = 0; q < 16; ++q) xn[q] = sn[rowoff(2 * q + ak)];
    }
    tcm_v16f p1 = {0}, p2 = {0}, p3 = {0};
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      p1 = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[q], x[q].x, p1, 0, 0, 0);
      p2 = __builtin_amdgcn_mfma_f32_32x32x2f32(ui[q], x[q].y, p2, 0, 0, 0);
      p3 = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[q] + ui[q], x[q].x + x[q].y, p3, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r >> 2) + 4 * ak + (r & 3);
      s[rowoff(row)] = make_float2(p1[r] - p2[r], p3[r] - p1[r] - p2[r]);
    }
    if (more) {
#pragma unroll
      for (int q = 0; q < 16; ++q) x[q] = xn[q];
      s = sn;
    }
  }
}

// f(0), f(1), ..., f(N - 1) with compile-time indices (register arrays
// indexed by them stay in registers)
template <int N, typename F>
__device__ __forceinline__ void unrolled(F&& f) {
  if constexpr (N > 0) {
    unrolled<N - 1>(f);
    f(std::integral_constant<int, N - 1>{});
  }
}

// One tile pass of dense 5-qubit blocks on FP32 matrix cores.  A tile is the
// 2^K amplitudes whose index bits are the pass's tile bits; each block
// multiplies the tile, viewed as 32 rows (the block's 5 tile bits, first bit
// most significant) x 2^(K-5) columns (the other tile bits, ascending), by
// its 32x32 complex matrix U.  Wave w takes the 32-column chunk w and forms
// Y = U X from three real products (Gauss's complex multiplication):
//   P1 = Ur Xr,  P2 = Ui Xi,  P3 = (Ur + Ui)(Xr + Xi),
//   Yr = P1 - P2,  Yi = P3 - P1 - P2,
// 48 v_mfma_f32_32x32x2f32 per chunk instead of the 64 of four products.
// Waves own disjoint columns, so only the block boundary needs a barrier.
// The kernel is persistent: a block stages the A operands of all the pass's
// blocks in LDS once, then runs tiles blockIdx.x, blockIdx.x + gridDim.x, ...
// and loads the next tile's amplitudes into registers while the matrix
// cores work on the current one.  Row offsets are XORs of the 5 gate bits'
// swizzled offsets (the swizzle is linear over XOR).  512 threads; LDS =
// 8 << K bytes (tile) + nmop * (8 KiB A operands, [block][K-step][lane], +
// the gate-bit and column offsets).
template <int K>
__global__ void __launch_bounds__(512) k_mfma_tile(const DPass* __restrict__ passes, int pi,
                                                   const DOp* __restrict__ dop, const int32_t* __restrict__ ukbits,
                                                   const uint32_t* __restrict__ runoff, const float* __restrict__ ct,
                                                   float2* __restrict__ st, int64_t sstride, int64_t ntiles) {
  // each thread moves PER2 pairs of adjacent amplitudes (16-byte accesses)
  constexpr int NTHR = 512, PER2 = (1 << K) / (2 * NTHR), CHUNKS = 1 << (K - 10), NCOL = 1 << (K - 5);
  extern __shared__ __align__(16) unsigned char smem[];
  float2* tile = reinterpret_cast<float2*>(smem);
  const DPass* P = uniform_ptr(passes + pi);
  const int L = P->L, nbl = P->nblk_log, nmop = P->nmop;
  float2* amat = tile + (1 << K);
  // per block: the swizzled offsets of its 5 gate bits, then the swizzled
  // offsets of its NCOL columns
  uint32_t* ebits = reinterpret_cast<uint32_t*>(amat + nmop * 1024);
  uint32_t* xcol = ebits + nmop * 8;
  const int tid = threadIdx.x;
  int64_t t = blockIdx.x;
  if (t >= ntiles) return;
  const DOp* ops = dop + P->mop0;
  const uint32_t* ro = runoff + P->runoff0;
  // Every table read happens here, before the first store: reads inside the
  // tile loop would be vector loads queued behind the prefetch and the stores.
  // A operands: lane l of K-step q supplies U[l % 32][2q + l / 32]
  for (int idx = tid; idx < nmop * 1024; idx += NTHR) {
    const int o = idx >> 10, q = (idx >> 6) & 15, ln = idx & 63;
    const float* U = ct + (-ops[o].moff - 1);
    const int e = 2 * ((ln & 31) * 32 + 2 * q + (ln >> 5));
    amat[idx] = make_float2(U[e], U[e + 1]);
  }
  if (tid < nmop * 8) {
    const int32_t* ub = ukbits + ops[tid >> 3].mem0;
    const int i = tid & 7;
    ebits[tid] = i < 5 ? swz(1u << ub[i]) : 0u;
  }
  // column c of a block -> the tile bits outside the block, ascending
  for (int idx = tid; idx < nmop * NCOL; idx += NTHR) {
    const int32_t* ub = ukbits + ops[idx / NCOL].mem0;
    uint32_t bm = 0;
    for (int j = 0; j < 5; ++j) bm |= 1u << ub[j];
    const uint32_t col = (uint32_t)(idx % NCOL);
    uint32_t cb = 0;
    for (int b = 0, j = 0; b < K; ++b)
      if (!((bm >> b) & 1)) {
        if ((col >> j) & 1) cb |= 1u << b;
        ++j;
      }
    xcol[idx] = swz(cb);
  }
  uint64_t bmask = 0;
  for (int i = 0; i < nbl; ++i) bmask |= 1ull << P->blk_bits[i];
  // pair it of this thread: tile elements x = 2 (it * NTHR + tid) and x + 1,
  // adjacent in the state (the low L >= 4 tile bits are contiguous) and in
  // LDS, at slots swz(x) and swz(x) ^ 1: in reverse order when swz(x) is odd
  // (the swizzle folds bits 4, 8, 12 into bit 0)
  uint32_t roff[PER2];
  unrolled<PER2>([&](auto it) { roff[it] = tile_dep(ro, L, it * 2 * NTHR); });
  const uint32_t dlo = tile_dep(ro, L, 2 * tid);
  const uint32_t slo = swz(2 * tid);
  tcm_v4f* tile4 = reinterpret_cast<tcm_v4f*>(smem);
  auto put4 = [&](int it, tcm_v4f v) {
    const uint32_t i = swz(it * 2 * NTHR) ^ slo;
    if (i & 1) v = v.zwxy;
    tile4[i >> 1] = v;
  };
  auto get4 = [&](int it) -> tcm_v4f {
    const uint32_t i = swz(it * 2 * NTHR) ^ slo;
    tcm_v4f v = tile4[i >> 1];
    if (i & 1) v = v.zwxy;
    return v;
  };
  // tile tt: batch item tt >> nbl; its block bits (ascending) hold the low nbl bits of tt
  auto tile_ptr = [&](int64_t tt) {
    const uint64_t blk = (uint64_t)tt & ((1ull << nbl) - 1);
    uint64_t m = bmask, base = 0;
    for (int i = 0; i < nbl; ++i) {
      const uint64_t low = m & (~m + 1);
      if ((blk >> i) & 1) base |= low;
      m ^= low;
    }
    return st + ((uint64_t)tt >> nbl) * sstride + base;
  };
  auto glb4 = [&](float2* sp, int it) -> tcm_v4f* { return reinterpret_cast<tcm_v4f*>(sp + roff[it] + dlo); };
  const int lane = tid & 63, wave = tid >> 6;
  const int ai = lane & 31, ak = lane >> 5;
  // A operands of the pass's first two blocks (filled once staged)
  float2 a0[16], a1[16];
  // block o of the pass on the tile in LDS; geta(q) gives its A operand of K-step q
  auto block = [&](int o, auto geta) {
    if (wave < CHUNKS) {
      const uint32_t* ob = ebits + o * 8;
      uint32_t e[5];
#pragma unroll
      for (int i = 0; i < 5; ++i) e[i] = __builtin_amdgcn_readfirstlane(ob[i]);
      const uint32_t xc = xcol[o * NCOL + wave * 32 + ai];
      // read rows 2q + ak; written rows 8 (r / 4) + 4 ak + r % 4
      const uint32_t rb = xc ^ (ak ? e[4] : 0u);
      const uint32_t wb = xc ^ (ak ? e[2] : 0u);
      float2 a[16], x[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        a[q] = geta(q);
        const uint32_t g = ((q & 1) ? e[3] : 0u) ^ ((q & 2) ? e[2] : 0u) ^ ((q & 4) ? e[1] : 0u) ^
                           ((q & 8) ? e[0] : 0u);
        x[q] = tile[rb ^ g];
      }
      tcm_v16f p1 = {0}, p2 = {0}, p3 = {0};
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        p1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q].x, x[q].x, p1, 0, 0, 0);
        p2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q].y, x[q].y, p2, 0, 0, 0);
        p3 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q].x + a[q].y, x[q].x + x[q].y, p3, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const uint32_t h = ((r & 1) ? e[4] : 0u) ^ ((r & 2) ? e[3] : 0u) ^ ((r & 4) ? e[1] : 0u) ^
                           ((r & 8) ? e[0] : 0u);
        tile[wb ^ h] = make_float2(p1[r] - p2[r], p3[r] - p1[r] - p2[r]);
      }
    }
    __syncthreads();
  };
  // the pass's blocks in order, each followed by a barrier
  auto blocks = [&]() {
    block(0, [&](int q) { return a0[q]; });
    if (nmop > 1) block(1, [&](int q) { return a1[q]; });
    for (int o = 2; o < nmop; ++o) block(o, [&](int q) { return amat[o * 1024 + q * 64 + lane]; });
  };
  float2* s = tile_ptr(t);
  tcm_v4f pre[PER2];
  unrolled<PER2>([&](auto it) { pre[it] = *glb4(s, it); });
  unrolled<PER2>([&](auto it) { put4(it, pre[it]); });
  // the first two blocks' A operands stay in registers for every tile
  __syncthreads();
  unrolled<16>([&](auto q) {
    a0[q] = amat[q * 64 + lane];
    a1[q] = amat[(nmop > 1 ? 1024 : 0) + q * 64 + lane];
  });
  // every tile but the last: the next tile's loads are issued before this
  // tile's matrix-core work; each thread then stores the LDS slots it wrote
  // and refills them with the next tile (same slots: no barrier between)
  for (;;) {
    const int64_t tn = t + gridDim.x;
    if (tn >= ntiles) break;
    float2* sn = tile_ptr(tn);
    unrolled<PER2>([&](auto it) { pre[it] = *glb4(sn, it); });
    __syncthreads();  // the tile (and, first time, the staged operands) in LDS
    blocks();
    unrolled<PER2>([&](auto it) {
      *glb4(s, it) = get4(it);
      put4(it, pre[it]);
    });
    t = tn;
    s = sn;
  }
  __syncthreads();
  blocks();
  unrolled<PER2>([&](auto it) { *glb4(s, it) = get4(it); });
}

// Blocks of `kern` (512 threads, `lds` bytes of LDS) resident on the whole
// device at once: the grid of a persistent kernel.
template <typename Kern>
static int64_t resident_grid(Kern kern, size_t lds) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void*, size_t>, int64_t> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  std::lock_guard<std::mutex> lk(mu);
  const auto key = std::make_tuple(dev, (const void*)kern, lds);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 512, lds) != hipSuccess || per_cu < 1) per_cu = 1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1) ncu = 256;
  return cache[key] = (int64_t)per_cu * ncu;
}

template <typename K>
static void ensure_lds_attr(K kern, size_t bytes);

template <int K>
static void launch_mfma_tile_k(const DPass& P, const DPass* passes_dev, int pi, const DOp* dop_dev,
                               const int32_t* ukbits_dev, const uint32_t* runoff_dev, const void* ctab, void* state,
                               int n, int64_t batch, hipStream_t st) {
  const size_t lds = ((size_t)8 << K) + (size_t)P.nmop * (8192 + 32 + (4 << (K - 5)));
  ensure_lds_attr(k_mfma_tile<K>, lds);
  const int64_t ntiles = batch << P.nblk_log;
  const int64_t grid = std::min<int64_t>(ntiles, resident_grid(k_mfma_tile<K>, lds));
  hipLaunchKernelGGL(k_mfma_tile<K>, dim3((unsigned)grid), dim3(512), lds, st, passes_dev, pi, dop_dev, ukbits_dev,
                     runoff_dev, (const float*)ctab, (float2*)state, (int64_t)1 << n, ntiles);
}

bool launch_mfma_tile(const DPass& P, const DPass* passes_dev, int pi, const DOp* dop_dev, const int32_t* ukbits_dev,
                      const uint32_t* runoff_dev, const void* ctab, void* state, int n, int64_t batch,
                      hipStream_t st) {
  if (P.nmop < 1 || P.nmop > kMfmaTileMaxOps || P.nblk_log != n - P.k) return false;
  switch (P.k) {
    case 10: launch_mfma_tile_k<10>(P, passes_dev, pi, dop_dev, ukbits_dev, runoff_dev, ctab, state, n, batch, st); return true;
    case 11: launch_mfma_tile_k<11>(P, passes_dev, pi, dop_dev, ukbits_dev, runoff_dev, ctab, state, n, batch, st); return true;
    case 12: launch_mfma_tile_k<12>(P, passes_dev, pi, dop_dev, ukbits_dev, runoff_dev, ctab, state, n, batch, st); return true;
    case 13: launch_mfma_tile_k<13>(P, passes_dev, pi, dop_dev, ukbits_dev, runoff_dev, ctab, state, n, batch, st); return true;
    default: return false;
  }
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// out[b * ncomp + c] = sum_s part[b * bstride + s * ncomp + c]
__global__ void k_sum_slots(const double* __restrict__ part, int64_t bstride, int64_t nslots, int ncomp,
                            double* __restrict__ out) {
  const int64_t b = blockIdx.x;
  double a = 0.0, c = 0.0;
  for (int64_t s = threadIdx.x; s < nslots; s += blockDim.x) {
    a += part[b * bstride + s * ncomp];
    if (ncomp > 1) c += part[b * bstride + s * ncomp + 1];
  }
  block_sum2(a, c);
  if (threadIdx.x == 0) {
    out[b * ncomp] = a;
    if (ncomp > 1) out[b * ncomp + 1] = c;
  }
}

// gsum[b][gop] = sum over slots of part[b][gl][slot]; grid (ngop_pass, batch)
__global__ void k_reduce_gop(const double* __restrict__ part, int64_t pstride, int64_t nslots,
                             const int32_t* __restrict__ gop_list, int gop_single, double* __restrict__ gsum,
                             int ngop) {
  const int gl = blockIdx.x;
  const int64_t b = blockIdx.y;
  double a = 0.0, d = 0.0;
  for (int64_t s = threadIdx.x; s < nslots; s += blockDim.x) a += part[b * pstride + gl * nslots + s];
  block_sum2(a, d);
  if (threadIdx.x == 0) {
    const int gop = gop_list ? gop_list[gl] : gop_single;
    gsum[b * ngop + gop] = a;
  }
}

// grad[b][p] = mult * sum_{gop of p} gfac[gop] * gsum[b][gop]
__global__ void k_grad_assemble(const double* __restrict__ gsum, int ngop, const int32_t* __restrict__ p_ptr,
                                const int32_t* __restrict__ p_gops, const double* __restrict__ gfac, double mult,
                                double* __restrict__ grad, int nparams, int64_t batch) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)nparams * batch) return;
  const int64_t b = i / nparams;
  const int p = (int)(i % nparams);
  double a = 0.0;
  for (int q = p_ptr[p]; q < p_ptr[p + 1]; ++q) a += gfac[p_gops[q]] * gsum[b * ngop + p_gops[q]];
  grad[i] = mult * a;
}

// ---------------------------------------------------------------------------
// Pauli sums: one launch per flip group (terms sharing the X/Y positions).
// (P psi)[y] = i^{ny} (-1)^{popcount((y ^ flip) & yz)} psi[y ^ flip]
// (index rule of ps2coo_core, quantum.py:1461-1485)
// ---------------------------------------------------------------------------
struct PauliAcc {
  double re, im;
};

__device__ __forceinline__ PauliAcc pauli_coef(const DPTerm* __restrict__ terms, int t0, int nt, uint64_t y) {
  // sum_k w_k i^{ny_k} (-1)^{popcount(y & yz_k)}, y = the source index y ^ flip
  double re = 0.0, im = 0.0;
  for (int t = t0; t < t0 + nt; ++t) {
    const DPTerm tm = terms[t];
    double w = (__popcll(y & tm.yz) & 1) ? -tm.w : tm.w;
    switch (tm.ny & 3) {
      case 0: re += w; break;
      case 1: im += w; break;
      case 2: re -= w; break;
      default: im -= w; break;
    }
  }
  return {re, im};
}

template <typename T, bool ENERGY, bool APPLY>
__global__ void __launch_bounds__(256) k_pauli(DPGroup G, const DPTerm* __restrict__ terms, int n,
                                               const V2<T>* __restrict__ psi, int64_t sstride,
                                               V2<T>* __restrict__ out, int first, const double* __restrict__ alpha,
                                               double* __restrict__ epart, int64_t estride) {
  const int64_t b = blockIdx.y;
  const V2<T>* s = psi + b * sstride;
  V2<T>* o = APPLY ? out + b * sstride : nullptr;
  const double al = (APPLY && alpha) ? alpha[b] : 1.0;
  const int64_t N = 1ll << n;
  double er = 0.0, ei = 0.0;
  if (G.flip) {
    const int h = 63 - __clzll(G.flip);
    const int64_t items = N >> 1;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < items; idx += (int64_t)gridDim.x * blockDim.x) {
      const uint64_t y0 = ins0(idx, h), y1 = y0 ^ G.flip;
      const V2<T> p0 = s[y0], p1 = s[y1];
      const PauliAcc a0 = pauli_coef(terms, G.t0, G.nt, y1);  // (H psi)[y0] = a0 * psi[y1]
      const PauliAcc a1 = pauli_coef(terms, G.t0, G.nt, y0);
      const double h0r = a0.re * p1.x - a0.im * p1.y, h0i = a0.re * p1.y + a0.im * p1.x;
      const double h1r = a1.re * p0.x - a1.im * p0.y, h1i = a1.re * p0.y + a1.im * p0.x;
      if constexpr (ENERGY) {
        er += (double)p0.x * h0r + (double)p0.y * h0i + (double)p1.x * h1r + (double)p1.y * h1i;
        ei += (double)p0.x * h0i - (double)p0.y * h0r + (double)p1.x * h1i - (double)p1.y * h1r;
      }
      if constexpr (APPLY) {
        if (first) {
          o[y0] = mk<T>((T)(al * h0r), (T)(al * h0i));
          o[y1] = mk<T>((T)(al * h1r), (T)(al * h1i));
        } else {
          V2<T> q0 = o[y0], q1 = o[y1];
          o[y0] = mk<T>((T)(q0.x + al * h0r), (T)(q0.y + al * h0i));
          o[y1] = mk<T>((T)(q1.x + al * h1r), (T)(q1.y + al * h1i));
        }
      }
    }
  } else {
    for (int64_t y = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; y < N; y += (int64_t)gridDim.x * blockDim.x) {
      const V2<T> p = s[y];
      const PauliAcc a = pauli_coef(terms, G.t0, G.nt, (uint64_t)y);
      const double hr = a.re * p.x - a.im * p.y, hi = a.re * p.y + a.im * p.x;
      if constexpr (ENERGY) {
        er += (double)p.x * hr + (double)p.y * hi;
        ei += (double)p.x * hi - (double)p.y * hr;
      }
      if constexpr (APPLY) {
        if (first) o[y] = mk<T>((T)(al * hr), (T)(al * hi));
        else {
          V2<T> q = o[y];
          o[y] = mk<T>((T)(q.x + al * hr), (T)(q.y + al * hi));
        }
      }
    }
  }
  if constexpr (ENERGY) {
    block_sum2(er, ei);
    if (threadIdx.x == 0) {
      epart[b * estride + 2 * (G.slot0 + blockIdx.x)] = er;
      epart[b * estride + 2 * (G.slot0 + blockIdx.x) + 1] = ei;
    }
  }
}

// LDS-tiled Pauli pass.  For every amplitude a of the tile:
//   h[a] = sum_groups sum_terms w i^ny (-1)^{popcount((y ^ flip) & yz)} psi[y ^ flip]
// (the index rule of ps2coo_core, quantum.py:1461-1485), with the partner
// read from LDS; energy += conj(psi[a]) h[a] in fp64 and, when APPLY,
// lam_out[a] = lam_in[a] (or 0) + alpha * h[a].
template <typename T, bool ENERGY, bool APPLY>
__global__ void __launch_bounds__(256) k_pauli_tile(const DPPass* __restrict__ passes, int pi,
                                                    const DPGroupT* __restrict__ groups,
                                                    const DPTermT* __restrict__ terms,
                                                    const uint32_t* __restrict__ runoff,
                                                    const V2<T>* __restrict__ psi, const V2<T>* lam_in,
                                                    V2<T>* lam_out, int64_t sstride,
                                                    const double* __restrict__ alpha,
                                                    double* __restrict__ epart, int64_t estride) {
  extern __shared__ __align__(16) unsigned char smem[];
  V2<T>* tile = reinterpret_cast<V2<T>*>(smem);
  const DPPass* P = passes + pi;
  const int k = P->k, L = P->L, nbl = P->nblk_log;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const uint64_t blk = blockIdx.x & ((1ull << nbl) - 1);
  const uint64_t b = (uint64_t)blockIdx.x >> nbl;
  uint64_t base = 0;
  for (int i = 0; i < nbl; ++i)
    if ((blk >> i) & 1) base |= 1ull << P->blk_bits[i];
  const V2<T>* s = psi + b * sstride + base;
  const uint32_t* ro = runoff + P->runoff0;
  const uint32_t dlo = tile_dep(ro, L, tid);
  const uint32_t slo = swz(tid);
  const int per = (1 << k) / nthr;
  for (int it = 0; it < per; ++it) {
    const uint32_t ah = it * nthr;
    tile[swz(ah) ^ slo] = s[tile_dep(ro, L, ah) + dlo];
  }
  __syncthreads();
  const double al = (APPLY && alpha) ? alpha[b] : 1.0;
  double er = 0.0, ei = 0.0;
  for (int it = 0; it < per; ++it) {
    const uint32_t a = it * nthr + tid;
    const V2<T> p = tile[swz(a)];
    T hr = 0, hi = 0;
    for (int g = P->g0; g < P->g0 + P->ng; ++g) {
      const DPGroupT G = groups[g];
      const uint32_t src = a ^ G.flip_t;
      const V2<T> q = G.flip_t ? tile[swz(src)] : p;
      T cr = 0, ci = 0;
      for (int t = G.t0; t < G.t0 + G.nt; ++t) {
        const DPTermT tm = terms[t];
        const int par = (__popc(src & tm.yz_t) + __popcll(base & tm.yz_g)) & 1;
        const T w = par ? (T)-tm.w : (T)tm.w;
        switch (tm.ny & 3) {
          case 0: cr += w; break;
          case 1: ci += w; break;
          case 2: cr -= w; break;
          default: ci -= w; break;
        }
      }
      hr += cr * q.x - ci * q.y;
      hi += cr * q.y + ci * q.x;
    }
    if constexpr (ENERGY) {
      er += (double)p.x * hr + (double)p.y * hi;
      ei += (double)p.x * hi - (double)p.y * hr;
    }
    if constexpr (APPLY) {
      const uint64_t off = b * sstride + base + tile_dep(ro, L, it * nthr) + dlo;
      V2<T> o = lam_in ? lam_in[off] : mk<T>(0, 0);
      lam_out[off] = mk<T>((T)(o.x + al * hr), (T)(o.y + al * hi));
    }
  }
  if constexpr (ENERGY) {
    block_sum2(er, ei);
    if (threadIdx.x == 0) {
      epart[b * estride + 2 * (P->slot0 + blk)] = er;
      epart[b * estride + 2 * (P->slot0 + blk) + 1] = ei;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_inner(const V2<T>* __restrict__ a, const V2<T>* __restrict__ c, int n,
                                               int64_t sstride, double* __restrict__ part, int64_t pstride) {
  const int64_t b = blockIdx.y;
  const int64_t N = 1ll << n;
  double re = 0.0, im = 0.0;
  for (int64_t y = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; y < N; y += (int64_t)gridDim.x * blockDim.x) {
    const V2<T> x = a[b * sstride + y], z = c[b * sstride + y];
    re += (double)x.x * z.x + (double)x.y * z.y;
    im += (double)x.x * z.y - (double)x.y * z.x;
  }
  block_sum2(re, im);
  if (threadIdx.x == 0) {
    part[b * pstride + 2 * blockIdx.x] = re;
    part[b * pstride + 2 * blockIdx.x + 1] = im;
  }
}

template <typename T>
__global__ void k_copy(V2<T>* __restrict__ dst, const V2<T>* __restrict__ src, int64_t total) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ===========================================================================
// host launchers
// ===========================================================================
static inline int grid_for(int64_t work, int64_t cap = 1 << 20) {
  int64_t g = (work + 255) / 256;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

template <typename T>
void launch_prep(const Sched& S, const DeviceTables& D, const void* ctab, const void* params, int64_t pstride,
                 void* mtab, int64_t batch, hipStream_t st) {
  if (S.recipes.empty() || batch <= 0) return;
  // merged-diagonal (DG) recipes follow the base recipes; their widest group
  const int nrec = (int)S.recipes.size();
  int first_dg = nrec, dg_log = 0;
  for (int i = 0; i < nrec; ++i)
    if (S.recipes[i].kind < 0) {
      if (first_dg == nrec) first_dg = i;
      dg_log = std::max(dg_log, (int)S.recipes[i].R);
    }
  hipLaunchKernelGGL(k_prep<T>, dim3((unsigned)batch), dim3(256), 0, st, (const DRecipe*)D.recipes, nrec, first_dg,
                     dg_log, (const DFactor*)D.factors, (const T*)params, pstride, (const T*)ctab, (T*)mtab,
                     (int64_t)S.mtab_size);
}

template <typename T>
void launch_init_zero(void* state, int n, int64_t batch, hipStream_t st) {
  const int64_t total = (1ll << n) * batch;
  hipLaunchKernelGGL(k_init_zero<T>, dim3(grid_for(total)), dim3(256), 0, st, (V2<T>*)state, total, 1ll << n);
}

template <typename T>
void launch_copy(void* dst, const void* src, int64_t total, hipStream_t st) {
  hipLaunchKernelGGL(k_copy<T>, dim3(grid_for(total)), dim3(256), 0, st, (V2<T>*)dst, (const V2<T>*)src, total);
}

template <typename T, bool DAG>
void launch_plain_gate(const DOp& op, const int32_t* ukbits_dev, const void* mtab, int64_t mstride,
                       const void* ctab, void* state, int n, int64_t batch, hipStream_t st) {
  const T* mt = (const T*)mtab;
  const T* ct = (const T*)ctab;
  V2<T>* s = (V2<T>*)state;
  const int64_t N = 1ll << n;
  switch (op.kind) {
    case OK_U1: case OK_R1: case OK_D1: case OK_X1:
      hipLaunchKernelGGL((k_plain1<T, 0, DAG>), dim3(grid_for(batch * (N >> 1))), dim3(256), 0, st, op, mt, mstride,
                         ct, s, N, n, batch);
      break;
    case OK_U2: case OK_D2: case OK_CX: case OK_CZ: case OK_SWAP:
      hipLaunchKernelGGL((k_plain2<T, 0, DAG>), dim3(grid_for(batch * (N >> 2))), dim3(256), 0, st, op, mt, mstride,
                         ct, s, N, n, batch);
      break;
    case OK_UK:
      if constexpr (std::is_same<T, float>::value) {
        if (op.nmem == 5 && (op.flags & OP_F_MFMA) && n >= 10) {
          const int64_t waves = batch * (N >> 10);
          const int blocks = (int)std::min<int64_t>((waves + 3) / 4, 1 << 20);
          hipLaunchKernelGGL((k_mfma_u5<DAG>), dim3(blocks), dim3(256), 0, st, op, ukbits_dev, ct, (float2*)s,
                             N, n, batch);
          break;
        }
      }
      switch (op.nmem) {
        case 3: hipLaunchKernelGGL((k_plaink<T, 3, DAG>), dim3(grid_for(batch * (N >> 3))), dim3(256), 0, st, op, ukbits_dev, ct, s, N, n, batch); break;
        case 4: hipLaunchKernelGGL((k_plaink<T, 4, DAG>), dim3(grid_for(batch * (N >> 4))), dim3(256), 0, st, op, ukbits_dev, ct, s, N, n, batch); break;
        case 5: hipLaunchKernelGGL((k_plaink<T, 5, DAG>), dim3(grid_for(batch * (N >> 5))), dim3(256), 0, st, op, ukbits_dev, ct, s, N, n, batch); break;
      }
      break;
    default: break;
  }
}

template <typename T>
static size_t fused_lds_fwd(int k) { return sizeof(V2<T>) << k; }
template <typename T>
static size_t fused_lds_adj(int k, int ngop, int nthr) {
  return (2 * sizeof(V2<T>) << k) + sizeof(double) * (size_t)(ngop > 0 ? ngop : 1) * (nthr / 64);
}

// kernels that need more than 64 KiB of dynamic LDS must opt in once
template <typename K>
static void ensure_lds_attr(K kern, size_t bytes) {
  if (bytes <= 65536) return;
  static std::mutex mu;
  static std::map<const void*, size_t> done;
  std::lock_guard<std::mutex> lk(mu);
  size_t& d = done[(const void*)kern];
  if (bytes <= d) return;
  (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  d = bytes;
}

template <typename T, int R, int TB, bool FULL>
static void launch_fused_fwd_tb(const Sched& S, const DeviceTables& D, int pi, const void* mtab, const void* ctab,
                                void* state, int64_t batch, int n, void* ckout, hipStream_t st) {
  const DPass& P = S.dpass[pi];
  const size_t lds = fused_lds_fwd<T>(P.k);
  ensure_lds_attr(k_fused_fwd<T, R, TB, FULL>, lds);
  const int64_t blocks = batch << P.nblk_log;
  hipLaunchKernelGGL((k_fused_fwd<T, R, TB, FULL>), dim3((unsigned)blocks), dim3(1u << P.tthreads_log), lds, st,
                     (const DPass*)D.dpass, pi, (const DGroup*)D.dgroup, (const DOp*)D.dop,
                     (const uint32_t*)D.runoff, (const T*)mtab, (int64_t)S.mtab_size, (const T*)ctab,
                     (V2<T>*)state, (int64_t)1 << n, (V2<T>*)ckout);
}

template <typename T, int R>
void launch_fused_fwd(const Sched& S, const DeviceTables& D, int pi, const void* mtab, const void* ctab,
                      void* state, int64_t batch, int n, void* ckout, hipStream_t st) {
  const int thr = 1 << S.dpass[pi].tthreads_log;
  const bool full = S.dpass[pi].full != 0;
#define TCM_FWD(TB, F) launch_fused_fwd_tb<T, R, TB, F>(S, D, pi, mtab, ctab, state, batch, n, ckout, st)
  if (thr <= 256) { if (full) TCM_FWD(256, true); else TCM_FWD(256, false); }
  else if (thr <= 512) { if (full) TCM_FWD(512, true); else TCM_FWD(512, false); }
  else { if (full) TCM_FWD(1024, true); else TCM_FWD(1024, false); }
#undef TCM_FWD
}

template <typename T>
void launch_fused_fwd_multi(const MultiJobs& J, int64_t blocks, int threads, size_t lds, hipStream_t st) {
  if (threads <= 256) {
    ensure_lds_attr(k_fused_fwd_multi<T, 4, 256>, lds);
    hipLaunchKernelGGL((k_fused_fwd_multi<T, 4, 256>), dim3((unsigned)blocks), dim3(threads), lds, st, J);
  } else {
    ensure_lds_attr(k_fused_fwd_multi<T, 4, 512>, lds);
    hipLaunchKernelGGL((k_fused_fwd_multi<T, 4, 512>), dim3((unsigned)blocks), dim3(threads), lds, st, J);
  }
}

template <typename T, int R, int TB, bool FULL>
static void launch_fused_adj_tb(const Sched& S, const DeviceTables& D, int pi, const void* mtab, const void* ctab,
                                const void* psi_in, void* psi_out, const void* lam_in, void* lam_out, int64_t batch,
                                int n, double* part, int64_t pstride, hipStream_t st) {
  const DPass& P = S.dpass[pi];
  const size_t lds = fused_lds_adj<T>(P.k, P.ngop, 1 << P.tthreads_log);
  ensure_lds_attr(k_fused_adj<T, R, TB, FULL>, lds);
  const int64_t blocks = batch << P.nblk_log;
  hipLaunchKernelGGL((k_fused_adj<T, R, TB, FULL>), dim3((unsigned)blocks), dim3(1u << P.tthreads_log), lds, st,
                     (const DPass*)D.dpass, pi, (const DGroup*)D.dgroup, (const DOp*)D.dop,
                     (const DMember*)D.dmem, (const uint32_t*)D.runoff, (const T*)mtab, (int64_t)S.mtab_size,
                     (const T*)ctab, (const V2<T>*)psi_in, (V2<T>*)psi_out, (const V2<T>*)lam_in, (V2<T>*)lam_out,
                     (int64_t)1 << n, part, pstride);
}

template <typename T, int R>
void launch_fused_adj(const Sched& S, const DeviceTables& D, int pi, const void* mtab, const void* ctab,
                      const void* psi_in, void* psi_out, const void* lam_in, void* lam_out, int64_t batch, int n,
                      double* part, int64_t pstride, hipStream_t st) {
  const int thr = 1 << S.dpass[pi].tthreads_log;
  const bool full = S.dpass[pi].full != 0;
#define TCM_ADJ(TB, F) \
  launch_fused_adj_tb<T, R, TB, F>(S, D, pi, mtab, ctab, psi_in, psi_out, lam_in, lam_out, batch, n, part, pstride, st)
  if (thr <= 256) { if (full) TCM_ADJ(256, true); else TCM_ADJ(256, false); }
  else if (thr <= 512) { if (full) TCM_ADJ(512, true); else TCM_ADJ(512, false); }
  else { if (full) TCM_ADJ(1024, true); else TCM_ADJ(1024, false); }
#undef TCM_ADJ
}

template <typename T>
void launch_plain_adj(const DOp& op, const void* mtab, int64_t mstride, const void* ctab, const void* psi_in,
                      void* psi_out, const void* lam_in, void* lam_out, int64_t batch, int n, int slots, double* part,
                      int64_t pstride, hipStream_t st) {
  dim3 grid((unsigned)slots, (unsigned)batch);
  const int nq = (op.kind == OK_U1 || op.kind == OK_R1 || op.kind == OK_D1 || op.kind == OK_X1) ? 1 : 2;
  if (nq == 1)
    hipLaunchKernelGGL((k_plain_adj<T, 1>), grid, dim3(256), 0, st, op, (const T*)mtab, mstride, (const T*)ctab,
                       (const V2<T>*)psi_in, (V2<T>*)psi_out, (const V2<T>*)lam_in, (V2<T>*)lam_out,
                       (int64_t)1 << n, n, part, pstride);
  else
    hipLaunchKernelGGL((k_plain_adj<T, 2>), grid, dim3(256), 0, st, op, (const T*)mtab, mstride, (const T*)ctab,
                       (const V2<T>*)psi_in, (V2<T>*)psi_out, (const V2<T>*)lam_in, (V2<T>*)lam_out,
                       (int64_t)1 << n, n, part, pstride);
}

template <typename T>
void launch_plain_adj_uk(const DOp& op, const int32_t* ukbits_dev, const void* ctab, void* psi, void* lam,
                         int64_t batch, int n, hipStream_t st) {
  launch_plain_gate<T, true>(op, ukbits_dev, nullptr, 0, ctab, psi, n, batch, st);
  launch_plain_gate<T, true>(op, ukbits_dev, nullptr, 0, ctab, lam, n, batch, st);
}

template <typename T>
void launch_pauli(const DPGroup& G, const DPTerm* terms_dev, int n, const void* psi, void* out, int first,
                  const double* alpha, double* epart, int64_t estride, int64_t batch, bool energy, bool apply,
                  hipStream_t st) {
  dim3 grid((unsigned)G.nblk, (unsigned)batch);
  const V2<T>* p = (const V2<T>*)psi;
  V2<T>* o = (V2<T>*)out;
  const int64_t ss = 1ll << n;
  if (energy && apply)
    hipLaunchKernelGGL((k_pauli<T, true, true>), grid, dim3(256), 0, st, G, terms_dev, n, p, ss, o, first, alpha, epart, estride);
  else if (energy)
    hipLaunchKernelGGL((k_pauli<T, true, false>), grid, dim3(256), 0, st, G, terms_dev, n, p, ss, o, first, alpha, epart, estride);
  else
    hipLaunchKernelGGL((k_pauli<T, false, true>), grid, dim3(256), 0, st, G, terms_dev, n, p, ss, o, first, alpha, epart, estride);
}

template <typename T>
void launch_pauli_tile(const DPPass& P, const DPPass* passes_dev, int pi, const DPGroupT* groups_dev,
                       const DPTermT* terms_dev, const uint32_t* runoff_dev, const void* psi, const void* lam_in,
                       void* lam_out, int n, const double* alpha, double* epart, int64_t estride, int64_t batch,
                       bool energy, bool apply, hipStream_t st) {
  const int64_t blocks = batch << P.nblk_log;
  const size_t lds = sizeof(V2<T>) << P.k;
  // the kernel strides over the tile with blockDim.x threads (at most 256)
  const dim3 grid((unsigned)blocks), block(1u << std::min(P.nthr_log, 8));
  const V2<T>* ps = (const V2<T>*)psi;
  const V2<T>* li = (const V2<T>*)lam_in;
  V2<T>* lo = (V2<T>*)lam_out;
  const int64_t ss = 1ll << n;
  if (energy && apply) {
    ensure_lds_attr(k_pauli_tile<T, true, true>, lds);
    hipLaunchKernelGGL((k_pauli_tile<T, true, true>), grid, block, lds, st, passes_dev, pi, groups_dev, terms_dev,
                       runoff_dev, ps, li, lo, ss, alpha, epart, estride);
  } else if (energy) {
    ensure_lds_attr(k_pauli_tile<T, true, false>, lds);
    hipLaunchKernelGGL((k_pauli_tile<T, true, false>), grid, block, lds, st, passes_dev, pi, groups_dev, terms_dev,
                       runoff_dev, ps, li, lo, ss, alpha, epart, estride);
  } else {
    ensure_lds_attr(k_pauli_tile<T, false, true>, lds);
    hipLaunchKernelGGL((k_pauli_tile<T, false, true>), grid, block, lds, st, passes_dev, pi, groups_dev, terms_dev,
                       runoff_dev, ps, li, lo, ss, alpha, epart, estride);
  }
}

template <typename T>
void launch_inner(const void* a, const void* c, int n, int64_t batch, int nblk, double* part, int64_t pstride,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_inner<T>, dim3((unsigned)nblk, (unsigned)batch), dim3(256), 0, st, (const V2<T>*)a,
                     (const V2<T>*)c, n, (int64_t)1 << n, part, pstride);
}

void launch_sum_slots(const double* part, int64_t bstride, int64_t nslots, int ncomp, double* out, int64_t batch,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_sum_slots, dim3((unsigned)batch), dim3(256), 0, st, part, bstride, nslots, ncomp, out);
}

void launch_reduce_gop(const double* part, int64_t pstride, int64_t nslots, const int32_t* gop_list, int gop_single,
                       int ngop_pass, double* gsum, int ngop, int64_t batch, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_gop, dim3((unsigned)ngop_pass, (unsigned)batch), dim3(256), 0, st, part, pstride,
                     nslots, gop_list, gop_single, gsum, ngop);
}

void launch_grad_assemble(const double* gsum, int ngop, const int32_t* p_ptr, const int32_t* p_gops,
                          const double* gfac, double mult, double* grad, int nparams, int64_t batch,
                          hipStream_t st) {
  const int64_t total = (int64_t)nparams * batch;
  if (total <= 0) return;
  hipLaunchKernelGGL(k_grad_assemble, dim3(grid_for(total, 1 << 24)), dim3(256), 0, st, gsum, ngop, p_ptr, p_gops,
                     gfac, mult, grad, nparams, batch);
}

// ---------------------------------------------------------------------------
// Measurement: probability vector (basecircuit.py:510-524), draws from
// |psi|^2 by inverse CDF (backend.probability_sample,
// abstract_backend.py:1124-1157) and qubit-by-qubit conditional measurement
// (measure_jit, basecircuit.py:371-445).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_probability(const V2<T>* __restrict__ psi, T* __restrict__ out, int64_t total) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const V2<T> x = psi[i];
    out[i] = x.x * x.x + x.y * x.y;
  }
}

// weight of index i for the sampler: |psi_i|^2 of a state (PROB false) or
// p_i of a probability vector (PROB true)
template <typename T, bool PROB>
__device__ __forceinline__ double sample_weight(const void* src, int64_t i) {
  if constexpr (PROB) {
    return (double)static_cast<const T*>(src)[i];
  } else {
    const V2<T> x = static_cast<const V2<T>*>(src)[i];
    return (double)x.x * x.x + (double)x.y * x.y;
  }
}

// cum[b][c + 1] = fp64 sum of the weights over chunk c (2^chunk_log
// indices); grid (nchunks, batch)
template <typename T, bool PROB>
__global__ void __launch_bounds__(256) k_prob_chunks(const void* __restrict__ src, int64_t sstride, int chunk_log,
                                                     int64_t nchunks, double* __restrict__ cum) {
  const int64_t c = blockIdx.x, b = blockIdx.y;
  const int64_t off = b * sstride + (c << chunk_log);
  const int64_t len = 1ll << chunk_log;
  double a = 0.0, d = 0.0;
  for (int64_t i = threadIdx.x; i < len; i += blockDim.x) a += sample_weight<T, PROB>(src, off + i);
  block_sum2(a, d);
  if (threadIdx.x == 0) cum[b * (nchunks + 1) + c + 1] = a;
}

// in-place inclusive prefix sums of cum[b][1..nchunks], cum[b][0] = 0; one
// 1024-thread block per batch item, each thread owning a contiguous range
__global__ void __launch_bounds__(1024) k_scan_chunks(double* __restrict__ cum, int64_t nchunks) {
  __shared__ double sh[1024];
  double* c = cum + blockIdx.x * (nchunks + 1);
  const int t = threadIdx.x, nt = blockDim.x;
  const int64_t per = (nchunks + nt - 1) / nt;
  const int64_t lo = min(nchunks, (int64_t)t * per), hi = min(nchunks, lo + per);
  double s = 0.0;
  for (int64_t i = lo; i < hi; ++i) s += c[1 + i];
  sh[t] = s;
  __syncthreads();
  for (int off = 1; off < nt; off <<= 1) {
    const double v = t >= off ? sh[t - off] : 0.0;
    __syncthreads();
    sh[t] += v;
    __syncthreads();
  }
  double run = sh[t] - s;
  for (int64_t i = lo; i < hi; ++i) {
    run += c[1 + i];
    c[1 + i] = run;
  }
  if (t == 0) c[0] = 0.0;
}

// one wave per draw: r = total * (1 - u); the drawn index is the smallest i
// whose inclusive prefix sum of |psi|^2 reaches r (searchsorted, side left):
// binary search over the chunk prefixes, then a 64-wide running scan of the
// chunk.  grid (ceil(shots / 4), batch), 256 threads
template <typename T, bool PROB>
__global__ void __launch_bounds__(256) k_sample_draw(const void* __restrict__ src, int64_t sstride, int chunk_log,
                                                     int64_t nchunks, const double* __restrict__ cum,
                                                     const double* __restrict__ u, int64_t shots,
                                                     int64_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t s = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t b = blockIdx.y;
  if (s >= shots) return;  // uniform per wave
  const double* c = cum + b * (nchunks + 1);
  const double r = c[nchunks] * (1.0 - u[b * shots + s]);
  int64_t lo = 0, hi = nchunks - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (c[mid + 1] >= r) hi = mid;
    else lo = mid + 1;
  }
  const int64_t len = 1ll << chunk_log;
  const int64_t first = lo << chunk_log;
  const int64_t base = b * sstride + first;
  double run = c[lo];
  int64_t found = -1, lastnz = -1;
  for (int64_t j = 0; j < len; j += 64) {
    const double p = (j + lane < len) ? sample_weight<T, PROB>(src, base + j + lane) : 0.0;
    double incl = p;
    for (int off = 1; off < 64; off <<= 1) {
      const double v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    const unsigned long long valid = (len - j >= 64) ? ~0ull : ((1ull << (len - j)) - 1);
    const unsigned long long hit = __ballot(run + incl >= r) & valid;
    const unsigned long long nz = __ballot(p > 0.0) & valid;
    if (nz) lastnz = j + 63 - __clzll((long long)nz);
    if (hit) {
      found = j + __ffsll((long long)hit) - 1;
      break;
    }
    run += __shfl(incl, 63, 64);
  }
  // rounding left r just above the chunk's running total: the last index
  // of the chunk with nonzero probability
  if (found < 0) found = lastnz >= 0 ? lastnz : 0;
  if (lane == 0) out[b * shots + s] = first + found;
}

// partial sums of |psi_i|^2 over the indices that agree with the outcomes
// fixed so far (fixval[b] on the fixed bits), split by index bit M.bit; the
// free bits are enumerated by depositing a counter into M.run.
// grid (nblk, batch); part[b][blk][2]
template <typename T>
__global__ void __launch_bounds__(256) k_measure_partial(const V2<T>* __restrict__ psi, int64_t sstride,
                                                         MeasureStep M, const uint64_t* __restrict__ fixval,
                                                         double* __restrict__ part) {
  const int64_t b = blockIdx.y;
  const V2<T>* s = psi + b * sstride;
  const uint64_t fv = M.first ? 0ull : fixval[b];
  const int64_t total = 1ll << M.nfree;
  double a0 = 0.0, a1 = 0.0;
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < total; m += (int64_t)gridDim.x * blockDim.x) {
    uint64_t i = fv;
    for (int r = 0; r < M.nrun; ++r) {
      const int src = M.run[r] & 0xff, len = (M.run[r] >> 8) & 0xff, dst = M.run[r] >> 16;
      i |= (((uint64_t)m >> src) & ((1ull << len) - 1)) << dst;
    }
    const V2<T> x = s[i];
    const double p = (double)x.x * x.x + (double)x.y * x.y;
    if ((i >> M.bit) & 1) a1 += p;
    else a0 += p;
  }
  block_sum2(a0, a1);
  if (threadIdx.x == 0) {
    part[(b * gridDim.x + blockIdx.x) * 2] = a0;
    part[(b * gridDim.x + blockIdx.x) * 2 + 1] = a1;
  }
}

// measure_jit's decision for step k (basecircuit.py:419-431): pu = p0 / p,
// outcome 1 iff status - pu + eps > 0, p *= pu or (1 - pu); one block per
// batch item
__global__ void __launch_bounds__(256) k_measure_decide(const double* __restrict__ part, int nblk, int k, int nmeas,
                                                        int bit, const double* __restrict__ status,
                                                        uint64_t* __restrict__ fixval, double* __restrict__ prob,
                                                        double* __restrict__ outcome) {
  const int64_t b = blockIdx.x;
  double a0 = 0.0, a1 = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) a0 += part[(b * nblk + i) * 2];
  block_sum2(a0, a1);
  if (threadIdx.x == 0) {
    const double p = k == 0 ? 1.0 : prob[b];
    const double pu = a0 / p;
    const double r = status[b * nmeas + k];
    const double eps = 0.31415926 * 1e-12;
    const int o = (r - pu + eps > 0.0) ? 1 : 0;
    prob[b] = p * (o ? 1.0 - pu : pu);
    fixval[b] = (k == 0 ? 0ull : fixval[b]) | ((uint64_t)o << bit);
    outcome[b * nmeas + k] = (double)o;
  }
}

template <typename T>
void launch_probability(const void* psi, void* out, int64_t total, hipStream_t st) {
  hipLaunchKernelGGL(k_probability<T>, dim3(grid_for(total)), dim3(256), 0, st, (const V2<T>*)psi, (T*)out, total);
}

template <typename T>
void launch_sample(const void* src, bool prob, int n, int64_t batch, const double* u, int64_t shots, int64_t* out,
                   double* cum, hipStream_t st) {
  const int chunk_log = n < 10 ? n : 10;
  const int64_t nchunks = 1ll << (n - chunk_log);
  const int64_t sstride = 1ll << n;
  const dim3 gc((unsigned)nchunks, (unsigned)batch), gd((unsigned)((shots + 3) / 4), (unsigned)batch);
  if (prob)
    hipLaunchKernelGGL((k_prob_chunks<T, true>), gc, dim3(256), 0, st, src, sstride, chunk_log, nchunks, cum);
  else
    hipLaunchKernelGGL((k_prob_chunks<T, false>), gc, dim3(256), 0, st, src, sstride, chunk_log, nchunks, cum);
  hipLaunchKernelGGL(k_scan_chunks, dim3((unsigned)batch), dim3(1024), 0, st, cum, nchunks);
  if (prob)
    hipLaunchKernelGGL((k_sample_draw<T, true>), gd, dim3(256), 0, st, src, sstride, chunk_log, nchunks,
                       (const double*)cum, u, shots, out);
  else
    hipLaunchKernelGGL((k_sample_draw<T, false>), gd, dim3(256), 0, st, src, sstride, chunk_log, nchunks,
                       (const double*)cum, u, shots, out);
}

template <typename T>
void launch_measure_step(const void* psi, int n, int64_t batch, const MeasureStep& M, int k, int nmeas,
                         const double* status, uint64_t* fixval, double* prob, double* outcome, double* part,
                         int nblk, hipStream_t st) {
  hipLaunchKernelGGL(k_measure_partial<T>, dim3((unsigned)nblk, (unsigned)batch), dim3(256), 0, st,
                     (const V2<T>*)psi, (int64_t)1 << n, M, (const uint64_t*)fixval, part);
  hipLaunchKernelGGL(k_measure_decide, dim3((unsigned)batch), dim3(256), 0, st, (const double*)part, nblk, k, nmeas,
                     M.bit, status, fixval, prob, outcome);
}

// ---------------------------------------------------------------------------
// Sparse form of a Pauli sum (PauliStringSum2COO / ps2coo_core,
// quantum.py:1304-1485) and <psi|H|psi> over a CSR matrix
// (sparse_expectation, templates/measurements.py:177-191).
// ---------------------------------------------------------------------------
// entry t * N + y of term t (group g): row y, column y ^ flip,
// value w (-i)^ny (-1)^popc(y & yz); grid-stride over nterms * N
template <typename T>
__global__ void k_pauli_coo(const DPGroup* __restrict__ groups, int ngroups, const DPTerm* __restrict__ terms,
                            int n, int64_t* __restrict__ rows, int64_t* __restrict__ cols, V2<T>* __restrict__ vals,
                            int64_t total) {
  const int64_t N = 1ll << n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int t = (int)(e >> n);
    const uint64_t y = (uint64_t)(e & (N - 1));
    int g = 0;
    while (g + 1 < ngroups && groups[g + 1].t0 <= t) ++g;
    const DPTerm tm = terms[t];
    const double sg = (__popcll(y & tm.yz) & 1) ? -tm.w : tm.w;
    double re = 0.0, im = 0.0;
    switch (tm.ny & 3) {  // (-i)^ny
      case 0: re = sg; break;
      case 1: im = -sg; break;
      case 2: re = -sg; break;
      default: im = sg; break;
    }
    rows[e] = (int64_t)y;
    cols[e] = (int64_t)(y ^ groups[g].flip);
    vals[e] = mk<T>((T)re, (T)im);
  }
}

// part[blk] = sum over rows r of this block's grid-stride share of
// conj(psi_r) * sum_k vals_k psi_{cols_k} (fp64); one thread per row
template <typename T>
__global__ void __launch_bounds__(256) k_csr_expect(const int64_t* __restrict__ rowptr, const int64_t* __restrict__ cols,
                                                    const V2<T>* __restrict__ vals, const V2<T>* __restrict__ psi,
                                                    int64_t nrows, double* __restrict__ part) {
  double er = 0.0, ei = 0.0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * blockDim.x) {
    double hr = 0.0, hi = 0.0;
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      const V2<T> v = vals[k], x = psi[cols[k]];
      hr += (double)v.x * x.x - (double)v.y * x.y;
      hi += (double)v.x * x.y + (double)v.y * x.x;
    }
    const V2<T> p = psi[r];
    er += (double)p.x * hr + (double)p.y * hi;
    ei += (double)p.x * hi - (double)p.y * hr;
  }
  block_sum2(er, ei);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = er;
    part[2 * blockIdx.x + 1] = ei;
  }
}

template <typename T>
void launch_pauli_coo(const DPGroup* groups, int ngroups, const DPTerm* terms, int n, int64_t nterms, int64_t* rows,
                      int64_t* cols, void* vals, hipStream_t st) {
  const int64_t total = nterms << n;
  hipLaunchKernelGGL(k_pauli_coo<T>, dim3(grid_for(total)), dim3(256), 0, st, groups, ngroups, terms, n, rows, cols,
                     (V2<T>*)vals, total);
}

template <typename T>
void launch_csr_expect(const int64_t* rowptr, const int64_t* cols, const void* vals, const void* psi, int64_t nrows,
                       int nblk, double* part, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_csr_expect<T>, dim3((unsigned)nblk), dim3(256), 0, st, rowptr, cols, (const V2<T>*)vals,
                     (const V2<T>*)psi, nrows, part);
  hipLaunchKernelGGL(k_sum_slots, dim3(1), dim3(256), 0, st, (const double*)part, (int64_t)2 * nblk, (int64_t)nblk, 2,
                     out);
}

// explicit instantiations used by engine.cpp
#define TCM_INST(T)                                                                                            \
  template void launch_prep<T>(const Sched&, const DeviceTables&, const void*, const void*, int64_t, void*,   \
                               int64_t, hipStream_t);                                                          \
  template void launch_init_zero<T>(void*, int, int64_t, hipStream_t);                                         \
  template void launch_copy<T>(void*, const void*, int64_t, hipStream_t);                                      \
  template void launch_plain_gate<T, false>(const DOp&, const int32_t*, const void*, int64_t, const void*,    \
                                            void*, int, int64_t, hipStream_t);                                 \
  template void launch_plain_gate<T, true>(const DOp&, const int32_t*, const void*, int64_t, const void*,     \
                                           void*, int, int64_t, hipStream_t);                                  \
  template void launch_fused_fwd<T, 4>(const Sched&, const DeviceTables&, int, const void*, const void*,      \
                                       void*, int64_t, int, void*, hipStream_t);                               \
  template void launch_fused_adj<T, 4>(const Sched&, const DeviceTables&, int, const void*, const void*,      \
                                       const void*, void*, const void*, void*, int64_t, int, double*, int64_t, \
                                       hipStream_t);                                                           \
  template void launch_plain_adj<T>(const DOp&, const void*, int64_t, const void*, const void*, void*,        \
                                    const void*, void*, int64_t, int, int, double*, int64_t, hipStream_t);     \
  template void launch_plain_adj_uk<T>(const DOp&, const int32_t*, const void*, void*, void*, int64_t, int,   \
                                       hipStream_t);                                                           \
  template void launch_pauli<T>(const DPGroup&, const DPTerm*, int, const void*, void*, int, const double*,   \
                                double*, int64_t, int64_t, bool, bool, hipStream_t);                           \
  template void launch_inner<T>(const void*, const void*, int, int64_t, int, double*, int64_t, hipStream_t); \
  template void launch_fused_fwd_multi<T>(const MultiJobs&, int64_t, int, size_t, hipStream_t);                 \
  template void launch_pauli_tile<T>(const DPPass&, const DPPass*, int, const DPGroupT*, const DPTermT*,        \
                                     const uint32_t*, const void*, const void*, void*, int, const double*,       \
                                     double*, int64_t, int64_t, bool, bool, hipStream_t);                      \
  template void launch_probability<T>(const void*, void*, int64_t, hipStream_t);                               \
  template void launch_sample<T>(const void*, bool, int, int64_t, const double*, int64_t, int64_t*, double*,   \
                                 hipStream_t);                                                                 \
  template void launch_measure_step<T>(const void*, int, int64_t, const MeasureStep&, int, int, const double*,  \
                                       uint64_t*, double*, double*, double*, int, hipStream_t);                \
  template void launch_pauli_coo<T>(const DPGroup*, int, const DPTerm*, int, int64_t, int64_t*, int64_t*, void*,  \
                                    hipStream_t);                                                              \
  template void launch_csr_expect<T>(const int64_t*, const int64_t*, const void*, const void*, int64_t, int,      \
                                     double*, double*, hipStream_t);

TCM_INST(float)
TCM_INST(double)

}  // namespace tcm
