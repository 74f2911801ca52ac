// HIP kernels of libtcm for CDNA4 (gfx950), measurement helpers:
//  - per-term Pauli expectations  E[b][t] = w_t <psi_b|P_t|psi_b>  (fp64)
//    behind any_measurements / parameterized_local_measurements over many
//    structures at once (templates/measurements.py:18-153);
//  - weighted Pauli applications  out_b = sum_t a[b][t] w_t P_t psi_b, the
//    vector-Jacobian product of those expectations;
//  - CSR mat-vec  y_b = A x_b, the vector-Jacobian product of
//    sparse_expectation (templates/measurements.py:177-191).
//
// Pauli action (index rule of ps2coo_core, quantum.py:1461-1485):
//   (P psi)[a] = (-i)^ny (-1)^popcount(a & yz) psi[a ^ flip].
// State layout: [batch][2^n] complex, qubit 0 = most significant index bit.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace tcm {

namespace {

template <typename T>
__device__ __forceinline__ V2<T> make2(T a, T b) {
  V2<T> r;
  r.x = a;
  r.y = b;
  return r;
}

// fp64 sum of (a, b) over a block of 256 threads; result valid in thread 0
__device__ __forceinline__ void block_sum256(double& a, double& b) {
  __shared__ double red[2][4];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][w] = a;
    red[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    b = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// (re, im) * (-i)^ny
__device__ __forceinline__ void times_mi_pow(int ny, double& re, double& im) {
  const double r = re, i = im;
  switch (ny & 3) {
    case 0: break;
    case 1: re = i; im = -r; break;
    case 2: re = -r; im = -i; break;
    default: re = -i; im = r; break;
  }
}

}  // namespace

// Partial sums of E[b][t]: grid (nblk, nterms, batch) of 256 threads; block
// x of term t sums conj(psi[a]) (-1)^popc(a & yz) psi[a ^ flip] over its
// grid-stride share of a (both reads coalesced: the XOR with a fixed mask
// permutes within aligned runs), then scales by w (-i)^ny.
template <typename T>
__global__ void __launch_bounds__(256) k_pauli_terms(const DPTermF* __restrict__ terms, int n,
                                                     const V2<T>* __restrict__ psi, double* __restrict__ part,
                                                     int nblk) {
  const int t = blockIdx.y, nterms = gridDim.y;
  const int64_t b = blockIdx.z;
  const DPTermF tm = terms[t];
  const int64_t N = 1ll << n;
  const V2<T>* s = psi + b * N;
  double er = 0.0, ei = 0.0;
  for (int64_t a = blockIdx.x * 256ll + threadIdx.x; a < N; a += (int64_t)nblk * 256) {
    const V2<T> p = s[a], q = s[(uint64_t)a ^ tm.flip];
    double cr = (double)p.x * q.x + (double)p.y * q.y;
    double ci = (double)p.x * q.y - (double)p.y * q.x;
    if (__popcll((uint64_t)a & tm.yz) & 1) {
      cr = -cr;
      ci = -ci;
    }
    er += cr;
    ei += ci;
  }
  block_sum256(er, ei);
  if (threadIdx.x == 0) {
    times_mi_pow(tm.ny, er, ei);
    double* o = part + (((int64_t)b * nterms + t) * nblk + blockIdx.x) * 2;
    o[0] = tm.w * er;
    o[1] = tm.w * ei;
  }
}

// out[b][t] = sum of the nblk partials of (b, t); one wave per (b, t)
__global__ void __launch_bounds__(64) k_pauli_terms_sum(const double* __restrict__ part, int nblk,
                                                        double* __restrict__ out) {
  const int64_t bt = blockIdx.x;
  double a = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 64) {
    a += part[(bt * nblk + i) * 2];
    c += part[(bt * nblk + i) * 2 + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    c += __shfl_xor(c, off, 64);
  }
  if (threadIdx.x == 0) {
    out[2 * bt] = a;
    out[2 * bt + 1] = c;
  }
}

// out_b[a] = sum_t wts[b][t] w_t (-i)^ny_t (-1)^popc(a & yz_t) psi_b[a ^ flip_t]
// in fp64, stored in T; grid-stride over batch * 2^n amplitudes, the term
// loop uniform across the wave
template <typename T>
__global__ void __launch_bounds__(256) k_pauli_terms_apply(const DPTermF* __restrict__ terms, int nterms, int n,
                                                           const V2<T>* __restrict__ psi, V2<T>* __restrict__ out,
                                                           const double* __restrict__ wts, int64_t total) {
  const int64_t N = 1ll << n;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256) {
    const int64_t b = i >> n;
    const uint64_t a = (uint64_t)(i & (N - 1));
    const V2<T>* s = psi + b * N;
    const double* wb = wts + b * nterms;
    double hr = 0.0, hi = 0.0;
    for (int t = 0; t < nterms; ++t) {
      const DPTermF tm = terms[t];
      const double c = (__popcll(a & tm.yz) & 1) ? -tm.w * wb[t] : tm.w * wb[t];
      if (c == 0.0) continue;
      const V2<T> q = s[a ^ tm.flip];
      double re = c * (double)q.x, im = c * (double)q.y;
      times_mi_pow(tm.ny, re, im);
      hr += re;
      hi += im;
    }
    out[i] = make2<T>((T)hr, (T)hi);
  }
}

// y_b[r] = sum_k vals[k] x_b[cols[k]] over row r of a CSR matrix (fp64
// accumulation); one thread per (b, r)
template <typename T>
__global__ void __launch_bounds__(256) k_csr_matvec(const int64_t* __restrict__ rowptr,
                                                    const int64_t* __restrict__ cols, const V2<T>* __restrict__ vals,
                                                    const V2<T>* __restrict__ x, V2<T>* __restrict__ y,
                                                    int64_t nrows, int64_t total) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256) {
    const int64_t b = i / nrows, r = i - b * nrows;
    const V2<T>* xb = x + b * nrows;
    double hr = 0.0, hi = 0.0;
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      const V2<T> v = vals[k], q = xb[cols[k]];
      hr += (double)v.x * q.x - (double)v.y * q.y;
      hi += (double)v.x * q.y + (double)v.y * q.x;
    }
    y[i] = make2<T>((T)hr, (T)hi);
  }
}

namespace {
unsigned grid_of(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  return (unsigned)(g < 1 ? 1 : g);
}
}  // namespace

template <typename T>
void launch_pauli_terms(const DPTermF* terms_dev, int nterms, int n, const void* psi, int64_t batch, int nblk,
                        double* part, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_pauli_terms<T>, dim3((unsigned)nblk, (unsigned)nterms, (unsigned)batch), dim3(256), 0, st,
                     terms_dev, n, (const V2<T>*)psi, part, nblk);
  hipLaunchKernelGGL(k_pauli_terms_sum, dim3((unsigned)(batch * nterms)), dim3(64), 0, st, (const double*)part, nblk,
                     out);
}

template <typename T>
void launch_pauli_terms_apply(const DPTermF* terms_dev, int nterms, int n, const void* psi, void* out,
                              const double* weights, int64_t batch, hipStream_t st) {
  const int64_t total = batch << n;
  hipLaunchKernelGGL(k_pauli_terms_apply<T>, dim3(grid_of(total)), dim3(256), 0, st, terms_dev, nterms, n,
                     (const V2<T>*)psi, (V2<T>*)out, weights, total);
}

template <typename T>
void launch_csr_matvec(const int64_t* rowptr, const int64_t* cols, const void* vals, const void* x, void* y,
                       int64_t nrows, int64_t batch, hipStream_t st) {
  const int64_t total = batch * nrows;
  hipLaunchKernelGGL(k_csr_matvec<T>, dim3(grid_of(total)), dim3(256), 0, st, rowptr, cols, (const V2<T>*)vals,
                     (const V2<T>*)x, (V2<T>*)y, nrows, total);
}

#define TCM_AUX_INST(T)                                                                                       \
  template void launch_pauli_terms<T>(const DPTermF*, int, int, const void*, int64_t, int, double*, double*,  \
                                      hipStream_t);                                                           \
  template void launch_pauli_terms_apply<T>(const DPTermF*, int, int, const void*, void*, const double*,      \
                                            int64_t, hipStream_t);                                            \
  template void launch_csr_matvec<T>(const int64_t*, const int64_t*, const void*, const void*, void*, int64_t, \
                                     int64_t, hipStream_t);

TCM_AUX_INST(float)
TCM_AUX_INST(double)

}  // namespace tcm
