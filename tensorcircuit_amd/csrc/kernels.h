// Declarations shared by kernels.hip (device code + launchers) and
// engine.cpp (the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "tcm_internal.h"

namespace tcm {

template <typename T>
struct V2T;
template <>
struct V2T<float> {
  using type = float2;
};
template <>
struct V2T<double> {
  using type = double2;
};
template <typename T>
using V2 = typename V2T<T>::type;

// Pauli-sum device descriptors
struct DPGroup {
  uint64_t flip;   // X/Y positions (index bits)
  int32_t t0, nt;  // terms
  int32_t slot0;   // first energy-partial slot of this group
  int32_t nblk;    // blocks launched (= slots used)
};
struct DPTerm {
  uint64_t yz;     // Y/Z positions (index bits)
  int32_t ny;      // number of Y factors
  int32_t pad;
  double w;        // weight
};

// LDS-tiled Pauli passes: every term whose X/Y positions lie in the tile is
// evaluated from one read of psi
struct DPPass {
  int32_t k, L, nblk_log, runoff0;
  int32_t g0, ng;      // groups of this pass
  int32_t nthr_log;    // threads per block (log2)
  int32_t slot0;       // first energy-partial slot
  int32_t blk_bits[32];
};
// one Pauli term in the caller's order (per-term expectations / weighted
// applications): P = w (-i)^ny X^flip with the sign (-1)^popcount(row & yz)
struct DPTermF {
  uint64_t flip;   // X/Y positions (index bits)
  uint64_t yz;     // Y/Z positions (index bits)
  int32_t ny;      // number of Y factors
  int32_t pad;
  double w;        // weight
};

struct DPGroupT {
  uint32_t flip_t;     // tile-local X/Y mask
  int32_t t0, nt, pad;
};
struct DPTermT {
  uint64_t yz_g;       // Y/Z positions outside the tile (global bits)
  uint32_t yz_t;       // Y/Z positions inside the tile (tile-local bits)
  int32_t ny;
  double w;
};

// one program's pass inside a multi-program launch (passed by value)
struct MultiJob {
  const void *dpass, *dgroup, *dop, *runoff, *ctab, *mtab;
  void* state;
  int64_t mstride, sstride;
  int64_t block0;   // first block of this job in the launch
  int32_t pass, pad;
};
constexpr int kMaxMultiJobs = 24;
struct MultiJobs {
  int32_t njobs, pad;
  MultiJob job[kMaxMultiJobs];
};

// device copies of one schedule's descriptors (owned by the program)
struct DeviceTables {
  void* dpass = nullptr;
  void* dgroup = nullptr;
  void* dop = nullptr;
  void* dmem = nullptr;
  void* ukbits = nullptr;
  void* runoff = nullptr;
  void* recipes = nullptr;
  void* factors = nullptr;
  void* gop_list = nullptr;
  // program-wide (shared by both schedules of a dtype, owned by mode 0)
  void* ctab = nullptr;    // constant matrices in T
  void* p_ptr = nullptr;
  void* p_gops = nullptr;
  void* gfac = nullptr;
  // pass compiler: loaded modules and kernel handles (hipModule_t / hipFunction_t)
  std::vector<void*> jit_mod;
  std::vector<void*> jit_fn;
};

// launch compiled pass kernels (jit.cpp)
int jit_load(const Sched& S, DeviceTables& D);
// zinit: the pass starts from |0..0> instead of reading `state`
int jit_launch_fwd(const Sched& S, const DeviceTables& D, int pi, int dt, const void* mtab, const void* ctab,
                   void* state, int64_t batch, int n, void* ckout, hipStream_t st, int zinit = 0);
int jit_launch_adj(const Sched& S, const DeviceTables& D, int pi, int dt, const void* mtab, const void* ctab,
                   const void* psi_in, void* psi_out, const void* lam_in, void* lam_out, int64_t batch, int n,
                   double* part, int64_t pstride, hipStream_t st);

template <typename T>
void launch_prep(const Sched& S, const DeviceTables& D, const void* ctab, const void* params, int64_t pstride,
                 void* mtab, int64_t batch, hipStream_t st);
template <typename T>
void launch_init_zero(void* state, int n, int64_t batch, hipStream_t st);
template <typename T>
void launch_copy(void* dst, const void* src, int64_t total, hipStream_t st);
template <typename T, bool DAG>
void launch_plain_gate(const DOp& op, const int32_t* ukbits_dev, const void* mtab, int64_t mstride,
                       const void* ctab, void* state, int n, int64_t batch, hipStream_t st);
template <typename T, int R>
void launch_fused_fwd(const Sched& S, const DeviceTables& D, int pi, const void* mtab, const void* ctab,
                      void* state, int64_t batch, int n, void* ckout, hipStream_t st);
template <typename T, int R>
void launch_fused_adj(const Sched& S, const DeviceTables& D, int pi, const void* mtab, const void* ctab,
                      const void* psi_in, void* psi_out, const void* lam_in, void* lam_out, int64_t batch, int n,
                      double* part, int64_t pstride, hipStream_t st);
template <typename T>
void launch_plain_adj(const DOp& op, const void* mtab, int64_t mstride, const void* ctab, const void* psi_in,
                      void* psi_out, const void* lam_in, void* lam_out, int64_t batch, int n, int slots, double* part,
                      int64_t pstride, hipStream_t st);
template <typename T>
void launch_plain_adj_uk(const DOp& op, const int32_t* ukbits_dev, const void* ctab, void* psi, void* lam,
                         int64_t batch, int n, hipStream_t st);
template <typename T>
void launch_pauli(const DPGroup& G, const DPTerm* terms_dev, int n, const void* psi, void* out, int first,
                  const double* alpha, double* epart, int64_t estride, int64_t batch, bool energy, bool apply,
                  hipStream_t st);
template <typename T>
void launch_pauli_tile(const DPPass& P, const DPPass* passes_dev, int pi, const DPGroupT* groups_dev,
                       const DPTermT* terms_dev, const uint32_t* runoff_dev, const void* psi, const void* lam_in,
                       void* lam_out, int n, const double* alpha, double* epart, int64_t estride, int64_t batch,
                       bool energy, bool apply, hipStream_t st);
template <typename T>
void launch_fused_fwd_multi(const MultiJobs& J, int64_t blocks, int threads, size_t lds, hipStream_t st);
template <typename T>
void launch_inner(const void* a, const void* c, int n, int64_t batch, int nblk, double* part, int64_t pstride,
                  hipStream_t st);
void launch_sum_slots(const double* part, int64_t bstride, int64_t nslots, int ncomp, double* out, int64_t batch,
                      hipStream_t st);
// measurement
struct MeasureStep {
  int32_t nfree;     // free (not yet measured) index bits, enumerated
  int32_t nrun;      // runs of consecutive free bits
  int32_t bit;       // index bit of the qubit measured in this step
  int32_t first;     // 1 on the first step: no bits fixed yet
  int32_t run[34];   // counter bits src .. src+len-1 -> index bits dst ..
};
// a tile pass of dense 5-qubit blocks on FP32 matrix cores (complex64);
// false if the pass does not fit the kernel (tile bits 10..13, at most
// kMfmaTileMaxOps blocks)
bool launch_mfma_tile(const DPass& P, const DPass* passes_dev, int pi, const DOp* dop_dev, const int32_t* ukbits_dev,
                      const uint32_t* runoff_dev, const void* ctab, void* state, int n, int64_t batch,
                      hipStream_t st);
template <typename T>
void launch_probability(const void* psi, void* out, int64_t total, hipStream_t st);
template <typename T>
void launch_pauli_coo(const DPGroup* groups, int ngroups, const DPTerm* terms, int n, int64_t nterms, int64_t* rows,
                      int64_t* cols, void* vals, hipStream_t st);
// per-term Pauli expectations and weighted applications, CSR mat-vec
// (kernels_aux.hip)
template <typename T>
void launch_pauli_terms(const DPTermF* terms_dev, int nterms, int n, const void* psi, int64_t batch, int nblk,
                        double* part, double* out, hipStream_t st);
template <typename T>
void launch_pauli_terms_apply(const DPTermF* terms_dev, int nterms, int n, const void* psi, void* out,
                              const double* weights, int64_t batch, hipStream_t st);
template <typename T>
void launch_csr_matvec(const int64_t* rowptr, const int64_t* cols, const void* vals, const void* x, void* y,
                       int64_t nrows, int64_t batch, hipStream_t st);
template <typename T>
void launch_csr_expect(const int64_t* rowptr, const int64_t* cols, const void* vals, const void* psi, int64_t nrows,
                       int nblk, double* part, double* out, hipStream_t st);
// src: states [batch][2^n] (prob false) or probability vectors (prob true)
template <typename T>
void launch_sample(const void* src, bool prob, int n, int64_t batch, const double* u, int64_t shots, int64_t* out,
                   double* cum, hipStream_t st);
template <typename T>
void launch_measure_step(const void* psi, int n, int64_t batch, const MeasureStep& M, int k, int nmeas,
                         const double* status, uint64_t* fixval, double* prob, double* outcome, double* part,
                         int nblk, hipStream_t st);
void launch_reduce_gop(const double* part, int64_t pstride, int64_t nslots, const int32_t* gop_list, int gop_single,
                       int ngop_pass, double* gsum, int ngop, int64_t batch, hipStream_t st);
void launch_grad_assemble(const double* gsum, int ngop, const int32_t* p_ptr, const int32_t* p_gops,
                          const double* gfac, double mult, double* grad, int nparams, int64_t batch,
                          hipStream_t st);

}  // namespace tcm
