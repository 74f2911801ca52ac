/*
 * tcm.h -- C ABI of libtcm, the MI355X (gfx950) dense statevector engine
 * behind tensorcircuit_amd's Circuit API.
 *
 * The reference (tensorcircuit 0.12.1) has no native code: its statevector
 * arithmetic is the tensordot loop of the contractor (cons.py:523-631,
 * plain_contractor cons.py:334-370) reached from Circuit.wavefunction
 * (circuit.py:792-810) and Circuit.expectation (circuit.py:914-990), with
 * gradients from the torch backend's autograd (backends/pytorch_backend.py:
 * 497-709).  Every entry point below replaces one of those Python-level
 * operations; the replaced reference interface is cited on each.
 *
 * Conventions
 *  - States are caller-owned device buffers of interleaved (re, im) complex
 *    numbers, fp32 pairs for TCM_C64 and fp64 pairs for TCM_C128, laid out
 *    [batch][2^n], qubit 0 = most significant bit of the amplitude index
 *    (circuit.py:803-810).
 *  - Parameters are device buffers [batch][nparams] of float (TCM_C64) or
 *    double (TCM_C128).
 *  - Energies, expectations and gradients are returned in fp64.
 *  - All calls enqueue on `stream` (a hipStream_t, NULL = legacy default) and
 *    never synchronise the device; results are valid once the stream is.
 *  - Return value: 0 ok; < 0 argument error (TCM_E_*); > 0 a hipError_t.
 *    tcm_last_error() gives a thread-local message for the last failure.
 *  - The library allocates device memory only for program / Pauli-sum
 *    descriptor tables (uploaded on first use per device, freed by
 *    tcm_program_free / tcm_pauli_free) and never frees caller memory.
 *    Scratch comes from the caller's workspace (tcm_*_workspace_bytes).
 */
#ifndef TCM_H_
#define TCM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCM_ABI_VERSION 1

enum tcm_dtype { TCM_C64 = 0, TCM_C128 = 1 };

enum tcm_status {
  TCM_OK = 0,
  TCM_E_ARG = -1,       /* bad argument (null pointer, size, dtype) */
  TCM_E_QUBIT = -2,     /* qubit index out of range */
  TCM_E_DUPQUBIT = -3,  /* repeated qubit in one gate / term */
  TCM_E_WORKSPACE = -4, /* workspace too small */
  TCM_E_UNSUPPORTED = -5,
  TCM_E_NODEVICE = -6,  /* no HIP device */
  TCM_E_NONFINITE = -7  /* NaN / Inf in a host-side matrix or coefficient */
};

/* Gate kinds.  A "fixed" gate has param < 0 and uses `theta`; a trainable
 * gate has param >= 0 and uses angle = scale * params[param].
 * Matrices are complex, row-major, index[0] most significant
 * (gates.py:248-336, backends/abstract_backend.py:390-400). */
enum tcm_gate_kind {
  TCM_G_MAT = 0,     /* constant 2^nq x 2^nq matrix in `mat` (1 <= nq <= 5) */
  TCM_G_RX = 1,      /* exp(-i a/2 X)            gates.py:579-594 */
  TCM_G_RY = 2,      /* exp(-i a/2 Y)            gates.py:600-615 */
  TCM_G_RZ = 3,      /* exp(-i a/2 Z)            gates.py:621-636 */
  TCM_G_RXX = 4,     /* exp(-i a/2 XX)           gates.py:863-865 */
  TCM_G_RYY = 5,     /* exp(-i a/2 YY) */
  TCM_G_RZZ = 6,     /* exp(-i a/2 ZZ) */
  TCM_G_PHASE = 7,   /* diag(1, e^{i a})         gates.py:460-478 */
  TCM_G_EXP1 = 8,    /* cos(a) I - i sin(a) G, G = `mat` with G^2 = I,
                        nq in {1,2}              gates.py:825-858 */
  TCM_G_BARRIER = 9, /* layer boundary (checkpoint candidate), no qubits */
  TCM_G_EXPG = 10    /* exp(-i a G) = (I - G^2) + cos(a) G^2 - i sin(a) G for
                        a Hermitian G = `mat` with G^3 = G (eigenvalues in
                        {-1, 0, 1}), nq in {1,2}: controlled rotations
                        (G = |1><1| (x) P / 2 scaled, gates.py:291-326),
                        controlled phase (G = |11><11|), iswap(theta)
                        (G = (XX + YY) / 2, gates.py:675-705) */
};

typedef struct tcm_gate {
  int32_t kind;        /* tcm_gate_kind */
  int32_t nq;          /* number of qubits (0 for a barrier) */
  int32_t q[5];        /* qubit indices, 0 <= q < nqubits, distinct */
  int32_t param;       /* parameter slot, or -1 */
  double theta;        /* fixed angle when param < 0 */
  double scale;        /* angle = scale * params[param] */
  const double* mat;   /* host complex matrix (re,im interleaved), or NULL */
} tcm_gate;

/* Engine options (all-zero = defaults). */
typedef struct tcm_opts {
  int32_t fuse;        /* 0: one kernel per gate ("plain");
                          1: LDS-tiled passes, several gates per HBM pass */
  int32_t tile_bits;   /* forward tile size log2 for fused passes; 0 = auto */
  int32_t adj_tile_bits; /* adjoint tile size log2; 0 = auto */
  int32_t ckpt_every;  /* adjoint: checkpoint the forward state every
                          `ckpt_every` barriers (0 = none) */
  int32_t reg_slots;   /* register-resident qubits per thread; 0 = auto (5 for
                          compiled complex64 passes on >= 20 qubits, else 4) */
  int32_t mfma;        /* 1: dense 5-qubit gates run on FP32 matrix cores,
                          one streaming pass per gate; 2: consecutive such
                          gates grouped into LDS tile passes of 2^k
                          amplitudes (k = tile_bits clamped to 10..13,
                          default 13; at most 8 gates per pass; fewer HBM
                          passes, 1.6x faster than 1 on a 30-qubit depth-20
                          brickwork); complex64 only, complex128 keeps the
                          VALU kernel */
  int32_t jit;         /* 1: compile every tile pass into its own kernel
                          (straight-line code, hiprtc) on first use of a
                          dtype/mode; identical passes share a kernel and
                          code objects are cached on disk */
  int32_t greedy;      /* 1: first-fit choice of each pass's tile qubits and
                          of each register group; 0: first-fit improved by
                          hill climbing on the number of gates a pass (or
                          group) takes, which needs fewer passes on layered
                          circuits */
} tcm_opts;

typedef struct tcm_program tcm_program; /* opaque, caller-owned */
typedef struct tcm_pauli tcm_pauli;     /* opaque, caller-owned */

typedef struct tcm_program_info {
  int32_t nqubits, nparams, ngates, npasses;
  int32_t nckpt;              /* checkpoints stored by the adjoint */
  int32_t tile_bits;          /* tile chosen for fused passes (0 = plain) */
  int32_t reserved[2];
} tcm_program_info;

/* ---- library ------------------------------------------------------------ */

int tcm_abi_version(void);
const char* tcm_last_error(void);

/* ---- gate programs --------------------------------------------------------
 * Replaces the per-call graph building + contraction-path search of
 * BaseCircuit.apply_general_gate (basecircuit.py:120-245) and the contractor
 * (cons.py:634-663): the gate list (the reference's QIR, basecircuit.py:
 * 131-142) is lowered once, fused into LDS-tiled passes and cached. */
int tcm_program_build(const tcm_gate* gates, int32_t ngates, int32_t nqubits,
                      int32_t nparams, const tcm_opts* opts, tcm_program** out);
int tcm_program_free(tcm_program* prog);
int tcm_program_get_info(const tcm_program* prog, tcm_program_info* info);

/* Introspection (no reference counterpart): writes a JSON description of the
 * schedule used for dtype (TCM_C64/TCM_C128) and mode (0 forward, 1 adjoint)
 * into buf (NUL-terminated, truncated to cap); returns the full length. */
int64_t tcm_program_describe(const tcm_program* prog, int32_t dtype,
                             int32_t mode, char* buf, int64_t cap);

/* Pass compiler (opts.jit = 1, no reference counterpart): compile the
 * schedule for dtype and mode (0 forward, 1 adjoint) now instead of on first
 * use; needs no device (targets the current device's architecture, else
 * gfx950).  *seconds (nullable) receives the compile time, 0 when the code
 * objects came from the on-disk cache. */
int tcm_program_compile(tcm_program* prog, int32_t dtype, int32_t mode,
                        double* seconds);

/* Scratch bytes tcm_forward needs for `batch` states (per-item gate
 * matrices evaluated from the parameters). */
int tcm_forward_workspace_bytes(const tcm_program* prog, int32_t dtype,
                                int64_t batch, int64_t* bytes);

/* Forward simulation.  Replaces Circuit.wavefunction (circuit.py:792-810).
 * If init_zero != 0 every state in the batch is first set to |0..0>
 * (all_zero_nodes, basecircuit.py:47-60); otherwise `state` holds the input
 * states (Circuit(n, inputs=...), circuit.py:86-96).  param_stride = number
 * of parameter values between consecutive batch items (0 = shared).
 * `ckpts` (nullable) receives info.nckpt x batch x 2^n amplitudes: the
 * forward states at the checkpoint boundaries, for tcm_adjoint_vjp. */
int tcm_forward(tcm_program* prog, int32_t dtype, const void* params,
                int64_t param_stride, void* state, int64_t batch,
                int32_t init_zero, void* ckpts, void* workspace,
                int64_t ws_bytes, void* stream);

/* ---- Pauli sums -----------------------------------------------------------
 * codes: host int8 [nterms][nqubits], 0=I 1=X 2=Y 3=Z (quantum.py:1025-1068);
 * coeffs: host real weights [nterms]. */
int tcm_pauli_build(const int8_t* codes, const double* coeffs, int32_t nterms,
                    int32_t nqubits, tcm_pauli** out);
int tcm_pauli_free(tcm_pauli* ham);

/* Pass compiler for Pauli sums (no reference counterpart): compile the
 * tiled passes of `ham` for dtype into straight-line kernels, used by every
 * later expectation / apply / adjoint call with this sum; needs no device.
 * *seconds (nullable): compile time, 0 when cached. */
int tcm_pauli_compile(tcm_pauli* ham, int32_t dtype, double* seconds);

/* Scratch bytes for expectation / apply over `batch` states. */
int tcm_pauli_workspace_bytes(const tcm_pauli* ham, int32_t dtype,
                              int64_t batch, int64_t* bytes);

/* out[b] = sum_k w_k <psi_b|P_k|psi_b> as complex fp64 (re, im) pairs,
 * unnormalised.  Replaces Circuit.expectation_ps (abstractcircuit.py:
 * 1208-1288) summed over terms by the caller (vqe_tc.py:75-81), with fp64
 * accumulation instead of the reference's state-dtype sum. */
int tcm_expect_pauli_sum(const tcm_pauli* ham, int32_t dtype, const void* state,
                         int64_t batch, double* out, void* workspace,
                         int64_t ws_bytes, void* stream);

/* The same result as tcm_expect_pauli_sum without a workspace and without
 * any device allocation.  No counterpart in the reference: for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The partial sums live in a static
 * device array of 8 MiB per GPU (held until the process ends); a call whose
 * batch needs more (tcm_pauli_workspace_bytes > 8 MiB) returns
 * TCM_E_WORKSPACE.  `ham` must already be prepared on the current device
 * for `dtype` (one tcm_expect_pauli_sum call with it), else TCM_E_ARG.
 * Calls that overlap on one GPU share the array and race. */
int tcm_expect_pauli_sum_noalloc(const tcm_pauli* ham, int32_t dtype,
                                 const void* state, int64_t batch, double* out,
                                 void* stream);

/* out_state = alpha_b * (H psi_b), alpha given per batch item on the device
 * (fp64 [batch], NULL = 1).  The vector-Jacobian product of expectation
 * (backward of the reference's autograd through the expectation network,
 * basecircuit.py:267-319). */
int tcm_apply_pauli_sum(const tcm_pauli* ham, int32_t dtype, const void* state,
                        void* out_state, int64_t batch, const double* alpha,
                        void* stream);

/* Per-term expectations: out[b][t] = w_t <psi_b|P_t|psi_b> as complex fp64
 * (re, im) pairs, terms in the order given to tcm_pauli_build, all terms
 * and states in one launch.  Replaces the reference's vmap of
 * any_measurements over Pauli structures (templates/measurements.py:18-83,
 * one expectation contraction per structure) and
 * parameterized_local_measurements (templates/measurements.py:86-153).
 * nterms and batch are each at most 65535.  Workspace:
 * tcm_pauli_terms_workspace_bytes. */
int tcm_pauli_terms_workspace_bytes(const tcm_pauli* ham, int64_t batch,
                                    int64_t* bytes);
int tcm_expect_pauli_terms(const tcm_pauli* ham, int32_t dtype,
                           const void* state, int64_t batch, double* out,
                           void* workspace, int64_t ws_bytes, void* stream);

/* out_state_b = sum_t weights[b][t] w_t P_t psi_b (weights: device fp64
 * [batch][nterms]) -- the vector-Jacobian product of
 * tcm_expect_pauli_terms (the backward of any_measurements through the
 * reference's autograd, basecircuit.py:267-319). */
int tcm_apply_pauli_terms(const tcm_pauli* ham, int32_t dtype,
                          const void* state, void* out_state, int64_t batch,
                          const double* weights, void* stream);

/* Sparse form of the sum: *nnz = nterms * 2^nqubits entries, written by
 * tcm_pauli_coo as rows / cols (device int64 [nnz]) and values (device
 * complex [nnz] of dtype): term t's entry t * 2^n + y sits at (y, y ^ flip)
 * with value w_t (-i)^ny (-1)^popcount(y & yzmask) -- the reference's
 * PauliString2COO / ps2coo_core (quantum.py:1415-1485); duplicate
 * coordinates of different terms add (PauliStringSum2COO,
 * quantum.py:1304-1356). */
int tcm_pauli_nnz(const tcm_pauli* ham, int64_t* nnz);
int tcm_pauli_coo(const tcm_pauli* ham, int32_t dtype, int64_t* rows,
                  int64_t* cols, void* vals, void* stream);

/* out = <psi|A|psi> (complex fp64 (re, im)) for a 2^n x 2^n CSR matrix A
 * (rowptr int64 [2^n + 1], cols int64, vals complex of dtype, all device),
 * as sparse_expectation (templates/measurements.py:177-191). */
int tcm_csr_expect_workspace_bytes(int32_t nqubits, int64_t* bytes);
int tcm_csr_expect(int32_t nqubits, int32_t dtype, const int64_t* rowptr,
                   const int64_t* cols, const void* vals, const void* state,
                   double* out, void* workspace, int64_t ws_bytes,
                   void* stream);

/* y_b = A x_b for a 2^n x 2^n CSR matrix A (as tcm_csr_expect) and `batch`
 * vectors x (device [batch][2^n] complex of dtype); x and y must not
 * alias.  The vector-Jacobian product of sparse_expectation
 * (templates/measurements.py:177-191): dL/dpsi = conj(g) A psi +
 * g A^dagger psi, the second product with the CSR form of A^dagger. */
int tcm_csr_matvec(int32_t nqubits, int32_t dtype, const int64_t* rowptr,
                   const int64_t* cols, const void* vals, const void* x,
                   void* y, int64_t batch, void* stream);

/* out[b] = <a_b|b_b> (complex fp64) -- expectation of general operators after
 * they were applied to a copy of the state (circuit.py:914-990). */
int tcm_inner_workspace_bytes(int32_t nqubits, int64_t batch, int64_t* bytes);
int tcm_inner(int32_t nqubits, int32_t dtype, const void* a, const void* b,
              int64_t batch, double* out, void* workspace, int64_t ws_bytes,
              void* stream);

/* ---- measurement ----------------------------------------------------------
 * out[b][i] = |psi_b[i]|^2 in the state's real type (float for complex64):
 * Circuit.probability (basecircuit.py:510-524). */
int tcm_probability(int32_t nqubits, int32_t dtype, const void* state,
                    int64_t batch, void* out, void* stream);

/* Draws from p_i = |psi_b[i]|^2 by inverse CDF, the reference's
 * backend.probability_sample (abstract_backend.py:1124-1157) behind
 * Circuit.sample(allow_state=True) and quantum.measurement_counts
 * (basecircuit.py:586-604, quantum.py:2218-2322): for uniforms u (device
 * fp64 [batch][shots], each in [0, 1]) r = total * (1 - u) and
 * out[b][s] (device int64) is the smallest index whose inclusive prefix sum
 * of p reaches r.  Prefix sums in fp64 (the reference sums the normalised
 * vector in the state dtype).  Workspace: tcm_sample_workspace_bytes. */
int tcm_sample_workspace_bytes(int32_t nqubits, int64_t batch, int64_t* bytes);
int tcm_sample(int32_t nqubits, int32_t dtype, const void* state, int64_t batch,
               const double* uniforms, int64_t shots, int64_t* out,
               void* workspace, int64_t ws_bytes, void* stream);
/* The same draws from probability vectors p (device [batch][2^nqubits],
 * float for dtype TCM_C64, double for TCM_C128), normalised by their sum:
 * backend.probability_sample on a given p (measurement_counts with
 * is_prob=True, quantum.py:2290-2292). */
int tcm_sample_prob(int32_t nqubits, int32_t dtype, const void* p, int64_t batch,
                    const double* uniforms, int64_t shots, int64_t* out,
                    void* workspace, int64_t ws_bytes, void* stream);

/* Sequential measurement of `qubits` (host int32 [nmeas], distinct, in
 * order) as measure_jit (basecircuit.py:371-445) with external randomness
 * status (device fp64 [batch][nmeas]): at step k, pu = P(qubit k = 0 and
 * the outcomes so far) / p, outcome 1 iff status - pu + 0.31415926e-12 > 0,
 * p *= pu (outcome 0) or 1 - pu, p starting at 1.  outcomes: device fp64
 * [batch][nmeas] (0.0 / 1.0); prob: device fp64 [batch] = final p.
 * perfect_sampling (basecircuit.py:359-369) is all qubits in order.
 * Workspace: tcm_measure_workspace_bytes. */
int tcm_measure_workspace_bytes(int32_t nqubits, int64_t batch, int64_t* bytes);
int tcm_measure(int32_t nqubits, int32_t dtype, const void* state, int64_t batch,
                const int32_t* qubits, int32_t nmeas, const double* status,
                double* outcomes, double* prob, void* workspace,
                int64_t ws_bytes, void* stream);

/* ---- gradients ------------------------------------------------------------
 * Replaces backend.value_and_grad / vvag over an energy function
 * (backends/pytorch_backend.py:497-544, 673-709): one forward, E and
 * lambda = H psi in fp64-accumulated kernels, then a reverse sweep
 * (adjoint method) with forward checkpoints every opts.ckpt_every barriers.
 * mode 0 = tcm_adjoint_energy_grad, mode 1 = tcm_adjoint_vjp. */
int tcm_adjoint_workspace_bytes(const tcm_program* prog, const tcm_pauli* ham,
                                int32_t dtype, int64_t batch, int32_t mode,
                                int64_t* bytes);

/* energy[b] = <psi_b|H|psi_b>, grad[b][p] = dE/dparams[b][p] (both fp64).
 * `inputs` (nullable) are the initial states; NULL = |0..0>.  All states,
 * checkpoints and partial sums live in the workspace. */
int tcm_adjoint_energy_grad(tcm_program* prog, int32_t dtype,
                            const void* params, int64_t param_stride,
                            const tcm_pauli* ham, const void* inputs,
                            int64_t batch, double* energy, double* grad,
                            void* workspace, int64_t ws_bytes, void* stream);

/* Vector-Jacobian product of the forward map params -> psi (torch
 * convention, grad_state = dL/dRe psi + i dL/dIm psi):
 * grad_params[b][p] = Re <grad_state_b | d psi_b / d params[b][p]> (fp64),
 * grad_input (nullable) = U^dagger grad_state.  `state` holds the final
 * states produced by tcm_forward with the same params (read only) and
 * `ckpts` the checkpoints that call stored (required when info.nckpt > 0). */
int tcm_adjoint_vjp(tcm_program* prog, int32_t dtype, const void* params,
                    int64_t param_stride, const void* ckpts, const void* state,
                    const void* grad_state, int64_t batch, double* grad_params,
                    void* grad_input, void* workspace, int64_t ws_bytes,
                    void* stream);

/* ---- batched uneven circuits ----------------------------------------------
 * Many circuits of different programs / qubit counts in few launches (the
 * reference's vmap over circuit structures, backends/pytorch_backend.py:
 * 605-617, and chunked adaptive_vmap, experimental.py:17-74): descriptor i
 * runs its program on `batch` states at `state` with parameters at `params`
 * (param_stride apart).  Pass r of every program goes into one launch. */
typedef struct tcm_multi_desc {
  tcm_program* prog;
  const void* params;
  int64_t param_stride;
  void* state;
  int64_t batch;
} tcm_multi_desc;

int tcm_multi_workspace_bytes(const tcm_multi_desc* descs, int32_t ndesc,
                              int32_t dtype, int64_t* bytes);
int tcm_forward_multi(const tcm_multi_desc* descs, int32_t ndesc,
                      int32_t dtype, int32_t init_zero, void* workspace,
                      int64_t ws_bytes, void* stream);

/* Compiled multi-program launches (no reference counterpart beyond the
 * same vmap over structures): for a fixed list of programs, pass r of every
 * program is generated into one kernel (straight-line code per program, a
 * block running as many of a program's tiles side by side as its thread
 * count allows) and compiled once (hiprtc, cached on disk).  A step of the
 * reference's vmapped uneven circuits is then one launch per pass index
 * instead of one per program and pass.  Passes that cannot be compiled
 * (plain gates, matrix-core passes) run as single-program launches in the
 * same order, as do passes whose tiles side by side would need more than
 * 64 KiB of LDS per block.  `batches` (nullable): the expected batch of
 * each program, weighting the choice of threads per block toward the
 * programs with the most amplitudes.  At most 48 programs per plan;
 * *seconds (nullable): compile time, 0 when cached. */
typedef struct tcm_multi_plan tcm_multi_plan; /* opaque, caller-owned */
int tcm_multi_plan_build(tcm_program* const* progs, const int64_t* batches,
                         int32_t nprog, int32_t dtype, tcm_multi_plan** out,
                         double* seconds);
int tcm_multi_plan_free(tcm_multi_plan* plan);
/* Forward of every program of the plan on its descriptor (descs[i].prog
 * must be the plan's program i); workspace as tcm_multi_workspace_bytes. */
int tcm_forward_multi_plan(tcm_multi_plan* plan, const tcm_multi_desc* descs,
                           int32_t ndesc, int32_t init_zero, void* workspace,
                           int64_t ws_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TCM_H_ */
