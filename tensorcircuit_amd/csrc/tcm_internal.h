// Internal structures of libtcm shared by the host scheduler (program.cpp)
// and the HIP kernels (kernels.hip).
#pragma once

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tcm.h"

namespace tcm {

// ---------------------------------------------------------------------------
// Device op kinds.  An op acts on register slots (fused passes) or on global
// bit positions (plain per-gate kernels).  Matrix layouts (T = float|double,
// complex numbers as (re, im) pairs, row-major, first qubit most significant):
//   U1 8T, R1 4T (real 2x2), D1 4T (2 diagonal entries), U2 32T, D2 8T,
//   DG 2*2^R T (diagonal over all slots of a register group), UK 2*4^k T.
// ---------------------------------------------------------------------------
enum OpKind : int32_t {
  OK_U1 = 0,
  OK_R1 = 1,
  OK_D1 = 2,
  OK_X1 = 3,    // Pauli X (register swap, no matrix)
  OK_U2 = 4,
  OK_D2 = 5,
  OK_CX = 6,    // s0 control, s1 target (no matrix)
  OK_CZ = 7,    // no matrix
  OK_SWAP = 8,  // no matrix
  OK_DG = 9,    // fused diagonal over a register group
  OK_UK = 10,   // dense k-qubit, k = 3..5 (applied directly on the LDS tile)
  OK_NKINDS
};

// Generators for gradients: a trainable gate is U(a) = exp(-i c a G) with
// Hermitian G, so dU/da = -i c G U and dE/da = 2 c Im<lambda|G|psi_after>.
enum GenKind : int32_t {
  GN_NONE = 0,
  GN_X = 1,
  GN_Y = 2,
  GN_Z = 3,
  GN_XX = 4,
  GN_YY = 5,
  GN_ZZ = 6,
  GN_P1 = 7,  // |1><1| (phase gate; c = -1)
  GN_M1 = 8,  // general Hermitian 2x2 in the constant table
  GN_M2 = 9   // general Hermitian 4x4 in the constant table
};

// One device op.  `moff` >= 0: offset of the matrix in the per-batch table
// (T units); `moff` < 0: offset (-moff-1) in the constant table.
struct DOp {
  int32_t kind;
  int32_t s0, s1;      // slots (fused) or bit positions (plain)
  int32_t moff;
  int32_t gen;         // GenKind (adjoint)
  int32_t goff;        // generator matrix (GN_M1/GN_M2) constant-table offset
  int32_t gop;         // gradient-op index, -1 if none
  int32_t nmem;        // OK_DG: number of member gradients; OK_UK: k
  int32_t mem0;        // OK_DG: first entry in the member table; OK_UK: first
                       // entry in the bit table
  int32_t gloc;        // index of `gop` among the gradient ops of its pass
  int32_t flags;       // OP_F_* bits
};

// DOp flags
// OP_F_SHEAR: an ry rotation whose table entry also holds the three-shear
// form of its reduced angle (see k_prep): entries 4, 5, 6 = t, s, sigma
enum : int32_t { OP_F_MFMA = 1, OP_F_SHEAR = 2 };

// Gradient member of an OK_DG op (diagonal generator on slots of the group).
struct DMember {
  int32_t gen;   // GN_Z, GN_ZZ, GN_P1
  int32_t s0, s1;
  int32_t gop;
  int32_t gloc;  // index among the gradient ops of the pass
};

// Register group inside a fused pass: each thread holds 2^R amplitudes of the
// LDS tile at tile addresses swz(tbase(tid)) ^ S[j].
struct DGroup {
  int32_t cb[20];    // tile-local bit positions of the thread bits (k - R)
  int32_t nrun;      // cb as runs of consecutive bits: run[r] = src | len << 8 | dst << 16
  int32_t run[20];   // (thread bits src .. src+len-1 -> tile bits dst ..)
  int32_t S[32];     // swizzled tile offsets of the 2^R register slots
  int32_t op0, nops;
  int32_t lds_op;    // 1: group holds a single OK_UK op applied on the LDS tile
  int32_t ukbits[5]; // tile-local bits of that op (first = most significant)
};

// dense 5-qubit blocks per matrix-core tile pass (their A operands share the
// LDS with the tile: 8 KiB each)
constexpr int kMfmaTileMaxOps = 8;

struct DPass {
  int32_t k;          // tile bits
  int32_t L;          // contiguous low bits inside the tile (run length 2^L)
  int32_t nblk_log;   // n - k
  int32_t tthreads_log;  // log2(threads) = k - R
  int32_t group0, ngroups;
  int32_t runoff0;    // first entry of this pass in the run-offset table
  int32_t ngop;       // gradient ops in this pass
  int32_t gop_list0;  // first entry of this pass in Sched::gop_list
  int32_t plain_op;   // >= 0: not a tile pass but one plain gate (Sched::dop index)
  int32_t full;       // 1: the pass holds dense 4x4 ops / generators (full kernel)
  int32_t blk_bits[32];  // global bit positions of the block bits (ascending)
  int32_t mfma;       // 1: a tile pass of dense 5-qubit blocks on FP32 matrix cores
  int32_t mop0, nmop; // its blocks: Sched::dop entries (OK_UK, tile-local bits)
  int32_t pad;
};

// Matrix recipes evaluated on the device per batch item (prep kernel).
struct DRecipe {
  int32_t kind;    // tcm_gate_kind (RX..EXP1), or -1 = DG product
  int32_t param;   // parameter slot or -1
  int32_t out;     // offset in the per-batch table
  int32_t layout;  // OK_R1 / OK_D1 / OK_U1 / OK_D2 / OK_U2 output layout
  int32_t goff;    // EXP1 generator (constant table), or DG: first factor
  int32_t nfac;    // DG: number of factors
  int32_t R;       // DG: register slots
  int32_t pad;
  double theta, scale;
};

// Table entries of an ry recipe: [c, -s, s, c] of the exact rotation by
// phi = angle / 2, then [t, s', sigma, 0]: phi reduced to phi' in
// [-pi/2, pi/2] with R(phi) = sigma R(phi'), t = tan(phi' / 2), s' = sin phi'
constexpr int kRySlot = 8;

// Factor of a DG recipe: a diagonal (D1 or D2 layout, per-batch or constant)
// evaluated on slots s0 (, s1) of the group index.
struct DFactor {
  int32_t src;     // >= 0 per-batch offset; < 0 constant offset (-src-1)
  int32_t layout;  // OK_D1, OK_D2, OK_CZ
  int32_t s0, s1;
};

// ---------------------------------------------------------------------------
// Host-side program
// ---------------------------------------------------------------------------
struct LGate {
  int32_t src;        // tcm_gate_kind
  int32_t nq;
  int32_t q[5];
  int32_t bits[5];    // global bit positions, n-1-q
  int32_t param;
  double theta, scale;
  std::vector<double> mat;   // MAT: the matrix; EXP1: the generator
  int32_t okind;      // OpKind for application
  int32_t gen;        // GenKind
  double gfac;        // dE/dparam = gfac * Im<lambda|G|psi>  (energy mode, x2)
  int32_t mslot;      // per-batch matrix offset (T units) or -1
  int32_t cslot;      // constant matrix offset or -1 (MAT and UK kinds)
  int32_t goff;       // generator constant offset or -1
  int32_t gop;        // gradient op index or -1
  bool barrier;       // layer marker
  int32_t layer;      // number of barriers before this gate
};

struct HGroup {
  std::vector<int32_t> slots;   // tile-local bits held in registers, R entries
  std::vector<int32_t> gates;   // indices into LGate list, program order
  bool lds_op = false;
};

struct HPass {
  std::vector<int32_t> tile;    // global bit positions, ascending (k); empty = plain gate
  int32_t L = 0;
  std::vector<HGroup> groups;
  std::vector<int32_t> gates;   // all gates of the pass, program order
  bool mfma = false;            // dense 5-qubit blocks on matrix cores (no groups)
};

struct DeviceTables;  // per-device uploaded tables (kernels.hip)

// Compiled pass kernels (pass compiler): kernel index per pass (-1 = the
// generic fused kernel runs it) and the pass's per-item table offset base.
struct JitPass {
  int32_t kernel = -1;
  int32_t mbase = 0;  // smallest base-recipe offset the pass reads
  int32_t dbase = 0;  // smallest merged-diagonal table offset it reads
};

// One schedule: a sequence of passes for one dtype and one mode (forward /
// adjoint), plus the flattened device descriptors and matrix recipes.
struct Sched {
  int32_t k = 0;               // tile bits, 0 = plain (one kernel per gate)
  int32_t R = 4;               // register slots per thread
  int32_t Lwant = 4;           // low bits always in a tile (coalescing)
  std::vector<HPass> passes;
  std::vector<int32_t> ckpt_after;   // pass index after which checkpoint c is taken
  // device descriptors
  std::vector<DPass> dpass;
  std::vector<DGroup> dgroup;
  std::vector<DOp> dop;         // fused: per group; plain: one per gate
  std::vector<DMember> dmem;
  std::vector<int32_t> ukbits;  // bit tables of OK_UK ops
  std::vector<uint32_t> runoff;
  std::vector<DRecipe> recipes; // base recipes + DG recipes of this schedule
  std::vector<DFactor> factors;
  int32_t mtab_size = 0;        // per-batch matrix table size (T units)
  int32_t mtab_base = 0;        // end of the base recipes; merged diagonal
                                // tables of this schedule follow
  int32_t max_ngop_pass = 0;    // largest number of gradient ops in a pass
  int32_t max_slots = 0;        // largest number of gradient partial slots per gop
  std::vector<int32_t> gop_slots;  // per gop: number of partial slots (blocks)
  std::vector<int32_t> gop_list;   // per pass (from DPass::gop_list0): gops by gloc
  std::vector<DeviceTables*> dev;  // per device
  // pass compiler output (opts.jit): 0 not built, 1 built, -1 failed
  int32_t jit_state = 0;
  std::vector<JitPass> jit;
  std::vector<std::vector<char>> jit_code;  // code objects, one per chunk
  std::vector<int32_t> jit_chunk;           // kernel -> chunk
  double jit_seconds = 0.0;                 // compile time (0 when cached)
  std::string jit_prefix;                   // kernel names: prefix + index
};

struct Program {
  int32_t n = 0, nparams = 0;
  tcm_opts opts{};
  std::vector<LGate> gates;      // without barriers
  int32_t nlayers = 0;           // barriers seen
  int32_t ngop = 0;              // trainable gate count
  std::vector<int32_t> gop_gate; // gop -> gate index
  // parameter -> gops (CSR) for the gradient reduction
  std::vector<int32_t> p_ptr, p_gops;
  std::vector<double> gfac;      // per gop: energy-mode gradient factor
  std::vector<double> ctab;      // constant matrices (double; cast on upload)
  std::vector<int32_t> ckpt_gate;  // gate count before checkpoint c
  Sched sched[2][2];             // [dtype][mode 0 fwd / 1 adjoint]
  std::mutex mu;
};

// pass compiler (passgen.cpp, jit.cpp)
// one program's part of a multi-program forward launch (passgen.cpp)
struct MultiPassJob {
  int32_t prog;       // index in the plan's program list
  int32_t nblk_log;   // tiles per state = 2^nblk_log
  int32_t per_block;  // tiles per block of the launch (sub-blocks)
  int32_t mbase, dbase;
};
struct MultiPassInfo {
  std::vector<MultiPassJob> jobs;
  int32_t ntb = 64;   // threads per block of the launch
  size_t lds = 0;     // LDS bytes per block
};
bool multi_pass_ok(const Sched& S, int r);
std::string gen_multi_prelude(int dt);
std::string gen_multi_source(const std::vector<const Program*>& progs, const std::vector<double>& weight, int dt,
                             int r, int kid, MultiPassInfo* info);
std::string gen_pass_source(const Program& P, const Sched& S, int dt, int mode, std::vector<JitPass>& jp,
                            std::vector<std::string>& bodies);
int jit_build(const Program& P, Sched& S, int dt, int mode);

// host scheduling (program.cpp)
int build_program(const tcm_gate* gates, int32_t ngates, int32_t n, int32_t nparams,
                  const tcm_opts* opts, Program** out);

// thread-local error reporting (abi.cpp)
int set_error(int code, const char* fmt, ...);

// swizzle of tile addresses in LDS: XOR of the low 4 bits with higher nibbles.
// Linear over GF(2), so swz(a ^ b) = swz(a) ^ swz(b).
inline uint32_t swz_host(uint32_t t) {
  return t ^ (((t >> 4) ^ (t >> 8) ^ (t >> 12) ^ (t >> 16)) & 0xFu);
}

}  // namespace tcm
