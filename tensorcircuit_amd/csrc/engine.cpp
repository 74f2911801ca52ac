// C ABI of libtcm (declared in include/tcm.h): argument checking, device
// table upload, workspace carving and the launch sequences for forward
// simulation, Pauli-sum energies and the adjoint reverse sweep.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "kernels.h"

using namespace tcm;

struct tcm_program {
  Program p;
};

// compiled multi-program forward launches (tcm_multi_plan_build)
struct tcm_multi_plan {
  std::vector<tcm_program*> progs;
  int dtype = 0;
  int maxp = 0;                          // most passes of any program
  std::vector<tcm::MultiPassInfo> info;  // per pass index r
  std::vector<int32_t> kid;              // per r: kernel index, -1 = none
  std::vector<std::vector<char>> code;
  std::vector<int32_t> chunk_of;
  double seconds = 0.0;
  std::mutex mu;
  std::vector<std::vector<void*>> mods, fns;  // per device
};

struct PauliDev {
  void* terms = nullptr;
  void* flat = nullptr;    // DPTermF per term in the caller's order
  void* groups = nullptr;  // DPGroup per flip mask (sparse form only)
  void* tpasses[2] = {nullptr, nullptr};
  void* tgroups[2] = {nullptr, nullptr};
  void* tterms[2] = {nullptr, nullptr};
  void* truns[2] = {nullptr, nullptr};
  // compiled tile passes (per dtype): modules and one kernel per pass
  std::vector<void*> jmods[2], jfns[2];
};

// compiled tile passes of a Pauli sum for one dtype
struct PauliJit {
  int state = 0;  // 0 none, 1 built, -1 failed
  std::vector<std::vector<char>> code;
  std::vector<int32_t> chunk;
  double seconds = 0.0;
};

// LDS-tiled schedule of a Pauli sum for one dtype
struct PauliTileSched {
  bool ok = false;                 // false: every flip group runs the per-group kernel
  std::vector<DPPass> passes;
  std::vector<DPGroupT> groups;
  std::vector<DPTermT> terms;
  std::vector<uint32_t> runoff;
  std::vector<int> legacy;         // flip groups wider than a tile
  std::vector<int> legacy_slot0;
  int nslots = 0;                  // energy-partial slots of this dtype
};

struct tcm_pauli {
  int n = 0;
  std::vector<DPGroup> groups;     // by flip mask, diagonal first
  std::vector<DPTerm> terms;
  std::vector<DPTermF> flat;       // every term in the caller's order
  int nslots = 0;                  // slots of the per-group kernels
  PauliTileSched ts[2];
  PauliJit jit[2];
  std::mutex mu;
  std::vector<PauliDev> dev;
};

namespace tcm {
std::string describe_sched(const Program& P, int dtype, int mode);
std::string gen_pauli_source(int n, int dt, const std::vector<DPPass>& passes, const std::vector<DPGroupT>& groups,
                             const std::vector<DPTermT>& terms, std::vector<std::string>& bodies);
int compile_bodies(const std::string& prelude, const std::vector<std::string>& bodies, const std::string& dump_tag,
                   std::vector<std::vector<char>>& code, std::vector<int32_t>& chunk_of, double* seconds);
int load_modules(const std::vector<std::vector<char>>& code, const std::vector<int32_t>& chunk_of,
                 const std::string& prefix, std::vector<void*>& mods_out, std::vector<void*>& fns_out);

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
}  // namespace tcm

namespace {

constexpr int64_t kAlign = 256;
inline int64_t align_up(int64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

int hip_fail(hipError_t e, const char* what) {
  return set_error((int)e, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                        \
  do {                                             \
    hipError_t e_ = (expr);                        \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

int current_device(int* dev) {
  int cnt = 0;
  hipError_t e = hipGetDeviceCount(&cnt);
  if (e != hipSuccess || cnt <= 0) {
    (void)hipGetLastError();
    return set_error(TCM_E_NODEVICE, "no HIP device available");
  }
  HIP_TRY(hipGetDevice(dev), "hipGetDevice");
  return 0;
}

template <typename V>
int upload(const std::vector<V>& v, void** out) {
  *out = nullptr;
  if (v.empty()) return 0;
  HIP_TRY(hipMalloc(out, v.size() * sizeof(V)), "hipMalloc(tables)");
  HIP_TRY(hipMemcpy(*out, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice), "hipMemcpy(tables)");
  return 0;
}

void free_tables(DeviceTables* D) {
  if (!D) return;
  void* ptrs[] = {D->dpass, D->dgroup, D->dop, D->dmem, D->ukbits, D->runoff, D->recipes, D->factors,
                  D->gop_list, D->ctab, D->p_ptr, D->p_gops, D->gfac};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (void* m : D->jit_mod)
    if (m) (void)hipModuleUnload((hipModule_t)m);
  delete D;
}

int ensure_tables(Program& P, int dtype, int mode, DeviceTables** out) {
  int dev = 0;
  if (int rc = current_device(&dev)) return rc;
  std::lock_guard<std::mutex> lk(P.mu);
  Sched& S = P.sched[dtype][mode];
  if (S.k > 0 && S.dpass.size() != S.passes.size())
    return set_error(TCM_E_ARG, "internal: schedule has %zu passes but %zu descriptors", S.passes.size(), S.dpass.size());
  if (P.opts.jit && S.k > 0 && S.jit_state == 0)
    if (int rc = jit_build(P, S, dtype, mode)) return rc;
  if (S.jit_state < 0) return set_error(TCM_E_UNSUPPORTED, "pass compiler failed for this program");
  if ((int)S.dev.size() <= dev) S.dev.resize(dev + 1, nullptr);
  if (S.dev[dev]) {
    *out = S.dev[dev];
    return 0;
  }
  DeviceTables* D = new DeviceTables();
  int rc = 0;
  rc = rc ? rc : upload(S.dpass, &D->dpass);
  rc = rc ? rc : upload(S.dgroup, &D->dgroup);
  rc = rc ? rc : upload(S.dop, &D->dop);
  rc = rc ? rc : upload(S.dmem, &D->dmem);
  rc = rc ? rc : upload(S.ukbits, &D->ukbits);
  rc = rc ? rc : upload(S.runoff, &D->runoff);
  rc = rc ? rc : upload(S.recipes, &D->recipes);
  rc = rc ? rc : upload(S.factors, &D->factors);
  rc = rc ? rc : upload(S.gop_list, &D->gop_list);
  if (!rc) {
    if (dtype == TCM_C64) {
      std::vector<float> cf(P.ctab.begin(), P.ctab.end());
      if (cf.empty()) cf.push_back(0.f);
      rc = upload(cf, &D->ctab);
    } else {
      std::vector<double> cd(P.ctab);
      if (cd.empty()) cd.push_back(0.0);
      rc = upload(cd, &D->ctab);
    }
  }
  rc = rc ? rc : upload(P.p_ptr, &D->p_ptr);
  rc = rc ? rc : upload(P.p_gops, &D->p_gops);
  rc = rc ? rc : upload(P.gfac, &D->gfac);
  rc = rc ? rc : jit_load(S, *D);
  if (rc) {
    free_tables(D);
    return rc;
  }
  S.dev[dev] = D;
  *out = D;
  return 0;
}

inline size_t tsize(int dtype) { return dtype == TCM_C64 ? sizeof(float) : sizeof(double); }

struct Carve {
  char* base;
  int64_t off = 0;
  explicit Carve(void* b) : base((char*)b) {}
  void* take(int64_t bytes) {
    void* p = base ? base + off : nullptr;
    off += align_up(bytes);
    return p;
  }
};

// ---- forward ---------------------------------------------------------------
int64_t forward_ws(const Program& P, int dtype, int64_t batch) {
  const Sched& S = P.sched[dtype][0];
  return align_up(batch * (int64_t)std::max(S.mtab_size, 1) * (int64_t)tsize(dtype));
}

// one forward pass pi (tiled, matrix-core or plain); ckout: checkpoint copy
// written by a tiled pass (nullable)
template <typename T>
int run_one_pass(Program& P, const Sched& S, const DeviceTables& D, int pi, const void* mtab, void* state,
                 int64_t batch, void* ckout, hipStream_t st, bool zinit = false) {
  const int n = P.n;
  const bool mfma = S.k > 0 && S.dpass[pi].mfma;
  const bool tiled = S.k > 0 && S.dpass[pi].plain_op < 0 && !mfma;
  if (mfma) {
    if constexpr (std::is_same<T, float>::value) {
      if (!launch_mfma_tile(S.dpass[pi], (const DPass*)D.dpass, pi, (const DOp*)D.dop, (const int32_t*)D.ukbits,
                            (const uint32_t*)D.runoff, D.ctab, state, n, batch, st))
        return set_error(TCM_E_UNSUPPORTED, "matrix-core tile pass %d: %d blocks on %d tile bits not supported", pi,
                         S.dpass[pi].nmop, S.dpass[pi].k);
    } else
      return set_error(TCM_E_UNSUPPORTED, "matrix-core passes are complex64 only");
  } else if (tiled) {
    if (!S.jit.empty() && S.jit[pi].kernel >= 0) {
      if (int rc = jit_launch_fwd(S, D, pi, sizeof(T) == 8, mtab, D.ctab, state, batch, n, ckout, st, zinit ? 1 : 0))
        return rc;
    } else if (S.R == 4) {
      launch_fused_fwd<T, 4>(S, D, pi, mtab, D.ctab, state, batch, n, ckout, st);
    } else {
      return set_error(TCM_E_UNSUPPORTED, "register slots R=%d not built", S.R);
    }
  } else {
    const DOp& op = S.k > 0 ? S.dop[S.dpass[pi].plain_op] : S.dop[pi];
    launch_plain_gate<T, false>(op, (const int32_t*)D.ukbits, mtab, S.mtab_size, D.ctab, state, n, batch, st);
  }
  return 0;
}

// forward passes; init_zero: the state starts as |0..0>, generated by a
// compiled first pass itself (no separate write and read of the state)
// unless the first pass is of another kind or the initial state is a
// checkpoint
template <typename T>
int run_forward_passes(Program& P, const Sched& S, const DeviceTables& D, const void* mtab, void* state,
                       int64_t batch, void* ckpts, hipStream_t st, bool init_zero = false) {
  const int n = P.n;
  const int64_t N = 1ll << n;
  const int64_t ck_stride = batch * N;
  auto ckpt_ptr = [&](size_t c) { return (void*)((char*)ckpts + (int64_t)c * ck_stride * 2 * sizeof(T)); };
  bool ck0 = false;
  if (ckpts)
    for (int a : S.ckpt_after) ck0 = ck0 || a < 0;
  const bool zfirst = init_zero && !ck0 && !S.passes.empty() && S.k > 0 && !S.dpass[0].mfma &&
                      S.dpass[0].plain_op < 0 && !S.jit.empty() && S.jit[0].kernel >= 0;
  if (init_zero && !zfirst) launch_init_zero<T>(state, n, batch, st);
  // checkpoints of the initial state (before any pass)
  if (ckpts)
    for (size_t c = 0; c < S.ckpt_after.size(); ++c)
      if (S.ckpt_after[c] < 0) launch_copy<T>(ckpt_ptr(c), state, ck_stride, st);
  const int np = (int)S.passes.size();
  for (int pi = 0; pi < np; ++pi) {
    const bool mfma = S.k > 0 && S.dpass[pi].mfma;
    const bool tiled = S.k > 0 && S.dpass[pi].plain_op < 0 && !mfma;
    void* ckout = nullptr;
    std::vector<size_t> extra;
    if (ckpts)
      for (size_t c = 0; c < S.ckpt_after.size(); ++c)
        if (S.ckpt_after[c] == pi) {
          if (!ckout && tiled) ckout = ckpt_ptr(c);
          else extra.push_back(c);
        }
    if (int rc = run_one_pass<T>(P, S, D, pi, mtab, state, batch, ckout, st, zfirst && pi == 0)) return rc;
    for (size_t c : extra) launch_copy<T>(ckpt_ptr(c), state, ck_stride, st);
  }
  return 0;
}

template <typename T>
int forward_impl(Program& P, int dtype, const void* params, int64_t pstride, void* state, int64_t batch,
                 int init_zero, void* ckpts, void* ws, int64_t ws_bytes, hipStream_t st) {
  DeviceTables* D = nullptr;
  if (int rc = ensure_tables(P, dtype, 0, &D)) return rc;
  const Sched& S = P.sched[dtype][0];
  Carve cv(ws);
  void* mtab = cv.take(batch * (int64_t)std::max(S.mtab_size, 1) * sizeof(T));
  if (cv.off > ws_bytes) return set_error(TCM_E_WORKSPACE, "workspace %lld < %lld bytes", (long long)ws_bytes, (long long)cv.off);
  launch_prep<T>(S, *D, D->ctab, params, pstride, mtab, batch, st);
  if (int rc = run_forward_passes<T>(P, S, *D, mtab, state, batch, ckpts, st, init_zero != 0)) return rc;
  HIP_TRY(hipGetLastError(), "forward launch");
  return 0;
}

// ---- Pauli -----------------------------------------------------------------
int ensure_pauli(tcm_pauli* H, int dt, PauliDev** out) {
  int dev = 0;
  if (int rc = current_device(&dev)) return rc;
  std::lock_guard<std::mutex> lk(H->mu);
  if ((int)H->dev.size() <= dev) H->dev.resize(dev + 1);
  PauliDev& pd = H->dev[dev];
  if (!pd.terms) {
    if (int rc = upload(H->terms, &pd.terms)) return rc;
  }
  if (!pd.groups) {
    if (int rc = upload(H->groups, &pd.groups)) return rc;
  }
  if (!pd.flat) {
    if (int rc = upload(H->flat, &pd.flat)) return rc;
  }
  const PauliTileSched& S = H->ts[dt];
  if (S.ok && !pd.tpasses[dt]) {
    int rc = upload(S.passes, &pd.tpasses[dt]);
    rc = rc ? rc : upload(S.groups, &pd.tgroups[dt]);
    rc = rc ? rc : upload(S.terms, &pd.tterms[dt]);
    rc = rc ? rc : upload(S.runoff, &pd.truns[dt]);
    if (rc) return rc;
  }
  if (H->jit[dt].state == 1 && pd.jfns[dt].empty() && !H->jit[dt].chunk.empty())
    if (int rc = load_modules(H->jit[dt].code, H->jit[dt].chunk, "tcm_pauli_k", pd.jmods[dt], pd.jfns[dt])) return rc;
  *out = &pd;
  return 0;
}

// compile the tiled passes of H for dtype dt (caller holds no lock)
int pauli_jit_build(tcm_pauli* H, int dt, double* seconds) {
  std::lock_guard<std::mutex> lk(H->mu);
  PauliJit& J = H->jit[dt];
  if (seconds) *seconds = 0.0;
  if (J.state == 1) return 0;
  if (J.state < 0) return set_error(TCM_E_UNSUPPORTED, "pass compiler failed for this Pauli sum");
  const PauliTileSched& S = H->ts[dt];
  if (!S.ok) {
    J.state = 1;  // nothing tiled: the per-group kernels run
    return 0;
  }
  std::vector<std::string> bodies;
  const std::string prelude = gen_pauli_source(H->n, dt, S.passes, S.groups, S.terms, bodies);
  if (int rc = compile_bodies(prelude, bodies, "pauli", J.code, J.chunk, &J.seconds)) {
    J.state = -1;
    return rc;
  }
  J.state = 1;
  if (seconds) *seconds = J.seconds;
  return 0;
}

int pauli_slots(const tcm_pauli* H, int dt) { return H->ts[dt].ok ? H->ts[dt].nslots : H->nslots; }

int pauli_blocks(int n, bool flip) {
  const int64_t items = flip ? (1ll << (n - 1)) : (1ll << n);
  int64_t b = items / 2048;
  return (int)std::max<int64_t>(1, std::min<int64_t>(1024, b));
}

int64_t pauli_ws(const tcm_pauli* H, int dt, int64_t batch) {
  return align_up(batch * (int64_t)pauli_slots(H, dt) * 2 * sizeof(double)) + align_up(batch * 2 * sizeof(double));
}

template <typename T>
int pauli_energy_apply(tcm_pauli* H, int dt, const void* psi, void* out, const double* alpha, int64_t batch,
                       double* epart, bool energy, bool apply, hipStream_t st, PauliDev* ready = nullptr) {
  // `ready`: device objects already checked present (no upload or load)
  PauliDev* pd = ready;
  if (!pd)
    if (int rc = ensure_pauli(H, dt, &pd)) return rc;
  const DPTerm* td = (const DPTerm*)pd->terms;
  const int64_t estride = (int64_t)pauli_slots(H, dt) * 2;
  const PauliTileSched& S = H->ts[dt];
  if (!S.ok) {
    for (size_t g = 0; g < H->groups.size(); ++g)
      launch_pauli<T>(H->groups[g], td, H->n, psi, out, g == 0 ? 1 : 0, alpha, epart, estride, batch, energy, apply,
                      st);
    return 0;
  }
  bool first = true;
  const bool jit = !pd->jfns[dt].empty();
  for (size_t pi = 0; pi < S.passes.size(); ++pi) {
    if (jit) {
      // parameter list of tcm_pauli_k<pi> (passgen.cpp)
      const void* a_psi = psi;
      const void* a_lin = first ? nullptr : out;
      void* a_lout = out;
      long long sstride = 1ll << H->n, es = estride;
      const double* a_al = alpha;
      double* a_ep = epart;
      long long total = batch << S.passes[pi].nblk_log;
      int flags = (energy ? 1 : 0) | (apply ? 2 : 0);
      void* args[] = {&a_psi, &a_lin, &a_lout, &sstride, &a_al, &a_ep, &es, &total, &flags};
      const hipError_t e = hipModuleLaunchKernel((hipFunction_t)pd->jfns[dt][pi], (unsigned)total, 1, 1,
                                                 1u << S.passes[pi].nthr_log, 1, 1, 0, st, args, nullptr);
      if (e != hipSuccess) return set_error((int)e, "hipModuleLaunchKernel(pauli): %s", hipGetErrorString(e));
    } else {
      launch_pauli_tile<T>(S.passes[pi], (const DPPass*)pd->tpasses[dt], (int)pi, (const DPGroupT*)pd->tgroups[dt],
                           (const DPTermT*)pd->tterms[dt], (const uint32_t*)pd->truns[dt], psi,
                           first ? nullptr : out, out, H->n, alpha, epart, estride, batch, energy, apply, st);
    }
    first = false;
  }
  for (size_t i = 0; i < S.legacy.size(); ++i) {
    DPGroup G = H->groups[S.legacy[i]];
    G.slot0 = S.legacy_slot0[i];
    launch_pauli<T>(G, td, H->n, psi, out, first ? 1 : 0, alpha, epart, estride, batch, energy, apply, st);
    first = false;
  }
  return 0;
}

// pext: gather the bits of x at positions pos[0..] into bits 0..
uint32_t pext_bits(uint64_t x, const std::vector<int>& pos) {
  uint32_t r = 0;
  for (size_t i = 0; i < pos.size(); ++i)
    if ((x >> pos[i]) & 1) r |= 1u << i;
  return r;
}

// Pack flip groups into LDS tiles: a tile = the low bits (coalescing) plus
// the X/Y positions of the groups it evaluates, padded with the lowest free
// bits; a group wider than a tile keeps the per-group kernel.
void pauli_tile_sched(tcm_pauli* H, int dt) {
  PauliTileSched& S = H->ts[dt];
  const int n = H->n;
  if (n < 10) {
    S.ok = false;
    return;
  }
  S.ok = true;
  const uint64_t all = (n >= 64) ? ~0ull : ((1ull << n) - 1);
  // launches of the first-fit packing below with k-bit tiles (a group too
  // wide for a tile is a launch of its own)
  auto npasses = [&](int k, uint64_t lmask) {
    std::vector<uint64_t> rem;
    int np = 0;
    for (const DPGroup& G : H->groups)
      if (__builtin_popcountll(G.flip | lmask) <= k) rem.push_back(G.flip);
      else ++np;
    while (!rem.empty()) {
      uint64_t tm = (n <= k) ? all : lmask;
      std::vector<uint64_t> rest;
      for (uint64_t f : rem)
        if (__builtin_popcountll(tm | f) <= k) tm |= f;
        else rest.push_back(f);
      rem.swap(rest);
      ++np;
    }
    return np;
  };
  // complex64 tiles of 12 bits (32 KiB of LDS, several blocks per CU), or of
  // 14 bits (128 KiB, one block of 1024 threads per CU) where that saves a
  // pass: a 24-qubit sum of single X terms then takes 2 passes instead of 3
  int k = std::min(n, dt == 0 ? 12 : 11);
  const int Lw = std::min(dt == 0 ? 4 : 3, k);
  const uint64_t lmask = (1ull << Lw) - 1;
  if (dt == 0 && n > k) {
    const int k14 = std::min(n, 14);
    if (npasses(k14, lmask) < npasses(k, lmask)) k = k14;
  }
  std::vector<int> rem;
  for (int g = 0; g < (int)H->groups.size(); ++g) {
    if (__builtin_popcountll(H->groups[g].flip | lmask) > k) S.legacy.push_back(g);
    else rem.push_back(g);
  }
  int slot = 0;
  while (!rem.empty()) {
    uint64_t tm = (n <= k) ? all : lmask;
    std::vector<int> take, rest;
    for (int g : rem) {
      uint64_t f = H->groups[g].flip;
      if (__builtin_popcountll(tm | f) <= k) {
        tm |= f;
        take.push_back(g);
      } else {
        rest.push_back(g);
      }
    }
    for (int b = 0; b < n && __builtin_popcountll(tm) < k; ++b) tm |= 1ull << b;
    std::vector<int> tile;
    for (int b = 0; b < n; ++b)
      if ((tm >> b) & 1) tile.push_back(b);
    DPPass pp{};
    pp.k = (int)tile.size();
    int L = 0;
    while (L < pp.k && tile[L] == L) ++L;
    pp.L = L;
    pp.nblk_log = n - pp.k;
    int nb = 0;
    for (int b = 0; b < n; ++b)
      if (!((tm >> b) & 1)) pp.blk_bits[nb++] = b;
    pp.runoff0 = (int)S.runoff.size();
    for (int r = 0; r < (1 << (pp.k - L)); ++r) {
      uint32_t off = 0;
      for (int i = 0; i < pp.k - L; ++i)
        if ((r >> i) & 1) off |= 1u << tile[L + i];
      S.runoff.push_back(off);
    }
    pp.g0 = (int)S.groups.size();
    pp.ng = (int)take.size();
    // 16 elements per thread (up to 1024 threads)
    pp.nthr_log = std::min(10, pp.k - 4);
    pp.slot0 = slot;
    slot += 1 << pp.nblk_log;
    for (int g : take) {
      const DPGroup& G = H->groups[g];
      DPGroupT gt{};
      gt.flip_t = pext_bits(G.flip, tile);
      gt.t0 = (int)S.terms.size();
      gt.nt = G.nt;
      for (int t = G.t0; t < G.t0 + G.nt; ++t) {
        const DPTerm& tm0 = H->terms[t];
        DPTermT tt{};
        tt.yz_t = pext_bits(tm0.yz & tm, tile);
        tt.yz_g = tm0.yz & ~tm;
        tt.ny = tm0.ny;
        tt.w = tm0.w;
        S.terms.push_back(tt);
      }
      S.groups.push_back(gt);
    }
    S.passes.push_back(pp);
    rem.swap(rest);
  }
  for (int g : S.legacy) {
    S.legacy_slot0.push_back(slot);
    slot += H->groups[g].nblk;
  }
  S.nslots = slot;
}

// ---- adjoint ---------------------------------------------------------------
struct AdjWs {
  void *mtab_f, *mtab_a, *psi, *lam, *ckpt;
  double *part, *gsum, *epart, *e2;
  int64_t bytes;
};

AdjWs carve_adj(const Program& P, const tcm_pauli* H, int dtype, int64_t batch, int mode, void* base) {
  const Sched& F = P.sched[dtype][0];
  const Sched& A = P.sched[dtype][1];
  const int64_t N = 1ll << P.n;
  const int64_t ts = tsize(dtype);
  Carve cv(base);
  AdjWs w{};
  w.mtab_f = mode == 0 ? cv.take(batch * (int64_t)std::max(F.mtab_size, 1) * ts) : nullptr;
  w.mtab_a = cv.take(batch * (int64_t)std::max(A.mtab_size, 1) * ts);
  w.psi = cv.take(batch * N * 2 * ts);
  w.lam = cv.take(batch * N * 2 * ts);
  w.ckpt = mode == 0 ? cv.take((int64_t)P.ckpt_gate.size() * batch * N * 2 * ts) : nullptr;
  const int64_t ngp = std::max(A.max_ngop_pass, 1);
  w.part = (double*)cv.take(batch * ngp * (int64_t)A.max_slots * sizeof(double));
  w.gsum = (double*)cv.take(batch * (int64_t)std::max(P.ngop, 1) * sizeof(double));
  if (mode == 0 && H) {
    w.epart = (double*)cv.take(batch * (int64_t)pauli_slots(H, dtype) * 2 * sizeof(double));
    w.e2 = (double*)cv.take(batch * 2 * sizeof(double));
  } else {
    w.epart = w.e2 = nullptr;
  }
  w.bytes = cv.off;
  return w;
}

template <typename T>
int reverse_sweep(Program& P, int dtype, const AdjWs& w, const void* psi_src, const void* lam_src, const void* ckpts,
                  void* grad_input, int64_t batch, hipStream_t st) {
  DeviceTables* D = nullptr;
  if (int rc = ensure_tables(P, dtype, 1, &D)) return rc;
  const Sched& A = P.sched[dtype][1];
  const int n = P.n;
  const int64_t N = 1ll << n;
  const int64_t tot = batch * N;
  auto ckpt_ptr = [&](size_t c) { return (const void*)((const char*)ckpts + (int64_t)c * tot * 2 * sizeof(T)); };
  const int64_t ngp = std::max(A.max_ngop_pass, 1);
  const int64_t pstride = ngp * (int64_t)A.max_slots;
  const int np = (int)A.passes.size();
  // checkpoint restoring psi before pass p: the one taken after pass p-1
  auto ckpt_before = [&](int p) -> int {
    for (size_t c = 0; c < A.ckpt_after.size(); ++c)
      if (A.ckpt_after[c] == p - 1) return (int)c;
    return -1;
  };
  if (A.k == 0) {
    // plain: work in place on the workspace copies
    if (psi_src != w.psi) launch_copy<T>(w.psi, psi_src, tot, st);
    if (lam_src != w.lam) launch_copy<T>(w.lam, lam_src, tot, st);
    for (int p = np - 1; p >= 0; --p) {
      const DOp& op = A.dop[p];
      if (op.kind == OK_UK) {
        launch_plain_adj_uk<T>(op, (const int32_t*)D->ukbits, D->ctab, w.psi, w.lam, batch, n, st);
      } else {
        const int slots = op.gop >= 0 ? A.gop_slots[op.gop] : std::min<int64_t>(512, std::max<int64_t>(1, (N >> 2) / 256));
        launch_plain_adj<T>(op, w.mtab_a, A.mtab_size, D->ctab, w.psi, w.psi, w.lam, w.lam, batch, n, slots, w.part,
                            pstride, st);
        if (op.gop >= 0) launch_reduce_gop(w.part, pstride, slots, nullptr, op.gop, 1, w.gsum, P.ngop, batch, st);
      }
      int c = ckpt_before(p);
      if (c >= 0 && ckpts) launch_copy<T>(w.psi, ckpt_ptr(c), tot, st);
    }
    if (grad_input) launch_copy<T>(grad_input, w.lam, tot, st);
  } else {
    const void* pin = psi_src;
    const void* lin = lam_src;
    for (int p = np - 1; p >= 0; --p) {
      const DPass& dp = A.dpass[p];
      int c = ckpts ? ckpt_before(p) : -1;
      if (dp.plain_op >= 0) {
        // >=3-qubit constant gate: undo in place on the workspace copies
        if (pin != w.psi) launch_copy<T>(w.psi, pin, tot, st);
        if (lin != w.lam) launch_copy<T>(w.lam, lin, tot, st);
        launch_plain_adj_uk<T>(A.dop[dp.plain_op], (const int32_t*)D->ukbits, D->ctab, w.psi, w.lam, batch, n, st);
        if (p == 0 && grad_input) launch_copy<T>(grad_input, w.lam, tot, st);
        pin = c >= 0 ? ckpt_ptr(c) : w.psi;
        lin = w.lam;
        continue;
      }
      // psi after the last pass (p = 0) is not read again; a compiled pass
      // also skips lambda there unless it is the input gradient
      const bool jitp = !A.jit.empty() && A.jit[p].kernel >= 0;
      void* pout = (c >= 0 || p == 0) ? nullptr : w.psi;
      void* lout = p > 0 ? w.lam : grad_input ? grad_input : jitp ? nullptr : w.lam;
      if (jitp) {
        if (int rc = jit_launch_adj(A, *D, p, sizeof(T) == 8, w.mtab_a, D->ctab, pin, pout, lin, lout, batch, n,
                                    w.part, pstride, st))
          return rc;
      } else if (A.R == 4) {
        launch_fused_adj<T, 4>(A, *D, p, w.mtab_a, D->ctab, pin, pout, lin, lout, batch, n, w.part, pstride, st);
      } else {
        return set_error(TCM_E_UNSUPPORTED, "register slots R=%d not built", A.R);
      }
      if (dp.ngop > 0)
        launch_reduce_gop(w.part, pstride, 1ll << dp.nblk_log, (const int32_t*)D->gop_list + dp.gop_list0, 0,
                          dp.ngop, w.gsum, P.ngop, batch, st);
      pin = c >= 0 ? ckpt_ptr(c) : w.psi;
      lin = lout;
    }
    if (np == 0 && grad_input) launch_copy<T>(grad_input, lam_src, tot, st);
  }
  return 0;
}

template <typename T>
int energy_grad_impl(Program& P, int dtype, const void* params, int64_t pstride, tcm_pauli* H, const void* inputs,
                     int64_t batch, double* energy, double* grad, void* ws, int64_t ws_bytes, hipStream_t st) {
  AdjWs w = carve_adj(P, H, dtype, batch, 0, nullptr);
  if (w.bytes > ws_bytes) return set_error(TCM_E_WORKSPACE, "workspace %lld < %lld bytes", (long long)ws_bytes, (long long)w.bytes);
  w = carve_adj(P, H, dtype, batch, 0, ws);
  DeviceTables* DF = nullptr;
  DeviceTables* DA = nullptr;
  if (int rc = ensure_tables(P, dtype, 0, &DF)) return rc;
  if (int rc = ensure_tables(P, dtype, 1, &DA)) return rc;
  // a program built with the pass compiler compiles the Hamiltonian too
  if (P.opts.jit)
    if (int rc = pauli_jit_build(H, dtype, nullptr)) return rc;
  const Sched& F = P.sched[dtype][0];
  const Sched& A = P.sched[dtype][1];
  const int64_t N = 1ll << P.n;
  // forward with checkpoints
  launch_prep<T>(F, *DF, DF->ctab, params, pstride, w.mtab_f, batch, st);
  if (inputs) launch_copy<T>(w.psi, inputs, batch * N, st);
  if (int rc = run_forward_passes<T>(P, F, *DF, w.mtab_f, w.psi, batch, w.ckpt, st, inputs == nullptr)) return rc;
  // energy and lambda = H psi
  if (int rc = pauli_energy_apply<T>(H, dtype, w.psi, w.lam, nullptr, batch, w.epart, true, true, st)) return rc;
  launch_sum_slots(w.epart, (int64_t)pauli_slots(H, dtype) * 2, pauli_slots(H, dtype), 2, w.e2, batch, st);
  // reverse sweep
  launch_prep<T>(A, *DA, DA->ctab, params, pstride, w.mtab_a, batch, st);
  if (int rc = reverse_sweep<T>(P, dtype, w, w.psi, w.lam, w.ckpt, nullptr, batch, st)) return rc;
  launch_grad_assemble(w.gsum, P.ngop, (const int32_t*)DA->p_ptr, (const int32_t*)DA->p_gops,
                       (const double*)DA->gfac, 1.0, grad, P.nparams, batch, st);
  // energy real parts
  HIP_TRY(hipMemcpy2DAsync(energy, sizeof(double), w.e2, 2 * sizeof(double), sizeof(double), batch,
                           hipMemcpyDeviceToDevice, st),
          "energy copy");
  HIP_TRY(hipGetLastError(), "adjoint launch");
  return 0;
}

template <typename T>
int vjp_impl(Program& P, int dtype, const void* params, int64_t pstride, const void* ckpts, const void* state,
             const void* grad_state, int64_t batch, double* grad_params, void* grad_input, void* ws, int64_t ws_bytes,
             hipStream_t st) {
  AdjWs w = carve_adj(P, nullptr, dtype, batch, 1, nullptr);
  if (w.bytes > ws_bytes) return set_error(TCM_E_WORKSPACE, "workspace %lld < %lld bytes", (long long)ws_bytes, (long long)w.bytes);
  if (!P.ckpt_gate.empty() && !ckpts) return set_error(TCM_E_ARG, "program stores checkpoints: ckpts required");
  w = carve_adj(P, nullptr, dtype, batch, 1, ws);
  DeviceTables* DA = nullptr;
  if (int rc = ensure_tables(P, dtype, 1, &DA)) return rc;
  const Sched& A = P.sched[dtype][1];
  launch_prep<T>(A, *DA, DA->ctab, params, pstride, w.mtab_a, batch, st);
  if (int rc = reverse_sweep<T>(P, dtype, w, state, grad_state, ckpts, grad_input, batch, st)) return rc;
  if (grad_params)
    launch_grad_assemble(w.gsum, P.ngop, (const int32_t*)DA->p_ptr, (const int32_t*)DA->p_gops,
                         (const double*)DA->gfac, 0.5, grad_params, P.nparams, batch, st);
  HIP_TRY(hipGetLastError(), "vjp launch");
  return 0;
}

bool valid_dtype(int d) { return d == TCM_C64 || d == TCM_C128; }

// ---- batched uneven circuits ------------------------------------------------
int64_t multi_ws(const tcm_multi_desc* descs, int ndesc, int dtype) {
  int64_t total = 0;
  for (int i = 0; i < ndesc; ++i)
    total += align_up(descs[i].batch * (int64_t)std::max(descs[i].prog->p.sched[dtype][0].mtab_size, 1) *
                      (int64_t)tsize(dtype));
  return total;
}

template <typename T>
int forward_multi_impl(const tcm_multi_desc* descs, int ndesc, int dtype, int init_zero, void* ws, hipStream_t st) {
  std::vector<DeviceTables*> D(ndesc);
  std::vector<void*> mtab(ndesc);
  Carve cv(ws);
  int maxp = 0;
  for (int i = 0; i < ndesc; ++i) {
    Program& P = descs[i].prog->p;
    if (int rc = ensure_tables(P, dtype, 0, &D[i])) return rc;
    const Sched& S = P.sched[dtype][0];
    if (S.k > 0 && S.R != 4)
      return set_error(TCM_E_UNSUPPORTED, "multi-program launches run 4 register slots (program %d has %d)", i, S.R);
    mtab[i] = cv.take(descs[i].batch * (int64_t)std::max(S.mtab_size, 1) * sizeof(T));
    if (descs[i].batch <= 0) continue;
    launch_prep<T>(S, *D[i], D[i]->ctab, descs[i].params, descs[i].param_stride, mtab[i], descs[i].batch, st);
    if (init_zero) launch_init_zero<T>(descs[i].state, P.n, descs[i].batch, st);
    maxp = std::max(maxp, (int)S.passes.size());
  }
  for (int r = 0; r < maxp; ++r) {
    MultiJobs J{};
    int64_t blocks = 0;
    int threads = 64;
    size_t lds = 0;
    auto flush = [&]() {
      if (J.njobs == 0) return;
      launch_fused_fwd_multi<T>(J, blocks, threads, lds, st);
      J = MultiJobs{};
      blocks = 0;
      threads = 64;
      lds = 0;
    };
    for (int i = 0; i < ndesc; ++i) {
      const Program& P = descs[i].prog->p;
      const Sched& S = P.sched[dtype][0];
      if (descs[i].batch <= 0 || r >= (int)S.passes.size()) continue;
      if (S.k == 0 || S.dpass[r].plain_op >= 0) {
        const DOp& op = S.k > 0 ? S.dop[S.dpass[r].plain_op] : S.dop[r];
        launch_plain_gate<T, false>(op, (const int32_t*)D[i]->ukbits, mtab[i], S.mtab_size, D[i]->ctab,
                                    descs[i].state, P.n, descs[i].batch, st);
        continue;
      }
      if (S.R != 4) return set_error(TCM_E_UNSUPPORTED, "register slots R=%d not built", S.R);
      if (S.dpass[r].mfma) return set_error(TCM_E_UNSUPPORTED, "matrix-core passes are not run in multi launches");
      const DPass& dp = S.dpass[r];
      MultiJob& q = J.job[J.njobs++];
      q.dpass = D[i]->dpass;
      q.dgroup = D[i]->dgroup;
      q.dop = D[i]->dop;
      q.runoff = D[i]->runoff;
      q.ctab = D[i]->ctab;
      q.mtab = mtab[i];
      q.state = descs[i].state;
      q.mstride = S.mtab_size;
      q.sstride = 1ll << P.n;
      q.block0 = blocks;
      q.pass = r;
      blocks += descs[i].batch << dp.nblk_log;
      threads = std::max(threads, 1 << dp.tthreads_log);
      lds = std::max(lds, sizeof(V2<T>) << dp.k);
      if (J.njobs == kMaxMultiJobs) flush();
    }
    flush();
  }
  return 0;
}

// kernel argument of tcm_multi_k<kid>: one job per program of the launch
// (the layout of TcmMJob in the generated source)
struct MJobArg {
  const void* state;
  long long sstride;
  const void* mtab;
  long long mstride, mbase, dbase;
  const void* ctab;
  unsigned long long block0, ntiles;
};
constexpr int kMaxPlanPrograms = 48;  // 48 x 72-byte jobs fit the kernel-argument segment

int plan_load(tcm_multi_plan* M, std::vector<void*>** fns) {
  int dev = 0;
  if (int rc = current_device(&dev)) return rc;
  std::lock_guard<std::mutex> lk(M->mu);
  if ((int)M->fns.size() <= dev) {
    M->fns.resize(dev + 1);
    M->mods.resize(dev + 1);
  }
  if (M->fns[dev].empty() && !M->chunk_of.empty())
    if (int rc = load_modules(M->code, M->chunk_of, "tcm_multi_k", M->mods[dev], M->fns[dev])) return rc;
  *fns = &M->fns[dev];
  return 0;
}

template <typename T>
int forward_plan_impl(tcm_multi_plan* M, const tcm_multi_desc* descs, int ndesc, int dtype, int init_zero, void* ws,
                      hipStream_t st) {
  std::vector<void*>* fns = nullptr;
  if (int rc = plan_load(M, &fns)) return rc;
  std::vector<DeviceTables*> D(ndesc);
  std::vector<void*> mtab(ndesc);
  Carve cv(ws);
  for (int i = 0; i < ndesc; ++i) {
    Program& P = descs[i].prog->p;
    if (int rc = ensure_tables(P, dtype, 0, &D[i])) return rc;
    const Sched& S = P.sched[dtype][0];
    mtab[i] = cv.take(descs[i].batch * (int64_t)std::max(S.mtab_size, 1) * sizeof(T));
    if (descs[i].batch <= 0) continue;
    launch_prep<T>(S, *D[i], D[i]->ctab, descs[i].params, descs[i].param_stride, mtab[i], descs[i].batch, st);
    if (init_zero) launch_init_zero<T>(descs[i].state, P.n, descs[i].batch, st);
  }
  for (int r = 0; r < M->maxp; ++r) {
    std::vector<char> in_kernel(ndesc, 0);
    if (M->kid[r] >= 0) {
      const tcm::MultiPassInfo& I = M->info[r];
      std::vector<MJobArg> jobs(I.jobs.size());
      unsigned long long blocks = 0;
      for (size_t q = 0; q < I.jobs.size(); ++q) {
        const tcm::MultiPassJob& J = I.jobs[q];
        const tcm_multi_desc& d = descs[J.prog];
        const Sched& S = d.prog->p.sched[dtype][0];
        MJobArg& a = jobs[q];
        a.state = d.state;
        a.sstride = 1ll << d.prog->p.n;
        a.mtab = mtab[J.prog];
        a.mstride = S.mtab_size;
        a.mbase = J.mbase;
        a.dbase = J.dbase;
        a.ctab = D[J.prog]->ctab;
        a.block0 = blocks;
        a.ntiles = (unsigned long long)std::max<int64_t>(d.batch, 0) << J.nblk_log;
        // an empty batch keeps one (clamped) block out of the launch: its
        // job gets no blocks and the search never selects it
        blocks += (a.ntiles + J.per_block - 1) / J.per_block;
        in_kernel[J.prog] = d.batch > 0;
      }
      if (blocks > 0) {
        // the kernel's argument is the fixed-size job array, by value
        void* args[] = {jobs.data()};
        size_t asz = jobs.size() * sizeof(MJobArg);
        void* cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, jobs.data(), HIP_LAUNCH_PARAM_BUFFER_SIZE, &asz,
                       HIP_LAUNCH_PARAM_END};
        (void)args;
        const hipError_t e = hipModuleLaunchKernel((hipFunction_t)(*fns)[M->kid[r]], (unsigned)blocks, 1, 1,
                                                   (unsigned)I.ntb, 1, 1, 0, st, nullptr, cfg);
        if (e != hipSuccess) return set_error((int)e, "hipModuleLaunchKernel(multi): %s", hipGetErrorString(e));
      }
    }
    for (int i = 0; i < ndesc; ++i) {
      Program& P = descs[i].prog->p;
      const Sched& S = P.sched[dtype][0];
      if (in_kernel[i] || descs[i].batch <= 0 || r >= (int)S.passes.size()) continue;
      if (int rc = run_one_pass<T>(P, S, *D[i], r, mtab[i], descs[i].state, descs[i].batch, nullptr, st)) return rc;
    }
  }
  return 0;
}

}  // namespace

// ===========================================================================
// ---- expectation with caller-provided scratch, no allocation ---------------
namespace tcm {
// The launches of tcm_expect_pauli_sum with `scratch` (device, scratch_bytes)
// as the partial-sum buffer, for a Pauli sum whose device tables (and
// compiled passes) already exist on the current device; never uploads,
// loads or allocates.
// argument and size checks of tcm_expect_pauli_sum_noalloc (host only, made
// before anything touches the device)
int pauli_expect_check(const tcm_pauli* ham, int32_t dtype, const void* state, int64_t batch, double* out,
                       int64_t scratch_bytes) {
  if (!ham || !valid_dtype(dtype) || batch < 0 || (batch > 0 && (!state || !out)))
    return set_error(TCM_E_ARG, "bad arguments to tcm_expect_pauli_sum_noalloc");
  const int64_t need = pauli_ws(ham, dtype, batch);
  if (need > scratch_bytes)
    return set_error(TCM_E_WORKSPACE,
                     "tcm_expect_pauli_sum_noalloc: %lld states need %lld bytes of partial sums, the static scratch "
                     "holds %lld; split the batch",
                     (long long)batch, (long long)need, (long long)scratch_bytes);
  return 0;
}

int pauli_expect_prepared(const tcm_pauli* ham, int32_t dtype, const void* state, int64_t batch, double* out,
                          void* scratch, int64_t scratch_bytes, void* stream) {
  if (int rc = pauli_expect_check(ham, dtype, state, batch, out, scratch_bytes)) return rc;
  if (!scratch) return set_error(TCM_E_ARG, "tcm_expect_pauli_sum_noalloc: no scratch");
  if (batch == 0) return 0;
  int dev = 0;
  if (int rc = current_device(&dev)) return rc;
  tcm_pauli* H = const_cast<tcm_pauli*>(ham);
  PauliDev* pd = nullptr;
  {
    std::lock_guard<std::mutex> lk(H->mu);
    if ((int)H->dev.size() > dev) {
      PauliDev& d = H->dev[dev];
      const bool tables = d.terms && d.groups && (!H->ts[dtype].ok || d.tpasses[dtype]);
      const bool kernels = !(H->jit[dtype].state == 1 && !H->jit[dtype].chunk.empty() && d.jfns[dtype].empty());
      if (tables && kernels) pd = &d;
    }
  }
  if (!pd)
    return set_error(TCM_E_ARG,
                     "tcm_expect_pauli_sum_noalloc: the Pauli sum is not prepared on device %d (make one "
                     "tcm_expect_pauli_sum call with it for this dtype first)",
                     dev);
  hipStream_t st = (hipStream_t)stream;
  double* epart = (double*)scratch;
  int rc = dtype == TCM_C64 ? pauli_energy_apply<float>(H, dtype, state, nullptr, nullptr, batch, epart, true, false, st, pd)
                            : pauli_energy_apply<double>(H, dtype, state, nullptr, nullptr, batch, epart, true, false, st, pd);
  if (rc) return rc;
  const int ns = pauli_slots(H, dtype);
  launch_sum_slots(epart, (int64_t)ns * 2, ns, 2, out, batch, st);
  HIP_TRY(hipGetLastError(), "expect launch");
  return 0;
}
}  // namespace tcm

extern "C" {

int tcm_abi_version(void) { return TCM_ABI_VERSION; }

const char* tcm_last_error(void) { return g_err.c_str(); }

int tcm_program_build(const tcm_gate* gates, int32_t ngates, int32_t nqubits, int32_t nparams, const tcm_opts* opts,
                      tcm_program** out) {
  if (!out) return set_error(TCM_E_ARG, "out is NULL");
  *out = nullptr;
  Program* P = nullptr;
  int rc = build_program(gates, ngates, nqubits, nparams, opts, &P);
  if (rc) return rc;
  tcm_program* tp = new tcm_program();
  // move-construct the program into the handle
  tp->p.n = P->n;
  tp->p.nparams = P->nparams;
  tp->p.opts = P->opts;
  tp->p.gates = std::move(P->gates);
  tp->p.nlayers = P->nlayers;
  tp->p.ngop = P->ngop;
  tp->p.gop_gate = std::move(P->gop_gate);
  tp->p.p_ptr = std::move(P->p_ptr);
  tp->p.p_gops = std::move(P->p_gops);
  tp->p.gfac = std::move(P->gfac);
  tp->p.ctab = std::move(P->ctab);
  tp->p.ckpt_gate = std::move(P->ckpt_gate);
  for (int d = 0; d < 2; ++d)
    for (int m = 0; m < 2; ++m) tp->p.sched[d][m] = std::move(P->sched[d][m]);
  delete P;
  *out = tp;
  return 0;
}

int tcm_program_free(tcm_program* prog) {
  if (!prog) return 0;
  int cur = -1;
  bool have = hipGetDevice(&cur) == hipSuccess;
  for (int d = 0; d < 2; ++d)
    for (int m = 0; m < 2; ++m) {
      auto& v = prog->p.sched[d][m].dev;
      for (size_t dev = 0; dev < v.size(); ++dev)
        if (v[dev]) {
          (void)hipSetDevice((int)dev);
          free_tables(v[dev]);
          v[dev] = nullptr;
        }
    }
  if (have) (void)hipSetDevice(cur);
  (void)hipGetLastError();
  delete prog;
  return 0;
}

int tcm_program_get_info(const tcm_program* prog, tcm_program_info* info) {
  if (!prog || !info) return set_error(TCM_E_ARG, "NULL argument");
  std::memset(info, 0, sizeof *info);
  const Program& P = prog->p;
  info->nqubits = P.n;
  info->nparams = P.nparams;
  info->ngates = (int)P.gates.size();
  info->npasses = (int)P.sched[0][0].passes.size();
  info->nckpt = (int)P.ckpt_gate.size();
  info->tile_bits = P.sched[0][0].k;
  return 0;
}

int tcm_program_compile(tcm_program* prog, int32_t dtype, int32_t mode, double* seconds) {
  if (!prog || !valid_dtype(dtype) || (mode != 0 && mode != 1)) return set_error(TCM_E_ARG, "bad compile args");
  Program& P = prog->p;
  std::lock_guard<std::mutex> lk(P.mu);
  Sched& S = P.sched[dtype][mode];
  if (!P.opts.jit) return set_error(TCM_E_UNSUPPORTED, "program built without opts.jit");
  if (S.k > 0 && S.jit_state == 0)
    if (int rc = jit_build(P, S, dtype, mode)) return rc;
  if (S.jit_state < 0) return set_error(TCM_E_UNSUPPORTED, "pass compiler failed for this program");
  if (seconds) *seconds = S.jit_seconds;
  return 0;
}

int64_t tcm_program_describe(const tcm_program* prog, int32_t dtype, int32_t mode, char* buf, int64_t cap) {
  if (!prog || !valid_dtype(dtype) || (mode != 0 && mode != 1)) return set_error(TCM_E_ARG, "bad describe args");
  std::string s = describe_sched(prog->p, dtype, mode);
  if (buf && cap > 0) {
    int64_t m = std::min<int64_t>(cap - 1, (int64_t)s.size());
    std::memcpy(buf, s.data(), m);
    buf[m] = 0;
  }
  return (int64_t)s.size();
}

int tcm_forward_workspace_bytes(const tcm_program* prog, int32_t dtype, int64_t batch, int64_t* bytes) {
  if (!prog || !bytes || !valid_dtype(dtype) || batch < 0) return set_error(TCM_E_ARG, "bad arguments");
  *bytes = forward_ws(prog->p, dtype, batch);
  return 0;
}

int tcm_forward(tcm_program* prog, int32_t dtype, const void* params, int64_t param_stride, void* state,
                int64_t batch, int32_t init_zero, void* ckpts, void* workspace, int64_t ws_bytes, void* stream) {
  if (!prog || !valid_dtype(dtype) || batch < 0 || (batch > 0 && !state))
    return set_error(TCM_E_ARG, "bad arguments to tcm_forward");
  if (batch == 0) return 0;
  if (prog->p.nparams > 0 && !params) return set_error(TCM_E_ARG, "params required");
  if (param_stride < 0) return set_error(TCM_E_ARG, "param_stride < 0");
  hipStream_t st = (hipStream_t)stream;
  return dtype == TCM_C64
             ? forward_impl<float>(prog->p, dtype, params, param_stride, state, batch, init_zero, ckpts, workspace, ws_bytes, st)
             : forward_impl<double>(prog->p, dtype, params, param_stride, state, batch, init_zero, ckpts, workspace, ws_bytes, st);
}

int tcm_pauli_build(const int8_t* codes, const double* coeffs, int32_t nterms, int32_t nqubits, tcm_pauli** out) {
  if (!out) return set_error(TCM_E_ARG, "out is NULL");
  *out = nullptr;
  if (nqubits < 1 || nqubits > 32) return set_error(TCM_E_ARG, "nqubits=%d out of range", nqubits);
  if (nterms < 1 || !codes || !coeffs) return set_error(TCM_E_ARG, "need at least one term");
  std::map<uint64_t, std::vector<DPTerm>> by_flip;
  std::vector<DPTermF> flat;
  for (int t = 0; t < nterms; ++t) {
    if (!std::isfinite(coeffs[t])) return set_error(TCM_E_NONFINITE, "term %d: non-finite coefficient", t);
    uint64_t flip = 0, yz = 0;
    int ny = 0;
    for (int q = 0; q < nqubits; ++q) {
      int c = codes[(int64_t)t * nqubits + q];
      if (c < 0 || c > 3) return set_error(TCM_E_ARG, "term %d: code %d at qubit %d not in 0..3", t, c, q);
      uint64_t bit = 1ull << (nqubits - 1 - q);
      if (c == 1) flip |= bit;
      if (c == 2) { flip |= bit; yz |= bit; ++ny; }
      if (c == 3) yz |= bit;
    }
    DPTerm tm{};
    tm.yz = yz;
    tm.ny = ny;
    tm.w = coeffs[t];
    by_flip[flip].push_back(tm);
    DPTermF tf{};
    tf.flip = flip;
    tf.yz = yz;
    tf.ny = ny;
    tf.w = coeffs[t];
    flat.push_back(tf);
  }
  tcm_pauli* H = new tcm_pauli();
  H->n = nqubits;
  H->flat = std::move(flat);
  int slot = 0;
  for (auto& kv : by_flip) {  // std::map: flip 0 (diagonal) first
    DPGroup g{};
    g.flip = kv.first;
    g.t0 = (int)H->terms.size();
    g.nt = (int)kv.second.size();
    g.nblk = pauli_blocks(nqubits, kv.first != 0);
    g.slot0 = slot;
    slot += g.nblk;
    H->terms.insert(H->terms.end(), kv.second.begin(), kv.second.end());
    H->groups.push_back(g);
  }
  H->nslots = slot;
  pauli_tile_sched(H, 0);
  pauli_tile_sched(H, 1);
  *out = H;
  return 0;
}

int tcm_pauli_free(tcm_pauli* ham) {
  if (!ham) return 0;
  int cur = -1;
  bool have = hipGetDevice(&cur) == hipSuccess;
  for (size_t d = 0; d < ham->dev.size(); ++d) {
    PauliDev& pd = ham->dev[d];
    (void)hipSetDevice((int)d);
    if (pd.terms) (void)hipFree(pd.terms);
    if (pd.flat) (void)hipFree(pd.flat);
    if (pd.groups) (void)hipFree(pd.groups);
    for (int dt = 0; dt < 2; ++dt) {
      void* ps[] = {pd.tpasses[dt], pd.tgroups[dt], pd.tterms[dt], pd.truns[dt]};
      for (void* q : ps)
        if (q) (void)hipFree(q);
      for (void* m : pd.jmods[dt])
        if (m) (void)hipModuleUnload((hipModule_t)m);
    }
  }
  if (have) (void)hipSetDevice(cur);
  (void)hipGetLastError();
  delete ham;
  return 0;
}

int tcm_pauli_compile(tcm_pauli* ham, int32_t dtype, double* seconds) {
  if (!ham || !valid_dtype(dtype)) return set_error(TCM_E_ARG, "bad compile args");
  return pauli_jit_build(ham, dtype, seconds);
}

int tcm_pauli_workspace_bytes(const tcm_pauli* ham, int32_t dtype, int64_t batch, int64_t* bytes) {
  if (!ham || !bytes || !valid_dtype(dtype) || batch < 0) return set_error(TCM_E_ARG, "bad arguments");
  *bytes = pauli_ws(ham, dtype, batch);
  return 0;
}

int tcm_expect_pauli_sum(const tcm_pauli* ham, int32_t dtype, const void* state, int64_t batch, double* out,
                         void* workspace, int64_t ws_bytes, void* stream) {
  if (!ham || !valid_dtype(dtype) || batch < 0 || (batch > 0 && (!state || !out)))
    return set_error(TCM_E_ARG, "bad arguments to tcm_expect_pauli_sum");
  if (batch == 0) return 0;
  if (pauli_ws(ham, dtype, batch) > ws_bytes) return set_error(TCM_E_WORKSPACE, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  tcm_pauli* H = const_cast<tcm_pauli*>(ham);
  double* epart = (double*)workspace;
  int rc = dtype == TCM_C64 ? pauli_energy_apply<float>(H, dtype, state, nullptr, nullptr, batch, epart, true, false, st)
                            : pauli_energy_apply<double>(H, dtype, state, nullptr, nullptr, batch, epart, true, false, st);
  if (rc) return rc;
  const int ns = pauli_slots(H, dtype);
  launch_sum_slots(epart, (int64_t)ns * 2, ns, 2, out, batch, st);
  HIP_TRY(hipGetLastError(), "expect launch");
  return 0;
}

// ---- measurement ----------------------------------------------------------
namespace {
// chunk prefix sums of the sampler: batch x (2^(n - min(n, 10)) + 1) doubles
int64_t sample_ws(int n, int64_t batch) {
  const int cl = n < 10 ? n : 10;
  return batch * (((int64_t)1 << (n - cl)) + 1) * (int64_t)sizeof(double);
}
// blocks of the measurement partial sums and workspace: fixval (u64 per
// batch item), then part[batch][kMeasBlocks][2]
constexpr int kMeasBlocks = 1024;
int64_t measure_ws(int64_t batch) {
  return batch * (int64_t)sizeof(uint64_t) + batch * kMeasBlocks * 2 * (int64_t)sizeof(double);
}
}  // namespace

int tcm_probability(int32_t n, int32_t dtype, const void* state, int64_t batch, void* out, void* stream) {
  if (n < 1 || n > 40 || !valid_dtype(dtype) || batch < 0 || (batch > 0 && (!state || !out)))
    return set_error(TCM_E_ARG, "bad arguments to tcm_probability");
  if (batch == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const int64_t total = batch << n;
  if (dtype == TCM_C64) launch_probability<float>(state, out, total, st);
  else launch_probability<double>(state, out, total, st);
  HIP_TRY(hipGetLastError(), "probability launch");
  return 0;
}

int tcm_sample_workspace_bytes(int32_t n, int64_t batch, int64_t* bytes) {
  if (!bytes || n < 1 || n > 40 || batch < 0) return set_error(TCM_E_ARG, "bad arguments to tcm_sample_workspace_bytes");
  *bytes = sample_ws(n, batch);
  return 0;
}

namespace {
int sample_impl(const char* what, bool prob, int32_t n, int32_t dtype, const void* state, int64_t batch,
                const double* uniforms, int64_t shots, int64_t* out, void* workspace, int64_t ws_bytes,
                void* stream) {
  if (n < 1 || n > 40 || !valid_dtype(dtype) || batch < 0 || shots < 0 ||
      (batch > 0 && shots > 0 && (!state || !uniforms || !out || !workspace)))
    return set_error(TCM_E_ARG, "bad arguments to %s", what);
  if (batch == 0 || shots == 0) return 0;
  if (batch > 65535 || (shots + 3) / 4 > 0x7fffffffll || ((int64_t)1 << (n - (n < 10 ? n : 10))) > 0x7fffffffll)
    return set_error(TCM_E_ARG, "%s: batch, shots or state too large for one launch", what);
  if (ws_bytes < sample_ws(n, batch)) return set_error(TCM_E_WORKSPACE, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TCM_C64) launch_sample<float>(state, prob, n, batch, uniforms, shots, out, (double*)workspace, st);
  else launch_sample<double>(state, prob, n, batch, uniforms, shots, out, (double*)workspace, st);
  HIP_TRY(hipGetLastError(), "sample launch");
  return 0;
}
}  // namespace

int tcm_sample(int32_t n, int32_t dtype, const void* state, int64_t batch, const double* uniforms, int64_t shots,
               int64_t* out, void* workspace, int64_t ws_bytes, void* stream) {
  return sample_impl("tcm_sample", false, n, dtype, state, batch, uniforms, shots, out, workspace, ws_bytes, stream);
}

int tcm_sample_prob(int32_t n, int32_t dtype, const void* p, int64_t batch, const double* uniforms, int64_t shots,
                    int64_t* out, void* workspace, int64_t ws_bytes, void* stream) {
  return sample_impl("tcm_sample_prob", true, n, dtype, p, batch, uniforms, shots, out, workspace, ws_bytes, stream);
}

int tcm_measure_workspace_bytes(int32_t n, int64_t batch, int64_t* bytes) {
  if (!bytes || n < 1 || n > 40 || batch < 0) return set_error(TCM_E_ARG, "bad arguments to tcm_measure_workspace_bytes");
  *bytes = measure_ws(batch);
  return 0;
}

int tcm_measure(int32_t n, int32_t dtype, const void* state, int64_t batch, const int32_t* qubits, int32_t nmeas,
                const double* status, double* outcomes, double* prob, void* workspace, int64_t ws_bytes,
                void* stream) {
  if (n < 1 || n > 40 || !valid_dtype(dtype) || batch < 0 || nmeas < 0 || nmeas > n ||
      (nmeas > 0 && !qubits) || (batch > 0 && nmeas > 0 && (!state || !status || !outcomes || !prob || !workspace)))
    return set_error(TCM_E_ARG, "bad arguments to tcm_measure");
  uint64_t seen = 0;
  for (int k = 0; k < nmeas; ++k) {
    if (qubits[k] < 0 || qubits[k] >= n) return set_error(TCM_E_ARG, "tcm_measure: qubit %d out of range", qubits[k]);
    const uint64_t bitm = 1ull << (n - 1 - qubits[k]);
    if (seen & bitm) return set_error(TCM_E_ARG, "tcm_measure: qubit %d repeated", qubits[k]);
    seen |= bitm;
  }
  if (batch == 0 || nmeas == 0) return 0;
  if (batch > 65535) return set_error(TCM_E_ARG, "tcm_measure: batch too large for one launch");
  if (ws_bytes < measure_ws(batch)) return set_error(TCM_E_WORKSPACE, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  uint64_t* fixval = (uint64_t*)workspace;
  double* part = (double*)((char*)workspace + batch * (int64_t)sizeof(uint64_t));
  uint64_t fixed = 0;
  for (int k = 0; k < nmeas; ++k) {
    MeasureStep M{};
    M.bit = n - 1 - qubits[k];
    M.first = k == 0 ? 1 : 0;
    // runs of consecutive free index bits, counter bits in ascending order
    int src = 0;
    for (int b = 0; b < n;) {
      if (fixed >> b & 1) {
        ++b;
        continue;
      }
      int len = 0;
      while (b + len < n && !(fixed >> (b + len) & 1)) ++len;
      M.run[M.nrun++] = src | len << 8 | b << 16;
      src += len;
      b += len;
    }
    M.nfree = src;
    const int64_t work = (int64_t)1 << M.nfree;
    const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(kMeasBlocks, work / 2048));
    if (dtype == TCM_C64)
      launch_measure_step<float>(state, n, batch, M, k, nmeas, status, fixval, prob, outcomes, part, nblk, st);
    else
      launch_measure_step<double>(state, n, batch, M, k, nmeas, status, fixval, prob, outcomes, part, nblk, st);
    fixed |= 1ull << M.bit;
  }
  HIP_TRY(hipGetLastError(), "measure launch");
  return 0;
}

int tcm_pauli_nnz(const tcm_pauli* ham, int64_t* nnz) {
  if (!ham || !nnz) return set_error(TCM_E_ARG, "bad arguments to tcm_pauli_nnz");
  *nnz = (int64_t)ham->terms.size() << ham->n;
  return 0;
}

int tcm_pauli_coo(const tcm_pauli* ham, int32_t dtype, int64_t* rows, int64_t* cols, void* vals, void* stream) {
  if (!ham || !valid_dtype(dtype) || !rows || !cols || !vals)
    return set_error(TCM_E_ARG, "bad arguments to tcm_pauli_coo");
  if (ham->terms.empty()) return 0;
  if (ham->n > 34) return set_error(TCM_E_ARG, "tcm_pauli_coo: 2^%d rows per term is too many entries", ham->n);
  tcm_pauli* H = const_cast<tcm_pauli*>(ham);
  PauliDev* pd = nullptr;
  if (int rc = ensure_pauli(H, dtype, &pd)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t nt = (int64_t)H->terms.size();
  if (dtype == TCM_C64)
    launch_pauli_coo<float>((const DPGroup*)pd->groups, (int)H->groups.size(), (const DPTerm*)pd->terms, H->n, nt,
                            rows, cols, vals, st);
  else
    launch_pauli_coo<double>((const DPGroup*)pd->groups, (int)H->groups.size(), (const DPTerm*)pd->terms, H->n, nt,
                             rows, cols, vals, st);
  HIP_TRY(hipGetLastError(), "pauli_coo launch");
  return 0;
}

namespace {
constexpr int kCsrBlocks = 1024;
}

int tcm_csr_expect_workspace_bytes(int32_t n, int64_t* bytes) {
  if (!bytes || n < 1 || n > 40) return set_error(TCM_E_ARG, "bad arguments to tcm_csr_expect_workspace_bytes");
  *bytes = 2 * kCsrBlocks * (int64_t)sizeof(double);
  return 0;
}

int tcm_csr_expect(int32_t n, int32_t dtype, const int64_t* rowptr, const int64_t* cols, const void* vals,
                   const void* state, double* out, void* workspace, int64_t ws_bytes, void* stream) {
  if (n < 1 || n > 40 || !valid_dtype(dtype) || !rowptr || !cols || !vals || !state || !out || !workspace)
    return set_error(TCM_E_ARG, "bad arguments to tcm_csr_expect");
  if (ws_bytes < 2 * kCsrBlocks * (int64_t)sizeof(double)) return set_error(TCM_E_WORKSPACE, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nrows = (int64_t)1 << n;
  const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(kCsrBlocks, nrows / 256));
  if (dtype == TCM_C64)
    launch_csr_expect<float>(rowptr, cols, vals, state, nrows, nblk, (double*)workspace, out, st);
  else
    launch_csr_expect<double>(rowptr, cols, vals, state, nrows, nblk, (double*)workspace, out, st);
  HIP_TRY(hipGetLastError(), "csr_expect launch");
  return 0;
}

int tcm_apply_pauli_sum(const tcm_pauli* ham, int32_t dtype, const void* state, void* out_state, int64_t batch,
                        const double* alpha, void* stream) {
  if (!ham || !valid_dtype(dtype) || batch < 0 || (batch > 0 && (!state || !out_state)) || state == out_state)
    return set_error(TCM_E_ARG, "bad arguments to tcm_apply_pauli_sum");
  if (batch == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  tcm_pauli* H = const_cast<tcm_pauli*>(ham);
  int rc = dtype == TCM_C64 ? pauli_energy_apply<float>(H, dtype, state, out_state, alpha, batch, nullptr, false, true, st)
                            : pauli_energy_apply<double>(H, dtype, state, out_state, alpha, batch, nullptr, false, true, st);
  if (rc) return rc;
  HIP_TRY(hipGetLastError(), "apply launch");
  return 0;
}

// ---- per-term Pauli expectations / weighted applications ------------------
namespace {
int pauli_terms_blocks(int n) { return pauli_blocks(n, false); }
int64_t pauli_terms_ws(const tcm_pauli* H, int64_t batch) {
  return align_up(batch * (int64_t)H->flat.size() * pauli_terms_blocks(H->n) * 2 * (int64_t)sizeof(double));
}
}  // namespace

int tcm_pauli_terms_workspace_bytes(const tcm_pauli* ham, int64_t batch, int64_t* bytes) {
  if (!ham || !bytes || batch < 0) return set_error(TCM_E_ARG, "bad arguments to tcm_pauli_terms_workspace_bytes");
  *bytes = pauli_terms_ws(ham, batch);
  return 0;
}

int tcm_expect_pauli_terms(const tcm_pauli* ham, int32_t dtype, const void* state, int64_t batch, double* out,
                           void* workspace, int64_t ws_bytes, void* stream) {
  if (!ham || !valid_dtype(dtype) || batch < 0 || (batch > 0 && (!state || !out)))
    return set_error(TCM_E_ARG, "bad arguments to tcm_expect_pauli_terms");
  if (batch == 0) return 0;
  const int64_t nt = (int64_t)ham->flat.size();
  if (nt > 65535 || batch > 65535)
    return set_error(TCM_E_ARG, "tcm_expect_pauli_terms: %lld terms x %lld states exceeds 65535 per axis",
                     (long long)nt, (long long)batch);
  if (!workspace || pauli_terms_ws(ham, batch) > ws_bytes) return set_error(TCM_E_WORKSPACE, "workspace too small");
  tcm_pauli* H = const_cast<tcm_pauli*>(ham);
  PauliDev* pd = nullptr;
  if (int rc = ensure_pauli(H, dtype, &pd)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int nblk = pauli_terms_blocks(H->n);
  if (dtype == TCM_C64)
    launch_pauli_terms<float>((const DPTermF*)pd->flat, (int)nt, H->n, state, batch, nblk, (double*)workspace, out, st);
  else
    launch_pauli_terms<double>((const DPTermF*)pd->flat, (int)nt, H->n, state, batch, nblk, (double*)workspace, out,
                               st);
  HIP_TRY(hipGetLastError(), "pauli_terms launch");
  return 0;
}

int tcm_apply_pauli_terms(const tcm_pauli* ham, int32_t dtype, const void* state, void* out_state, int64_t batch,
                          const double* weights, void* stream) {
  if (!ham || !valid_dtype(dtype) || batch < 0 || (batch > 0 && (!state || !out_state || !weights)) ||
      state == out_state)
    return set_error(TCM_E_ARG, "bad arguments to tcm_apply_pauli_terms");
  if (batch == 0) return 0;
  tcm_pauli* H = const_cast<tcm_pauli*>(ham);
  PauliDev* pd = nullptr;
  if (int rc = ensure_pauli(H, dtype, &pd)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int nt = (int)H->flat.size();
  if (dtype == TCM_C64)
    launch_pauli_terms_apply<float>((const DPTermF*)pd->flat, nt, H->n, state, out_state, weights, batch, st);
  else
    launch_pauli_terms_apply<double>((const DPTermF*)pd->flat, nt, H->n, state, out_state, weights, batch, st);
  HIP_TRY(hipGetLastError(), "pauli_terms_apply launch");
  return 0;
}

int tcm_csr_matvec(int32_t n, int32_t dtype, const int64_t* rowptr, const int64_t* cols, const void* vals,
                   const void* x, void* y, int64_t batch, void* stream) {
  if (n < 1 || n > 40 || !valid_dtype(dtype) || batch < 0 ||
      (batch > 0 && (!rowptr || !cols || !vals || !x || !y)) || x == y)
    return set_error(TCM_E_ARG, "bad arguments to tcm_csr_matvec");
  if (batch == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const int64_t nrows = (int64_t)1 << n;
  if (dtype == TCM_C64)
    launch_csr_matvec<float>(rowptr, cols, vals, x, y, nrows, batch, st);
  else
    launch_csr_matvec<double>(rowptr, cols, vals, x, y, nrows, batch, st);
  HIP_TRY(hipGetLastError(), "csr_matvec launch");
  return 0;
}

int tcm_inner_workspace_bytes(int32_t nqubits, int64_t batch, int64_t* bytes) {
  if (!bytes || nqubits < 1 || nqubits > 32 || batch < 0) return set_error(TCM_E_ARG, "bad arguments");
  *bytes = align_up(batch * (int64_t)pauli_blocks(nqubits, false) * 2 * sizeof(double));
  return 0;
}

int tcm_inner(int32_t nqubits, int32_t dtype, const void* a, const void* b, int64_t batch, double* out,
              void* workspace, int64_t ws_bytes, void* stream) {
  if (!valid_dtype(dtype) || nqubits < 1 || nqubits > 32 || batch < 0 || (batch > 0 && (!a || !b || !out)))
    return set_error(TCM_E_ARG, "bad arguments to tcm_inner");
  if (batch == 0) return 0;
  const int nblk = pauli_blocks(nqubits, false);
  if (align_up(batch * (int64_t)nblk * 2 * sizeof(double)) > ws_bytes) return set_error(TCM_E_WORKSPACE, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  double* part = (double*)workspace;
  if (dtype == TCM_C64) launch_inner<float>(a, b, nqubits, batch, nblk, part, (int64_t)nblk * 2, st);
  else launch_inner<double>(a, b, nqubits, batch, nblk, part, (int64_t)nblk * 2, st);
  launch_sum_slots(part, (int64_t)nblk * 2, nblk, 2, out, batch, st);
  HIP_TRY(hipGetLastError(), "inner launch");
  return 0;
}

int tcm_adjoint_workspace_bytes(const tcm_program* prog, const tcm_pauli* ham, int32_t dtype, int64_t batch,
                                int32_t mode, int64_t* bytes) {
  if (!prog || !bytes || !valid_dtype(dtype) || batch < 0 || (mode != 0 && mode != 1) || (mode == 0 && !ham))
    return set_error(TCM_E_ARG, "bad arguments");
  *bytes = carve_adj(prog->p, ham, dtype, batch, mode, nullptr).bytes;
  return 0;
}

int tcm_adjoint_energy_grad(tcm_program* prog, int32_t dtype, const void* params, int64_t param_stride,
                            const tcm_pauli* ham, const void* inputs, int64_t batch, double* energy, double* grad,
                            void* workspace, int64_t ws_bytes, void* stream) {
  if (!prog || !ham || !valid_dtype(dtype) || batch < 0 || (batch > 0 && (!energy || (!grad && prog->p.nparams > 0))))
    return set_error(TCM_E_ARG, "bad arguments to tcm_adjoint_energy_grad");
  if (ham->n != prog->p.n) return set_error(TCM_E_ARG, "Hamiltonian on %d qubits, program on %d", ham->n, prog->p.n);
  if (batch == 0) return 0;
  if (prog->p.nparams > 0 && !params) return set_error(TCM_E_ARG, "params required");
  hipStream_t st = (hipStream_t)stream;
  tcm_pauli* H = const_cast<tcm_pauli*>(ham);
  return dtype == TCM_C64
             ? energy_grad_impl<float>(prog->p, dtype, params, param_stride, H, inputs, batch, energy, grad, workspace, ws_bytes, st)
             : energy_grad_impl<double>(prog->p, dtype, params, param_stride, H, inputs, batch, energy, grad, workspace, ws_bytes, st);
}

int tcm_multi_workspace_bytes(const tcm_multi_desc* descs, int32_t ndesc, int32_t dtype, int64_t* bytes) {
  if (!bytes || !valid_dtype(dtype) || ndesc < 0 || (ndesc > 0 && !descs)) return set_error(TCM_E_ARG, "bad arguments");
  for (int i = 0; i < ndesc; ++i)
    if (!descs[i].prog || descs[i].batch < 0) return set_error(TCM_E_ARG, "descriptor %d: bad program or batch", i);
  *bytes = multi_ws(descs, ndesc, dtype);
  return 0;
}

int tcm_forward_multi(const tcm_multi_desc* descs, int32_t ndesc, int32_t dtype, int32_t init_zero, void* workspace,
                      int64_t ws_bytes, void* stream) {
  if (!valid_dtype(dtype) || ndesc < 0 || (ndesc > 0 && !descs)) return set_error(TCM_E_ARG, "bad arguments");
  for (int i = 0; i < ndesc; ++i) {
    const tcm_multi_desc& d = descs[i];
    if (!d.prog || d.batch < 0 || (d.batch > 0 && !d.state)) return set_error(TCM_E_ARG, "descriptor %d: bad arguments", i);
    if (d.prog->p.nparams > 0 && d.batch > 0 && !d.params) return set_error(TCM_E_ARG, "descriptor %d: params required", i);
  }
  if (multi_ws(descs, ndesc, dtype) > ws_bytes) return set_error(TCM_E_WORKSPACE, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  int rc = dtype == TCM_C64 ? forward_multi_impl<float>(descs, ndesc, dtype, init_zero, workspace, st)
                            : forward_multi_impl<double>(descs, ndesc, dtype, init_zero, workspace, st);
  if (rc) return rc;
  HIP_TRY(hipGetLastError(), "forward_multi launch");
  return 0;
}

int tcm_multi_plan_build(tcm_program* const* progs, const int64_t* batches, int32_t nprog, int32_t dtype,
                         tcm_multi_plan** out, double* seconds) {
  if (!out) return set_error(TCM_E_ARG, "out is NULL");
  *out = nullptr;
  if (seconds) *seconds = 0.0;
  if (!valid_dtype(dtype) || nprog < 1 || !progs) return set_error(TCM_E_ARG, "bad arguments to tcm_multi_plan_build");
  if (nprog > kMaxPlanPrograms)
    return set_error(TCM_E_ARG, "tcm_multi_plan_build: %d programs (at most %d per plan)", nprog, kMaxPlanPrograms);
  std::vector<const Program*> ps;
  std::vector<double> weight;
  for (int i = 0; i < nprog; ++i) {
    if (!progs[i]) return set_error(TCM_E_ARG, "program %d is NULL", i);
    if (batches && batches[i] < 0) return set_error(TCM_E_ARG, "program %d: negative batch", i);
    ps.push_back(&progs[i]->p);
    weight.push_back(std::ldexp(batches ? (double)batches[i] : 1.0, progs[i]->p.n));
  }
  tcm_multi_plan* M = new tcm_multi_plan();
  M->progs.assign(progs, progs + nprog);
  M->dtype = dtype;
  for (const Program* p : ps) M->maxp = std::max(M->maxp, (int)p->sched[dtype][0].passes.size());
  M->info.resize(M->maxp);
  M->kid.assign(M->maxp, -1);
  std::vector<std::string> bodies;
  for (int r = 0; r < M->maxp; ++r) {
    std::string src = gen_multi_source(ps, weight, dtype, r, (int)bodies.size(), &M->info[r]);
    if (src.empty()) continue;
    if (M->info[r].lds > 160u * 1024u) {
      const size_t lds = M->info[r].lds;
      delete M;
      return set_error(TCM_E_UNSUPPORTED, "multi-program pass %d needs %zu bytes of LDS per block", r, lds);
    }
    M->kid[r] = (int)bodies.size();
    bodies.push_back(std::move(src));
  }
  if (!bodies.empty())
    if (int rc = compile_bodies(gen_multi_prelude(dtype), bodies, "multi", M->code, M->chunk_of, &M->seconds)) {
      delete M;
      return rc;
    }
  if (seconds) *seconds = M->seconds;
  *out = M;
  return 0;
}

int tcm_multi_plan_free(tcm_multi_plan* plan) {
  if (!plan) return 0;
  int cur = -1;
  const bool have = hipGetDevice(&cur) == hipSuccess;
  for (size_t d = 0; d < plan->mods.size(); ++d) {
    if (plan->mods[d].empty()) continue;
    (void)hipSetDevice((int)d);
    for (void* m : plan->mods[d])
      if (m) (void)hipModuleUnload((hipModule_t)m);
  }
  if (have) (void)hipSetDevice(cur);
  (void)hipGetLastError();
  delete plan;
  return 0;
}

int tcm_forward_multi_plan(tcm_multi_plan* plan, const tcm_multi_desc* descs, int32_t ndesc, int32_t init_zero,
                           void* workspace, int64_t ws_bytes, void* stream) {
  if (!plan || ndesc != (int)plan->progs.size() || !descs)
    return set_error(TCM_E_ARG, "tcm_forward_multi_plan: %d descriptors for a plan of %zu programs", ndesc,
                     plan ? plan->progs.size() : (size_t)0);
  for (int i = 0; i < ndesc; ++i) {
    const tcm_multi_desc& d = descs[i];
    if (d.prog != plan->progs[i]) return set_error(TCM_E_ARG, "descriptor %d: not the plan's program %d", i, i);
    if (d.batch < 0 || (d.batch > 0 && !d.state)) return set_error(TCM_E_ARG, "descriptor %d: bad arguments", i);
    if (d.prog->p.nparams > 0 && d.batch > 0 && !d.params) return set_error(TCM_E_ARG, "descriptor %d: params required", i);
  }
  const int dtype = plan->dtype;
  if (multi_ws(descs, ndesc, dtype) > ws_bytes) return set_error(TCM_E_WORKSPACE, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  int rc = dtype == TCM_C64 ? forward_plan_impl<float>(plan, descs, ndesc, dtype, init_zero, workspace, st)
                            : forward_plan_impl<double>(plan, descs, ndesc, dtype, init_zero, workspace, st);
  if (rc) return rc;
  HIP_TRY(hipGetLastError(), "forward_multi_plan launch");
  return 0;
}

int tcm_adjoint_vjp(tcm_program* prog, int32_t dtype, const void* params, int64_t param_stride, const void* ckpts,
                    const void* state, const void* grad_state, int64_t batch, double* grad_params, void* grad_input,
                    void* workspace, int64_t ws_bytes, void* stream) {
  if (!prog || !valid_dtype(dtype) || batch < 0 || (batch > 0 && (!state || !grad_state)))
    return set_error(TCM_E_ARG, "bad arguments to tcm_adjoint_vjp");
  if (batch == 0) return 0;
  if (prog->p.nparams > 0 && !params) return set_error(TCM_E_ARG, "params required");
  hipStream_t st = (hipStream_t)stream;
  return dtype == TCM_C64
             ? vjp_impl<float>(prog->p, dtype, params, param_stride, ckpts, state, grad_state, batch, grad_params, grad_input, workspace, ws_bytes, st)
             : vjp_impl<double>(prog->p, dtype, params, param_stride, ckpts, state, grad_state, batch, grad_params, grad_input, workspace, ws_bytes, st);
}

}  // extern "C"
