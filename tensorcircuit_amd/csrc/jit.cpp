// Pass compiler, part 2: compile the generated pass kernels with hiprtc,
// cache the code objects on disk, load them per device and launch them.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "kernels.h"

namespace tcm {

int set_error(int code, const char* fmt, ...);

namespace {

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ull;
  }
  return h;
}

std::string cache_dir() {
  const char* d = getenv("TCM_CACHE_DIR");
  std::string base;
  if (d && *d) base = d;
  else {
    const char* h = getenv("HOME");
    if (!h || !*h) return std::string();
    base = std::string(h) + "/.cache/tensorcircuit_amd";
  }
  return base + "/jit";
}

void mkdirs(const std::string& p) {
  std::string cur;
  for (size_t i = 0; i < p.size(); ++i) {
    cur += p[i];
    if (p[i] == '/' && cur.size() > 1) mkdir(cur.c_str(), 0755);
  }
  mkdir(p.c_str(), 0755);
}

bool read_file(const std::string& path, std::vector<char>& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return !out.empty();
}

void write_file(const std::string& path, const std::vector<char>& data) {
  const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    if (!f) return;
    f.write(data.data(), (std::streamsize)data.size());
    if (!f) return;
  }
  std::rename(tmp.c_str(), path.c_str());
}

std::string device_arch() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    std::string a = prop.gcnArchName;
    const size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
  }
  (void)hipGetLastError();
  const char* e = getenv("TCM_OFFLOAD_ARCH");
  return (e && *e) ? std::string(e) : std::string("gfx950");
}

// one hiprtc compilation; empty `err` on success
void rtc_compile(const std::string& src, const std::vector<std::string>& opts, std::vector<char>* code,
                 std::string* err) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "tcm_passes.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    *err = "hiprtcCreateProgram failed";
    return;
  }
  std::vector<const char*> o;
  for (const auto& s : opts) o.push_back(s.c_str());
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)o.size(), o.data());
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    *err = std::string("hiprtc: ") + hiprtcGetErrorString(r) + "\n" + log.substr(0, 4000);
    hiprtcDestroyProgram(&prog);
    return;
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  code->resize(n);
  if (n) hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  if (!n) *err = "hiprtc produced no code";
}

}  // namespace

// Compile kernel bodies (each a complete __global__ definition) behind a
// shared prelude into code objects: bodies are dealt round-robin into up
// to 16 chunks compiled on parallel threads; each chunk is cached on disk
// under a hash of its text and the compiler options.  On success code[c]
// holds chunk c and chunk_of[i] the chunk of body i.
int compile_bodies(const std::string& prelude, const std::vector<std::string>& bodies, const std::string& dump_tag,
                   std::vector<std::vector<char>>& code, std::vector<int32_t>& chunk_of, double* seconds) {
  const auto t0 = std::chrono::steady_clock::now();
  *seconds = 0.0;
  code.clear();
  chunk_of.assign(bodies.size(), 0);
  if (bodies.empty()) return 0;
  const std::string arch = device_arch();
  const std::vector<std::string> opts = {"--offload-arch=" + arch, "-O3", "-std=c++17", "-ffp-contract=fast",
                                         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  std::string okey = arch;
  for (const auto& o : opts) okey += " " + o;
  unsigned hw = std::thread::hardware_concurrency();
  const int nch = (int)std::max<size_t>(1, std::min<size_t>({bodies.size(), (size_t)(hw ? hw : 4), (size_t)16}));
  std::vector<std::string> srcs(nch, prelude);
  for (size_t i = 0; i < bodies.size(); ++i) {
    chunk_of[i] = (int)(i % nch);
    srcs[i % nch] += bodies[i];
  }
  if (const char* dump = getenv("TCM_JIT_DUMP"))
    for (int c = 0; c < nch; ++c) {
      std::ofstream f(std::string(dump) + "/tcm_passes_" + dump_tag + "_" + std::to_string(c) + ".hip");
      f << srcs[c];
    }
  code.assign(nch, std::vector<char>());
  std::vector<std::string> errs(nch);
  const std::string dir = cache_dir();
  std::vector<std::string> paths(nch);
  std::vector<int> todo;
  for (int c = 0; c < nch; ++c) {
    if (!dir.empty()) {
      char h[32];
      snprintf(h, sizeof h, "%016llx", (unsigned long long)fnv1a(srcs[c], fnv1a(okey)));
      paths[c] = dir + "/" + h + ".co";
      if (read_file(paths[c], code[c])) continue;
    }
    todo.push_back(c);
  }
  std::vector<std::thread> th;
  for (int c : todo) th.emplace_back(rtc_compile, std::cref(srcs[c]), std::cref(opts), &code[c], &errs[c]);
  for (auto& t : th) t.join();
  for (int c : todo)
    if (!errs[c].empty()) return set_error(TCM_E_UNSUPPORTED, "pass compiler: %s", errs[c].c_str());
  if (!dir.empty() && !todo.empty()) {
    mkdirs(dir);
    for (int c : todo) write_file(paths[c], code[c]);
  }
  *seconds = todo.empty() ? 0.0 : std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int jit_build(const Program& P, Sched& S, int dt, int mode) {
  S.jit_prefix = mode == 0 ? "tcm_fwd_k" : "tcm_adj_k";
  std::vector<std::string> bodies;
  const std::string prelude = gen_pass_source(P, S, dt, mode, S.jit, bodies);
  if (int rc = compile_bodies(prelude, bodies, std::to_string(mode), S.jit_code, S.jit_chunk, &S.jit_seconds)) {
    S.jit_state = -1;
    return rc;
  }
  S.jit_state = 1;
  return 0;
}

// load code objects and look up kernels prefix + index (index i in chunk_of[i])
int load_modules(const std::vector<std::vector<char>>& code, const std::vector<int32_t>& chunk_of,
                 const std::string& prefix, std::vector<void*>& mods_out, std::vector<void*>& fns_out) {
  std::vector<hipModule_t> mods(code.size(), nullptr);
  for (size_t c = 0; c < code.size(); ++c) {
    const hipError_t e = hipModuleLoadData(&mods[c], code[c].data());
    if (e != hipSuccess) return set_error((int)e, "hipModuleLoadData: %s", hipGetErrorString(e));
    mods_out.push_back((void*)mods[c]);
  }
  fns_out.assign(chunk_of.size(), nullptr);
  for (size_t k = 0; k < chunk_of.size(); ++k) {
    hipFunction_t f;
    const std::string name = prefix + std::to_string(k);
    const hipError_t e = hipModuleGetFunction(&f, mods[chunk_of[k]], name.c_str());
    if (e != hipSuccess) return set_error((int)e, "hipModuleGetFunction(%s): %s", name.c_str(), hipGetErrorString(e));
    fns_out[k] = (void*)f;
  }
  return 0;
}

int jit_load(const Sched& S, DeviceTables& D) {
  if (S.jit_state != 1 || !D.jit_fn.empty()) return 0;
  return load_modules(S.jit_code, S.jit_chunk, S.jit_prefix, D.jit_mod, D.jit_fn);
}

int jit_launch_fwd(const Sched& S, const DeviceTables& D, int pi, int dt, const void* mtab, const void* ctab,
                   void* state, int64_t batch, int n, void* ckout, hipStream_t st, int zinit) {
  const DPass& P = S.dpass[pi];
  const JitPass& J = S.jit[pi];
  (void)dt;
  long long sstride = 1ll << n, mstride = S.mtab_size, mbase = J.mbase, dbase = J.dbase;
  void* a_st = state;
  const void* a_mt = mtab;
  const void* a_ct = ctab;
  void* a_ck = ckout;
  long long total = batch << P.nblk_log;
  // one entry per parameter of the generated signature, in its order
  int a_z = zinit;
  void* args[] = {&a_st, &sstride, &a_mt, &mstride, &mbase, &dbase, &a_ct, &a_ck, &total, &a_z};
  const unsigned blocks = (unsigned)total;
  const hipError_t e = hipModuleLaunchKernel((hipFunction_t)D.jit_fn[J.kernel], blocks, 1, 1,
                                             1u << P.tthreads_log, 1, 1, 0, st, args, nullptr);
  if (e != hipSuccess) return set_error((int)e, "hipModuleLaunchKernel: %s", hipGetErrorString(e));
  return 0;
}

int jit_launch_adj(const Sched& S, const DeviceTables& D, int pi, int dt, const void* mtab, const void* ctab,
                   const void* psi_in, void* psi_out, const void* lam_in, void* lam_out, int64_t batch, int n,
                   double* part, int64_t pstride, hipStream_t st) {
  const DPass& P = S.dpass[pi];
  const JitPass& J = S.jit[pi];
  (void)dt;
  long long sstride = 1ll << n, mstride = S.mtab_size, mbase = J.mbase, dbase = J.dbase, ps = pstride;
  const void* a_pi = psi_in;
  void* a_po = psi_out;
  const void* a_li = lam_in;
  void* a_lo = lam_out;
  const void* a_mt = mtab;
  const void* a_ct = ctab;
  double* a_part = part;
  long long total = batch << P.nblk_log;
  void* args[] = {&a_pi, &a_po, &a_li, &a_lo, &sstride, &a_mt, &mstride, &mbase, &dbase, &a_ct, &a_part, &ps, &total};
  const unsigned blocks = (unsigned)total;
  const hipError_t e = hipModuleLaunchKernel((hipFunction_t)D.jit_fn[J.kernel], blocks, 1, 1,
                                             1u << P.tthreads_log, 1, 1, 0, st, args, nullptr);
  if (e != hipSuccess) return set_error((int)e, "hipModuleLaunchKernel: %s", hipGetErrorString(e));
  return 0;
}

}  // namespace tcm
