// Pass compiler, part 1: HIP source text for the tile passes of a schedule.
//
// The generic fused kernels (kernels.hip) read each pass's gate list from
// device tables and branch on op kind and register slot per op.  Here the
// same pass is written out as straight-line code: tile/bit mappings,
// register slots and table offsets are compile-time constants, permutation
// gates (X, CNOT, SWAP) become register renames, and there is no per-op
// dispatch.  The arithmetic of every op and gradient is the generic
// kernel's, term for term.
//
// Passes with the same body (layered circuits repeat them with shifted
// per-item matrix offsets) share one kernel: offsets into the per-item
// table are emitted relative to the pass's smallest offset, which the
// launcher passes as `mbase`.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.h"

namespace tcm {

namespace {

std::string fmt(const char* f, ...) {
  va_list ap;
  va_start(ap, f);
  va_list ap2;
  va_copy(ap2, ap);
  const int len = vsnprintf(nullptr, 0, f, ap);
  va_end(ap);
  std::string out(len > 0 ? (size_t)len : 0, '\0');
  if (len > 0) vsnprintf(&out[0], (size_t)len + 1, f, ap2);
  va_end(ap2);
  return out;
}

// LDS swizzle of the compiled passes: the low 5 element bits XOR the higher
// 5-bit fields, so that 32 lanes (256 bytes of complex64, all 64 banks) land
// on distinct banks in more of the register-group exchanges than with the
// 4-bit swizzle of the generic kernels (swz_host); linear over GF(2)
uint32_t swz_h(uint32_t t) { return t ^ (((t >> 5) ^ (t >> 10) ^ (t >> 15)) & 0x1Fu); }
// register offsets of a group: dg.S holds them in the generic kernels'
// swizzle, which is its own inverse (it reads only bits it leaves alone)
uint32_t sreg(int32_t s) { return swz_h(swz_host((uint32_t)s)); }

// OR of bit fields: source bits src[i] -> destination bits dst[i], grouped
// into runs of consecutive positions.
std::string deposit_expr(const char* x, const std::vector<int>& src, const std::vector<int>& dst, const char* ty) {
  std::string e;
  size_t i = 0;
  while (i < src.size()) {
    size_t len = 1;
    while (i + len < src.size() && src[i + len] == src[i] + (int)len && dst[i + len] == dst[i] + (int)len) ++len;
    const unsigned long long mask = (len >= 64) ? ~0ull : ((1ull << len) - 1);
    const bool wide = std::string(ty).find("long") != std::string::npos;
    std::string term = fmt("(((%s)(%s) >> %d) & 0x%llx%s)", ty, x, src[i], mask, wide ? "ull" : "u");
    if (dst[i] > 0) term = fmt("(%s << %d)", term.c_str(), dst[i]);
    e += (e.empty() ? "" : " | ") + term;
    i += len;
  }
  return e.empty() ? std::string("0") : e;
}

const char* kPrelude = R"(
typedef REAL T;
struct __attribute__((aligned(2 * sizeof(REAL)))) cT { T x, y; };
__device__ __forceinline__ unsigned swz(unsigned t) {
  return t ^ (((t >> 5) ^ (t >> 10) ^ (t >> 15)) & 0x1Fu);
}
// LDS element at a byte offset (addresses are kept in bytes so the XOR
// swizzle needs no per-access shift)
#define LDSB(arr, off) (*reinterpret_cast<cT*>(reinterpret_cast<char*>(arr) + (off)))
// sign bit of a table entry (1 for negative)
__device__ __forceinline__ unsigned sgnb(float x) { return __float_as_uint(x) >> 31; }
__device__ __forceinline__ unsigned sgnb(double x) { return (unsigned)((unsigned long long)__double_as_longlong(x) >> 63); }
__device__ __forceinline__ T imcj(cT a, cT b) { return a.x * b.y - a.y * b.x; }
__device__ __forceinline__ T recj(cT a, cT b) { return a.x * b.x + a.y * b.y; }
union tcm_d2i { double d; int i[2]; };
// DPP moves inside each 16-lane row (row_shr, row_ror, quad_perm); with
// row_shr, lanes with no source read 0
template <int CTRL>
__device__ __forceinline__ double dpp_shr(double x) {
  tcm_d2i u, r;
  u.d = x;
  r.i[0] = __builtin_amdgcn_update_dpp(0, u.i[0], CTRL, 0xF, 0xF, true);
  r.i[1] = __builtin_amdgcn_update_dpp(0, u.i[1], CTRL, 0xF, 0xF, true);
  return r.d;
}
template <int CTRL>
__device__ __forceinline__ float dpp_shr(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, true));
}
__device__ __forceinline__ double read_lane(double x, int lane) {
  tcm_d2i u, r;
  u.d = x;
  r.i[0] = __builtin_amdgcn_readlane(u.i[0], lane);
  r.i[1] = __builtin_amdgcn_readlane(u.i[1], lane);
  return r.d;
}
// sum over each 16-lane row in the working precision (the row total lands
// in its last lane); a float row sum covers 256 amplitudes (16 per thread)
// and is widened to double by the caller
template <typename V>
__device__ __forceinline__ V rowsum(V x) {
  x += dpp_shr<0x111>(x);
  x += dpp_shr<0x112>(x);
  x += dpp_shr<0x114>(x);
  x += dpp_shr<0x118>(x);
  return x;
}
// row_bcast into the rows of ROWS; the other rows add the zero `old` value
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_rows(double x) {
  tcm_d2i u, r;
  u.d = x;
  r.i[0] = __builtin_amdgcn_update_dpp(0, u.i[0], CTRL, ROWS, 0xF, false);
  r.i[1] = __builtin_amdgcn_update_dpp(0, u.i[1], CTRL, ROWS, 0xF, false);
  return r.d;
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_rows(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROWS, 0xF, false));
}
// sum over each 16-lane row, in every lane of the row: quad_perm swaps of
// lanes (i ^ 1, then i ^ 2) give each lane its quad's sum, row_ror 4 and 8
// add the other three quads
template <typename V>
__device__ __forceinline__ V rowall(V x) {
  x += dpp_shr<0xB1>(x);
  x += dpp_shr<0x4E>(x);
  x += dpp_shr<0x124>(x);
  x += dpp_shr<0x128>(x);
  return x;
}
// sum over the wave, total in lane 63 only: row prefix sums by
// row_shr 1/2/4/8, then row_bcast:15 into rows 1 and 3 and row_bcast:31
// into rows 2 and 3
template <typename V>
__device__ __forceinline__ V wsum63(V x) {
  x = rowsum(x);
  x += dpp_rows<0x142, 0xA>(x);
  x += dpp_rows<0x143, 0xC>(x);
  return x;
}
// fp64 sum over a 64-lane wave (uniform result)
__device__ __forceinline__ double wsum(double x) {
  x += dpp_shr<0x111>(x);
  x += dpp_shr<0x112>(x);
  x += dpp_shr<0x114>(x);
  x += dpp_shr<0x118>(x);
  return (read_lane(x, 15) + read_lane(x, 31)) + (read_lane(x, 47) + read_lane(x, 63));
}
)";

struct Gen {
  const Sched& S;
  const DPass& dp;
  const HPass& hp;
  bool dag;       // adjoint: apply U^dagger, psi and lambda arrays
  int R, nreg;
  int mbase, dbase;
  std::string out;
  std::vector<int> where;  // logical register index -> variable index
  // sigma entries of the shear-form rotations applied so far in this pass:
  // each applies R(phi') = sigma R(phi), so the pass output is exact once
  // multiplied by the product of their sigmas (frame_fix)
  std::vector<std::string> frame;

  Gen(const Sched& s, const DPass& d, const HPass& h, bool adj) : S(s), dp(d), hp(h), dag(adj), R(s.R) {
    nreg = 1 << R;
    mbase = dbase = -1;
  }

  void line(const std::string& s) { out += s; out += '\n'; }

  // matrix element source for an op's table offset `moff` (+i): base
  // recipes relative to mbase, merged diagonal tables (built per schedule
  // after the base recipes) relative to dbase, so that passes repeating
  // with shifted tables share one kernel
  std::string mel(int moff, int i) const {
    if (moff >= S.mtab_base) return fmt("md[%d]", moff - dbase + i);
    if (moff >= 0) return fmt("mt[%d]", moff - mbase + i);
    return fmt("ct[%d]", -moff - 1 + i);
  }
  std::string var(const char* a, int j) const { return fmt("%s%d", a, where[j]); }

  // ---- ops on one register array named `a` (DAG per `dag`) ----
  void op_u1(const char* a, int S0, const std::string& pre) {
    const int m = 1 << S0;
    for (int j = 0; j < nreg; ++j) {
      if (j & m) continue;
      const std::string x = var(a, j), y = var(a, j | m);
      line(fmt("{ const cT x = %s, y = %s;", x.c_str(), y.c_str()));
      line(fmt("  %s = cT{%sar * x.x - %sai * x.y + %sbr * y.x - %sbi * y.y, %sar * x.y + %sai * x.x + %sbr * y.y + %sbi * y.x};",
               x.c_str(), pre.c_str(), pre.c_str(), pre.c_str(), pre.c_str(), pre.c_str(), pre.c_str(), pre.c_str(),
               pre.c_str()));
      line(fmt("  %s = cT{%scr * x.x - %sci * x.y + %sdr * y.x - %sdi * y.y, %scr * x.y + %sci * x.x + %sdr * y.y + %sdi * y.x}; }",
               y.c_str(), pre.c_str(), pre.c_str(), pre.c_str(), pre.c_str(), pre.c_str(), pre.c_str(), pre.c_str(),
               pre.c_str()));
    }
  }
  // real rotation as three shears per real 2-vector (x, y) of a pair:
  // x -= t y; y += s x; x -= t y (three fused multiply-adds instead of two
  // multiplies and two multiply-adds)
  void op_shear(const char* a, int S0, const std::string& pre) {
    const int m = 1 << S0;
    const char* P = pre.c_str();
    for (int j = 0; j < nreg; ++j) {
      if (j & m) continue;
      const std::string x = var(a, j), y = var(a, j | m);
      line(fmt("%s.x -= %st * %s.x; %s.y -= %st * %s.y;", x.c_str(), P, y.c_str(), x.c_str(), P, y.c_str()));
      line(fmt("%s.x += %ss * %s.x; %s.y += %ss * %s.y;", y.c_str(), P, x.c_str(), y.c_str(), P, x.c_str()));
      line(fmt("%s.x -= %st * %s.x; %s.y -= %st * %s.y;", x.c_str(), P, y.c_str(), x.c_str(), P, y.c_str()));
    }
  }
  // negate every register of the arrays when the product of the pass's
  // sigmas is -1 (sign bits XORed; block-uniform)
  void frame_fix(const std::vector<const char*>& arrs) {
    if (frame.empty()) return;
    std::string e;
    for (const auto& f : frame) e += fmt("%ssgnb(%s)", e.empty() ? "" : " ^ ", f.c_str());
    line(fmt("if (%s) {", e.c_str()));
    for (auto a : arrs)
      for (int j = 0; j < nreg; ++j) line(fmt("  %s%d = cT{-%s%d.x, -%s%d.y};", a, j, a, j, a, j));
    line("}");
  }
  void op_r1(const char* a, int S0, const std::string& pre) {
    const int m = 1 << S0;
    for (int j = 0; j < nreg; ++j) {
      if (j & m) continue;
      const std::string x = var(a, j), y = var(a, j | m);
      const char* P = pre.c_str();
      line(fmt("{ const cT x = %s, y = %s;", x.c_str(), y.c_str()));
      line(fmt("  %s = cT{%sa * x.x + %sb * y.x, %sa * x.y + %sb * y.y};", x.c_str(), P, P, P, P));
      line(fmt("  %s = cT{%sc * x.x + %sd * y.x, %sc * x.y + %sd * y.y}; }", y.c_str(), P, P, P, P));
    }
  }
  void op_d1(const char* a, int S0, const std::string& pre) {
    const int m = 1 << S0;
    const char* P = pre.c_str();
    for (int j = 0; j < nreg; ++j) {
      const std::string x = var(a, j);
      const char* d = (j & m) ? "1" : "0";
      line(fmt("{ const cT x = %s; %s = cT{%sd%sr * x.x - %sd%si * x.y, %sd%sr * x.y + %sd%si * x.x}; }", x.c_str(),
               x.c_str(), P, d, P, d, P, d, P, d));
    }
  }
  void op_d2(const char* a, int S0, int S1, const std::string& pre) {
    const char* P = pre.c_str();
    for (int j = 0; j < nreg; ++j) {
      const int r = (((j >> S0) & 1) << 1) | ((j >> S1) & 1);
      const std::string x = var(a, j);
      line(fmt("{ const cT x = %s; %s = cT{%sd%dr * x.x - %sd%di * x.y, %sd%dr * x.y + %sd%di * x.x}; }", x.c_str(),
               x.c_str(), P, r, P, r, P, r, P, r));
    }
  }
  void op_dg(const char* a, const std::string& pre) {
    const char* P = pre.c_str();
    for (int j = 0; j < nreg; ++j) {
      const std::string x = var(a, j);
      line(fmt("{ const cT x = %s; %s = cT{%sg%dr * x.x - %sg%di * x.y, %sg%dr * x.y + %sg%di * x.x}; }", x.c_str(),
               x.c_str(), P, j, P, j, P, j, P, j));
    }
  }
  void op_u2(const char* a, int S0, int S1, const std::string& pre) {
    const int m0 = 1 << S0, m1 = 1 << S1;
    const char* P = pre.c_str();
    for (int j = 0; j < nreg; ++j) {
      if (j & (m0 | m1)) continue;
      const int idx[4] = {j, j | m1, j | m0, j | m0 | m1};
      line(fmt("{ const cT x0 = %s, x1 = %s, x2 = %s, x3 = %s;", var(a, idx[0]).c_str(), var(a, idx[1]).c_str(),
               var(a, idx[2]).c_str(), var(a, idx[3]).c_str()));
      for (int r = 0; r < 4; ++r) {
        std::string re, im;
        for (int c = 0; c < 4; ++c) {
          re += fmt("%s%su%d%dr * x%d.x - %su%d%di * x%d.y", c ? " + " : "", P, r, c, c, P, r, c, c);
          im += fmt("%s%su%d%dr * x%d.y + %su%d%di * x%d.x", c ? " + " : "", P, r, c, c, P, r, c, c);
        }
        line(fmt("  %s = cT{%s, %s};", var(a, idx[r]).c_str(), re.c_str(), im.c_str()));
      }
      line("}");
    }
  }
  void op_cz(const char* a, int S0, int S1) {
    for (int j = 0; j < nreg; ++j)
      if ((j >> S0 & 1) && (j >> S1 & 1)) {
        const std::string x = var(a, j);
        line(fmt("%s = cT{-%s.x, -%s.y};", x.c_str(), x.c_str(), x.c_str()));
      }
  }

  // coefficient declarations for an op, prefixed `pre` (DAG-adjusted)
  void coeffs(const DOp& op, const std::string& pre) {
    const char* P = pre.c_str();
    switch (op.kind) {
      case OK_U1:
        if (!dag)
          line(fmt("const T %sar = %s, %sai = %s, %sbr = %s, %sbi = %s, %scr = %s, %sci = %s, %sdr = %s, %sdi = %s;", P,
                   mel(op.moff, 0).c_str(), P, mel(op.moff, 1).c_str(), P, mel(op.moff, 2).c_str(), P,
                   mel(op.moff, 3).c_str(), P, mel(op.moff, 4).c_str(), P, mel(op.moff, 5).c_str(), P,
                   mel(op.moff, 6).c_str(), P, mel(op.moff, 7).c_str()));
        else  // U^dagger: a* c* / b* d*
          line(fmt("const T %sar = %s, %sai = -%s, %sbr = %s, %sbi = -%s, %scr = %s, %sci = -%s, %sdr = %s, %sdi = -%s;",
                   P, mel(op.moff, 0).c_str(), P, mel(op.moff, 1).c_str(), P, mel(op.moff, 4).c_str(), P,
                   mel(op.moff, 5).c_str(), P, mel(op.moff, 2).c_str(), P, mel(op.moff, 3).c_str(), P,
                   mel(op.moff, 6).c_str(), P, mel(op.moff, 7).c_str()));
        break;
      case OK_R1:
        if (op.flags & OP_F_SHEAR)  // R(phi')^dagger = R(-phi'): t and s' change sign
          line(fmt("const T %st = %s%s, %ss = %s%s;", P, dag ? "-" : "", mel(op.moff, 4).c_str(), P, dag ? "-" : "",
                   mel(op.moff, 5).c_str()));
        else
          line(fmt("const T %sa = %s, %sb = %s, %sc = %s, %sd = %s;", P, mel(op.moff, 0).c_str(), P,
                   mel(op.moff, dag ? 2 : 1).c_str(), P, mel(op.moff, dag ? 1 : 2).c_str(), P,
                   mel(op.moff, 3).c_str()));
        break;
      case OK_D1:
        line(fmt("const T %sd0r = %s, %sd0i = %s%s, %sd1r = %s, %sd1i = %s%s;", P, mel(op.moff, 0).c_str(), P,
                 dag ? "-" : "", mel(op.moff, 1).c_str(), P, mel(op.moff, 2).c_str(), P, dag ? "-" : "",
                 mel(op.moff, 3).c_str()));
        break;
      case OK_D2:
        for (int r = 0; r < 4; ++r)
          line(fmt("const T %sd%dr = %s, %sd%di = %s%s;", P, r, mel(op.moff, 2 * r).c_str(), P, r, dag ? "-" : "",
                   mel(op.moff, 2 * r + 1).c_str()));
        break;
      case OK_DG:
        for (int j = 0; j < nreg; ++j)
          line(fmt("const T %sg%dr = %s, %sg%di = %s%s;", P, j, mel(op.moff, 2 * j).c_str(), P, j, dag ? "-" : "",
                   mel(op.moff, 2 * j + 1).c_str()));
        break;
      case OK_U2:
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) {
            const int e = dag ? (c * 4 + r) : (r * 4 + c);
            line(fmt("const T %su%d%dr = %s, %su%d%di = %s%s;", P, r, c, mel(op.moff, 2 * e).c_str(), P, r, c,
                     dag ? "-" : "", mel(op.moff, 2 * e + 1).c_str()));
          }
        break;
      default:
        break;
    }
  }

  // apply op to the arrays in `arrs` (one or two names)
  void apply(const DOp& op, const std::vector<const char*>& arrs, int id) {
    const std::string pre = fmt("c%d_", id);
    coeffs(op, pre);
    switch (op.kind) {
      case OK_U1: for (auto a : arrs) op_u1(a, op.s0, pre); break;
      case OK_R1:
        if (op.flags & OP_F_SHEAR) {
          for (auto a : arrs) op_shear(a, op.s0, pre);
          frame.push_back(mel(op.moff, 6));
        } else {
          for (auto a : arrs) op_r1(a, op.s0, pre);
        }
        break;
      case OK_D1: for (auto a : arrs) op_d1(a, op.s0, pre); break;
      case OK_D2: for (auto a : arrs) op_d2(a, op.s0, op.s1, pre); break;
      case OK_DG: for (auto a : arrs) op_dg(a, pre); break;
      case OK_U2: for (auto a : arrs) op_u2(a, op.s0, op.s1, pre); break;
      case OK_CZ: for (auto a : arrs) op_cz(a, op.s0, op.s1); break;
      case OK_X1: {
        const int m = 1 << op.s0;
        for (int j = 0; j < nreg; ++j)
          if (!(j & m)) std::swap(where[j], where[j | m]);
      } break;
      case OK_CX: {
        const int c = 1 << op.s0, t = 1 << op.s1;
        for (int j = 0; j < nreg; ++j)
          if ((j & c) && !(j & t)) std::swap(where[j], where[j | t]);
      } break;
      case OK_SWAP: {
        const int a0 = 1 << op.s0, a1 = 1 << op.s1;
        for (int j = 0; j < nreg; ++j)
          if ((j & a0) && !(j & a1)) std::swap(where[j], where[(j ^ a0) | a1]);
      } break;
      default:
        break;
    }
  }

  std::string pv(int j) const { return var("p", j); }
  std::string lv(int j) const { return var("l", j); }

  // ga += (or -=) Im(conj(a) b) = a.x b.y - a.y b.x, and Re(conj(a) b)
  void imcj_acc(const std::string& a, const std::string& b, bool plus) {
    line(fmt("ga %s= %s.x * %s.y; ga %s= %s.y * %s.x;", plus ? "+" : "-", a.c_str(), b.c_str(), plus ? "-" : "+",
             a.c_str(), b.c_str()));
  }
  void recj_acc(const std::string& a, const std::string& b, bool plus) {
    const char* sg = plus ? "+" : "-";
    line(fmt("ga %s= %s.x * %s.x; ga %s= %s.y * %s.y;", sg, a.c_str(), b.c_str(), sg, a.c_str(), b.c_str()));
  }

  // gradient Im<l|G|p> of a generator on the registers; returns expression code into `ga`
  bool grad(int gen, int s0, int s1, int goff) {
    line("T ga = 0;");
    switch (gen) {
      // every term is one `ga += a * b` / `ga -= a * b` statement, so the
      // compiler emits one fused multiply-add per product (a chain into ga)
      // instead of forming each complex product separately
      case GN_X: {
        const int m = 1 << s0;
        for (int j = 0; j < nreg; ++j)
          if (!(j & m)) {
            imcj_acc(lv(j), pv(j | m), true);
            imcj_acc(lv(j | m), pv(j), true);
          }
      } break;
      case GN_Y: {
        const int m = 1 << s0;
        for (int j = 0; j < nreg; ++j)
          if (!(j & m)) {
            recj_acc(lv(j | m), pv(j), true);
            recj_acc(lv(j), pv(j | m), false);
          }
      } break;
      case GN_Z:
        for (int j = 0; j < nreg; ++j) imcj_acc(lv(j), pv(j), !(j >> s0 & 1));
        break;
      case GN_P1:
        for (int j = 0; j < nreg; ++j)
          if (j >> s0 & 1) imcj_acc(lv(j), pv(j), true);
        break;
      case GN_XX: {
        const int m = (1 << s0) | (1 << s1);
        for (int j = 0; j < nreg; ++j) imcj_acc(lv(j), pv(j ^ m), true);
      } break;
      case GN_YY: {
        const int m = (1 << s0) | (1 << s1);
        for (int j = 0; j < nreg; ++j) {
          const int par = ((j >> s0) ^ (j >> s1)) & 1;
          imcj_acc(lv(j), pv(j ^ m), par != 0);
        }
      } break;
      case GN_ZZ:
        for (int j = 0; j < nreg; ++j) {
          const int par = ((j >> s0) ^ (j >> s1)) & 1;
          imcj_acc(lv(j), pv(j), par == 0);
        }
        break;
      case GN_M1: {
        const int m = 1 << s0;
        std::string g[8];
        for (int i = 0; i < 8; ++i) g[i] = fmt("ct[%d]", goff + i);
        for (int j = 0; j < nreg; ++j) {
          if (j & m) continue;
          line(fmt("{ const cT x = %s, y = %s;", pv(j).c_str(), pv(j | m).c_str()));
          line(fmt("  const cT u = cT{%s * x.x - %s * x.y + %s * y.x - %s * y.y, %s * x.y + %s * x.x + %s * y.y + %s * y.x};",
                   g[0].c_str(), g[1].c_str(), g[2].c_str(), g[3].c_str(), g[0].c_str(), g[1].c_str(), g[2].c_str(),
                   g[3].c_str()));
          line(fmt("  const cT w = cT{%s * x.x - %s * x.y + %s * y.x - %s * y.y, %s * x.y + %s * x.x + %s * y.y + %s * y.x};",
                   g[4].c_str(), g[5].c_str(), g[6].c_str(), g[7].c_str(), g[4].c_str(), g[5].c_str(), g[6].c_str(),
                   g[7].c_str()));
          line(fmt("  ga += imcj(%s, u) + imcj(%s, w); }", lv(j).c_str(), lv(j | m).c_str()));
        }
      } break;
      case GN_M2: {
        const int m0 = 1 << s0, m1 = 1 << s1;
        for (int j = 0; j < nreg; ++j) {
          if (j & (m0 | m1)) continue;
          const int idx[4] = {j, j | m1, j | m0, j | m0 | m1};
          line("{");
          for (int r = 0; r < 4; ++r) {
            std::string re, im;
            for (int c = 0; c < 4; ++c) {
              const int e = r * 4 + c;
              re += fmt("%sct[%d] * %s.x - ct[%d] * %s.y", c ? " + " : "", goff + 2 * e, pv(idx[c]).c_str(),
                        goff + 2 * e + 1, pv(idx[c]).c_str());
              im += fmt("%sct[%d] * %s.y + ct[%d] * %s.x", c ? " + " : "", goff + 2 * e, pv(idx[c]).c_str(),
                        goff + 2 * e + 1, pv(idx[c]).c_str());
            }
            line(fmt("  ga += imcj(%s, cT{%s, %s});", lv(idx[r]).c_str(), re.c_str(), im.c_str()));
          }
          line("}");
        }
      } break;
      default:
        return false;
    }
    return true;
  }
};

// Gradient partials per trainable gate and block in an adjoint pass: one
// per 16-lane row (fewest instructions) when that LDS fits as many blocks
// per CU as one per wave does, else one per wave.
// LDS slots of gradient partials: per row, gates are stored 16 at a time
// (one lane each), so the count rounds up to a multiple of 16
size_t red_slots(const DPass& dp, int per) {
  const int NT = 1 << dp.tthreads_log;
  const size_t g = (size_t)std::max(1, dp.ngop);
  return (per == NT / 16 ? (g + 15) / 16 * 16 : g) * (size_t)per;
}

int adj_partials(int dt, const DPass& dp) {
  const int NT = 1 << dp.tthreads_log;
  const size_t tile = (size_t)(dt == 0 ? 8 : 16) << dp.k;
  const size_t rb = dt == 0 ? 4 : 8;  // bytes per partial (working precision)
  auto blocks = [&](int per) { return (160u * 1024u) / (tile + rb * red_slots(dp, per)); };
  return blocks(NT / 16) == blocks(NT / 64) ? NT / 16 : NT / 64;
}

bool pass_supported(const Sched& S, const DPass& dp) {
  if (dp.plain_op >= 0 || dp.k <= 0 || dp.mfma) return false;
  if ((1 << dp.tthreads_log) < 64 || dp.tthreads_log > 10) return false;
  for (int g = 0; g < dp.ngroups; ++g) {
    const DGroup& dg = S.dgroup[dp.group0 + g];
    if (dg.lds_op) return false;
    for (int o = dg.op0; o < dg.op0 + dg.nops; ++o) {
      const int k = S.dop[o].kind;
      if (k == OK_UK || (k != OK_DG && (S.dop[o].s0 < 0))) return false;
    }
  }
  return true;
}

// body of one pass kernel (everything after the signature); mbase / dbase
// are the smallest base-recipe / merged-diagonal table offsets it reads
// multi: the forward body as a device function of a multi-program launch:
// block and thread indices come in as `bid` / `tid_` and the LDS tile as the
// pointer `tile` (the kernel shares one LDS buffer among its programs'
// bodies); every block is live (no early return: the caller clamps)
std::string pass_body(const Program& P, const Sched& S, int pi, bool adj, int dt, int* mbase_out, int* dbase_out,
                      bool multi = false) {
  const DPass& dp = S.dpass[pi];
  const HPass& hp = S.passes[pi];
  Gen G(S, dp, hp, adj);
  const int R = S.R, nreg = 1 << R;
  const int k = dp.k, tl = dp.tthreads_log, NT = 1 << tl, nbl = dp.nblk_log;
  const bool wide = P.n > 31;
  const uint32_t cb = dt == 0 ? 8u : 16u;  // bytes per complex amplitude
  // gradient partials: per 16-lane row (rowsum, last lane of each row) or
  // per wave (wsum63, lane 63)
  const int PR = adj ? adj_partials(dt, dp) : 1;
  const bool rows = PR == NT / 16;
  const char* IT = wide ? "unsigned long long" : "unsigned";
  int mb = -1, db = -1;
  for (int g = 0; g < dp.ngroups; ++g) {
    const DGroup& dg = S.dgroup[dp.group0 + g];
    for (int o = dg.op0; o < dg.op0 + dg.nops; ++o) {
      const DOp& op = S.dop[o];
      if (op.moff < 0 || op.kind == OK_X1 || op.kind == OK_CX || op.kind == OK_CZ || op.kind == OK_SWAP) continue;
      int& base = op.moff >= S.mtab_base ? db : mb;
      base = base < 0 ? op.moff : std::min(base, op.moff);
    }
  }
  if (mb < 0) mb = 0;
  if (db < 0) db = S.mtab_base;
  G.mbase = mb;
  G.dbase = db;
  *mbase_out = mb;
  *dbase_out = db;

  // block base and thread / element deposits
  std::vector<int> bs, bd, ts, td;
  for (int i = 0; i < nbl; ++i) {
    bs.push_back(i);
    bd.push_back(dp.blk_bits[i]);
  }
  for (int t = 0; t < tl; ++t) {
    ts.push_back(t);
    td.push_back(hp.tile[t]);
  }
  auto dep_const = [&](uint64_t a) {
    uint64_t o = 0;
    for (int t = 0; t < k; ++t)
      if (a >> t & 1) o |= 1ull << hp.tile[t];
    return o;
  };
  if (!multi) G.line("if ((unsigned long long)blockIdx.x >= (unsigned long long)total) return;");
  G.line(multi ? "const unsigned tid = tid_;" : "const unsigned tid = threadIdx.x;");
  const char* BI = multi ? "bid" : "(unsigned long long)blockIdx.x";
  G.line(fmt("const unsigned long long blk = %s & 0x%llxull;", BI, (nbl >= 64) ? ~0ull : ((1ull << nbl) - 1)));
  G.line(fmt("const unsigned long long b = %s >> %d;", BI, nbl));
  G.line(fmt("const unsigned long long base = %s;", deposit_expr("blk", bs, bd, "unsigned long long").c_str()));
  G.line(fmt("%s dlo = %s;", IT, deposit_expr("tid", ts, td, IT).c_str()));
  G.line("unsigned slo = swz(tid) * (unsigned)sizeof(cT);");
  G.line("const T* __restrict__ mt = mtab + b * mstride + mbase;");
  G.line("const T* __restrict__ md = mtab + b * mstride + dbase;");
  const int nit = 1 << (k - tl);  // elements per thread
  auto hi = [&](int it) { return (unsigned long long)dep_const((uint64_t)it << tl); };

  int opid = 0;
  // gradient partials in emission order: slot e of `red` holds gate
  // gl_of[e].  Per-row partials: every lane of a row gets the row sum
  // (rowall), lane e % 16 keeps it in `gacc`, and each 16 gates are stored
  // with one unpredicated LDS write per lane, so no exec-mask branch splits
  // the straight-line gate code.  Per-wave partials: lane 63 stores.
  std::vector<int> gl_of;
  auto emit_grad = [&](const std::string& v, int gloc) {
    const int e = (int)gl_of.size();
    gl_of.push_back(gloc);
    if (rows) {
      const int j = e & 15;
      if (j == 0) G.line(fmt("gacc = rowall(%s);", v.c_str()));
      else G.line(fmt("{ const T gs = rowall(%s); gacc = lr == %du ? gs : gacc; }", v.c_str(), j));
      if (j == 15) G.line(fmt("red[rw + %du] = gacc;", (e - 15) * PR));
    } else {
      G.line(fmt("{ const T gs = wsum63(%s); if ((tid & 63u) == 63u) red[%d + (tid >> 6)] = gs; }", v.c_str(),
                 e * PR));
    }
  };
  // gradients and inverse gates of one group, in reverse program order, on
  // the register arrays p (psi) and l (lambda)
  auto emit_adj_ops = [&](const DGroup& dg) -> bool {
      for (int o = dg.op0 + dg.nops - 1; o >= dg.op0; --o) {
        const DOp& op = S.dop[o];
        if (op.kind == OK_DG && op.nmem > 0) {
          G.line("{");
          for (int j = 0; j < nreg; ++j) G.line(fmt("const T im%d = imcj(%s, %s);", j, G.lv(j).c_str(), G.pv(j).c_str()));
          // Z-generator sums sum_j (-1)^{bit b of j} im_j for every needed
          // bit b share one butterfly: level b splits the running pair sums
          // into sums (next level) and differences (summed into zs<b>)
          int zneed = 0;
          for (int q = 0; q < op.nmem; ++q)
            if (S.dmem[op.mem0 + q].gen == GN_Z) zneed |= 1 << S.dmem[op.mem0 + q].s0;
          if (zneed) {
            std::vector<std::string> v;
            for (int j = 0; j < nreg; ++j) v.push_back(fmt("im%d", j));
            for (int bit = 0; bit < R && (zneed >> bit); ++bit) {
              std::vector<std::string> nx;
              std::string zs;
              for (size_t j = 0; j + 1 < v.size(); j += 2) {
                if (zneed >> (bit + 1)) {
                  nx.push_back(fmt("bs%d_%d", bit, (int)(j / 2)));
                  G.line(fmt("const T %s = %s + %s;", nx.back().c_str(), v[j].c_str(), v[j + 1].c_str()));
                }
                if (zneed >> bit & 1) zs += fmt("%s(%s - %s)", zs.empty() ? "" : " + ", v[j].c_str(), v[j + 1].c_str());
              }
              if (zneed >> bit & 1) G.line(fmt("const T zs%d = %s;", bit, zs.c_str()));
              v.swap(nx);
            }
          }
          for (int q = 0; q < op.nmem; ++q) {
            const DMember& M = S.dmem[op.mem0 + q];
            if (M.gen == GN_Z) {
              emit_grad(fmt("zs%d", M.s0), M.gloc);
              continue;
            }
            std::string e;
            for (int j = 0; j < nreg; ++j) {
              int sg = 0;
              if (M.gen == GN_Z) sg = (j >> M.s0 & 1) ? -1 : 1;
              else if (M.gen == GN_P1) sg = (j >> M.s0 & 1) ? 1 : 0;
              else if (M.gen == GN_ZZ) sg = (((j >> M.s0) ^ (j >> M.s1)) & 1) ? -1 : 1;
              if (sg) e += fmt(" %s im%d", sg > 0 ? "+" : "-", j);
            }
            if (e.empty()) e = " + 0";
            G.line("{");
            G.line(fmt("const T ga = (T)0%s;", e.c_str()));
            emit_grad("ga", M.gloc);
            G.line("}");
          }
          G.line("}");
        } else if (op.gen != GN_NONE) {
          G.line("{");
          if (!G.grad(op.gen, op.s0, op.s1, op.goff)) return false;
          emit_grad("ga", op.gloc);
          G.line("}");
        }
        G.apply(op, {"p", "l"}, opid++);
        // a gate's gradient reduction may overlap its own inverse, not the
        // next gates: without this fence the scheduler stretches live
        // ranges across gates and spills
        G.line("__builtin_amdgcn_sched_barrier(0);");
      }
    return true;
  };
  if (adj) {
    // Adjoint pass.  One LDS tile shared by psi and lambda: each layout change moves psi
    // through it, then lambda (four barriers instead of one), for half the
    // LDS of two tiles and twice the resident blocks per CU.
    G.line(fmt("__shared__ cT tb[%d];", 1 << k));
    // gradient partials: one per 16-lane row (or wave) and trainable gate, in
    // the working precision; the block sum below is fp64
    G.line(fmt("__shared__ T red[%d];", (int)red_slots(dp, PR)));
    if (rows) {
      G.line("const unsigned lr = tid & 15u;");
      G.line(fmt("const unsigned rw = lr * %du + (tid >> 4);", PR));
      G.line("T gacc = 0;");
    }
    {
      std::string decl = "cT";
      for (int j = 0; j < nreg; ++j) decl += fmt("%s p%d, l%d", j ? "," : "", j, j);
      G.line(decl + ";");
    }
    G.line("const cT* sp = psi_in + b * sstride + base;");
    G.line("const cT* sl = lam_in + b * sstride + base;");
    for (int it = 0; it < nit; ++it) G.line(fmt("const cT rl%d = sl[dlo + %lluull];", it, hi(it)));
    G.line("{");
    for (int it = 0; it < nit; ++it) G.line(fmt("const cT rp%d = sp[dlo + %lluull];", it, hi(it)));
    for (int it = 0; it < nit; ++it) G.line(fmt("LDSB(tb, slo ^ %uu) = rp%d;", swz_h((uint32_t)it << tl) * cb, it));
    G.line("}");
    G.line("__syncthreads();");
    const int ng1 = dp.ngroups;
    int prev = -1;
    std::vector<int> prev_where;
    auto group_addr = [&](const DGroup& dg, int g) {
      std::vector<int> cs, cd;
      for (int r = 0; r < dg.nrun; ++r) {
        const int w = dg.run[r];
        const int src = w & 0xFF, len = (w >> 8) & 0xFF, dst = w >> 16;
        for (int i = 0; i < len; ++i) {
          cs.push_back(src + i);
          cd.push_back(dst + i);
        }
      }
      G.line(fmt("const unsigned ga%d = swz(%s) * (unsigned)sizeof(cT);", g,
                 deposit_expr("tid", cs, cd, "unsigned").c_str()));
    };
    auto write_regs = [&](const char* a, int g, const std::vector<int>& wh) {
      const DGroup& dg = S.dgroup[dp.group0 + g];
      for (int j = 0; j < nreg; ++j) G.line(fmt("LDSB(tb, ga%d ^ %uu) = %s%d;", g, sreg(dg.S[j]) * cb, a, wh[j]));
    };
    auto read_regs = [&](const char* a, int g) {
      const DGroup& dg = S.dgroup[dp.group0 + g];
      for (int j = 0; j < nreg; ++j) G.line(fmt("%s%d = LDSB(tb, ga%d ^ %uu);", a, j, g, sreg(dg.S[j]) * cb));
    };
    for (int gg = 0; gg < ng1; ++gg) {
      const int g = ng1 - 1 - gg;
      const DGroup& dg = S.dgroup[dp.group0 + g];
      group_addr(dg, g);
      if (prev < 0) {
        read_regs("p", g);
        G.line("__syncthreads();");
        for (int it = 0; it < nit; ++it) G.line(fmt("LDSB(tb, slo ^ %uu) = rl%d;", swz_h((uint32_t)it << tl) * cb, it));
        G.line("__syncthreads();");
        read_regs("l", g);
      } else {
        G.line("__syncthreads();");
        write_regs("p", prev, prev_where);
        G.line("__syncthreads();");
        read_regs("p", g);
        G.line("__syncthreads();");
        write_regs("l", prev, prev_where);
        G.line("__syncthreads();");
        read_regs("l", g);
      }
      G.where.assign(nreg, 0);
      for (int j = 0; j < nreg; ++j) G.where[j] = j;
      if (!emit_adj_ops(dg)) return std::string();
      prev = g;
      prev_where = G.where;
    }
    if ((int)gl_of.size() != dp.ngop) return std::string();
    if (rows && (gl_of.size() & 15)) G.line(fmt("red[rw + %du] = gacc;", (int)(gl_of.size() & ~(size_t)15) * PR));
    G.frame_fix({"p", "l"});
    G.line("asm volatile(\"\" : \"+v\"(dlo), \"+v\"(slo));");
    G.line("__syncthreads();");
    write_regs("p", prev, prev_where);
    G.line("__syncthreads();");
    G.line("if (psi_out) { cT* so = psi_out + b * sstride + base;");
    for (int it = 0; it < nit; ++it)
      G.line(fmt("  so[dlo + %lluull] = LDSB(tb, slo ^ %uu);", hi(it), swz_h((uint32_t)it << tl) * cb));
    G.line("}");
    G.line("__syncthreads();");
    write_regs("l", prev, prev_where);
    G.line("__syncthreads();");
    G.line("if (lam_out) { cT* so = lam_out + b * sstride + base;");
    for (int it = 0; it < nit; ++it)
      G.line(fmt("  so[dlo + %lluull] = LDSB(tb, slo ^ %uu);", hi(it), swz_h((uint32_t)it << tl) * cb));
    G.line("}");
    if (dp.ngop > 0) {
      std::string tab;
      for (int q : gl_of) tab += fmt("%s%d", tab.empty() ? "" : ", ", q);
      G.line(fmt("static const unsigned short gl_of[%d] = {%s};", dp.ngop, tab.c_str()));
      G.line(fmt("for (unsigned e = tid; e < %du; e += %du) {", dp.ngop, NT));
      G.line("  double sum = 0.0;");
      G.line(fmt("  for (int w = 0; w < %d; ++w) sum += (double)red[e * %d + w];", PR, PR));
      G.line(fmt("  part[b * pstride + (long long)gl_of[e] * %lldll + (long long)blk] = sum;", 1ll << nbl));
      G.line("}");
    }
    return G.out;
  }
  // forward pass: one tile
  if (!multi) G.line(fmt("__shared__ cT tile[%d];", 1 << k));
  G.line("cT* __restrict__ s = st + b * sstride + base;");
  if (multi) {
    for (int it = 0; it < nit; ++it) G.line(fmt("const cT r%d = s[dlo + %lluull];", it, hi(it)));
  } else {
    // zinit: the pass starts from |0..0> instead of reading the state (the
    // element index base + dlo + hi(it) is 0 only for it = 0, hi(0) = 0)
    for (int it = 0; it < nit; ++it) G.line(fmt("cT r%d;", it));
    G.line("if (zinit) {");
    G.line("  r0 = cT{(base | (unsigned long long)dlo) == 0ull ? (T)1 : (T)0, (T)0};");
    for (int it = 1; it < nit; ++it) G.line(fmt("  r%d = cT{(T)0, (T)0};", it));
    G.line("} else {");
    for (int it = 0; it < nit; ++it) G.line(fmt("  r%d = s[dlo + %lluull];", it, hi(it)));
    G.line("}");
  }
  for (int it = 0; it < nit; ++it) G.line(fmt("LDSB(tile, slo ^ %uu) = r%d;", swz_h((uint32_t)it << tl) * cb, it));
  G.line("__syncthreads();");
  for (int g = 0; g < dp.ngroups; ++g) {
    const DGroup& dg = S.dgroup[dp.group0 + g];
    G.where.assign(nreg, 0);
    for (int j = 0; j < nreg; ++j) G.where[j] = j;
    std::vector<int> cs, cd;
    for (int r = 0; r < dg.nrun; ++r) {
      const int w = dg.run[r];
      const int src = w & 0xFF, len = (w >> 8) & 0xFF, dst = w >> 16;
      for (int i = 0; i < len; ++i) {
        cs.push_back(src + i);
        cd.push_back(dst + i);
      }
    }
    G.line("{");
    G.line(fmt("const unsigned a0 = swz(%s) * (unsigned)sizeof(cT);", deposit_expr("tid", cs, cd, "unsigned").c_str()));
    for (int j = 0; j < nreg; ++j) G.line(fmt("cT v%d = LDSB(tile, a0 ^ %uu);", j, sreg(dg.S[j]) * cb));
    for (int o = dg.op0; o < dg.op0 + dg.nops; ++o) G.apply(S.dop[o], {"v"}, opid++);
    if (g == dp.ngroups - 1) G.frame_fix({"v"});
    G.line("unsigned a1 = a0; asm volatile(\"\" : \"+v\"(a1));");
    for (int j = 0; j < nreg; ++j) G.line(fmt("LDSB(tile, a1 ^ %uu) = v%d;", sreg(dg.S[j]) * cb, G.where[j]));
    G.line("}");
    G.line("__syncthreads();");
  }
  // recompute the store addresses rather than holding the load addresses live
  G.line("asm volatile(\"\" : \"+v\"(dlo), \"+v\"(slo));");
  for (int it = 0; it < nit; ++it) {
    G.line(fmt("{ const cT x = LDSB(tile, slo ^ %uu); s[dlo + %lluull] = x;", swz_h((uint32_t)it << tl) * cb, hi(it)));
    G.line(fmt("  if (ck) ck[b * sstride + base + dlo + %lluull] = x; }", hi(it)));
  }
  return G.out;
}

}  // namespace

std::string gen_pass_source(const Program& P, const Sched& S, int dt, int mode, std::vector<JitPass>& jp,
                            std::vector<std::string>& bodies) {
  jp.assign(S.dpass.size(), JitPass{});
  std::unordered_map<std::string, int> seen;
  std::string src;
  bodies.clear();
  for (int pi = 0; pi < (int)S.dpass.size(); ++pi) {
    const DPass& dp = S.dpass[pi];
    if (!pass_supported(S, dp)) continue;
    int mb = 0, db = 0;
    std::string body = pass_body(P, S, pi, mode == 1, dt, &mb, &db);
    if (body.empty()) continue;
    const int NT = 1 << dp.tthreads_log;
    const std::string key = fmt("%d|", NT) + body;
    auto it = seen.find(key);
    int id;
    if (it == seen.end()) {
      id = (int)bodies.size();
      seen.emplace(key, id);
      std::string k;
      // register target: as many waves per SIMD as the LDS tiles allow
      // (160 KiB per CU), so registers never cut occupancy below that
      const size_t cbytes = dt == 0 ? 8 : 16;
      const size_t lds = (cbytes << dp.k) + (mode == 0 ? 0 : (dt == 0 ? 4 : 8) * red_slots(dp, adj_partials(dt, dp)));
      const int blocks_cu = (int)std::max<size_t>(1, (160u * 1024u) / lds);
      const int waves = std::max(1, std::min(8, blocks_cu * (NT / 64) / 4));
      // parameter lists: jit_launch_fwd / jit_launch_adj pass exactly these
      if (mode == 0)
        k = fmt("extern \"C\" __global__ void __launch_bounds__(%d) __attribute__((amdgpu_waves_per_eu(%d))) "
                "tcm_fwd_k%d(cT* __restrict__ st, long long sstride, "
                "const T* __restrict__ mtab, long long mstride, long long mbase, long long dbase, "
                "const T* __restrict__ ct, cT* __restrict__ ck, long long total, int zinit) {\n",
                NT, waves, id);
      else
        k = fmt("extern \"C\" __global__ void __launch_bounds__(%d) __attribute__((amdgpu_waves_per_eu(%d))) "
                "tcm_adj_k%d(const cT* psi_in, "
                "cT* psi_out, const cT* lam_in, cT* lam_out, "
                "long long sstride, const T* __restrict__ mtab, long long mstride, long long mbase, long long dbase, "
                "const T* __restrict__ ct, double* __restrict__ part, long long pstride, long long total) {\n",
                NT, waves, id);
      bodies.push_back(k + body + "}\n");
    } else {
      id = it->second;
    }
    jp[pi].kernel = id;
    jp[pi].mbase = mb;
    jp[pi].dbase = db;
  }
  src = std::string("#define REAL ") + (dt == 0 ? "float" : "double") + "\n" + kPrelude;
  return src;
}

}  // namespace tcm

namespace tcm {

// ---------------------------------------------------------------------------
// Multi-program forward launches: pass r of every program of a plan in one
// kernel.  Each program's pass body is a device function; a block of NTB
// threads runs F = NTB / NT tiles of one program side by side (NT = that
// pass's threads per tile), each with its own slice of the kernel's LDS,
// all in lockstep through the same barriers.  Block -> program by the
// job table (block0 prefix sums); the tile index of a sub-block past the
// program's last tile is clamped to the last tile, which the same block
// also computes, so the duplicate writes equal values after the last
// barrier and no read races them.
// ---------------------------------------------------------------------------
constexpr size_t kMultiLds = 64u * 1024u;

bool multi_pass_ok(const Sched& S, int r) {
  return r < (int)S.dpass.size() && S.k > 0 && pass_supported(S, S.dpass[r]);
}

std::string gen_multi_prelude(int dt) {
  return std::string("#define REAL ") + (dt == 0 ? "float" : "double") + "\n" + kPrelude +
         "struct TcmMJob { const void* state; long long sstride; const void* mtab; long long mstride, mbase, "
         "dbase; const void* ctab; unsigned long long block0, ntiles; };\n";
}

// the kernel `tcm_multi_k<kid>` for pass r (device functions and kernel; the
// common prelude is gen_multi_prelude); empty when no program has a
// compilable pass r
std::string gen_multi_source(const std::vector<const Program*>& progs, const std::vector<double>& weight, int dt,
                             int r, int kid, MultiPassInfo* info) {
  info->jobs.clear();
  info->ntb = 64;
  info->lds = 0;
  const size_t cb = dt == 0 ? 8 : 16;
  std::vector<std::string> fns;
  std::unordered_map<std::string, int> seen;
  std::string calls;
  // threads per block: the most any included pass needs; a pass joins only
  // if its tiles side by side fit kMultiLds of LDS (two blocks per CU), so
  // one wide pass does not lower the occupancy of the whole launch; the
  // others run as single-program launches
  auto joins = [&](const Sched& S) {
    if (!multi_pass_ok(S, r)) return false;
    const DPass& dp = S.dpass[r];
    return (size_t)(info->ntb >> dp.tthreads_log) * (cb << dp.k) <= kMultiLds;
  };
  // the thread count that lets the most work (weight: amplitudes per
  // program) join
  {
    double best = -1.0;
    int best_ntb = 64;
    for (int ntb = 64; ntb <= 1024; ntb *= 2) {
      info->ntb = ntb;
      double w = 0.0;
      for (size_t i = 0; i < progs.size(); ++i) {
        const Sched& S = progs[i]->sched[dt][0];
        if (joins(S) && (1 << S.dpass[r].tthreads_log) <= ntb) w += weight[i];
      }
      if (w > best) {
        best = w;
        best_ntb = ntb;
      }
    }
    info->ntb = best_ntb;
  }
  for (size_t i = 0; i < progs.size(); ++i) {
    const Program& P = *progs[i];
    const Sched& S = P.sched[dt][0];
    if (!joins(S) || (1 << S.dpass[r].tthreads_log) > info->ntb) continue;
    const DPass& dp = S.dpass[r];
    int mb = 0, db = 0;
    const std::string body = pass_body(P, S, r, false, dt, &mb, &db, true);
    if (body.empty()) continue;
    int id;
    auto it = seen.find(body);
    if (it == seen.end()) {
      id = (int)fns.size();
      seen.emplace(body, id);
      fns.push_back(fmt("__device__ __forceinline__ void tcm_mb%d_%d(cT* __restrict__ st, long long sstride, "
                        "const T* __restrict__ mtab, long long mstride, long long mbase, long long dbase, "
                        "const T* __restrict__ ct, cT* __restrict__ ck, unsigned long long bid, unsigned tid_, "
                        "cT* __restrict__ tile) {\n",
                        kid, id) +
                    body + "}\n");
    } else {
      id = it->second;
    }
    const int NT = 1 << dp.tthreads_log;
    const int F = info->ntb / NT;
    const size_t tile = cb << dp.k;
    info->lds = std::max(info->lds, tile * (size_t)F);
    MultiPassJob j{};
    j.prog = (int)i;
    j.nblk_log = dp.nblk_log;
    j.per_block = F;
    j.mbase = mb;
    j.dbase = db;
    const int q = (int)info->jobs.size();
    info->jobs.push_back(j);
    calls += fmt("    case %d: {\n"
                 "      const unsigned sub = threadIdx.x / %du, t = threadIdx.x %% %du;\n"
                 "      const TcmMJob& jb = J.job[%d];\n"
                 "      unsigned long long tl = (gb - jb.block0) * %dull + sub;\n"
                 "      if (tl >= jb.ntiles) tl = jb.ntiles - 1;\n"
                 "      tcm_mb%d_%d((cT*)jb.state, jb.sstride, (const T*)jb.mtab, jb.mstride, jb.mbase, jb.dbase, "
                 "(const T*)jb.ctab, nullptr, tl, t, reinterpret_cast<cT*>(lds + sub * %zuu));\n"
                 "    } break;\n",
                 q, NT, NT, q, F, kid, id, tile);
  }
  if (info->jobs.empty()) return std::string();
  const int nj = (int)info->jobs.size();
  std::string src = fmt("struct TcmMJobs%d { TcmMJob job[%d]; };\n", kid, nj);
  for (const auto& f : fns) src += f;
  const size_t lds_b = info->lds;
  const int blocks_cu = (int)std::max<size_t>(1, (160u * 1024u) / std::max<size_t>(lds_b, 1));
  const int waves = std::max(1, std::min(8, blocks_cu * (info->ntb / 64) / 4));
  src += fmt("extern \"C\" __global__ void __launch_bounds__(%d) __attribute__((amdgpu_waves_per_eu(%d))) "
             "tcm_multi_k%d(TcmMJobs%d J) {\n"
             "  __shared__ __attribute__((aligned(16))) char lds[%zu];\n"
             "  const unsigned long long gb = blockIdx.x;\n"
             "  int q = 0;\n"
             "  while (q + 1 < %d && gb >= J.job[q + 1].block0) ++q;\n"
             "  switch (q) {\n%s    default: break;\n  }\n}\n",
             info->ntb, waves, kid, kid, lds_b, nj, calls.c_str());
  return src;
}

// ---------------------------------------------------------------------------
// Pauli-sum tile passes, compiled: for every pass of the tiled Pauli schedule
// a kernel that reads psi once into LDS and evaluates h = sum of the pass's
// terms (P psi)[y] = i^ny (-1)^popc((y ^ flip) & yz) psi[y ^ flip] with all
// masks compile-time.  Per term, the sign splits into a block-uniform part
// (bits outside the tile), a per-thread part (thread bits) and a per-element
// constant, so a thread sums each group's weights once per sign pattern and
// spends a few adds per element.  Semantics are those of k_pauli_tile.
// ---------------------------------------------------------------------------
std::string gen_pauli_source(int n, int dt, const std::vector<DPPass>& passes, const std::vector<DPGroupT>& groups,
                             const std::vector<DPTermT>& terms, std::vector<std::string>& bodies) {
  bodies.clear();
  const uint32_t cb = dt == 0 ? 8u : 16u;
  for (size_t pi = 0; pi < passes.size(); ++pi) {
    const DPPass& P = passes[pi];
    const int k = P.k, tl = P.nthr_log, NT = 1 << tl, nbl = P.nblk_log, NW = NT / 64;
    const int nit = 1 << (k - tl);
    // tile bits: the bits not in blk_bits, ascending
    std::vector<int> tile;
    {
      std::vector<char> isb(n, 0);
      for (int i = 0; i < nbl; ++i) isb[P.blk_bits[i]] = 1;
      for (int b = 0; b < n; ++b)
        if (!isb[b]) tile.push_back(b);
    }
    std::vector<int> bs, bd, ts, td;
    for (int i = 0; i < nbl; ++i) {
      bs.push_back(i);
      bd.push_back(P.blk_bits[i]);
    }
    for (int t = 0; t < tl; ++t) {
      ts.push_back(t);
      td.push_back(tile[t]);
    }
    auto hi = [&](int it) {
      unsigned long long o = 0;
      for (int t = tl; t < k; ++t)
        if (((uint64_t)it << tl) >> t & 1) o |= 1ull << tile[t];
      return o;
    };
    std::string o;
    auto L = [&](const std::string& s) { o += s; o += '\n'; };
    L(fmt("extern \"C\" __global__ void __launch_bounds__(%d) tcm_pauli_k%d(const cT* __restrict__ psi, "
          "const cT* lam_in, cT* lam_out, long long sstride, const double* __restrict__ alpha, "
          "double* __restrict__ epart, long long estride, long long total, int flags) {",
          NT, (int)pi));
    L("if ((unsigned long long)blockIdx.x >= (unsigned long long)total) return;");
    L(fmt("__shared__ cT tile[%d];", 1 << k));
    L(fmt("__shared__ double red[%d];", 2 * NW));
    L("const unsigned tid = threadIdx.x;");
    L(fmt("const unsigned long long blk = blockIdx.x & 0x%llxull;", nbl >= 64 ? ~0ull : ((1ull << nbl) - 1)));
    L(fmt("const unsigned long long b = (unsigned long long)blockIdx.x >> %d;", nbl));
    L(fmt("const unsigned long long base = %s;", deposit_expr("blk", bs, bd, "unsigned long long").c_str()));
    L(fmt("const unsigned long long dlo = %s;", deposit_expr("tid", ts, td, "unsigned long long").c_str()));
    L("const unsigned slo = swz(tid) * (unsigned)sizeof(cT);");
    L("const cT* __restrict__ s = psi + b * sstride + base;");
    for (int it = 0; it < nit; ++it) L(fmt("const cT r%d = s[dlo + %lluull];", it, hi(it)));
    for (int it = 0; it < nit; ++it) L(fmt("LDSB(tile, slo ^ %uu) = r%d;", swz_h((uint32_t)it << tl) * cb, it));
    L("__syncthreads();");
    for (int it = 0; it < nit; ++it) L(fmt("T hr%d = 0, hi%d = 0;", it, it));
    const uint32_t tmask = (uint32_t)NT - 1;
    for (int g = P.g0; g < P.g0 + P.ng; ++g) {
      const DPGroupT& G = groups[g];
      const uint32_t flo = G.flip_t & tmask, fhi = G.flip_t >> tl;
      L("{");
      // per-thread weight sums by element sign pattern: pattern -> (re, im)
      std::vector<uint32_t> pats;
      std::vector<std::string> sre, sim;
      for (int t = G.t0; t < G.t0 + G.nt; ++t) {
        const DPTermT& tm = terms[t];
        const uint32_t ylo = tm.yz_t & tmask, pat = tm.yz_t >> tl;
        size_t pidx = 0;
        while (pidx < pats.size() && pats[pidx] != pat) ++pidx;
        if (pidx == pats.size()) {
          pats.push_back(pat);
          sre.push_back("");
          sim.push_back("");
        }
        // sign bit: popc(base & yz_g) + popc((tid ^ flip_lo) & yz_lo)
        std::string sb = fmt("((__builtin_popcountll(base & 0x%llxull) + __builtin_popcount((tid ^ %uu) & %uu)) & 1)",
                             (unsigned long long)tm.yz_g, flo, ylo);
        const double w = tm.w;
        // i^ny * w: 0 -> +re, 1 -> +im, 2 -> -re, 3 -> -im
        const int q = tm.ny & 3;
        const double ww = (q >= 2) ? -w : w;
        const std::string term = fmt("(%s ? (T)(%.17g) : (T)(%.17g))", sb.c_str(), -ww, ww);
        if (q == 0 || q == 2) sre[pidx] += (sre[pidx].empty() ? "" : " + ") + term;
        else sim[pidx] += (sim[pidx].empty() ? "" : " + ") + term;
      }
      for (size_t pidx = 0; pidx < pats.size(); ++pidx)
        L(fmt("const T sr%d = %s, si%d = %s;", (int)pidx, sre[pidx].empty() ? "(T)0" : sre[pidx].c_str(), (int)pidx,
              sim[pidx].empty() ? "(T)0" : sim[pidx].c_str()));
      if (G.flip_t) L(fmt("const unsigned pa = swz(tid ^ %uu) * (unsigned)sizeof(cT);", flo));
      for (int it = 0; it < nit; ++it) {
        std::string cr, ci;
        for (size_t pidx = 0; pidx < pats.size(); ++pidx) {
          const bool neg = __builtin_popcount(((uint32_t)it ^ fhi) & pats[pidx]) & 1;
          cr += fmt(" %s sr%d", neg ? "-" : "+", (int)pidx);
          ci += fmt(" %s si%d", neg ? "-" : "+", (int)pidx);
        }
        std::string qv = G.flip_t ? fmt("LDSB(tile, pa ^ %uu)", swz_h(((uint32_t)it ^ fhi) << tl) * cb)
                                  : fmt("r%d", it);
        L(fmt("{ const cT q = %s; const T cr = (T)0%s, ci = (T)0%s;", qv.c_str(), cr.c_str(), ci.c_str()));
        L(fmt("  hr%d += cr * q.x - ci * q.y; hi%d += cr * q.y + ci * q.x; }", it, it));
      }
      L("}");
    }
    // energy partials and lambda update
    L("if (flags & 1) {");
    L("  double er = 0.0, ei = 0.0;");
    for (int it = 0; it < nit; ++it) {
      L(fmt("  er += (double)r%d.x * hr%d + (double)r%d.y * hi%d;", it, it, it, it));
      L(fmt("  ei += (double)r%d.x * hi%d - (double)r%d.y * hr%d;", it, it, it, it));
    }
    L("  er = wsum(er); ei = wsum(ei);");
    L("  if ((tid & 63u) == 0) { red[2 * (tid >> 6)] = er; red[2 * (tid >> 6) + 1] = ei; }");
    L("}");
    L("__syncthreads();");
    L("if ((flags & 1) && tid == 0) {");
    L("  double a = 0.0, c = 0.0;");
    L(fmt("  for (int w = 0; w < %d; ++w) { a += red[2 * w]; c += red[2 * w + 1]; }", NW));
    L(fmt("  epart[b * estride + 2 * (%d + blk)] = a;", P.slot0));
    L(fmt("  epart[b * estride + 2 * (%d + blk) + 1] = c;", P.slot0));
    L("}");
    L("if (flags & 2) {");
    L("  const double al = alpha ? alpha[b] : 1.0;");
    L("  const unsigned long long ob = b * sstride + base + dlo;");
    for (int it = 0; it < nit; ++it) {
      L(fmt("  { const unsigned long long off = ob + %lluull;", hi(it)));
      L("    const cT o = lam_in ? lam_in[off] : cT{(T)0, (T)0};");
      L(fmt("    lam_out[off] = cT{(T)(o.x + al * hr%d), (T)(o.y + al * hi%d)}; }", it, it));
    }
    L("}");
    L("}");
    bodies.push_back(o);
  }
  return std::string("#define REAL ") + (dt == 0 ? "float" : "double") + "\n" + kPrelude;
}

}  // namespace tcm
