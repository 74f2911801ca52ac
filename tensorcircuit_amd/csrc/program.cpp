// Host side of libtcm: lowering of the gate list (the reference's QIR,
// basecircuit.py:131-142) into device ops, and the scheduler that packs gates
// into LDS-tiled passes (several gates per HBM read+write of the state) and
// register groups inside each pass.
//
// The reference contracts a tensor network with an opt_einsum path
// (cons.py:523-663); a dense statevector engine instead needs an order of
// gate application and a memory schedule.  Gates are applied in program
// order except that gates on disjoint qubits may be reordered (they commute),
// which is what the greedy clustering below exploits.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <functional>
#include <sstream>

#include "tcm_internal.h"

namespace tcm {

namespace {

inline int popc64(uint64_t x) { return __builtin_popcountll(x); }

bool finite_all(const double* m, size_t cnt) {
  for (size_t i = 0; i < cnt; ++i)
    if (!std::isfinite(m[i])) return false;
  return true;
}

// complex matrix helpers on interleaved (re, im) row-major d x d matrices
bool is_diag(const std::vector<double>& m, int d, double tol = 0.0) {
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c)
      if (r != c && (std::fabs(m[2 * (r * d + c)]) > tol || std::fabs(m[2 * (r * d + c) + 1]) > tol))
        return false;
  return true;
}
bool is_real(const std::vector<double>& m) {
  for (size_t i = 1; i < m.size(); i += 2)
    if (m[i] != 0.0) return false;
  return true;
}
bool equals(const std::vector<double>& m, const double* ref, int cnt, double tol = 0.0) {
  for (int i = 0; i < cnt; ++i)
    if (std::fabs(m[i] - ref[i]) > tol) return false;
  return true;
}

// reference matrices (interleaved complex)
const double kX[8] = {0, 0, 1, 0, 1, 0, 0, 0};
const double kY[8] = {0, 0, 0, -1, 0, 1, 0, 0};
const double kZ[8] = {1, 0, 0, 0, 0, 0, -1, 0};
double kCX[32], kCZ[32], kSWAP[32], kXX[32], kYY[32], kZZ[32];

void kron2(const double* a, const double* b, double* out) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      int ra = r >> 1, rb = r & 1, ca = c >> 1, cb = c & 1;
      double ar = a[2 * (ra * 2 + ca)], ai = a[2 * (ra * 2 + ca) + 1];
      double br = b[2 * (rb * 2 + cb)], bi = b[2 * (rb * 2 + cb) + 1];
      out[2 * (r * 4 + c)] = ar * br - ai * bi;
      out[2 * (r * 4 + c) + 1] = ar * bi + ai * br;
    }
}

struct RefInit {
  RefInit() {
    std::memset(kCX, 0, sizeof kCX);
    std::memset(kCZ, 0, sizeof kCZ);
    std::memset(kSWAP, 0, sizeof kSWAP);
    auto set = [](double* m, int r, int c, double v) { m[2 * (r * 4 + c)] = v; };
    set(kCX, 0, 0, 1); set(kCX, 1, 1, 1); set(kCX, 2, 3, 1); set(kCX, 3, 2, 1);
    set(kCZ, 0, 0, 1); set(kCZ, 1, 1, 1); set(kCZ, 2, 2, 1); set(kCZ, 3, 3, -1);
    set(kSWAP, 0, 0, 1); set(kSWAP, 1, 2, 1); set(kSWAP, 2, 1, 1); set(kSWAP, 3, 3, 1);
    kron2(kX, kX, kXX);
    kron2(kY, kY, kYY);
    kron2(kZ, kZ, kZZ);
  }
} ref_init;

// G^2 == I check for exp1 generators
// a * b for interleaved complex d x d matrices
std::vector<double> matmul_c(const std::vector<double>& a, const std::vector<double>& b, int d) {
  std::vector<double> out(2 * d * d, 0.0);
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c) {
      double sr = 0, si = 0;
      for (int k = 0; k < d; ++k) {
        const double ar = a[2 * (r * d + k)], ai = a[2 * (r * d + k) + 1];
        const double br = b[2 * (k * d + c)], bi = b[2 * (k * d + c) + 1];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      out[2 * (r * d + c)] = sr;
      out[2 * (r * d + c) + 1] = si;
    }
  return out;
}
bool close_all(const std::vector<double>& a, const std::vector<double>& b, double tol) {
  for (size_t i = 0; i < a.size(); ++i)
    if (std::fabs(a[i] - b[i]) > tol) return false;
  return true;
}
bool is_hermitian(const std::vector<double>& m, int d) {
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c)
      if (std::fabs(m[2 * (r * d + c)] - m[2 * (c * d + r)]) > 1e-9 ||
          std::fabs(m[2 * (r * d + c) + 1] + m[2 * (c * d + r) + 1]) > 1e-9)
        return false;
  return true;
}

bool is_involution(const std::vector<double>& g, int d) {
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c) {
      double sr = 0, si = 0;
      for (int k = 0; k < d; ++k) {
        double ar = g[2 * (r * d + k)], ai = g[2 * (r * d + k) + 1];
        double br = g[2 * (k * d + c)], bi = g[2 * (k * d + c) + 1];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      double er = (r == c) ? 1.0 : 0.0;
      if (std::fabs(sr - er) > 1e-9 || std::fabs(si) > 1e-9) return false;
    }
  return true;
}

int layout_size(int layout) {
  switch (layout) {
    case OK_U1: return 8;
    case OK_R1: return 4;
    case OK_D1: return 4;
    case OK_U2: return 32;
    case OK_D2: return 8;
    default: return 0;
  }
}

// Clustering with commutation by disjointness.  Each cluster takes items in
// program order: an item joins if its bits lie inside the cluster's bit set
// (the `fixed` bits plus at most `cap` free bits) and every earlier item
// sharing a bit with it joined too (an item that stays behind holds its
// bits for everything after it).  Items that stay behind keep their
// relative order.  `solo` items form a cluster of their own.
//
// Choice of the free-bit set: first fit (the bits of the items in the order
// they come, while they fit), then hill climbing (swap one bit in for one
// out, or add one while below `cap`, whenever more items join).  With
// `climb` set, the climbed set and its best neighbours are further compared
// by a rollout of the next few clusters under the climbing rule, and the
// candidate that finishes the items in fewer clusters (else covers more
// items within the rollout) wins.  On a layered circuit whose two-qubit
// gates form a chain (hardware-efficient ansatz) first fit stalls on the
// chain's next link, climbing alone grabs the widest window and strands the
// lagging layers, and the rollout keeps the staircase of windows that moves
// every layer forward together.
struct ClusterScorer {
  int cap = 0;
  uint64_t fixed = 0;
  // current item list (positions 0..G-1), its masks and predecessor lists
  std::vector<int> rem;
  std::vector<uint64_t> mask;
  std::vector<char> solo;
  std::vector<int32_t> pred_ptr, pred;
  std::vector<uint64_t> acc;

  void set_items(const std::vector<int>& items, const std::vector<uint64_t>& gmask,
                 const std::vector<char>& gsolo) {
    rem = items;
    const size_t G = rem.size();
    mask.resize(G);
    solo.resize(G);
    pred_ptr.assign(G + 1, 0);
    pred.clear();
    int last[64];
    for (int b = 0; b < 64; ++b) last[b] = -1;
    for (size_t i = 0; i < G; ++i) {
      mask[i] = gmask[rem[i]];
      solo[i] = gsolo[rem[i]];
      for (uint64_t t = mask[i]; t; t &= t - 1) {
        const int b = __builtin_ctzll(t);
        if (last[b] >= 0 && std::find(pred.begin() + pred_ptr[i], pred.end(), last[b]) == pred.end())
          pred.push_back(last[b]);
        last[b] = (int)i;
      }
      pred_ptr[i + 1] = (int32_t)pred.size();
    }
    acc.resize(G);
  }

  // counts[c] = number of items taken by free-bit set sets[c], c < ns <= 64
  void score(const uint64_t* sets, int ns, int* counts) {
    uint64_t W[64];
    const uint64_t all = ns >= 64 ? ~0ull : ((1ull << ns) - 1);
    for (int b = 0; b < 64; ++b) W[b] = (fixed >> b & 1) ? all : 0;
    for (int c = 0; c < ns; ++c)
      for (uint64_t t = sets[c]; t; t &= t - 1) W[__builtin_ctzll(t)] |= 1ull << c;
    for (int c = 0; c < ns; ++c) counts[c] = 0;
    for (size_t i = 0; i < rem.size(); ++i) {
      uint64_t a = solo[i] ? 0 : all;
      for (uint64_t t = mask[i]; t && a; t &= t - 1) a &= W[__builtin_ctzll(t)];
      for (int32_t p = pred_ptr[i]; p < pred_ptr[i + 1] && a; ++p) a &= acc[pred[p]];
      acc[i] = a;
      for (; a; a &= a - 1) ++counts[__builtin_ctzll(a)];
    }
  }

  int count(uint64_t S) {
    int c;
    score(&S, 1, &c);
    return c;
  }

  uint64_t first_fit() const {
    uint64_t S = 0, blocked = 0;
    for (size_t i = 0; i < rem.size(); ++i) {
      const uint64_t m = mask[i];
      if (solo[i] || (m & blocked)) {
        blocked |= m;
        continue;
      }
      const uint64_t ns = S | (m & ~fixed);
      if (popc64(ns) <= cap) S = ns;
      else blocked |= m;
    }
    return S;
  }

  // hill climbing from S; `near` receives the final neighbourhood (count, set)
  uint64_t climb(uint64_t S, int* best_out, std::vector<std::pair<int, uint64_t>>* near) {
    uint64_t cand = 0;
    for (uint64_t m : mask) cand |= m;
    cand &= ~fixed;
    int best = count(S);
    std::vector<uint64_t> nb;
    std::vector<int> cnt;
    for (int iter = 0; iter < 256; ++iter) {
      nb.clear();
      for (uint64_t outs = cand & ~S; outs; outs &= outs - 1) {
        const uint64_t c = outs & (~outs + 1);
        if (popc64(S) < cap) nb.push_back(S | c);
        for (uint64_t ins = S; ins; ins &= ins - 1) nb.push_back((S & ~(ins & (~ins + 1))) | c);
      }
      cnt.assign(nb.size(), 0);
      for (size_t o = 0; o < nb.size(); o += 64)
        score(nb.data() + o, (int)std::min<size_t>(64, nb.size() - o), cnt.data() + o);
      size_t bi = nb.size();
      int bc = best;
      for (size_t j = 0; j < nb.size(); ++j)
        if (cnt[j] > bc) bc = cnt[j], bi = j;
      if (bi == nb.size()) {
        if (near) {
          near->clear();
          for (size_t j = 0; j < nb.size(); ++j) near->push_back({cnt[j], nb[j]});
        }
        break;
      }
      best = bc;
      S = nb[bi];
    }
    *best_out = best;
    return S;
  }

  // split the current items into those set S takes and the rest
  void split(uint64_t S, std::vector<int>& taken, std::vector<int>& rest) const {
    const uint64_t inside = S | fixed;
    uint64_t held = 0;
    taken.clear();
    rest.clear();
    for (size_t i = 0; i < rem.size(); ++i) {
      const uint64_t m = mask[i];
      if (!solo[i] && !(m & held) && !(m & ~inside)) taken.push_back(rem[i]);
      else {
        rest.push_back(rem[i]);
        held |= m;
      }
    }
  }
};

std::vector<std::vector<int>> cluster(const std::vector<int>& items,
                                      const std::function<uint64_t(int)>& mask,
                                      uint64_t fixed, int cap,
                                      const std::function<bool(int)>& solo,
                                      std::vector<uint64_t>* used_out, bool climb, int exact_max) {
  int maxid = 0;
  for (int it : items) maxid = std::max(maxid, it);
  std::vector<uint64_t> gmask(items.empty() ? 0 : maxid + 1, 0);
  std::vector<char> gsolo(gmask.size(), 0);
  for (int it : items) {
    gmask[it] = mask(it);
    gsolo[it] = solo(it) ? 1 : 0;
  }
  ClusterScorer sc;
  sc.cap = cap;
  sc.fixed = fixed;
  // a partial schedule: the clusters so far and the items left
  struct Node {
    std::vector<std::vector<int>> clusters;
    std::vector<uint64_t> sets;
    std::vector<int> rest;
    int covered = 0, score = 0;
    uint64_t next = 0;                             // climbed set for `rest`
    std::vector<std::pair<int, uint64_t>> near;    // its neighbourhood
  };
  // climbed set (and neighbourhood) for the first cluster of `n.rest`
  auto prepare = [&](Node& n) {
    n.score = n.covered;
    if (n.rest.empty()) return;
    sc.set_items(n.rest, gmask, gsolo);
    if (sc.solo[0]) {
      n.score += 1;
      return;
    }
    int best = 0;
    n.next = sc.first_fit();
    if (climb) n.next = sc.climb(n.next, &best, &n.near);
    else best = sc.count(n.next);
    n.score += best;
  };
  // beam width and candidates per node (pure first fit / climbing when
  // the item list is long)
  int beam = 8, ncand = 8;
  if (!climb || items.size() > 3000) beam = ncand = 1;
  std::vector<std::vector<int>> shorter_out;
  std::vector<Node> front(1);
  front[0].rest = items;
  prepare(front[0]);
  std::vector<int> taken, rest;
  // exact search for a schedule one cluster shorter than the beam's (of up
  // to exact_max clusters), over the climbed set, its best neighbours and
  // every window of `cap` consecutive free bits (a chain-structured
  // circuit's staircase windows are such windows); bounded by a node budget
  // (4000 for a target of up to 5 clusters, else 4000 per cluster: a
  // 26-qubit 10-layer HEA goes from 11 to 10 passes in under a second)
  bool any_solo = false;
  for (int it : items) any_solo |= gsolo[it] != 0;
  auto shorter = [&](const Node& done) -> bool {
    const int target = (int)done.clusters.size() - 1;
    if (!climb || any_solo || target < 1 || target > exact_max || items.size() > 4000) return false;
    uint64_t allbits = 0;
    for (int it : items) allbits |= gmask[it];
    allbits &= ~fixed;
    std::vector<int> fb;
    for (int b = 0; b < 64; ++b)
      if (allbits >> b & 1) fb.push_back(b);
    std::vector<uint64_t> windows;
    for (size_t i = 0; i + cap <= fb.size(); ++i) {
      uint64_t w = 0;
      for (int j = 0; j < cap; ++j) w |= 1ull << fb[i + j];
      windows.push_back(w);
    }
    int budget = target <= 5 ? 4000 : 4000 * target;
    std::vector<std::vector<int>> cl;
    std::vector<uint64_t> sets;
    std::function<bool(const std::vector<int>&, int)> dfs = [&](const std::vector<int>& rem, int left) -> bool {
      if (rem.empty()) return true;
      if (left == 0 || --budget < 0) return false;
      ClusterScorer loc;
      loc.cap = cap;
      loc.fixed = fixed;
      loc.set_items(rem, gmask, gsolo);
      uint64_t u = 0;
      for (int it : rem) u |= gmask[it];
      u &= ~fixed;
      std::vector<uint64_t> cands;
      if (popc64(u) <= cap) {
        cands.push_back(u);  // one cluster takes everything left
      } else {
        if (left == 1) return false;
        int best = 0;
        std::vector<std::pair<int, uint64_t>> near;
        const uint64_t c0 = loc.climb(loc.first_fit(), &best, &near);
        std::vector<std::pair<int, uint64_t>> all(near);
        all.push_back({best, c0});
        std::vector<int> wc(windows.size(), 0);
        for (size_t o = 0; o < windows.size(); o += 64)
          loc.score(windows.data() + o, (int)std::min<size_t>(64, windows.size() - o), wc.data() + o);
        for (size_t i = 0; i < windows.size(); ++i) all.push_back({wc[i], windows[i]});
        std::sort(all.begin(), all.end(), [](const std::pair<int, uint64_t>& a, const std::pair<int, uint64_t>& b) {
          return a.first > b.first || (a.first == b.first && a.second < b.second);
        });
        for (const auto& pr : all) {
          if (pr.first <= 0) break;
          if (std::find(cands.begin(), cands.end(), pr.second) == cands.end()) cands.push_back(pr.second);
          if (cands.size() >= 24) break;
        }
      }
      std::vector<int> tk, rs;
      for (uint64_t c : cands) {
        loc.split(c, tk, rs);
        if (tk.empty()) continue;
        uint64_t usedb = 0;
        for (int it : tk) usedb |= gmask[it];
        cl.push_back(tk);
        sets.push_back(c & usedb);
        if (dfs(rs, left - 1)) return true;
        cl.pop_back();
        sets.pop_back();
      }
      return false;
    };
    if (!dfs(items, target)) return false;
    if (used_out) *used_out = sets;
    shorter_out = cl;
    return true;
  };
  while (true) {
    for (const Node& n : front)
      if (n.rest.empty()) {
        if (shorter(n)) return shorter_out;
        if (used_out) *used_out = n.sets;
        return n.clusters;
      }
    std::vector<Node> kids;
    for (const Node& n : front) {
      sc.set_items(n.rest, gmask, gsolo);
      std::vector<uint64_t> cands;
      if (sc.solo[0]) {
        cands.push_back(~0ull);  // marker: the solo item alone
      } else {
        cands.push_back(n.next);
        std::vector<std::pair<int, uint64_t>> near = n.near;
        std::sort(near.begin(), near.end(),
                  [](const std::pair<int, uint64_t>& a, const std::pair<int, uint64_t>& b) {
                    return a.first > b.first || (a.first == b.first && a.second < b.second);
                  });
        for (const auto& pr : near) {
          if ((int)cands.size() >= ncand) break;
          if (pr.first > 0) cands.push_back(pr.second);
        }
      }
      for (uint64_t c : cands) {
        Node k;
        k.clusters = n.clusters;
        k.sets = n.sets;
        uint64_t S;
        if (c == ~0ull) {
          taken.assign(1, n.rest[0]);
          rest.assign(n.rest.begin() + 1, n.rest.end());
          S = gmask[n.rest[0]] & ~fixed;
        } else {
          sc.split(c, taken, rest);
          uint64_t usedb = 0;
          for (int it : taken) usedb |= gmask[it];
          S = c & usedb;
        }
        k.clusters.push_back(taken);
        k.sets.push_back(S);
        k.rest = rest;
        k.covered = n.covered + (int)taken.size();
        kids.push_back(std::move(k));
      }
    }
    for (Node& k : kids) prepare(k);
    // best `beam` distinct children: score = items covered plus what the
    // next climbed cluster would take
    std::stable_sort(kids.begin(), kids.end(), [](const Node& a, const Node& b) {
      if (a.rest.empty() != b.rest.empty()) return a.rest.empty();
      if (a.score != b.score) return a.score > b.score;
      return a.covered > b.covered;
    });
    front.clear();
    for (Node& k : kids) {
      bool dup = false;
      for (const Node& f : front)
        if (f.rest == k.rest) dup = true;
      if (dup) continue;
      front.push_back(std::move(k));
      if ((int)front.size() >= beam) break;
    }
  }
}

uint64_t gate_mask(const LGate& g) {
  uint64_t m = 0;
  for (int i = 0; i < g.nq; ++i) m |= (1ull << g.bits[i]);
  return m;
}

bool diag_kind(const LGate& g) {
  return (g.okind == OK_D1 || g.okind == OK_D2 || g.okind == OK_CZ) &&
         (g.gen == GN_NONE || g.gen == GN_Z || g.gen == GN_ZZ || g.gen == GN_P1);
}

}  // namespace

// ---------------------------------------------------------------------------
// lowering
// ---------------------------------------------------------------------------
static int lower(Program& P, const tcm_gate* gates, int32_t ngates) {
  const int n = P.n;
  int layer = 0;
  int mslot = 0;
  for (int gi = 0; gi < ngates; ++gi) {
    const tcm_gate& g = gates[gi];
    if (g.kind == TCM_G_BARRIER) {
      ++layer;
      continue;
    }
    LGate L{};
    L.src = g.kind;
    L.nq = g.nq;
    L.param = g.param;
    L.theta = g.theta;
    L.scale = g.scale;
    L.mslot = L.cslot = L.goff = L.gop = -1;
    L.gen = GN_NONE;
    L.gfac = 0.0;
    L.layer = layer;
    L.barrier = false;
    if (g.nq < 1 || g.nq > 5)
      return set_error(TCM_E_ARG, "gate %d: nq=%d out of range 1..5", gi, g.nq);
    for (int i = 0; i < g.nq; ++i) {
      if (g.q[i] < 0 || g.q[i] >= n)
        return set_error(TCM_E_QUBIT, "gate %d: qubit %d out of range [0,%d)", gi, g.q[i], n);
      for (int j = 0; j < i; ++j)
        if (g.q[j] == g.q[i])
          return set_error(TCM_E_DUPQUBIT, "gate %d: repeated qubit %d", gi, g.q[i]);
      L.q[i] = g.q[i];
      L.bits[i] = n - 1 - g.q[i];
    }
    if (g.param >= P.nparams)
      return set_error(TCM_E_ARG, "gate %d: param %d >= nparams %d", gi, g.param, P.nparams);
    if (g.param < 0 && !std::isfinite(g.theta))
      return set_error(TCM_E_NONFINITE, "gate %d: non-finite angle", gi);
    if (g.param >= 0 && !std::isfinite(g.scale))
      return set_error(TCM_E_NONFINITE, "gate %d: non-finite scale", gi);
    const int d = 1 << g.nq;
    const int msz = 2 * d * d;
    double c = 0.0;  // generator coefficient: U = exp(-i c a G)
    switch (g.kind) {
      case TCM_G_MAT: {
        if (!g.mat) return set_error(TCM_E_ARG, "gate %d: MAT without matrix", gi);
        if (!finite_all(g.mat, msz)) return set_error(TCM_E_NONFINITE, "gate %d: non-finite matrix", gi);
        L.mat.assign(g.mat, g.mat + msz);
        L.param = -1;
        if (g.nq == 1) {
          if (equals(L.mat, kX, 8)) L.okind = OK_X1;
          else if (is_diag(L.mat, 2)) L.okind = OK_D1;
          else if (is_real(L.mat)) L.okind = OK_R1;
          else L.okind = OK_U1;
        } else if (g.nq == 2) {
          if (equals(L.mat, kCX, 32)) L.okind = OK_CX;
          else if (equals(L.mat, kCZ, 32)) L.okind = OK_CZ;
          else if (equals(L.mat, kSWAP, 32)) L.okind = OK_SWAP;
          else if (is_diag(L.mat, 4)) L.okind = OK_D2;
          else L.okind = OK_U2;
        } else {
          L.okind = OK_UK;
        }
        // constant table entry in the op's layout
        std::vector<double> e;
        switch (L.okind) {
          case OK_D1: e = {L.mat[0], L.mat[1], L.mat[6], L.mat[7]}; break;
          case OK_R1: e = {L.mat[0], L.mat[2], L.mat[4], L.mat[6]}; break;
          case OK_D2:
            for (int r = 0; r < 4; ++r) {
              e.push_back(L.mat[2 * (r * 4 + r)]);
              e.push_back(L.mat[2 * (r * 4 + r) + 1]);
            }
            break;
          case OK_U1: case OK_U2: case OK_UK: e = L.mat; break;
          default: break;
        }
        if (!e.empty()) {
          L.cslot = (int)P.ctab.size();
          P.ctab.insert(P.ctab.end(), e.begin(), e.end());
        }
        break;
      }
      case TCM_G_RX: case TCM_G_RY: case TCM_G_RZ:
        if (g.nq != 1) return set_error(TCM_E_ARG, "gate %d: rotation needs 1 qubit", gi);
        L.okind = g.kind == TCM_G_RX ? OK_U1 : (g.kind == TCM_G_RY ? OK_R1 : OK_D1);
        L.gen = g.kind == TCM_G_RX ? GN_X : (g.kind == TCM_G_RY ? GN_Y : GN_Z);
        c = 0.5;
        break;
      case TCM_G_RXX: case TCM_G_RYY: case TCM_G_RZZ:
        if (g.nq != 2) return set_error(TCM_E_ARG, "gate %d: 2-qubit rotation needs 2 qubits", gi);
        L.okind = g.kind == TCM_G_RZZ ? OK_D2 : OK_U2;
        L.gen = g.kind == TCM_G_RXX ? GN_XX : (g.kind == TCM_G_RYY ? GN_YY : GN_ZZ);
        c = 0.5;
        break;
      case TCM_G_PHASE:
        if (g.nq != 1) return set_error(TCM_E_ARG, "gate %d: phase needs 1 qubit", gi);
        L.okind = OK_D1;
        L.gen = GN_P1;
        c = -1.0;
        break;
      case TCM_G_EXP1: {
        if (g.nq > 2) return set_error(TCM_E_UNSUPPORTED, "gate %d: exp1 on %d qubits", gi, g.nq);
        if (!g.mat) return set_error(TCM_E_ARG, "gate %d: EXP1 without generator", gi);
        if (!finite_all(g.mat, msz)) return set_error(TCM_E_NONFINITE, "gate %d: non-finite generator", gi);
        L.mat.assign(g.mat, g.mat + msz);
        bool dg = is_diag(L.mat, d);
        L.okind = g.nq == 1 ? (dg ? OK_D1 : OK_U1) : (dg ? OK_D2 : OK_U2);
        if (g.nq == 1) {
          L.gen = equals(L.mat, kX, 8) ? GN_X : equals(L.mat, kY, 8) ? GN_Y
                : equals(L.mat, kZ, 8) ? GN_Z : GN_M1;
        } else {
          L.gen = equals(L.mat, kXX, 32) ? GN_XX : equals(L.mat, kYY, 32) ? GN_YY
                : equals(L.mat, kZZ, 32) ? GN_ZZ : GN_M2;
        }
        L.goff = (int)P.ctab.size();
        P.ctab.insert(P.ctab.end(), L.mat.begin(), L.mat.end());
        if (g.param >= 0 && !is_involution(L.mat, d))
          return set_error(TCM_E_UNSUPPORTED,
                           "gate %d: trainable exp1 needs G^2 = I for an exact gradient", gi);
        c = 1.0;
        break;
      }
      case TCM_G_EXPG: {
        // exp(-i a G), G Hermitian with G^3 = G; the constant table holds G
        // then Q = G^2 (the matrix is (I - Q) + cos(a) Q - i sin(a) G and
        // dU/da = -i G U, so G is the gradient generator)
        if (g.nq > 2) return set_error(TCM_E_UNSUPPORTED, "gate %d: generator gate on %d qubits", gi, g.nq);
        if (!g.mat) return set_error(TCM_E_ARG, "gate %d: EXPG without generator", gi);
        if (!finite_all(g.mat, msz)) return set_error(TCM_E_NONFINITE, "gate %d: non-finite generator", gi);
        L.mat.assign(g.mat, g.mat + msz);
        std::vector<double> q = matmul_c(L.mat, L.mat, d);
        if (!is_hermitian(L.mat, d) || !close_all(matmul_c(q, L.mat, d), L.mat, 1e-9))
          return set_error(TCM_E_ARG, "gate %d: EXPG generator must be Hermitian with G^3 = G", gi);
        const bool dg = is_diag(L.mat, d);
        L.okind = g.nq == 1 ? (dg ? OK_D1 : OK_U1) : (dg ? OK_D2 : OK_U2);
        L.gen = g.nq == 1 ? GN_M1 : GN_M2;
        L.goff = (int)P.ctab.size();
        P.ctab.insert(P.ctab.end(), L.mat.begin(), L.mat.end());
        P.ctab.insert(P.ctab.end(), q.begin(), q.end());
        c = 1.0;
        break;
      }
      default:
        return set_error(TCM_E_ARG, "gate %d: unknown kind %d", gi, g.kind);
    }
    if (g.kind != TCM_G_MAT) {
      L.mslot = mslot;
      mslot += L.src == TCM_G_RY ? kRySlot : layout_size(L.okind);
      if (L.param >= 0) {
        L.gop = P.ngop++;
        L.gfac = 2.0 * c * L.scale;
        P.gop_gate.push_back((int)P.gates.size());
        P.gfac.push_back(L.gfac);
      } else {
        L.gen = GN_NONE;
      }
    }
    P.gates.push_back(std::move(L));
  }
  P.nlayers = layer;
  // parameter -> gops CSR
  P.p_ptr.assign(P.nparams + 1, 0);
  for (int gop = 0; gop < P.ngop; ++gop) P.p_ptr[P.gates[P.gop_gate[gop]].param + 1]++;
  for (int p = 0; p < P.nparams; ++p) P.p_ptr[p + 1] += P.p_ptr[p];
  P.p_gops.assign(P.ngop, 0);
  std::vector<int> fill(P.p_ptr.begin(), P.p_ptr.end() - 1);
  for (int gop = 0; gop < P.ngop; ++gop) P.p_gops[fill[P.gates[P.gop_gate[gop]].param]++] = gop;
  // checkpoints: after every ckpt_every-th barrier (not after the last gate)
  const int ce = P.opts.ckpt_every;
  if (ce > 0) {
    int nb = 0;
    int count = 0;
    for (int gi = 0; gi < ngates; ++gi) {
      if (gates[gi].kind == TCM_G_BARRIER) {
        ++nb;
        if (nb % ce == 0 && count < (int)P.gates.size() &&
            (P.ckpt_gate.empty() || P.ckpt_gate.back() != count))
          P.ckpt_gate.push_back(count);
      } else {
        ++count;
      }
    }
  }
  return 0;
}

// base recipes (one per non-constant gate), shared by all schedules
static void base_recipes(const Program& P, std::vector<DRecipe>& out, int* size) {
  int sz = 0;
  for (const LGate& g : P.gates) {
    if (g.mslot < 0) continue;
    DRecipe r{};
    r.kind = g.src;
    r.param = g.param;
    r.out = g.mslot;
    r.layout = g.okind;
    r.goff = g.goff;
    r.theta = g.theta;
    r.scale = g.scale;
    out.push_back(r);
    sz = std::max(sz, g.mslot + (g.src == TCM_G_RY ? kRySlot : layout_size(g.okind)));
  }
  *size = sz;
}

// ---------------------------------------------------------------------------
// schedules
// ---------------------------------------------------------------------------
static void sched_plain(const Program& P, Sched& S) {
  S.k = 0;
  base_recipes(P, S.recipes, &S.mtab_size);
  S.mtab_base = S.mtab_size;
  size_t ci = 0;
  for (int gi = 0; gi < (int)P.gates.size(); ++gi) {
    while (ci < P.ckpt_gate.size() && P.ckpt_gate[ci] == gi) {
      S.ckpt_after.push_back(gi - 1);
      ++ci;
    }
    const LGate& g = P.gates[gi];
    HPass hp;
    hp.gates.push_back(gi);
    S.passes.push_back(hp);
    DOp op{};
    op.kind = g.okind;
    op.s0 = g.bits[0];
    op.s1 = g.nq > 1 ? g.bits[1] : -1;
    op.moff = g.mslot >= 0 ? g.mslot : (g.cslot >= 0 ? -g.cslot - 1 : 0);
    op.gen = g.gen;
    op.goff = g.goff;
    op.gop = g.gop;
    op.gloc = g.gop >= 0 ? 0 : -1;
    if (g.okind == OK_UK) {
      op.nmem = g.nq;
      op.mem0 = (int)S.ukbits.size();
      for (int i = 0; i < g.nq; ++i) S.ukbits.push_back(g.bits[i]);
      if (g.nq == 5 && P.opts.mfma) op.flags |= OP_F_MFMA;
    }
    S.dop.push_back(op);
  }
  while (ci < P.ckpt_gate.size()) {
    S.ckpt_after.push_back((int)P.gates.size() - 1);
    ++ci;
  }
}

static void build_group_ops(const Program& P, Sched& S, const HGroup& hg,
                            const std::vector<int>& local_of_bit,
                            const std::vector<int>& gloc_of_gop, DGroup& dg) {
  // slot of a tile-local bit inside this group
  auto slot_of = [&](int global_bit) {
    int lb = local_of_bit[global_bit];
    for (size_t s = 0; s < hg.slots.size(); ++s)
      if (hg.slots[s] == lb) return (int)s;
    return -1;
  };
  dg.op0 = (int)S.dop.size();
  const auto& gl = hg.gates;
  size_t i = 0;
  while (i < gl.size()) {
    // maximal run of mergeable diagonal gates
    size_t j = i;
    while (j < gl.size() && diag_kind(P.gates[gl[j]])) ++j;
    if (j - i >= 2) {
      DRecipe r{};
      r.kind = -1;
      r.out = S.mtab_size;
      r.layout = OK_DG;
      r.goff = (int)S.factors.size();
      r.nfac = (int)(j - i);
      r.R = S.R;
      S.mtab_size += 2 << S.R;
      DOp op{};
      op.kind = OK_DG;
      op.moff = r.out;
      op.gop = -1;
      op.gen = GN_NONE;
      op.mem0 = (int)S.dmem.size();
      for (size_t t = i; t < j; ++t) {
        const LGate& g = P.gates[gl[t]];
        DFactor f{};
        f.src = g.mslot >= 0 ? g.mslot : (g.cslot >= 0 ? -g.cslot - 1 : 0);
        f.layout = g.okind;
        f.s0 = slot_of(g.bits[0]);
        f.s1 = g.nq > 1 ? slot_of(g.bits[1]) : -1;
        S.factors.push_back(f);
        if (g.gop >= 0) {
          DMember m{};
          m.gen = g.gen;
          m.s0 = f.s0;
          m.s1 = f.s1;
          m.gop = g.gop;
          m.gloc = gloc_of_gop[g.gop];
          S.dmem.push_back(m);
        }
      }
      op.nmem = (int)S.dmem.size() - op.mem0;
      S.recipes.push_back(r);
      S.dop.push_back(op);
      i = j;
      continue;
    }
    const LGate& g = P.gates[gl[i]];
    DOp op{};
    op.kind = g.okind;
    op.moff = g.mslot >= 0 ? g.mslot : (g.cslot >= 0 ? -g.cslot - 1 : 0);
    op.gen = g.gen;
    if (g.src == TCM_G_RY && g.okind == OK_R1 && g.mslot >= 0) op.flags |= OP_F_SHEAR;
    op.goff = g.goff;
    op.gop = g.gop;
    op.gloc = g.gop >= 0 ? gloc_of_gop[g.gop] : -1;
    if (g.okind == OK_UK) {
      op.nmem = g.nq;
      op.s0 = op.s1 = -1;
    } else {
      op.s0 = slot_of(g.bits[0]);
      op.s1 = g.nq > 1 ? slot_of(g.bits[1]) : -1;
    }
    S.dop.push_back(op);
    ++i;
  }
  dg.nops = (int)S.dop.size() - dg.op0;
}

static void tile_run(const Program& P, Sched& S, const std::vector<int>& items, uint64_t lmask, int cap) {
  const int n = P.n;
  const int k = S.k, R = S.R;
  std::vector<std::vector<int>> clusters;
  std::vector<uint64_t> used;
  if (n <= k) {
    clusters.push_back(items);
    used.push_back(0);
  } else {
    // the longer exact search (schedules of up to 12 passes) where a pass
    // moves a state of 2^22 amplitudes or more per item; 5 elsewhere, where
    // the host time would outweigh a pass
    clusters = cluster(
        items, [&](int gi) { return gate_mask(P.gates[gi]); }, lmask, cap,
        [](int) { return false; }, &used, !P.opts.greedy, n >= 22 ? 12 : 5);
  }
  for (size_t ci = 0; ci < clusters.size(); ++ci) {
    HPass hp;
    hp.gates = clusters[ci];
    uint64_t tmask;
    if (n <= k) {
      tmask = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    } else {
      tmask = lmask | used[ci];
      // pad with the lowest free bits (longer contiguous runs)
      for (int b = 0; b < n && popc64(tmask) < k; ++b) tmask |= (1ull << b);
    }
    for (int b = 0; b < n; ++b)
      if (tmask >> b & 1) hp.tile.push_back(b);
    int L = 0;
    while (L < (int)hp.tile.size() && hp.tile[L] == L) ++L;
    hp.L = L;
    // tile-local index of each global bit
    std::vector<int> local_of_bit(n, -1);
    for (int i = 0; i < (int)hp.tile.size(); ++i) local_of_bit[hp.tile[i]] = i;
    auto lmask_of = [&](int gi) {
      const LGate& g = P.gates[gi];
      uint64_t m = 0;
      for (int i = 0; i < g.nq; ++i) m |= 1ull << local_of_bit[g.bits[i]];
      return m;
    };
    std::vector<uint64_t> gused;
    auto groups = cluster(
        hp.gates, lmask_of, 0, R, [&](int gi) { return P.gates[gi].nq >= 3; }, &gused, !P.opts.greedy, 5);
    for (size_t q = 0; q < groups.size(); ++q) {
      HGroup hg;
      hg.gates = groups[q];
      hg.lds_op = groups[q].size() == 1 && P.gates[groups[q][0]].nq >= 3;
      if (hg.lds_op) {
        const LGate& g = P.gates[groups[q][0]];
        for (int i = 0; i < g.nq; ++i) hg.slots.push_back(local_of_bit[g.bits[i]]);
      } else {
        uint64_t m = gused[q];
        for (int b = 0; b < k; ++b)
          if (m >> b & 1) hg.slots.push_back(b);
        // pad with the highest unused tile bits (keeps low bits on lanes)
        for (int b = k - 1; b >= 0 && (int)hg.slots.size() < R; --b)
          if (!(m >> b & 1)) {
            hg.slots.push_back(b);
            m |= 1ull << b;
          }
      }
      hp.groups.push_back(hg);
    }
    S.passes.push_back(hp);
  }
}

// Consecutive dense 5-qubit gates (matrix-core blocks) grouped into tile
// passes: each pass stages a tile of kM bits (the low Lwant index bits plus
// the blocks' qubits) and applies its blocks in order from LDS.
static void mfma_tile_run(const Program& P, Sched& S, const std::vector<int>& items, uint64_t lmask, int kM) {
  const int n = P.n;
  const int cap = kM - popc64(lmask);
  std::vector<uint64_t> used;
  auto clusters = cluster(
      items, [&](int gi) { return gate_mask(P.gates[gi]); }, lmask, cap, [](int) { return false; }, &used,
      !P.opts.greedy, 5);
  for (size_t ci = 0; ci < clusters.size(); ++ci) {
    HPass hp;
    hp.mfma = true;
    uint64_t tmask = lmask | used[ci];
    for (int b = 0; b < n && popc64(tmask) < kM; ++b) tmask |= (1ull << b);
    for (int b = 0; b < n; ++b)
      if (tmask >> b & 1) hp.tile.push_back(b);
    int L = 0;
    while (L < (int)hp.tile.size() && hp.tile[L] == L) ++L;
    hp.L = L;
    // at most kMfmaTileMaxOps blocks per pass (their matrices share the LDS)
    const std::vector<int>& c = clusters[ci];
    for (size_t g0 = 0; g0 < c.size(); g0 += kMfmaTileMaxOps) {
      hp.gates.assign(c.begin() + g0, c.begin() + std::min(c.size(), g0 + (size_t)kMfmaTileMaxOps));
      S.passes.push_back(hp);
    }
  }
}

static void sched_fused(const Program& P, Sched& S, int k, int R, int Lwant, bool mfma_tiles) {
  const int n = P.n;
  k = std::min(k, n);
  S.k = k;
  S.R = R;
  S.Lwant = std::min(Lwant, k);
  base_recipes(P, S.recipes, &S.mtab_size);
  S.mtab_base = S.mtab_size;
  const uint64_t lmask = (S.Lwant >= 64) ? ~0ull : ((1ull << S.Lwant) - 1);
  const int cap = k - S.Lwant;
  // segments between checkpoints
  std::vector<int> bounds;
  bounds.push_back(0);
  for (int c : P.ckpt_gate) bounds.push_back(c);
  bounds.push_back((int)P.gates.size());
  for (size_t sgi = 0; sgi + 1 < bounds.size(); ++sgi) {
    int g0 = bounds[sgi], g1 = bounds[sgi + 1];
    if (sgi > 0) S.ckpt_after.push_back((int)S.passes.size() - 1);
    if (g1 <= g0) continue;
    // runs of <=2-qubit gates are tiled; a >=3-qubit gate is a plain pass,
    // except runs of dense 5-qubit gates on matrix cores (complex64), which
    // share tile passes of 2^13 amplitudes
    std::vector<int> run, mrun;
    const int kM = std::min(P.opts.tile_bits > 0 ? std::max(10, std::min(P.opts.tile_bits, 13)) : 13, n);
    auto flush = [&]() {
      if (!run.empty()) tile_run(P, S, run, lmask, cap);
      run.clear();
    };
    auto mflush = [&]() {
      if (!mrun.empty()) mfma_tile_run(P, S, mrun, lmask, kM);
      mrun.clear();
    };
    for (int gi = g0; gi < g1; ++gi) {
      const LGate& g = P.gates[gi];
      if (mfma_tiles && g.nq == 5 && g.okind == OK_UK && g.cslot >= 0 && n >= kM && kM >= 10) {
        flush();
        mrun.push_back(gi);
      } else if (g.nq >= 3) {
        flush();
        mflush();
        HPass hp;
        hp.gates.push_back(gi);
        S.passes.push_back(hp);
      } else {
        mflush();
        run.push_back(gi);
      }
    }
    flush();
    mflush();
  }
  // flatten to device descriptors
  for (HPass& hp : S.passes) {
    DPass dp{};
    dp.plain_op = -1;
    if (hp.mfma) {
      dp.k = (int)hp.tile.size();
      dp.L = hp.L;
      dp.nblk_log = n - dp.k;
      dp.tthreads_log = 9;
      dp.mfma = 1;
      dp.gop_list0 = (int)S.gop_list.size();
      dp.runoff0 = (int)S.runoff.size();
      int nb = 0;
      for (int b = 0; b < n; ++b)
        if (std::find(hp.tile.begin(), hp.tile.end(), b) == hp.tile.end()) dp.blk_bits[nb++] = b;
      for (int r = 0; r < (1 << (dp.k - dp.L)); ++r) {
        uint32_t off = 0;
        for (int i = 0; i < dp.k - dp.L; ++i)
          if (r >> i & 1) off |= 1u << hp.tile[dp.L + i];
        S.runoff.push_back(off);
      }
      std::vector<int> local_of_bit(n, -1);
      for (int i = 0; i < dp.k; ++i) local_of_bit[hp.tile[i]] = i;
      dp.mop0 = (int)S.dop.size();
      for (int gi : hp.gates) {
        const LGate& g = P.gates[gi];
        DOp op{};
        op.kind = OK_UK;
        op.s0 = op.s1 = -1;
        op.moff = -g.cslot - 1;
        op.gen = GN_NONE;
        op.gop = -1;
        op.gloc = -1;
        op.nmem = 5;
        op.mem0 = (int)S.ukbits.size();
        op.flags = OP_F_MFMA;
        for (int i = 0; i < 5; ++i) S.ukbits.push_back(local_of_bit[g.bits[i]]);
        S.dop.push_back(op);
      }
      dp.nmop = (int)hp.gates.size();
      S.dpass.push_back(dp);
      continue;
    }
    if (hp.tile.empty()) {
      // a >=3-qubit gate applied by the plain kernel (no gradient)
      const LGate& g = P.gates[hp.gates[0]];
      DOp op{};
      op.kind = g.okind;
      op.s0 = op.s1 = -1;
      op.moff = g.cslot >= 0 ? -g.cslot - 1 : 0;
      op.gen = GN_NONE;
      op.gop = -1;
      op.gloc = -1;
      op.nmem = g.nq;
      op.mem0 = (int)S.ukbits.size();
      for (int i = 0; i < g.nq; ++i) S.ukbits.push_back(g.bits[i]);
      if (g.nq == 5 && P.opts.mfma) op.flags |= OP_F_MFMA;
      dp.plain_op = (int)S.dop.size();
      S.dop.push_back(op);
      dp.gop_list0 = (int)S.gop_list.size();
      S.dpass.push_back(dp);
      continue;
    }
    dp.k = (int)hp.tile.size();
    dp.L = hp.L;
    dp.nblk_log = n - dp.k;
    dp.tthreads_log = dp.k - R;
    dp.group0 = (int)S.dgroup.size();
    dp.ngroups = (int)hp.groups.size();
    dp.runoff0 = (int)S.runoff.size();
    int nb = 0;
    for (int b = 0; b < n; ++b)
      if (std::find(hp.tile.begin(), hp.tile.end(), b) == hp.tile.end()) dp.blk_bits[nb++] = b;
    const int nruns = 1 << (dp.k - dp.L);
    for (int r = 0; r < nruns; ++r) {
      uint32_t off = 0;
      for (int i = 0; i < dp.k - dp.L; ++i)
        if (r >> i & 1) off |= 1u << hp.tile[dp.L + i];
      S.runoff.push_back(off);
    }
    std::vector<int> local_of_bit(n, -1);
    for (int i = 0; i < dp.k; ++i) local_of_bit[hp.tile[i]] = i;
    int ngop = 0;
    std::vector<int> gloc_of_gop(P.ngop, -1);
    dp.gop_list0 = (int)S.gop_list.size();
    for (int gi : hp.gates)
      if (P.gates[gi].gop >= 0) {
        gloc_of_gop[P.gates[gi].gop] = ngop++;
        S.gop_list.push_back(P.gates[gi].gop);
      }
    dp.ngop = ngop;
    S.max_ngop_pass = std::max(S.max_ngop_pass, ngop);
    for (const HGroup& hg : hp.groups) {
      DGroup dg{};
      std::memset(&dg, 0, sizeof dg);
      uint64_t sm = 0;
      for (int s : hg.slots) sm |= 1ull << s;
      int nc = 0;
      for (int b = 0; b < dp.k; ++b)
        if (!(sm >> b & 1)) dg.cb[nc++] = b;
      // thread bits in runs of consecutive tile bits (one bit-field move each)
      for (int i = 0; i < nc;) {
        int len = 1;
        while (i + len < nc && dg.cb[i + len] == dg.cb[i] + len) ++len;
        dg.run[dg.nrun++] = i | len << 8 | dg.cb[i] << 16;
        i += len;
      }
      if (hg.lds_op) {
        dg.lds_op = 1;
        for (size_t i = 0; i < hg.slots.size(); ++i) dg.ukbits[i] = hg.slots[i];
      } else {
        for (int j = 0; j < (1 << R); ++j) {
          uint32_t t = 0;
          for (int s = 0; s < R; ++s)
            if (j >> s & 1) t |= 1u << hg.slots[s];
          dg.S[j] = (int32_t)swz_host(t);
        }
      }
      build_group_ops(P, S, hg, local_of_bit, gloc_of_gop, dg);
      for (int o = dg.op0; o < dg.op0 + dg.nops; ++o)
        if (S.dop[o].kind == OK_U2 || S.dop[o].gen == GN_M2) dp.full = 1;
      S.dgroup.push_back(dg);
    }
    S.dpass.push_back(dp);
  }
}

static void finish_sched(const Program& P, Sched& S) {
  // gradient partial slots per gop
  S.gop_slots.assign(P.ngop, 1);
  S.max_slots = 1;
  if (S.k == 0) {
    for (int gi = 0; gi < (int)P.gates.size(); ++gi) {
      const LGate& g = P.gates[gi];
      if (g.gop < 0) continue;
      int64_t items = (1ll << P.n) >> g.nq;
      int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(512, items / 256));
      S.gop_slots[g.gop] = (int)blocks;
      S.max_slots = std::max(S.max_slots, (int)blocks);
    }
    S.max_ngop_pass = P.ngop > 0 ? 1 : 0;
  } else {
    int slots = 1 << (P.n - S.k);
    for (int g = 0; g < P.ngop; ++g) S.gop_slots[g] = slots;
    S.max_slots = slots;
  }
}

int build_program(const tcm_gate* gates, int32_t ngates, int32_t n, int32_t nparams,
                  const tcm_opts* opts, Program** out) {
  if (!out) return set_error(TCM_E_ARG, "out is NULL");
  *out = nullptr;
  if (n < 1 || n > 32) return set_error(TCM_E_ARG, "nqubits=%d out of range 1..32", n);
  if (ngates < 0 || (ngates > 0 && !gates)) return set_error(TCM_E_ARG, "bad gate list");
  if (nparams < 0) return set_error(TCM_E_ARG, "nparams < 0");
  Program* P = new Program();
  P->n = n;
  P->nparams = nparams;
  if (opts) P->opts = *opts;
  int rc = lower(*P, gates, ngates);
  if (rc) {
    delete P;
    return rc;
  }
  const tcm_opts& o = P->opts;
  // generic kernels are built for 4 register-resident qubits; compiled
  // passes take 3 or 4
  if (o.reg_slots != 0 && o.reg_slots != 4 && !(o.jit && o.reg_slots >= 3 && o.reg_slots <= 5)) {
    delete P;
    return set_error(TCM_E_UNSUPPORTED, "reg_slots=%d: 4 register-resident qubits (3 to 5 with jit)", o.reg_slots);
  }
  for (int dt = 0; dt < 2; ++dt) {
    // register-resident qubits: 5 for compiled complex64 passes (32
    // amplitudes per thread: fewer LDS exchanges per gate; measured faster
    // on the 24-qubit HEA despite 2 waves per SIMD in the adjoint), else 4
    // 5 register slots for compiled complex64 passes on states of >= 20
    // qubits (24-qubit HEA batch: 8 % fewer VALU instructions, 1.5 % faster);
    // 4 otherwise (1024 x 16-qubit batch: 4.80 ms with 4, 5.59 ms with 5)
    const int R = o.reg_slots ? o.reg_slots : ((o.jit && dt == 0 && n >= 20) ? 5 : 4);
    const int Lw = dt == 0 ? 4 : 3;
    // defaults measured on MI355X: 32 KiB forward tiles (several blocks per
    // CU overlap HBM and compute); the generic adjoint holds two tiles.
    // Compiled passes (jit) hold fewer registers and take 64 KiB tiles, and
    // their adjoint stages psi and lambda through one tile: fewer passes at
    // about the same per-pass bandwidth (24-qubit HEA: 44 + 44 passes
    // instead of 56 + 68)
    int kf = o.tile_bits > 0 ? o.tile_bits : (dt == 0 ? 12 : 11) + (o.jit ? 1 : 0);
    int ka = o.adj_tile_bits > 0 ? o.adj_tile_bits : (o.jit ? (dt == 0 ? 13 : 12) : (dt == 0 ? 11 : 10));
    // a tile must fit one CU's 160 KiB of LDS with room for the adjoint's
    // partial sums: 2^14 complex64 or 2^13 complex128 amplitudes (128 KiB)
    const int kmax = dt == 0 ? 14 : 13;
    kf = std::min(kf, kmax);
    ka = std::min(ka, kmax);
    // fused passes need at least one full wave (64 threads) per tile
    bool fused = o.fuse != 0 && n >= R + 6;
    for (int mode = 0; mode < 2; ++mode) {
      Sched& S = P->sched[dt][mode];
      int k = mode == 0 ? kf : ka;
      if (fused && std::min(k, n) >= R + 6 && k >= std::min(n, Lw + 5)) {
        sched_fused(*P, S, k, R, Lw, dt == 0 && mode == 0 && o.mfma == 2);
      } else {
        sched_plain(*P, S);
      }
      finish_sched(*P, S);
    }
  }
  *out = P;
  return 0;
}

static void json_sched(std::ostringstream& os, const Program& P, const Sched& S) {
  os << "{\"k\":" << S.k << ",\"R\":" << S.R << ",\"mtab\":" << S.mtab_size
     << ",\"ckpt_after\":[";
  for (size_t i = 0; i < S.ckpt_after.size(); ++i) os << (i ? "," : "") << S.ckpt_after[i];
  os << "],\"passes\":[";
  for (size_t pi = 0; pi < S.passes.size(); ++pi) {
    const HPass& hp = S.passes[pi];
    os << (pi ? "," : "") << "{\"tile\":[";
    for (size_t i = 0; i < hp.tile.size(); ++i) os << (i ? "," : "") << hp.tile[i];
    os << "],\"L\":" << hp.L << ",\"mfma\":" << (hp.mfma ? 1 : 0) << ",\"gates\":[";
    for (size_t i = 0; i < hp.gates.size(); ++i) os << (i ? "," : "") << hp.gates[i];
    os << "],\"groups\":[";
    for (size_t q = 0; q < hp.groups.size(); ++q) {
      const HGroup& hg = hp.groups[q];
      os << (q ? "," : "") << "{\"slots\":[";
      for (size_t i = 0; i < hg.slots.size(); ++i) os << (i ? "," : "") << hg.slots[i];
      os << "],\"lds\":" << (hg.lds_op ? 1 : 0) << ",\"gates\":[";
      for (size_t i = 0; i < hg.gates.size(); ++i) os << (i ? "," : "") << hg.gates[i];
      os << "]}";
    }
    os << "]}";
  }
  os << "],\"ndpass\":" << S.dpass.size() << ",\"ndg\":";
  int ndg = 0;
  for (const DOp& d : S.dop) ndg += d.kind == OK_DG;
  os << ndg << "}";
  (void)P;
}

std::string describe_sched(const Program& P, int dtype, int mode) {
  std::ostringstream os;
  os << "{\"n\":" << P.n << ",\"ngates\":" << P.gates.size() << ",\"ngop\":" << P.ngop
     << ",\"nckpt\":" << P.ckpt_gate.size() << ",\"sched\":";
  json_sched(os, P, P.sched[dtype][mode]);
  os << "}";
  return os.str();
}

}  // namespace tcm
