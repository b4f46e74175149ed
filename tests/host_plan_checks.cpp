// Host-side checks of the device schedule's planning code (hsrq_sched.cpp),
// built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_plan.py and run on the CPU, no GPU involved.  Random problem
// sets and context settings; for each:
//   * lane assignment: every shift of every problem in exactly one piece,
//     pieces of a problem contiguous, tiles multiples of 64 in [64, 256];
//   * descriptors and workspace layout: every array inside the problem's
//     workspace, 256-byte aligned, no two arrays overlapping;
//   * start steps: each problem's steps inside the lane's;
//   * item tables over all steps: each (tile column, shift group) sweep item
//     once, each (tile column, row block above the next diagonal tile,
//     column block) panel-pass item once and nothing else, row blocks inside
//     their tile rows, init and back-transform items covering every shift.
// Prints "ok <cases>" and exits 0, or names the failed check and exits 1.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <tuple>
#include <vector>

#include "../hsrq3in_amd/csrc/hsrq_host.h"

using namespace hsrq;

static int fails = 0;
#define CHECK(c)                                                              \
  do {                                                                        \
    if (!(c)) {                                                               \
      if (fails++ < 20) std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
    }                                                                         \
  } while (0)

struct Range {
  uintptr_t lo, hi;
};

static void check_layout(const ProbDev& P, const char* ws, size_t bytes) {
  const int n = P.n, m = P.m, T = P.T, nrb = P.nrb;
  const size_t nm = (size_t)n * m * 8;
  std::vector<Range> r;
  auto add = [&](const void* p, size_t sz) {
    const uintptr_t a = (uintptr_t)p;
    CHECK(a % kAlign == 0);
    CHECK(a >= (uintptr_t)ws && a + sz <= (uintptr_t)ws + bytes);
    if (sz > 0) r.push_back({a, a + sz});   // (empty arrays may share an address)
  };
  add(P.Rt[0], nm);
  add(P.Rt[1], nm);
  add(P.C, nm);
  add(P.S, nm);
  add(P.Bm[0], (size_t)P.b * 2 * m * 8);
  add(P.Bm[1], (size_t)P.b * 2 * m * 8);
  add(P.aux[0], (size_t)4 * m * 8);
  add(P.aux[1], (size_t)4 * m * 8);
  add(P.Ynb[0], (size_t)nrb * m * 8);
  add(P.Ynb[1], (size_t)nrb * m * 8);
  add(P.rowinf, (size_t)T * 8);
  add(P.diagn, (size_t)T * 8);
  add(P.part, (size_t)nrb * n * 8);
  add(P.scal, 4 * 8);
  add(P.Ep[0], (size_t)T * m * 4);
  add(P.Ep[1], (size_t)T * m * 4);
  add(P.Ef, (size_t)T * m * 4);
  add(P.status, (size_t)std::max(m, 1) * 4);
  std::sort(r.begin(), r.end(), [](const Range& a, const Range& b) { return a.lo < b.lo; });
  for (size_t i = 1; i < r.size(); ++i) CHECK(r[i].lo >= r[i - 1].hi);
}

static void check_partition(const ProbDev& P) {
  CHECK(P.T >= 1 && P.b0 >= 1 && P.b0 <= P.b);
  CHECK(P.b0 + (P.T - 1) * P.b == P.n);
  int rows = 0;
  for (int rb = 0; rb < P.nrb; ++rb) {
    int t, r0, nr;
    rowblock_range(P, rb, t, r0, nr);
    CHECK(nr >= 1 && nr <= kRowBlock);
    CHECK(r0 == rows);
    CHECK(t >= 0 && t < P.T);
    CHECK(r0 >= tile_first_row(P, t) && r0 + nr <= tile_first_row(P, t) + tile_rows(P, t));
    CHECK(rb >= tile_first_rowblock(P, t) &&
          rb < tile_first_rowblock(P, t) + tile_num_rowblocks(P, t));
    rows += nr;
  }
  CHECK(rows == P.n);
}

int main(int argc, char** argv) {
  const int cases = argc > 1 ? std::atoi(argv[1]) : 200;
  std::mt19937 rng(12345);
  static double dummy[8];
  for (int c = 0; c < cases; ++c) {
    hsrq_context ctx;   // only the planning fields are read
    const int tiles[] = {0, 0, 64, 128, 192, 256};
    ctx.tile = tiles[rng() % 6];
    ctx.lanes = (int)(rng() % 5);
    ctx.pack = (int)((rng() % 3) == 0 ? 4 : (rng() % 2));
    const int count = (rng() % 4 == 0) ? 1 + (int)(rng() % 300) : 1 + (int)(rng() % 9);
    const int tlsz = (rng() % 3 == 0) ? 4 * (1 + (int)(rng() % 80)) : 0;
    std::vector<hsrq_problem> in(count);
    for (auto& q : in) {
      // batches of many problems: smaller orders and shift counts
      const int nmax = count > 9 ? 1500 : 3000;
      const int n = (rng() % 5 == 0) ? 1 + (int)(rng() % 70) : 1 + (int)(rng() % nmax);
      const int mmax = count > 9 ? std::max(1, n / 5) : n;
      const int m = (rng() % 7 == 0) ? 0 : 1 + (int)(rng() % mmax);
      q = hsrq_problem{n, dummy, n + (int)(rng() % 3), m, dummy, dummy, n};
      CHECK(validate_problem(q) == HSRQ_OK);
    }
    std::vector<std::vector<Piece>> groups;
    std::vector<int> tl;
    assign_lanes(&ctx, count, in.data(), tlsz, groups, tl);
    CHECK(groups.size() == tl.size());
    // every shift once, pieces of a problem contiguous
    std::vector<std::vector<std::pair<int, int>>> cover(count);
    for (size_t g = 0; g < groups.size(); ++g) {
      CHECK(tl[g] % 64 == 0 && tl[g] >= 64 && tl[g] <= 256);
      for (const Piece& p : groups[g]) {
        CHECK(p.prob >= 0 && p.prob < count && p.k0 >= 0 && p.m >= 0);
        cover[p.prob].push_back({p.k0, p.k0 + p.m});
      }
    }
    for (int i = 0; i < count; ++i) {
      auto& v = cover[i];
      std::sort(v.begin(), v.end());
      int at = 0;
      for (auto& x : v) {
        CHECK(x.first == at);
        at = x.second;
      }
      CHECK(at == in[i].m);
    }
    // per lane: descriptors, layout, start steps, item tables
    for (size_t g = 0; g < groups.size(); ++g) {
      const int b = tl[g];
      std::vector<ProbDev> probs;
      // the descriptors only compute addresses: a workspace base that is
      // never dereferenced (no allocation)
      char* const base = reinterpret_cast<char*>((uintptr_t)1 << 40);
      for (const Piece& p : groups[g]) {
        hsrq_problem q = in[p.prob];
        q.m = p.m;
        const size_t bytes = prob_workspace_bytes(q.n, q.m, b);
        ProbDev P;
        prob_describe(P, q, b, -1, 0, base, nullptr);
        check_partition(P);
        check_layout(P, base, bytes);
        probs.push_back(P);
      }
      int steps = 0;
      for (const ProbDev& P : probs)
        if (P.m > 0) steps = std::max(steps, P.T);
      const std::vector<int> off = step_offsets(probs, steps);
      CHECK(off.size() == probs.size());
      const int pack = choose_pack(ctx.pack, b, 1 << 20);
      std::set<std::tuple<int, int, int>> k1seen;
      std::set<std::tuple<int, int, int, int>> gbseen;
      for (int s = 0; s < steps; ++s) {
        StepItems st;
        for (size_t i = 0; i < probs.size(); ++i)
          add_step_items(probs[i], (int)i, probs[i].T - 1 - (s - off[i]), pack, st);
        for (const K1Item& it : st.k1) {
          const ProbDev& P = probs[it.prob];
          CHECK(it.J >= 0 && it.J < P.T && it.k0 >= 0 && it.k0 < P.m);
          CHECK(P.T - 1 - it.J == s - off[it.prob]);
          CHECK(k1seen.insert({it.prob, it.J, it.k0}).second);
        }
        for (const GemmItem& it : st.gb) {
          const ProbDev& P = probs[it.prob];
          CHECK(it.J >= 1 && it.J < P.T);
          CHECK(it.rb >= 0 && it.rb + it.nrbk <= tile_first_rowblock(P, it.J - 1));
          CHECK(it.cb >= 0 && it.cb * kGemmCols < 2 * P.m);
          CHECK(gbseen.insert({it.prob, it.J, it.rb, it.cb}).second);
        }
      }
      const int k1step = pack == 4 ? 16 : kShiftsPerK1;
      size_t k1want = 0, gbwant = 0;
      for (size_t i = 0; i < probs.size(); ++i) {
        const ProbDev& P = probs[i];
        if (P.m == 0) continue;
        CHECK(off[i] >= 0 && off[i] + P.T <= steps);
        k1want += (size_t)P.T * ((P.m + k1step - 1) / k1step);
        const int ncb = (2 * P.m + kGemmCols - 1) / kGemmCols;
        for (int J = 2; J < P.T; ++J) gbwant += (size_t)tile_first_rowblock(P, J - 1) * ncb;
      }
      CHECK(k1seen.size() == k1want);
      CHECK(gbseen.size() == gbwant);
      // init and back-transform items cover every shift of every problem
      std::vector<RowItem> init;
      std::vector<ShiftItem> bt;
      for (size_t i = 0; i < probs.size(); ++i) {
        add_init_items(probs[i], (int)i, init);
        add_bt_items(probs[i], (int)i, bt);
      }
      size_t want_bt = 0, want_init = 0;
      for (const ProbDev& P : probs) {
        want_bt += P.m;
        want_init += (size_t)P.nrb * ((P.m + kShiftsPerK1 - 1) / kShiftsPerK1);
      }
      CHECK(bt.size() == want_bt && init.size() == want_init);
      for (const ShiftItem& it : bt) CHECK(it.k >= 0 && it.k < probs[it.prob].m);
    }
  }
  if (fails) {
    std::fprintf(stderr, "%d checks failed\n", fails);
    return 1;
  }
  std::printf("ok %d\n", cases);
  return 0;
}
