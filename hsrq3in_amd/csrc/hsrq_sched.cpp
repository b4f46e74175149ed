// hsrq_sched.cpp -- per-problem workspace layout, descriptors and the item
// tables of the device schedule, shared by plans and the request server.
//
// Tiling follows the reference's partition (partition.c:7-17): T = ceil(n/b)
// tiles with the FIRST tile ragged, b0 = n - (T-1) b.  A problem's sweep runs
// over tile columns J = T-1 ... 0; step J consists of the diagonal sweep of
// tile J (all shifts), the scaling exponents of the panel pass, and the panel
// pass over rows 0 .. l_J - 1 (two parts: the rows of tile J-1, which gate
// the next sweep, and all rows above).
#include <algorithm>
#include <cstring>
#include <utility>

#include "hsrq_host.h"

namespace hsrq {

hipError_t& last_hip_error() {
  static thread_local hipError_t e = hipSuccess;
  return e;
}

int validate_problem(const hsrq_problem& q) {
  if (q.n <= 0 || q.m < 0 || q.ldH < q.n || q.ldvr < q.n || !q.H) return HSRQ_EINVAL;
  if (q.m > 0 && (!q.wr || !q.vr)) return HSRQ_EINVAL;
  return HSRQ_OK;
}

int choose_tile(const hsrq_context* ctx, int tlsz, int n) {
  int t = tlsz > 0 ? tlsz : (ctx ? ctx->tile : 0);
  if (t <= 0) t = n >= 4096 ? 128 : 64;
  t = ((t + kRowBlock - 1) / kRowBlock) * kRowBlock;
  return std::min(256, std::max(64, t));
}

int choose_pack(int requested, int b, int total_shifts) {
  // the packed sweep needs 16 lanes x 2^k rows per lane
  if (b % 64 != 0 || ((b / 16) & (b / 16 - 1)) != 0) return 1;
  if (requested == 1 || requested == 4) return requested;
  return total_shifts >= kPackedMinShifts ? 4 : 1;
}

// Per SIMD 512 VGPRs, per CU 160 KiB LDS.  The looping panel pass takes 128
// VGPRs per wave and 40 KiB of LDS per block; the 128-row k_diag 128 VGPRs
// and 33 KiB, so three panel-pass blocks fit beside one sweep block (n =
// 8000, m = 800: 6.56 ms with two blocks per CU, 6.31 ms with three; with
// four the sweeps wait for room: 7.33 ms).  With more
// than one sweep block per CU (above 4 shifts x 256 CUs) no panel-pass block
// fits beside the sweeps: the lookahead then only delays the passes (n =
// 20000, m = 2000: 94.0 ms with 2 blocks per CU, 64.3 ms without).  The
// packed sweeps run where many shifts fill the GPU by themselves: there a
// capped panel pass only slows the plan (256-problem batch in one lane: 31.4
// ms without, 41.7 ms with one block per CU).  Otherwise no lookahead.
int lookahead_blocks_per_cu(int pack, int b, int shifts, int cus) {
  (void)b;
  return pack == 1 && shifts <= kShiftsPerK1 * std::max(cus, 1) ? 3 : 0;
}

namespace {
struct Lay {
  size_t rt0, rt1, c, s, bm0, bm1, aux0, aux1, ynb0, ynb1, rowinf, diagn, part, scal, ep0, ep1,
      ef, st, ccol, rrow, be0, be1, h32, b320, b321;
  size_t total;
};

Lay layout(int n, int m, int b, bool eq) {
  const int T = (n + b - 1) / b;
  const int b0 = n - (T - 1) * b;
  const int nrb = (b0 + kRowBlock - 1) / kRowBlock + (T - 1) * (b / kRowBlock);
  const size_t nm = (size_t)n * m * sizeof(double);
  const size_t Tm = (size_t)T * m;
  Lay L;
  size_t x = 0;
  auto take = [&](size_t bytes) {
    const size_t o = x;
    x += align_up(bytes);
    return o;
  };
  L.rt0 = take(nm);
  L.rt1 = take(nm);
  L.c = take(nm);
  L.s = take(nm);
  L.bm0 = take((size_t)b * 2 * m * sizeof(double));
  L.bm1 = take((size_t)b * 2 * m * sizeof(double));
  L.aux0 = take((size_t)4 * m * sizeof(double));
  L.aux1 = take((size_t)4 * m * sizeof(double));
  L.ynb0 = take((size_t)nrb * m * sizeof(double));
  L.ynb1 = take((size_t)nrb * m * sizeof(double));
  L.rowinf = take((size_t)T * sizeof(double));
  L.diagn = take((size_t)T * sizeof(double));
  L.part = take((size_t)nrb * n * sizeof(double));
  L.scal = take(4 * sizeof(double));
  L.ep0 = take(Tm * sizeof(int));
  L.ep1 = take(Tm * sizeof(int));
  L.ef = take(Tm * sizeof(int));
  L.st = take((size_t)std::max(m, 1) * sizeof(int));
  L.ccol = take((size_t)n * sizeof(int));
  L.rrow = take((size_t)n * sizeof(int));
  L.be0 = take((size_t)std::max(2 * m, 1) * sizeof(int));
  L.be1 = take((size_t)std::max(2 * m, 1) * sizeof(int));
  L.h32 = L.b320 = L.b321 = 0;
  if (eq) {   // FP32 operands of the reduced-precision panel pass
    L.h32 = take((size_t)n * n * sizeof(float));
    L.b320 = take((size_t)b * 2 * m * sizeof(float));
    L.b321 = take((size_t)b * 2 * m * sizeof(float));
  }
  L.total = x;
  return L;
}
}  // namespace

size_t prob_workspace_bytes(int n, int m, int b, bool eq) { return layout(n, m, b, eq).total; }

void prob_describe(ProbDev& P, const hsrq_problem& q, int b, int its, int exact, char* ws,
                   const int* its_k, bool eq) {
  std::memset(&P, 0, sizeof(P));
  P.H = q.H;
  P.ldH = q.ldH;
  P.n = q.n;
  P.m = q.m;
  P.b = b;
  P.T = (q.n + b - 1) / b;
  P.b0 = q.n - (P.T - 1) * b;
  P.nb0blk = (P.b0 + kRowBlock - 1) / kRowBlock;
  P.nrb = P.nb0blk + (P.T - 1) * (b / kRowBlock);
  P.wr = q.wr;
  P.Y = q.vr;
  P.ldY = q.ldvr;
  P.its = its;
  P.its_k = its_k;
  P.force_exact = exact;
  const Lay L = layout(q.n, q.m, b, eq);
  P.Rt[0] = (double*)(ws + L.rt0);
  P.Rt[1] = (double*)(ws + L.rt1);
  P.C = (double*)(ws + L.c);
  P.S = (double*)(ws + L.s);
  P.Bm[0] = (double*)(ws + L.bm0);
  P.Bm[1] = (double*)(ws + L.bm1);
  P.aux[0] = (double*)(ws + L.aux0);
  P.aux[1] = (double*)(ws + L.aux1);
  P.Ynb[0] = (double*)(ws + L.ynb0);
  P.Ynb[1] = (double*)(ws + L.ynb1);
  P.rowinf = (double*)(ws + L.rowinf);
  P.diagn = (double*)(ws + L.diagn);
  P.part = (double*)(ws + L.part);
  P.scal = (double*)(ws + L.scal);
  P.Ep[0] = (int*)(ws + L.ep0);
  P.Ep[1] = (int*)(ws + L.ep1);
  P.Ef = (int*)(ws + L.ef);
  P.status = (int*)(ws + L.st);
  if (eq) {
    P.ccol = (int*)(ws + L.ccol);
    P.rrow = (int*)(ws + L.rrow);
    P.Be[0] = (int*)(ws + L.be0);
    P.Be[1] = (int*)(ws + L.be1);
    P.H32 = (float*)(ws + L.h32);
    P.B32[0] = (float*)(ws + L.b320);
    P.B32[1] = (float*)(ws + L.b321);
  }
}

// Tile of the automatic policy.  A single problem below n = 4096 takes
// 64-row tiles (half the step latency of 128); in a batch 128-row tiles win:
// 30.3 ms vs 33.2-34.4 ms for 64-row tiles on the small problems and 44-57
// ms with 256-row tiles (MI355X, 256 problems of n in [200, 4000]).  A single
// problem with more shifts than one-shift sweep workgroups fit on the CUs
// (no lookahead with 128-row tiles: the panel passes set the pace) and
// fewer than the packed sweeps take takes 256-row tiles: half the steps, so
// half the panel passes' R~/y traffic and prologues (MI355X, FP64, 128 /
// 256: n = 20000, m = 2000: 64.5 / 56.4 ms; n = 16000, m = 1600: 36.9 /
// 33.9; n = 12000, m = 2000: 26.7 / 24.4; n = 8000, m = 2000: 13.6 / 13.1;
// with the lookahead, m = 800: n = 8000 6.1 / 8.3, n = 20000 27.2 / 31.5).
int auto_tile(int n, long long shifts, int count, int cus) {
  if (count == 1 && n < 4096) return 64;
  if (count == 1 && shifts > (long long)kShiftsPerK1 * std::max(cus, 1) &&
      shifts < kPackedMinShifts)
    return 256;
  return 128;
}


// Lane assignment.  Problems are grouped by tile.  Few problems (< 8): each
// problem's shifts are cut into `nl` pieces, one per lane, so the diagonal
// sweeps of one lane overlap the panel passes of another.  Many problems: a
// group with many shifts is split into up to 3 lanes of whole problems,
// balanced by panel-pass work m n^2, largest first (256-problem batch: 1
// lane 34.7 ms, 2 lanes 30.8, 3 lanes 30.3, 4 lanes 34.3).
void assign_lanes(const hsrq_context* ctx, int count, const hsrq_problem* in, int tlsz,
                  std::vector<std::vector<Piece>>& groups, std::vector<int>& tiles) {
  const bool fixed = tlsz > 0 || ctx->tile > 0;
  std::vector<int> tile_of(count);
  int nmax = 0;
  for (int i = 0; i < count; ++i) nmax = std::max(nmax, in[i].n);
  for (int i = 0; i < count; ++i)
    tile_of[i] = fixed ? choose_tile(ctx, tlsz, nmax)
                       : choose_tile(ctx, auto_tile(in[i].n, in[i].m, count, ctx->cus), 0);
  std::vector<int> distinct(tile_of);
  std::sort(distinct.begin(), distinct.end());
  distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
  for (int b : distinct) {
    std::vector<int> g;
    long long shifts = 0;
    for (int i = 0; i < count; ++i)
      if (tile_of[i] == b) {
        g.push_back(i);
        shifts += in[i].m;
      }
    if (count < 8) {
      int nl = ctx->lanes > 0 ? ctx->lanes : (shifts >= kSplitMinShifts ? 2 : 1);
      std::vector<std::vector<Piece>> h(nl);
      for (int i : g) {
        // cut in multiples of 16 shifts (whole packed workgroups)
        const int m = in[i].m, per = ((m + nl - 1) / nl + 15) / 16 * 16;
        for (int l = 0; l < nl; ++l) {
          const int k0 = std::min(m, l * per), k1 = std::min(m, k0 + per);
          if (k1 > k0 || (l == 0 && m == 0)) h[l].push_back({i, k0, k1 - k0});
        }
      }
      for (auto& x : h)
        if (!x.empty()) {
          groups.push_back(x);
          tiles.push_back(b);
        }
      continue;
    }
    const int nl = ctx->lanes > 0
                       ? ctx->lanes
                       : (int)std::min<long long>(3, std::max<long long>(1, shifts / (2 * kPackedMinShifts)));
    std::vector<int> ord(g);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) {
      return (double)in[a].m * in[a].n * in[a].n > (double)in[c].m * in[c].n * in[c].n;
    });
    std::vector<std::vector<Piece>> h(nl);
    std::vector<double> load(nl, 0.0);
    for (int i : ord) {
      const int t = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      h[t].push_back({i, 0, in[i].m});
      load[t] += (double)in[i].m * in[i].n * in[i].n;
    }
    for (auto& x : h) {
      std::sort(x.begin(), x.end(), [](const Piece& u, const Piece& v) { return u.prob < v.prob; });
      if (!x.empty()) {
        groups.push_back(x);
        tiles.push_back(b);
      }
    }
  }
}

// Start steps of a lane's problems: largest T m first, each where it raises
// the per-step maximum of the shifts in flight least (then their sum of
// squares).  Weighting the steps by their panel-pass work as well (rows above
// the tile) measured no better on the 256-problem batch.
std::vector<int> step_offsets(const std::vector<ProbDev>& probs, int steps) {
  const int cnt = (int)probs.size();
  std::vector<int> off(cnt, 0);
  if (cnt < 2) return off;
  std::vector<int> order;
  for (int i = 0; i < cnt; ++i)
    if (probs[i].m > 0) order.push_back(i);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (int64_t)probs[a].T * probs[a].m > (int64_t)probs[b].T * probs[b].m;
  });
  std::vector<double> load(steps, 0.0);
  for (int i : order) {
    const int T = probs[i].T;
    const double w = probs[i].m;
    int best = 0;
    double bmax = 0.0, bsq = 0.0;
    for (int o = 0; o + T <= steps; ++o) {
      double mx = 0.0, sq = 0.0;
      for (int s = o; s < o + T; ++s) {
        const double x = load[s] + w;
        mx = std::max(mx, x);
        sq += x * x;
      }
      if (o == 0 || mx < bmax || (mx == bmax && sq < bsq)) {
        best = o;
        bmax = mx;
        bsq = sq;
      }
    }
    off[i] = best;
    for (int s = best; s < best + T; ++s) load[s] += w;
  }
  return off;
}

}  // namespace hsrq
