// Newton (SNESNEWTONLS + backtracking line search + SNESConvergedDefault) and
// GMRES(m) with block-Jacobi ILU(0) driven from the host on device vectors,
// plus the time loops: TS backward Euler (TSBEULER), fixed-dt SNES stepping
// (demo/transient/transient_snes_mpfaof90.F90:406-465) and
// TDyTimeIntegratorRunToTime (src/tdyti.c:71-196).
#include <algorithm>
#include <cmath>
#include <deque>

#include <unordered_map>

#include "krylov.hip"
#include "coarse.hip"
#include "linesearch.h"
#include "tdyhip_internal.h"

namespace tdyhip {

struct HostBJ {
  int nblocks = 0, bs = 0;
  std::vector<int32_t> row, jpos;
  std::vector<uint8_t> diag, nslot;
  std::vector<uint16_t> lcol, levL, levU, cL, cU;
  std::vector<uint32_t> pair;
  std::vector<uint8_t> npair;
  std::vector<int32_t> nlev;
  std::vector<int32_t> nlu;   // nblocks: most strictly lower (or upper) in-block entries of a row
  std::vector<int32_t> perm;  // nc: block-layout position of each cell
  std::vector<int32_t> bw;    // nblocks: longest row of each block (ELL slots in use)
  std::vector<int32_t> ecol;  // W*nblocks*bs: global column of slot k at block position p (k*NB + p)
  std::vector<int16_t> ecol16;  // the same as offsets from the row, when every offset fits in int16
  // Deduplicated index patterns (int16 offsets only): blocks whose L/U local
  // columns and column offsets are identical share one copy.  When set, cL/cU
  // hold npat patterns and ecol16 is pattern-major: slot k of thread t of
  // pattern q at (q*W + k)*bs + t.
  std::vector<int32_t> pat;     // nblocks: pattern of each block (empty: one copy per block)
  int npat = 0;
};

// Breadth-first partition of the cells into blocks of at most `bs` cells over
// the Jacobian's vertex-adjacency graph, then the in-block ILU(0) slots and
// level schedules.
static std::string build_bj(const std::vector<int32_t> &jrp, const std::vector<int32_t> &jcols, int nc, int W, int bs,
                            const std::vector<int32_t> &part, HostBJ &out) {
  if (bs < 1 || bs > kBlkMax) return "block size must be in [1,512]";
  const int S = bs <= 64 ? 64 : bs <= 256 ? 256 : 512;  // per-block stride = launch width
  out.bs = S;
  std::vector<int32_t> blk(nc, -1), local(nc, -1);
  std::vector<std::vector<int32_t>> members;
  std::vector<int32_t> queue;
  queue.reserve(4096);
  int seed = 0;
  if (!part.empty()) {
    int nbk = 0;
    for (int c = 0; c < nc; c++) {
      if (part[c] < 0) return "negative block id in the preconditioner partition";
      nbk = std::max(nbk, part[c] + 1);
    }
    members.assign(nbk, {});
    for (int c = 0; c < nc; c++) { blk[c] = part[c]; members[part[c]].push_back(c); }
    for (auto &mem : members)
      if ((int)mem.size() > bs) return "preconditioner block larger than pc_block";
    members.erase(std::remove_if(members.begin(), members.end(), [](const std::vector<int32_t> &m) { return m.empty(); }),
                  members.end());
    for (size_t b = 0; b < members.size(); b++)
      for (int c : members[b]) blk[c] = (int32_t)b;
    seed = nc;
  }
  while (true) {
    while (seed < nc && blk[seed] >= 0) seed++;
    if (seed >= nc) break;
    const int b = (int)members.size();
    members.emplace_back();
    auto &mem = members.back();
    queue.clear();
    queue.push_back(seed);
    blk[seed] = b;
    size_t head = 0;
    while (head < queue.size() && (int)mem.size() < bs) {
      const int c = queue[head++];
      mem.push_back(c);
      for (int k = jrp[c]; k < jrp[c + 1]; k++) {
        const int nb = jcols[k];
        if (blk[nb] < 0 && (int)queue.size() < bs) {
          blk[nb] = b;
          queue.push_back(nb);
        }
      }
    }
    // cells queued but not taken (cannot happen: queue is capped at bs) -> release
    for (size_t q = head; q < queue.size(); q++)
      if ((int)mem.size() < bs) mem.push_back(queue[q]);
      else blk[queue[q]] = -1;
    std::sort(mem.begin(), mem.end());
  }
  const int nb = (int)members.size();
  out.nblocks = nb;
  const int64_t NB = (int64_t)nb * S;  // rows of the block layout
  if (NB * W > INT32_MAX) return "mesh too large for the block-Jacobi layout";
  out.row.assign((size_t)nb * S, -1);
  out.jpos.assign((size_t)nb * kW * S, -1);
  out.lcol.assign((size_t)nb * kW * S, 0xffff);
  out.diag.assign((size_t)nb * S, 0);
  out.nslot.assign((size_t)nb * S, 0);
  out.levL.assign((size_t)nb * S, 0);
  out.levU.assign((size_t)nb * S, 0);
  out.cL.assign((size_t)nb * kLU * S, (uint16_t)S);
  out.cU.assign((size_t)nb * kLU * S, (uint16_t)S);
  out.npair.assign((size_t)nb * kLU * S, 0);
  out.pair.assign((size_t)nb * kPair * S, 0);
  out.nlev.assign((size_t)nb * 2, 0);
  out.nlu.assign((size_t)nb, 0);
  for (int b = 0; b < nb; b++) {
    const auto &mem = members[b];
    const int m = (int)mem.size();
    for (int i = 0; i < m; i++) local[mem[i]] = i;
    int maxL = 0, maxU = 0;
    for (int i = 0; i < m; i++) {
      const int r = mem[i];
      out.row[(size_t)b * S + i] = r;
      int ns = 0;
      for (int k = jrp[r]; k < jrp[r + 1]; k++) {
        const int c = jcols[k];
        if (blk[c] != b) continue;
        if (ns >= kW) return "more than 27 in-block entries in a row";
        // position of (r, slot k-jrp[r]) in the block-contiguous Jacobian, from
        // the block's first entry (block b's W slots start at b*W*S)
        out.jpos[(size_t)b * kW * S + (size_t)ns * S + i] = (int32_t)((k - jrp[r]) * S + i);
        out.lcol[(size_t)b * kW * S + (size_t)ns * S + i] = (uint16_t)local[c];
        if (c == r) out.diag[(size_t)b * S + i] = (uint8_t)ns;
        ns++;
      }
      out.nslot[(size_t)b * S + i] = (uint8_t)ns;
    }
    // levels: L forward, U backward
    for (int i = 0; i < m; i++) {
      int lv = 0;
      const int di = out.diag[(size_t)b * S + i];
      for (int p = 0; p < di; p++) {
        const int k = out.lcol[(size_t)b * kW * S + (size_t)p * S + i];
        lv = std::max(lv, (int)out.levL[(size_t)b * S + k] + 1);
      }
      out.levL[(size_t)b * S + i] = (uint16_t)lv;
      maxL = std::max(maxL, lv + 1);
    }
    for (int i = m - 1; i >= 0; i--) {
      int lv = 0;
      const int di = out.diag[(size_t)b * S + i], ns = out.nslot[(size_t)b * S + i];
      for (int p = di + 1; p < ns; p++) {
        const int k = out.lcol[(size_t)b * kW * S + (size_t)p * S + i];
        lv = std::max(lv, (int)out.levU[(size_t)b * S + k] + 1);
      }
      out.levU[(size_t)b * S + i] = (uint16_t)lv;
      maxU = std::max(maxU, lv + 1);
    }
    out.nlev[2 * b] = maxL;
    out.nlev[2 * b + 1] = maxU;
    // split L / U column tables (padding -> the zero entry S)
    for (int i = 0; i < m; i++) {
      const int di = out.diag[(size_t)b * S + i], ns = out.nslot[(size_t)b * S + i];
      if (di > kLU || ns - di - 1 > kLU) return "more than 13 lower or upper in-block entries in a row";
      out.nlu[b] = std::max(out.nlu[b], std::max(di, ns - di - 1));
      // ILU(0) update pairs: row i slot d -= l_ik * row k slot s, for the entries of row k
      // after its diagonal whose column is in row i's pattern
      for (int p = 0; p < di; p++) {
        const int k = out.lcol[(size_t)b * kW * S + (size_t)p * S + i];
        const int dk = out.diag[(size_t)b * S + k], nk = out.nslot[(size_t)b * S + k];
        int np = 0, pi = p + 1;
        for (int q = dk + 1; q < nk; q++) {
          const int cq = out.lcol[(size_t)b * kW * S + (size_t)q * S + k];
          while (pi < ns && out.lcol[(size_t)b * kW * S + (size_t)pi * S + i] < cq) pi++;
          if (pi < ns && out.lcol[(size_t)b * kW * S + (size_t)pi * S + i] == cq) {
            if (np >= kLU) return "too many ILU(0) update pairs in a row";
            out.pair[(size_t)b * kPair * S + (size_t)(p * kPairW + np / 2) * S + i] |=
                (uint32_t)(pi | (q << 8)) << (16 * (np & 1));
            np++;
          }
        }
        out.npair[(size_t)b * kLU * S + (size_t)p * S + i] = (uint8_t)np;
      }
      for (int p = 0; p < kLU; p++) {
        out.cL[(size_t)b * kLU * S + (size_t)p * S + i] =
            p < di ? out.lcol[(size_t)b * kW * S + (size_t)p * S + i] : (uint16_t)S;
        const int q = di + 1 + p;
        out.cU[(size_t)b * kLU * S + (size_t)p * S + i] =
            q < ns ? out.lcol[(size_t)b * kW * S + (size_t)q * S + i] : (uint16_t)S;
      }
    }
    for (int i = 0; i < m; i++) local[mem[i]] = -1;
  }
  // block-ordered ELL columns (padding slots point at the row itself with a 0 value)
  out.perm.assign(nc, -1);
  out.ecol.assign((size_t)W * NB, 0);
  for (int64_t p = 0; p < NB; p++) {
    const int r = out.row[p];
    if (r < 0) continue;
    out.perm[r] = (int32_t)p;
    const int len = jrp[r + 1] - jrp[r];
    for (int k = 0; k < W; k++) out.ecol[(size_t)k * NB + p] = k < len ? jcols[jrp[r] + k] : r;
  }
  for (int c = 0; c < nc; c++)
    if (out.perm[c] < 0) return "cell missing from the block-Jacobi partition";
  out.bw.assign(out.nblocks, 0);
  for (int64_t p = 0; p < NB; p++) {
    const int r = out.row[p];
    if (r >= 0) out.bw[p / S] = std::max(out.bw[p / S], jrp[r + 1] - jrp[r]);
  }
  bool fits = true;
  for (int64_t p = 0; p < NB && fits; p++) {
    const int r = out.row[p];
    for (int k = 0; k < W && fits; k++) {
      const int64_t off = r < 0 ? 0 : (int64_t)out.ecol[(size_t)k * NB + p] - r;
      fits = off >= INT16_MIN && off <= INT16_MAX;
    }
  }
  if (fits) {
    out.ecol16.resize(out.ecol.size());
    for (int64_t p = 0; p < NB; p++) {
      const int r = out.row[p];
      for (int k = 0; k < W; k++)
        out.ecol16[(size_t)k * NB + p] = r < 0 ? 0 : (int16_t)(out.ecol[(size_t)k * NB + p] - r);
    }
    out.ecol.clear();
    // Blocks of a structured mesh repeat a few index patterns (an interior
    // brick's in-block columns and column offsets are those of every other
    // interior brick): keep one copy per pattern when that at least halves
    // them, so the matvec streams values, not indices.
    const size_t lu = (size_t)kLU * S, sw = (size_t)kW * S, pw = (size_t)kPair * S;
    std::unordered_map<std::string, int> seen;
    std::vector<int> rep;
    std::vector<int32_t> pat(nb);
    std::string key;
    auto add = [&](const auto &v, size_t off, size_t n) {
      key.append((const char *)(v.data() + off), n * sizeof(v[0]));
    };
    for (int b = 0; b < nb; b++) {
      key.clear();
      add(out.cL, b * lu, lu); add(out.cU, b * lu, lu);
      add(out.jpos, b * sw, sw); add(out.lcol, b * sw, sw);
      add(out.npair, b * lu, lu); add(out.pair, b * pw, pw);
      add(out.diag, (size_t)b * S, S); add(out.nslot, (size_t)b * S, S);
      add(out.levL, (size_t)b * S, S); add(out.levU, (size_t)b * S, S);
      add(out.nlev, (size_t)2 * b, 2);
      add(out.nlu, (size_t)b, 1);
      for (int k = 0; k < W; k++) key.append((const char *)(out.ecol16.data() + (size_t)k * NB + (size_t)b * S), (size_t)S * 2);
      auto it = seen.find(key);
      if (it == seen.end()) {
        it = seen.emplace(key, (int)rep.size()).first;
        rep.push_back(b);
      }
      pat[b] = it->second;
    }
    const int npat = (int)rep.size();
    if (2 * npat <= nb) {
      // keep the representative block of every pattern, in pattern order
      auto shrink = [&](auto &v, size_t per) {
        std::remove_reference_t<decltype(v)> w(npat * per);
        for (int q = 0; q < npat; q++)
          std::copy(v.begin() + (size_t)rep[q] * per, v.begin() + (size_t)(rep[q] + 1) * per, w.begin() + (size_t)q * per);
        v.swap(w);
      };
      shrink(out.cL, lu); shrink(out.cU, lu); shrink(out.jpos, sw); shrink(out.lcol, sw);
      shrink(out.npair, lu); shrink(out.pair, pw); shrink(out.diag, S); shrink(out.nslot, S);
      shrink(out.levL, S); shrink(out.levU, S); shrink(out.nlev, 2); shrink(out.nlu, 1);
      std::vector<int16_t> e16((size_t)npat * W * S);
      for (int q = 0; q < npat; q++)
        for (int k = 0; k < W; k++)
          std::copy(out.ecol16.begin() + (size_t)k * NB + (size_t)rep[q] * S,
                    out.ecol16.begin() + (size_t)k * NB + (size_t)(rep[q] + 1) * S, e16.begin() + ((size_t)q * W + k) * S);
      out.ecol16.swap(e16);
      out.pat.swap(pat);
      out.npat = npat;
    }
  }
  return std::string();
}

struct HostCoarse {
  int nb = 0, ng = 0, gsz = 0;
  std::vector<int32_t> nbr, grp, gmem;
  std::vector<uint8_t> cq;
};

// Level-1 / level-2 structure of the coarse correction: neighbour blocks of
// each block, the slot -> neighbour map, and groups of blocks built by
// breadth-first aggregation over block adjacency (group size grown until at
// most kNG groups remain).  Returns an empty string when the mesh is outside
// the correction's limits (then it is not used).
static std::string build_coarse(const HostBJ &hb, const std::vector<int32_t> &jrp, const std::vector<int32_t> &jcols,
                                int nc, int W, HostCoarse &out) {
  const int S = hb.bs, nb = hb.nblocks;
  const int64_t NB = (int64_t)nb * S;
  std::vector<int32_t> blk(nc);
  for (int c = 0; c < nc; c++) blk[c] = hb.perm[c] / S;
  out.nb = nb;
  out.nbr.assign((size_t)nb * kCQ, -1);
  std::vector<int32_t> tmp;
  for (int b = 0; b < nb; b++) {
    tmp.clear();
    for (int i = 0; i < S; i++) {
      const int r = hb.row[(size_t)b * S + i];
      if (r < 0) continue;
      for (int k = jrp[r]; k < jrp[r + 1]; k++) tmp.push_back(blk[jcols[k]]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    if ((int)tmp.size() > kCQ) return "a block has more than 32 neighbour blocks";
    for (size_t q = 0; q < tmp.size(); q++) out.nbr[(size_t)b * kCQ + q] = tmp[q];
  }
  out.cq.assign((size_t)W * NB, 255);
  for (int64_t p = 0; p < NB; p++) {
    const int r = hb.row[p];
    if (r < 0) continue;
    const int b = (int)(p / S);
    const int32_t *nl = &out.nbr[(size_t)b * kCQ];
    for (int k = 0; k < jrp[r + 1] - jrp[r]; k++) {
      const int cb = blk[jcols[jrp[r] + k]];
      for (int q = 0; q < kCQ && nl[q] >= 0; q++)
        if (nl[q] == cb) { out.cq[(size_t)k * NB + p] = (uint8_t)q; break; }
    }
  }
  // groups
  for (int G = std::max(1, (nb + kNG - 1) / kNG); G <= kGMax; G++) {
    std::vector<int32_t> grp(nb, -1), mem;
    std::vector<std::vector<int32_t>> groups;
    std::vector<int32_t> queue;
    for (int seed = 0; seed < nb; seed++) {
      if (grp[seed] >= 0) continue;
      const int g = (int)groups.size();
      groups.emplace_back();
      queue.assign(1, seed);
      grp[seed] = g;
      size_t head = 0;
      while (head < queue.size() && (int)groups[g].size() < G) {
        const int b = queue[head++];
        groups[g].push_back(b);
        for (int q = 0; q < kCQ; q++) {
          const int nbk = out.nbr[(size_t)b * kCQ + q];
          if (nbk < 0) break;
          if (grp[nbk] < 0 && (int)queue.size() < G) { grp[nbk] = g; queue.push_back(nbk); }
        }
      }
      for (size_t q = head; q < queue.size(); q++) {  // queued beyond the group size: release
        if ((int)groups[g].size() < G) groups[g].push_back(queue[q]);
        else grp[queue[q]] = -1;
      }
    }
    if ((int)groups.size() > kNG) continue;
    out.ng = (int)groups.size();
    out.gsz = G;
    out.grp = grp;
    out.gmem.assign((size_t)out.ng * G, -1);
    for (int g = 0; g < out.ng; g++)
      for (size_t i = 0; i < groups[g].size(); i++) out.gmem[(size_t)g * G + i] = groups[g][i];
    return std::string();
  }
  return "too many blocks for the coarse correction";
}

struct SolverWS {
  int n = 0, restart = 0, fp32 = -1, bs = 0, part_version = -1;
  int64_t nnz = 0;
  int64_t bytes = 0;  // bytes held in `allocs`
  BJLayout bj;
  std::vector<void *> allocs;
  // Newton vectors (FP64)
  double *F = nullptr, *Y = nullptr, *W = nullptr, *G = nullptr, *JY = nullptr, *Udot = nullptr, *J = nullptr;
  double *Uprev_ts = nullptr, *Usave = nullptr;
  // Krylov (TV)
  void *V = nullptr, *w = nullptr, *z = nullptr, *b = nullptr, *x = nullptr, *Jk = nullptr, *fac = nullptr,
       *partial = nullptr;
  double *dpartial = nullptr, *h = nullptr;
  double *RV = nullptr;  // (restart + 1) * nblocks: block sums of the basis vectors (coarse correction on)
  double *hs = nullptr;  // Hessenberg, Givens cs/sn, g, y of the device GMRES cycle
  GmresDev *gdev = nullptr;
  int *status = nullptr;
  int nchunk = 0;
  int64_t NB = 0;                // rows of the block layout (nblocks * bs)
  int32_t *d_perm = nullptr;     // nc: block-layout position of each cell
  int32_t *d_ecol = nullptr;     // W*NB block-ordered ELL columns (int32), or
  int16_t *d_ecol16 = nullptr;   // the same as int16 offsets from the row (one of the two is set)
  // coarse correction (when requested and the mesh is within its limits)
  bool coarse = false;
  int coarse_req = -1;
  double lu_mean = 0.0;          // mean per-block L/U slot count the sweeps read
  double w_mean = 0.0;           // mean per-block ELL slot count the SpMV reads
  CoarseLayout cl;
  // the coarse setup's own stream and its fork/join events: it runs beside
  // the ILU(0) factor, as both only read the Jacobian
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // one GMRES restart cycle captured as a graph, per Krylov precision
  // (index 0 FP64, 1 FP32); its kernels only touch workspace arrays
  hipGraphExec_t cycle_exec[2] = {nullptr, nullptr};
  // Newton iterations (and tails) captured as graphs
  struct NewtonGraph {
    const double *X;
    int type, fp32, cycles;
    hipGraphExec_t exec;
  };
  std::vector<NewtonGraph> newton_graphs;
  int cycles_hint = 1;            // restart cycles the last Newton iteration's solve needed
  struct NewtonDev *nd = nullptr; // the Newton iteration's outcome on the device
  double *nh = nullptr;           // 4 norms / dots of the Newton tail
  double *dyn = nullptr;          // dtime, shift (1 / dt of the TS form) read by the solver's launches
};

constexpr int kMaxGraphCycles = 8;    // restart cycles inside one Newton-iteration graph
constexpr size_t kMaxNewtonGraphs = 16;

static void ws_free(SolverWS &ws) {
  for (hipGraphExec_t &e : ws.cycle_exec)
    if (e) (void)hipGraphExecDestroy(e);
  for (auto &e : ws.newton_graphs) (void)hipGraphExecDestroy(e.exec);
  if (ws.side) (void)hipStreamSynchronize(ws.side);
  if (ws.fork_ev) (void)hipEventDestroy(ws.fork_ev);
  if (ws.join_ev) (void)hipEventDestroy(ws.join_ev);
  if (ws.side) (void)hipStreamDestroy(ws.side);
  for (void *p : ws.allocs) (void)hipFree(p);
  ws.allocs.clear();
  ws = SolverWS();
}

template <class T>
static int ws_alloc(SolverWS &ws, T **p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc (solver workspace): ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? TDYHIP_ERR_MEM : TDYHIP_ERR_HIP;
  }
  ws.allocs.push_back(*p);
  ws.bytes += (int64_t)(n * sizeof(T));
  return 0;
}

// an ELL table built slot-major over the block layout (slot k of position p
// at k*NB + p) in the device's block-contiguous order ((b*W + k)*bs + t)
template <class T>
static std::vector<T> block_contiguous(const std::vector<T> &in, int W, int bs, int nblocks) {
  const int64_t NB = (int64_t)nblocks * bs;
  std::vector<T> out(in.size());
  for (int64_t b = 0; b < nblocks; b++)
    for (int k = 0; k < W; k++)
      std::copy(in.begin() + (size_t)(k * NB + b * bs), in.begin() + (size_t)(k * NB + (b + 1) * bs),
                out.begin() + (size_t)((b * W + k) * bs));
  return out;
}

static int ensure_ws(tdyhip_ctx *ctx, const tdyhip_solver_opts &o) {
  if (!ctx->ws) ctx->ws = new SolverWS();
  SolverWS &ws = *ctx->ws;
  const int fp32 = o.krylov_fp32 ? 1 : 0;
  // one workspace serves both Krylov precisions (buffers sized for FP64 plus
  // the FP32 Jacobian copy), so switching precision costs nothing
  (void)fp32;
  if (ws.n == ctx->nc && ws.restart == o.ksp_restart && ws.bs == o.pc_block &&
      ws.part_version == ctx->pc_part_version && ws.coarse_req == (o.pc_coarse ? 1 : 0))
    return 0;
  ws_free(ws);
  HostBJ hb;
  std::string msg = build_bj(ctx->h_jrp, ctx->h_jcols, ctx->nc, ctx->jell_w, o.pc_block, ctx->pc_part, hb);
  if (!msg.empty()) { set_error(msg); return TDYHIP_ERR_ARG; }
  const int n = ctx->nc;
  const int64_t nnz = (int64_t)ctx->jell_w * hb.nblocks * hb.bs;  // block-ordered slot-major ELL Jacobian
  ws.n = n; ws.restart = o.ksp_restart; ws.fp32 = fp32; ws.bs = o.pc_block; ws.nnz = nnz;
  ws.nchunk = (int)((n + kChunk - 1) / kChunk);
  int rc = 0;
  // Krylov buffers are sized for FP64 in both modes: an FP32 solve that fails
  // is repeated in FP64 on the same buffers.
  const size_t tv = sizeof(double);
  auto up = [&](auto **p, const auto &vec) -> int {
    int r = ws_alloc(ws, p, vec.size());
    if (r) return r;
    if (hipMemcpyAsync(*p, vec.data(), vec.size() * sizeof(vec[0]), hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      set_error("hipMemcpy (block-Jacobi layout)");
      return TDYHIP_ERR_HIP;
    }
    return 0;
  };
  ws.bj.nblocks = hb.nblocks;
  ws.bj.bs = hb.bs;
  ws.NB = (int64_t)hb.nblocks * hb.bs;
  if ((rc = up(&ws.d_perm, hb.perm))) return rc;
  ws.bj.d_pat = nullptr;
  ws.bj.npat = hb.npat;
  ws.bj.d_nlu = nullptr;
  if (std::any_of(hb.nlu.begin(), hb.nlu.end(), [](int32_t v) { return v < kLU; }) && (rc = up(&ws.bj.d_nlu, hb.nlu)))
    return rc;
  {
    double sum = 0.0;  // mean stored L/U width over the blocks (byte model)
    for (size_t b = 0; b < (size_t)hb.nblocks; b++) sum += hb.nlu[hb.pat.empty() ? b : (size_t)hb.pat[b]];
    ws.lu_mean = hb.nblocks ? sum / hb.nblocks : 0.0;
  }
  // per-block ELL widths pay off where rows are wide and uneven (prism meshes); hexahedral
  // rows fill all kW slots, and their FP32 SpMV unrolls for exactly kW
  ws.bj.d_bw = nullptr;
  ws.w_mean = ctx->jell_w;
  if (ctx->jell_w > kW && std::any_of(hb.bw.begin(), hb.bw.end(), [&](int32_t v) { return v < ctx->jell_w; })) {
    if ((rc = up(&ws.bj.d_bw, hb.bw))) return rc;
    double sum = 0.0;
    for (int32_t v : hb.bw) sum += v;
    ws.w_mean = hb.nblocks ? sum / hb.nblocks : 0.0;
  }
  if (!hb.pat.empty() && (rc = up(&ws.bj.d_pat, hb.pat))) return rc;
  ws.bj.ell_w = ctx->jell_w;
  // the host builds the ELL tables slot-major over the layout (k*NB + p); the
  // device keeps them block-contiguous ((b*W + k)*bs + t), as the Jacobian
  const int Wd = ctx->jell_w;
  if (!hb.ecol16.empty()) {
    if ((rc = up(&ws.d_ecol16, hb.pat.empty() ? block_contiguous(hb.ecol16, Wd, hb.bs, hb.nblocks) : hb.ecol16)))
      return rc;
  } else if ((rc = up(&ws.d_ecol, block_contiguous(hb.ecol, Wd, hb.bs, hb.nblocks)))) {
    return rc;
  }
  ws.part_version = ctx->pc_part_version;
  if ((rc = up(&ws.bj.d_row, hb.row)) || (rc = up(&ws.bj.d_jpos, hb.jpos)) || (rc = up(&ws.bj.d_lcol, hb.lcol)) ||
      (rc = up(&ws.bj.d_diag, hb.diag)) || (rc = up(&ws.bj.d_nslot, hb.nslot)) || (rc = up(&ws.bj.d_levL, hb.levL)) ||
      (rc = up(&ws.bj.d_levU, hb.levU)) || (rc = up(&ws.bj.d_nlev, hb.nlev)) || (rc = up(&ws.bj.d_cL, hb.cL)) ||
      (rc = up(&ws.bj.d_cU, hb.cU)) || (rc = up(&ws.bj.d_npair, hb.npair)) || (rc = up(&ws.bj.d_pair, hb.pair)))
    return rc;
  for (double **p : {&ws.F, &ws.Y, &ws.W, &ws.G, &ws.JY, &ws.Udot, &ws.Uprev_ts, &ws.Usave})
    if ((rc = ws_alloc(ws, p, n))) return rc;
  if ((rc = ws_alloc(ws, &ws.J, nnz))) return rc;
  char **cp[] = {(char **)&ws.V, (char **)&ws.w, (char **)&ws.z, (char **)&ws.b, (char **)&ws.x, (char **)&ws.fac,
                 (char **)&ws.partial};
  size_t cnt[] = {(size_t)(o.ksp_restart + 1) * n, (size_t)n, (size_t)n, (size_t)n, (size_t)n,
                  (size_t)hb.nblocks * kW * hb.bs, (size_t)(o.ksp_restart + 2) * (ws.nchunk + kNG)};
  for (int i = 0; i < 7; i++)
    if ((rc = ws_alloc(ws, cp[i], cnt[i] * tv))) return rc;
  if ((rc = ws_alloc(ws, (float **)&ws.Jk, nnz))) return rc;  // FP32 copy of the Jacobian
  if ((rc = ws_alloc(ws, &ws.dpartial, ws.nchunk))) return rc;
  if ((rc = ws_alloc(ws, &ws.h, o.ksp_restart + 2))) return rc;
  if ((rc = ws_alloc(ws, &ws.hs, (size_t)(o.ksp_restart + 1) * o.ksp_restart + 4 * (o.ksp_restart + 1)))) return rc;
  if ((rc = ws_alloc(ws, &ws.gdev, 1))) return rc;
  if ((rc = ws_alloc(ws, &ws.status, 1))) return rc;
  if ((rc = ws_alloc(ws, (char **)&ws.nd, 256))) return rc;  // a NewtonDev (defined with the solver below)
  if ((rc = ws_alloc(ws, &ws.nh, 4))) return rc;
  if ((rc = ws_alloc(ws, &ws.dyn, 2))) return rc;
  ws.coarse_req = o.pc_coarse ? 1 : 0;
  if (o.pc_coarse) {
    HostCoarse hc;
    // (the Galerkin row kernel handles up to 32 slots per row)
    if (ctx->jell_w <= 32 && build_coarse(hb, ctx->h_jrp, ctx->h_jcols, ctx->nc, ctx->jell_w, hc).empty()) {
      CoarseLayout &C = ws.cl;
      C.nb = hc.nb; C.ng = hc.ng; C.gsz = hc.gsz;
      if ((rc = up(&C.nbr, hc.nbr)) || (rc = up(&C.cq, block_contiguous(hc.cq, ctx->jell_w, hb.bs, hb.nblocks))) ||
          (rc = up(&C.grp, hc.grp)) ||
          (rc = up(&C.gmem, hc.gmem)))
        return rc;
      if ((rc = ws_alloc(ws, &C.A1, (size_t)hc.nb * kCQ)) ||
          (rc = ws_alloc(ws, &C.G1, (size_t)hc.ng * hc.gsz * hc.gsz)) ||
          (rc = ws_alloc(ws, &C.A2, (size_t)hc.ng * hc.ng)) || (rc = ws_alloc(ws, &C.rc1, (size_t)hc.nb)) ||
          (rc = ws_alloc(ws, &C.z1, (size_t)hc.nb)) ||
          (rc = ws_alloc(ws, &ws.RV, (size_t)(o.ksp_restart + 1) * hc.nb)))
        return rc;
      C.status = ws.status;
      if (hipStreamCreateWithFlags(&ws.side, hipStreamNonBlocking) != hipSuccess) ws.side = nullptr;
      if (!ws.side || hipEventCreateWithFlags(&ws.fork_ev, hipEventDisableTiming) != hipSuccess ||
          hipEventCreateWithFlags(&ws.join_ev, hipEventDisableTiming) != hipSuccess) {
        set_error("hipStreamCreate / hipEventCreate (coarse setup stream) failed");
        return TDYHIP_ERR_HIP;
      }
      ws.coarse = true;
    }
  }
  return 0;
}

#define SCHK(expr)                                                                      \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                    \
      return TDYHIP_ERR_HIP;                                                            \
    }                                                                                   \
  } while (0)

static int dev_dot(tdyhip_ctx *ctx, const double *x, const double *y, double *out, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  KT(ctx, st, "k_dot", (double)ws.n * (x == y ? 8 : 16), k_dot<double, double><<<ws.nchunk, kRedBlock, 0, st>>>(ws.n, x, y, ws.dpartial));
  k_reduce_partials<double><<<1, kRedBlock, 0, st>>>(1, ws.nchunk, ws.dpartial, ws.h);
  SCHK(hipMemcpyAsync(out, ws.h, sizeof(double), hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  return 0;
}

static int dev_norm(tdyhip_ctx *ctx, const double *x, double *out, hipStream_t st) {
  double s;
  int rc = dev_dot(ctx, x, x, &s, st);
  *out = std::sqrt(s);
  return rc;
}

// Index arrays shared by pattern count once per launch (npat patterns of
// W*bs offsets), not once per row.
static double bytes_spmv(const tdyhip_ctx *ctx, const SolverWS &ws, size_t tv, bool idx16) {
  const double idx = ws.bj.npat ? (double)ws.bj.npat * ctx->jell_w * ws.bj.bs * 2 + (double)ws.bj.nblocks * 4
                                : ws.w_mean * ctx->nc * (idx16 ? 2 : 4);
  return ws.w_mean * ctx->nc * tv + idx + 2.0 * ctx->nc * tv;
}

// y = J x with the FP64 Jacobian in the block layout
static void spmv_blk64(tdyhip_ctx *ctx, const SolverWS &ws, const double *x, double *y, hipStream_t st,
                       const int *stop = nullptr) {
  const unsigned g = (unsigned)((ws.NB + 255) / 256);
  const bool i16 = ws.d_ecol16 != nullptr;
  if (i16)
    KT(ctx, st, "k_spmv_blk", bytes_spmv(ctx, ws, sizeof(double), true),
       k_spmv_blk<double, int16_t><<<g, 256, 0, st>>>(ws.NB, ctx->jell_w, ws.bj.d_row, ws.d_ecol16, ws.bj.d_pat, ws.bj.bs, ws.bj.d_bw, ws.J, x, y, stop));
  else
    KT(ctx, st, "k_spmv_blk", bytes_spmv(ctx, ws, sizeof(double), false),
       k_spmv_blk<double, int32_t><<<g, 256, 0, st>>>(ws.NB, ctx->jell_w, ws.bj.d_row, ws.d_ecol, nullptr, ws.bj.bs, ws.bj.d_bw, ws.J, x, y, stop));
}
static double bytes_apply(const SolverWS &ws, size_t tv) {
  const double rows = (double)ws.bj.nblocks * ws.bj.bs;
  const double idx_rows = ws.bj.npat ? (double)ws.bj.npat * ws.bj.bs : rows;
  return rows * ((2 * ws.lu_mean + 1) * tv + 4) + idx_rows * (2 * ws.lu_mean * 2 + 2 + 2) + 2.0 * ws.n * tv;
}
static double bytes_factor(const SolverWS &ws, size_t tv) {
  const double rows = (double)ws.bj.nblocks * ws.bj.bs;
  const double idx_rows = ws.bj.npat ? (double)ws.bj.npat * ws.bj.bs : rows;
  return rows * (kW * 8 + (2 * kLU + 1) * tv) + idx_rows * (kW * (4 + 2) + kLU + kPair * 4 + 1 + 1 + 2);
}

// ---------------------------------------------------------------- GMRES in launch pieces
// GMRES(m), left preconditioned, classical Gram-Schmidt, on ws.J (FP64) or its
// FP32 copy, solving J y = b with b and y FP64.  The solve is launched in
// pieces that the host-driven path and the captured Newton iteration share:
//   gmres_prep   factor the preconditioner, y = 0, r0 = M^-1 b and the first
//                cycle's start (a solve may be converged before any step);
//   gmres_cycle  one restart cycle (m Arnoldi steps that stop early on the
//                device, the Hessenberg solve, the update of y), then the next
//                cycle's start: w = M^-1 (b - J y) with the FP64 Jacobian and
//                its norm.  Once the solve has finished (converged, iteration
//                cap or a non-finite residual) every kernel of a later cycle
//                returns at once, so any number of cycles may run in a row.
template <class TV>
struct KBufs {
  TV *V, *w, *z, *b, *fac, *Jk, *part;
  double *dH, *dcs, *dsn, *dg, *dy, *rc1;
  GmresDev *gd;
  int m, nch;
  int64_t N;
};

template <class TV>
static KBufs<TV> kbufs(tdyhip_ctx *ctx, int m) {
  SolverWS &ws = *ctx->ws;
  KBufs<TV> k;
  k.V = (TV *)ws.V; k.w = (TV *)ws.w; k.z = (TV *)ws.z; k.b = (TV *)ws.b; k.fac = (TV *)ws.fac;
  k.Jk = std::is_same<TV, double>::value ? (TV *)ws.J : (TV *)ws.Jk;
  k.part = (TV *)ws.partial;
  k.dH = ws.hs; k.dcs = k.dH + (size_t)(m + 1) * m; k.dsn = k.dcs + m; k.dg = k.dsn + m; k.dy = k.dg + (m + 1);
  k.rc1 = ws.coarse ? ws.cl.rc1 : nullptr;
  k.gd = ws.gdev;
  k.m = m;
  k.N = ws.n;
  k.nch = (int)((k.N + Red<TV>::chunk - 1) / Red<TV>::chunk);  // chunks of the multi-vector kernels
  return k;
}

// level-1/2 solve of the coarse correction: z1 per block
static CoarseDot coarse_dot(const SolverWS &ws) {
  const CoarseLayout &C = ws.cl;
  CoarseDot cd;
  cd.ng = C.ng; cd.gsz = C.gsz; cd.gmem = C.gmem; cd.G1 = C.G1; cd.A2 = C.A2; cd.rc1 = C.rc1; cd.z1 = C.z1;
  cd.RV = ws.RV; cd.rv_ld = C.nb;
  return cd;
}

static void coarse_solve(tdyhip_ctx *ctx, const int *stp, hipStream_t st) {
  const CoarseLayout &C = ctx->ws->cl;
  KT(ctx, st, "k_coarse_level1", (double)C.ng * C.gsz * C.gsz * 8 + (double)C.nb * 24 + (double)C.ng * C.ng * 8,
     k_coarse_level1<<<C.ng, kL1Threads, 0, st>>>(coarse_dot(*ctx->ws), stp));
}

// out = M^-1 in (block-Jacobi ILU(0), plus the coarse correction when on)
template <class TV>
static void pc_apply(tdyhip_ctx *ctx, const KBufs<TV> &k, const TV *in, TV *out, const int *stp, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  if (ws.bj.bs == 64) KT(ctx, st, "k_bj_apply", bytes_apply(ws, sizeof(TV)), k_bj_apply<TV, 64><<<ws.bj.nblocks, 64, 0, st>>>(ws.bj, k.fac, in, out, stp, k.rc1));
  else if (ws.bj.bs == 256) KT(ctx, st, "k_bj_apply", bytes_apply(ws, sizeof(TV)), k_bj_apply<TV, 256><<<ws.bj.nblocks, 256, 0, st>>>(ws.bj, k.fac, in, out, stp, k.rc1));
  else KT(ctx, st, "k_bj_apply", bytes_apply(ws, sizeof(TV)), k_bj_apply<TV, 512><<<ws.bj.nblocks, 512, 0, st>>>(ws.bj, k.fac, in, out, stp, k.rc1));
  if (!ws.coarse) return;
  const CoarseLayout &C = ws.cl;
  coarse_solve(ctx, stp, st);
  const double bpro = (double)ws.NB * (4 + 2 * sizeof(TV)) + (double)C.nb * 8;
  if (ws.bj.bs == 64) KT(ctx, st, "k_coarse_prolong", bpro, k_coarse_prolong<TV, 64><<<C.nb, 64, 0, st>>>(C, ws.bj.d_row, out, stp));
  else if (ws.bj.bs == 256) KT(ctx, st, "k_coarse_prolong", bpro, k_coarse_prolong<TV, 256><<<C.nb, 256, 0, st>>>(C, ws.bj.d_row, out, stp));
  else KT(ctx, st, "k_coarse_prolong", bpro, k_coarse_prolong<TV, 512><<<C.nb, 512, 0, st>>>(C, ws.bj.d_row, out, stp));
}

// out = M^-1 (b - J y) for a restart: the block-Jacobi part in one fused
// pass (k_bj_resid_apply), then the coarse correction as in pc_apply
template <class TV>
static void resid_pc_apply(tdyhip_ctx *ctx, const KBufs<TV> &k, const double *bvec, const double *y, TV *out,
                           const int *stp, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  const int W = ctx->jell_w;
  const bool i16 = ws.d_ecol16 != nullptr;
  const double by = bytes_spmv(ctx, ws, sizeof(double), i16) + bytes_apply(ws, sizeof(TV)) + (double)ws.n * 8;
  const unsigned nb = ws.bj.nblocks;
#define TDY_RESID(TI, EC)                                                                                           \
  do {                                                                                                             \
    if (ws.bj.bs == 64)                                                                                            \
      KT(ctx, st, "k_bj_resid_apply", by, k_bj_resid_apply<TV, 64, TI><<<nb, 64, 0, st>>>(ws.bj, W, EC, ws.J, y, bvec, k.fac, out, stp, k.rc1)); \
    else if (ws.bj.bs == 256)                                                                                      \
      KT(ctx, st, "k_bj_resid_apply", by, k_bj_resid_apply<TV, 256, TI><<<nb, 256, 0, st>>>(ws.bj, W, EC, ws.J, y, bvec, k.fac, out, stp, k.rc1)); \
    else                                                                                                           \
      KT(ctx, st, "k_bj_resid_apply", by, k_bj_resid_apply<TV, 512, TI><<<nb, 512, 0, st>>>(ws.bj, W, EC, ws.J, y, bvec, k.fac, out, stp, k.rc1)); \
  } while (0)
  if (i16) TDY_RESID(int16_t, ws.d_ecol16);
  else TDY_RESID(int32_t, ws.d_ecol);
#undef TDY_RESID
  if (!ws.coarse) return;
  const CoarseLayout &C = ws.cl;
  coarse_solve(ctx, stp, st);
  const double bpro = (double)ws.NB * (4 + 2 * sizeof(TV)) + (double)C.nb * 8;
  if (ws.bj.bs == 64) KT(ctx, st, "k_coarse_prolong", bpro, k_coarse_prolong<TV, 64><<<C.nb, 64, 0, st>>>(C, ws.bj.d_row, out, stp));
  else if (ws.bj.bs == 256) KT(ctx, st, "k_coarse_prolong", bpro, k_coarse_prolong<TV, 256><<<C.nb, 256, 0, st>>>(C, ws.bj.d_row, out, stp));
  else KT(ctx, st, "k_coarse_prolong", bpro, k_coarse_prolong<TV, 512><<<C.nb, 512, 0, st>>>(C, ws.bj.d_row, out, stp));
}

// V_j = u / ||u|| and out = M^-1 A V_j in one pass (block-ordered SpMV fused
// into the ILU sweeps); with na, ||u|| comes from the previous step's norm
// partials and the launch also finishes that step's Arnoldi column
template <class TV>
static void spmv_apply(tdyhip_ctx *ctx, const KBufs<TV> &k, const TV *u, TV *vj, TV *out, const int *stp,
                       const NormArgs *na, double *rv, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  const int W = ctx->jell_w;
  const double by = bytes_spmv(ctx, ws, sizeof(TV), ws.d_ecol16 != nullptr) + bytes_apply(ws, sizeof(TV));
  const unsigned nb = ws.bj.nblocks;
  const NormArgs nv = na ? *na : NormArgs{};
  GmresDev *gd = k.gd;
  const TV *Jk = k.Jk, *fac = k.fac;
  double *rc1 = k.rc1;
#define TDY_SPMV(TI, EC, NRM)                                                                                      \
  do {                                                                                                             \
    if (ws.bj.bs == 64)                                                                                            \
      KT(ctx, st, "k_bj_spmv_apply", by, k_bj_spmv_apply<TV, 64, TI, NRM><<<nb, 64, 0, st>>>(ws.bj, ws.NB, W, EC, Jk, u, gd, vj, fac, out, stp, rc1, nv, rv)); \
    else if (ws.bj.bs == 256)                                                                                      \
      KT(ctx, st, "k_bj_spmv_apply", by, k_bj_spmv_apply<TV, 256, TI, NRM><<<nb, 256, 0, st>>>(ws.bj, ws.NB, W, EC, Jk, u, gd, vj, fac, out, stp, rc1, nv, rv)); \
    else                                                                                                           \
      KT(ctx, st, "k_bj_spmv_apply", by, k_bj_spmv_apply<TV, 512, TI, NRM><<<nb, 512, 0, st>>>(ws.bj, ws.NB, W, EC, Jk, u, gd, vj, fac, out, stp, rc1, nv, rv)); \
  } while (0)
  if (ws.d_ecol16) {
    if (na) TDY_SPMV(int16_t, ws.d_ecol16, true);
    else TDY_SPMV(int16_t, ws.d_ecol16, false);
  } else {
    if (na) TDY_SPMV(int32_t, ws.d_ecol, true);
    else TDY_SPMV(int32_t, ws.d_ecol, false);
  }
#undef TDY_SPMV
  // the coarse solve runs in the multi-dot launch that follows (its first
  // workgroups), and the update kernel adds the prolongation
}

// ws.h[0] = ||v||^2
template <class TV>
static void norm2_to_h0(tdyhip_ctx *ctx, const TV *v, const int *stp, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  k_dot<TV, TV><<<ws.nchunk, kRedBlock, 0, st>>>(ws.n, v, v, ws.dpartial, stp);
  k_reduce_partials<double><<<1, kRedBlock, 0, st>>>(1, ws.nchunk, ws.dpartial, ws.h, stp);
}

template <class TV>
static void gmres_prep(tdyhip_ctx *ctx, const tdyhip_solver_opts &o, const double *bvec, double *yvec,
                       hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  const KBufs<TV> k = kbufs<TV>(ctx, o.ksp_restart);
  const unsigned g256 = (unsigned)((k.N + 255) / 256);
  if (!std::is_same<TV, double>::value)
    KT(ctx, st, "k_convert_J", (double)ws.nnz * (8 + sizeof(TV)),
       k_convert<double, TV><<<(unsigned)((ws.nnz + 255) / 256), 256, 0, st>>>(ws.nnz, ws.J, k.Jk));
  k_gmres_init<<<1, 1, 0, st>>>(k.gd, ws.status, o.ksp_max_it);
  // coarse correction: Galerkin rows, group inverses and the level-2 inverse
  // for this Jacobian; each application then adds its prolongation to the
  // block-Jacobi result (the restriction is fused into the ILU sweeps).  Its
  // latency-bound kernels run on the side stream, beside the factor, and
  // join before the first preconditioner application.  With the FP64
  // factor, whose workgroups leave no CU room for the group inverses' LDS,
  // the rows and group inverses run first and only the one-workgroup level-2
  // inverse goes beside it.
  const CoarseLayout &C = ws.cl;
  auto rows_groups = [&](hipStream_t cs) {
    const int W = ctx->jell_w;
    const double brows = (double)ws.NB * W * (8 + 1) + (double)C.nb * kCQ * (4 + 8);
    if (ws.bj.bs == 64) KT(ctx, cs, "k_coarse_rows", brows, k_coarse_rows<64><<<C.nb, 64, 0, cs>>>(C, ws.NB, W, ws.J));
    else if (ws.bj.bs == 256) KT(ctx, cs, "k_coarse_rows", brows, k_coarse_rows<256><<<C.nb, 256, 0, cs>>>(C, ws.NB, W, ws.J));
    else KT(ctx, cs, "k_coarse_rows", brows, k_coarse_rows<512><<<C.nb, 512, 0, cs>>>(C, ws.NB, W, ws.J));
    KT(ctx, cs, "k_coarse_groups", (double)C.nb * kCQ * 16, k_coarse_groups<<<C.ng, kGroupThreads, 0, cs>>>(C));
  };
  const bool early = ws.coarse && sizeof(TV) == 8;
  if (early) rows_groups(st);
  const bool side = ws.coarse && hipEventRecord(ws.fork_ev, st) == hipSuccess &&
                    hipStreamWaitEvent(ws.side, ws.fork_ev, 0) == hipSuccess;
  if (ws.coarse) {
    const hipStream_t cs = side ? ws.side : st;
    if (!early) rows_groups(cs);
    KT(ctx, cs, "k_coarse_invert", (double)C.ng * C.ng * 16, k_coarse_invert<<<1, 1024, 0, cs>>>(C));
  }
  if (ws.bj.bs == 64) KT(ctx, st, "k_bj_factor", bytes_factor(ws, sizeof(TV)), k_bj_factor<TV, 64><<<ws.bj.nblocks, FactorLanes<64>::NT, 0, st>>>(ws.bj, ws.J, k.fac, ws.status));
  else if (ws.bj.bs == 256) KT(ctx, st, "k_bj_factor", bytes_factor(ws, sizeof(TV)), k_bj_factor<TV, 256><<<ws.bj.nblocks, FactorLanes<256>::NT, 0, st>>>(ws.bj, ws.J, k.fac, ws.status));
  else KT(ctx, st, "k_bj_factor", bytes_factor(ws, sizeof(TV)), k_bj_factor<TV, 512><<<ws.bj.nblocks, FactorLanes<512>::NT, 0, st>>>(ws.bj, ws.J, k.fac, ws.status));
  k_convert<double, TV><<<g256, 256, 0, st>>>(k.N, bvec, k.b);
  // The solution accumulates in FP64 in place (yvec) in both precisions, and
  // every restart recomputes the true residual b - J y with the FP64 Jacobian
  // (mixed-precision restarted GMRES): the FP32 Arnoldi cycles only ever
  // solve for a correction, so the FP32 form stops on the same residual as
  // the FP64 one.
  k_scale_copy<double><<<g256, 256, 0, st>>>(k.N, 0.0, bvec, yvec);
  if (side && (hipEventRecord(ws.join_ev, ws.side) != hipSuccess || hipStreamWaitEvent(st, ws.join_ev, 0) != hipSuccess)) {
    set_error("coarse setup join failed");
    (void)hipStreamSynchronize(ws.side);  // outside a capture this still orders it; a capture fails to end
  }
  pc_apply<TV>(ctx, k, k.b, k.w, nullptr, st);  // r0 = M^-1 b  (y0 = 0)
  norm2_to_h0<TV>(ctx, k.w, nullptr, st);
  k_cycle_start<<<1, 1, 0, st>>>(k.gd, ws.h, k.dg, k.m, o.ksp_rtol, o.ksp_atol, 1);
}

// On meshes of up to this many reduction chunks, step j > 0 finishes column
// j-1 (its norm and Givens update) inside the matvec: every workgroup re-reads
// the norm partials, which costs less than a launch while they are few.  (At
// 1M rows, 2048 partials per workgroup cost as much as they save.)
constexpr int kFuseNormChunks = 512;

template <class TV>
static void gmres_cycle(tdyhip_ctx *ctx, const tdyhip_solver_opts &o, const double *bvec, double *yvec,
                        hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  const KBufs<TV> k = kbufs<TV>(ctx, o.ksp_restart);
  const int m = k.m, nch = k.nch;
  const int64_t N = k.N;
  const unsigned g256 = (unsigned)((N + 255) / 256);
  GmresDev *gd = k.gd;
  const int *stop = &gd->done;
  // With the coarse correction, w = w_bj + P z1 (w_bj the block-Jacobi part
  // the fused matvec leaves): the multi-dot's first cd.ng workgroups solve for
  // z1 and add the coarse terms RV_i . z1 of h_i = V_i . w (RV_i the block sums
  // of V_i, formed by the matvec that formed V_i), the other workgroups the
  // terms V_i . w_bj; the update kernel adds the prolongation P z1 to w.
  Prolong pro;
  CoarseDot cd;
  if (ws.coarse) {
    pro.z1 = ws.cl.z1;
    pro.pos = ws.d_perm;
    pro.bs_log2 = ws.bj.bs == 64 ? 6 : ws.bj.bs == 256 ? 8 : 9;
    cd = coarse_dot(ws);
  }
  const int pst = nch + cd.ng;  // partials per vector: the coarse groups', then the row chunks'
  const unsigned gmd = (unsigned)pst;
  // The unnormalised Arnoldi vector alternates between w and z: step j reads
  // it from one and leaves the next in the other.
  for (int j = 0; j < m; j++) {
    TV *u = (j % 2 == 0) ? k.w : k.z, *ov = (j % 2 == 0) ? k.z : k.w;
    const NormArgs na{k.part, nch, j - 1, m, ws.h, k.dH, k.dcs, k.dsn, k.dg};
    const bool fuse_norm = nch <= kFuseNormChunks;
    // timer tags: the launches of step j work only if the cycle reached
    // column j; a fused matvec also finishes column j-1, so it counts with it
    ctx->kt.tag = (fuse_norm && j > 0) ? j - 1 : j;
    spmv_apply<TV>(ctx, k, u, k.V + (int64_t)j * N, ov, stop, (fuse_norm && j > 0) ? &na : nullptr,
                   ws.coarse ? ws.RV + (int64_t)j * ws.cl.nb : nullptr, st);
    ctx->kt.tag = j;
    // h = V^T o, o -= V h, ||o||^2, Givens update of column j.  (A single
    // kernel per reduction with a last-block hand-off is much slower here:
    // the device-scope fences it needs write back the per-XCD L2s.)
    const Prolong none;
    const double bmd = (double)(j + 2) * N * sizeof(TV);
    // FP32: two passes of 16 beyond 16 vectors (67 VGPRs against 83 for one of 32)
    if (j < 8) KT(ctx, st, "k_mdot", bmd, k_mdot<TV, 8><<<gmd, kRedBlock, 0, st>>>(N, j + 1, k.V, N, ov, k.part, pst, none, stop, cd));
    else if (j < 16 || sizeof(TV) == 4) KT(ctx, st, "k_mdot", bmd, k_mdot<TV, 16><<<gmd, kRedBlock, 0, st>>>(N, j + 1, k.V, N, ov, k.part, pst, none, stop, cd));
    else KT(ctx, st, "k_mdot", bmd, k_mdot<TV, 32><<<gmd, kRedBlock, 0, st>>>(N, j + 1, k.V, N, ov, k.part, pst, none, stop, cd));
    k_reduce_partials<TV><<<j + 1, kRedBlock, 0, st>>>(j + 1, pst, k.part, ws.h, stop);
    const double bmx = (double)(j + 3) * N * sizeof(TV) + (pro.z1 ? (double)N * 4 : 0.0);
    KT(ctx, st, "k_maxpy_norm", bmx, k_maxpy_norm<TV><<<nch, kRedBlock, 0, st>>>(N, j + 1, k.V, N, ws.h, ov, k.part, stop, pro));
    if (!fuse_norm || j == m - 1) k_norm_arnoldi<TV><<<1, kRedBlock, 0, st>>>(nch, k.part, gd, j, m, ws.h, k.dH, k.dcs, k.dsn, k.dg);
  }
  ctx->kt.tag = -1;
  if (m <= kHessLdsMaxM)
    k_hess_solve_lds<<<1, 256, sizeof(double) * (size_t)((m + 1) * m + m), st>>>(gd, m, k.dH, k.dg, k.dy);
  else
    k_hess_solve<<<1, 1, 0, st>>>(gd, m, k.dH, k.dg, k.dy);
  // (booked at the solution's read and write only: the basis vectors it reads depend on the cycle's length)
  KT(ctx, st, "k_gemv_update", (double)N * 16, k_gemv_update_dev<TV, double><<<g256, 256, 0, st>>>(N, m, k.V, N, k.dy, gd, yvec));
  // the next cycle's start, w = M^-1 (b - J y) in FP64 and its norm; nothing
  // once the solve has finished
  const int *fin = &gd->finished;
  ctx->kt.tag = kTagRestart;
  resid_pc_apply<TV>(ctx, k, bvec, yvec, k.w, fin, st);
  norm2_to_h0<TV>(ctx, k.w, fin, st);
  ctx->kt.tag = -1;
  k_cycle_start<<<1, 1, 0, st>>>(gd, ws.h, k.dg, m, o.ksp_rtol, o.ksp_atol, 0);
}

// The one-cycle graph of precision TV (captured on first use, replayed by
// the host-driven loop); its kernels only touch workspace arrays.
template <class TV>
static hipGraphExec_t cycle_graph(tdyhip_ctx *ctx, const tdyhip_solver_opts &o, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  hipGraphExec_t &cyc = ws.cycle_exec[std::is_same<TV, double>::value ? 0 : 1];
  if (cyc) return cyc;
  hipGraph_t gr = nullptr;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) return nullptr;
  gmres_cycle<TV>(ctx, o, ws.F, ws.Y, st);
  if (hipStreamEndCapture(st, &gr) != hipSuccess) return nullptr;
  if (hipGraphInstantiate(&cyc, gr, nullptr, nullptr, 0) != hipSuccess) cyc = nullptr;
  (void)hipGraphDestroy(gr);
  return cyc;
}

// Host-driven solve: the preparation, then one cycle at a time (replayed as
// the one-cycle graph when allowed) with a read of the cycle's state after
// each, until the solve has finished.
template <class TV>
static int gmres(tdyhip_ctx *ctx, const tdyhip_solver_opts &o, const double *bvec, double *yvec, int *its_out,
                 hipStream_t st, bool allow_graph) {
  SolverWS &ws = *ctx->ws;
  gmres_prep<TV>(ctx, o, bvec, yvec, st);
  SCHK(hipGetLastError());
  const bool graph = allow_graph && o.hip_graphs && !ctx->kt.on && bvec == ws.F && yvec == ws.Y;
  hipGraphExec_t cyc = graph ? cycle_graph<TV>(ctx, o, st) : nullptr;
  if (graph && !cyc) { set_error("GMRES cycle graph capture failed"); return TDYHIP_ERR_HIP; }
  GmresDev hg;
  int status = 0;
  for (int cycle = 0;; cycle++) {
    if (cyc) SCHK(hipGraphLaunch(cyc, st));
    else gmres_cycle<TV>(ctx, o, bvec, yvec, st);
    SCHK(hipGetLastError());
    SCHK(hipMemcpyAsync(&hg, ws.gdev, sizeof(GmresDev), hipMemcpyDeviceToHost, st));
    if (cycle == 0) SCHK(hipMemcpyAsync(&status, ws.status, sizeof(int), hipMemcpyDeviceToHost, st));
    SCHK(hipStreamSynchronize(st));
    kt_settle(ctx, hg.jlast, !hg.finished);
    if (status) { set_error("zero pivot in block-Jacobi ILU(0)"); return TDYHIP_ERR_SINGULAR; }
    if (hg.finished) break;
    if (ctx->cancel.load(std::memory_order_relaxed)) { set_error("solve cancelled"); return TDYHIP_ERR_CANCELLED; }
  }
  *its_out = hg.its;
  // KSP_DIVERGED_ITS: PETSc's SNES counts it as a linear-solve failure
  // (SNESCheckKSPSolve, maxLinearSolveFailures 1 -> SNES_DIVERGED_LINEAR_SOLVE)
  if (hg.diverged) { set_error("GMRES residual is not finite"); return TDYHIP_ERR_NOCONV; }
  if (!hg.converged) { set_error("GMRES reached ksp_max_it"); return TDYHIP_ERR_NOCONV; }
  return 0;
}

}  // namespace tdyhip

namespace tdyhip {

struct Form {
  int type;             // 0 TS backward Euler, 1 SNES
  const double *Uprev;  // TS only
  double dt;            // TS only
};

// Udot = (U - Uprev) / dt with 1 / dt read from device memory (dyn[1])
__global__ void k_udot(int64_t n, const double *__restrict__ dyn, const double *__restrict__ U,
                       const double *__restrict__ Up, double *__restrict__ Ud, const int *__restrict__ stop) {
  if (stop && *stop) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Ud[i] = (U[i] - Up[i]) * dyn[1];
}

static int fuser_eval(tdyhip_ctx *ctx, int kind, const double *X, double *out, const int *stop);

// The solver's residual and Jacobian launches read dtime and shift from
// ws.dyn (so a replayed graph follows dt) and honour the given stop flag.  A
// context attached to an evaluation fuser hands its SNES-form evaluations to
// the fuser's fused launches instead.
static int eval_F(tdyhip_ctx *ctx, const Form &fm, const double *X, double *F, hipStream_t st,
                  const int *stop = nullptr) {
  if (ctx->fuser && fm.type == 1) return fuser_eval(ctx, 0, X, F, stop);
  SolverWS &ws = *ctx->ws;
  ctx->launch_dyn = ws.dyn;
  ctx->launch_stop = stop;
  int rc;
  if (fm.type == 0) {
    k_udot<<<nblk(ctx->nc, 256), 256, 0, st>>>(ctx->nc, ws.dyn, X, fm.Uprev, ws.Udot, stop);
    rc = launch_state(ctx, 0, X, ws.Udot, st);
    if (!rc) rc = launch_flux(ctx, 0, 0.0, X, ws.Udot, F, nullptr, st);
  } else {
    rc = launch_state(ctx, 1, X, nullptr, st);
    if (!rc) rc = launch_flux(ctx, 1, 0.0, X, nullptr, F, nullptr, st);
  }
  ctx->launch_dyn = nullptr;
  ctx->launch_stop = nullptr;
  return rc;
}

static int eval_J(tdyhip_ctx *ctx, const Form &fm, const double *X, double *J, hipStream_t st) {
  if (ctx->fuser && fm.type == 1) return fuser_eval(ctx, 1, X, J, nullptr);
  SolverWS &ws = *ctx->ws;
  ctx->launch_dyn = ws.dyn;
  int rc;
  if (fm.type == 0) {
    k_udot<<<nblk(ctx->nc, 256), 256, 0, st>>>(ctx->nc, ws.dyn, X, fm.Uprev, ws.Udot, nullptr);
    rc = launch_flux(ctx, 0, 0.0, X, ws.Udot, nullptr, J, st, true, ws.d_perm, ws.bj.bs, ws.bj.d_row, ws.NB);
  } else {
    rc = launch_flux(ctx, 1, 0.0, X, nullptr, nullptr, J, st, true, ws.d_perm, ws.bj.bs, ws.bj.d_row, ws.NB);
  }
  ctx->launch_dyn = nullptr;
  return rc;
}

// ---------------------------------------------------------------- Newton tail on the device
// After the linear solve of a Newton iteration: the step length checks, the
// line search's full step and its sufficient-decrease test, and the step
// itself, with no host round trip (PETSc 3.17 SNESNEWTONLS + SNESLINESEARCHBT,
// the first trial of the search).  The host reads NewtonDev once per
// iteration; a rejected full step is backtracked on the host.
struct NewtonDev {
  double fnorm;      // ||F|| at the current iterate, kept on the device across iterations
  double ynorm;      // ||Y|| after the maxstep cut
  double yscale;     // factor applied to Y (1, or maxstep / ||Y||)
  double initslope;  // -(F, J Y) with PETSc's guards
  double gnorm;      // ||F(X - Y)||
  double xnorm;      // ||X|| once the iteration is over
  int skip;          // the tail does nothing: the linear solve failed, or a zero or non-finite step
  int not_accepted;  // 0 once the full step passed the sufficient-decrease test and was taken
  int lin_ok;        // the linear solve converged
  int ystate;        // 0 a step, 1 a zero step, 2 a non-finite step
  int its, finished, converged, diverged, pc_status;  // the linear solve's outcome
  int pad;
};
static_assert(sizeof(NewtonDev) <= 256, "NewtonDev must fit the workspace buffer allocated for it");

__global__ void k_nt_gate(const GmresDev *__restrict__ gd, const int *__restrict__ status, NewtonDev *nd) {
  nd->lin_ok = gd->finished && gd->converged && *status == 0;
  nd->skip = !nd->lin_ok;
  nd->not_accepted = 1;
  nd->ystate = 0;
  nd->yscale = 1.0;
}

__global__ void k_nt_ynorm(NewtonDev *nd, const double *__restrict__ h, double maxstep) {
  if (nd->skip) return;
  const double yn = sqrt(h[0]);
  if (!isfinite(yn)) { nd->ystate = 2; nd->skip = 1; return; }
  if (yn == 0.0) { nd->ystate = 1; nd->skip = 1; nd->ynorm = 0.0; return; }
  nd->yscale = yn > maxstep ? maxstep / yn : 1.0;
  nd->ynorm = yn > maxstep ? maxstep : yn;
}

__global__ void k_nt_scale(int64_t n, const NewtonDev *__restrict__ nd, double *__restrict__ Y) {
  if (nd->skip || nd->yscale == 1.0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Y[i] = nd->yscale * Y[i];
}

__global__ void k_nt_slope(NewtonDev *nd, const double *__restrict__ h) {
  if (nd->skip) return;
  nd->initslope = bt_initslope(h[0]);
}

__global__ void k_nt_accept(NewtonDev *nd, const double *__restrict__ h, double alpha) {
  if (nd->skip) return;
  const double g = sqrt(h[0]);
  nd->gnorm = g;
  if (bt_sufficient(nd->fnorm, g, 1.0, nd->initslope, alpha)) {
    nd->not_accepted = 0;
    nd->fnorm = g;
  }
}

__global__ void k_copy_gated(int64_t n, const double *__restrict__ src, double *__restrict__ dst,
                             const int *__restrict__ stop) {
  if (*stop) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void k_nt_final(NewtonDev *nd, const double *__restrict__ h, const GmresDev *__restrict__ gd,
                           const int *__restrict__ status) {
  nd->xnorm = sqrt(h[0]);
  nd->its = gd->its;
  nd->finished = gd->finished;
  nd->converged = gd->converged;
  nd->diverged = gd->diverged;
  nd->pc_status = *status;
}

__global__ void k_nt_set_fnorm(NewtonDev *nd, const double *__restrict__ h) { nd->fnorm = sqrt(h[0]); }
__global__ void k_set_scalar(double *p, double v) { *p = v; }
__global__ void k_set_dyn(double *dyn, double d0, double d1) {
  dyn[0] = d0;
  dyn[1] = d1;
}

// nh[slot] = <x, y> (skipped with the stop flag)
static void dot_to(tdyhip_ctx *ctx, const double *x, const double *y, int slot, const int *stop, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  KT(ctx, st, "k_dot", (double)ws.n * (x == y ? 8 : 16), k_dot<double, double><<<ws.nchunk, kRedBlock, 0, st>>>(ws.n, x, y, ws.dpartial, stop));
  k_reduce_partials<double><<<1, kRedBlock, 0, st>>>(1, ws.nchunk, ws.dpartial, ws.nh + slot, stop);
}

constexpr double kMaxStep = 1e8, kBTAlpha = 1e-4;  // PETSc 3.17 SNESLINESEARCHBT defaults

// The tail in three parts: up to the full step W = X + Y, the residual at W
// (ws.G, skipped when the tail is), and from its norm on.
static void newton_tail_pre(tdyhip_ctx *ctx, double *X, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  NewtonDev *nd = ws.nd;
  const int64_t n = ctx->nc;
  const unsigned g = nblk(n, 256);
  k_nt_gate<<<1, 1, 0, st>>>(ws.gdev, ws.status, nd);
  dot_to(ctx, ws.Y, ws.Y, 0, &nd->skip, st);
  k_nt_ynorm<<<1, 1, 0, st>>>(nd, ws.nh, kMaxStep);
  k_nt_scale<<<g, 256, 0, st>>>(n, nd, ws.Y);
  spmv_blk64(ctx, ws, ws.Y, ws.JY, st, &nd->skip);
  dot_to(ctx, ws.F, ws.JY, 1, &nd->skip, st);
  k_nt_slope<<<1, 1, 0, st>>>(nd, ws.nh + 1);
  KT(ctx, st, "k_waxpy", (double)n * 24, k_waxpy<<<g, 256, 0, st>>>(n, 1.0, X, ws.Y, ws.W, &nd->skip));
}

static void newton_tail_post(tdyhip_ctx *ctx, double *X, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  NewtonDev *nd = ws.nd;
  const int64_t n = ctx->nc;
  const unsigned g = nblk(n, 256);
  dot_to(ctx, ws.G, ws.G, 2, &nd->skip, st);
  k_nt_accept<<<1, 1, 0, st>>>(nd, ws.nh + 2, kBTAlpha);
  k_copy_gated<<<g, 256, 0, st>>>(n, ws.W, X, &nd->not_accepted);
  k_copy_gated<<<g, 256, 0, st>>>(n, ws.G, ws.F, &nd->not_accepted);
  dot_to(ctx, X, X, 3, nullptr, st);
  k_nt_final<<<1, 1, 0, st>>>(nd, ws.nh + 3, ws.gdev, ws.status);
}

static int newton_tail(tdyhip_ctx *ctx, const Form &fm, double *X, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  newton_tail_pre(ctx, X, st);
  int rc = eval_F(ctx, fm, ws.W, ws.G, st, &ws.nd->skip);
  if (rc) return rc;
  newton_tail_post(ctx, X, st);
  return hipGetLastError() == hipSuccess ? 0 : TDYHIP_ERR_HIP;
}

// The whole Newton iteration as one launch sequence: Jacobian, preparation,
// `cycles` restart cycles (those past the end of the solve return at once)
// and the tail.
static int newton_iter_launch(tdyhip_ctx *ctx, const Form &fm, double *X, const tdyhip_solver_opts &ok, int fp32,
                              int cycles, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  int rc = eval_J(ctx, fm, X, ws.J, st);
  if (rc) return rc;
  if (fp32) {
    gmres_prep<float>(ctx, ok, ws.F, ws.Y, st);
    for (int c = 0; c < cycles; c++) gmres_cycle<float>(ctx, ok, ws.F, ws.Y, st);
  } else {
    gmres_prep<double>(ctx, ok, ws.F, ws.Y, st);
    for (int c = 0; c < cycles; c++) gmres_cycle<double>(ctx, ok, ws.F, ws.Y, st);
  }
  return newton_tail(ctx, fm, X, st);
}

// Captured graphs keyed by the iterate's pointer, the form, the Krylov
// precision and the cycle count (cycles = 0: the tail alone).
static hipGraphExec_t newton_graph(tdyhip_ctx *ctx, const Form &fm, double *X, const tdyhip_solver_opts &ok,
                                   int fp32, int cycles, bool tail_only, hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  const int key_c = tail_only ? 0 : cycles;
  for (auto &e : ws.newton_graphs)
    if (e.X == X && e.type == fm.type && e.fp32 == fp32 && e.cycles == key_c) return e.exec;
  if (ws.newton_graphs.size() >= kMaxNewtonGraphs) {
    (void)hipStreamSynchronize(st);  // the oldest graph may still be in flight
    (void)hipGraphExecDestroy(ws.newton_graphs.front().exec);
    ws.newton_graphs.erase(ws.newton_graphs.begin());
  }
  hipGraph_t gr = nullptr;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) return nullptr;
  const int rc = tail_only ? newton_tail(ctx, fm, X, st) : newton_iter_launch(ctx, fm, X, ok, fp32, cycles, st);
  const hipError_t ec = hipStreamEndCapture(st, &gr);
  if (rc || ec != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    return nullptr;
  }
  hipGraphExec_t ex = nullptr;
  if (hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) != hipSuccess) ex = nullptr;
  (void)hipGraphDestroy(gr);
  if (ex) ws.newton_graphs.push_back({X, fm.type, fp32, key_c, ex});
  return ex;
}

static int read_newton(tdyhip_ctx *ctx, NewtonDev &h, hipStream_t st) {
  SCHK(hipMemcpyAsync(&h, ctx->ws->nd, sizeof(NewtonDev), hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  return 0;
}

// FP32 Krylov: the FP32 attempt gets at most kFp32MaxIt iterations; if it
// does not converge (or its ILU(0) meets a zero pivot), the same Newton
// correction is solved again in FP64 from scratch.
constexpr int kFp32MaxIt = 1000;

// One Newton iteration up to and including the full step's test: on return h
// holds the linear solve's outcome, the step length and, when the solve
// converged, the residual norm at the full step and whether it was taken.
static int newton_iteration(tdyhip_ctx *ctx, const Form &fm, double *X, const tdyhip_solver_opts &o,
                            tdyhip_solver_stats &stats, hipStream_t st, NewtonDev &h) {
  SolverWS &ws = *ctx->ws;
  const int fp32 = o.krylov_fp32 ? 1 : 0;
  tdyhip_solver_opts ok = o;
  if (fp32) ok.ksp_max_it = std::min(o.ksp_max_it, kFp32MaxIt);
  // (a context on an evaluation fuser has its evaluations launched by the
  // fuser's thread, outside any graph of its own)
  const bool graphs = o.hip_graphs && !ctx->kt.on && !ctx->fuser;
  int rc;
  if (graphs) {
    // Jacobian, preparation, the cycles the last iteration needed and the
    // tail in one replay; a solve that needs more cycles continues one cycle
    // graph at a time, then replays the tail alone.
    const int C = std::max(1, std::min(ws.cycles_hint, kMaxGraphCycles));
    hipGraphExec_t ex = newton_graph(ctx, fm, X, ok, fp32, C, false, st);
    if (!ex) { set_error("Newton iteration graph capture failed"); return TDYHIP_ERR_HIP; }
    SCHK(hipGraphLaunch(ex, st));
    if ((rc = read_newton(ctx, h, st))) return rc;
    if (!h.finished && !h.pc_status) {
      hipGraphExec_t cyc = fp32 ? cycle_graph<float>(ctx, ok, st) : cycle_graph<double>(ctx, ok, st);
      if (!cyc) { set_error("GMRES cycle graph capture failed"); return TDYHIP_ERR_HIP; }
      GmresDev hg;
      do {
        if (ctx->cancel.load(std::memory_order_relaxed)) { set_error("solve cancelled"); return TDYHIP_ERR_CANCELLED; }
        SCHK(hipGraphLaunch(cyc, st));
        SCHK(hipMemcpyAsync(&hg, ws.gdev, sizeof(GmresDev), hipMemcpyDeviceToHost, st));
        SCHK(hipStreamSynchronize(st));
      } while (!hg.finished);
      hipGraphExec_t tail = newton_graph(ctx, fm, X, ok, fp32, 0, true, st);
      if (!tail) { set_error("Newton tail graph capture failed"); return TDYHIP_ERR_HIP; }
      SCHK(hipGraphLaunch(tail, st));
      if ((rc = read_newton(ctx, h, st))) return rc;
    }
  } else {
    if ((rc = eval_J(ctx, fm, X, ws.J, st))) return rc;
    int its = 0;
    rc = fp32 ? gmres<float>(ctx, ok, ws.F, ws.Y, &its, st, false) : gmres<double>(ctx, ok, ws.F, ws.Y, &its, st, false);
    if (rc && rc != TDYHIP_ERR_SINGULAR && rc != TDYHIP_ERR_NOCONV) return rc;
    if ((rc = newton_tail(ctx, fm, X, st)) || (rc = read_newton(ctx, h, st))) return rc;
  }
  int lits = h.its;
  ws.cycles_hint = std::max(1, (h.its + o.ksp_restart - 1) / o.ksp_restart);
  if (fp32 && !h.lin_ok) {
    // the FP32 attempt failed: the same correction in FP64, host-driven
    stats.krylov_fallbacks++;
    int its = 0;
    rc = gmres<double>(ctx, o, ws.F, ws.Y, &its, st, true);
    if (rc && rc != TDYHIP_ERR_SINGULAR && rc != TDYHIP_ERR_NOCONV) return rc;
    if (graphs) {
      hipGraphExec_t tail = newton_graph(ctx, fm, X, o, 0, 0, true, st);
      if (!tail) { set_error("Newton tail graph capture failed"); return TDYHIP_ERR_HIP; }
      SCHK(hipGraphLaunch(tail, st));
    } else if ((rc = newton_tail(ctx, fm, X, st))) {
      return rc;
    }
    if ((rc = read_newton(ctx, h, st))) return rc;
    lits += its;
  }
  stats.linear_its += lits;
  return 0;
}

// SNESSolve_NEWTONLS with SNESLINESEARCHBT and SNESConvergedDefault
// (PETSc 3.17 defaults: atol 1e-50, stol 1e-8, maxit 50, divtol 1e4,
// line-search alpha 1e-4, minlambda 1e-12, maxstep 1e8).  Each iteration is
// one replayed graph and one read of its outcome (newton_iteration); only a
// full step that fails the sufficient-decrease test is backtracked from the
// host.
static int newton(tdyhip_ctx *ctx, const Form &fm, double *X, const tdyhip_solver_opts &o, tdyhip_solver_stats &stats,
                  hipStream_t st) {
  SolverWS &ws = *ctx->ws;
  const int n = ctx->nc;
  const unsigned g = nblk(n, 256);
  const double minlambda = 1e-12, divtol = 1e4;
  int rc;
  k_set_dyn<<<1, 1, 0, st>>>(ws.dyn, ctx->dtime, fm.type == 0 ? 1.0 / fm.dt : 0.0);
  if ((rc = eval_F(ctx, fm, X, ws.F, st))) return rc;
  dot_to(ctx, ws.F, ws.F, 0, nullptr, st);
  k_nt_set_fnorm<<<1, 1, 0, st>>>(ws.nd, ws.nh);
  double fnorm2;
  SCHK(hipMemcpyAsync(&fnorm2, ws.nh, sizeof(double), hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  double fnorm = std::sqrt(fnorm2);
  const double fnorm0 = fnorm, ttol = o.snes_rtol * fnorm;
  int status = 0, it = 0;
  if (!std::isfinite(fnorm)) { stats.status = -4; return 0; }  // SNES_DIVERGED_FNORM_NAN
  if (fnorm < o.snes_atol) { stats.status = 0; stats.last_fnorm = fnorm; return 0; }
  for (it = 1; it <= o.snes_max_it; it++) {
    if (ctx->cancel.load(std::memory_order_relaxed)) { set_error("solve cancelled"); return TDYHIP_ERR_CANCELLED; }
    NewtonDev h;
    if ((rc = newton_iteration(ctx, fm, X, o, stats, st, h))) return rc;
    // a failed or non-finite linear solve (KSP_DIVERGED_*: SNES_DIVERGED_LINEAR_SOLVE)
    if (!h.lin_ok || h.ystate == 2) { status = -3; break; }
    if (h.ystate == 1) break;  // a zero step: converged
    double ynorm = h.ynorm, lambda = 1.0, gnorm = h.gnorm, xnorm = h.xnorm;
    if (h.not_accepted) {
      // backtracking from the full step; a NaN or Inf residual ends it as a
      // failure (SNES_LINESEARCH_FAILED_NANORINF)
      auto trial = [&](double lam, double *gn) -> int {
        KT(ctx, st, "k_waxpy", (double)n * 24, k_waxpy<<<g, 256, 0, st>>>(n, lam, X, ws.Y, ws.W));
        int r2 = eval_F(ctx, fm, ws.W, ws.G, st);
        return r2 ? r2 : dev_norm(ctx, ws.G, gn, st);
      };
      auto stop = [&]() -> int {
        if (!ctx->cancel.load(std::memory_order_relaxed)) return 0;
        set_error("solve cancelled");
        return TDYHIP_ERR_CANCELLED;
      };
      BTParams btp;
      btp.alpha = kBTAlpha;
      btp.minlambda = minlambda;
      BTResult bt;
      if ((rc = bt_search(fnorm, h.initslope, btp, true, h.gnorm, trial, stop, bt))) return rc;
      if (!bt.accepted) {
        // SNES_CONVERGED_SNORM_RELATIVE when the step is negligible, else
        // SNES_DIVERGED_LINE_SEARCH (X keeps the last accepted iterate)
        status = (ynorm < o.snes_stol * h.xnorm) ? 0 : -6;
        break;
      }
      lambda = bt.lambda;
      gnorm = bt.gnorm;
      SCHK(hipMemcpyAsync(X, ws.W, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      SCHK(hipMemcpyAsync(ws.F, ws.G, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      k_set_scalar<<<1, 1, 0, st>>>(&ws.nd->fnorm, gnorm);
      if ((rc = dev_norm(ctx, X, &xnorm, st))) return rc;
    }
    fnorm = gnorm;
    ynorm *= lambda;
    if (!std::isfinite(fnorm)) { status = -4; break; }
    if (fnorm < o.snes_atol) break;
    if (fnorm <= ttol) break;
    if (ynorm < o.snes_stol * xnorm) break;
    if (fnorm > divtol * fnorm0) { status = -2; break; }
  }
  if (it > o.snes_max_it) { status = -5; it = o.snes_max_it; }
  stats.newton_its += it;
  stats.status = status;
  stats.last_fnorm = fnorm;
  return 0;
}

}  // namespace tdyhip

extern "C" {

void tdyhip_solver_defaults(tdyhip_solver_opts *o) {
  if (!o) return;
  o->snes_rtol = 1e-8; o->snes_atol = 1e-50; o->snes_stol = 1e-8; o->snes_max_it = 50;
  o->ksp_rtol = 1e-5; o->ksp_atol = 1e-50; o->ksp_max_it = 10000; o->ksp_restart = 30;
  o->krylov_fp32 = 0; o->pc_block = 256; o->hip_graphs = 1; o->pc_coarse = 0;
}

static int check_opts(const tdyhip_solver_opts *o) {
  if (!o) { tdyhip::set_error("null solver options"); return TDYHIP_ERR_ARG; }
  if (o->ksp_restart < 1 || o->ksp_restart > 200 || o->pc_block < 1 || o->pc_block > tdyhip::kBlkMax ||
      o->snes_max_it < 0 || o->ksp_max_it < 1) {
    tdyhip::set_error("bad solver options");
    return TDYHIP_ERR_ARG;
  }
  return 0;
}

int tdyhip_device_bytes(const tdyhip_ctx *ctx, int64_t *bytes) {
  using namespace tdyhip;
  if (!ctx || !bytes) { set_error("bad argument"); return TDYHIP_ERR_ARG; }
  *bytes = ctx->dev_bytes + (ctx->ws ? ctx->ws->bytes : 0);
  return 0;
}

int tdyhip_kernel_timers(tdyhip_ctx *ctx, int32_t enable) {
  using namespace tdyhip;
  if (!ctx) { set_error("null context"); return TDYHIP_ERR_ARG; }
  kt_collect(ctx);
  if (enable) ctx->kt.recs.clear();
  ctx->kt.on = enable != 0;
  return 0;
}

int tdyhip_kernel_timer_report(tdyhip_ctx *ctx, char *buf, int64_t buflen) {
  using namespace tdyhip;
  if (!ctx || !buf || buflen < 3) { set_error("bad argument"); return TDYHIP_ERR_ARG; }
  kt_collect(ctx);
  std::string js = "[";
  char line[512];
  for (size_t i = 0; i < ctx->kt.recs.size(); i++) {
    const auto &r = ctx->kt.recs[i];
    snprintf(line, sizeof line, "%s{\"kernel\":\"%s\",\"launches\":%lld,\"ms\":%.6f,\"bytes\":%.6e}", i ? "," : "",
             r.first.c_str(), (long long)r.second.launches, r.second.ms, r.second.bytes);
    js += line;
  }
  js += "]";
  if ((int64_t)js.size() + 1 > buflen) { set_error("report buffer too small"); return TDYHIP_ERR_ARG; }
  std::memcpy(buf, js.c_str(), js.size() + 1);
  return 0;
}

int tdyhip_set_pc_partition(tdyhip_ctx *ctx, const int32_t *block_of_cell) {
  using namespace tdyhip;
  if (!ctx) { set_error("null context"); return TDYHIP_ERR_ARG; }
  if (block_of_cell) {
    for (int c = 0; c < ctx->nc; c++)
      if (block_of_cell[c] < 0 || block_of_cell[c] >= ctx->nc) { set_error("block id out of range"); return TDYHIP_ERR_ARG; }
    ctx->pc_part.assign(block_of_cell, block_of_cell + ctx->nc);
  } else {
    ctx->pc_part.clear();
  }
  ctx->pc_part_version++;
  return 0;
}

int tdyhip_newton_snes(tdyhip_ctx *ctx, double *U, const tdyhip_solver_opts *o, tdyhip_solver_stats *st) {
  using namespace tdyhip;
  if (!ctx || !U || !st) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  int rc = check_opts(o);
  if (rc || (rc = ensure_ws(ctx, *o))) return rc;
  *st = tdyhip_solver_stats{};
  Form fm{1, nullptr, 0.0};
  rc = newton(ctx, fm, U, *o, *st, ctx->stream);
  if (!rc) { SCHK(hipStreamSynchronize(ctx->stream)); }
  return rc;
}

int tdyhip_run_snes_steps(tdyhip_ctx *ctx, double *U, double dt, int32_t nsteps, const tdyhip_solver_opts *o,
                          tdyhip_solver_stats *st) {
  using namespace tdyhip;
  if (!ctx || !U || !st || nsteps < 0) { set_error("bad argument"); return TDYHIP_ERR_ARG; }
  int rc = check_opts(o);
  if (rc || (rc = ensure_ws(ctx, *o))) return rc;
  if ((rc = tdyhip_set_dtime(ctx, dt))) return rc;
  *st = tdyhip_solver_stats{};
  Form fm{1, nullptr, 0.0};
  for (int s = 0; s < nsteps; s++) {
    if ((rc = tdyhip_snes_presolve(ctx, U, ctx->stream))) return rc;
    if ((rc = newton(ctx, fm, U, *o, *st, ctx->stream))) return rc;
    st->steps++;
    if (st->status) { set_error("Newton did not converge"); return TDYHIP_ERR_NOCONV; }
  }
  SCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int tdyhip_run_ts_be(tdyhip_ctx *ctx, double *U, double dt, int32_t nsteps, const tdyhip_solver_opts *o,
                     tdyhip_solver_stats *st) {
  using namespace tdyhip;
  if (!ctx || !U || !st || nsteps < 0 || !(dt > 0)) { set_error("bad argument"); return TDYHIP_ERR_ARG; }
  int rc = check_opts(o);
  if (rc || (rc = ensure_ws(ctx, *o))) return rc;
  *st = tdyhip_solver_stats{};
  SolverWS &ws = *ctx->ws;
  Form fm{0, ws.Uprev_ts, dt};
  for (int s = 0; s < nsteps; s++) {
    SCHK(hipMemcpyAsync(ws.Uprev_ts, U, sizeof(double) * ctx->nc, hipMemcpyDeviceToDevice, ctx->stream));
    if ((rc = newton(ctx, fm, U, *o, *st, ctx->stream))) return rc;
    st->steps++;
    if (st->status) { set_error("Newton did not converge"); return TDYHIP_ERR_NOCONV; }
  }
  SCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int tdyhip_cancel(tdyhip_ctx *ctx, int32_t on) {
  if (!ctx) { set_error("null context"); return TDYHIP_ERR_ARG; }
  ctx->cancel.store(on ? 1 : 0, std::memory_order_relaxed);
  return 0;
}

int tdyhip_run_tdyti(tdyhip_ctx *ctx, double *U, double dt_init, double growth, double dt_max, double final_time,
                     const tdyhip_solver_opts *o, tdyhip_solver_stats *st) {
  using namespace tdyhip;
  if (!ctx || !U || !st || !(dt_init > 0)) { set_error("bad argument"); return TDYHIP_ERR_ARG; }
  int rc = check_opts(o);
  if (rc || (rc = ensure_ws(ctx, *o))) return rc;
  *st = tdyhip_solver_stats{};
  SolverWS &ws = *ctx->ws;
  Form fm{1, nullptr, 0.0};
  double time = 0.0, dt = dt_init, dt_save = 0.0;
  int icut = 0;
  while (time < final_time) {
    const double dt_for_sync = final_time - time;
    if (dt > dt_for_sync) { dt_save = dt; dt = dt_for_sync; }
    if ((rc = tdyhip_set_dtime(ctx, dt))) return rc;
    bool conv = false;
    while (!conv) {
      SCHK(hipMemcpyAsync(ws.Usave, U, sizeof(double) * ctx->nc, hipMemcpyDeviceToDevice, ctx->stream));
      if ((rc = tdyhip_snes_presolve(ctx, U, ctx->stream))) return rc;
      if ((rc = newton(ctx, fm, U, *o, *st, ctx->stream))) return rc;
      if (st->status < 0) {
        icut++;
        st->cuts++;
        dt *= 0.25;
        if ((rc = tdyhip_set_dtime(ctx, dt))) return rc;
        SCHK(hipMemcpyAsync(U, ws.Usave, sizeof(double) * ctx->nc, hipMemcpyDeviceToDevice, ctx->stream));
        if (icut == 25) { set_error("time step cut 25 times"); return TDYHIP_ERR_NOCONV; }
      } else {
        conv = true;
        time += dt;
        st->steps++;
      }
    }
    if (dt_save > 0.) { time = final_time; dt = dt_save; dt_save = 0.; }
    dt *= growth;
    if (dt > dt_max) dt = dt_max;
  }
  SCHK(hipStreamSynchronize(ctx->stream));
  return 0;
}

// Test hook, not part of the interface in include/tdyhip.h: Ainv = A^-1 for
// one dense n x n row-major matrix by the coarse correction's register-tiled
// Gauss-Jordan at width 64 (the group inverses, n <= 64) or 128 (the level-2
// inverse, n <= 128), on the current device.  A zero pivot returns
// TDYHIP_ERR_SINGULAR.
int tdyhip_test_dense_inverse(const double *A, int32_t n, int32_t width, double *Ainv) {
  using namespace tdyhip;
  if (!A || !Ainv || (width != kGMax && width != kNG) || n < 1 || n > width) { set_error("bad argument"); return TDYHIP_ERR_ARG; }
  const size_t bytes = sizeof(double) * (size_t)n * n;
  double *dA = nullptr;
  int *dst = nullptr;
  int rc = 0, status = 0;
  if (hipMalloc((void **)&dA, bytes) != hipSuccess) { set_error("hipMalloc failed"); return TDYHIP_ERR_MEM; }
  if (hipMalloc((void **)&dst, sizeof(int)) != hipSuccess) {
    (void)hipFree(dA);
    set_error("hipMalloc failed");
    return TDYHIP_ERR_MEM;
  }
  if (hipMemcpy(dA, A, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dst, &status, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    rc = TDYHIP_ERR_HIP;
  } else {
    if (width == kGMax) k_test_dense_invert<kGMax><<<1, TiledGJ<kGMax>::NT>>>(dA, n, dst);
    else k_test_dense_invert<kNG><<<1, TiledGJ<kNG>::NT>>>(dA, n, dst);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(&status, dst, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(Ainv, dA, bytes, hipMemcpyDeviceToHost) != hipSuccess)
      rc = TDYHIP_ERR_HIP;
  }
  (void)hipFree(dA);
  (void)hipFree(dst);
  if (rc) { set_error("dense inverse failed"); return rc; }
  if (status) { set_error("singular matrix (zero pivot)"); return TDYHIP_ERR_SINGULAR; }
  return 0;
}

}  // extern "C"
