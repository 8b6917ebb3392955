// Newton's inner linear solve on the device (subsystem (d)): ELL SpMV,
// block-Jacobi ILU(0) and the BLAS-1 pieces of GMRES(m), templated on the
// Krylov precision.
//
// The reference leaves this to PETSc (KSPGMRES, restart 30, left
// preconditioning, classical Gram-Schmidt; PCILU(0) serial / PCBJACOBI+ILU(0)
// parallel; driven from src/tdyti.c:135,188).
//
// Block-Jacobi ILU(0): cells are partitioned (host) into compact blocks of at
// most BS cells (BS = 64: one wavefront per block; BS = 256 or 512: one 4- or
// 8-wave workgroup per block; at 512 the FP64 factor holds 137 KB of LDS).  Inside a block rows keep their global order;
// entries coupling different blocks are dropped.  The factorization (IKJ)
// and the two triangular solves advance level by level (levels of the
// in-block dependency DAG, computed on the host) with the block's rows in
// LDS, so no grid-wide synchronisation is ever needed.  The Jacobian is held
// slot-major (slot s of row i at s*n + i) so every sweep is coalesced.
#include "tdyhip_internal.h"

namespace tdyhip {

constexpr int kBlkMax = 512;  // max rows per block
constexpr int kW = 27;        // max in-block entries per row (hexahedral vertex adjacency)
constexpr int kLU = 13;       // max strictly lower (upper) entries per row
// Factored block layout per row (slot-major, stride BS): slots 0..12 L
// multipliers, 13..25 U entries, 26 the inverse diagonal; padded L/U slots hold
// 0 and point at the zero entry x[BS] of the apply kernel's LDS vector.

// Per-block arrays are strided by the block width BS (64, 256 or 512).
struct BJLayout {
  int nblocks = 0;
  int bs = 0;                  // stride / launch width (64, 256 or 512)
  int32_t *d_row = nullptr;    // nblocks*bs global row (-1 pad)
  int32_t *d_jpos = nullptr;   // nblocks*kW*bs position in the ELL Jacobian less the block's first entry (-1 pad)
  int ell_w = 0;               // ELL width W: block b's entries start at b*W*bs
  uint16_t *d_lcol = nullptr;  // nblocks*kW*bs local column (0xffff pad)
  uint8_t *d_diag = nullptr;   // nblocks*bs slot of the diagonal
  uint8_t *d_nslot = nullptr;  // nblocks*bs number of slots
  uint16_t *d_levL = nullptr;  // nblocks*bs
  uint16_t *d_levU = nullptr;  // nblocks*bs
  int32_t *d_nlev = nullptr;   // nblocks*2 (L, U level counts)
  int32_t *d_nlu = nullptr;    // nblocks: L/U slots in use (the factor writes and the sweeps read only
                               // these); nullptr when every block uses all kLU
  int32_t *d_bw = nullptr;     // nblocks: ELL slots in use (the block's longest row; the SpMV reads only
                               // these); nullptr when every block reads all W
  uint16_t *d_cL = nullptr;    // nblocks*kLU*bs local column of L slot (bs = zero entry)
  uint16_t *d_cU = nullptr;    // nblocks*kLU*bs local column of U slot
  // ILU(0) update lists: for row i and lower slot p (row k), pairs (d, s) with
  // a_i[d] -= l_ik * a_k[s]; npair[p] pairs per p, each packed d | s << 8 in
  // 16 bits, pairs q and q+1 of slot p sharing the 32-bit word p*kPairW + q/2.
  uint8_t *d_npair = nullptr;  // nblocks*kLU*bs
  uint32_t *d_pair = nullptr;  // nblocks*kPair*bs (kPair = kLU*kPairW words per row)
  // Shared index patterns: every index table above except d_row (and the
  // solver's int16 column offsets) holds one entry per pattern, block b using
  // pattern d_pat[b]; nullptr: every block has its own.
  int32_t *d_pat = nullptr;    // nblocks
  int npat = 0;
};

constexpr int kPairW = (kLU + 1) / 2;  // 32-bit words per lower slot (two 16-bit pairs each)
constexpr int kPair = kLU * kPairW;    // pair words per row

// ---------------------------------------------------------------- device-resident GMRES cycle
// State shared by the kernels of one restart cycle; `done` makes every later
// kernel of the cycle return at once, so the host launches a whole cycle and
// synchronises once.
struct GmresDev {
  double tol;      // stopping tolerance on the preconditioned residual
  double inv_h;    // 1/h(j+1,j) or 1/beta for the next basis vector
  double resid;    // current residual estimate |g(j+1)|
  int done;        // cycle finished (converged, breakdown, or iteration cap)
  int converged;
  int its;         // total Arnoldi steps
  int jcur;        // basis vectors completed in this cycle
  int maxit;
  int finished;    // the solve is over: converged, iteration cap, or a non-finite residual
  int diverged;    // a non-finite residual estimate (KSP_DIVERGED_NANORINF)
  int jlast;       // columns of the last cycle that ended (its Hessenberg solve)
};

// ---------------------------------------------------------------- ILU(0) factor
// TPR lanes (adjacent, so in one wavefront) share a row: each computes the
// row's multiplier for lower slot p and applies its share of the slot's
// update pairs (the targets of one slot are distinct), which shortens the
// sequential elimination chain of each level about TPR times.  Lanes of one
// wavefront issue their LDS operations in program order, so the writes of
// slot p are seen by the reads of slot p+1 without a barrier.
template <int BS>
struct FactorLanes {
  static constexpr int TPR = BS <= 256 ? 4 : 2;  // lanes per row (workgroup <= 1024 threads)
  static constexpr int NT = BS * TPR;
  static constexpr int WPL = (kPairW + TPR - 1) / TPR;  // pair words per lane
  static constexpr int AHEAD = 2;                        // lower slots whose update lists are in flight
  // waves per SIMD the registers are budgeted for (at most 64 VGPRs at 8):
  // two 16-wave workgroups per CU, as many as the LDS holds at BS <= 256
  static constexpr int WAVES = BS <= 256 ? 8 : 4;
};

template <class TV, int BS>
__global__ void __launch_bounds__(FactorLanes<BS>::NT) __attribute__((amdgpu_waves_per_eu(FactorLanes<BS>::WAVES)))
k_bj_factor(BJLayout L, const double *__restrict__ J,
                                                                   TV *__restrict__ fac, int *__restrict__ status) {
  constexpr int TPR = FactorLanes<BS>::TPR, NT = FactorLanes<BS>::NT, WPL = FactorLanes<BS>::WPL;
  __shared__ TV a[kW][BS];
  __shared__ uint16_t lc[kW][BS];
  __shared__ uint8_t dg[BS];
  __shared__ uint8_t ns[BS];
  const int b = blockIdx.x, tid = threadIdx.x;
  // index tables of the block's pattern (shared by identical blocks); values per block
  const int pb = L.d_pat ? L.d_pat[b] : b;
  const int64_t base = (int64_t)pb * BS;
  const int64_t ibase = (int64_t)pb * kW * BS;
  const int64_t sbase = (int64_t)b * kW * BS;
  {
    // all of the block's value loads in flight together
    constexpr int NE = (kW * BS + NT - 1) / NT;
    const double *Jb = J + (int64_t)b * L.ell_w * BS;
    int32_t pos[NE];
    uint16_t lcv[NE];
#pragma unroll
    for (int i = 0; i < NE; i++) {
      const int e = tid + i * NT;
      pos[i] = e < kW * BS ? L.d_jpos[ibase + e] : -1;  // relative to the block's first position
      lcv[i] = e < kW * BS ? L.d_lcol[ibase + e] : (uint16_t)0;
    }
#pragma unroll
    for (int i = 0; i < NE; i++) {
      const int e = tid + i * NT;
      if (e < kW * BS) {
        a[e / BS][e % BS] = (pos[i] >= 0) ? (TV)Jb[pos[i]] : (TV)0;
        lc[e / BS][e % BS] = lcv[i];
      }
    }
  }
  for (int r = tid; r < BS; r += NT) {
    ns[r] = L.d_nslot[base + r];
    dg[r] = L.d_diag[base + r];
  }
  const int t = tid / TPR, sub = tid % TPR;  // row of the block, lane within the row
  const int lev = L.d_levL[base + t];
  const int nlev = L.d_nlev[2 * pb];
  const int nsl = L.d_nslot[base + t];
  const int di0 = L.d_diag[base + t];
  // The update lists of lower slots p+1 .. p+AHEAD (count and this lane's
  // packed pair words) are in flight while slot p is eliminated, the first
  // ones fetched before the level loop, so global-memory latency is off the
  // elimination chain.  The slot loop is unrolled and the loads carry no
  // condition (the tables hold kLU slots per row; a lane's words past the
  // row's kPairW re-read the last one and are never used), so each list
  // lands in registers of its own and the wait before slot p is for slot p's
  // loads alone.  (A rotating queue in a rolled loop copies registers whose
  // loads are pending, which waits for every load at every slot.  At 1M
  // cells the FP32 factor went from 1579 to 1372 us; FP64, 1709 -> 1744 us,
  // is held by its LDS chain instead.)
  // A slot's table rows are addressed as a uniform base plus this lane's
  // 32-bit offset.  The base takes an opaque zero made where it is used, so
  // the compiler cannot hoist the thirteen slots' addresses out of the level
  // loop, where they would hold 64-bit addresses in vector registers.
  const uint32_t *__restrict__ pairs = L.d_pair + (int64_t)pb * kPair * BS;
  const uint8_t *__restrict__ npair = L.d_npair + (int64_t)pb * kLU * BS;
  uint32_t woff[WPL];
#pragma unroll
  for (int i = 0; i < WPL; i++) woff[i] = (uint32_t)(min(sub + TPR * i, kPairW - 1) * BS + t);
  auto fetch = [&](int p, int &cnt, uint32_t (&w)[WPL]) {
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    cnt = (npair + (p * BS + z))[(uint32_t)t];
    const uint32_t *__restrict__ pp = pairs + (p * kPairW * BS + z);
#pragma unroll
    for (int i = 0; i < WPL; i++) w[i] = pp[woff[i]];
  };
  constexpr int kAhead = FactorLanes<BS>::AHEAD;
  int nq[kLU];
  uint32_t qw[kLU][WPL];
#pragma unroll
  for (int d = 0; d < kAhead; d++) fetch(d, nq[d], qw[d]);
  __syncthreads();
  for (int l = 0; l < nlev; l++) {
    if (lev == l && nsl > 0) {
      // A slot's LDS reads go out as one batch: its own value, the pivot of
      // row k (finished at an earlier level), the pair sources in row k and
      // the pair targets in row t (distinct slots above p, so no read of the
      // batch waits for a write of the same slot), and the next slot's
      // column; the next column's diagonal slot is read behind the writes.
      // (the first lists pass through an empty asm here, so nothing derived
      // from them is hoisted out of the level loop into spilled registers)
#pragma unroll
      for (int d = 0; d < kAhead; d++) {
        asm volatile("" : "+v"(nq[d]));
#pragma unroll
        for (int i = 0; i < WPL; i++) asm volatile("" : "+v"(qw[d][i]));
      }
      int k = di0 > 0 ? lc[0][t] : t;
      int dk = dg[k];
#pragma unroll
      for (int p = 0; p < kLU; p++) {  // lower entries, increasing column
        if (p + kAhead < kLU) fetch(p + kAhead, nq[p + kAhead], qw[p + kAhead]);
        if (p < di0) {
          const int ncur = nq[p];
          const uint32_t *cur = qw[p];
          TV src[2 * WPL], tgt[2 * WPL];
#pragma unroll
          for (int i = 0; i < WPL; i++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
              const int q = 2 * (sub + TPR * i) + h;  // pair index of this lane's half-word
              const uint32_t w = (q < ncur) ? (h ? (cur[i] >> 16) : (cur[i] & 0xffffu)) : 0u;  // unused pairs read slot 0
              src[2 * i + h] = a[w >> 8][k];
              tgt[2 * i + h] = a[w & 0xff][t];
            }
          }
          const TV pv = a[dk][k];
          const TV ap = a[p][t];
          const int kn = (p + 1 < di0) ? lc[p + 1][t] : k;
          const TV lik = ap / pv;
          __builtin_amdgcn_wave_barrier();  // every lane has read a[p][t] before it is overwritten
          if (sub == 0) a[p][t] = lik;
#pragma unroll
          for (int i = 0; i < WPL; i++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
              const int q = 2 * (sub + TPR * i) + h;
              const uint32_t w = h ? (cur[i] >> 16) : (cur[i] & 0xffffu);
              if (q < ncur) a[w & 0xff][t] = tgt[2 * i + h] - lik * src[2 * i + h];
            }
          }
          const int dkn = dg[kn];
          __builtin_amdgcn_wave_barrier();
          k = kn;
          dk = dkn;
        }
      }
      if (sub == 0 && a[dg[t]][t] == (TV)0) atomicOr(status, 1);
    }
    __syncthreads();
  }
  // split layout: L multipliers, U entries, inverse diagonal
  const int di = dg[t];
  const int nlu = L.d_nlu ? L.d_nlu[pb] : kLU;
  for (int p = sub; p < nlu; p += TPR) {
    fac[sbase + (int64_t)p * BS + t] = (p < di) ? a[p][t] : (TV)0;
    const int q = di + 1 + p;
    fac[sbase + (int64_t)(kLU + p) * BS + t] = (q < nsl) ? a[q][t] : (TV)0;
  }
  if (sub == 0) fac[sbase + (int64_t)(2 * kLU) * BS + t] = nsl > 0 ? (TV)1 / a[di][t] : (TV)0;
}

// This thread's share of p[0..n) in a workgroup of NT threads: the sum of
// p[c] over c = t, t + NT, ... in increasing c, its loads issued eight at a
// time (a rolled loop would wait for each load before the next).
template <int NT, class TV>
__device__ inline TV strided_sum(const TV *__restrict__ p, int n) {
  TV s = 0;
  int c = threadIdx.x;
  for (; c + 7 * NT < n; c += 8 * NT) {
    TV v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = p[c + i * NT];
#pragma unroll
    for (int i = 0; i < 8; i++) s += v[i];
  }
  for (; c < n; c += NT) s += p[c];
  return s;
}

// Sum over an NT-thread workgroup; every thread gets the result.
template <int NT, class TV>
__device__ inline TV block_sum_n(TV v, TV *sh) {
  const int t = threadIdx.x;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((t & 63) == 0) sh[t >> 6] = v;
  __syncthreads();
  TV s = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; w++) s += sh[w];
  __syncthreads();
  return s;
}

// Restriction of the block's right-hand side to one value (the coarse
// correction's R1^T r), written by thread 0; all threads call it.
template <class TV, int BS>
__device__ inline void restrict_block(const TV *x, double *rc1, int b, int t) {
  __shared__ TV sh[BS / 64];
  const TV s = block_sum_n<BS>(x[t], sh);
  if (t == 0) rc1[b] = (double)s;
}

// The same restriction, and in the same pass the block sum of v (this
// thread's entry of the basis vector) into rv[b]; rc1[b] is the value
// restrict_block gives.
template <class TV, int BS>
__device__ inline void restrict_block2(const TV *x, double *rc1, TV v, double *rv, int b, int t) {
  __shared__ TV sh[2][BS / 64];
  TV a = x[t];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    v += __shfl_down(v, off, 64);
  }
  if ((t & 63) == 0) {
    sh[0][t >> 6] = a;
    sh[1][t >> 6] = v;
  }
  __syncthreads();
  if (t == 0) {
    TV s = 0, q = 0;
#pragma unroll
    for (int w = 0; w < BS / 64; w++) {
      s += sh[0][w];
      q += sh[1][w];
    }
    rc1[b] = (double)s;
    rv[b] = (double)q;
  }
  __syncthreads();
}

// ---------------------------------------------------------------- ILU(0) apply: y = (LU)^-1 r
// Forward and backward sweeps over the block's levels with x (BS+1 entries,
// x[BS] == 0 the target of padded slots) in LDS.  Every thread of the block
// must call this; x[t] holds the right-hand side on entry and the solution on
// return.
template <class TV, int BS>
__device__ inline void bj_sweeps(const BJLayout &L, const TV *__restrict__ fac, TV *x, int b, int t) {
  const int pb = L.d_pat ? L.d_pat[b] : b;  // index tables of the block's pattern
  const int64_t base = (int64_t)pb * BS;
  const int64_t sbase = (int64_t)b * kW * BS;
  const int64_t cbase = (int64_t)pb * kLU * BS;
  TV vl[kLU], vu[kLU];
  uint16_t cl[kLU], cu[kLU];
  const int nlu = L.d_nlu ? L.d_nlu[pb] : kLU;  // slots past it hold 0 against the zero entry x[BS]
#pragma unroll
  for (int p = 0; p < kLU; p++) {
    const bool in = p < nlu;
    vl[p] = in ? fac[sbase + (int64_t)p * BS + t] : (TV)0;
    vu[p] = in ? fac[sbase + (int64_t)(kLU + p) * BS + t] : (TV)0;
    cl[p] = in ? L.d_cL[cbase + (int64_t)p * BS + t] : (uint16_t)BS;
    cu[p] = in ? L.d_cU[cbase + (int64_t)p * BS + t] : (uint16_t)BS;
  }
  const TV dinv = fac[sbase + (int64_t)(2 * kLU) * BS + t];
  const int levL = L.d_levL[base + t], levU = L.d_levU[base + t];
  const int nL = L.d_nlev[2 * pb], nU = L.d_nlev[2 * pb + 1];
  __syncthreads();
  // forward: branch-free inner loop, all LDS reads of a level in flight together
  for (int l = 0; l < nL; l++) {
    if (levL == l) {
      TV s = x[t];
#pragma unroll
      for (int p = 0; p < kLU; p++) s -= vl[p] * x[cl[p]];
      x[t] = s;
    }
    __syncthreads();
  }
  for (int l = 0; l < nU; l++) {
    if (levU == l) {
      TV s = x[t];
#pragma unroll
      for (int p = 0; p < kLU; p++) s -= vu[p] * x[cu[p]];
      x[t] = s * dinv;
    }
    __syncthreads();
  }
}

template <class TV, int BS>
__global__ void __launch_bounds__(BS) k_bj_apply(BJLayout L, const TV *__restrict__ fac, const TV *__restrict__ r,
                                                 TV *__restrict__ y, const int *__restrict__ stop = nullptr,
                                                 double *__restrict__ rc1 = nullptr) {
  __shared__ TV x[BS + 1];
  if (stop && *stop) return;  // uniform over the workgroup
  const int b = blockIdx.x, t = threadIdx.x;
  const int row = L.d_row[(int64_t)b * BS + t];
  x[t] = row >= 0 ? r[row] : (TV)0;
  if (t == 0) x[BS] = (TV)0;
  if (rc1) restrict_block<TV, BS>(x, rc1, b, t);  // rc1 is uniform
  bj_sweeps<TV, BS>(L, fac, x, b, t);
  if (row >= 0) y[row] = x[t];
}

// ---------------------------------------------------------------- block-ordered SpMV
// The solver's Jacobian is stored block-contiguous over the block layout:
// slot k of position p = b*bs + t at a[(b*W + k)*bs + t], its global column at
// the same place of ecol (or of the block pattern's table, see ecol_addr);
// the global row of p is L.d_row[p] (-1 for padding positions).

// Column of slot k at block position p: an int32 global column, or an int16
// offset from the row (used when every offset of the mesh fits, which halves
// the index bytes of every SpMV).
template <class TI>
__device__ inline int blk_col(const TI *__restrict__ ecol, int64_t q, int row) {
  if (sizeof(TI) == 2) return row + (int)ecol[q];
  return (int)ecol[q];
}

// UNROLL: issue a hexahedral row's 27 loads at once (more registers, more
// memory-level parallelism); worth it where occupancy is set by other phases.
// Slot k's column index sits at ecol[eb + k*es]: eb = p, es = NB in the
// slot-major layout; eb = (pattern*W)*bs + t, es = bs with shared patterns.
template <class TV, class TI, bool UNROLL = (sizeof(TV) == 4)>
__device__ inline TV blk_row(int64_t ab, int64_t as, int W, int row, const TI *__restrict__ ecol, int64_t eb,
                             int64_t es, const TV *__restrict__ a, const TV *__restrict__ x) {
  TV s = 0;
  if (UNROLL && W == kW) {
    int c[kW];
    TV v[kW];
#pragma unroll
    for (int k = 0; k < kW; k++) {
      c[k] = blk_col(ecol, eb + k * es, row);
      v[k] = a[ab + k * as];
    }
#pragma unroll
    for (int k = 0; k < kW; k++) s += v[k] * x[c[k]];
  } else {
    for (int k = 0; k < W; k++) s += a[ab + k * as] * x[blk_col(ecol, eb + k * es, row)];
  }
  return s;
}

// The solver's ELL arrays (Jacobian values, column indices, coarse slot
// map) are block-contiguous: slot k of position p = b*bs + t (block b, row t
// of the block) at (b*W + k)*bs + t, so a block's W slots are one contiguous
// run of W*bs entries (one DRAM stream per workgroup instead of W streams
// NB apart).  Index tables shared by the blocks of one pattern sit at
// (pattern*W + k)*bs + t.
__device__ inline int64_t ell_base(int64_t p, int bs, int W) {
  const int64_t b = p / bs;
  return (b * W) * bs + (p - b * bs);
}

// Base and stride of block position p's column indices (see blk_row).
__device__ inline void ecol_addr(const int32_t *__restrict__ pat, int bs, int W, int64_t p, int64_t &eb, int64_t &es) {
  if (pat) {
    const int64_t b = p / bs;
    eb = ((int64_t)pat[b] * W) * bs + (p - b * bs);
  } else {
    eb = ell_base(p, bs, W);
  }
  es = bs;
}

// y[row(p)] = (A x)[row(p)]
template <class TV, class TI>
__global__ void __launch_bounds__(256) k_spmv_blk(int64_t NB, int W, const int32_t *__restrict__ rowof,
                                                  const TI *__restrict__ ecol, const int32_t *__restrict__ pat,
                                                  int bs, const int32_t *__restrict__ bw, const TV *__restrict__ a,
                                                  const TV *__restrict__ x, TV *__restrict__ y,
                                                  const int *__restrict__ stop = nullptr) {
  if (stop && *stop) return;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NB) return;
  const int row = rowof[p];
  if (row < 0) return;
  int64_t eb, es;
  ecol_addr(pat, bs, W, p, eb, es);
  y[row] = blk_row<TV, TI, true>(ell_base(p, bs, W), bs, bw ? bw[p / bs] : W, row, ecol, eb, es, a, x);
}

// Arnoldi column j: Hessenberg entries from h[0..j] (dots) and hj1 = h[j+1]
// (||w||^2), previous Givens rotations, the new rotation, residual estimate
// and stop test.  Rotation i turns (entry i, entry i+1) into (final entry i,
// the entry i+1 that rotation i+1 reads), so the running entry stays in a
// register and nothing written is read back from memory.
// hr, csr, snr: where h[0..j], cs[0..j-1] and sn[0..j-1] are read (the
// arrays themselves, or copies in LDS); cs[j] and sn[j] are written to cs, sn.
__device__ inline void arnoldi_update(GmresDev *gd, int j, int m, const double *hr, double hj1,
                                      double *__restrict__ H, double *__restrict__ cs, double *__restrict__ sn,
                                      double *__restrict__ g, const double *csr, const double *snr) {
  const int ld = m + 1;
  const double hn = sqrt(hj1);
  double a = hr[0];
  for (int i = 0; i < j; i++) {
    const double c = hr[i + 1], ci = csr[i], si = snr[i];
    H[i + j * ld] = ci * a + si * c;
    a = -si * a + ci * c;
  }
  const double c = hn;
  const double d = sqrt(a * a + c * c);
  const double csj = d > 0 ? a / d : 1.0, snj = d > 0 ? c / d : 0.0;
  cs[j] = csj;
  sn[j] = snj;
  H[j + j * ld] = d;
  H[j + 1 + j * ld] = 0.0;
  const double gj = g[j], gj1 = -snj * gj;
  g[j + 1] = gj1;
  g[j] = csj * gj;
  const int its = gd->its + 1;
  gd->its = its;
  gd->jcur = j + 1;
  const double resid = fabs(gj1);
  gd->resid = resid;
  gd->inv_h = hn > 0.0 ? 1.0 / hn : 0.0;
  const bool bad = !isfinite(resid) || !isfinite(hn);
  const int conv = !bad && (resid <= gd->tol || hn == 0.0);
  gd->converged = conv;
  if (bad) gd->diverged = 1;
  const bool fin = conv || bad || its >= gd->maxit;  // finished is 0 on entry (the step runs only while not done)
  if (fin) gd->finished = 1;
  if (fin || j + 1 == m) gd->done = 1;
}

// ---------------------------------------------------------------- fused Arnoldi matvec
// The previous Arnoldi column, finished inside the next matvec (NORM mode):
// ||u||^2 is the sum of the nchunk norm partials, reduced by every workgroup
// in the same fixed order; workgroup 0 also stores it and runs the Givens
// update of column j.
struct NormArgs {
  const void *part = nullptr;
  int nchunk = 0, j = 0, m = 0;
  double *h = nullptr, *H = nullptr, *cs = nullptr, *sn = nullptr, *g = nullptr;
};

// One GMRES step's  V_j = s u  and  out = M^-1 A V_j = s M^-1 A u  (s = 1/||u||
// read on the device, or with NORM formed here from the norm partials of the
// previous step): the block's rows of A u go straight into LDS and the
// triangular sweeps follow, so neither A u nor a separate scaling pass
// touches HBM; while some workgroups sweep their levels, others stream their
// rows of A.  u and out must be different vectors.
template <class TV, int BS, class TI, bool NORM>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) k_bj_spmv_apply(BJLayout L, int64_t NB, int W, const TI *__restrict__ ecol,
                                                      const TV *__restrict__ a, const TV *__restrict__ u,
                                                      GmresDev *__restrict__ gd, TV *__restrict__ vj,
                                                      const TV *__restrict__ fac, TV *__restrict__ out,
                                                      const int *__restrict__ stop, double *__restrict__ rc1,
                                                      NormArgs na, double *__restrict__ rv) {
  __shared__ TV x[BS + 1];
  __shared__ TV shn[BS / 64];
  __shared__ int stop_s;
  const int b = blockIdx.x, t = threadIdx.x;
  TV sc;
  if (NORM) {
    // workgroup 0 of this same launch may set the stop flag: read it once
    if (t == 0) stop_s = stop ? *stop : 0;
    __syncthreads();
    if (stop_s) return;  // uniform over the workgroup
    const TV *part = static_cast<const TV *>(na.part);
    TV s = 0;
    s = strided_sum<BS>(part, na.nchunk);
    s = block_sum_n<BS>(s, shn);
    const double hn = sqrt((double)s);
    sc = (TV)(hn > 0.0 ? 1.0 / hn : 0.0);
    if (b == 0 && t == 0) {
      na.h[na.j + 1] = (double)s;
      arnoldi_update(gd, na.j, na.m, na.h, (double)s, na.H, na.cs, na.sn, na.g, na.cs, na.sn);
    }
  } else {
    if (stop && *stop) return;  // uniform over the workgroup
    sc = (TV)gd->inv_h;
  }
  const int64_t p = (int64_t)b * BS + t;
  const int row = L.d_row[p];
  // the sweeps below hold ~100 registers anyway: the unrolled row costs no occupancy here
  int64_t eb, es;
  ecol_addr(L.d_pat, BS, W, p, eb, es);
  const int wb = L.d_bw ? L.d_bw[b] : W;  // slots past the block's longest row are padding
  x[t] = row >= 0 ? sc * blk_row<TV, TI, true>(((int64_t)b * W) * BS + t, BS, wb, row, ecol, eb, es, a, u) : (TV)0;
  const TV vv = row >= 0 ? sc * u[row] : (TV)0;
  if (row >= 0) vj[row] = vv;
  if (t == 0) x[BS] = (TV)0;
  // rc1 and rv are uniform; rv: the block sums of V_j, for the coarse terms of the multi-dot
  if (rc1 && rv) restrict_block2<TV, BS>(x, rc1, vv, rv, b, t);
  else if (rc1) restrict_block<TV, BS>(x, rc1, b, t);
  bj_sweeps<TV, BS>(L, fac, x, b, t);
  if (row >= 0) out[row] = x[t];
}

// A restart's preconditioned residual out = M_BJ^-1 (b - J y) in one pass:
// the block's rows of b - J y (FP64, the difference taken in FP64 and then
// rounded to TV, as a separate product and difference would give) go into
// LDS, restricted for the coarse level when rc1 is set, and the sweeps follow.
template <class TV, int BS, class TI>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) k_bj_resid_apply(BJLayout L, int W, const TI *__restrict__ ecol,
                                                       const double *__restrict__ J, const double *__restrict__ y,
                                                       const double *__restrict__ bvec, const TV *__restrict__ fac,
                                                       TV *__restrict__ out, const int *__restrict__ stop,
                                                       double *__restrict__ rc1) {
  __shared__ TV x[BS + 1];
  if (stop && *stop) return;  // uniform over the workgroup
  const int b = blockIdx.x, t = threadIdx.x;
  const int64_t p = (int64_t)b * BS + t;
  const int row = L.d_row[p];
  int64_t eb, es;
  ecol_addr(L.d_pat, BS, W, p, eb, es);
  const int wb = L.d_bw ? L.d_bw[b] : W;
  x[t] = row >= 0 ? (TV)(bvec[row] - blk_row<double, TI, true>(((int64_t)b * W) * BS + t, BS, wb, row, ecol, eb, es, J, y))
                  : (TV)0;
  if (t == 0) x[BS] = (TV)0;
  if (rc1) restrict_block<TV, BS>(x, rc1, b, t);  // rc1 is uniform
  bj_sweeps<TV, BS>(L, fac, x, b, t);
  if (row >= 0) out[row] = x[t];
}

// ---------------------------------------------------------------- BLAS-1 with deterministic reductions
constexpr int kRedBlock = 256;
constexpr int kRedPerThread = 2;
constexpr int kChunk = kRedBlock * kRedPerThread;
// Rows per thread of the multi-vector kernels.  (Four FP32 rows per thread,
// to match FP64's bytes in flight, halves the workgroups and measured slower:
// multi-dot 27 -> 30 us, update 14.5 -> 21 us at 1M rows.)
template <class TV>
struct Red {
  static constexpr int E = kRedPerThread;
  static constexpr int chunk = kRedBlock * E;
};

template <class TV>
__device__ inline TV block_sum(TV v, TV *sh) {
  const int t = threadIdx.x;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((t & 63) == 0) sh[t >> 6] = v;
  __syncthreads();
  TV s = 0;
  if (t == 0)
    for (int w = 0; w < kRedBlock / 64; w++) s += sh[w];
  __syncthreads();
  return s;
}

// Sum of 32 per-lane values over a wavefront by recursive halving: at each
// step a lane keeps one half of its values, adds the partner's copy of that
// half and hands over the other half (16 + 8 + 4 + 2 + 1 exchanges), then one
// more exchange completes the pair.  Afterwards lanes 2j and 2j+1 both hold
// the wavefront sum of value j.
template <int HALF, class TV>
__device__ inline void halve_step(TV (&v)[32], int lane) {
  const bool upper = (lane & (2 * HALF)) != 0;
#pragma unroll
  for (int k = 0; k < HALF; k++) {
    const TV send = upper ? v[k] : v[HALF + k];
    const TV keep = upper ? v[HALF + k] : v[k];
    v[k] = keep + __shfl_xor(send, 2 * HALF, 64);
  }
}

// NS (8, 16 or 32) values: the halving steps use lane bits log2(NS)..1, so
// value j ends in lanes with (lane >> 1) % NS == j; the lane bits left over
// (bit 0, and bits 5 / 4 for fewer values) are summed by plain exchanges.
template <int NS, class TV>
__device__ inline TV wave_reduce_scatter(TV (&v)[32]) {
  const int lane = threadIdx.x & 63;
  if (NS >= 32) halve_step<16>(v, lane);
  if (NS >= 16) halve_step<8>(v, lane);
  halve_step<4>(v, lane);
  halve_step<2>(v, lane);
  halve_step<1>(v, lane);
  TV r = v[0] + __shfl_xor(v[0], 1, 64);
  if (NS <= 16) r += __shfl_xor(r, 32, 64);
  if (NS <= 8) r += __shfl_xor(r, 16, 64);
  return r;
}

// The coarse correction's level-1/2 solve run by the first ng workgroups of
// a multi-dot launch (ng = 0: none), with the group's terms of the dots.
struct CoarseDot {
  int ng = 0, gsz = 0;
  const int32_t *gmem = nullptr;                // ng*gsz blocks of each group (-1 pad)
  const double *G1 = nullptr, *A2 = nullptr;    // group inverses, inverted level-2 operator
  const double *rc1 = nullptr;                  // restricted vector (one value per block)
  const double *RV = nullptr;                   // block sums of the basis vectors, rv_ld per vector
  int64_t rv_ld = 0;
  double *z1 = nullptr;                         // coarse correction per block (out)
};
template <class TV>
__device__ void coarse_level1_dots(const CoarseDot &cd, int g, int nv, TV *partial, int pstride);

// Coarse prolongation applied on the fly: w[i] += z1[pos[i] >> bs_log2]
// (pos = block-layout position of row i).  z1 == nullptr: none.
struct Prolong {
  const double *z1 = nullptr;
  const int32_t *pos = nullptr;
  int bs_log2 = 0;
};

// Two adjacent rows per lane, read as one 16-byte (FP64) or 8-byte (FP32)
// access when the rows' addresses allow it (ldv and the row even); the last
// odd row of a vector is read alone.
template <class TV> struct Pair2;
template <> struct Pair2<double> { using T = double2; };
template <> struct Pair2<float> { using T = float2; };
template <class TV>
__device__ inline void load2(const TV *__restrict__ p, int64_t r, int64_t n, bool vec, TV &a, TV &b) {
  if (vec && r + 1 < n) {
    const typename Pair2<TV>::T q = *reinterpret_cast<const typename Pair2<TV>::T *>(p + r);
    a = q.x;
    b = q.y;
  } else {
    a = r < n ? p[r] : (TV)0;
    b = r + 1 < n ? p[r + 1] : (TV)0;
  }
}

// The row pairs (r, r + 1) of vectors V_j0 .. V_j0+cnt-1 (cnt <= NS, the
// same for the whole workgroup) passed to f(j, a, b) in order.  The loads
// carry only workgroup-uniform guards, so all of them are issued before the
// first product instead of one at a time.  The chunk must hold both rows and
// the pair be aligned (ldv and r even).
template <class TV, int NS, class F>
__device__ inline void load_pairs(const TV *__restrict__ V, int64_t ldv, int64_t r, int j0, int cnt, F &&f) {
  using P = typename Pair2<TV>::T;
  const TV *p = V + (int64_t)j0 * ldv + r;
  P q[NS];
#pragma unroll
  for (int j = 0; j < NS; j++)
    if (j < cnt) q[j] = *reinterpret_cast<const P *>(p + (int64_t)j * ldv);
#pragma unroll
  for (int j = 0; j < NS; j++)
    if (j < cnt) f(j, q[j].x, q[j].y);
}

// partial[j*nchunk + chunk] = <w, V_j> over the chunk, j = 0..nv-1.  Each
// thread holds two adjacent rows and accumulates up to NS dots in registers;
// one reduce-scatter per wavefront and one barrier per group of NS vectors
// finish the chunk.  With pro.z1 set, w first receives the coarse
// prolongation (written back).
template <class TV, int NS>
__global__ void __launch_bounds__(kRedBlock) k_mdot(int64_t n, int nv, const TV *__restrict__ V, int64_t ldv,
                                                    TV *__restrict__ w, TV *__restrict__ partial, int nchunk,
                                                    Prolong pro, const int *__restrict__ stop = nullptr,
                                                    CoarseDot cd = CoarseDot{}) {
  static_assert(Red<TV>::E == 2, "two rows per thread");
  __shared__ TV red[kRedBlock / 64][NS];
  if (stop && *stop) return;
  // partial[j*nchunk + c]: c < cd.ng the coarse workgroups' terms, then one per row chunk
  if ((int)blockIdx.x < cd.ng) {
    coarse_level1_dots<TV>(cd, blockIdx.x, nv, partial, nchunk);
    return;
  }
  const int chunk = blockIdx.x - cd.ng;
  const int t = threadIdx.x, lane = t & 63, wv_id = t >> 6;
  const int64_t r0 = (int64_t)chunk * Red<TV>::chunk + 2 * t;
  const bool vec = (ldv & 1) == 0;
  TV w0, w1;
  load2(w, r0, n, true, w0, w1);
  if (pro.z1) {
    if (r0 < n) { w0 += (TV)pro.z1[pro.pos[r0] >> pro.bs_log2]; w[r0] = w0; }
    if (r0 + 1 < n) { w1 += (TV)pro.z1[pro.pos[r0 + 1] >> pro.bs_log2]; w[r0 + 1] = w1; }
  }
  // FP32: every chunk but a ragged last one issues its NS loads together
  // (8-byte accesses: 22.4 -> 19.4 us at 1M rows).  FP64 keeps one load at
  // a time, which measured faster there (30.8 against 32.7 us): its 16-byte
  // loads already keep enough bytes in flight at the higher occupancy.
  const bool full = sizeof(TV) == 4 && vec && (int64_t)(chunk + 1) * Red<TV>::chunk <= n;
  for (int j0 = 0; j0 < nv; j0 += NS) {
    TV acc[32];
#pragma unroll
    for (int j = 0; j < 32; j++) acc[j] = 0;
    if (full) {
      load_pairs<TV, NS>(V, ldv, r0, j0, min(NS, nv - j0), [&](int j, TV v0, TV v1) {
        acc[j] += w0 * v0;
        acc[j] += w1 * v1;
      });
    } else {
#pragma unroll
      for (int j = 0; j < NS; j++) {
        if (j0 + j < nv) {
          TV v0, v1;
          load2(V + (int64_t)(j0 + j) * ldv, r0, n, vec, v0, v1);
          acc[j] += w0 * v0;
          acc[j] += w1 * v1;
        }
      }
    }
    const TV r = wave_reduce_scatter<NS>(acc);
    if ((lane & 1) == 0 && (lane >> 1) < NS) red[wv_id][lane >> 1] = r;
    __syncthreads();
    if (t < NS && j0 + t < nv) {
      TV s = 0;
#pragma unroll
      for (int q = 0; q < kRedBlock / 64; q++) s += red[q][t];
      partial[(int64_t)(j0 + t) * nchunk + cd.ng + chunk] = s;
    }
    __syncthreads();
  }
}

// out[j] = sum_c partial[j*nchunk + c] (fixed order)
template <class TV>
__global__ void k_reduce_partials(int nv, int nchunk, const TV *__restrict__ partial, double *__restrict__ out,
                                  const int *__restrict__ stop = nullptr) {
  __shared__ TV sh[kRedBlock / 64];
  if (stop && *stop) return;
  const int j = blockIdx.x;
  TV s = 0;
  s = strided_sum<kRedBlock>(partial + (int64_t)j * nchunk, nchunk);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[j] = (double)s;
}

// w -= sum_j h[j] V_j ; partial[chunk] = ||w||^2 over the chunk (two
// adjacent rows per thread, as k_mdot)
template <class TV>
__global__ void __launch_bounds__(kRedBlock) k_maxpy_norm(int64_t n, int nv, const TV *__restrict__ V, int64_t ldv,
                                                          const double *__restrict__ h, TV *__restrict__ w,
                                                          TV *__restrict__ partial,
                                                          const int *__restrict__ stop = nullptr,
                                                          Prolong pro = Prolong{}) {
  __shared__ TV sh[kRedBlock / 64];
  if (stop && *stop) return;
  const int64_t r0 = (int64_t)blockIdx.x * Red<TV>::chunk + 2 * threadIdx.x;
  const bool vec = (ldv & 1) == 0;
  TV x0, x1;
  load2(w, r0, n, true, x0, x1);
  if (pro.z1) {  // the coarse prolongation, when the multi-dot left it out
    if (r0 < n) x0 += (TV)pro.z1[pro.pos[r0] >> pro.bs_log2];
    if (r0 + 1 < n) x1 += (TV)pro.z1[pro.pos[r0 + 1] >> pro.bs_log2];
  }
  for (int j = 0; j < nv; j++) {
    TV v0, v1;
    load2(V + (int64_t)j * ldv, r0, n, vec, v0, v1);
    const TV hj = (TV)h[j];
    x0 -= hj * v0;
    x1 -= hj * v1;
  }
  TV s = 0;
  if (r0 + 1 < n) {
    *reinterpret_cast<typename Pair2<TV>::T *>(w + r0) = typename Pair2<TV>::T{x0, x1};
    s += x0 * x0;
    s += x1 * x1;
  } else if (r0 < n) {
    w[r0] = x0;
    s += x0 * x0;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// partial[chunk] = <x, y>
template <class TX, class TY>
__global__ void __launch_bounds__(kRedBlock) k_dot(int64_t n, const TX *__restrict__ x, const TY *__restrict__ y,
                                                   double *__restrict__ partial, const int *__restrict__ stop = nullptr) {
  __shared__ double sh[kRedBlock / 64];
  if (stop && *stop) return;
  const int64_t i0 = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  double s = 0;
#pragma unroll
  for (int e = 0; e < kRedPerThread; e++) {
    const int64_t i = i0 + (int64_t)e * kRedBlock;
    if (i < n) s += (double)x[i] * (double)y[i];
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <class TV>
__global__ void k_scale_copy(int64_t n, TV a, const TV *__restrict__ x, TV *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

template <class TA, class TB>
__global__ void k_convert(int64_t n, const TA *__restrict__ a, TB *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (TB)a[i];
}

// out = a - b
template <class TV>
__global__ void k_sub(int64_t n, const TV *__restrict__ a, const TV *__restrict__ b, TV *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] - b[i];
}



template <class TV>
__global__ void __launch_bounds__(256) k_spmv_ell_stop(int n, int W, const int32_t *__restrict__ ecol,
                                                       const TV *__restrict__ a, const TV *__restrict__ x,
                                                       TV *__restrict__ y, const GmresDev *__restrict__ gd) {
  if (gd->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = ell_row<TV>(n, W, i, ecol, a, x);
}

// y = s * x with s read on the device (1/beta or 1/h(j+1,j))
template <class TV>
__global__ void k_scale_dev(int64_t n, const TV *__restrict__ x, TV *__restrict__ y, const GmresDev *__restrict__ gd) {
  if (gd->done) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (TV)gd->inv_h * x[i];
}

__global__ void k_gmres_init(GmresDev *gd, int *status, int maxit) {
  gd->its = 0;
  gd->done = 0;
  gd->converged = 0;
  gd->finished = 0;
  gd->diverged = 0;
  gd->jcur = 0;
  gd->maxit = maxit;
  *status = 0;
}

// Start of a cycle: beta = sqrt(h[0]) from a norm reduction; sets tol on the first cycle.
// After the solve has finished, a later cycle is empty: no columns, stop flag set.
__global__ void k_cycle_start(GmresDev *gd, const double *__restrict__ h, double *__restrict__ g, int m, double rtol,
                              double atol, int first) {
  if (!first && gd->finished) {
    gd->jcur = 0;
    gd->done = 1;
    return;
  }
  const double beta = sqrt(h[0]);
  if (first) gd->tol = fmax(rtol * beta, atol);
  for (int i = 0; i <= m; i++) g[i] = 0.0;
  g[0] = beta;
  gd->resid = beta;
  gd->jcur = 0;
  gd->inv_h = beta > 0.0 ? 1.0 / beta : 0.0;
  const bool bad = !isfinite(beta);
  const int conv = !bad && !(beta > gd->tol);
  gd->converged = conv;
  if (bad) gd->diverged = 1;
  gd->finished = conv || bad || gd->its >= gd->maxit;
  gd->done = gd->finished;
}

// h[j+1] = sum of the nchunk norm partials (fixed order, as k_reduce_partials)
// and the Arnoldi/Givens update of column j, in one workgroup.
template <class TV>
__global__ void __launch_bounds__(kRedBlock) k_norm_arnoldi(int nchunk, const TV *__restrict__ partial, GmresDev *gd,
                                                            int j, int m, double *__restrict__ h,
                                                            double *__restrict__ H, double *__restrict__ cs,
                                                            double *__restrict__ sn, double *__restrict__ g) {
  __shared__ TV sh[kRedBlock / 64];
  __shared__ double hs[kRedBlock], css[kRedBlock], sns[kRedBlock];  // h, cs, sn read by the update
  if (gd->done) return;
  const int t = threadIdx.x;
  const bool staged = j < kRedBlock;  // the columns fit: the update reads them from LDS
  if (staged) {
    if (t <= j) hs[t] = h[t];
    if (t < j) { css[t] = cs[t]; sns[t] = sn[t]; }
  }
  TV s = 0;
  s = strided_sum<kRedBlock>(partial, nchunk);
  s = block_sum(s, sh);  // its barriers also order the LDS copies above
  if (t == 0) {
    h[j + 1] = (double)s;
    if (staged) arnoldi_update(gd, j, m, hs, (double)s, H, cs, sn, g, css, sns);
    else arnoldi_update(gd, j, m, h, (double)s, H, cs, sn, g, cs, sn);
  }
}

// y = H(1:j,1:j)^-1 g(1:j) for the j = jcur columns of the cycle
__global__ void k_hess_solve(GmresDev *gd, int m, const double *__restrict__ H, const double *__restrict__ g,
                             double *__restrict__ y) {
  const int ld = m + 1, j = gd->jcur;
  gd->jlast = j;
  for (int i = j - 1; i >= 0; i--) {
    double s = g[i];
    for (int k = i + 1; k < j; k++) s -= H[i + k * ld] * y[k];
    y[i] = s / H[i + i * ld];
  }
}

// The same with the cycle's Hessenberg columns staged in LDS by the
// workgroup and y formed in LDS, so the one thread's back substitution waits
// on no global load or store (same operations, same order).  Dynamic LDS:
// ((m + 1) * m + m) doubles.
constexpr int kHessLdsMaxM = 64;
__global__ void __launch_bounds__(256) k_hess_solve_lds(GmresDev *gd, int m, const double *__restrict__ H,
                                                        const double *__restrict__ g, double *__restrict__ y) {
  extern __shared__ double hs[];
  const int ld = m + 1, j = gd->jcur;
  for (int e = threadIdx.x; e < ld * j; e += blockDim.x) hs[e] = H[e];
  __syncthreads();
  if (threadIdx.x != 0) return;
  gd->jlast = j;
  double *ys = hs + (size_t)ld * m;
  for (int i = j - 1; i >= 0; i--) {
    double s = g[i];
    for (int k = i + 1; k < j; k++) s -= hs[i + k * ld] * ys[k];
    ys[i] = s / hs[i + i * ld];
  }
  for (int i = 0; i < j; i++) y[i] = ys[i];
}

// x += sum_{k<jcur} y[k] V_k
template <class TV, class TX>
__global__ void k_gemv_update_dev(int64_t n, int mmax, const TV *__restrict__ V, int64_t ldv,
                                  const double *__restrict__ c, const GmresDev *__restrict__ gd, TX *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int nv = gd->jcur < mmax ? gd->jcur : mmax;
  TV s = 0;
  for (int j = 0; j < nv; j++) s += (TV)c[j] * V[j * ldv + i];
  x[i] += (TX)s;
}

// w = x - lambda*y
__global__ void k_waxpy(int64_t n, double lambda, const double *__restrict__ x, const double *__restrict__ y,
                        double *__restrict__ w, const int *__restrict__ stop = nullptr) {
  if (stop && *stop) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = x[i] - lambda * y[i];
}

}  // namespace tdyhip
