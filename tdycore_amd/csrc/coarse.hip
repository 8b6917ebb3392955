// Coarse correction for the block-Jacobi ILU(0) preconditioner (optional;
// PETSc's default that the reference runs with is one-level block Jacobi).
//
//   M^-1 r = BJ^-1 r + R1 [ G1^-1 + R2 A2^-1 R2^T ] R1^T r
//
// Level 1 has one unknown per ILU block (R1 sums a block's rows), with the
// Galerkin operator A1 = R1^T J R1; its blocks are aggregated into groups of
// at most kGMax, each group's diagonal block of A1 is inverted exactly (G1),
// and level 2 (one unknown per group, A2 = R2^T A1 R2, at most kNG of them) is
// inverted exactly as a dense matrix.  Everything is additive, linear and
// fixed for one Newton step, so GMRES sees a constant preconditioner.
// Restriction to level 1 is fused into the ILU sweep kernels (one block sum
// per workgroup); the level-1/2 work is a few thousand values; prolongation
// adds one value per block to the preconditioned vector.
//
// All reductions run in a fixed order, so the correction is deterministic.
#include "tdyhip_internal.h"

namespace tdyhip {

constexpr int kCQ = 32;     // neighbour blocks of a block (itself included)
constexpr int kGMax = 64;   // blocks per level-1 group
constexpr int kNG = 128;    // level-2 unknowns (dense inverse in LDS)

struct CoarseLayout {
  int nb = 0, ng = 0, gsz = 0;    // blocks, groups, max blocks per group
  int32_t *nbr = nullptr;         // nb*kCQ neighbour block ids, ascending (-1 pad)
  uint8_t *cq = nullptr;          // W*NB, block-contiguous as the ELL: position in nbr[block(p)] of slot k's column block (255 none)
  int32_t *grp = nullptr;         // nb: group of each block
  int32_t *gmem = nullptr;        // ng*gsz: blocks of each group (-1 pad)
  double *A1 = nullptr;           // nb*kCQ: Galerkin rows (column = neighbour position)
  double *G1 = nullptr;           // ng*gsz*gsz: inverse of each group's diagonal block (column-major)
  double *A2 = nullptr;           // ng*ng: level-2 operator, inverted in place
  double *rc1 = nullptr;          // nb: restricted vector
  double *z1 = nullptr;           // nb: coarse correction per block (levels 1 and 2)
  int *status = nullptr;          // nonzero: singular coarse block
};

// A1[b][q] = sum over the block's rows and slots whose column lies in block
// nbr[b][q] of J (one workgroup per block, one thread per row).  Each thread
// forms its 32 per-neighbour sums, one register reduce-scatter per wavefront
// gives the wavefront's 32 sums (combining lanes in the same xor pairs as a
// shuffle-down tree), and thread q adds the wavefronts' sums in order.
template <int BS>
__global__ void __launch_bounds__(BS) k_coarse_rows(CoarseLayout C, int64_t NB, int W, const double *__restrict__ J) {
  static_assert(kCQ == 32, "one reduce-scatter value per neighbour");
  constexpr int NW = BS / 64;
  __shared__ double wsum[NW][kCQ];
  const int b = blockIdx.x, t = threadIdx.x;
  const int64_t e0 = ((int64_t)b * W) * BS + t;  // block-contiguous ELL: slot k at e0 + k*BS
  uint8_t q_of[32];
  double v_of[32];
#pragma unroll
  for (int k = 0; k < 32; k++) {
    q_of[k] = 255;
    v_of[k] = 0.0;
    if (k < W) {
      q_of[k] = C.cq[e0 + k * BS];
      v_of[k] = q_of[k] != 255 ? J[e0 + k * BS] : 0.0;
    }
  }
  int nq = 0;  // neighbour blocks (uniform over the workgroup)
  while (nq < kCQ && C.nbr[b * kCQ + nq] >= 0) nq++;
  double v[kCQ];
#pragma unroll
  for (int q = 0; q < kCQ; q++) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 32; k++) s += (q_of[k] == q) ? v_of[k] : 0.0;
    v[q] = s;
  }
  const double r = wave_reduce_scatter<kCQ>(v);  // lanes 2q and 2q+1 hold neighbour q's sum
  const int lane = t & 63;
  if ((lane & 1) == 0) wsum[t >> 6][lane >> 1] = r;
  __syncthreads();
  if (t < nq) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; w++) s += wsum[w][t];
    C.A1[b * kCQ + t] = s;
  }
}

// Gauss-Jordan inverse with row pivoting of an n x n matrix (n <= N), held
// as 4x4 register tiles, one per thread of an NT = (N/4)^2-thread
// workgroup: thread t holds rows 4*(t/(N/4)).. and columns 4*(t%(N/4))..
// (zero outside n x n).  The pivot of column k is the first row i >= k of
// largest |M[i][k]|; row k is swapped with it and scaled,
//   M[k][j] = (j == k ? 1 : M[k][j]) / M[k][k] (as inv or M[k][j] * inv),
// and every other row becomes M[i][j] = (j == k ? 0 : M[i][j]) - M[i][k] M[k][j];
// each step publishes the pivot row, the row it replaces and the pivot
// column through LDS (s).  On true, column c < n of the tiles is column
// s.dst[c] of the inverse.  false (for every thread): a zero or NaN pivot.
template <int N>
struct TiledGJ {
  static constexpr int TB = 4, NTC = N / TB, NT = NTC * NTC;
  static_assert(NTC <= 32, "the pivot search reduces over one wavefront");
  double rowp[N], rowk[N], colk[N];
  double cv[NTC];
  int ci[NTC];
  int piv[N], src[N], dst[N];
  int flag, pk;
  double dkk;
};

template <int N>
__device__ inline bool gj_invert_tiles(double (&m)[4][4], int n, TiledGJ<N> &s) {
  constexpr int TB = TiledGJ<N>::TB, NTC = TiledGJ<N>::NTC;
  const int t = threadIdx.x;
  const int rb = t / NTC, cb = t % NTC, r0 = rb * TB, c0 = cb * TB;
  if (t == 0) s.flag = 0;
  for (int k = 0; k < n; k++) {
    const int kb = k / TB, ka = k % TB;
    // pivot of column k: the first row i >= k of largest |M[i][k]|
    if (cb == kb) {
      double mx = -1.0;
      int p = n;
#pragma unroll
      for (int a = 0; a < TB; a++) {
        const int i = r0 + a;
        const double v = fabs(m[a][ka]);
        if (i >= k && i < n && v > mx) { mx = v; p = i; }
      }
      s.cv[rb] = mx;
      s.ci[rb] = p;
      if (rb == kb) s.dkk = m[ka][ka];
    }
    __syncthreads();
    if (t < 64) {
      double mx = t < NTC ? s.cv[t] : -1.0;
      int p = t < NTC ? s.ci[t] : n;
#pragma unroll
      for (int off = 16; off > 0; off >>= 1) {
        const double om = __shfl_xor(mx, off, 64);
        const int op = __shfl_xor(p, off, 64);
        if (om > mx || (om == mx && op < p)) { mx = om; p = op; }
      }
      if (t == 0) {
        s.piv[k] = p;
        s.pk = p;
        if (!(mx > 0.0) || s.dkk != s.dkk) s.flag = 1;
      }
    }
    __syncthreads();
    if (s.flag) return false;  // uniform over the workgroup
    const int p = s.pk, pb = p / TB, pa = p % TB;
    // publish row p (the new row k) and row k (the new row p)
    if (rb == pb)
#pragma unroll
      for (int b = 0; b < TB; b++) s.rowp[c0 + b] = m[pa][b];
    if (rb == kb)
#pragma unroll
      for (int b = 0; b < TB; b++) s.rowk[c0 + b] = m[ka][b];
    __syncthreads();
    const double inv = 1.0 / s.rowp[k];
    double sr[TB];  // row k after the swap and scaling: inv at column k
#pragma unroll
    for (int b = 0; b < TB; b++) sr[b] = (c0 + b == k) ? inv : s.rowp[c0 + b] * inv;
    if (p != k && rb == pb)
#pragma unroll
      for (int b = 0; b < TB; b++) m[pa][b] = s.rowk[c0 + b];
    if (rb == kb)
#pragma unroll
      for (int b = 0; b < TB; b++) m[ka][b] = sr[b];
    // column k after the swap (0 at row k)
    if (cb == kb)
#pragma unroll
      for (int a = 0; a < TB; a++) s.colk[r0 + a] = (r0 + a == k) ? 0.0 : m[a][ka];
    __syncthreads();
#pragma unroll
    for (int a = 0; a < TB; a++) {
      const int i = r0 + a;
      if (i == k) continue;
      const double c = s.colk[i];
#pragma unroll
      for (int b = 0; b < TB; b++) m[a][b] = ((c0 + b == k) ? 0.0 : m[a][b]) - c * sr[b];
    }
  }
  // undo the row interchanges as column interchanges, last first: the
  // result's column j is column src[j] of the tiles
  __syncthreads();
  if (t == 0) {
    for (int j = 0; j < n; j++) s.src[j] = j;
    for (int k = n - 1; k >= 0; k--) {
      const int q = s.src[k];
      s.src[k] = s.src[s.piv[k]];
      s.src[s.piv[k]] = q;
    }
  }
  __syncthreads();
  if (t < n) s.dst[s.src[t]] = t;
  __syncthreads();
  return true;
}

// One workgroup per group (TiledGJ<kGMax>::NT threads): invert the group's
// diagonal block of A1 into G1 and form the group's row of A2.  The
// members' neighbour lists (block, group, A1 value) are staged in LDS first,
// all their loads in flight together; the block is assembled in LDS and
// inverted in registers.
constexpr int kGroupThreads = TiledGJ<kGMax>::NT;
__global__ void __launch_bounds__(kGroupThreads) k_coarse_groups(CoarseLayout C) {
  constexpr int LD = kGMax + 1;  // odd row pitch: column reads do not hit one LDS bank
  constexpr int NQ = kGMax * kCQ;
  constexpr int PER = NQ / kGroupThreads;  // neighbour entries per thread
  constexpr int TB = TiledGJ<kGMax>::TB, NTC = TiledGJ<kGMax>::NTC;
  __shared__ double M[kGMax * LD];
  __shared__ TiledGJ<kGMax> gs;
  __shared__ int mem[kGMax];
  __shared__ int gq[NQ];     // group of each member's neighbour q (-1 none)
  __shared__ int nq[NQ];     // the neighbour block
  __shared__ double aq[NQ];  // its A1 entry
  const int g = blockIdx.x, t = threadIdx.x, G = C.gsz;
  if (t < G) mem[t] = C.gmem[g * G + t];
  __syncthreads();
  int n = 0;
  while (n < G && mem[n] >= 0) n++;
  {
    int nb[PER];
    double av[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int e = t + r * kGroupThreads, i = e / kCQ, q = e % kCQ;
      nb[r] = i < n ? C.nbr[mem[i] * kCQ + q] : -1;
      av[r] = i < n ? C.A1[mem[i] * kCQ + q] : 0.0;
    }
    int gv[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) gv[r] = nb[r] >= 0 ? C.grp[nb[r]] : -1;
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int e = t + r * kGroupThreads;
      nq[e] = nb[r];
      gq[e] = gv[r];
      aq[e] = nb[r] >= 0 ? av[r] : 0.0;
    }
  }
  for (int e = t; e < n * LD; e += kGroupThreads) M[e] = 0.0;
  __syncthreads();
  // row i of the diagonal block: A1 of member i at the columns of member j
  // (a neighbour list holds each block once)
  for (int e = t; e < n * kCQ; e += kGroupThreads) {
    if (gq[e] != g) continue;
    const int i = e / kCQ, bj = nq[e];
    int j = 0;
    while (j < n && mem[j] != bj) j++;
    if (j < n) M[i * LD + j] = aq[e];
  }
  __syncthreads();
  const int r0 = (t / NTC) * TB, c0 = (t % NTC) * TB;
  double m[TB][TB];
#pragma unroll
  for (int a = 0; a < TB; a++)
#pragma unroll
    for (int b = 0; b < TB; b++) m[a][b] = (r0 + a < n && c0 + b < n) ? M[(r0 + a) * LD + c0 + b] : 0.0;
  if (!gj_invert_tiles<kGMax>(m, n, gs)) {
    if (t == 0) atomicOr(C.status, 1);
    return;
  }
  // G1 column-major per group (entry (i, j) at j*G + i, zero outside n x n):
  // level 1 reads it coalesced
  double *g1 = C.G1 + (int64_t)g * G * G;
#pragma unroll
  for (int a = 0; a < TB; a++)
#pragma unroll
    for (int b = 0; b < TB; b++) {
      const int i = r0 + a, c = c0 + b;
      if (i >= G || c >= G) continue;
      if (c < n) g1[gs.dst[c] * G + i] = i < n ? m[a][b] : 0.0;
      else g1[c * G + i] = 0.0;
    }
  // A2[g][h] = sum over members and their neighbour entries in group h, in
  // the fixed order of the list
  for (int h = t; h < C.ng; h += kGroupThreads) {
    double sum = 0.0;
    for (int e = 0; e < n * kCQ; e++) sum += (gq[e] == h) ? aq[e] : 0.0;
    C.A2[(int64_t)g * C.ng + h] = sum;
  }
}

// Single workgroup of 1024 threads: A2 <- A2^-1 (ng <= kNG), the matrix in
// registers (gj_invert_tiles).
__global__ void __launch_bounds__(TiledGJ<kNG>::NT) k_coarse_invert(CoarseLayout C) {
  constexpr int TB = TiledGJ<kNG>::TB, NTC = TiledGJ<kNG>::NTC;
  static_assert(TiledGJ<kNG>::NT == 1024, "one 4x4 tile per thread of 1024");
  __shared__ TiledGJ<kNG> gs;
  const int n = C.ng, t = threadIdx.x;
  const int r0 = (t / NTC) * TB, c0 = (t % NTC) * TB;
  double m[TB][TB];
#pragma unroll
  for (int a = 0; a < TB; a++)
#pragma unroll
    for (int b = 0; b < TB; b++)
      m[a][b] = (r0 + a < n && c0 + b < n) ? C.A2[(int64_t)(r0 + a) * n + c0 + b] : 0.0;
  if (!gj_invert_tiles<kNG>(m, n, gs)) {
    if (t == 0) atomicOr(C.status, 1);
    return;
  }
#pragma unroll
  for (int a = 0; a < TB; a++)
#pragma unroll
    for (int b = 0; b < TB; b++)
      if (r0 + a < n && c0 + b < n) C.A2[(int64_t)(r0 + a) * n + gs.dst[c0 + b]] = m[a][b];
}

// Test kernel: A <- A^-1 for one dense n x n row-major matrix (n <= N) by
// gj_invert_tiles<N> in one workgroup; N = kGMax is the group inverses'
// width, kNG the level-2 inverse's.  *status set on a zero pivot.
template <int N>
__global__ void __launch_bounds__(TiledGJ<N>::NT) k_test_dense_invert(double *A, int n, int *status) {
  constexpr int TB = TiledGJ<N>::TB, NTC = TiledGJ<N>::NTC;
  __shared__ TiledGJ<N> gs;
  const int t = threadIdx.x;
  const int r0 = (t / NTC) * TB, c0 = (t % NTC) * TB;
  double m[TB][TB];
#pragma unroll
  for (int a = 0; a < TB; a++)
#pragma unroll
    for (int b = 0; b < TB; b++) m[a][b] = (r0 + a < n && c0 + b < n) ? A[(r0 + a) * n + c0 + b] : 0.0;
  if (!gj_invert_tiles<N>(m, n, gs)) {
    if (t == 0) *status = 1;
    return;
  }
#pragma unroll
  for (int a = 0; a < TB; a++)
#pragma unroll
    for (int b = 0; b < TB; b++)
      if (r0 + a < n && c0 + b < n) A[(r0 + a) * n + gs.dst[c0 + b]] = m[a][b];
}

// Level 1 and 2 of an application for group g (one workgroup of kL1Threads):
//   z1[b] = (G1 rc1)[b] + (A2^-1 r2)[g]  for the blocks b of g,
// with r2[h] the sum of rc1 over the blocks of group h.  Every workgroup
// forms all of r2 itself (ng*gsz reads, in a fixed order), so one launch
// finishes the coarse solve and the prolongation only adds z1[block].  The
// loads that do not depend on rc1 (this thread's G1 entries, its A2 entry)
// go first, the member lists' loads eight at a time.  z1 of the group's
// blocks is left in zs (LDS, zero past the members) and written to cd.z1;
// the members in ms.  Every thread calls it.
constexpr int kL1Threads = 256;
constexpr int kL1Chunks = kL1Threads / kGMax;  // column chunks of the G1 product
constexpr int kL1Cols = kGMax / kL1Chunks;     // columns per chunk
static_assert(kL1Threads >= 2 * kNG, "two threads per group for r2");
__device__ inline void coarse_z1(const CoarseDot &cd, int g, double *zs, int *ms) {
  __shared__ double r2s[kNG];
  __shared__ double rg[kGMax];
  __shared__ double zc[kL1Chunks][kGMax];
  __shared__ double sh[kL1Threads / 64];
  const int t = threadIdx.x, G = cd.gsz, ng = cd.ng;
  const int i = t % kGMax, ch = t / kGMax;  // row of G1, column chunk
  double gv[kL1Cols];
  const double *g1 = cd.G1 + (int64_t)g * G * G;  // column-major per group: entry (i, j) at j*G + i
#pragma unroll
  for (int u = 0; u < kL1Cols; u++) {
    const int j = ch * kL1Cols + u;
    gv[u] = (i < G && j < G) ? g1[j * G + i] : 0.0;
  }
  const double a2 = t < ng ? cd.A2[(int64_t)g * ng + t] : 0.0;
  // r2[h]: threads 2h and 2h+1 sum the two halves of group h's members in
  // order, then half 0 + half 1
  {
    const int h = t >> 1, hf = t & 1, HG = (G + 1) / 2;
    const int m0 = hf * HG, nm = h < ng ? max(0, min(G, m0 + HG) - m0) : 0;
    double s = 0.0;
    for (int k0 = 0; k0 < nm; k0 += 8) {
      int bm[8];
      double rv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) bm[u] = k0 + u < nm ? cd.gmem[h * G + m0 + k0 + u] : -1;
#pragma unroll
      for (int u = 0; u < 8; u++) rv[u] = bm[u] >= 0 ? cd.rc1[bm[u]] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (bm[u] >= 0) s += rv[u];
    }
    const double o = __shfl_xor(s, 1, 64);
    if (h < ng && hf == 0) r2s[h] = s + o;
  }
  const int bt = t < G ? cd.gmem[g * G + t] : -1;
  if (t < kGMax) {
    rg[t] = bt >= 0 ? cd.rc1[bt] : 0.0;
    ms[t] = bt;
  }
  __syncthreads();
  // z2 = row g of A2^-1 times r2
  const double z2 = block_sum_n<kL1Threads>(t < ng ? a2 * r2s[t] : 0.0, sh);
  // G1 rg by column chunks, the chunks added in order; padded members have
  // zero rows, columns and rg
  double z = 0.0;
#pragma unroll
  for (int u = 0; u < kL1Cols; u++) z += gv[u] * rg[ch * kL1Cols + u];
  zc[ch][i] = z;
  __syncthreads();
  if (t < kGMax) {
    double zz = 0.0;
    if (t < G) {
      zz = zc[0][t];
#pragma unroll
      for (int c = 1; c < kL1Chunks; c++) zz += zc[c][t];
      zz += z2;
      if (bt >= 0) cd.z1[bt] = zz;
    }
    zs[t] = bt >= 0 ? zz : 0.0;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kL1Threads) k_coarse_level1(CoarseDot cd, const int *__restrict__ stop) {
  __shared__ double zs[kGMax];
  __shared__ int ms[kGMax];
  if (stop && *stop) return;
  coarse_z1(cd, blockIdx.x, zs, ms);
}

// In a multi-dot launch: z1 of group g as above, then the group's coarse
// terms of the dots with the nv basis vectors.  V_v . (w_bj + P z1) splits
// into V_v . w_bj (the row-chunk workgroups) and RV_v . z1 with RV_v the
// block sums of V_v; this group adds sum over its blocks of RV_v[b] z1[b],
// eight members per thread and the eight chunks in order, to partial[v *
// pstride + g].
template <class TV>
__device__ void coarse_level1_dots(const CoarseDot &cd, int g, int nv, TV *partial, int pstride) {
  static_assert(kL1Threads == kRedBlock, "the multi-dot's workgroup size");
  __shared__ double zs[kGMax];
  __shared__ int ms[kGMax];
  __shared__ double cps[kGMax / 8][32];
  coarse_z1(cd, g, zs, ms);
  const int t = threadIdx.x, v = t % 32, c = t / 32;  // vector, chunk of eight members
  for (int v0 = 0; v0 < nv; v0 += 32) {
    double q = 0.0;
    if (v0 + v < nv) {
      const double *rv = cd.RV + (int64_t)(v0 + v) * cd.rv_ld;
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int b = ms[c * 8 + u];
        x[u] = b >= 0 ? rv[b] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) q += x[u] * zs[c * 8 + u];
    }
    cps[c][v] = q;
    __syncthreads();
    if (t < 32 && v0 + t < nv) {
      double p = cps[0][t];
#pragma unroll
      for (int k = 1; k < kGMax / 8; k++) p += cps[k][t];
      partial[(int64_t)(v0 + t) * pstride + g] = (TV)p;
    }
    __syncthreads();
  }
}

// Prolongation: out[row] += z1[b] for the rows of block b.
template <class TV, int BS>
__global__ void __launch_bounds__(BS) k_coarse_prolong(CoarseLayout C, const int32_t *__restrict__ rowof,
                                                       TV *__restrict__ out, const int *__restrict__ stop) {
  if (stop && *stop) return;
  const int b = blockIdx.x, t = threadIdx.x;
  const int row = rowof[(int64_t)b * BS + t];
  if (row >= 0) out[row] += (TV)C.z1[b];
}

}  // namespace tdyhip
