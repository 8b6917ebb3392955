// Transmissibility setup for general meshes (hexahedra and right prisms):
// interaction regions of uneven size, bucketed by size.
//
// A vertex of a prism mesh with d triangles around it has 2d cells and 3d
// faces, so its local system is 3d x 3d (tdympfao.c:1126-1513 sized by the
// region): 18 x 18 at d = 6, up to 36 x 36 and more on a random Delaunay
// triangulation, against 12 x 12 for an interior hexahedral vertex.  The host
// sorts the vertices into buckets by max(unknowns, columns) <= 16, 32, 64; a
// bucket's regions take a team of LANES lanes each, 64 / LANES regions per
// one-wavefront workgroup step, so a small region does not pay for the
// largest one, and one persistent launch (k_ir_persistent) drains every
// bucket.
//
// Per region the team assembles [A | B] in LDS (lane = equation), eliminates
// with partial pivoting applied to B as it goes (the same operations, in the
// same order, as LU factorisation followed by the permuted forward solve),
// back-substitutes one column of B per lane, and writes the region's T block
// (rows: internal faces, then Dirichlet faces; columns: cells, then Dirichlet
// faces) row-major, one column per lane.  The sums follow the oracle's order
// and nothing here is contracted into fused multiply-adds, so the results
// agree with the oracle's bit for bit.
#include "tdyhip_internal.h"

namespace tdyhip {

struct GenArgs {
  const int32_t *vlist, *vdesc, *csub, *csck, *fa, *fb;
  const double *Gsub;
  double *T;
  int *status;
  int n;
  int maxU, pitch;  // the bucket's largest unknown count; LDS row pitch of [A | B] (odd)
};

template <int LANES>
__device__ inline void team_argmax(double &v, int &i) {
  // largest |value|, first index among equals (the serial scan's choice)
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, LANES);
    const int oi = __shfl_xor(i, off, LANES);
    if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
  }
}

// C row of subcell (i, l): the three (column, -G) pairs sorted by column, and
// F = -G0 - G1 - G2 (ComputeCandFmatrix, tdympfao.c:1000-1076).
__device__ inline double gen_crow(const GenArgs &a, int oc, int of, int nU, int i, int l, int (&col)[3],
                                  double (&val)[3]) {
#pragma clang fp contract(off)
  const double *G = a.Gsub + (size_t)a.csub[oc + i] * 9 + l * 3;
  const int pk = a.csck[oc + i];
  for (int m = 0; m < 3; m++) {
    const int km = (pk >> (8 * m)) & 0xff;
    const int fam = a.fa[of + km];
    const int role = fam & 3, fidx = (fam >> 8) & 0xff;
    col[m] = role == 2 ? nU + fidx : fidx;
    val[m] = -G[m];
  }
  const double F = -G[0] - G[1] - G[2];
  // sort the three pairs by column (the columns are distinct)
  if (col[0] > col[1]) { int t = col[0]; col[0] = col[1]; col[1] = t; double u = val[0]; val[0] = val[1]; val[1] = u; }
  if (col[1] > col[2]) { int t = col[1]; col[1] = col[2]; col[2] = t; double u = val[1]; val[1] = val[2]; val[2] = u; }
  if (col[0] > col[1]) { int t = col[0]; col[0] = col[1]; col[1] = t; double u = val[0]; val[0] = val[1]; val[1] = u; }
  return F;
}

__device__ inline double crow_at(const int (&col)[3], const double (&val)[3], int j) {
  return col[0] == j ? val[0] : col[1] == j ? val[1] : col[2] == j ? val[2] : 0.0;
}

// One workgroup step (64 lanes, one wavefront): regions blk * TEAMS + team of
// the bucket, one team of LANES lanes each.  [A | B] of a region sits in LDS,
// rows of a.pitch doubles (odd: column reads spread over the banks), sized by
// the bucket's largest region.
template <int LANES>
__device__ inline void ir_block(const GenArgs &a, int blk, double *lds) {
#pragma clang fp contract(off)
  constexpr int TEAMS = 64 / LANES;
  const int P = a.pitch;
  const int team = threadIdx.x / LANES, lane = threadIdx.x % LANES;
  const int vi = blk * TEAMS + team;
  const bool act = vi < a.n;  // idle teams still reach every barrier
  double *A = lds + (size_t)team * a.maxU * P, *B = A + a.maxU;
  int ncell = 0, nface = 0, nU = 0, ndir = 0, oc = 0, of = 0, ot = 0;
  if (act) {
    const int32_t *d = a.vdesc + (size_t)a.vlist[vi] * 8;
    ncell = d[0]; nface = d[1]; nU = d[2]; ndir = d[3]; oc = d[4]; of = d[5]; ot = d[6];
  }
  const int nB = ncell + ndir;
  for (int e = lane; e < a.maxU * P; e += LANES) A[e] = 0.0;
  __syncthreads();
  // assemble: one vertex face (equation) per lane
  for (int k = lane; k < nface; k += LANES) {
    const int fak = a.fa[of + k], fbk = a.fb[of + k];
    const int role = fak & 3;
    if (role == 2) continue;
    const int row = (fak >> 8) & 0xff, ui = (fak >> 16) & 0xff, ul = (fak >> 24) & 0xff;
    int cu[3], cd[3] = {-1, -1, -1};
    double vu[3], vd[3] = {0.0, 0.0, 0.0};
    const double Fu = gen_crow(a, oc, of, nU, ui, ul, cu, vu);
    double Fd = 0.0;
    int di = -1;
    if (role == 0) {
      di = fbk & 0xff;
      Fd = gen_crow(a, oc, of, nU, di, (fbk >> 8) & 0xff, cd, vd);
    }
    // A row = -C_up + C_dn over the unknowns (role 0), -C_up (no-flow)
    for (int m = 0; m < 3; m++) {
      if (cu[m] < nU) A[row * P + cu[m]] = -vu[m] + crow_at(cd, vd, cu[m]);
      if (role == 0 && cd[m] < nU && cd[m] != cu[0] && cd[m] != cu[1] && cd[m] != cu[2])
        A[row * P + cd[m]] = -0.0 + vd[m];
    }
    B[row * P + ui] += -Fu;
    if (role == 0) B[row * P + di] += Fd;
    for (int j = 0; j < ndir; j++) {
      const double u = crow_at(cu, vu, nU + j);
      B[row * P + ncell + j] = role == 0 ? u - crow_at(cd, vd, nU + j) : u;
    }
  }
  __syncthreads();
  // elimination with partial pivoting on [A | B]; every team runs maxU
  // steps (the workgroup's barriers), a region's steps past its nU are empty
  int bad = 0;
  for (int k = 0; k < a.maxU; k++) {
    const bool step = k < nU && !bad;
    double mx = (lane >= k && lane < nU) ? fabs(A[lane * P + k]) : -1.0;
    int p = (lane >= k && lane < nU) ? lane : LANES;
    team_argmax<LANES>(mx, p);
    if (step && !(mx > 0.0)) bad = 1;  // uniform over the team
    __syncthreads();
    if (step && !bad && p != k) {
      for (int j = lane; j < nU; j += LANES) { const double t = A[k * P + j]; A[k * P + j] = A[p * P + j]; A[p * P + j] = t; }
      for (int j = lane; j < nB; j += LANES) { const double t = B[k * P + j]; B[k * P + j] = B[p * P + j]; B[p * P + j] = t; }
    }
    __syncthreads();
    if (step && !bad && lane > k && lane < nU) {
      const double inv = 1.0 / A[k * P + k];
      const double l = A[lane * P + k] * inv;
      A[lane * P + k] = l;
      for (int j = k + 1; j < nU; j++) {
        const double akj = A[k * P + j];
        if (akj != 0.0) A[lane * P + j] -= l * akj;
      }
      for (int j = 0; j < nB; j++) B[lane * P + j] -= l * B[k * P + j];
    }
    __syncthreads();
  }
  if (act && bad) {
    if (lane == 0) atomicOr(a.status, 1);
  }
  // back substitution, one column of B per lane: X = U^-1 B
  if (!bad && lane < nB) {
    for (int k = nU - 1; k >= 0; k--) {
      B[k * P + lane] /= A[k * P + k];
      const double s = B[k * P + lane];
      for (int i = 0; i < k; i++) B[i * P + lane] -= A[i * P + k] * s;
    }
  }
  __syncthreads();
  if (act && !bad) {
  // T = C_up X - F_up (+ C_up,dir) for internal and Dirichlet faces
  for (int k = 0; k < nface; k++) {
    const int fak = a.fa[of + k], fbk = a.fb[of + k];
    const int trow = (fbk >> 16) - 1;
    if (trow < 0) continue;
    const int ui = (fak >> 16) & 0xff, ul = (fak >> 24) & 0xff;
    int cu[3];
    double vu[3];
    const double Fs = gen_crow(a, oc, of, nU, ui, ul, cu, vu);
    if (lane < nB) {
      double s = 0.0;
      for (int m = 0; m < 3; m++)
        if (cu[m] < nU) s += vu[m] * B[cu[m] * P + lane];
      const double t = lane < ncell ? s + ((lane == ui) ? -Fs : 0.0) : s + crow_at(cu, vu, nU + lane - ncell);
      a.T[(int64_t)ot + (int64_t)trow * nB + lane] = t;
    }
  }
  }
  __syncthreads();  // the LDS is reused by the workgroup's next step
}

// One persistent launch over every bucket.  Work items are chunks of `chunk`
// workgroup steps of one bucket, laid out segment by segment in the
// schedule's bucket order; each one-wavefront workgroup takes the next item
// from a device counter (zeroed before the launch) until none is left, which
// every workgroup reaches.  The LDS holds the largest bucket's tiles.
struct IRSchedule {
  GenArgs ga[3];      // per bucket: region list, count, largest region, pitch
  int lanes[3];       // team width of each bucket (16, 32, 64)
  int nseg;           // buckets in schedule order
  int seg_bucket[3];
  int seg_end[3];     // items up to the end of each segment
  int chunk;          // workgroup steps per item
  int *counter;
};

__global__ void __launch_bounds__(64) k_ir_persistent(IRSchedule sc) {
  extern __shared__ double lds[];
  const int total = sc.seg_end[sc.nseg - 1];
  for (;;) {
    int item = 0;
    if (threadIdx.x == 0) item = atomicAdd(sc.counter, 1);
    item = __shfl(item, 0);
    if (item >= total) return;
    int s = 0;
    while (item >= sc.seg_end[s]) s++;
    const int b = sc.seg_bucket[s];
    const GenArgs &a = sc.ga[b];
    const int first = (item - (s ? sc.seg_end[s - 1] : 0)) * sc.chunk;
    const int L = sc.lanes[b];
    const int nblk = (a.n + 64 / L - 1) / (64 / L);
    const int last = min(first + sc.chunk, nblk);
    for (int blk = first; blk < last; blk++) {
      if (L == 16) ir_block<16>(a, blk, lds);
      else if (L == 32) ir_block<32>(a, blk, lds);
      else ir_block<64>(a, blk, lds);
    }
  }
}

}  // namespace tdyhip
