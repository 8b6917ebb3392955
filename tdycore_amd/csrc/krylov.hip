// Newton's inner linear solve on the device (subsystem (d)): CSR SpMV,
// block-Jacobi ILU(0) and GMRES(m), templated on the Krylov precision.
//
// The reference leaves this to PETSc (KSPGMRES, restart 30, left
// preconditioning, classical Gram-Schmidt; PCILU(0) serial / PCBJACOBI+ILU(0)
// parallel; driven from src/tdyti.c:135,188).
//
// Block-Jacobi ILU(0): cells are partitioned (host, breadth-first over the
// vertex-adjacency graph) into compact blocks of at most kBlk cells.  Inside a
// block rows keep their global order; entries coupling different blocks are
// dropped.  One workgroup owns one block: the factorization (IKJ) and the two
// triangular solves advance level by level (levels of the in-block
// dependency DAG, computed on the host) with the block's rows in LDS, so no
// grid-wide synchronisation is ever needed.
#include "tdyhip_internal.h"

namespace tdyhip {

constexpr int kBlk = 256;  // max rows per block (one workgroup)
constexpr int kW = 27;     // max in-block entries per row (hexahedral vertex adjacency)

struct BJLayout {
  int nblocks = 0;
  int maxlev = 0;
  int32_t *d_row = nullptr;    // nblocks*kBlk global row (-1 pad)
  int32_t *d_jpos = nullptr;   // nblocks*kW*kBlk position in J values (-1 pad), slot-major
  uint8_t *d_lcol = nullptr;   // nblocks*kW*kBlk local column (255 pad)
  uint8_t *d_diag = nullptr;   // nblocks*kBlk slot of the diagonal
  uint8_t *d_nslot = nullptr;  // nblocks*kBlk number of slots
  uint16_t *d_levL = nullptr;  // nblocks*kBlk
  uint16_t *d_levU = nullptr;  // nblocks*kBlk
  int32_t *d_nlev = nullptr;   // nblocks*2 (L, U level counts)
};

// ---------------------------------------------------------------- SpMV
template <class TV>
__global__ void __launch_bounds__(256) k_spmv(int n, const int32_t *__restrict__ rp, const int32_t *__restrict__ cj,
                                              const TV *__restrict__ a, const TV *__restrict__ x, TV *__restrict__ y) {
  // one row per 4-lane group would help long rows; 27-entry rows: one row per lane
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TV s = 0;
  const int e = rp[i + 1];
  for (int k = rp[i]; k < e; k++) s += a[k] * x[cj[k]];
  y[i] = s;
}

// ---------------------------------------------------------------- ILU(0) factor
template <class TV>
__global__ void __launch_bounds__(kBlk) k_bj_factor(BJLayout L, const double *__restrict__ J, TV *__restrict__ fac,
                                                    int *__restrict__ status) {
  __shared__ TV a[kW][kBlk];
  __shared__ uint8_t lc[kW][kBlk];
  __shared__ uint8_t dg[kBlk];
  __shared__ uint8_t ns[kBlk];
  const int b = blockIdx.x, t = threadIdx.x;
  const int64_t base = (int64_t)b * kBlk;
  const int64_t sbase = (int64_t)b * kW * kBlk;
  const int nsl = L.d_nslot[base + t];
  ns[t] = (uint8_t)nsl;
  dg[t] = L.d_diag[base + t];
  for (int p = 0; p < kW; p++) {
    const int32_t pos = L.d_jpos[sbase + (int64_t)p * kBlk + t];
    a[p][t] = (pos >= 0) ? (TV)J[pos] : (TV)0;
    lc[p][t] = L.d_lcol[sbase + (int64_t)p * kBlk + t];
  }
  const int lev = L.d_levL[base + t];
  const int nlev = L.d_nlev[2 * b];
  __syncthreads();
  for (int l = 0; l < nlev; l++) {
    if (lev == l && nsl > 0) {
      const int di = dg[t];
      for (int p = 0; p < di; p++) {  // lower entries, increasing column
        const int k = lc[p][t];
        const TV ukk = a[dg[k]][k];
        const TV lik = a[p][t] / ukk;
        a[p][t] = lik;
        // row i entries after p, merged with row k entries after its diagonal
        int pi = p + 1;
        const int nk = ns[k];
        for (int q = dg[k] + 1; q < nk; q++) {
          const int cq = lc[q][k];
          while (pi < nsl && lc[pi][t] < cq) pi++;
          if (pi < nsl && lc[pi][t] == cq) a[pi][t] -= lik * a[q][k];
        }
      }
      if (a[dg[t]][t] == (TV)0) atomicOr(status, 1);
    }
    __syncthreads();
  }
  for (int p = 0; p < kW; p++) fac[sbase + (int64_t)p * kBlk + t] = a[p][t];
}

// ---------------------------------------------------------------- ILU(0) apply: y = (LU)^-1 r
template <class TV>
__global__ void __launch_bounds__(kBlk) k_bj_apply(BJLayout L, const TV *__restrict__ fac, const TV *__restrict__ r,
                                                   TV *__restrict__ y) {
  __shared__ TV x[kBlk];
  const int b = blockIdx.x, t = threadIdx.x;
  const int64_t base = (int64_t)b * kBlk;
  const int64_t sbase = (int64_t)b * kW * kBlk;
  const int row = L.d_row[base + t];
  const int nsl = L.d_nslot[base + t];
  const int di = L.d_diag[base + t];
  TV v[kW];
  uint8_t c[kW];
#pragma unroll
  for (int p = 0; p < kW; p++) {
    v[p] = fac[sbase + (int64_t)p * kBlk + t];
    c[p] = L.d_lcol[sbase + (int64_t)p * kBlk + t];
  }
  x[t] = row >= 0 ? r[row] : (TV)0;
  const int levL = L.d_levL[base + t], levU = L.d_levU[base + t];
  const int nL = L.d_nlev[2 * b], nU = L.d_nlev[2 * b + 1];
  __syncthreads();
  for (int l = 0; l < nL; l++) {
    if (levL == l && row >= 0) {
      TV s = x[t];
#pragma unroll
      for (int p = 0; p < kW; p++)
        if (p < di) s -= v[p] * x[c[p]];
      x[t] = s;
    }
    __syncthreads();
  }
  for (int l = 0; l < nU; l++) {
    if (levU == l && row >= 0) {
      TV s = x[t];
#pragma unroll
      for (int p = 0; p < kW; p++)
        if (p > di && p < nsl) s -= v[p] * x[c[p]];
      TV dv = v[0];
#pragma unroll
      for (int p = 0; p < kW; p++)
        if (p == di) dv = v[p];
      x[t] = s / dv;
    }
    __syncthreads();
  }
  if (row >= 0) y[row] = x[t];
}

// ---------------------------------------------------------------- BLAS-1 with deterministic reductions
constexpr int kRedBlock = 256;
constexpr int kRedPerThread = 4;
constexpr int kChunk = kRedBlock * kRedPerThread;

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

// partial[j*nchunk + chunk] = <w, V_j> over the chunk, j = 0..nv-1
template <class TV>
__global__ void __launch_bounds__(kRedBlock) k_mdot(int64_t n, int nv, const TV *__restrict__ V, int64_t ldv,
                                                    const TV *__restrict__ w, TV *__restrict__ partial, int nchunk) {
  __shared__ TV sh[kRedBlock / 64];
  const int64_t i0 = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  TV wv[kRedPerThread];
#pragma unroll
  for (int e = 0; e < kRedPerThread; e++) {
    const int64_t i = i0 + (int64_t)e * kRedBlock;
    wv[e] = i < n ? w[i] : (TV)0;
  }
  for (int j = 0; j < nv; j++) {
    TV s = 0;
#pragma unroll
    for (int e = 0; e < kRedPerThread; e++) {
      const int64_t i = i0 + (int64_t)e * kRedBlock;
      if (i < n) s += wv[e] * V[j * ldv + i];
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[(int64_t)j * nchunk + blockIdx.x] = s;
  }
}

// out[j] = sum_c partial[j*nchunk + c] (fixed order)
template <class TV>
__global__ void k_reduce_partials(int nv, int nchunk, const TV *__restrict__ partial, double *__restrict__ out) {
  __shared__ TV sh[kRedBlock / 64];
  const int j = blockIdx.x;
  TV s = 0;
  for (int c = threadIdx.x; c < nchunk; c += kRedBlock) s += partial[(int64_t)j * nchunk + c];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[j] = (double)s;
}

// w -= sum_j h[j] V_j ; partial[chunk] = ||w||^2 over the chunk
template <class TV>
__global__ void __launch_bounds__(kRedBlock) k_maxpy_norm(int64_t n, int nv, const TV *__restrict__ V, int64_t ldv,
                                                          const double *__restrict__ h, TV *__restrict__ w,
                                                          TV *__restrict__ partial) {
  __shared__ TV sh[kRedBlock / 64];
  const int64_t i0 = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  TV s = 0;
#pragma unroll
  for (int e = 0; e < kRedPerThread; e++) {
    const int64_t i = i0 + (int64_t)e * kRedBlock;
    if (i < n) {
      TV x = w[i];
      for (int j = 0; j < nv; j++) x -= (TV)h[j] * V[j * ldv + i];
      w[i] = x;
      s += x * x;
    }
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// partial[chunk] = <x, y>
template <class TX, class TY>
__global__ void __launch_bounds__(kRedBlock) k_dot(int64_t n, const TX *__restrict__ x, const TY *__restrict__ y,
                                                   double *__restrict__ partial) {
  __shared__ double sh[kRedBlock / 64];
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

// x += sum_j c[j] V_j
template <class TV, class TX>
__global__ void k_gemv_update(int64_t n, int nv, const TV *__restrict__ V, int64_t ldv, const double *__restrict__ c,
                              TX *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TV s = 0;
  for (int j = 0; j < nv; j++) s += (TV)c[j] * V[j * ldv + i];
  x[i] += (TX)s;
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

// w = x - lambda*y
__global__ void k_waxpy(int64_t n, double lambda, const double *__restrict__ x, const double *__restrict__ y,
                        double *__restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = x[i] - lambda * y[i];
}

}  // namespace tdyhip
