// Tile-level device building blocks for the FP64 BTA kernels on gfx950.
//
// A workgroup is 256 threads = 4 wave64s and owns one 64x64 output tile;
// wave w computes the 32x32 quadrant (w>>1, w&1) as 2x2 blocks of the
// v_mfma_f64_16x16x4 instruction.  Operands are streamed straight from
// global memory (L2) in chunks of 16 along k: every lane fetches 4
// consecutive k of one row, and the 4 MFMAs of a chunk use the k-permuted
// order {4g+m}, which leaves the contraction unchanged because A and B use
// the same permutation.
#pragma once
#include <hip/hip_runtime.h>

#define DH_TILE 64
#define DH_NT 256

typedef double dh_d4 __attribute__((ext_vector_type(4)));

// Operand access modes: every operand is read as opM[row][k], where opM is
// op(A) for the left operand and op(B)^T for the right one.
enum {
  M_ROW = 0,      // opM[row][k] = P[row*ld + k]
  M_COL = 1,      // opM[row][k] = P[k*ld + row]
  M_UPK_ROW = 2,  // packed inverse factor, opM[row][k] = P[row][k] (k>row), dinv[row] (k==row), 0 (k<row)
  M_UPK_COL = 3,  // packed inverse factor, opM[row][k] = P[k][row] (k<row), dinv[row] (k==row), 0 (k>row)
};

struct Opnd {
  const double* p;
  long ld;
  int mode;
  const double* dinv;
};

__device__ __forceinline__ void dh_load4(const Opnd& o, int row, int k, double v[4]) {
  if (o.mode == M_ROW || o.mode == M_UPK_ROW) {
    const double2* q = reinterpret_cast<const double2*>(o.p + (long)row * o.ld + k);
    double2 x0 = q[0], x1 = q[1];
    v[0] = x0.x; v[1] = x0.y; v[2] = x1.x; v[3] = x1.y;
    if (o.mode == M_UPK_ROW) {
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        int kk = k + m;
        v[m] = kk > row ? v[m] : (kk == row ? o.dinv[row] : 0.0);
      }
    }
  } else {
    const double* q = o.p + (long)k * o.ld + row;
    v[0] = q[0]; v[1] = q[o.ld]; v[2] = q[2 * o.ld]; v[3] = q[3 * o.ld];
    if (o.mode == M_UPK_COL) {
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        int kk = k + m;
        v[m] = kk < row ? v[m] : (kk == row ? o.dinv[row] : 0.0);
      }
    }
  }
}

__device__ __forceinline__ dh_d4 dh_mfma(double a, double b, dh_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void dh_zero(dh_d4 acc[2][2]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = dh_d4{0.0, 0.0, 0.0, 0.0};
}

// acc += opA[i0:i0+64, klo:khi] * opB[klo:khi, j0:j0+64] for this wave's quadrant.
__device__ __forceinline__ void dh_mma(dh_d4 acc[2][2], const Opnd& A, int i0, const Opnd& B, int j0,
                                       int klo, int khi) {
  if (klo >= khi) return;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int ra = i0 + (w >> 1) * 32 + (lane & 15);
  const int rb = j0 + (w & 1) * 32 + (lane & 15);
  const int kq = (lane >> 4) * 4;
  double a0[4], a1[4], b0[4], b1[4];
  dh_load4(A, ra, klo + kq, a0);
  dh_load4(A, ra + 16, klo + kq, a1);
  dh_load4(B, rb, klo + kq, b0);
  dh_load4(B, rb + 16, klo + kq, b1);
  for (int k = klo; k < khi; k += 16) {
    double na0[4], na1[4], nb0[4], nb1[4];
    const int kn = k + 16;
    if (kn < khi) {
      dh_load4(A, ra, kn + kq, na0);
      dh_load4(A, ra + 16, kn + kq, na1);
      dh_load4(B, rb, kn + kq, nb0);
      dh_load4(B, rb + 16, kn + kq, nb1);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      acc[0][0] = dh_mfma(a0[m], b0[m], acc[0][0]);
      acc[0][1] = dh_mfma(a0[m], b1[m], acc[0][1]);
      acc[1][0] = dh_mfma(a1[m], b0[m], acc[1][0]);
      acc[1][1] = dh_mfma(a1[m], b1[m], acc[1][1]);
    }
    if (kn < khi) {
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        a0[m] = na0[m]; a1[m] = na1[m]; b0[m] = nb0[m]; b1[m] = nb1[m];
      }
    }
  }
}

// C[i0.., j0..] = alpha*acc + beta*C (trans: C[j][i] instead of C[i][j]).
// Lane l, item r of the 16x16 f64 MFMA result holds row 4*r + l/16 and
// column l%16 (measured on gfx950 with one-hot operands).
__device__ __forceinline__ void dh_store(double* C, long ldc, int i0, int j0, const dh_d4 acc[2][2],
                                         double alpha, double beta, bool trans) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int rbase = i0 + (w >> 1) * 32 + (lane >> 4);
  const int cbase = j0 + (w & 1) * 32 + (lane & 15);
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = rbase + bi * 16 + 4 * r;
        const int j = cbase + bj * 16;
        double* p = trans ? (C + (long)j * ldc + i) : (C + (long)i * ldc + j);
        double v = alpha * acc[bi][bj][r];
        if (beta != 0.0) v += beta * (*p);
        *p = v;
      }
}

// Unblocked Cholesky of a 64x64 tile plus the inverse of its factor.
// On exit the tile holds L in its lower triangle (diagonal included) and
// inv(L)^T in its strict upper triangle; dinv[k] = 1/L[k][k]; when inv_out
// is given the dense inv(L) (zeros above the diagonal, ld 64) is written
// there too.  A pivot that is not > 0 records 1 + row_base + k in *info
// (atomic min) and lets NaN propagate.  LDS: one 64 x 65 tile whose strict
// upper triangle receives inv(L)^T while the lower triangle holds L.
__device__ void dh_leaf_potrf(double* tile, long ld, double* dinv, double* inv_out, int* info,
                              int row_base) {
  __shared__ double S[64][65];
  __shared__ double dg[64];
  __shared__ double vd[64];
  const int tid = threadIdx.x;
  for (int e = tid; e < 4096; e += DH_NT) {
    int r = e >> 6, c = e & 63;
    S[r][c] = tile[(long)r * ld + c];
  }
  __syncthreads();
  // right-looking elimination with deferred column scaling: one barrier per column
  for (int k = 0; k < 63; ++k) {
    const double rp = 1.0 / S[k][k];
    const int m = 63 - k;  // trailing order
    for (int e = tid; e < m * m; e += DH_NT) {
      const int i = k + 1 + e / m;
      const int j = k + 1 + e % m;
      if (j <= i) S[i][j] -= S[i][k] * S[j][k] * rp;
    }
    __syncthreads();
  }
  if (tid < 64) {
    const double p = S[tid][tid];
    if (!(p > 0.0)) atomicMin(info, 1 + row_base + tid);
    dg[tid] = sqrt(p);
  }
  __syncthreads();
  for (int e = tid; e < 4096; e += DH_NT) {
    int i = e >> 6, k = e & 63;
    if (i > k) S[i][k] = S[i][k] / dg[k];
    else if (i == k) S[i][k] = dg[k];
  }
  __syncthreads();
  // inverse by forward substitution, column j by 4 lanes of one wave;
  // V[i][j] (i > j) is kept at S[j][i].  Every lane of the group writes the
  // same value, so each lane only reads entries it wrote itself.
  {
    const int j = tid >> 2, sub = tid & 3;
    const double vjj = 1.0 / S[j][j];
    if (sub == 0) vd[j] = vjj;
    for (int i = j + 1; i < 64; ++i) {
      double part = (sub == 0) ? S[i][j] * vjj : 0.0;
      for (int k = j + 1 + sub; k < i; k += 4) part += S[i][k] * S[j][k];
      part += __shfl_xor(part, 1);
      part += __shfl_xor(part, 2);
      S[j][i] = -part / S[i][i];
    }
  }
  __syncthreads();
  for (int e = tid; e < 4096; e += DH_NT) {
    int r = e >> 6, c = e & 63;
    tile[(long)r * ld + c] = S[r][c];
    if (inv_out) inv_out[r * 64 + c] = (r > c) ? S[c][r] : (r == c ? vd[r] : 0.0);
  }
  if (tid < 64) dinv[tid] = vd[tid];
}
