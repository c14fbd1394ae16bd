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

template <int MODE>
__device__ __forceinline__ void dh_load4(const Opnd& o, int row, int k, double v[4]) {
  if (MODE == M_ROW || MODE == M_UPK_ROW) {
    const double2* q = reinterpret_cast<const double2*>(o.p + (long)row * o.ld + k);
    double2 x0 = q[0], x1 = q[1];
    v[0] = x0.x; v[1] = x0.y; v[2] = x1.x; v[3] = x1.y;
    if (MODE == M_UPK_ROW) {
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        int kk = k + m;
        v[m] = kk > row ? v[m] : (kk == row ? o.dinv[row] : 0.0);
      }
    }
  } else {
    const double* q = o.p + (long)k * o.ld + row;
    v[0] = q[0]; v[1] = q[o.ld]; v[2] = q[2 * o.ld]; v[3] = q[3 * o.ld];
    if (MODE == M_UPK_COL) {
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
// k advances in chunks of 32 (klo, khi multiples of 32); the loads of the
// next chunk are issued before the 32 MFMAs of the current one.
template <int MA, int MB>
__device__ __forceinline__ void dh_load_chunk(const Opnd& A, const Opnd& B, int ra, int rb, int k,
                                              double a[2][2][4], double b[2][2][4]) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    dh_load4<MA>(A, ra, k + 16 * h, a[h][0]);
    dh_load4<MA>(A, ra + 16, k + 16 * h, a[h][1]);
    dh_load4<MB>(B, rb, k + 16 * h, b[h][0]);
    dh_load4<MB>(B, rb + 16, k + 16 * h, b[h][1]);
  }
}

template <int MA, int MB>
__device__ __forceinline__ void dh_mma(dh_d4 acc[2][2], const Opnd& A, int i0, const Opnd& B, int j0,
                                       int klo, int khi) {
  if (klo >= khi) return;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int ra = i0 + (w >> 1) * 32 + (lane & 15);
  const int rb = j0 + (w & 1) * 32 + (lane & 15);
  const int kq = (lane >> 4) * 4;
  double a[2][2][4], b[2][2][4];
  dh_load_chunk<MA, MB>(A, B, ra, rb, klo + kq, a, b);
  for (int k = klo; k < khi; k += 32) {
    double na[2][2][4], nb[2][2][4];
    const bool more = k + 32 < khi;
    if (more) dh_load_chunk<MA, MB>(A, B, ra, rb, k + 32 + kq, na, nb);
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        acc[0][0] = dh_mfma(a[h][0][m], b[h][0][m], acc[0][0]);
        acc[0][1] = dh_mfma(a[h][0][m], b[h][1][m], acc[0][1]);
        acc[1][0] = dh_mfma(a[h][1][m], b[h][0][m], acc[1][0]);
        acc[1][1] = dh_mfma(a[h][1][m], b[h][1][m], acc[1][1]);
      }
    if (more) {
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            a[h][x][m] = na[h][x][m];
            b[h][x][m] = nb[h][x][m];
          }
    }
  }
}

// 128x128 workgroup tile: wave w computes the 64x64 quadrant (w>>1, w&1) as
// 4x4 MFMA blocks; one 16-wide k chunk is prefetched ahead of 64 MFMAs.
template <int MA, int MB>
__device__ __forceinline__ void dh_mma128(dh_d4 acc[4][4], const Opnd& A, int i0, const Opnd& B, int j0,
                                          int klo, int khi) {
  if (klo >= khi) return;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int ra = i0 + (w >> 1) * 64 + (lane & 15);
  const int rb = j0 + (w & 1) * 64 + (lane & 15);
  const int kq = (lane >> 4) * 4;
  double a[4][4], b[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    dh_load4<MA>(A, ra + 16 * x, klo + kq, a[x]);
    dh_load4<MB>(B, rb + 16 * x, klo + kq, b[x]);
  }
  for (int k = klo; k < khi; k += 16) {
    double na[4][4], nb[4][4];
    const bool more = k + 16 < khi;
    if (more) {
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        dh_load4<MA>(A, ra + 16 * x, k + 16 + kq, na[x]);
        dh_load4<MB>(B, rb + 16 * x, k + 16 + kq, nb[x]);
      }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = dh_mfma(a[x][m], b[y][m], acc[x][y]);
    if (more) {
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          a[x][m] = na[x][m];
          b[x][m] = nb[x][m];
        }
    }
  }
}

__device__ __forceinline__ void dh_store128(double* C, long ldc, int i0, int j0, const dh_d4 acc[4][4],
                                            double alpha, double beta) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int rbase = i0 + (w >> 1) * 64 + (lane >> 4);
  const int cbase = j0 + (w & 1) * 64 + (lane & 15);
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double* p = C + (long)(rbase + 16 * x + 4 * r) * ldc + cbase + 16 * y;
        double v = alpha * acc[x][y][r];
        if (beta != 0.0) v += beta * (*p);
        *p = v;
      }
}

// Operand registers of opA[i0:i0+64, 0:64] for dh_mma_a64 (loaded ahead of
// the rest of the product).
template <int MA>
__device__ __forceinline__ void dh_load_a64(const Opnd& A, int i0, double a[2][2][2][4]) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int ra = i0 + (w >> 1) * 32 + (lane & 15);
  const int kq = (lane >> 4) * 4;
#pragma unroll
  for (int kc = 0; kc < 2; ++kc)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      dh_load4<MA>(A, ra, 32 * kc + 16 * h + kq, a[kc][h][0]);
      dh_load4<MA>(A, ra + 16, 32 * kc + 16 * h + kq, a[kc][h][1]);
    }
}

// acc += opA[i0:i0+64, 0:64] opB[j0:j0+64, 0:64]^T with opA from dh_load_a64.
template <int MB>
__device__ __forceinline__ void dh_mma_a64(dh_d4 acc[2][2], const double a[2][2][2][4], const Opnd& B, int j0) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int rb = j0 + (w & 1) * 32 + (lane & 15);
  const int kq = (lane >> 4) * 4;
  double b[2][2][2][4];
#pragma unroll
  for (int kc = 0; kc < 2; ++kc)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      dh_load4<MB>(B, rb, 32 * kc + 16 * h + kq, b[kc][h][0]);
      dh_load4<MB>(B, rb + 16, 32 * kc + 16 * h + kq, b[kc][h][1]);
    }
#pragma unroll
  for (int kc = 0; kc < 2; ++kc)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        acc[0][0] = dh_mfma(a[kc][h][0][m], b[kc][h][0][m], acc[0][0]);
        acc[0][1] = dh_mfma(a[kc][h][0][m], b[kc][h][1][m], acc[0][1]);
        acc[1][0] = dh_mfma(a[kc][h][1][m], b[kc][h][0][m], acc[1][0]);
        acc[1][1] = dh_mfma(a[kc][h][1][m], b[kc][h][1][m], acc[1][1]);
      }
}

// The inverse of dh_store's mapping: v = C[i0.., j0..] in accumulator layout.
__device__ __forceinline__ void dh_load_acc(const double* C, long ldc, int i0, int j0, dh_d4 v[2][2]) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int rbase = i0 + (w >> 1) * 32 + (lane >> 4);
  const int cbase = j0 + (w & 1) * 32 + (lane & 15);
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) v[bi][bj][r] = C[(long)(rbase + bi * 16 + 4 * r) * ldc + cbase + bj * 16];
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

// Cholesky of a 64x64 tile fused with the inverse of its factor.
// On exit the tile holds L in its lower triangle (diagonal included) and
// inv(L)^T in its strict upper triangle; dinv[k] = 1/L[k][k]; when inv_out
// is given the dense inv(L) (zeros above the diagonal, ld 64) is written
// there too.  A pivot that is not > 0 records 1 + row_base + k in *info
// (atomic min) and lets NaN propagate.
//
// Blocked right-looking elimination, 4 columns per barrier.  Thread
// (ty, tx) = (tid/16, tid%16) keeps the entries (ty+16u, tx+16v) in
// registers.  Step c (columns K0 = 4c .. K0+3) publishes those 4 columns of
// the trailing matrix S and rows K0..K0+3 of the working inverse W (L W = I
// is solved alongside) through LDS; after one barrier every thread factors
// the 4x4 pivot block P = L_B L_B^T and inverts it (V_B = inv(L_B))
// redundantly, then
//   L[i][K0+b] = sum_m S[i][K0+m] V_B[b][m]            (rows i >= K0+4)
//   V[K0+a][j] = sum_m V_B[a][m] W[K0+m][j]            (columns j < K0)
//   S[i][j]   -= sum_b L[i][K0+b] L[j][K0+b]            (i, j >= K0+4)
//   W[i][j]   -= sum_b L[i][K0+b] V[K0+b][j]            (i >= K0+4, j < K0+4)
// Entries above the diagonal carry unused values.  No global store is issued
// inside the loop (a barrier waits for outstanding stores): L goes to LDS,
// inv(L) stays in registers, and both are written once at the end.
struct LeafShm {
  double colb[2][4][64];  // columns K0..K0+3 of S (rows >= K0)
  double rowb[2][4][64];  // rows K0..K0+3 of W (columns < K0)
  double Ls[64][65];
};

__device__ __forceinline__ double dh_rsq_nr(double p) {
  double r = __builtin_amdgcn_rsq(p);
  r = r * (1.5 - 0.5 * p * r * r);
  r = r * (1.5 - 0.5 * p * r * r);
  return r;
}

template <int KV>
__device__ __forceinline__ void dh_leaf_block(double R[4][4], LeafShm& sh, int ty, int tx, int tid, int* info,
                                              int row_base, int c4) {
  const int kq = 4 * c4;
  const int K0 = 16 * KV + kq;
  const int bsel = c4 & 1;
  const bool colown = tx >= kq && tx < kq + 4;
  const bool rowown = ty >= kq && ty < kq + 4;
  if (colown) {
#pragma unroll
    for (int u = KV; u < 4; ++u) {
      sh.colb[bsel][tx - kq][ty + 16 * u] = R[u][KV];
      if (u > KV || ty >= kq + 4) R[u][KV] = 0.0;  // becomes W[i][K0+jj] below the block
    }
  }
  if (rowown) {
#pragma unroll
    for (int v = 0; v <= KV; ++v)
      if (v < KV || tx < kq) sh.rowb[bsel][ty - kq][tx + 16 * v] = R[KV][v];
  }
  __syncthreads();
  // 4x4 pivot block: Cholesky and inverse (every thread)
  double P[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b <= a; ++b) P[a][b] = sh.colb[bsel][b][K0 + a];
  double Lb[4][4], Vb[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) Lb[a][b] = Vb[a][b] = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double piv = P[k][k];
#pragma unroll
    for (int m = 0; m < k; ++m) piv -= Lb[k][m] * Lb[k][m];
    const double r = dh_rsq_nr(piv);
    if (tid == 0 && !(piv > 0.0)) atomicMin(info, 1 + row_base + K0 + k);
    Lb[k][k] = piv * r;
    Vb[k][k] = r;
#pragma unroll
    for (int a = k + 1; a < 4; ++a) {
      double x = P[a][k];
#pragma unroll
      for (int m = 0; m < k; ++m) x -= Lb[a][m] * Lb[k][m];
      Lb[a][k] = x * r;
    }
  }
#pragma unroll
  for (int a = 1; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < a; ++b) {
      double x = 0.0;
#pragma unroll
      for (int m = b; m < a; ++m) x += Lb[a][m] * Vb[m][b];
      Vb[a][b] = -Vb[a][a] * x;
    }
  // L columns K0..K0+3 into LDS (thread tid < 64 owns row tid)
  if (tid < 64) {
    const int i = tid;
    if (i >= K0 + 4) {
      double c[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) c[m] = sh.colb[bsel][m][i];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        double x = 0.0;
#pragma unroll
        for (int m = 0; m <= b; ++m) x += c[m] * Vb[b][m];
        sh.Ls[i][K0 + b] = x;
      }
    } else if (i >= K0) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        double x = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a) x = (a == i - K0) ? Lb[a][b] : x;
        sh.Ls[i][K0 + b] = x;
      }
    } else {
#pragma unroll
      for (int b = 0; b < 4; ++b) sh.Ls[i][K0 + b] = 0.0;
    }
  }
  // per column: S-part factors (L[j][K0+b]) or W-part factors (V[K0+b][j])
  double gj[4][4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int j = tx + 16 * v;
    if (v > KV || (v == KV && tx >= kq + 4)) {  // S part: j >= K0+4
      double c[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) c[m] = sh.colb[bsel][m][j];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        double x = 0.0;
#pragma unroll
        for (int m = 0; m <= b; ++m) x += c[m] * Vb[b][m];
        gj[v][b] = x;
      }
    } else if (v < KV || tx < kq) {  // W part: j < K0
      double w[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) w[m] = sh.rowb[bsel][m][j];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        double x = 0.0;
#pragma unroll
        for (int m = 0; m <= b; ++m) x += Vb[b][m] * w[m];
        gj[v][b] = x;
      }
    } else {  // in-block column j = K0 + (tx - kq): V[K0+b][j] = Vb[b][tx-kq]
      const int jj = tx - kq;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        double x = 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m) x = (m == jj && b >= m) ? Vb[b][m] : x;
        gj[v][b] = x;
      }
    }
  }
  // rank-4 updates of the rows below the block (lower and diagonal 16x16 sub-blocks only)
#pragma unroll
  for (int u = KV; u < 4; ++u) {
    const int i = ty + 16 * u;
    const bool below = (u > KV) || (ty >= kq + 4);
    double c[4], li[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) c[m] = below ? sh.colb[bsel][m][i] : 0.0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double x = 0.0;
#pragma unroll
      for (int m = 0; m <= b; ++m) x += c[m] * Vb[b][m];
      li[b] = x;
    }
#pragma unroll
    for (int v = 0; v <= u; ++v)
#pragma unroll
      for (int b = 0; b < 4; ++b) R[u][v] = fma(-li[b], gj[v][b], R[u][v]);
  }
  // rows of the block are final: row K0+aa of inv(L)
  if (rowown) {
    const int aa = ty - kq;
#pragma unroll
    for (int v = 0; v <= KV; ++v) {
      double x = 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b) x = (b == aa) ? gj[v][b] : x;
      R[KV][v] = x;
    }
  }
}

template <int KV>
__device__ __forceinline__ void dh_leaf_blocks(double R[4][4], LeafShm& sh, int ty, int tx, int tid, int* info,
                                               int row_base) {
#pragma unroll 1
  for (int c4 = 0; c4 < 4; ++c4) dh_leaf_block<KV>(R, sh, ty, tx, tid, info, row_base, c4);
}

// Factor the tile read from src (leading dimension lds; may be LDS) into
// tile (ld); src == tile factors in place.
__device__ void dh_leaf_potrf2(const double* src, long lds, double* tile, long ld, double* dinv, double* inv_out,
                               int* info, int row_base) {
  __shared__ LeafShm sh;
  const int tid = threadIdx.x;
  const int ty = tid >> 4, tx = tid & 15;
  double R[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = ty + 16 * u, j = tx + 16 * v;
      R[u][v] = (j <= i) ? src[(long)i * lds + j] : 0.0;
    }
  dh_leaf_blocks<0>(R, sh, ty, tx, tid, info, row_base);
  dh_leaf_blocks<1>(R, sh, ty, tx, tid, info, row_base);
  dh_leaf_blocks<2>(R, sh, ty, tx, tid, info, row_base);
  dh_leaf_blocks<3>(R, sh, ty, tx, tid, info, row_base);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = ty + 16 * u, j = tx + 16 * v;
      if (j <= i) {
        tile[(long)i * ld + j] = sh.Ls[i][j];
        if (j < i) tile[(long)j * ld + i] = R[u][v];  // inv(L)^T above the diagonal
        else dinv[i] = R[u][v];
      }
      if (inv_out) inv_out[i * 64 + j] = (j <= i) ? R[u][v] : 0.0;
    }
}

__device__ __forceinline__ void dh_leaf_potrf(double* tile, long ld, double* dinv, double* inv_out, int* info,
                                              int row_base) {
  dh_leaf_potrf2(tile, ld, tile, ld, dinv, inv_out, info, row_base);
}
