// Tile-level device building blocks for the FP64 BTA kernels on gfx950.
//
// A workgroup is 256 threads = 4 wave64s and owns one 64x64 output tile;
// wave w computes the 32x32 quadrant (w>>1, w&1) as 2x2 blocks of the
// v_mfma_f64_16x16x4 instruction.  Operands are streamed straight from
// global memory (L2) in chunks of 16 along k: lane group g = lane/16 holds
// the k offsets {2g, 2g+1, 2g+8, 2g+9} of its row, so that each load
// instruction of a row-major operand reads 64 contiguous bytes per row, and
// MFMA m of a chunk contracts over offset m of every group.  A and B use
// the same permutation of k, which leaves the contraction unchanged.
#pragma once
#include <hip/hip_runtime.h>

#define DH_TILE 64
#ifndef DH_LEAF_TS
#define DH_LEAF_TS(k)  // phase hook of the leaf steps (timing probes define it)
#endif
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

// k offset of item m relative to the group base (chunk start + 2g).
__device__ __forceinline__ constexpr int dh_koff(int m) { return (m & 1) + 8 * (m >> 1); }

// v[m] = opM[row][k + dh_koff(m)], k = chunk start + 2g.
template <int MODE>
__device__ __forceinline__ void dh_load4(const Opnd& o, int row, int k, double v[4]) {
  if (MODE == M_ROW || MODE == M_UPK_ROW) {
    const double* q = o.p + (long)row * o.ld + k;
    const double2 x0 = *reinterpret_cast<const double2*>(q);
    const double2 x1 = *reinterpret_cast<const double2*>(q + 8);
    v[0] = x0.x; v[1] = x0.y; v[2] = x1.x; v[3] = x1.y;
    if (MODE == M_UPK_ROW) {  // one unconditional load, then selects (no lane-divergent branch)
      const double d = o.dinv[row];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int kk = k + dh_koff(m);
        v[m] = kk > row ? v[m] : (kk == row ? d : 0.0);
      }
    }
  } else {
    const double* q = o.p + (long)k * o.ld + row;
    v[0] = q[0]; v[1] = q[o.ld]; v[2] = q[8 * o.ld]; v[3] = q[9 * o.ld];
    if (MODE == M_UPK_COL) {
      const double d = o.dinv[row];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int kk = k + dh_koff(m);
        v[m] = kk < row ? v[m] : (kk == row ? d : 0.0);
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
// k advances in chunks of 16 (klo, khi multiples of 16) held in a ring of S
// register buffers: the loads of chunk q + S - 1 are issued before the 16
// MFMAs of chunk q.  One chunk ahead (S = 2) keeps the register count low
// enough for three waves per SIMD, whose interleaving hides the rest of the
// L2 latency; the persistent factorization uses the same depth (deeper rings
// measured slower).
#ifndef DH_MMA_STAGES
#define DH_MMA_STAGES 2
#endif
template <int MA, int MB>
__device__ __forceinline__ void dh_load16(const Opnd& A, const Opnd& B, int ra, int rb, int k, double a[2][4],
                                          double b[2][4]) {
  dh_load4<MA>(A, ra, k, a[0]);
  dh_load4<MA>(A, ra + 16, k, a[1]);
  dh_load4<MB>(B, rb, k, b[0]);
  dh_load4<MB>(B, rb + 16, k, b[1]);
}

// DESC: chunks in decreasing k (from khi down).  Tiles of a triangular operand whose k ranges
// share their upper end then read the same chunks at the same time, as tiles with a common
// lower end do in increasing order, and share them in L2.
template <int MA, int MB, int S = DH_MMA_STAGES, bool DESC = false>
__device__ __forceinline__ void dh_mma(dh_d4 acc[2][2], const Opnd& A, int i0, const Opnd& B, int j0,
                                       int klo, int khi) {
  if (klo >= khi) return;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int ra = i0 + (w >> 1) * 32 + (lane & 15);
  const int rb = j0 + (w & 1) * 32 + (lane & 15);
  const int kq = (lane >> 4) * 2;  // group base offset within a 16-chunk
  const int nck = (khi - klo) >> 4;
  const int k0 = (DESC ? khi - 16 : klo) + kq;
  const int dk = DESC ? -16 : 16;
  double a[S][2][4], b[S][2][4];
#pragma unroll
  for (int s = 0; s < S - 1; ++s)
    if (s < nck) dh_load16<MA, MB>(A, B, ra, rb, k0 + dk * s, a[s], b[s]);
  for (int q = 0; q < nck; q += S) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int ck = q + s;
      if (ck < nck) {
        if (ck + S - 1 < nck)
          dh_load16<MA, MB>(A, B, ra, rb, k0 + dk * (ck + S - 1), a[(s + S - 1) % S], b[(s + S - 1) % S]);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          acc[0][0] = dh_mfma(a[s][0][m], b[s][0][m], acc[0][0]);
          acc[0][1] = dh_mfma(a[s][0][m], b[s][1][m], acc[0][1]);
          acc[1][0] = dh_mfma(a[s][1][m], b[s][0][m], acc[1][0]);
          acc[1][1] = dh_mfma(a[s][1][m], b[s][1][m], acc[1][1]);
        }
      }
    }
  }
}

// Two 64x64 tiles that share one operand panel: acc0 += opA[i0..] . opB[j0..] and acc1 += opA[i1..] .
// opB[j1..] over k in [klo, khi), with i0 == i1 (SHARE_A: one row panel of opA against two of opB) or
// j0 == j1 (the converse).  The shared panel's chunk is loaded once for both tiles: 24 loads per lane
// for 32 MFMAs per wave and chunk, against 16 for 16 with dh_mma.  Same k order and ring as dh_mma,
// so each tile's sum is bitwise the one dh_mma forms.
template <int MA, int MB, bool SHARE_A>
__device__ __forceinline__ void dh_mma_pair(dh_d4 acc0[2][2], dh_d4 acc1[2][2], const Opnd& A, int i0, int i1,
                                            const Opnd& B, int j0, int j1, int klo, int khi) {
  if (klo >= khi) return;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int ra0 = i0 + (w >> 1) * 32 + (lane & 15), ra1 = i1 + (w >> 1) * 32 + (lane & 15);
  const int rb0 = j0 + (w & 1) * 32 + (lane & 15), rb1 = j1 + (w & 1) * 32 + (lane & 15);
  const int kq = (lane >> 4) * 2;
  const int nck = (khi - klo) >> 4;
  const int k0 = klo + kq;
  // x: the shared panel (A rows ra0, or B rows rb0), y0 / y1: the two others
  double x[2][2][4], y0[2][2][4], y1[2][2][4];
  auto load = [&](int st, int k) {
    if (SHARE_A) {
      dh_load4<MA>(A, ra0, k, x[st][0]);
      dh_load4<MA>(A, ra0 + 16, k, x[st][1]);
      dh_load4<MB>(B, rb0, k, y0[st][0]);
      dh_load4<MB>(B, rb0 + 16, k, y0[st][1]);
      dh_load4<MB>(B, rb1, k, y1[st][0]);
      dh_load4<MB>(B, rb1 + 16, k, y1[st][1]);
    } else {
      dh_load4<MB>(B, rb0, k, x[st][0]);
      dh_load4<MB>(B, rb0 + 16, k, x[st][1]);
      dh_load4<MA>(A, ra0, k, y0[st][0]);
      dh_load4<MA>(A, ra0 + 16, k, y0[st][1]);
      dh_load4<MA>(A, ra1, k, y1[st][0]);
      dh_load4<MA>(A, ra1 + 16, k, y1[st][1]);
    }
  };
  load(0, k0);
  for (int q = 0; q < nck; q += 2) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int ck = q + s;
      if (ck < nck) {
        if (ck + 1 < nck) load(s ^ 1, k0 + 16 * (ck + 1));
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
          for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int v = 0; v < 2; ++v) {
              if (SHARE_A) {
                acc0[u][v] = dh_mfma(x[s][u][m], y0[s][v][m], acc0[u][v]);
                acc1[u][v] = dh_mfma(x[s][u][m], y1[s][v][m], acc1[u][v]);
              } else {
                acc0[u][v] = dh_mfma(y0[s][u][m], x[s][v][m], acc0[u][v]);
                acc1[u][v] = dh_mfma(y1[s][u][m], x[s][v][m], acc1[u][v]);
              }
            }
        }
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
  const int kq = (lane >> 4) * 2;  // group base offset within a 16-chunk
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

// C = alpha*acc + beta*C on a 128x128 tile.
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
  const int kq = (lane >> 4) * 2;  // group base offset within a 16-chunk
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
  const int kq = (lane >> 4) * 2;  // group base offset within a 16-chunk
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
// Blocked right-looking elimination, 4 columns per step.  Thread
// (ty, tx) = (tid/16, tid%16) keeps the entries (ty+16u, tx+16v) in
// registers.  Step c (columns K0 = 4c .. K0+3) publishes through LDS the 4x4
// pivot block P of the trailing matrix S and, for every index j, the
// 4-vector U[j][m] = S[j][K0+m] (j >= K0+4), delta(m, j-K0) (j in the block)
// or W[K0+m][j] (j < K0), W being the working inverse (L W = I is solved
// alongside).  After a barrier every thread factors P = L_B L_B^T and inverts
// it (V_B = inv(L_B)) redundantly, and one thread per index j forms
//   G[j][b] = sum_m U[j][m] V_B[b][m]
// (= L[j][K0+b] below the block, V[K0+b][j] above and in it) in LDS; after a
// second barrier every thread applies
//   S[i][j] -= sum_b G[i][b] G[j][b]  (i, j >= K0+4),  W[i][j] -= sum_b G[i][b] G[j][b]  (j < K0+4)
// to its own entries of the rows below the block, and the owners of rows
// K0..K0+3 replace them by the final inverse rows V[K0+a][j] = G[j][a].
// FP64 VALU instructions cost about 9 cycles each per wave on gfx950, so the
// step avoids per-lane selections and divergent paths.  Entries above the
// diagonal carry unused values.  No global store is issued inside the loop
// (a barrier waits for outstanding stores): L goes to LDS, inv(L) stays in
// registers, and both are written once at the end.
template <int NB>  // tile of 16*NB rows
struct __align__(16) LeafShmT {
  double U[16 * NB][4];  // column factor operands of the step
  double G[16 * NB][4];  // column factors of the step
  double Pb[4][4];       // pivot block
  double Ls[16 * NB][16 * NB + 1];
};
typedef LeafShmT<4> LeafShm;

__device__ __forceinline__ double dh_rsq_nr(double p) {
  // v_rsq_f64 is accurate to about 2^-24 (relative); one third-order
  // correction r (1 + e/2 + 3e^2/8), e = 1 - p r^2, brings it to double precision.
  const double r = __builtin_amdgcn_rsq(p);
  const double e = fma(-p * r, r, 1.0);
  return fma(r * e, fma(e, 0.375, 0.5), r);
}

__device__ __forceinline__ void dh_ld4(const double* q, double v[4]) {
  const double2 x0 = reinterpret_cast<const double2*>(q)[0];
  const double2 x1 = reinterpret_cast<const double2*>(q)[1];
  v[0] = x0.x; v[1] = x0.y; v[2] = x1.x; v[3] = x1.y;
}

__device__ __forceinline__ void dh_st4(double* q, const double v[4]) {
  reinterpret_cast<double2*>(q)[0] = double2{v[0], v[1]};
  reinterpret_cast<double2*>(q)[1] = double2{v[2], v[3]};
}

template <int NB, int KV>
__device__ __forceinline__ void dh_leaf_block(double R[NB][NB], LeafShmT<NB>& sh, int ty, int tx, int tid,
                                              int* info, int row_base, int c4) {
  const int kq = 4 * c4;
  const int K0 = 16 * KV + kq;
  DH_LEAF_TS(0);
  if (tx >= kq && tx < kq + 4) {  // owners of columns K0..K0+3
    const int m = tx - kq;
#pragma unroll
    for (int u = KV; u < NB; ++u) {
      const int i = ty + 16 * u;
      if (u > KV || ty >= kq + 4) {
        sh.U[i][m] = R[u][KV];
        R[u][KV] = 0.0;  // becomes W[i][K0+m] below the block
      } else if (ty >= kq) {
        sh.Pb[ty - kq][m] = R[u][KV];
        sh.U[i][m] = (ty - kq == m) ? 1.0 : 0.0;
      }
    }
  }
  if (ty >= kq && ty < kq + 4) {  // owners of rows K0..K0+3: W rows left of the block
#pragma unroll
    for (int v = 0; v <= KV; ++v)
      if (v < KV || tx < kq) sh.U[tx + 16 * v][ty - kq] = R[KV][v];
  }
  __syncthreads();
  DH_LEAF_TS(1);
  // 4x4 pivot block: Cholesky and inverse (every thread)
  double P[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a) dh_ld4(&sh.Pb[a][0], P[a]);
  double Lb[4][4], Vb[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) Lb[a][b] = Vb[a][b] = 0.0;
  int bad = 4;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double piv = P[k][k];
#pragma unroll
    for (int m = 0; m < k; ++m) piv -= Lb[k][m] * Lb[k][m];
    const double r = dh_rsq_nr(piv);
    bad = (bad == 4 && !(piv > 0.0)) ? k : bad;
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
  if (tid == 0 && bad < 4) atomicMin(info, 1 + row_base + K0 + bad);
#pragma unroll
  for (int a = 1; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < a; ++b) {
      double x = 0.0;
#pragma unroll
      for (int m = b; m < a; ++m) x += Lb[a][m] * Vb[m][b];
      Vb[a][b] = -Vb[a][a] * x;
    }
  DH_LEAF_TS(2);
  if (tid < 16 * NB) {  // one index j per thread
    const int j = tid;
    double c[4], g[4];
    dh_ld4(&sh.U[j][0], c);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double x = 0.0;
#pragma unroll
      for (int m = 0; m <= b; ++m) x = fma(c[m], Vb[b][m], x);
      g[b] = x;
    }
    dh_st4(&sh.G[j][0], g);
    if (j >= K0 + 4) {
#pragma unroll
      for (int b = 0; b < 4; ++b) sh.Ls[j][K0 + b] = g[b];
    }
    if (tid == 0) {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b <= a; ++b) sh.Ls[K0 + a][K0 + b] = Lb[a][b];
    }
  }
  __syncthreads();
  DH_LEAF_TS(3);
  double gj[NB][4];
#pragma unroll
  for (int v = 0; v < NB; ++v) dh_ld4(&sh.G[tx + 16 * v][0], gj[v]);
  DH_LEAF_TS(4);
  // rank-4 updates of the rows below the block (lower and diagonal 16x16 sub-blocks only)
#pragma unroll
  for (int u = KV; u < NB; ++u) {
    const bool below = (u > KV) || (ty >= kq + 4);
    if (below) {
      double li[4];
      dh_ld4(&sh.G[ty + 16 * u][0], li);
#pragma unroll
      for (int v = 0; v <= u; ++v)
#pragma unroll
        for (int b = 0; b < 4; ++b) R[u][v] = fma(-li[b], gj[v][b], R[u][v]);
    }
  }
  // rows of the block are final: row K0+a of inv(L)
  if (ty >= kq && ty < kq + 4) {
    const int a = ty - kq;
#pragma unroll
    for (int v = 0; v <= KV; ++v) R[KV][v] = sh.G[tx + 16 * v][a];
  }
}

template <int NB, int KV>
__device__ __forceinline__ void dh_leaf_blocks(double R[NB][NB], LeafShmT<NB>& sh, int ty, int tx, int tid,
                                               int* info, int row_base) {
#pragma unroll 1
  for (int c4 = 0; c4 < 4; ++c4) dh_leaf_block<NB, KV>(R, sh, ty, tx, tid, info, row_base, c4);
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
  dh_leaf_blocks<4, 0>(R, sh, ty, tx, tid, info, row_base);
  dh_leaf_blocks<4, 1>(R, sh, ty, tx, tid, info, row_base);
  dh_leaf_blocks<4, 2>(R, sh, ty, tx, tid, info, row_base);
  dh_leaf_blocks<4, 3>(R, sh, ty, tx, tid, info, row_base);
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

// ---------------------------------------------------------------------------
// 64x64 Cholesky + inverse from LDS: blocked 32x32 leaves on the two diagonal
// halves, the off-diagonal half of L and of inv(L) and the Schur update as
// 32-deep MFMA products, all in LDS.
//   L21 = A21 inv(L11)^T,  A22 -= L21 L21^T,  L22 = chol(A22),
//   inv(L)21 = -inv(L22) (L21 inv(L11)).
// ---------------------------------------------------------------------------
#define DH_LDS_LD 66  // leading dimension of the LDS tiles (16-byte aligned rows)
__shared__ __align__(16) double dh_lS[64 * DH_LDS_LD];  // the tile; L on exit (lower)
__shared__ __align__(16) double dh_lV[64 * DH_LDS_LD];  // inv(L) on exit (lower); scratch above

// One 16x16 block per wave: acc[r] (lane l) = sum_{k<32} A[i][k] B(k, j) at
// i = 4r + l/16, j = l%16, with A row-major (ld DH_LDS_LD) and B(k, j) =
// B[j][k] (TB) or B[k][j].
template <bool TB>
__device__ __forceinline__ dh_d4 dh_lds_mm16(const double* A, const double* B) {
  const int l = threadIdx.x & 63;
  const int i = l & 15, kk = l >> 4;
  dh_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int k = 4 * s + kk;
    const double a = A[i * DH_LDS_LD + k];
    const double b = TB ? B[i * DH_LDS_LD + k] : B[k * DH_LDS_LD + i];
    acc = dh_mfma(a, b, acc);
  }
  return acc;
}

// 32x32 blocked leaf on S (ld DH_LDS_LD): L into S's lower triangle, inv(L)
// into V (zeros above the diagonal).
__device__ __forceinline__ void dh_leaf32_lds(double* S, double* V, LeafShmT<2>& sh, int* info, int row_base) {
  const int tid = threadIdx.x;
  const int ty = tid >> 4, tx = tid & 15;
  double R[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const int i = ty + 16 * u, j = tx + 16 * v;
      R[u][v] = (j <= i) ? S[i * DH_LDS_LD + j] : 0.0;
    }
  dh_leaf_blocks<2, 0>(R, sh, ty, tx, tid, info, row_base);
  dh_leaf_blocks<2, 1>(R, sh, ty, tx, tid, info, row_base);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const int i = ty + 16 * u, j = tx + 16 * v;
      if (j <= i) S[i * DH_LDS_LD + j] = sh.Ls[i][j];
      V[i * DH_LDS_LD + j] = (j <= i) ? R[u][v] : 0.0;
    }
}

// Factor the 64x64 tile held in dh_lS (lower part): on exit dh_lS holds L in
// its lower triangle and dh_lV the dense inv(L) (lower triangle); the stores
// below write them out.  Same error convention as dh_leaf_potrf.
__device__ void dh_leaf64_core(int* info, int row_base) {
  __shared__ LeafShmT<2> sh;
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int bi = w >> 1, bj = w & 1;
  const int orow = 4 * 0 + (l >> 4), ocol = l & 15;  // output row (plus 4r) and column of a 16x16 block
  double* S = dh_lS;
  double* V = dh_lV;
  dh_leaf32_lds(S, V, sh, info, row_base);
  __syncthreads();
  // L21 = A21 inv(L11)^T, block (bi, bj) of 16x16 per wave
  dh_d4 x = dh_lds_mm16<true>(S + (32 + 16 * bi) * DH_LDS_LD, V + (16 * bj) * DH_LDS_LD);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) S[(32 + 16 * bi + 4 * r + orow) * DH_LDS_LD + 16 * bj + ocol] = x[r];
  __syncthreads();
  // A22 -= L21 L21^T
  x = dh_lds_mm16<true>(S + (32 + 16 * bi) * DH_LDS_LD, S + (32 + 16 * bj) * DH_LDS_LD);
#pragma unroll
  for (int r = 0; r < 4; ++r) S[(32 + 16 * bi + 4 * r + orow) * DH_LDS_LD + 32 + 16 * bj + ocol] -= x[r];
  __syncthreads();
  dh_leaf32_lds(S + 32 * DH_LDS_LD + 32, V + 32 * DH_LDS_LD + 32, sh, info, row_base + 32);
  __syncthreads();
  // T = L21 inv(L11), staged in V's upper-right block
  x = dh_lds_mm16<false>(S + (32 + 16 * bi) * DH_LDS_LD, V + 16 * bj);
#pragma unroll
  for (int r = 0; r < 4; ++r) V[(16 * bi + 4 * r + orow) * DH_LDS_LD + 32 + 16 * bj + ocol] = x[r];
  __syncthreads();
  // inv(L)21 = -inv(L22) T
  x = dh_lds_mm16<false>(V + (32 + 16 * bi) * DH_LDS_LD + 32, V + 32 + 16 * bj);
#pragma unroll
  for (int r = 0; r < 4; ++r) V[(32 + 16 * bi + 4 * r + orow) * DH_LDS_LD + 16 * bj + ocol] = -x[r];
  __syncthreads();
}

// Dense inv(L) of the tile factored by dh_leaf64_core (zeros above the diagonal, ld 64).
__device__ __forceinline__ void dh_leaf64_store_inv(double* inv_out) {
  const double* V = dh_lV;
#pragma unroll 2
  for (int e0 = 0; e0 < 64 * 64; e0 += 512) {  // two entries per thread, uniform trip count
    const int e = e0 + 2 * threadIdx.x;
    const int i = e >> 6, j = e & 63;
    double2 v;
    v.x = (j <= i) ? V[i * DH_LDS_LD + j] : 0.0;
    v.y = (j + 1 <= i) ? V[i * DH_LDS_LD + j + 1] : 0.0;
    *reinterpret_cast<double2*>(inv_out + e) = v;
  }
}

// L (lower triangle), inv(L)^T (strict upper triangle) and dinv of the tile factored by
// dh_leaf64_core.
__device__ __forceinline__ void dh_leaf64_store_tile(double* tile, long ld, double* dinv) {
  const double* S = dh_lS;
  const double* V = dh_lV;
#pragma unroll 2
  for (int e0 = 0; e0 < 64 * 64; e0 += 512) {
    const int e = e0 + 2 * threadIdx.x;
    const int i = e >> 6, j = e & 63;
    double2 t;
    t.x = (j <= i) ? S[i * DH_LDS_LD + j] : V[j * DH_LDS_LD + i];  // inv(L)^T above the diagonal
    t.y = (j + 1 <= i) ? S[i * DH_LDS_LD + j + 1] : V[(j + 1) * DH_LDS_LD + i];
    *reinterpret_cast<double2*>(tile + (long)i * ld + j) = t;
    if (j == i) dinv[i] = V[i * DH_LDS_LD + i];
    if (j + 1 == i) dinv[i] = V[i * DH_LDS_LD + i];
  }
}
