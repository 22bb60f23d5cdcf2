// Moment reduction and Maxwellian tables of the q1-line layout [q2][p][q1]
// (moments.py:5-22, collision_operator.py:52-65 and 102-116).
#pragma once
#include "fvm_math.h"

namespace bx {

// Moments of f in the q1-line layout: workgroup (q2 row, 64-wide q1 block,
// p chunk); wave w sums a contiguous quarter of the chunk's p range for its
// 64 q1 cells (loads issued eight at a time), and the four waves are added in
// a fixed order -> part[chunk][5][Q].
template <typename T>
__global__ __launch_bounds__(256) void mline_kernel(const T* f, int n1, int n2, int P, int nchunk,
                                                    int n_p1, int n_p2, double p1s, double dp1,
                                                    double p2s, double dp2, double p3s, double dp3,
                                                    double* part) {
  __shared__ double red[5][4][64];
  // the wave index is uniform across the wave: keep it (and all that derives
  // from it: p-cell, velocities, row pointers) in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nblk = (n1 + 63) / 64;
  int b = blockIdx.x;
  const int blk = b % nblk;
  b /= nblk;
  const int i2 = b % n2;
  const int chunk = b / n2;
  const int i1 = blk * 64 + lane;
  const bool live = i1 < n1;
  const int per = (P + 4 * nchunk - 1) / (4 * nchunk);
  const int pa = min(P, (chunk * 4 + wave) * per), pe = min(P, pa + per);
  const T* col = f + (long)i2 * P * n1 + (live ? i1 : 0);
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
  int j1 = pa % n_p1, j2 = (pa / n_p1) % n_p2, j3 = pa / (n_p1 * n_p2);
  double v2 = centre(p2s, j2, dp2), v3 = centre(p3s, j3, dp3);
  for (int p = pa; p < pe; p += 8) {
    double x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = (p + k < pe && live) ? double(col[(long)(p + k) * n1]) : 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (p + k < pe) {
        const double v1 = centre(p1s, j1, dp1);
        const double e = 0.5 * ((v1 * v1 + v2 * v2) + v3 * v3);
        a0 += x[k];
        a1 += x[k] * v1;
        a2 += x[k] * v2;
        a3 += x[k] * v3;
        a4 += x[k] * e;
        if (++j1 == n_p1) {
          j1 = 0;
          if (++j2 == n_p2) {
            j2 = 0;
            ++j3;
            v3 = centre(p3s, j3, dp3);
          }
          v2 = centre(p2s, j2, dp2);
        }
      }
    }
  }
  red[0][wave][lane] = a0;
  red[1][wave][lane] = a1;
  red[2][wave][lane] = a2;
  red[3][wave][lane] = a3;
  red[4][wave][lane] = a4;
  __syncthreads();
  if (wave == 0 && live) {
    const long Q = (long)n1 * n2;
    const long qc = (long)i2 * n1 + i1;
    for (int k = 0; k < 5; ++k)
      part[((long)chunk * 5 + k) * Q + qc] =
          ((red[k][0][lane] + red[k][1][lane]) + red[k][2][lane]) + red[k][3][lane];
  }
}

// Moments of f in the q1-line layout when a q1 line is a whole number of
// 16-byte vectors and 256 vectors cover a whole number of lines (n1 = 64,
// 128 or 256).  The (q2 row, p chunk) span of f is contiguous, so a workgroup
// streams it in 16-byte vectors: every thread keeps the same VEC q1 cells for
// the whole span and steps D = 256 VEC / n1 p-cells at a time, four loads in
// flight.  Threads holding the same q1 cells are added in a fixed order ->
// part[chunk][5][Q].
template <typename T>
__global__ __launch_bounds__(256) void mline_vec_kernel(const T* f, int n1, int n2, int P,
                                                        int nchunk, int n_p1, int n_p2,
                                                        double p1s, double dp1, double p2s,
                                                        double dp2, double p3s, double dp3,
                                                        double* part) {
  constexpr int VEC = 16 / sizeof(T);
  typedef T V __attribute__((ext_vector_type(VEC)));
  __shared__ double red[5][VEC][256];
  const int tid = threadIdx.x;
  const int i2 = blockIdx.x % n2, chunk = blockIdx.x / n2;
  const int lpl = n1 / VEC;  // threads per q1 line
  const int D = 256 / lpl;   // p-cells per workgroup step
  const int q1 = (tid % lpl) * VEC;
  const int per = (P + nchunk - 1) / nchunk;
  const int pa = min(P, chunk * per), pe = min(P, pa + per);
  const T* base = f + (long)i2 * P * n1 + q1;
  double acc[5][VEC];
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int c = 0; c < VEC; ++c) acc[k][c] = 0.0;
  int p = pa + tid / lpl;
  int j1 = p % n_p1, j2 = (p / n_p1) % n_p2, j3 = p / (n_p1 * n_p2);
  while (p < pe) {
    V x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int pu = p + u * D;
      x[u] = pu < pe ? __builtin_nontemporal_load((const V*)(base + (long)pu * n1)) : V(0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (p + u * D < pe) {
        const double v1 = centre(p1s, j1, dp1), v2 = centre(p2s, j2, dp2);
        const double v3 = centre(p3s, j3, dp3);
        const double e = 0.5 * ((v1 * v1 + v2 * v2) + v3 * v3);
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
          const double xv = double(x[u][c]);
          acc[0][c] += xv;
          acc[1][c] += xv * v1;
          acc[2][c] += xv * v2;
          acc[3][c] += xv * v3;
          acc[4][c] += xv * e;
        }
      }
      j1 += D;
      while (j1 >= n_p1) {
        j1 -= n_p1;
        if (++j2 == n_p2) {
          j2 = 0;
          ++j3;
        }
      }
    }
    p += 4 * D;
  }
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int c = 0; c < VEC; ++c) red[k][c][tid] = acc[k][c];
  __syncthreads();
  if (tid < n1) {
    const int l = tid / VEC, c = tid % VEC;
    const long Q = (long)n1 * n2;
    const long qc = (long)i2 * n1 + tid;
    for (int k = 0; k < 5; ++k) {
      double sum = 0.0;
      for (int g = 0; g < D; ++g) sum += red[k][c][g * lpl + l];
      part[((long)chunk * 5 + k) * Q + qc] = sum;
    }
  }
}

// Whether mline_vec_kernel takes lines of n1 cells of T.
template <typename T>
__host__ inline bool mline_vec_applies(int n1) {
  constexpr int VEC = 16 / sizeof(T);
  if (n1 % VEC != 0) return false;
  const int lpl = n1 / VEC;
  return lpl <= 256 && 256 % lpl == 0 && n1 <= 256;
}

// Sum the chunk partials -> moment block (times dp3*dp2*dp1).
__global__ void ml_finalize_kernel(const double* part, int nchunk, int Q, double measure,
                                   double* mom, double mass, double kB, int p_dim) {
  const int qc = blockIdx.x * blockDim.x + threadIdx.x;
  if (qc >= Q) return;
  double m[5];
  sum_moment_partials(part + qc, nchunk, Q, m);
  for (int k = 0; k < 5; ++k) m[k] *= measure;
  write_moment_block(mom, Q, qc, m, mass, kB, p_dim);
}

// Maxwellian tables of every q-cell from its moment block
// (collision_operator.py:52-65): tab[j1][q] = prefactor * exp(-m (v1-u1)^2/2kT),
// tab[n_p1 + j2][q] = exp(-m (v2-u2)^2/2kT), tab[n_p1 + n_p2 + j3][q] likewise.
// Grid: x over q-cells, y = table row j (no 64-bit index arithmetic).
template <typename T>
__global__ void bgk_table_kernel(const double* mom, int Q, int n_p1, int n_p2, int n_p3,
                                 double p1s, double dp1, double p2s, double dp2, double p3s,
                                 double dp3, double mass, double kB, T* tab) {
  const int qc = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;  // uniform over the workgroup
  if (qc >= Q) return;
  const double* b = mom + MOM_BGK * Q + qc;
  const double Tq = b[4 * Q];
  double val;
  if (j < n_p1) val = b[0] * maxwell_factor(centre(p1s, j, dp1), b[Q], Tq, mass, kB);
  else if (j < n_p1 + n_p2)
    val = maxwell_factor(centre(p2s, j - n_p1, dp2), b[2 * Q], Tq, mass, kB);
  else val = maxwell_factor(centre(p3s, j - n_p1 - n_p2, dp3), b[3 * Q], Tq, mass, kB);
  tab[(long)j * Q + qc] = T(val);
}

}  // namespace bx
