// Launch of the q1-line stage kernel, compiled once per storage type in its
// own translation unit (ql_f64.hip, ql_f32.hip) so the many specialisations
// of qline_kernel build in parallel.
#pragma once
#include "qline.h"

namespace bx {

// Launches qline_kernel over nb workgroups for reconstruction ``recon``,
// stage form ``mode`` (enum QlMode) and ``pd`` Maxwellian factors per cell
// (0 = no source).  WENO5 has a kernel per (mode, pd); the other
// reconstructions use the run-time form.
template <typename T>
void ql_launch(const QLParams& q, int recon, int nb, int mode, int pd, hipStream_t s);

// Resident workgroups per CU of the kernel ql_launch picks for these
// arguments (0 if the runtime cannot say).
template <typename T>
int ql_blocks_per_cu(int n1, int recon, int mode, int pd);

#ifdef BX_QL_INSTANTIATE
using QlKernel = void (*)(QLParams);

template <typename T, int NR, int M>
static QlKernel ql_weno_pd(int pd) {
  switch (pd) {
    case 0: return qline_kernel<T, RC_WENO5, NR, M, 0>;
    case 1: return qline_kernel<T, RC_WENO5, NR, M, 1>;
    case 2: return qline_kernel<T, RC_WENO5, NR, M, 2>;
    default: return qline_kernel<T, RC_WENO5, NR, M, 3>;
  }
}

template <typename T, int NR>
static QlKernel ql_nr(int recon, int mode, int pd) {
  switch (recon) {
    case RC_PC: return qline_kernel<T, RC_PC, NR, QL_GENERAL, -1>;
    case RC_MINMOD: return qline_kernel<T, RC_MINMOD, NR, QL_GENERAL, -1>;
    case RC_PPM: return qline_kernel<T, RC_PPM, NR, QL_GENERAL, -1>;
    default: break;
  }
  switch (mode) {
    case QL_WU_PLAIN: return ql_weno_pd<T, NR, QL_WU_PLAIN>(pd);
    case QL_WU_COMB: return ql_weno_pd<T, NR, QL_WU_COMB>(pd);
    case QL_W_PLAIN: return ql_weno_pd<T, NR, QL_W_PLAIN>(pd);
    default: return qline_kernel<T, RC_WENO5, NR, QL_GENERAL, -1>;
  }
}

template <typename T>
static QlKernel ql_pick(int n1, int recon, int mode, int pd) {
  if (n1 <= 64) return ql_nr<T, 1>(recon, mode, pd);
  if (n1 <= 128) return ql_nr<T, 2>(recon, mode, pd);
  return ql_nr<T, 4>(recon, mode, pd);
}

template <typename T>
void ql_launch(const QLParams& q, int recon, int nb, int mode, int pd, hipStream_t s) {
  hipLaunchKernelGGL(ql_pick<T>(q.n1, recon, mode, pd), dim3(nb), dim3(256), 0, s, q);
}

template <typename T>
int ql_blocks_per_cu(int n1, int recon, int mode, int pd) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ql_pick<T>(n1, recon, mode, pd), 256, 0) !=
      hipSuccess)
    return 0;
  return n;
}
#endif

}  // namespace bx
