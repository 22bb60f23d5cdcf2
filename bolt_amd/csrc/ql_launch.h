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

#ifdef BX_QL_INSTANTIATE
template <typename T, int R, int NR, int M, int PD>
static void ql_go(const QLParams& q, int nb, hipStream_t s) {
  hipLaunchKernelGGL((qline_kernel<T, R, NR, M, PD>), dim3(nb), dim3(256), 0, s, q);
}

template <typename T, int NR, int M>
static void ql_weno_pd(const QLParams& q, int nb, int pd, hipStream_t s) {
  switch (pd) {
    case 0: ql_go<T, RC_WENO5, NR, M, 0>(q, nb, s); break;
    case 1: ql_go<T, RC_WENO5, NR, M, 1>(q, nb, s); break;
    case 2: ql_go<T, RC_WENO5, NR, M, 2>(q, nb, s); break;
    default: ql_go<T, RC_WENO5, NR, M, 3>(q, nb, s);
  }
}

template <typename T, int NR>
static void ql_nr(const QLParams& q, int recon, int nb, int mode, int pd, hipStream_t s) {
  switch (recon) {
    case RC_PC: ql_go<T, RC_PC, NR, QL_GENERAL, -1>(q, nb, s); return;
    case RC_MINMOD: ql_go<T, RC_MINMOD, NR, QL_GENERAL, -1>(q, nb, s); return;
    case RC_PPM: ql_go<T, RC_PPM, NR, QL_GENERAL, -1>(q, nb, s); return;
    default: break;
  }
  switch (mode) {
    case QL_WU_PLAIN: ql_weno_pd<T, NR, QL_WU_PLAIN>(q, nb, pd, s); break;
    case QL_WU_COMB: ql_weno_pd<T, NR, QL_WU_COMB>(q, nb, pd, s); break;
    case QL_W_PLAIN: ql_weno_pd<T, NR, QL_W_PLAIN>(q, nb, pd, s); break;
    default: ql_go<T, RC_WENO5, NR, QL_GENERAL, -1>(q, nb, s);
  }
}

template <typename T>
void ql_launch(const QLParams& q, int recon, int nb, int mode, int pd, hipStream_t s) {
  if (q.n1 <= 64) ql_nr<T, 1>(q, recon, nb, mode, pd, s);
  else if (q.n1 <= 128) ql_nr<T, 2>(q, recon, nb, mode, pd, s);
  else ql_nr<T, 4>(q, recon, nb, mode, pd, s);
}
#endif

}  // namespace bx
