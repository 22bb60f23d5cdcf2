// FP64 specialisations of the q1-line stage kernel.
#define BX_QL_INSTANTIATE
#include "ql_launch.h"

namespace bx {
template void ql_launch<double>(const QLParams&, int, int, int, int, hipStream_t);
template int ql_blocks_per_cu<double>(int, int, int, int);
}
