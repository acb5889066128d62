// Instantiates the device BFS kernels for the packed puzzle family.
#include "bfs_launch.h"

namespace ps {
PS_DEFINE_BFS_FAMILY(packed, FAM_PACKED)
}  // namespace ps
