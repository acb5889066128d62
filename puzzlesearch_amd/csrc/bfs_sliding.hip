// Instantiates the device BFS kernels for the sliding puzzle family.
#include "bfs_launch.h"

namespace ps {
PS_DEFINE_BFS_FAMILY(sliding, FAM_SLIDING)
}  // namespace ps
