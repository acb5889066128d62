// Instantiates the device BFS kernels for the hanoi puzzle family.
#include "bfs_launch.h"

namespace ps {
PS_DEFINE_BFS_FAMILY(hanoi, FAM_HANOI)
}  // namespace ps
