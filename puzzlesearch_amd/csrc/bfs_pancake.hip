// Instantiates the device BFS kernels for the pancake puzzle family.
#include "bfs_launch.h"

namespace ps {
PS_DEFINE_BFS_FAMILY(pancake, FAM_PANCAKE)
}  // namespace ps
