// Instantiates the device BFS and segment-step kernels for the pancake puzzle family.
#include "bfs_launch.h"
#include "fs_launch.h"

namespace ps {
PS_DEFINE_BFS_FAMILY(pancake, FAM_PANCAKE)
PS_DEFINE_FS_FAMILY(pancake, FAM_PANCAKE)
}  // namespace ps
