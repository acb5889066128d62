// Instantiates the device BFS and segment-step kernels for the hanoi puzzle family.
#include "bfs_launch.h"
#include "fs_launch.h"
#include "seg3_launch.h"

namespace ps {
PS_DEFINE_BFS_FAMILY(hanoi, FAM_HANOI)
PS_DEFINE_FS_FAMILY(hanoi, FAM_HANOI)
PS_DEFINE_S3_FAMILY(hanoi, FAM_HANOI)
}  // namespace ps
