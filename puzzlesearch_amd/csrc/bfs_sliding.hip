// Instantiates the device BFS and segment-step kernels for the sliding puzzle family.
#include "bfs_launch.h"
#include "fs_launch.h"
#include "seg3_launch.h"

namespace ps {
PS_DEFINE_BFS_FAMILY(sliding, FAM_SLIDING)
PS_DEFINE_FS_FAMILY(sliding, FAM_SLIDING)
PS_DEFINE_S3_FAMILY(sliding, FAM_SLIDING)
}  // namespace ps
