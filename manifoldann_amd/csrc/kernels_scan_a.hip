// Instantiations of the first-pass scan for tile configurations 0, 7, 8, 9
// (one translation unit per group of configurations, so that the library's
// build compiles them in parallel).
#include "scan_impl.h"

namespace mann {

bool launch_scan_group_a(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st) {
  switch (a.cfg) {
    case 0: launch_cfg<2, 2, 2, 2, 3>(a, mode, grid, kind, st); return true;
    case 7: launch_cfg<2, 2, 2, 2, 4>(a, mode, grid, kind, st); return true;
    case 8: launch_cfg<2, 1, 2, 2, 3>(a, mode, grid, kind, st); return true;
    case 9: launch_cfg<2, 1, 2, 1, 3>(a, mode, grid, kind, st); return true;
    default: return false;
  }
}

}  // namespace mann
