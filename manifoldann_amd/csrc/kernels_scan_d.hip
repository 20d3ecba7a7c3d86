// Instantiations of the first-pass scan for tile configurations 5, 6
// (one translation unit per group of configurations, so that the library's
// build compiles them in parallel).
#include "scan_impl.h"

namespace mann {

bool launch_scan_group_d(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st) {
  switch (a.cfg) {
    case 5: launch_cfg<2, 4, 4, 2, 3>(a, mode, grid, kind, st); return true;
    case 6: launch_cfg<2, 4, 4, 2, 4>(a, mode, grid, kind, st); return true;
    default: return false;
  }
}

}  // namespace mann
