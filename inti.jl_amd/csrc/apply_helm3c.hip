// Direct-sum kernels of the 3D Helmholtz operators with a complex
// wavenumber (kHelmCK): FP64, one direction, one, four or sixteen
// right-hand sides.
#include "apply_kernel.hpp"

namespace inti {
ApplyFn apply_pick_helm3c(int kind, int nr) {
    if (nr == 1) return apply_kind<kHelmCK, 3, false, 1>(kind);
    if (nr == 4) return apply_kind<kHelmCK, 3, false, 4>(kind);
    if (nr == 16) return apply_kind<kHelmCK, 3, false, 16>(kind);
    return nullptr;
}
}  // namespace inti
