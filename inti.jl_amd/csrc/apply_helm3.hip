// Direct-sum kernels instantiated for one (PDE, dimension): INTI_HELMHOLTZ, 3D.
#include "apply_kernel.hpp"
#include "apply_sym.hpp"

namespace inti {
ApplyFn apply_pick_helm3(int kind, int mixed, int nr) { return apply_pick_pd<INTI_HELMHOLTZ, 3>(kind, mixed, nr); }
BatchFn batch_pick_helm3(int kind, int mixed, int nr) { return batch_pick_pd<INTI_HELMHOLTZ, 3>(kind, mixed, nr); }
SymFn sym_pick_helm3(int kind, int mixed, int G) { return sym_pick_pd<INTI_HELMHOLTZ, 3>(kind, mixed, G); }
SymFn sym_trans_pick_helm3(int kind, int G) { return sym_trans_pick_pd<INTI_HELMHOLTZ, 3>(kind, G); }
ApplyFn apply_trans_pick_helm3(int kind) { return apply_trans_pick_pd<INTI_HELMHOLTZ, 3>(kind); }
}  // namespace inti
