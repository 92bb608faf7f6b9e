// Direct-sum kernels instantiated for one (PDE, dimension): INTI_LAPLACE, 2D.
#include "apply_kernel.hpp"
#include "apply_sym.hpp"

namespace inti {
ApplyFn apply_pick_lap2(int kind, int mixed, int nr) { return apply_pick_pd<INTI_LAPLACE, 2>(kind, mixed, nr); }
BatchFn batch_pick_lap2(int kind, int mixed, int nr) { return batch_pick_pd<INTI_LAPLACE, 2>(kind, mixed, nr); }
SymFn sym_pick_lap2(int kind, int mixed, int G) { return sym_pick_pd<INTI_LAPLACE, 2>(kind, mixed, G); }
SymFn sym_trans_pick_lap2(int kind, int G) { return sym_trans_pick_pd<INTI_LAPLACE, 2>(kind, G); }
ApplyFn apply_trans_pick_lap2(int kind) { return apply_trans_pick_pd<INTI_LAPLACE, 2>(kind); }
}  // namespace inti
