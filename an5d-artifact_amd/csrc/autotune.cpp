// Autotuner: placeholder lookup until the search lands.
#include "internal.hpp"

namespace an5d {
bool tuned_lookup(const an5d_problem*, TunedChoice*) { return false; }
}  // namespace an5d
