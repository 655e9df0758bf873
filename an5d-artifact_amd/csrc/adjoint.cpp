// Adjoint entry points (filled in by the adjoint sweep).
#include "internal.hpp"
