// Batched entry points (filled in by the persistent work-queue kernel).
#include "internal.hpp"
