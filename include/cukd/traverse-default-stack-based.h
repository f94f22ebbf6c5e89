// cukd/traverse-default-stack-based.h -- the reference header of the same name: the
// stack-based traversal (traverse-default-stack-based.h:24-108): traverse_default.
// All three traversals are implemented in cukd/traverse.h.
#pragma once

#include "cukd/traverse.h"
