// cukd/traverse-stack-free.h -- the reference header of the same name: the
// stack-free traversal (traverse-stack-free.h:24-117): traverse_stack_free.
// All three traversals are implemented in cukd/traverse.h.
#pragma once

#include "cukd/traverse.h"
