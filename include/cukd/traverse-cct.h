// cukd/traverse-cct.h -- the reference header of the same name: the
// closest-corner-tracking traversal (traverse-cct.h:25-99): traverse_cct.
// All three traversals are implemented in cukd/traverse.h.
#pragma once

#include "cukd/traverse.h"
