// cukd/kdtree.h -- everything: builders, queries, spatial k-d tree,
// forests, multi-round lists and point files.
#pragma once

#include "cukd/builder.h"
#include "cukd/knn.h"
#include "cukd/spatial-kdtree.h"
#include "cukd/forest.h"
#include "cukd/io.h"
