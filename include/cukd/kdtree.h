// cukd/kdtree.h -- everything: builders, queries, spatial k-d tree.
#pragma once

#include "cukd/builder.h"
#include "cukd/knn.h"
#include "cukd/spatial-kdtree.h"
