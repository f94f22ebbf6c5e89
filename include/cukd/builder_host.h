// cukd/builder_host.h -- the reference's builder_host.h: buildTree_host is
// declared with the thrust builder (builder_thrust.h).
#pragma once

#include "cukd/builder_thrust.h"
