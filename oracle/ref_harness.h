// Test-infrastructure only: the handle the reference harness
// (oracle/ref_harness.cpp) hands out for a compiled-reference tree, shared
// with the drop-in adapter harness (oracle/adapter_harness.cpp).
#pragma once
#include <memory>

#include "grid.hpp"

struct RefTree {
    std::unique_ptr<grid::amrtree> tree;
    char err[256] = {0};
};
