// brdc_tree.h -- merge-tree geometry shared by host and device code.
//
// The tree is the one DLAED0_BR builds (dlaed0_br.f:76-93): a node of size m
// splits into a left child of floor(m/2) rows and a right child of
// ceil(m/2) rows, and every node is split until the rightmost (largest) node
// of a depth fits in a leaf.  All leaves therefore sit at the same depth, and
// node (depth, index) is found by walking the bits of index from the root, so
// no plan array is needed on the device.
#pragma once

#include <hip/hip_runtime.h>

namespace brdc {

__host__ __device__ inline void node_geom(int n, int depth, int idx, int* start, int* size) {
    int s = 0, sz = n;
    for (int b = depth - 1; b >= 0; b--) {
        const int left = sz / 2;
        if ((idx >> b) & 1) {
            s += left;
            sz -= left;
        } else {
            sz = left;
        }
    }
    *start = s;
    *size = sz;
}

// Number of halvings until ceil(n / 2^levels) <= leaf_max.
__host__ __device__ inline int tree_levels(int n, int leaf_max) {
    int levels = 0;
    while (((n + (1 << levels) - 1) >> levels) > leaf_max) levels++;
    return levels;
}

// Largest node size at a depth: ceil(n / 2^depth).
__host__ __device__ inline int max_node_size(int n, int depth) {
    return (n + (1 << depth) - 1) >> depth;
}

}  // namespace brdc
