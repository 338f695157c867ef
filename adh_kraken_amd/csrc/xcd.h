/*
 * xcd.h -- workgroup order for gathers on the eight XCDs of gfx950.
 *
 * The dispatcher deals the workgroups of a launch to the XCDs round-robin (workgroup w runs on XCD
 * w % 8), and each XCD has its own L2.  Kernels whose neighbouring workgroups read the same element or
 * node data (the assembly gathers, the element kernels, the matrix preparation passes) take their
 * logical workgroup index from xcd_block(): XCD x then runs one contiguous range of the grid, the data
 * shared across a range boundary is fetched by two L2s instead of every L2, and neighbours find each
 * other's lines in their own L2.  The map is a permutation of [0, gridDim.x), so what a kernel
 * computes does not change.
 */
#ifndef ADHK_XCD_H
#define ADHK_XCD_H

#include <hip/hip_runtime.h>

#define ADHK_NXCD 8

__device__ __forceinline__ int xcd_block(void)
{
    const int G = gridDim.x, w = blockIdx.x;
    const int per = G / ADHK_NXCD, rem = G % ADHK_NXCD, x = w % ADHK_NXCD;
    /* XCD x holds per + (x < rem) of the workgroups; give it logical blocks [x per + min(x, rem), ...) */
    return x * per + min(x, rem) + w / ADHK_NXCD;
}

#endif
