// Abstract k-d tree interface, source-compatible with the reference's
// FKDTree<TYPE, dim> (FKDTree.h:6-52) and its free validator
// test_correct_search (FKDTree.h:54-86).  Backends derive from it; the GPU
// one is FKDTree_HIP (FKDTree_hip.h).
//
// The tree is the flat heap-ordered left-complete k-d tree of the reference:
// node i has sons 2i+1 and 2i+2 and splits on dimension depth % dim.
#ifndef FKDTREE_AMD_FKDTREE_H_
#define FKDTREE_AMD_FKDTREE_H_

#include <algorithm>
#include <iostream>
#include <vector>

#include "FKDPoint.h"

template <class TYPE, unsigned int numberOfDimensions>
class FKDTree {
 public:
  FKDTree(const std::vector<FKDPoint<TYPE, numberOfDimensions>>& /*points*/) {}
  FKDTree() {}
  virtual ~FKDTree() {}

  // Ids of the points inside the closed box [minPoint, maxPoint].
  virtual std::vector<unsigned int> search_in_the_box(const FKDPoint<TYPE, numberOfDimensions>& minPoint,
                                                      const FKDPoint<TYPE, numberOfDimensions>& maxPoint) = 0;

  // One result per box; the reference's only implementation (the OpenCL
  // backend, FKDTree_opencl.h:123-164) returns the number of points per box.
  virtual std::vector<unsigned int> search_in_the_box_multiple(
      const std::vector<FKDPoint<TYPE, numberOfDimensions>>& minPoints,
      const std::vector<FKDPoint<TYPE, numberOfDimensions>>& maxPoints) = 0;

  virtual void build() = 0;
  virtual std::vector<TYPE> const& getDimensionVector(const int dimension) = 0;
  virtual std::vector<unsigned int> const& getIdVector() = 0;

  // Split-order check of the built tree: every node is >= its left son and
  // <= its right son on its split dimension, recursively.  (The reference
  // evaluates the same comparisons but discards the recursive results,
  // FKDTree.h:35,44; this version combines them.)
  bool test_correct_build(unsigned int index = 0, int dimension = 0) {
    const std::vector<TYPE>& split = getDimensionVector(dimension);
    const unsigned int n = static_cast<unsigned int>(getDimensionVector(0).size());
    const unsigned int left = 2 * index + 1, right = left + 1;
    const int next = (dimension + 1) % static_cast<int>(numberOfDimensions);
    bool ok = true;
    if (left < n) ok = ok && split[index] >= split[left] && test_correct_build(left, next);
    if (right < n) ok = ok && split[index] <= split[right] && test_correct_build(right, next);
    return ok;
  }
};

// Brute-force check of one search result: every point is reported iff it
// lies in the closed box (membership as in FKDTree.h:54-86).
template <typename TYPE, int numberOfDimensions>
bool test_correct_search(const std::vector<FKDPoint<TYPE, numberOfDimensions>> points,
                         const std::vector<unsigned int> foundPoints,
                         const FKDPoint<TYPE, numberOfDimensions>& minPoint,
                         const FKDPoint<TYPE, numberOfDimensions>& maxPoint) {
  std::vector<unsigned int> found(foundPoints);
  std::sort(found.begin(), found.end());
  bool good = true;
  for (const auto& p : points) {
    bool inside = true;
    for (int d = 0; d < numberOfDimensions; ++d) inside = inside && p[d] <= maxPoint[d] && p[d] >= minPoint[d];
    const bool reported = std::binary_search(found.begin(), found.end(), p.getId());
    if (reported != inside) {
      std::cout << "Point " << p.getId()
                << (reported ? " was wrongly found to be in the box." : " was wrongly found to be outside the box.")
                << std::endl;
      good = false;
    }
  }
  return good;
}

#endif  // FKDTREE_AMD_FKDTREE_H_
