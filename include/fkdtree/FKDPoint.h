// Point record of the FKDTree interface: `dim` coordinates and an unsigned
// id.  Source-compatible with the reference's FKDPoint (FKDPoint.h:13-104)
// and its make_FKDPoint helpers (FKDPoint.h:107-147), so code written
// against that header compiles unchanged.  For TYPE = float the record is
// exactly 4 * (dim + 1) bytes (coordinates, then the id), the layout the
// GPU backend uploads as is.
#ifndef FKDTREE_AMD_FKDPOINT_H_
#define FKDTREE_AMD_FKDPOINT_H_

#include <array>
#include <iostream>

template <class TYPE, int numberOfDimensions>
class FKDPoint {
 public:
  FKDPoint() : theElements(), theId(0) {}
  FKDPoint(const FKDPoint& other) = default;
  FKDPoint& operator=(const FKDPoint& other) = default;

  FKDPoint(TYPE x, TYPE y, unsigned int id) : theElements(), theId(id) {
    set2(x, y);
  }
  FKDPoint(TYPE x, TYPE y, TYPE z, unsigned int id) : theElements(), theId(id) {
    set2(x, y);
    at(2) = z;
  }
  FKDPoint(TYPE x, TYPE y, TYPE z, TYPE w, unsigned int id) : theElements(), theId(id) {
    set2(x, y);
    at(2) = z;
    at(3) = w;
  }

  // As in the reference, the caller keeps i < numberOfDimensions.
  TYPE& operator[](int const i) { return theElements[i]; }
  TYPE const& operator[](int const i) const { return theElements[i]; }

  void setDimension(int i, const TYPE& value) { theElements[i] = value; }
  void setId(const unsigned int id) { theId = id; }
  unsigned int getId() const { return theId; }

  void print() const {
    std::cout << "point id: " << theId << std::endl;
    for (const auto& v : theElements) std::cout << v << std::endl;
  }

 private:
  TYPE& at(int i) { return theElements[i < numberOfDimensions ? i : 0]; }
  void set2(TYPE x, TYPE y) {
    at(0) = x;
    at(1) = y;
  }

  std::array<TYPE, numberOfDimensions> theElements;
  unsigned int theId;
};

template <typename TYPE>
FKDPoint<TYPE, 1> make_FKDPoint(TYPE x, unsigned int id) {
  FKDPoint<TYPE, 1> p;
  p.setDimension(0, x);
  p.setId(id);
  return p;
}

template <typename TYPE>
FKDPoint<TYPE, 2> make_FKDPoint(TYPE x, TYPE y, unsigned int id) {
  FKDPoint<TYPE, 2> p;
  p.setDimension(0, x);
  p.setDimension(1, y);
  p.setId(id);
  return p;
}

template <typename TYPE>
FKDPoint<TYPE, 3> make_FKDPoint(TYPE x, TYPE y, TYPE z, unsigned int id) {
  FKDPoint<TYPE, 3> p;
  p.setDimension(0, x);
  p.setDimension(1, y);
  p.setDimension(2, z);
  p.setId(id);
  return p;
}

template <typename TYPE>
FKDPoint<TYPE, 4> make_FKDPoint(TYPE x, TYPE y, TYPE z, TYPE w, unsigned int id) {
  FKDPoint<TYPE, 4> p;
  p.setDimension(0, x);
  p.setDimension(1, y);
  p.setDimension(2, z);
  p.setDimension(3, w);
  p.setId(id);
  return p;
}

#endif  // FKDTREE_AMD_FKDPOINT_H_
