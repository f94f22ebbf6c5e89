// cukd/io.h -- the reference's point-file format and loader
// (cukd::common::loadPoints, common.h:245-266): a file is a size_t element
// count followed by the raw elements.
#pragma once

#include <hip/hip_runtime.h>

#include <fstream>
#include <stdexcept>
#include <string>

namespace cukd {

// Reads count and the elements of fileName into managed memory (as the
// reference does); throws std::runtime_error when the file is missing or
// shorter than its count says.
template <typename T>
inline T *loadPoints(const std::string &fileName, size_t &count)
{
  std::ifstream in(fileName, std::ios::binary);
  if (!in) throw std::runtime_error("cukd::loadPoints: cannot open " + fileName);
  count = 0;
  in.read((char *)&count, sizeof(count));
  if (!in) throw std::runtime_error("cukd::loadPoints: no count in " + fileName);
  T *points = nullptr;
  if (hipMallocManaged((void **)&points, (count ? count : 1) * sizeof(T)) != hipSuccess)
    throw std::runtime_error("cukd::loadPoints: out of memory for " + fileName);
  in.read((char *)points, (std::streamsize)(count * sizeof(T)));
  if ((size_t)in.gcount() != count * sizeof(T)) {
    (void)hipFree(points);
    throw std::runtime_error("cukd::loadPoints: " + fileName + " is shorter than its count");
  }
  return points;
}

template <typename T>
inline T *loadPoints(const std::string &fileName, int &count)
{
  size_t c = 0;
  T *p = loadPoints<T>(fileName, c);
  count = (int)c;
  return p;
}

// Writes count host-readable elements in the same format.
template <typename T>
inline void savePoints(const std::string &fileName, const T *points, size_t count)
{
  std::ofstream out(fileName, std::ios::binary);
  if (!out) throw std::runtime_error("cukd::savePoints: cannot create " + fileName);
  out.write((const char *)&count, sizeof(count));
  out.write((const char *)points, (std::streamsize)(count * sizeof(T)));
  if (!out) throw std::runtime_error("cukd::savePoints: write failed for " + fileName);
}

}  // namespace cukd
