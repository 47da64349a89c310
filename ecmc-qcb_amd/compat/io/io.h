// io.h -- the parts of src/io/io.h on the flow-measurement path.
#pragma once
#include <string>
#include <vector>

namespace io {
// fixed-point to_string (io.cpp:529-533)
std::string format_double(double val, int precision);
// append "t Q E" lines to dirpath/filename/filename_topo.txt (io.cpp:109-136)
void save_topo(const std::vector<double>& tQE, const std::string& filename,
               const std::string& dirpath, int precision);
}  // namespace io
