// Legacy-VTK writer and results-file naming (Model::saveVTK).
#pragma once

#include <string>

#include "../../../include/efem_gpu.h"

namespace efem {

// points X + u (u over the free DOFs in node order, or null), cells as the
// element node lists; false with msg on error
bool writeVtk(const efem_problem& pb, const double* u, const char* path, std::string& msg);

// <workingDir>/results/<date>/<date-time-ms>-<name>.vtk, directories created
bool resultsPath(const char* workingDir, const char* name, std::string& out, std::string& msg);

}  // namespace efem
