// Legacy-VTK output of a (deformed) configuration and the results-file
// naming of Model::saveVTK (src/model/Model.cpp:50-117).
//
// The reference hands a vtkUnstructuredGrid to vtkUnstructuredGridWriter:
// one point per node at position + dx stored as vtkPoints' default FP32,
// one cell per element with the element's node list unchanged and the cell
// type chosen by the node count (4 VTK_TETRA = 10, 10 VTK_QUADRATIC_TETRA =
// 24, 8 VTK_HEXAHEDRON = 12, 20 VTK_QUADRATIC_HEXAHEDRON = 25).  This writer
// emits the same grid as ASCII legacy VTK (DataFile version 4.2, the
// CELLS n size form every legacy reader accepts) without the VTK library:
// the points are the same FP32 values, printed with 9 significant digits so
// that reading them back as FP32 recovers them exactly.
#include "vtk.hpp"

#include <sys/stat.h>

#include <cerrno>
#include <chrono>
#include <cstdio>
#include <ctime>
#include <string>
#include <vector>

namespace efem {

namespace {

int cellType(int nen)
{
    switch (nen) {
    case 4: return 10;
    case 10: return 24;
    case 8: return 12;
    case 20: return 25;
    default: return -1;
    }
}

bool makeDirs(const std::string& dir, std::string& msg)
{
    // IO::ensureDir: every missing component is created
    for (size_t i = 1; i <= dir.size(); ++i) {
        if (i < dir.size() && dir[i] != '/') continue;
        const std::string part = dir.substr(0, i);
        if (part.empty()) continue;
        if (mkdir(part.c_str(), 0777) != 0 && errno != EEXIST) {
            msg = "cannot create directory " + part;
            return false;
        }
    }
    return true;
}

}  // namespace

bool writeVtk(const efem_problem& pb, const double* u, const char* path, std::string& msg)
{
    if (!path || !pb.X || (pb.n_elems > 0 && !pb.conn) || pb.n_nodes < 0 || pb.n_elems < 0) {
        msg = "efem_write_vtk: invalid problem or path";
        return false;
    }
    const int type = cellType(pb.nen);
    if (type < 0) {
        msg = "efem_write_vtk: nen must be 4, 8, 10 or 20";
        return false;
    }
    for (long long i = 0; i < static_cast<long long>(pb.n_elems) * pb.nen; ++i) {
        if (pb.conn[i] < 0 || pb.conn[i] >= pb.n_nodes) {
            msg = "efem_write_vtk: element node index out of range";
            return false;
        }
    }
    FILE* f = std::fopen(path, "w");
    if (!f) {
        msg = std::string("efem_write_vtk: cannot open ") + path;
        return false;
    }
    std::fprintf(f, "# vtk DataFile Version 4.2\nvtk output\nASCII\nDATASET UNSTRUCTURED_GRID\n");
    std::fprintf(f, "POINTS %d float\n", pb.n_nodes);
    // u is over the free DOFs, numbered in node order (fixed nodes keep X)
    long long fid = 0;
    for (int a = 0; a < pb.n_nodes; ++a) {
        const bool fixed = pb.fixed && pb.fixed[a];
        float p[3];
        for (int d = 0; d < 3; ++d) {
            double x = pb.X[3 * a + d];
            if (u && !fixed) x += u[3 * fid + d];
            p[d] = static_cast<float>(x);
        }
        if (!fixed) ++fid;
        std::fprintf(f, "%.9g %.9g %.9g\n", static_cast<double>(p[0]), static_cast<double>(p[1]),
                     static_cast<double>(p[2]));
    }
    std::fprintf(f, "\nCELLS %d %lld\n", pb.n_elems, static_cast<long long>(pb.n_elems) * (pb.nen + 1));
    for (int e = 0; e < pb.n_elems; ++e) {
        std::fprintf(f, "%d", pb.nen);
        for (int a = 0; a < pb.nen; ++a) std::fprintf(f, " %d", pb.conn[static_cast<size_t>(e) * pb.nen + a]);
        std::fputc('\n', f);
    }
    std::fprintf(f, "\nCELL_TYPES %d\n", pb.n_elems);
    for (int e = 0; e < pb.n_elems; ++e) std::fprintf(f, "%d\n", type);
    const bool ok = std::ferror(f) == 0;
    if (std::fclose(f) != 0 || !ok) {
        msg = std::string("efem_write_vtk: write failed for ") + path;
        return false;
    }
    return true;
}

bool resultsPath(const char* workingDir, const char* name, std::string& out, std::string& msg)
{
    if (!workingDir || !name) {
        msg = "efem_results_path: null argument";
        return false;
    }
    // <workingDir>results/YYYY-MM-DD/YYYY-MM-DD-HH-MM-SS-mmm-<name>.vtk in
    // local time; the working directory string ends with '/'
    // (Config::workingDir), one is added when missing
    std::string dir = workingDir;
    if (!dir.empty() && dir.back() != '/') dir += '/';
    const auto now = std::chrono::system_clock::now();
    const auto secs = std::chrono::duration_cast<std::chrono::seconds>(now.time_since_epoch());
    const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(now.time_since_epoch()) - secs;
    const std::time_t t = static_cast<std::time_t>(secs.count());
    std::tm tm{};
    localtime_r(&t, &tm);
    char day[16], stamp[32];
    std::strftime(day, sizeof(day), "%Y-%m-%d", &tm);
    std::strftime(stamp, sizeof(stamp), "%Y-%m-%d-%H-%M-%S", &tm);
    dir += "results/" + std::string(day) + "/";
    if (!makeDirs(dir, msg)) return false;
    char milli[8];
    std::snprintf(milli, sizeof(milli), "-%03d", static_cast<int>(ms.count()));
    out = dir + stamp + milli + "-" + name + ".vtk";
    return true;
}

}  // namespace efem
