// Synthetic workloads: structured meshes with fixed faces and pressure
// faces, in the reference's node orders (VTK order, positively oriented),
// with a project-defined RNG so the inputs are reproducible bit for bit.
#pragma once
#include <cstdint>
#include <vector>

namespace efem {

// SplitMix64 -> 53-bit uniform in [0,1); Box-Muller normals
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal();
};

struct Mesh {
    std::vector<double> X;          // [n][3]
    std::vector<int> conn;          // [E][nen]
    int nen = 8;
    int reduced = 0;
    int law = 0;
    std::vector<double> params;     // [4] or [E][4]
    int params_per_elem = 0;
    std::vector<unsigned char> fixed;
    std::vector<int> faces;         // [F][fnn]
    int fnn = 4;
    double pressure = 0.0;
    int nNodes() const { return static_cast<int>(X.size() / 3); }
    int nElems() const { return static_cast<int>(conn.size() / nen); }
    int nFaces() const { return static_cast<int>(faces.size() / fnn); }
};

// box [0,L]^3 of nx*ny*nz cells
struct BoxSpec {
    int nx = 20, ny = 4, nz = 4;
    double lx = 0.10, ly = 0.01, lz = 0.01;
    int kind = 8;          // 8: C3D8 hexes, 4: Kuhn C3D4, 10: Kuhn C3D10, 20: C3D20
    double jitter = 0.0;   // interior-node jitter, fraction of the cell size
    uint64_t seed = 1;
};

Mesh makeBox(const BoxSpec& s);
// all nodes on the plane coord[axis] == value are fixed
void fixPlane(Mesh& m, int axis, double value);
// faces of the elements lying on the plane coord[axis] == value, optionally
// restricted to lo <= coord[raxis] <= hi for every face node
void pressurePlane(Mesh& m, int axis, double value, double p, int raxis = -1, double lo = 0,
                   double hi = 0);

// the workloads (see efem_mesh_build in efem_gpu.h)
Mesh makeWorkload(int which, const double* opts, int n_opts);

}  // namespace efem
