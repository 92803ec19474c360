// Host-side mesh set-up for the device-resident SIMPLE engine.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "cfd_solver.h"

namespace cfd {

// One (sub)mesh: the reference's Vert/Face/Cell/Bnd arrays
// (src/BasicType.h:22-67) as structure-of-arrays, plus the derived lists the
// device kernels gather through.
struct HostMesh {
    int nvrt = 0, ncel = 0, nfac = 0, nbnd = 0, ncb = 0;
    std::vector<double> vert;                 // [V][3]
    std::vector<int> cell_vert, cell_rid;     // [C][8], [C]
    std::vector<int> cell_nface, cell_face;   // [C], [C][6] slot order (-1 pad)
    std::vector<int> cell_cell;               // [C][6] neighbour per slot
    std::vector<int> face_vert;               // [F][4]
    std::vector<int> face_c1, face_c2, face_bnd;
    std::vector<int> bnd_vert, bnd_rid, bnd_face;
    std::vector<double> face_x, face_n, face_xpac, face_xnac;  // [F][3]
    std::vector<double> face_area, face_lambda, face_rlencos;
    std::vector<double> cell_x, cell_vol;     // [C][3], [C]
    std::vector<double> bnd_dist;
    std::vector<int> colour;                  // greedy first-fit over cells
    int ncolour = 0;
    // colour of each cell's block of kIluBlock consecutive cells for the
    // ILU(0) order (colour_cells), -1 everywhere when the mesh has none
    std::vector<int> bcolour;
    int nbcolour = 0;
};

// cells per block of the ILU(0) block-colour order (the 6 tetrahedra of a
// split hexahedron)
constexpr int kIluBlock = 6;

int load_raw(const cfd_raw_mesh &raw, HostMesh &m, std::string &err);
int build_faces(HostMesh &m, std::string &err);
int build_geometry(HostMesh &m, std::string &err);
int split_cells(const HostMesh &full, const std::vector<int> &cells,
                int coupled_id, HostMesh &out, std::string &err);
void colour_cells(HostMesh &m);
int concat_meshes(const std::vector<HostMesh> &parts, HostMesh &out, std::vector<int> off[4],
                  std::string &err);
int build_parts(const cfd_raw_mesh &raw, const cfd_params &p, HostMesh &fluid, HostMesh &solid,
                bool &has_solid, int &coupled_id, std::string &err);

}  // namespace cfd
