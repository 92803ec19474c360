// ORACLE -- TEST INFRASTRUCTURE ONLY (see ref.h).
// Restates face construction, geometry and the conjugate-heat-transfer split.
#include <deque>
#include <cmath>
#include <cstdio>
#include <map>
#include <stdexcept>

#include "ref.h"

namespace oref {

void load_raw(const cfd_raw_mesh &m, Mesh &o) {
    o = Mesh();
    o.nvrt = m.nvrt;
    o.ncel = m.ncel;
    o.nbnd = m.nbnd;
    o.vert.resize(m.nvrt);
    for (int i = 0; i < m.nvrt; ++i)
        for (int k = 0; k < 3; ++k) o.vert[i][k] = m.vert[3 * i + k];
    o.cell.resize(m.ncel);
    for (int i = 0; i < m.ncel; ++i) {
        for (int k = 0; k < 8; ++k) o.cell[i].v[k] = m.cell_vert[8 * i + k];
        o.cell[i].rid = m.cell_rid[i];
    }
    o.bnd.resize(m.nbnd);
    for (int i = 0; i < m.nbnd; ++i) {
        for (int k = 0; k < 4; ++k) o.bnd[i].v[k] = m.bnd_vert[4 * i + k];
        o.bnd[i].rid = m.bnd_rid[i];
    }
}

// FindFace (src/Preprocessor.cpp:238-299): degenerate-slot rule, search of
// the first vertex's face list in insertion order, subset match.
static int find_face(Mesh &m, std::vector<std::vector<int>> &node_face, int ic,
                     int n1, int n2, int n3, int n4) {
    int rep = 0;
    if (n1 == n2 || n1 == n3 || n1 == n4) rep++;
    if (n2 == n3 || n2 == n4) rep++;
    if (n3 == n4) rep++;
    if (rep > 1) return 0;
    auto in_q = [&](int v) { return v == n1 || v == n2 || v == n3 || v == n4; };
    int fid = -100;
    for (int f : node_face[n1]) {
        const Face &F = m.face[f];
        if (in_q(F.v[0]) && in_q(F.v[1]) && in_q(F.v[2]) && in_q(F.v[3])) {
            fid = f;
            break;
        }
    }
    Cell &C = m.cell[ic];
    if (fid < 0) {
        Face F;
        F.v[0] = n1; F.v[1] = n2; F.v[2] = n3; F.v[3] = n4;
        F.c1 = ic;
        F.c2 = -10;  // the reference leaves this uninitialised (realloc)
        F.bnd = -10;
        int nf = (int)m.face.size();
        m.face.push_back(F);
        if (C.nface >= 6) return -1;
        C.face[C.nface++] = nf;
        for (int k = 0; k < 4; ++k) node_face[F.v[k]].push_back(nf);
    } else {
        Face &F = m.face[fid];
        if (F.bnd >= 0) {
            if (F.c1 == -10) F.c1 = ic;
            else if (F.c2 == -10) F.c2 = ic;
            else return -2;  // assert(false) in the reference
        } else {
            F.c2 = ic;
        }
        if (C.nface >= 6) return -1;
        C.face[C.nface++] = fid;
    }
    return 0;
}

// CreateFaces (src/Preprocessor.cpp:181-236)
int create_faces(Mesh &m) {
    std::vector<std::vector<int>> node_face(m.nvrt);
    for (auto &c : m.cell) c.nface = 0;
    m.face.clear();
    m.face.reserve((size_t)m.nbnd + 4 * (size_t)m.ncel);
    for (int i = 0; i < m.nbnd; ++i) {
        Face F;
        m.bnd[i].face = (int)m.face.size();
        for (int j = 0; j < 4; ++j) F.v[j] = m.bnd[i].v[j];
        F.bnd = i;
        F.c1 = -10;
        F.c2 = -10;
        for (int j = 0; j < 4; ++j) node_face[F.v[j]].push_back((int)m.face.size());
        m.face.push_back(F);
    }
    for (int i = 0; i < m.ncel; ++i) {
        const int *v = m.cell[i].v;
        int rc = 0;
        rc |= find_face(m, node_face, i, v[0], v[1], v[2], v[3]);
        rc |= find_face(m, node_face, i, v[4], v[5], v[6], v[7]);
        rc |= find_face(m, node_face, i, v[0], v[1], v[5], v[4]);
        rc |= find_face(m, node_face, i, v[1], v[2], v[6], v[5]);
        rc |= find_face(m, node_face, i, v[3], v[2], v[6], v[7]);
        rc |= find_face(m, node_face, i, v[0], v[3], v[7], v[4]);
        if (rc) return -1;
    }
    m.nfac = (int)m.face.size();
    return 0;
}

// vec_* helpers (src/tools.cpp:9-50): sequential accumulation
static inline void vminus(double *a, const double *b, const double *c) {
    for (int i = 0; i < 3; ++i) a[i] = b[i] - c[i];
}
static inline double vdot(const double *a, const double *b) {
    double s = 0.;
    for (int i = 0; i < 3; ++i) s += a[i] * b[i];
    return s;
}
static inline double vlen(const double *a) {
    double s = 0.;
    for (int i = 0; i < 3; ++i) s += a[i] * a[i];
    return std::sqrt(s);
}
static inline void vcross(double *a, const double *b, const double *c) {
    a[0] = b[1] * c[2] - b[2] * c[1];
    a[1] = -b[0] * c[2] + b[2] * c[0];
    a[2] = b[0] * c[1] - b[1] * c[0];
}
// Tri3DArea (src/BasicSolver.cpp:296-305)
static double tri_area(const double *x1, const double *x2, const double *x3) {
    double v1[3], v2[3];
    vminus(v1, x2, x1);
    vminus(v2, x3, x1);
    double a1 = v1[1] * v2[2] - v1[2] * v2[1];
    double a2 = -v1[0] * v2[2] + v1[2] * v2[0];
    double a3 = v1[0] * v2[1] - v1[1] * v2[0];
    return 0.5 * std::sqrt(a1 * a1 + a2 * a2 + a3 * a3);
}
// TetVolum (src/BasicSolver.cpp:306-314)
static double tet_vol(const double *x1, const double *x2, const double *x3,
                      const double *x4) {
    double v1[3], d1[3], d2[3], d3[3];
    vminus(d1, x2, x1);
    vminus(d2, x3, x1);
    vminus(d3, x4, x1);
    vcross(v1, d1, d2);
    return 1. / 6. * std::fabs(vdot(v1, d3));
}

// CellFaceInfo (src/BasicSolver.cpp:317-512)
int cell_face_info(Mesh &m) {
    auto V = [&](int i) { return m.vert[i].data(); };
    // rough centre: mean of the 8 stored vertices (:326-337)
    for (auto &c : m.cell)
        for (int j = 0; j < 3; ++j)
            c.x[j] = 1. / 8. *
                     (V(c.v[0])[j] + V(c.v[1])[j] + V(c.v[2])[j] + V(c.v[3])[j] +
                      V(c.v[4])[j] + V(c.v[5])[j] + V(c.v[6])[j] + V(c.v[7])[j]);
    // faces (:340-402)
    for (int i = 0; i < m.nfac; ++i) {
        Face &F = m.face[i];
        const double *p1 = V(F.v[0]), *p2 = V(F.v[1]), *p3 = V(F.v[2]),
                     *p4 = V(F.v[3]);
        double a1 = tri_area(p1, p2, p4);
        double a2 = tri_area(p2, p3, p4);
        double xc1[3], xc2[3], d1[3], d2[3], dx[3];
        for (int j = 0; j < 3; ++j) {
            xc1[j] = 1. / 3 * (p1[j] + p2[j] + p4[j]);
            xc2[j] = 1. / 3 * (p2[j] + p3[j] + p4[j]);
        }
        for (int j = 0; j < 3; ++j) F.x[j] = (a1 * xc1[j] + a2 * xc2[j]) / (a1 + a2);
        vminus(d1, p1, p3);
        vminus(d2, p2, p4);
        vcross(F.n, d1, d2);
        for (int j = 0; j < 3; ++j) F.n[j] = 0.5 * F.n[j];
        vminus(dx, F.x, m.cell[F.c1].x);
        if (vdot(F.n, dx) < 0.)
            for (int j = 0; j < 3; ++j) F.n[j] = -F.n[j];
        F.area = vlen(F.n);
        if (std::fabs((a1 + a2 - F.area) / F.area) > 1.e-4) return -1;
        if (F.c2 >= 0) {
            vminus(d1, F.x, m.cell[F.c1].x);
            vminus(d2, F.x, m.cell[F.c2].x);
            double r1 = vlen(d1), r2 = vlen(d2);
            F.lambda = r2 / (r1 + r2);
        } else {
            F.lambda = 1.;
        }
    }
    // cells (:407-459)
    for (int i = 0; i < m.ncel; ++i) {
        Cell &c = m.cell[i];
        double x[8][3];
        for (int k = 0; k < 8; ++k)
            for (int j = 0; j < 3; ++j) x[k][j] = V(c.v[k])[j];
        double v1 = tet_vol(x[0], x[1], x[3], x[5]);
        double v2 = tet_vol(x[0], x[4], x[3], x[5]);
        double v3 = tet_vol(x[3], x[4], x[7], x[5]);
        double v4 = tet_vol(x[1], x[2], x[3], x[6]);
        double v5 = tet_vol(x[1], x[5], x[3], x[6]);
        double v6 = tet_vol(x[7], x[3], x[5], x[6]);
        c.vol = v1 + v2 + v3 + v4 + v5 + v6;
        double xc[6][3];
        for (int j = 0; j < 3; ++j) {
            xc[0][j] = 0.25 * (x[0][j] + x[1][j] + x[3][j] + x[5][j]);
            xc[1][j] = 0.25 * (x[0][j] + x[4][j] + x[3][j] + x[5][j]);
            xc[2][j] = 0.25 * (x[3][j] + x[4][j] + x[7][j] + x[5][j]);
            xc[3][j] = 0.25 * (x[1][j] + x[2][j] + x[3][j] + x[6][j]);
            xc[4][j] = 0.25 * (x[1][j] + x[5][j] + x[3][j] + x[6][j]);
            xc[5][j] = 0.25 * (x[7][j] + x[3][j] + x[5][j] + x[6][j]);
        }
        double vs = v1 + v2 + v3 + v4 + v5 + v6;
        for (int j = 0; j < 3; ++j)
            c.x[j] = 1. / vs *
                     (v1 * xc[0][j] + v2 * xc[1][j] + v3 * xc[2][j] +
                      v4 * xc[3][j] + v5 * xc[4][j] + v6 * xc[5][j]);
        if (std::fabs((c.vol - (v1 + v2 + v3 + v4 + v5 + v6)) / c.vol) > 1.e-3)
            return -2;
        for (int j = 0; j < c.nface; ++j) {
            const Face &F = m.face[c.face[j]];
            int nc = F.c1;
            if (nc == i) nc = F.c2;
            c.cell[j] = nc;
        }
    }
    // boundary distance (:463-474)
    for (int i = 0; i < m.nbnd; ++i) {
        const Face &F = m.face[m.bnd[i].face];
        double dx[3];
        vminus(dx, F.x, m.cell[F.c1].x);
        m.bnd[i].distance = vdot(dx, F.n) / F.area;
        if (m.bnd[i].distance < 0.) return -3;
    }
    // rlencos, Xpac, Xnac (:477-509)
    for (int i = 0; i < m.nfac; ++i) {
        Face &F = m.face[i];
        double dx[3];
        if (F.c2 < 0) vminus(dx, F.x, m.cell[F.c1].x);
        else vminus(dx, m.cell[F.c2].x, m.cell[F.c1].x);
        F.rlencos = F.area / (vdot(dx, F.n) / F.area);
        if (F.c2 >= 0) {
            double d[3], ss;
            vminus(d, F.x, m.cell[F.c1].x);
            ss = vdot(d, F.n) / F.area;
            for (int j = 0; j < 3; ++j) d[j] = ss * F.n[j] / F.area;
            vminus(F.xpac, F.x, d);
            vminus(d, F.x, m.cell[F.c2].x);
            ss = vdot(d, F.n) / F.area;
            for (int j = 0; j < 3; ++j) d[j] = ss * F.n[j] / F.area;
            vminus(F.xnac, F.x, d);
        }
    }
    return 0;
}

// findCoupledBoundary (src/Preprocessor.cpp:308-438): sub-mesh of the cells
// new2old, vertices and faces renumbered in first-use order, original
// boundaries first, then faces cut by the split as coupled boundaries.
int split_region(const Mesh &full, const std::vector<int> &new2old,
                 int coupled_id, Mesh &o) {
    o = Mesh();
    int nc = (int)new2old.size();
    o.cell.resize(nc);
    std::map<int, int> v_old2new, f_old2new;
    for (int i = 0; i < nc; ++i) {
        o.cell[i] = full.cell[new2old[i]];
        for (int j = 0; j < 8; ++j) {
            int iv = o.cell[i].v[j];
            auto it = v_old2new.find(iv);
            if (it != v_old2new.end()) {
                o.cell[i].v[j] = it->second;
            } else {
                int nv = (int)o.vert.size();
                o.vert.push_back(full.vert[iv]);
                o.cell[i].v[j] = nv;
                v_old2new[iv] = nv;
            }
        }
    }
    for (int i = 0; i < nc; ++i) {
        Cell &c = o.cell[i];
        for (int j = 0; j < c.nface; ++j) {
            int iface = c.face[j];
            auto it = f_old2new.find(iface);
            if (it != f_old2new.end()) {
                o.face[it->second].c2 = i;
                c.face[j] = it->second;
            } else {
                Face F = full.face[iface];
                F.c1 = i;
                F.c2 = -10;
                for (int k = 0; k < 4; ++k) F.v[k] = v_old2new.at(F.v[k]);
                int nf = (int)o.face.size();
                c.face[j] = nf;
                f_old2new[iface] = nf;
                o.face.push_back(F);
            }
        }
    }
    int nb = 0;
    for (size_t i = 0; i < o.face.size(); ++i) {
        Face &F = o.face[i];
        if (F.c2 < 0 && F.bnd >= 0) {
            Bnd b = full.bnd[F.bnd];
            b.face = (int)i;
            o.bnd.push_back(b);
            F.bnd = nb++;
        }
    }
    int ncb = nb;
    for (size_t i = 0; i < o.face.size(); ++i) {
        Face &F = o.face[i];
        if (F.c2 < 0 && F.bnd < 0) {
            Bnd b;
            for (int k = 0; k < 4; ++k) b.v[k] = F.v[k];
            b.face = (int)i;
            b.rid = coupled_id;
            o.bnd.push_back(b);
            F.bnd = nb++;
        }
    }
    o.nvrt = (int)o.vert.size();
    o.ncel = nc;
    o.nfac = (int)o.face.size();
    o.nbnd = nb;
    o.ncb = ncb;
    return 0;
}

// Greedy first-fit colouring in natural order (no reference counterpart:
// it orders the multicolour SSOR / ILU(0) sweeps of the GPU backend).
std::vector<int> greedy_colour(int n, const std::vector<std::vector<int>> &adj) {
    std::vector<int> col(n, -1);
    std::vector<int> mark;
    for (int i = 0; i < n; ++i) {
        mark.assign(adj[i].size() + 2, 0);
        for (int j : adj[i])
            if (j < i && col[j] >= 0 && col[j] < (int)mark.size()) mark[col[j]] = 1;
        int c = 0;
        while (mark[c]) ++c;
        col[i] = c;
    }
    return col;
}

// Restates the product's block colouring rule for its ILU(0) order (the
// reference has none; this is the ordering the GPU factors in). Eligible
// meshes: every cell has 4 faces, n is a multiple of B, no triangle i-j-k in
// the adjacency. Blocks are cells [bB, (b+1)B); the block graph joins blocks
// with a face between them. Bipartite: colour by breadth-first parity from
// the lowest uncoloured block of each component (colour 0); otherwise
// greedy first-fit over blocks in order.
std::vector<int> block_colour(int n, const std::vector<int> &nface,
                              const std::vector<std::vector<int>> &adj, int B) {
    if (n == 0 || n % B) return {};
    for (int i = 0; i < n; ++i)
        if (nface[i] != 4) return {};
    std::vector<char> mark(n, 0);
    for (int i = 0; i < n; ++i) {  // triangle test through a neighbour mark
        for (int j : adj[i]) mark[j] = 1;
        bool tri = false;
        for (int j : adj[i])
            for (int k : adj[j])
                if (k != i && mark[k]) tri = true;
        for (int j : adj[i]) mark[j] = 0;
        if (tri) return {};
    }
    const int nb = n / B;
    std::vector<std::vector<int>> badj(nb);
    for (int i = 0; i < n; ++i)
        for (int j : adj[i])
            if (j / B != i / B) badj[i / B].push_back(j / B);
    std::vector<int> bc(nb, -1);
    bool bip = true;
    for (int s0 = 0; s0 < nb && bip; ++s0) {
        if (bc[s0] >= 0) continue;
        std::deque<int> q{s0};
        bc[s0] = 0;
        while (!q.empty() && bip) {
            int u = q.front();
            q.pop_front();
            for (int v : badj[u]) {
                if (bc[v] < 0) {
                    bc[v] = 1 - bc[u];
                    q.push_back(v);
                } else if (bc[v] == bc[u]) {
                    bip = false;
                }
            }
        }
    }
    if (!bip) bc = greedy_colour(nb, badj);
    std::vector<int> out(n);
    for (int i = 0; i < n; ++i) out[i] = bc[i / B];
    return out;
}

}  // namespace oref
