// Host-side mesh set-up: synthetic generator, gmsh 2.2 I/O, face
// construction, geometry, conjugate-heat-transfer split, cell colouring.
// Compiled with -ffp-contract=off: geometry is computed in the operation
// order of CellFaceInfo (src/BasicSolver.cpp:317-512) so that every FP64
// value is bit-identical to the reference's.
#include "mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace cfd {

int load_raw(const cfd_raw_mesh &r, HostMesh &m, std::string &err) {
    if (r.nvrt <= 0 || r.ncel <= 0 || r.nbnd < 0 || !r.vert || !r.cell_vert) {
        err = "empty or incomplete raw mesh";
        return -1;
    }
    m = HostMesh();
    m.nvrt = r.nvrt;
    m.ncel = r.ncel;
    m.nbnd = r.nbnd;
    m.vert.assign(r.vert, r.vert + 3 * (size_t)r.nvrt);
    m.cell_vert.assign(r.cell_vert, r.cell_vert + 8 * (size_t)r.ncel);
    m.cell_rid.assign(r.cell_rid, r.cell_rid + r.ncel);
    m.bnd_vert.assign(r.bnd_vert, r.bnd_vert + 4 * (size_t)r.nbnd);
    m.bnd_rid.assign(r.bnd_rid, r.bnd_rid + r.nbnd);
    for (int v : m.cell_vert)
        if (v < 0 || v >= m.nvrt) { err = "cell vertex out of range"; return -1; }
    for (int v : m.bnd_vert)
        if (v < 0 || v >= m.nvrt) { err = "boundary vertex out of range"; return -1; }
    return 0;
}

// ------------------------------------------------------------------ faces --
// Face identity is the set of distinct vertex ids. Open-addressing table on a
// 128-bit key (four sorted ids, -1 padded). For conforming meshes this finds
// the same face as FindFace's node-list subset search
// (src/Preprocessor.cpp:248-264), and the face numbering is identical:
// boundary faces first in boundary order (:198-212), then interior faces in
// order of first discovery over cells x the six hexahedron slots (:214-230).
namespace {
struct Key {
    int32_t v[4];
    bool operator==(const Key &o) const {
        return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2] && v[3] == o.v[3];
    }
};
inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline Key make_key(int a, int b, int c, int d) {
    int t[4] = {a, b, c, d};
    std::sort(t, t + 4);
    Key k;
    int n = 0;
    for (int i = 0; i < 4; ++i)
        if (i == 0 || t[i] != t[i - 1]) k.v[n++] = t[i];
    for (; n < 4; ++n) k.v[n] = -1;
    return k;
}
struct FaceTable {
    std::vector<Key> keys;
    std::vector<int> val;
    uint64_t mask = 0;
    explicit FaceTable(size_t expect) {
        size_t cap = 16;
        while (cap < 2 * expect + 16) cap <<= 1;
        keys.resize(cap);
        val.assign(cap, -1);
        mask = cap - 1;
    }
    // returns existing value, or inserts v and returns -1
    int find_or_insert(const Key &k, int v) {
        uint64_t h = mix64((uint64_t)(uint32_t)k.v[0] | ((uint64_t)(uint32_t)k.v[1] << 32)) ^
                     mix64((uint64_t)(uint32_t)k.v[2] | ((uint64_t)(uint32_t)k.v[3] << 32));
        for (uint64_t p = h & mask;; p = (p + 1) & mask) {
            if (val[p] < 0) {
                keys[p] = k;
                val[p] = v;
                return -1;
            }
            if (keys[p] == k) return val[p];
        }
    }
};
}  // namespace

int build_faces(HostMesh &m, std::string &err) {
    const int C = m.ncel, B = m.nbnd;
    m.cell_nface.assign(C, 0);
    m.cell_face.assign((size_t)C * 6, -1);
    m.face_vert.clear();
    m.face_c1.clear();
    m.face_c2.clear();
    m.face_bnd.clear();
    m.bnd_face.assign(B, -1);
    size_t expect = (size_t)B + 4 * (size_t)C;
    m.face_vert.reserve(4 * expect);
    m.face_c1.reserve(expect);
    m.face_c2.reserve(expect);
    m.face_bnd.reserve(expect);
    FaceTable tab(expect);
    for (int i = 0; i < B; ++i) {
        const int *v = &m.bnd_vert[4 * (size_t)i];
        int f = (int)m.face_c1.size();
        if (tab.find_or_insert(make_key(v[0], v[1], v[2], v[3]), f) >= 0) {
            err = "duplicate boundary face";
            return -1;
        }
        m.bnd_face[i] = f;
        m.face_vert.insert(m.face_vert.end(), v, v + 4);
        m.face_c1.push_back(-10);
        m.face_c2.push_back(-10);
        m.face_bnd.push_back(i);
    }
    // hexahedron face slots (src/Preprocessor.cpp:224-229)
    static const int slot[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                   {1, 2, 6, 5}, {3, 2, 6, 7}, {0, 3, 7, 4}};
    for (int ic = 0; ic < C; ++ic) {
        const int *cv = &m.cell_vert[8 * (size_t)ic];
        for (int s = 0; s < 6; ++s) {
            int n1 = cv[slot[s][0]], n2 = cv[slot[s][1]], n3 = cv[slot[s][2]],
                n4 = cv[slot[s][3]];
            int rep = (n1 == n2 || n1 == n3 || n1 == n4) + (n2 == n3 || n2 == n4) + (n3 == n4);
            if (rep > 1) continue;  // collapsed slot (src/Preprocessor.cpp:242-246)
            int nf = (int)m.face_c1.size();
            int f = tab.find_or_insert(make_key(n1, n2, n3, n4), nf);
            if (f < 0) {
                m.face_vert.push_back(n1);
                m.face_vert.push_back(n2);
                m.face_vert.push_back(n3);
                m.face_vert.push_back(n4);
                m.face_c1.push_back(ic);
                m.face_c2.push_back(-10);
                m.face_bnd.push_back(-10);
                f = nf;
            } else if (m.face_bnd[f] >= 0) {
                if (m.face_c1[f] == -10) m.face_c1[f] = ic;
                else if (m.face_c2[f] == -10) m.face_c2[f] = ic;
                else { err = "boundary face shared by three cells"; return -1; }
            } else {
                if (m.face_c2[f] != -10) { err = "interior face shared by three cells"; return -1; }
                m.face_c2[f] = ic;
            }
            m.cell_face[6 * (size_t)ic + m.cell_nface[ic]++] = f;
        }
    }
    m.nfac = (int)m.face_c1.size();
    m.ncb = 0;  // NCoupledBnd stays 0 unless a split sets it (src/BasicSolver.cpp:14)
    return 0;
}

// --------------------------------------------------------------- geometry --
namespace {
inline void g_sub(double *a, const double *b, const double *c) {
    a[0] = b[0] - c[0];
    a[1] = b[1] - c[1];
    a[2] = b[2] - c[2];
}
inline double g_dot(const double *a, const double *b) {
    double s = 0.;
    s += a[0] * b[0];
    s += a[1] * b[1];
    s += a[2] * b[2];
    return s;
}
inline void g_cross(double *a, const double *b, const double *c) {
    a[0] = b[1] * c[2] - b[2] * c[1];
    a[1] = -b[0] * c[2] + b[2] * c[0];
    a[2] = b[0] * c[1] - b[1] * c[0];
}
inline double g_tri(const double *p, const double *q, const double *r) {
    double u[3], w[3], a[3];
    g_sub(u, q, p);
    g_sub(w, r, p);
    g_cross(a, u, w);
    return 0.5 * std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}
inline double g_tet(const double *p, const double *q, const double *r, const double *s) {
    double d1[3], d2[3], d3[3], c[3];
    g_sub(d1, q, p);
    g_sub(d2, r, p);
    g_sub(d3, s, p);
    g_cross(c, d1, d2);
    return 1. / 6. * std::fabs(g_dot(c, d3));
}
}  // namespace

int build_geometry(HostMesh &m, std::string &err) {
    const int C = m.ncel, F = m.nfac, B = m.nbnd;
    const double *V = m.vert.data();
    auto P = [&](int v) { return V + 3 * (size_t)v; };
    m.cell_x.assign(3 * (size_t)C, 0.);
    m.cell_vol.assign(C, 0.);
    m.face_x.assign(3 * (size_t)F, 0.);
    m.face_n.assign(3 * (size_t)F, 0.);
    m.face_xpac.assign(3 * (size_t)F, 0.);
    m.face_xnac.assign(3 * (size_t)F, 0.);
    m.face_area.assign(F, 0.);
    m.face_lambda.assign(F, 0.);
    m.face_rlencos.assign(F, 0.);
    m.bnd_dist.assign(B, 0.);
    // vertex-mean centre, used for orientation and interpolation weights
    std::vector<double> xr(3 * (size_t)C);
    for (int i = 0; i < C; ++i) {
        const int *cv = &m.cell_vert[8 * (size_t)i];
        for (int j = 0; j < 3; ++j)
            xr[3 * i + j] = 1. / 8. * (P(cv[0])[j] + P(cv[1])[j] + P(cv[2])[j] + P(cv[3])[j] +
                                      P(cv[4])[j] + P(cv[5])[j] + P(cv[6])[j] + P(cv[7])[j]);
    }
    for (int f = 0; f < F; ++f) {
        const int *fv = &m.face_vert[4 * (size_t)f];
        const double *p1 = P(fv[0]), *p2 = P(fv[1]), *p3 = P(fv[2]), *p4 = P(fv[3]);
        double a1 = g_tri(p1, p2, p4), a2 = g_tri(p2, p3, p4);
        double *x = &m.face_x[3 * (size_t)f], *n = &m.face_n[3 * (size_t)f];
        for (int j = 0; j < 3; ++j) {
            double c1 = 1. / 3 * (p1[j] + p2[j] + p4[j]);
            double c2 = 1. / 3 * (p2[j] + p3[j] + p4[j]);
            x[j] = (a1 * c1 + a2 * c2) / (a1 + a2);
        }
        double d1[3], d2[3], dx[3];
        g_sub(d1, p1, p3);
        g_sub(d2, p2, p4);
        g_cross(n, d1, d2);
        for (int j = 0; j < 3; ++j) n[j] = 0.5 * n[j];
        int c1 = m.face_c1[f], c2 = m.face_c2[f];
        if (c1 < 0) { err = "face without a cell (unlisted boundary?)"; return -1; }
        g_sub(dx, x, &xr[3 * (size_t)c1]);
        if (g_dot(n, dx) < 0.)
            for (int j = 0; j < 3; ++j) n[j] = -n[j];
        double area = std::sqrt(g_dot(n, n));
        m.face_area[f] = area;
        if (std::fabs((a1 + a2 - area) / area) > 1.e-4) {
            err = "face area check failed (non-planar or crossed face)";
            return -1;
        }
        if (c2 >= 0) {
            g_sub(d1, x, &xr[3 * (size_t)c1]);
            g_sub(d2, x, &xr[3 * (size_t)c2]);
            double r1 = std::sqrt(g_dot(d1, d1)), r2 = std::sqrt(g_dot(d2, d2));
            m.face_lambda[f] = r2 / (r1 + r2);
        } else {
            m.face_lambda[f] = 1.;
        }
    }
    m.cell_cell.assign(6 * (size_t)C, -1);
    for (int i = 0; i < C; ++i) {
        const int *cv = &m.cell_vert[8 * (size_t)i];
        const double *x1 = P(cv[0]), *x2 = P(cv[1]), *x3 = P(cv[2]), *x4 = P(cv[3]),
                     *x5 = P(cv[4]), *x6 = P(cv[5]), *x7 = P(cv[6]), *x8 = P(cv[7]);
        double v1 = g_tet(x1, x2, x4, x6), v2 = g_tet(x1, x5, x4, x6), v3 = g_tet(x4, x5, x8, x6),
               v4 = g_tet(x2, x3, x4, x7), v5 = g_tet(x2, x6, x4, x7), v6 = g_tet(x8, x4, x6, x7);
        double vol = v1 + v2 + v3 + v4 + v5 + v6;
        m.cell_vol[i] = vol;
        double *cx = &m.cell_x[3 * (size_t)i];
        for (int j = 0; j < 3; ++j) {
            double c1 = 0.25 * (x1[j] + x2[j] + x4[j] + x6[j]);
            double c2 = 0.25 * (x1[j] + x5[j] + x4[j] + x6[j]);
            double c3 = 0.25 * (x4[j] + x5[j] + x8[j] + x6[j]);
            double c4 = 0.25 * (x2[j] + x3[j] + x4[j] + x7[j]);
            double c5 = 0.25 * (x2[j] + x6[j] + x4[j] + x7[j]);
            double c6 = 0.25 * (x8[j] + x4[j] + x6[j] + x7[j]);
            cx[j] = 1. / vol * (v1 * c1 + v2 * c2 + v3 * c3 + v4 * c4 + v5 * c5 + v6 * c6);
        }
        if (!(vol > 0.)) { err = "non-positive cell volume"; return -1; }
        for (int j = 0; j < m.cell_nface[i]; ++j) {
            int f = m.cell_face[6 * (size_t)i + j];
            int nc = m.face_c1[f];
            if (nc == i) nc = m.face_c2[f];
            m.cell_cell[6 * (size_t)i + j] = nc;
        }
    }
    for (int b = 0; b < B; ++b) {
        int f = m.bnd_face[b], c = m.face_c1[f];
        double dx[3];
        g_sub(dx, &m.face_x[3 * (size_t)f], &m.cell_x[3 * (size_t)c]);
        m.bnd_dist[b] = g_dot(dx, &m.face_n[3 * (size_t)f]) / m.face_area[f];
        if (m.bnd_dist[b] < 0.) { err = "wall distance is negative"; return -1; }
    }
    for (int f = 0; f < F; ++f) {
        int c1 = m.face_c1[f], c2 = m.face_c2[f];
        const double *x = &m.face_x[3 * (size_t)f], *n = &m.face_n[3 * (size_t)f];
        double area = m.face_area[f], dx[3];
        if (c2 < 0) g_sub(dx, x, &m.cell_x[3 * (size_t)c1]);
        else g_sub(dx, &m.cell_x[3 * (size_t)c2], &m.cell_x[3 * (size_t)c1]);
        m.face_rlencos[f] = area / (g_dot(dx, n) / area);
        if (c2 >= 0) {
            double d[3], ss;
            g_sub(d, x, &m.cell_x[3 * (size_t)c1]);
            ss = g_dot(d, n) / area;
            for (int j = 0; j < 3; ++j) d[j] = ss * n[j] / area;
            g_sub(&m.face_xpac[3 * (size_t)f], x, d);
            g_sub(d, x, &m.cell_x[3 * (size_t)c2]);
            ss = g_dot(d, n) / area;
            for (int j = 0; j < 3; ++j) d[j] = ss * n[j] / area;
            g_sub(&m.face_xnac[3 * (size_t)f], x, d);
        }
    }
    return 0;
}

// ------------------------------------------------------------------ split --
// Sub-mesh of `cells` with the semantics of findCoupledBoundary
// (src/Preprocessor.cpp:308-438): vertices and faces renumbered in first-use
// order, surviving boundary faces first (original boundary data kept), then
// the faces cut by the split as coupled boundaries with region coupled_id.
int split_cells(const HostMesh &full, const std::vector<int> &cells,
                int coupled_id, HostMesh &o, std::string &err) {
    o = HostMesh();
    const int nc = (int)cells.size();
    std::vector<int> vmap(full.nvrt, -1), fmap(full.nfac, -1);
    o.cell_vert.resize(8 * (size_t)nc);
    o.cell_rid.resize(nc);
    o.cell_nface.resize(nc);
    o.cell_face.assign(6 * (size_t)nc, -1);
    for (int i = 0; i < nc; ++i) {
        int oc = cells[i];
        o.cell_rid[i] = full.cell_rid[oc];
        o.cell_nface[i] = full.cell_nface[oc];
        for (int j = 0; j < 8; ++j) {
            int v = full.cell_vert[8 * (size_t)oc + j];
            if (vmap[v] < 0) {
                vmap[v] = (int)(o.vert.size() / 3);
                o.vert.insert(o.vert.end(), &full.vert[3 * (size_t)v], &full.vert[3 * (size_t)v] + 3);
            }
            o.cell_vert[8 * (size_t)i + j] = vmap[v];
        }
    }
    for (int i = 0; i < nc; ++i) {
        int oc = cells[i];
        for (int j = 0; j < full.cell_nface[oc]; ++j) {
            int f = full.cell_face[6 * (size_t)oc + j];
            if (fmap[f] >= 0) {
                o.face_c2[fmap[f]] = i;
            } else {
                fmap[f] = (int)o.face_c1.size();
                for (int k = 0; k < 4; ++k) o.face_vert.push_back(vmap[full.face_vert[4 * (size_t)f + k]]);
                o.face_c1.push_back(i);
                o.face_c2.push_back(-10);
                o.face_bnd.push_back(full.face_bnd[f]);
            }
            o.cell_face[6 * (size_t)i + j] = fmap[f];
        }
    }
    int nf = (int)o.face_c1.size(), nb = 0;
    for (int f = 0; f < nf; ++f)
        if (o.face_c2[f] < 0 && o.face_bnd[f] >= 0) {
            int ob = o.face_bnd[f];
            o.bnd_rid.push_back(full.bnd_rid[ob]);
            o.bnd_vert.insert(o.bnd_vert.end(), &full.bnd_vert[4 * (size_t)ob], &full.bnd_vert[4 * (size_t)ob] + 4);
            o.bnd_face.push_back(f);
            o.face_bnd[f] = nb++;
        }
    o.ncb = nb;
    for (int f = 0; f < nf; ++f)
        if (o.face_c2[f] < 0 && o.face_bnd[f] < 0) {
            o.bnd_rid.push_back(coupled_id);
            o.bnd_vert.insert(o.bnd_vert.end(), &o.face_vert[4 * (size_t)f], &o.face_vert[4 * (size_t)f] + 4);
            o.bnd_face.push_back(f);
            o.face_bnd[f] = nb++;
        }
    o.nvrt = (int)(o.vert.size() / 3);
    o.ncel = nc;
    o.nfac = nf;
    o.nbnd = nb;
    if (nc == 0) { err = "empty sub-mesh"; return -1; }
    return 0;
}

// Fluid / solid parts as buildSolver produces them (src/Preprocessor.cpp:
// 441-515), each with faces and geometry. Without conjugate heat transfer the
// fluid part is the whole mesh.
int build_parts(const cfd_raw_mesh &raw, const cfd_params &p, HostMesh &fluid, HostMesh &solid,
                bool &has_solid, int &coupled_id, std::string &err) {
    HostMesh full;
    if (load_raw(raw, full, err) || build_faces(full, err)) return -1;
    has_solid = false;
    coupled_id = -1;
    if (!p.cht) {
        fluid = std::move(full);
    } else {
        auto find = [&](int id) -> const cfd_region * {
            for (int r = 0; r < p.nregion; ++r)
                if (p.region[r].id == id) return &p.region[r];
            return nullptr;
        };
        for (int i = 0; i != p.nregion + 1; ++i)
            if (!find(i)) { coupled_id = i; break; }
        std::vector<int> fl, so;
        for (int i = 0; i < full.ncel; ++i) {
            const cfd_region *r = find(full.cell_rid[i]);
            if (!r || r->type1 != 5) { err = "cell region is not a body"; return -1; }
            (r->type2 == 0 ? fl : so).push_back(i);
        }
        if (split_cells(full, fl, coupled_id, fluid, err)) return -1;
        if (split_cells(full, so, coupled_id, solid, err)) return -1;
        has_solid = true;
    }
    if (build_geometry(fluid, err)) return -1;
    colour_cells(fluid);
    if (has_solid) {
        if (build_geometry(solid, err)) return -1;
        colour_cells(solid);
    }
    return 0;
}

// Disjoint union of independent meshes (batched cases): vertices, cells,
// faces and boundaries of part k follow those of part k-1 with every index
// offset; negative sentinels stay as they are. Each part keeps its own face
// order (boundary faces first), so off[k] tells where part k starts in each
// kind of array: off = {vert, cell, face, bnd} x (nparts + 1).
int concat_meshes(const std::vector<HostMesh> &parts, HostMesh &o, std::vector<int> off[4],
                  std::string &err) {
    o = HostMesh();
    for (int t = 0; t < 4; ++t) off[t].assign(1, 0);
    auto shift = [](std::vector<int> &dst, const std::vector<int> &src, int by) {
        for (int v : src) dst.push_back(v >= 0 ? v + by : v);
    };
    for (const HostMesh &m : parts) {
        const int v0 = o.nvrt, c0 = o.ncel, f0 = o.nfac, b0 = o.nbnd;
        if ((long long)c0 + m.ncel > 2000000000LL || (long long)f0 + m.nfac > 2000000000LL) {
            err = "batched meshes too large";
            return -1;
        }
        o.vert.insert(o.vert.end(), m.vert.begin(), m.vert.end());
        shift(o.cell_vert, m.cell_vert, v0);
        o.cell_rid.insert(o.cell_rid.end(), m.cell_rid.begin(), m.cell_rid.end());
        o.cell_nface.insert(o.cell_nface.end(), m.cell_nface.begin(), m.cell_nface.end());
        shift(o.cell_face, m.cell_face, f0);
        shift(o.cell_cell, m.cell_cell, c0);
        shift(o.face_vert, m.face_vert, v0);
        shift(o.face_c1, m.face_c1, c0);
        shift(o.face_c2, m.face_c2, c0);
        shift(o.face_bnd, m.face_bnd, b0);
        shift(o.bnd_vert, m.bnd_vert, v0);
        o.bnd_rid.insert(o.bnd_rid.end(), m.bnd_rid.begin(), m.bnd_rid.end());
        shift(o.bnd_face, m.bnd_face, f0);
        for (auto pr : {std::make_pair(&o.face_x, &m.face_x), std::make_pair(&o.face_n, &m.face_n),
                        std::make_pair(&o.face_xpac, &m.face_xpac),
                        std::make_pair(&o.face_xnac, &m.face_xnac),
                        std::make_pair(&o.face_area, &m.face_area),
                        std::make_pair(&o.face_lambda, &m.face_lambda),
                        std::make_pair(&o.face_rlencos, &m.face_rlencos),
                        std::make_pair(&o.cell_x, &m.cell_x),
                        std::make_pair(&o.cell_vol, &m.cell_vol),
                        std::make_pair(&o.bnd_dist, &m.bnd_dist)})
            pr.first->insert(pr.first->end(), pr.second->begin(), pr.second->end());
        o.colour.insert(o.colour.end(), m.colour.begin(), m.colour.end());
        o.ncolour = std::max(o.ncolour, m.ncolour);
        o.nvrt += m.nvrt;
        o.ncel += m.ncel;
        o.nfac += m.nfac;
        o.nbnd += m.nbnd;
        o.ncb += m.ncb;
        off[0].push_back(o.nvrt);
        off[1].push_back(o.ncel);
        off[2].push_back(o.nfac);
        off[3].push_back(o.nbnd);
    }
    // the union's cells need not come in whole blocks: greedy colouring only
    o.bcolour.assign(o.ncel, -1);
    o.nbcolour = 0;
    return 0;
}

namespace {
// Block colouring for the ILU(0) order (the product's own ordering; the
// reference factors in natural order). Defined for tetrahedral meshes (every
// cell has 4 faces) whose cells come in whole blocks of kIluBlock
// consecutive cells and whose adjacency has no triangles (no two neighbours
// of a cell are neighbours of each other). Blocks are coloured so that no
// two blocks of one colour share a face: by breadth-first parity when the
// block graph is bipartite (each connected component from its lowest block,
// colour 0), otherwise greedy first-fit in block order. Otherwise -1.
void colour_blocks(HostMesh &m) {
    const int n = m.ncel, B = kIluBlock;
    m.bcolour.assign(n, -1);
    m.nbcolour = 0;
    if (n == 0 || n % B) return;
    auto nbr = [&](int i, int a) { return m.cell_cell[6 * (size_t)i + a]; };
    for (int i = 0; i < n; ++i)
        if (m.cell_nface[i] != 4) return;
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 4; ++a) {
            const int j = nbr(i, a);
            if (j < 0) continue;
            for (int b = 0; b < 4; ++b) {
                const int k = nbr(j, b);
                if (k < 0 || k == i) continue;
                for (int c = 0; c < 4; ++c)
                    if (nbr(i, c) == k) return;  // triangle i, j, k
            }
        }
    const int nb = n / B;
    std::vector<std::vector<int>> badj(nb);
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 4; ++a) {
            const int j = nbr(i, a);
            if (j >= 0 && j / B != i / B) badj[i / B].push_back(j / B);
        }
    std::vector<int> bc(nb, -1), queue;
    bool bipartite = true;
    for (int s = 0; s < nb && bipartite; ++s) {
        if (bc[s] >= 0) continue;
        bc[s] = 0;
        queue.assign(1, s);
        for (size_t h = 0; h < queue.size() && bipartite; ++h) {
            const int u = queue[h];
            for (int v : badj[u]) {
                if (bc[v] < 0) {
                    bc[v] = 1 - bc[u];
                    queue.push_back(v);
                } else if (bc[v] == bc[u]) {
                    bipartite = false;
                }
            }
        }
    }
    if (!bipartite) {
        std::vector<char> used;
        for (int b = 0; b < nb; ++b) {
            used.assign(badj[b].size() + 2, 0);
            for (int v : badj[b])
                if (v < b && bc[v] < (int)used.size()) used[bc[v]] = 1;
            int c = 0;
            while (used[c]) ++c;
            bc[b] = c;
        }
    }
    int maxc = 0;
    for (int i = 0; i < n; ++i) {
        m.bcolour[i] = bc[i / B];
        maxc = std::max(maxc, m.bcolour[i]);
    }
    m.nbcolour = maxc + 1;
}
}  // namespace

// Greedy first-fit colouring over the interior-face adjacency, natural
// order: colour(i) = least colour not used by a neighbour j < i. Then the
// block colouring of the ILU(0) order (colour_blocks).
void colour_cells(HostMesh &m) {
    m.colour.assign(m.ncel, 0);
    int maxc = 0;
    for (int i = 0; i < m.ncel; ++i) {
        unsigned used = 0;
        for (int j = 0; j < m.cell_nface[i]; ++j) {
            int nb = m.cell_cell[6 * (size_t)i + j];
            if (nb >= 0 && nb < i && m.colour[nb] < 32) used |= 1u << m.colour[nb];
        }
        int c = 0;
        while (used & (1u << c)) ++c;
        m.colour[i] = c;
        maxc = std::max(maxc, c);
    }
    m.ncolour = maxc + 1;
    colour_blocks(m);
}

}  // namespace cfd

// ======================================================== C entry points ==
using cfd::HostMesh;

namespace cfd {
std::string &last_error();
}

struct cfd_mesh_s {
    HostMesh m;
};

extern "C" {

int cfd_mesh_box_counts(int nx, int ny, int nz, int tets, int *nvrt, int *ncel, int *nbnd) {
    if (nx < 1 || ny < 1 || nz < 1) {
        cfd::last_error() = "box dimensions must be positive";
        return -1;
    }
    long long nv = (long long)(nx + 1) * (ny + 1) * (nz + 1);
    long long nc = (long long)nx * ny * nz * (tets ? 6 : 1);
    long long nb = 2LL * ((long long)nx * ny + (long long)ny * nz + (long long)nx * nz) * (tets ? 2 : 1);
    if (nv > 2000000000LL || nc > 2000000000LL) {
        cfd::last_error() = "box too large";
        return -1;
    }
    *nvrt = (int)nv;
    *ncel = (int)nc;
    *nbnd = (int)nb;
    return 0;
}

static inline uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int cfd_mesh_box(int nx, int ny, int nz, double lx, double ly, double lz, int tets,
                 double jitter, uint64_t seed, const int side_rid[6], int fluid_rid,
                 int solid_rid, const double *block, double *vert, int *cell_vert,
                 int *cell_rid, int *bnd_vert, int *bnd_rid) {
    int nv, nc, nb;
    if (cfd_mesh_box_counts(nx, ny, nz, tets, &nv, &nc, &nb)) return -1;
    if (!(jitter >= 0. && jitter < 0.25)) {
        cfd::last_error() = "jitter must be in [0, 0.25)";
        return -1;
    }
    const double h[3] = {lx / nx, ly / ny, lz / nz};
    const int N[3] = {nx, ny, nz};
    auto vid = [&](int i, int j, int k) { return i + (nx + 1) * (j + (ny + 1) * k); };
    uint64_t s = seed;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                int id[3] = {i, j, k};
                double *p = vert + 3 * (size_t)vid(i, j, k);
                bool interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
                for (int a = 0; a < 3; ++a) {
                    p[a] = (id[a] == N[a]) ? (a == 0 ? lx : a == 1 ? ly : lz) : id[a] * h[a];
                    if (interior && jitter > 0.) {
                        double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
                        p[a] += (2. * u - 1.) * jitter * h[a];
                    }
                }
            }
    // Kuhn split: axis-increment paths from corner (0,0,0) to (1,1,1)
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    size_t c = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double cx = (i + 0.5) * h[0], cy = (j + 0.5) * h[1], cz = (k + 0.5) * h[2];
                bool solid = block && cx > block[0] && cx < block[1] && cy > block[2] &&
                             cy < block[3] && cz > block[4] && cz < block[5];
                int rid = solid ? solid_rid : fluid_rid;
                if (!tets) {
                    int q[8] = {vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j + 1, k), vid(i, j + 1, k),
                                vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i + 1, j + 1, k + 1),
                                vid(i, j + 1, k + 1)};
                    std::memcpy(cell_vert + 8 * c, q, sizeof(q));
                    cell_rid[c++] = rid;
                    continue;
                }
                for (int t = 0; t < 6; ++t) {
                    int o[3] = {i, j, k}, v[4];
                    v[0] = vid(o[0], o[1], o[2]);
                    o[perm[t][0]]++;
                    v[1] = vid(o[0], o[1], o[2]);
                    o[perm[t][1]]++;
                    v[2] = vid(o[0], o[1], o[2]);
                    v[3] = vid(i + 1, j + 1, k + 1);
                    int q[8] = {v[0], v[1], v[2], v[2], v[3], v[3], v[3], v[3]};
                    std::memcpy(cell_vert + 8 * c, q, sizeof(q));
                    cell_rid[c++] = rid;
                }
            }
    // boundary: sides x-, x+, y-, y+, z-, z+; in-plane axes (u < v), v outer
    size_t b = 0;
    for (int side = 0; side < 6; ++side) {
        int ax = side / 2, hi = side % 2;
        int u = ax == 0 ? 1 : 0, w = ax == 2 ? 1 : 2;
        for (int jv = 0; jv < N[w]; ++jv)
            for (int ju = 0; ju < N[u]; ++ju) {
                int q[4];
                int corner[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
                for (int t = 0; t < 4; ++t) {
                    int id[3];
                    id[ax] = hi ? N[ax] : 0;
                    id[u] = ju + corner[t][0];
                    id[w] = jv + corner[t][1];
                    q[t] = vid(id[0], id[1], id[2]);
                }
                if (!tets) {
                    std::memcpy(bnd_vert + 4 * b, q, sizeof(q));
                    bnd_rid[b++] = side_rid[side];
                } else {
                    int t1[4] = {q[0], q[1], q[2], q[2]}, t2[4] = {q[0], q[2], q[3], q[3]};
                    std::memcpy(bnd_vert + 4 * b, t1, sizeof(t1));
                    bnd_rid[b++] = side_rid[side];
                    std::memcpy(bnd_vert + 4 * b, t2, sizeof(t2));
                    bnd_rid[b++] = side_rid[side];
                }
            }
    }
    return 0;
}

int cfd_gmsh_write(const char *path, const cfd_raw_mesh *m) {
    FILE *f = std::fopen(path, "w");
    if (!f) {
        cfd::last_error() = std::string("cannot open ") + path;
        return -1;
    }
    std::fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n", m->nvrt);
    for (int i = 0; i < m->nvrt; ++i)
        std::fprintf(f, "%d %.17g %.17g %.17g\n", i + 1, m->vert[3 * i], m->vert[3 * i + 1],
                     m->vert[3 * i + 2]);
    std::fprintf(f, "$EndNodes\n$Elements\n%d\n", m->nbnd + m->ncel);
    int id = 1;
    for (int i = 0; i < m->nbnd; ++i) {
        const int *v = m->bnd_vert + 4 * i;
        if (v[2] == v[3])
            std::fprintf(f, "%d 2 2 %d %d %d %d %d\n", id++, m->bnd_rid[i], m->bnd_rid[i], v[0] + 1,
                         v[1] + 1, v[2] + 1);
        else
            std::fprintf(f, "%d 3 2 %d %d %d %d %d %d\n", id++, m->bnd_rid[i], m->bnd_rid[i],
                         v[0] + 1, v[1] + 1, v[2] + 1, v[3] + 1);
    }
    for (int i = 0; i < m->ncel; ++i) {
        const int *v = m->cell_vert + 8 * i;
        bool tet = v[2] == v[3] && v[4] == v[5] && v[5] == v[6] && v[6] == v[7];
        if (tet)
            std::fprintf(f, "%d 4 2 %d %d %d %d %d %d\n", id++, m->cell_rid[i], m->cell_rid[i],
                         v[0] + 1, v[1] + 1, v[2] + 1, v[4] + 1);
        else
            std::fprintf(f, "%d 5 2 %d %d %d %d %d %d %d %d %d %d\n", id++, m->cell_rid[i],
                         m->cell_rid[i], v[0] + 1, v[1] + 1, v[2] + 1, v[3] + 1, v[4] + 1,
                         v[5] + 1, v[6] + 1, v[7] + 1);
    }
    std::fprintf(f, "$EndElements\n");
    std::fclose(f);
    return 0;
}

// Reader for the gmsh 2.2 ASCII subset of ReadGridFile
// (src/Preprocessor.cpp:13-146), with the same 8-vertex expansion of
// tetrahedra, prisms and pyramids. Returns counts when arrays are null.
static int gmsh_read(const char *path, int *nvrt, int *ncel, int *nbnd, double *vert,
                     int *cell_vert, int *cell_rid, int *bnd_vert, int *bnd_rid) {
    std::ifstream in(path);
    if (!in) {
        cfd::last_error() = std::string("grid file not found: ") + path;
        return -1;
    }
    std::string tok;
    for (int i = 0; i < 6; ++i) in >> tok;  // $MeshFormat 2.2 0 8 $EndMeshFormat $Nodes
    if (tok != "$Nodes") {
        cfd::last_error() = "unsupported gmsh layout (expected $Nodes as 6th token)";
        return -1;
    }
    int nv = 0;
    in >> nv;
    if (nv <= 0) { cfd::last_error() = "no vertices"; return -1; }
    for (int i = 0; i < nv; ++i) {
        int id;
        double x, y, z;
        in >> id >> x >> y >> z;
        if (vert) { vert[3 * i] = x; vert[3 * i + 1] = y; vert[3 * i + 2] = z; }
    }
    in >> tok >> tok;
    int ne = 0;
    in >> ne;
    static const int nn[8] = {0, 2, 3, 4, 4, 8, 6, 5};
    int c = 0, b = 0;
    for (int e = 0; e < ne; ++e) {
        int id, type, ntag, tag[10], v[8];
        in >> id >> type >> ntag;
        if (!in || type < 2 || type > 7 || ntag < 2 || ntag > 3) {
            cfd::last_error() = "unsupported element";
            return -1;
        }
        for (int t = 0; t < ntag; ++t) in >> tag[t];
        for (int p = 0; p < nn[type]; ++p) { in >> v[p]; v[p]--; }
        if (type == 2 || type == 3) {
            if (bnd_vert) {
                int *q = bnd_vert + 4 * b;
                q[0] = v[0]; q[1] = v[1]; q[2] = v[2]; q[3] = type == 2 ? v[2] : v[3];
                bnd_rid[b] = tag[0];
            }
            ++b;
        } else {
            if (cell_vert) {
                int *q = cell_vert + 8 * c;
                if (type == 4) { int t[8] = {v[0], v[1], v[2], v[2], v[3], v[3], v[3], v[3]}; std::memcpy(q, t, sizeof t); }
                else if (type == 5) std::memcpy(q, v, sizeof(int) * 8);
                else if (type == 6) { int t[8] = {v[0], v[1], v[2], v[2], v[3], v[4], v[5], v[5]}; std::memcpy(q, t, sizeof t); }
                else { int t[8] = {v[0], v[1], v[2], v[3], v[4], v[4], v[4], v[4]}; std::memcpy(q, t, sizeof t); }
                cell_rid[c] = tag[0];
            }
            ++c;
        }
    }
    if (nvrt) *nvrt = nv;
    if (ncel) *ncel = c;
    if (nbnd) *nbnd = b;
    return 0;
}

int cfd_gmsh_read_counts(const char *path, int *nvrt, int *ncel, int *nbnd) {
    return gmsh_read(path, nvrt, ncel, nbnd, nullptr, nullptr, nullptr, nullptr, nullptr);
}
int cfd_gmsh_read(const char *path, double *vert, int *cell_vert, int *cell_rid, int *bnd_vert,
                  int *bnd_rid) {
    return gmsh_read(path, nullptr, nullptr, nullptr, vert, cell_vert, cell_rid, bnd_vert, bnd_rid);
}

cfd_mesh *cfd_mesh_build(const cfd_raw_mesh *raw) {
    cfd_mesh *h = new cfd_mesh;
    std::string err;
    if (cfd::load_raw(*raw, h->m, err) || cfd::build_faces(h->m, err) ||
        cfd::build_geometry(h->m, err)) {
        cfd::last_error() = err;
        delete h;
        return nullptr;
    }
    cfd::colour_cells(h->m);
    return h;
}
cfd_mesh *cfd_mesh_build_part(const cfd_raw_mesh *raw, const cfd_params *p, int part) {
    cfd_mesh *h = new cfd_mesh;
    HostMesh fluid, solid;
    bool has_solid;
    int cid;
    std::string err;
    if (cfd::build_parts(*raw, *p, fluid, solid, has_solid, cid, err)) {
        cfd::last_error() = err;
        delete h;
        return nullptr;
    }
    if (part == 1 && !has_solid) {
        cfd::last_error() = "no solid part without conjugate heat transfer";
        delete h;
        return nullptr;
    }
    h->m = part == 1 ? std::move(solid) : std::move(fluid);
    return h;
}
void cfd_mesh_free(cfd_mesh *h) { delete h; }
int cfd_mesh_counts(const cfd_mesh *h, int c[5]) {
    c[0] = h->m.nvrt; c[1] = h->m.ncel; c[2] = h->m.nfac; c[3] = h->m.nbnd; c[4] = h->m.ncb;
    return 0;
}

}  // extern "C"

namespace cfd {
int mesh_get_int(const HostMesh &m, const char *name, int *out) {
    std::string s(name);
    const std::vector<int> *v = nullptr;
    if (s == "face_vert") v = &m.face_vert;
    else if (s == "face_cell1") v = &m.face_c1;
    else if (s == "face_cell2") v = &m.face_c2;
    else if (s == "face_bnd") v = &m.face_bnd;
    else if (s == "cell_nface") v = &m.cell_nface;
    else if (s == "cell_face") v = &m.cell_face;
    else if (s == "cell_cell") v = &m.cell_cell;
    else if (s == "cell_vert") v = &m.cell_vert;
    else if (s == "cell_rid") v = &m.cell_rid;
    else if (s == "bnd_face") v = &m.bnd_face;
    else if (s == "bnd_rid") v = &m.bnd_rid;
    else if (s == "cell_colour") v = &m.colour;
    else if (s == "cell_bcolour") v = &m.bcolour;
    if (!v) { last_error() = "unknown int array " + s; return -1; }
    std::memcpy(out, v->data(), sizeof(int) * v->size());
    return 0;
}
int mesh_get_dbl(const HostMesh &m, const char *name, double *out) {
    std::string s(name);
    const std::vector<double> *v = nullptr;
    if (s == "vert") v = &m.vert;
    else if (s == "face_x") v = &m.face_x;
    else if (s == "face_n") v = &m.face_n;
    else if (s == "face_area") v = &m.face_area;
    else if (s == "face_lambda") v = &m.face_lambda;
    else if (s == "face_rlencos") v = &m.face_rlencos;
    else if (s == "face_xpac") v = &m.face_xpac;
    else if (s == "face_xnac") v = &m.face_xnac;
    else if (s == "cell_x") v = &m.cell_x;
    else if (s == "cell_vol") v = &m.cell_vol;
    else if (s == "bnd_distance") v = &m.bnd_dist;
    if (!v) { last_error() = "unknown double array " + s; return -1; }
    std::memcpy(out, v->data(), sizeof(double) * v->size());
    return 0;
}
}  // namespace cfd

extern "C" {
int cfd_mesh_get_int(const cfd_mesh *h, const char *name, int *out) {
    return cfd::mesh_get_int(h->m, name, out);
}
int cfd_mesh_get_dbl(const cfd_mesh *h, const char *name, double *out) {
    return cfd::mesh_get_dbl(h->m, name, out);
}
}
