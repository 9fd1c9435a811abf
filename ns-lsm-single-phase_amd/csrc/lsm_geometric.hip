// Geometric (hybrid) reinitialisation, hybrid_reinitialisation = .true. and
// GlobalMassCorrection = .false. (the reference's default configuration):
// geometric_reinitialisation.F90:1-93 with GetTetrahedraList.F90,
// GetIsosurfaceAndDistances.F90, NeighboursPhiCorrection.F90 and the
// marching-tetrahedra / distance routines of TetrahedronMethods.F90.
//
// The reference builds lists (interface nodes, their K-tetrahedra, the
// tetrahedra of each node) and walks them node by node.  Here the same
// quantities are positional: a cell knows its K-tetrahedra (bit mask) and the
// slot of their isosurfaces, a node its interface / neighbour flag, and the
// K-tetrahedra of a node are enumerated from its eight cells in the order the
// reference appends them (cells in k, j, i order, tetrahedra 1..6), so every
// running minimum sees its candidates in the reference's order and the
// results are bit-identical.
//
//   geo_cells     per cell: K-tetrahedra and their vertices (pnodes), slot
//   geo_iso       per cell with K-tetrahedra: their isosurface triangles
//   geo_nodes     per node: interface node (a vertex of a K-tetrahedron)
//   geo_mark      per node: neighbour (an interface node within 4 nodes)
//   geo_distance  per interface node: signed distance to the isosurfaces
//                 of the K-tetrahedra of the interface nodes within 4 nodes;
//                 per neighbour node: min(|phicorrected|, that distance)
#include <mutex>

#include "lsm_internal.hpp"

namespace lsm {
namespace {

// cell vertex q -> node offset (i, j, k), GetTetrahedraList.F90:240-247
__constant__ int8_t kCellVtx[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                      {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
// tetrahedra of a cell, TriangulationCase 1 (GetTetrahedraList.F90:118-123)
__constant__ int8_t kCellTets[6][4] = {{0, 5, 1, 6}, {0, 1, 2, 6}, {0, 2, 3, 6},
                                       {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 4, 5, 6}};
// cell vertices of each tetrahedron as a bit mask
__constant__ uint8_t kTetHas[6] = {0x63, 0x47, 0x4d, 0xc9, 0xd1, 0x71};
// tetrahedron edges and the marching-tetrahedra tables, TetrahedronMethods.F90:71-135
__constant__ int8_t kTetEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
__constant__ uint8_t kTetEdgeFlags[16] = {0x00, 0x0d, 0x13, 0x1e, 0x26, 0x2b, 0x35, 0x38,
                                          0x38, 0x35, 0x2b, 0x26, 0x1e, 0x13, 0x0d, 0x00};
__constant__ int8_t kTetTriangles[16][6] = {
    {-1, -1, -1, -1, -1, -1}, {0, 3, 2, -1, -1, -1}, {0, 1, 4, -1, -1, -1}, {1, 4, 2, 2, 4, 3},
    {1, 2, 5, -1, -1, -1}, {0, 3, 5, 0, 5, 1}, {0, 2, 5, 0, 5, 4}, {5, 4, 3, -1, -1, -1},
    {3, 4, 5, -1, -1, -1}, {4, 5, 0, 5, 2, 0}, {1, 5, 0, 5, 3, 0}, {5, 2, 1, -1, -1, -1},
    {3, 4, 2, 2, 4, 1}, {4, 1, 0, -1, -1, -1}, {2, 3, 0, -1, -1, -1}, {-1, -1, -1, -1, -1, -1}};

constexpr double kTiny = 2.2250738585072014e-308;   // tiny(eps_sims)

struct Geom {
    int ni, nj, nk;
    long sj, sk;
    long csj, csk;                 // cell strides: (ni-1), (ni-1)(nj-1)
    const double *x, *y, *z, *phi0;
};

// NORM2 as the Fortran runtime evaluates it for REAL(8): the largest |x| so
// far times sqrt(1 + sum of the others' squared ratios)
__device__ __forceinline__ double fnorm2(double a0, double a1, double a2)
{
    const double v[3] = {a0, a1, a2};
    double mx = 0.0, sum = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double a = fabs(v[q]);
        if (mx == 0.0) {
            mx = a;
        } else if (a > mx) {
            const double t = mx / a, tsq = t * t;
            sum = sum * tsq;
            sum = sum + tsq;
            mx = a;
        } else {
            const double t = a / mx;
            sum = sum + t * t;
        }
    }
    return mx * sqrt(1.0 + sum);
}

struct V3 {
    double x, y, z;
};

__device__ __forceinline__ V3 sub(const V3 &a, const V3 &b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot(const V3 &a, const V3 &b) { return ((0.0 + a.x * b.x) + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross(const V3 &a, const V3 &b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double nrm(const V3 &a) { return fnorm2(a.x, a.y, a.z); }

// MINVAL order: a later candidate replaces the minimum only when smaller
__device__ __forceinline__ double fmin_keep(double r, double x) { return x < r ? x : r; }

// DistancePointLine, TetrahedronMethods.F90:1356-1399
__device__ double dist_point_line(const V3 &A, const V3 &B, const V3 &P)
{
    const V3 AB = sub(B, A), AP = sub(P, A), BA = sub(A, B), BP = sub(P, B);
    bool valid = true;
    if (nrm(AB) < kEpsSims) valid = false;
    if (dot(AB, AP) < 0.0) valid = false;
    if (dot(BA, BP) < 0.0) valid = false;
    if (valid) return nrm(cross(AP, BP)) / nrm(AB);
    const double da = nrm(AP), db = nrm(BP);
    return db < da ? db : da;
}

// DistancePointTriangle with CheckProjectionPointOnATriangle3,
// TetrahedronMethods.F90:1008-1070,1593-1684 (the 'projection' weights P1,
// P2, P3 with alpha, beta, gamma as the reference forms them)
__device__ double dist_point_triangle(const double *T, const V3 &P)
{
    const V3 A{T[0], T[1], T[2]}, B{T[3], T[4], T[5]}, C{T[6], T[7], T[8]};
    double r = nrm(sub(P, A));
    r = fmin_keep(r, nrm(sub(P, B)));
    r = fmin_keep(r, nrm(sub(P, C)));
    r = fmin_keep(r, dist_point_line(A, B, P));
    r = fmin_keep(r, dist_point_line(B, C, P));
    r = fmin_keep(r, dist_point_line(C, A, P));
    if (nrm(sub(A, B)) <= kEpsSims || nrm(sub(A, C)) <= kEpsSims || nrm(sub(B, C)) <= kEpsSims) return r;
    const V3 u = sub(B, A), v = sub(C, A), w = sub(P, A);
    const V3 N = cross(u, v);
    const double alpha = dot(cross(u, w), N) / dot(N, N);
    const double beta = dot(cross(w, v), N) / dot(N, N);
    const double gamma = (1.0 - beta) - alpha;
    const V3 pp{(alpha * A.x + beta * B.x) + gamma * C.x, (alpha * A.y + beta * B.y) + gamma * C.y,
                (alpha * A.z + beta * B.z) + gamma * C.z};
    const double dpp = nrm(sub(P, pp));
    if (alpha > 0.0 && alpha < 1.0 && beta > 0.0 && beta < 1.0 && gamma > 0.0 && gamma < 1.0)
        r = fmin_keep(r, dpp);
    return r;
}

// GetTriangleArea, TetrahedronMethods.F90:691-707
__device__ __forceinline__ double triangle_area(const double *T)
{
    const V3 A{T[0], T[1], T[2]}, B{T[3], T[4], T[5]}, C{T[6], T[7], T[8]};
    return 0.5 * nrm(cross(sub(B, A), sub(C, A)));
}

// CellPnodeTetrahedronIdentification, TetrahedronMethods.F90:2124-2361:
// bits 0..5 = K-tetrahedra (phase change), *pn = their cell vertices
__device__ int cell_ktets(const double ph[8], int &pn)
{
    pn = 0;
    bool all_pos = true, all_neg = true;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (!(ph[q] > kEpsSims)) all_pos = false;
        if (!(ph[q] < -kEpsSims)) all_neg = false;
    }
    if (all_pos || all_neg) return 0;
    bool pcell = false;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (fabs(ph[q]) < kEpsSims) pcell = true;
    if (!pcell) {
        const double sg = copysign(1.0, ph[0]);
        for (int q = 1; q < 8; ++q)
            if (sg * ph[q] < 0.0) { pcell = true; break; }
    }
    if (!pcell) return 0;
    int tets = 0;
    for (int t = 0; t < 6; ++t) {
        bool pct = false, second = true;
        double pa = ph[kCellTets[t][0]];
        if (pa <= kEpsSims) second = false;
        if (fabs(pa) < kEpsSims) pct = true;
        const double sg = copysign(1.0, pa);
        for (int q = 1; q < 4; ++q) {
            pa = ph[kCellTets[t][q]];
            if (sg * pa < 0.0 || fabs(pa) < kEpsSims) {
                if (pa <= kEpsSims) second = false;
                pct = true;
            }
        }
        if (pct && !second) {
            tets |= 1 << t;
            pn |= kTetHas[t];
        }
    }
    return tets;
}

__device__ __forceinline__ void cell_of(const Geom &g, long c, int &i, int &j, int &k)
{
    i = (int)(c % g.csj);
    const long r = c / g.csj;
    j = (int)(r % (g.nj - 1));
    k = (int)(r / (g.nj - 1));
}

__global__ __launch_bounds__(256) void geo_cells(Geom g, long ncell, uint8_t *__restrict__ ctets,
                                                 uint8_t *__restrict__ cpn, int32_t *__restrict__ cslot,
                                                 int32_t *__restrict__ count)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int i, j, k;
    cell_of(g, c, i, j, k);
    double ph[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
        ph[q] = g.phi0[(long)(i + kCellVtx[q][0]) + g.sj * (j + kCellVtx[q][1]) + g.sk * (k + kCellVtx[q][2])];
    int pn;
    const int tm = cell_ktets(ph, pn);
    ctets[c] = (uint8_t)tm;
    cpn[c] = (uint8_t)pn;
    cslot[c] = tm ? atomicAdd(count, 1) : -1;
}

// marching-tetrahedra isosurface (TetrahedronMethods.F90:26-664): the number
// of triangles whose distances are used (ntriangles > 0 and area > 0, as
// GetIsosurfaceAndDistances.F90:182-184 filters them) and their vertices
__device__ int tet_isosurface(const double v[4][3], const double ph[4], double *tri)
{
    int sg[4], npos = 0, nneg = 0, nzero = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        sg[q] = 0;
        if (ph[q] > kTiny) sg[q] = 1;
        if (ph[q] < -kTiny) sg[q] = -1;
        npos += sg[q] == 1; nneg += sg[q] == -1; nzero += sg[q] == 0;
    }
    if (npos == 4 || nneg == 4) return 0;
    if (nzero > 0 && (nneg == 0 || npos == 0)) {
        if (nzero != 3) return 0;            // a vertex or an edge (ntriangles < 0), or no isosurface
        int r = 0;
        for (int q = 0; q < 4; ++q)
            if (sg[q] == 0) { tri[3 * r] = v[q][0]; tri[3 * r + 1] = v[q][1]; tri[3 * r + 2] = v[q][2]; ++r; }
        return triangle_area(tri) > 0.0 ? 1 : 0;
    }
    int flag = 0;
    for (int q = 0; q < 4; ++q)
        if (sg[q] == -1) flag |= 1 << q;
    const int ef = kTetEdgeFlags[flag];
    if (ef == 0) return 0;
    double E[6][3];
    for (int e = 0; e < 6; ++e) {
        E[e][0] = E[e][1] = E[e][2] = -9.9;
        if (!(ef & (1 << e))) continue;
        const int s0 = kTetEdge[e][0], s1 = kTetEdge[e][1];
        const double den = ph[s1] - ph[s0];
        double s = 0.0;
        if (!(fabs(den) < kEpsSims * kEpsSims * kEpsSims)) s = ph[s1] / den;
        for (int a = 0; a < 3; ++a) E[e][a] = v[s1][a] + s * (v[s0][a] - v[s1][a]);
    }
    int ntri = 0;
    double area = 0.0;
    for (int t = 0; t < 2; ++t) {
        if (kTetTriangles[flag][3 * t] < 0) break;
        ++ntri;
        for (int q = 0; q < 3; ++q)
            for (int a = 0; a < 3; ++a) tri[9 * t + 3 * q + a] = E[kTetTriangles[flag][3 * t + q]][a];
        area = area + triangle_area(tri + 9 * t);
    }
    return area > 0.0 ? ntri : 0;
}

// bounding sphere (centroid, radius rounded up) and shortest edge (rounded
// down) of a triangle: the distance bound that lets the search skip it
__device__ void tri_bounds(const double *T, double *bd)
{
    const double cx = (T[0] + T[3] + T[6]) / 3.0, cy = (T[1] + T[4] + T[7]) / 3.0, cz = (T[2] + T[5] + T[8]) / 3.0;
    double r = 0.0;
    for (int q = 0; q < 3; ++q) {
        const double dx = T[3 * q] - cx, dy = T[3 * q + 1] - cy, dz = T[3 * q + 2] - cz;
        r = fmax(r, sqrt(dx * dx + dy * dy + dz * dz));
    }
    double emin = 1e300;
    for (int q = 0; q < 3; ++q) {
        const int p2 = (q + 1) % 3;
        const double dx = T[3 * q] - T[3 * p2], dy = T[3 * q + 1] - T[3 * p2 + 1], dz = T[3 * q + 2] - T[3 * p2 + 2];
        emin = fmin(emin, sqrt(dx * dx + dy * dy + dz * dz));
    }
    bd[0] = cx; bd[1] = cy; bd[2] = cz;
    bd[3] = r * (1.0 + 1e-12);
    bd[4] = emin * (1.0 - 1e-12);
}

__global__ __launch_bounds__(256) void geo_iso(Geom g, long ncell, const uint8_t *__restrict__ ctets,
                                               const int32_t *__restrict__ cslot, int8_t *__restrict__ nuse,
                                               double *__restrict__ tris, double *__restrict__ bnd)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int tm = ctets[c];
    if (!tm) return;
    int i, j, k;
    cell_of(g, c, i, j, k);
    const long slot = (long)cslot[c] * 6;
    for (int t = 0; t < 6; ++t) {
        if (!(tm & (1 << t))) continue;
        double v[4][3], ph[4];
        for (int q = 0; q < 4; ++q) {
            const int vq = kCellTets[t][q];
            const long l = (long)(i + kCellVtx[vq][0]) + g.sj * (j + kCellVtx[vq][1]) + g.sk * (k + kCellVtx[vq][2]);
            v[q][0] = g.x[l]; v[q][1] = g.y[l]; v[q][2] = g.z[l];
            ph[q] = g.phi0[l];
        }
        double tri[18];
        const int nu = tet_isosurface(v, ph, tri);
        const long rec = slot + t;
        nuse[rec] = (int8_t)nu;
        for (int q = 0; q < 9 * nu; ++q) tris[rec * 18 + q] = tri[q];
        for (int u = 0; u < nu; ++u) tri_bounds(tri + 9 * u, bnd + rec * 10 + 5 * u);
    }
}

__device__ __forceinline__ void node_of(const Geom &g, long l, int &i, int &j, int &k)
{
    i = (int)(l % g.ni);
    const long r = l / g.ni;
    j = (int)(r % g.nj);
    k = (int)(r / g.nj);
}

// appends the flagged lanes' values to list (one atomic per wavefront); the
// whole wavefront must call it
__device__ __forceinline__ int wave_append(bool flag, int value, int32_t *list, int32_t *count)
{
    const unsigned long long bal = __ballot(flag);
    if (!bal) return -1;
    const int lane = threadIdx.x & 63;
    const int first = __ffsll((long long)bal) - 1;
    int base = 0;
    if (lane == first) base = atomicAdd(count, (int32_t)__popcll(bal));
    base = __shfl(base, first);
    if (!flag) return -1;
    const int pos = base + (int)__popcll(bal & ((1ull << lane) - 1ull));
    list[pos] = value;
    return pos;
}

// a node is an interface node when it is a vertex of a K-tetrahedron of one
// of its cells (GetTetrahedraList.F90:262-269); interface nodes are listed
// and mid maps each to its list position
__global__ __launch_bounds__(256) void geo_nodes(Geom g, long n, const uint8_t *__restrict__ cpn,
                                                 int8_t *__restrict__ nid, int32_t *__restrict__ mid,
                                                 int32_t *__restrict__ ilist, int32_t *__restrict__ count)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool f = false;
    if (l < n) {
        int i, j, k;
        node_of(g, l, i, j, k);
        for (int dk = -1; dk <= 0; ++dk)
            for (int dj = -1; dj <= 0; ++dj)
                for (int di = -1; di <= 0; ++di) {
                    const int ci = i + di, cj = j + dj, ck = k + dk;
                    if (ci < 0 || cj < 0 || ck < 0 || ci > g.ni - 2 || cj > g.nj - 2 || ck > g.nk - 2) continue;
                    const int pn = cpn[(long)ci + g.csj * cj + g.csk * ck];
                    // vertex index of this node in that cell
                    const int ox = -di, oy = -dj, oz = -dk;
                    const int vq = oz * 4 + (oy ? (ox ? 2 : 3) : ox);
                    if (pn & (1 << vq)) f = true;
                }
        nid[l] = f ? 1 : 0;
    }
    const int pos = wave_append(f, (int)l, ilist, count);
    if (f) mid[l] = pos;
}

__device__ __forceinline__ void box(int c, int n, int &lo, int &hi)
{
    lo = c - 4 < 0 ? 0 : c - 4;
    hi = c + 4 > n - 1 ? n - 1 : c + 4;
}

// neighbours: the nodes the search boxes of GetIsosurfaceAndDistances.F90:153-240
// visit without being interface nodes (InterfaceNodesID = -1), listed
__global__ __launch_bounds__(256) void geo_mark(Geom g, long n, const int8_t *__restrict__ nid,
                                                uint8_t *__restrict__ nmark, int32_t *__restrict__ nlist,
                                                int32_t *__restrict__ count)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool m = false;
    if (l < n) {
        if (nid[l] == 0) {
            int i, j, k;
            node_of(g, l, i, j, k);
            int i0, i1, j0, j1, k0, k1;
            box(i, g.ni, i0, i1); box(j, g.nj, j0, j1); box(k, g.nk, k0, k1);
            for (int ks = k0; ks <= k1 && !m; ++ks)
                for (int js = j0; js <= j1 && !m; ++js)
                    for (int is = i0; is <= i1; ++is)
                        if (nid[(long)is + g.sj * js + g.sk * ks] == 1) { m = true; break; }
        }
        nmark[l] = m ? 1 : 0;
    }
    (void)wave_append(m, (int)l, nlist, count);
}

// per interface node: its K-tetrahedra in the order GetTetrahedraList.F90:346-418
// appends them (cells in k, j, i order, tetrahedra 1..6), their number, and
// a sphere about the node holding all their triangles (radius up, shortest
// edge down)
__global__ __launch_bounds__(256) void geo_iface(Geom g, long nint, const int32_t *__restrict__ ilist,
                                                 const uint8_t *__restrict__ ctets, const int32_t *__restrict__ cslot,
                                                 const int8_t *__restrict__ nuse, const double *__restrict__ bnd,
                                                 int32_t *__restrict__ mrec, int8_t *__restrict__ nkt,
                                                 double *__restrict__ mrad, double *__restrict__ memin)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nint) return;
    const long l = ilist[m];
    int i, j, k;
    node_of(g, l, i, j, k);
    const double X = g.x[l], Y = g.y[l], Z = g.z[l];
    int q = 0;
    double R = 0.0, em = 1e300;
    bool any = false;
    for (int dk = -1; dk <= 0; ++dk)
        for (int dj = -1; dj <= 0; ++dj)
            for (int di = -1; di <= 0; ++di) {
                const int ci = i + di, cj = j + dj, ck = k + dk;
                if (ci < 0 || cj < 0 || ck < 0 || ci > g.ni - 2 || cj > g.nj - 2 || ck > g.nk - 2) continue;
                const long c = (long)ci + g.csj * cj + g.csk * ck;
                const int tm = ctets[c];
                if (!tm) continue;
                const int ox = -di, oy = -dj, oz = -dk;
                const int vq = oz * 4 + (oy ? (ox ? 2 : 3) : ox);
                const long slot = (long)cslot[c] * 6;
                for (int t = 0; t < 6; ++t) {
                    if (!((tm >> t) & 1) || !((kTetHas[t] >> vq) & 1)) continue;
                    const long rec = slot + t;
                    mrec[m * 24 + q] = (int32_t)rec;
                    ++q;
                    const int nu = nuse[rec];
                    for (int u = 0; u < nu; ++u) {
                        const double *bd = bnd + rec * 10 + 5 * u;
                        const double dx = bd[0] - X, dy = bd[1] - Y, dz = bd[2] - Z;
                        R = fmax(R, sqrt(dx * dx + dy * dy + dz * dz) + bd[3]);
                        em = fmin(em, bd[4]);
                        any = true;
                    }
                }
            }
    nkt[m] = (int8_t)q;
    mrad[m] = any ? R * (1.0 + 1e-12) : -1.0;   // -1: no usable triangle
    memin[m] = em;
}

struct DistArgs {
    Geom g;
    long nint, nnb;
    const int32_t *ilist, *nlist, *mid;
    const int8_t *nid;
    const int32_t *mrec;
    const int8_t *nkt;
    const double *mrad, *memin;
    const int8_t *nuse;
    const double *tris, *bnd;
    double *pc;
};

// True when every distance the reference's DistancePointTriangle can return
// for points of a sphere (centre c, radius r, shortest triangle edge emin)
// exceeds cmp: |P - c| - r bounds the exact distance from below, and the
// computed distance is within err of the exact one -- rounding of the
// coordinates (cmax), of the norms and of the cross-product line distance,
// whose error grows like |AP| |BP| / |AB| (L^2 / emin) -- with a margin of
// more than three orders of magnitude.  A triangle with a vanishing edge
// (emin = 0) is never skipped.
__device__ __forceinline__ bool beyond(double cx, double cy, double cz, double r, double emin, const V3 &P,
                                       double cmax, double cmp)
{
    const double dx = P.x - cx, dy = P.y - cy, dz = P.z - cz;
    const double dP = sqrt(dx * dx + dy * dy + dz * dz);
    const double L = dP + r;
    const double err = 1e-12 * (L + L * (L / emin) + cmax + L);
    return dP - r - err > cmp;
}

__device__ __forceinline__ double sd_update(double SD, double D, double sgn0)
{
    const double ex2 = (copysign(1.0, fabs(SD) - D) + 1.0) / 2.0;
    return sgn0 * (fabs(SD) + ex2 * (D - fabs(SD)));
}

__device__ __forceinline__ double pc_update(double pc, double D)
{
    const double ex2 = (copysign(1.0, fabs(pc) - D) + 1.0) / 2.0;
    const double aux = fabs(pc) + ex2 * (D - fabs(pc));
    return copysign(1.0, pc) * aux;
}

// One thread per listed node (interface nodes, then neighbours).  The node
// walks the sequence of GetIsosurfaceAndDistances.F90:153-240 (interface
// nodes: D and SD carried over the whole search box) or
// NeighboursPhiCorrection.F90:133-249 (neighbours: D restarted per interface
// node, phicorrected updated after every K-tetrahedron): the interface nodes
// of its box in k, j, i order, their K-tetrahedra, their triangles.  A
// triangle whose distance provably exceeds the running minimum D leaves D
// unchanged and is not evaluated; an interface node none of whose triangles
// can reach D (neighbours: |phicorrected|) leaves everything but the
// repeated SD updates unchanged.  Every update that does happen sees the
// values and the order of the reference, so the results are bit-identical.
// The walk is a cursor: lanes advance through the cheap steps until each has
// a triangle to evaluate, then evaluate together.
__global__ __launch_bounds__(256) void geo_distance(DistArgs a)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nint + a.nnb) return;
    const Geom &g = a.g;
    const bool iface = t < a.nint;
    const long l = iface ? a.ilist[t] : a.nlist[t - a.nint];
    int i, j, k;
    node_of(g, l, i, j, k);
    const V3 P{g.x[l], g.y[l], g.z[l]};
    const double cmax = fmax(fabs(P.x), fmax(fabs(P.y), fabs(P.z)));
    int i0, i1, j0, j1, k0, k1;
    box(i, g.ni, i0, i1); box(j, g.nj, j0, j1); box(k, g.nk, k0, k1);
    const double sgn0 = copysign(1.0, g.phi0[l]);
    double D = 10.0, SD = 99.9;
    double pc = iface ? 0.0 : a.pc[l];
    int is = i0, js = j0, ks = k0;
    bool box_left = true;
    long m = -1;
    int q = 0, nk = 0, tri = 0, nu = 0;
    long rec = -1;
    bool done = false;
    while (!done) {
        bool need = false;
        for (;;) {
            if (rec >= 0) {
                if (tri < nu) {
                    const double *bd = a.bnd + rec * 10 + 5 * tri;
                    if (beyond(bd[0], bd[1], bd[2], bd[3], bd[4], P, cmax, D)) { ++tri; continue; }
                    need = true;
                    break;
                }
                // end of a K-tetrahedron (GetIsosurfaceAndDistances.F90:221-229,
                // NeighboursPhiCorrection.F90:203-215)
                if (iface) SD = sd_update(SD, D, sgn0);
                else pc = pc_update(pc, D);
                rec = -1;
            }
            if (q < nk) {
                rec = a.mrec[m * 24 + q];
                ++q;
                tri = 0;
                nu = a.nuse[rec];
                continue;
            }
            if (!box_left) { done = true; break; }
            const long ls = (long)is + g.sj * js + g.sk * ks;
            if (++is > i1) {
                is = i0;
                if (++js > j1) {
                    js = j0;
                    if (++ks > k1) box_left = false;
                }
            }
            if (a.nid[ls] != 1) continue;
            const long mm = a.mid[ls];
            const int cnt = a.nkt[mm];
            if (!iface) D = 10.0;
            // a neighbour's updates with D >= 10 leave |phicorrected| <= 10
            // unchanged; D dips below a distance by at most a rounding of 10
            const double R = a.mrad[mm];
            bool skip;
            if (iface)
                skip = R < 0.0 || beyond(g.x[ls], g.y[ls], g.z[ls], R, a.memin[mm], P, cmax, D);
            else
                skip = fabs(pc) <= 10.0 &&
                       (R < 0.0 || beyond(g.x[ls], g.y[ls], g.z[ls], R, a.memin[mm], P, cmax, fabs(pc) + 1e-14));
            if (skip) {
                // D stays; SD takes its per-tetrahedron update cnt times
                if (iface)
                    for (int c = 0; c < cnt; ++c) {
                        const double s2 = sd_update(SD, D, sgn0);
                        if (s2 == SD) break;
                        SD = s2;
                    }
                continue;
            }
            m = mm;
            q = 0;
            nk = cnt;
        }
        if (need) {
            const double Da = dist_point_triangle(a.tris + rec * 18 + 9 * tri, P);
            const double ex = (copysign(1.0, D - Da) + 1.0) / 2.0;
            D = D + ex * (Da - D);
            ++tri;
        }
    }
    // NeighboursPhiCorrection.F90:89-130 with C = 0: phi = SD + 0
    a.pc[l] = iface ? SD + 0.0 : pc;
}

__global__ void geo_ids(long n, const int8_t *__restrict__ nid, const uint8_t *__restrict__ nmark,
                        int32_t *__restrict__ ids)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    ids[l] = nid[l] == 1 ? 1 : (nmark[l] ? -1 : 0);
}

__global__ void geo_zero(int32_t *c)
{
    if (threadIdx.x < 4) c[threadIdx.x] = 0;
}

__global__ void hybrid_init(const double *__restrict__ phi, const double *__restrict__ base,
                            double *__restrict__ pc, long n)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const double p = phi[l];
    double v = base ? base[l] : 0.0;
    if (p < -kEpsSims) v = -9.9;
    if (p > kEpsSims) v = 9.9;
    pc[l] = v;
}

// the isosurface store: one per device, grown on demand, kept until exit
std::mutex g_store_mu;
void *g_store[64];
size_t g_store_size[64];

int store_take(size_t bytes, void **p)
{
    int dev = 0;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= 64) return fail(LSM_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_store_mu);
    if (g_store_size[dev] < bytes) {
        if (g_store[dev]) {
            if ((rc = check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"))) return rc;
            (void)hipFree(g_store[dev]);
            g_store[dev] = nullptr;
            g_store_size[dev] = 0;
        }
        const size_t want = bytes + bytes / 4 + 4096;
        if ((rc = check_hip(hipMalloc(&g_store[dev], want), "hipMalloc(isosurface store)"))) return rc;
        g_store_size[dev] = want;
    }
    *p = g_store[dev];
    return LSM_OK;
}

unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

size_t geometric_work(const Blk &b)
{
    const size_t nc = (size_t)(b.e.ni - 1) * (b.e.nj - 1) * (b.e.nk - 1);
    const size_t n = (size_t)b.n;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    return al(n) + al(n) + al(nc) + al(nc) + al(nc * 4) + al(n * 4) * 3 + 256 + 2048;
}

int launch_hybrid_init(const double *phi, const double *base, double *pc, long n, hipStream_t s)
{
    hybrid_init<<<nblk(n), 256, 0, s>>>(phi, base, pc, n);
    return check_hip(hipGetLastError(), "hybrid_init");
}

int launch_geometric(const Blk &b, const double *phi0, double *pc, int32_t *ids, Carver &c, hipStream_t s)
{
    if (!b.x || !b.y || !b.z) return fail(LSM_ERR_ARG, "the geometric reinitialisation needs x, y, z");
    const long n = b.n;
    const long ncell = (long)(b.e.ni - 1) * (b.e.nj - 1) * (b.e.nk - 1);
    int8_t *nid = (int8_t *)c.take(n);
    uint8_t *nmark = (uint8_t *)c.take(n);
    uint8_t *ctets = (uint8_t *)c.take(ncell), *cpn = (uint8_t *)c.take(ncell);
    int32_t *cslot = (int32_t *)c.take(ncell * 4);
    int32_t *mid = (int32_t *)c.take(n * 4), *ilist = (int32_t *)c.take(n * 4), *nlist = (int32_t *)c.take(n * 4);
    int32_t *count = (int32_t *)c.take(256);
    if (!nid || !nmark || !ctets || !cpn || !cslot || !mid || !ilist || !nlist || !count)
        return fail(LSM_ERR_WORKSPACE, "workspace too small");
    if (n > 0x7fffffffL) return fail(LSM_ERR_ARG, "block of %ld nodes: at most 2^31 - 1", n);
    Geom g{b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, (long)(b.e.ni - 1), (long)(b.e.ni - 1) * (b.e.nj - 1),
           b.x, b.y, b.z, phi0};
    int rc;
    // counts: cells with K-tetrahedra, interface nodes, neighbours
    geo_zero<<<1, 64, 0, s>>>(count);
    if ((rc = check_hip(hipGetLastError(), "geo_zero"))) return rc;
    geo_cells<<<nblk(ncell), 256, 0, s>>>(g, ncell, ctets, cpn, cslot, count);
    if ((rc = check_hip(hipGetLastError(), "geo_cells"))) return rc;
    geo_nodes<<<nblk(n), 256, 0, s>>>(g, n, cpn, nid, mid, ilist, count + 1);
    if ((rc = check_hip(hipGetLastError(), "geo_nodes"))) return rc;
    geo_mark<<<nblk(n), 256, 0, s>>>(g, n, nid, nmark, nlist, count + 2);
    if ((rc = check_hip(hipGetLastError(), "geo_mark"))) return rc;
    // the counts size the isosurface store and the distance launch
    int32_t cnt[3] = {0, 0, 0};
    if ((rc = check_hip(hipMemcpyAsync(cnt, count, sizeof(cnt), hipMemcpyDeviceToHost, s), "D2H")) ||
        (rc = check_hip(hipStreamSynchronize(s), "hipStreamSynchronize")))
        return rc;
    const long nslots = cnt[0], nint = cnt[1], nnb = cnt[2];
    if (nint > 0) {
        const size_t nrec = (size_t)nslots * 6;
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t o_tris = al(nrec);
        const size_t o_bnd = o_tris + al(nrec * 18 * sizeof(double));
        const size_t o_mrec = o_bnd + al(nrec * 10 * sizeof(double));
        const size_t o_nkt = o_mrec + al((size_t)nint * 24 * sizeof(int32_t));
        const size_t o_rad = o_nkt + al((size_t)nint);
        const size_t o_emin = o_rad + al((size_t)nint * sizeof(double));
        const size_t total = o_emin + al((size_t)nint * sizeof(double));
        void *store = nullptr;
        if ((rc = store_take(total, &store))) return rc;
        char *sb = (char *)store;
        int8_t *nuse = (int8_t *)sb;
        double *tris = (double *)(sb + o_tris), *bnd = (double *)(sb + o_bnd);
        int32_t *mrec = (int32_t *)(sb + o_mrec);
        int8_t *nkt = (int8_t *)(sb + o_nkt);
        double *mrad = (double *)(sb + o_rad), *memin = (double *)(sb + o_emin);
        geo_iso<<<nblk(ncell), 256, 0, s>>>(g, ncell, ctets, cslot, nuse, tris, bnd);
        if ((rc = check_hip(hipGetLastError(), "geo_iso"))) return rc;
        geo_iface<<<nblk(nint), 256, 0, s>>>(g, nint, ilist, ctets, cslot, nuse, bnd, mrec, nkt, mrad, memin);
        if ((rc = check_hip(hipGetLastError(), "geo_iface"))) return rc;
        DistArgs a{g, nint, nnb, ilist, nlist, mid, nid, mrec, nkt, mrad, memin, nuse, tris, bnd, pc};
        {
            KTimer kt(LSM_KT_GEOMETRIC, s);
            geo_distance<<<nblk(nint + nnb), 256, 0, s>>>(a);
        }
        if ((rc = check_hip(hipGetLastError(), "geo_distance"))) return rc;
    }
    if (ids) {
        geo_ids<<<nblk(n), 256, 0, s>>>(n, nid, nmark, ids);
        if ((rc = check_hip(hipGetLastError(), "geo_ids"))) return rc;
    }
    return LSM_OK;
}

}  // namespace lsm
