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
//   geo_dilate,   per node: neighbour (an interface node within 4 nodes),
//   geo_mark      from the interface flags dilated along i, j, k
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
    Blank bl;                      // blanking regions (read by geo_cells only)
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

// Per-triangle record of the distance search (kTS doubles): the vertices,
// what DistancePointTriangle recomputes for every point although it depends on
// the triangle only (the edge norms, the normal and its square), and the
// bounds that let the search skip the triangle.
constexpr int kTS = 25;
enum {
    kA = 0, kB = 3, kC = 6,
    kN = 9, kNN = 12,
    kNAB = 13, kNBC = 14, kNCA = 15,   // NORM2(B-A), NORM2(C-B), NORM2(A-C)
    kInvN = 16,                        // 1/|N| of a well-shaped triangle, else 0
    kInvAB = 17, kInvBC = 18, kInvCA = 19,   // 1/|edge|^2
    kCx = 20, kCy = 21, kCz = 22, kR = 23, kEmin = 24   // sphere, shortest edge
};

// one edge of DistancePointLine (TetrahedronMethods.F90:1356-1399) folded
// into the running minimum: the perpendicular distance when the foot lies on
// the segment; otherwise the distance to the nearer end point, which is one
// of the vertex distances already in the minimum
__device__ __forceinline__ double edge_min(const V3 &X, const V3 &Y, const V3 &XP, const V3 &YP, double nXY,
                                           double r)
{
    if (nXY < kEpsSims) return r;
    if (dot(sub(Y, X), XP) < 0.0 || dot(sub(X, Y), YP) < 0.0) return r;
    return fmin_keep(r, nrm(cross(XP, YP)) / nXY);
}

// DistancePointTriangle with CheckProjectionPointOnATriangle3,
// TetrahedronMethods.F90:1008-1070,1593-1684 (the 'projection' weights P1,
// P2, P3 with alpha, beta, gamma as the reference forms them): the minimum
// of the same values, so the same result bit for bit
__device__ double tri_distance(const double *T, const V3 &P)
{
    const V3 A{T[kA], T[kA + 1], T[kA + 2]}, B{T[kB], T[kB + 1], T[kB + 2]}, C{T[kC], T[kC + 1], T[kC + 2]};
    const V3 AP = sub(P, A), BP = sub(P, B), CP = sub(P, C);
    double r = nrm(AP);
    r = fmin_keep(r, nrm(BP));
    r = fmin_keep(r, nrm(CP));
    r = edge_min(A, B, AP, BP, T[kNAB], r);
    r = edge_min(B, C, BP, CP, T[kNBC], r);
    r = edge_min(C, A, CP, AP, T[kNCA], r);
    if (T[kNAB] <= kEpsSims || T[kNCA] <= kEpsSims || T[kNBC] <= kEpsSims) return r;
    const V3 N{T[kN], T[kN + 1], T[kN + 2]};
    const double NN = T[kNN];
    const V3 u = sub(B, A), v = sub(C, A);
    const double alpha = dot(cross(u, AP), N) / NN;
    const double beta = dot(cross(AP, v), N) / NN;
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
    if (g.bl.n && g.bl.cell(i, j, k)) {
        // blanked cells are not searched (GetTetrahedraList.F90:209-227)
        ctets[c] = 0;
        cpn[c] = 0;
        cslot[c] = -1;
        return;
    }
    double ph[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
        ph[q] = g.phi0[(long)(i + kCellVtx[q][0]) + g.sj * (j + kCellVtx[q][1]) + g.sk * (k + kCellVtx[q][2])];
    int pn;
    const int tm = cell_ktets(ph, pn);
    ctets[c] = (uint8_t)tm;
    cpn[c] = (uint8_t)pn;
    cslot[c] = tm ? atomicAdd(count, __popc(tm)) : -1;   // first K-tetrahedron record
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

// the search record of a triangle (layout kTS above)
__device__ void tri_record(const double *tri, double *T)
{
    for (int q = 0; q < 9; ++q) T[q] = tri[q];
    const V3 A{tri[0], tri[1], tri[2]}, B{tri[3], tri[4], tri[5]}, C{tri[6], tri[7], tri[8]};
    const V3 u = sub(B, A), v = sub(C, A);
    const V3 N = cross(u, v);
    T[kN] = N.x; T[kN + 1] = N.y; T[kN + 2] = N.z;
    T[kNN] = dot(N, N);
    T[kNAB] = nrm(sub(B, A));
    T[kNBC] = nrm(sub(C, B));
    T[kNCA] = nrm(sub(A, C));
    // bounds: plain Euclidean arithmetic, rounded outward
    auto len2 = [](const V3 &a) { return a.x * a.x + a.y * a.y + a.z * a.z; };
    const V3 w = sub(C, B), e3 = sub(A, C);
    const double lu = sqrt(len2(u)), lv = sqrt(len2(v)), lw = sqrt(len2(w));
    const double nN = sqrt(len2(N));
    T[kInvN] = (nN > 0.0 && nN >= 1e-2 * lu * lv) ? 1.0 / nN : 0.0;
    T[kInvAB] = 1.0 / len2(u);
    T[kInvBC] = 1.0 / len2(w);
    T[kInvCA] = 1.0 / len2(e3);
    const double cx = (A.x + B.x + C.x) / 3.0, cy = (A.y + B.y + C.y) / 3.0, cz = (A.z + B.z + C.z) / 3.0;
    double r = 0.0;
    r = fmax(r, sqrt(len2(V3{A.x - cx, A.y - cy, A.z - cz})));
    r = fmax(r, sqrt(len2(V3{B.x - cx, B.y - cy, B.z - cz})));
    r = fmax(r, sqrt(len2(V3{C.x - cx, C.y - cy, C.z - cz})));
    T[kCx] = cx; T[kCy] = cy; T[kCz] = cz;
    T[kR] = r * (1.0 + 1e-12);
    T[kEmin] = fmin(lu, fmin(lv, lw)) * (1.0 - 1e-12);
}

__device__ __forceinline__ long ktet_rec(const int32_t *cslot, long c, int tm, int t)
{
    return (long)cslot[c] + __popc(tm & ((1 << t) - 1));
}

__global__ __launch_bounds__(256) void geo_iso(Geom g, long ncell, const uint8_t *__restrict__ ctets,
                                               const int32_t *__restrict__ cslot, int8_t *__restrict__ nuse,
                                               double *__restrict__ trec)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int tm = ctets[c];
    if (!tm) return;
    int i, j, k;
    cell_of(g, c, i, j, k);
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
        const long rec = ktet_rec(cslot, c, tm, t);
        nuse[rec] = (int8_t)nu;
        for (int u = 0; u < nu; ++u) tri_record(tri + 9 * u, trec + (rec * 2 + u) * kTS);
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
// visit without being interface nodes (InterfaceNodesID = -1).  The clamped
// 9^3 box is a product of three index ranges, so "an interface node in the
// box" is a dilation of the interface flags by 4 nodes along i, then j, then
// k; the last pass lists the neighbours.
__global__ __launch_bounds__(256) void geo_dilate(Geom g, long n, const uint8_t *__restrict__ src, int first,
                                                  uint8_t *__restrict__ dst, int axis)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    int idx[3];
    node_of(g, l, idx[0], idx[1], idx[2]);
    const int nn[3] = {g.ni, g.nj, g.nk};
    const long st[3] = {1, g.sj, g.sk};
    int lo, hi;
    box(idx[axis], nn[axis], lo, hi);
    const long base = l - (long)idx[axis] * st[axis];
    uint8_t m = 0;
    for (int c = lo; c <= hi; ++c) {
        const uint8_t v = src[base + (long)c * st[axis]];
        if (first ? v == 1 : v != 0) { m = 1; break; }
    }
    dst[l] = m;
}

__global__ __launch_bounds__(256) void geo_mark(Geom g, long n, const int8_t *__restrict__ nid,
                                                const uint8_t *__restrict__ dij, uint8_t *__restrict__ nmark,
                                                int32_t *__restrict__ nlist, int32_t *__restrict__ count)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool m = false;
    if (l < n) {
        if (nid[l] == 0) {
            int i, j, k;
            node_of(g, l, i, j, k);
            int k0, k1;
            box(k, g.nk, k0, k1);
            const long base = l - (long)k * g.sk;
            for (int c = k0; c <= k1; ++c)
                if (dij[base + (long)c * g.sk]) { m = true; break; }
        }
        nmark[l] = m ? 1 : 0;
    }
    (void)wave_append(m, (int)l, nlist, count);
}

// per interface node: its K-tetrahedra in the order GetTetrahedraList.F90:346-418
// appends them (cells in k, j, i order, tetrahedra 1..6), their number, and
// a sphere about the node holding all their triangles (radius up, shortest
// edge down)
// per interface node: frame (3 unit axes), extents along them, 4 vertices
constexpr int kMF = 0, kME = 9, kMV = 15, kMB = 27;

__global__ __launch_bounds__(256) void geo_iface(Geom g, long nint, const int32_t *__restrict__ ilist,
                                                 const uint8_t *__restrict__ ctets, const int32_t *__restrict__ cslot,
                                                 const int8_t *__restrict__ nuse, const double *__restrict__ trec,
                                                 int32_t *__restrict__ mrec, int8_t *__restrict__ nkt,
                                                 double *__restrict__ mrad, double *__restrict__ memin,
                                                 double *__restrict__ mbox)
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
                for (int t = 0; t < 6; ++t) {
                    if (!((tm >> t) & 1) || !((kTetHas[t] >> vq) & 1)) continue;
                    const long rec = ktet_rec(cslot, c, tm, t);
                    mrec[m * 24 + q] = (int32_t)rec;
                    ++q;
                    const int nu = nuse[rec];
                    for (int u = 0; u < nu; ++u) {
                        const double *T = trec + (rec * 2 + u) * kTS;
                        const double dx = T[kCx] - X, dy = T[kCy] - Y, dz = T[kCz] - Z;
                        R = fmax(R, sqrt(dx * dx + dy * dy + dz * dz) + T[kR]);
                        em = fmin(em, T[kEmin]);
                        any = true;
                    }
                }
            }
    nkt[m] = (int8_t)q;
    mrad[m] = any ? R * (1.0 + 1e-12) : -1.0;   // -1: no usable triangle
    memin[m] = em;
    if (!any) return;
    // a box about the node's triangles in a frame along their mean normal
    // (extents rounded outward), and four of their vertices
    double *bx = mbox + m * kMB;
    double nx = 0.0, ny = 0.0, nz = 0.0, rx = 0.0, ry = 0.0, rz = 0.0;
    int nv = 0;
    for (int qq = 0; qq < q; ++qq) {
        const long rec = mrec[m * 24 + qq];
        const int nu = nuse[rec];
        for (int u = 0; u < nu; ++u) {
            const double *T = trec + (rec * 2 + u) * kTS;
            const double a0 = T[kN], a1 = T[kN + 1], a2 = T[kN + 2];
            if (nv == 0) { rx = a0; ry = a1; rz = a2; }
            const double sg = (a0 * rx + a1 * ry + a2 * rz) < 0.0 ? -1.0 : 1.0;
            nx += sg * a0; ny += sg * a1; nz += sg * a2;
            if (nv < 4) { bx[kMV + 3 * nv] = T[kA]; bx[kMV + 3 * nv + 1] = T[kA + 1]; bx[kMV + 3 * nv + 2] = T[kA + 2]; }
            ++nv;
        }
    }
    for (int v = nv; v < 4; ++v)
        for (int c = 0; c < 3; ++c) bx[kMV + 3 * v + c] = bx[kMV + c];
    double nn = sqrt(nx * nx + ny * ny + nz * nz);
    if (!(nn > 0.0)) { nx = 0.0; ny = 0.0; nz = 1.0; nn = 1.0; }
    nx /= nn; ny /= nn; nz /= nn;
    // a second axis across the normal's smallest component
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (fabs(nx) <= fabs(ny) && fabs(nx) <= fabs(nz)) ax = 1.0;
    else if (fabs(ny) <= fabs(nz)) ay = 1.0;
    else az = 1.0;
    double t1x = ny * az - nz * ay, t1y = nz * ax - nx * az, t1z = nx * ay - ny * ax;
    const double t1n = sqrt(t1x * t1x + t1y * t1y + t1z * t1z);
    t1x /= t1n; t1y /= t1n; t1z /= t1n;
    const double t2x = ny * t1z - nz * t1y, t2y = nz * t1x - nx * t1z, t2z = nx * t1y - ny * t1x;
    const double F[9] = {nx, ny, nz, t1x, t1y, t1z, t2x, t2y, t2z};
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int qq = 0; qq < q; ++qq) {
        const long rec = mrec[m * 24 + qq];
        const int nu = nuse[rec];
        for (int u = 0; u < nu; ++u) {
            const double *T = trec + (rec * 2 + u) * kTS;
            for (int vq = 0; vq < 3; ++vq) {
                const double d0 = T[3 * vq] - X, d1 = T[3 * vq + 1] - Y, d2 = T[3 * vq + 2] - Z;
                for (int c = 0; c < 3; ++c) {
                    const double pr = d0 * F[3 * c] + d1 * F[3 * c + 1] + d2 * F[3 * c + 2];
                    lo[c] = fmin(lo[c], pr);
                    hi[c] = fmax(hi[c], pr);
                }
            }
        }
    }
    for (int c = 0; c < 9; ++c) bx[kMF + c] = F[c];
    for (int c = 0; c < 3; ++c) {
        const double w = 1e-12 * (fabs(lo[c]) + fabs(hi[c]) + R);
        bx[kME + 2 * c] = lo[c] - w;
        bx[kME + 2 * c + 1] = hi[c] + w;
    }
}

struct DistArgs {
    Geom g;
    long nint, nnb;
    const int32_t *ilist, *nlist, *mid;
    const int8_t *nid;
    const int32_t *mrec;
    const int8_t *nkt;
    const double *mrad, *memin, *mbox;
    const int8_t *nuse;
    const double *trec;
    double *pc;
};

// How far the reference's DistancePointTriangle may fall below the exact
// distance for points within L of P (shortest edge emin, coordinates up to
// cmax): rounding of the coordinates and of the norms, and the cross-product
// line distance, whose error grows like |AP| |BP| / |AB| <= L^2 / emin --
// with a margin of more than three orders of magnitude.  A vanishing edge
// makes it infinite.
__device__ __forceinline__ double dist_err(double L, double emin, double cmax)
{
    return 1e-12 * (L + L * (L / emin) + cmax + L);
}

__device__ __forceinline__ double seg_dist(const V3 &X, const V3 &e, double inv_e2, const V3 &P)
{
    const V3 w = sub(P, X);
    double t = (w.x * e.x + w.y * e.y + w.z * e.z) * inv_e2;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    const double dx = w.x - t * e.x, dy = w.y - t * e.y, dz = w.z - t * e.z;
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// A lower bound of the reference's distance from P to triangle T, error
// allowance subtracted.  The sphere bound; for a well-shaped triangle also the
// exact Euclidean distance: the distance to the triangle's plane when the foot
// of the perpendicular may lie in the triangle (the plane distance never
// exceeds the distance to the triangle, so an uncertain classification errs
// low), else the least distance to its three edges.  -inf when the
// allowance is not finite.
__device__ __forceinline__ double tri_lower(const double *T, const V3 &P, double cmax)
{
    const double dx = P.x - T[kCx], dy = P.y - T[kCy], dz = P.z - T[kCz];
    const double dP = sqrt(dx * dx + dy * dy + dz * dz);
    const double r = T[kR], emin = T[kEmin];
    const double L = dP + r;
    const double err = dist_err(L, emin, cmax);
    double lb = dP - r;
    if (!(err < 1e300)) return -1e300;
    const double invN = T[kInvN];
    if (invN != 0.0) {
        const V3 A{T[kA], T[kA + 1], T[kA + 2]}, B{T[kB], T[kB + 1], T[kB + 2]}, C{T[kC], T[kC + 1], T[kC + 2]};
        const V3 N{T[kN], T[kN + 1], T[kN + 2]};
        const double NN = T[kNN];
        const V3 u = sub(B, A), v = sub(C, A), w = sub(P, A);
        const V3 cu = cross(u, w), cv = cross(w, v);
        const double an = cu.x * N.x + cu.y * N.y + cu.z * N.z;
        const double bn = cv.x * N.x + cv.y * N.y + cv.z * N.z;
        const double tol = 1e-9 * NN * (1.0 + L / emin);
        double t;
        if (an >= -tol && bn >= -tol && an + bn <= NN + tol) {
            t = fabs(N.x * w.x + N.y * w.y + N.z * w.z) * invN;
        } else {
            t = seg_dist(A, u, T[kInvAB], P);
            t = fmin(t, seg_dist(B, sub(C, B), T[kInvBC], P));
            t = fmin(t, seg_dist(C, sub(A, C), T[kInvCA], P));
        }
        lb = fmax(lb, t);
    }
    return lb - err;
}

// An upper bound of the reference's distance from P to triangle T: its
// minimum includes the three vertex distances.
__device__ __forceinline__ double tri_upper(const double *T, const V3 &P)
{
    double m = 1e300;
    for (int q = 0; q < 3; ++q) {
        const double dx = P.x - T[3 * q], dy = P.y - T[3 * q + 1], dz = P.z - T[3 * q + 2];
        m = fmin(m, dx * dx + dy * dy + dz * dz);
    }
    return sqrt(m) * (1.0 + 1e-12) + 1e-300;
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

// the value a lane holds, the lane index being the same in every lane
// (v_readlane: no LDS round trip)
__device__ __forceinline__ double lane_value(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

constexpr int kBoxMax = 729;   // 9 x 9 x 9 search box
constexpr double kSlack = 2e-14;   // D and |phicorrected| exceed a bound by at most roundings of 10

// One wavefront (one workgroup) per listed node: interface nodes first, then
// neighbours.  The node's result follows the sequence of
// GetIsosurfaceAndDistances.F90:153-240 (interface nodes: D and SD carried over
// the whole search box) or NeighboursPhiCorrection.F90:133-249 (neighbours: D
// restarted per interface node, phicorrected updated after every
// K-tetrahedron): the interface nodes of the 9^3 box in k, j, i order, their
// K-tetrahedra, their triangles.
//  1. The lanes list the box's interface nodes in that order.
//  2. Per listed node, in parallel: a lower bound of the distances to all its
//     triangles and an upper bound of the least one; the prefix minimum of
//     the upper bounds bounds the comparand (D, or |phicorrected|) every
//     later node meets, so a node whose lower bound exceeds it cannot change
//     D (or phicorrected) and needs no distance.
//  3. In order, groups of remaining nodes: the lanes evaluate the groups'
//     triangle distances together, then every lane runs the reference's
//     updates over them in the reference's order (skipped nodes take their
//     unchanged-D updates), so the results are bit-identical.
__global__ __launch_bounds__(64) void geo_distance(DistArgs a)
{
    __shared__ int32_t s_m[kBoxMax];     // listed interface nodes
    __shared__ uint8_t s_cnt[kBoxMax];   // their K-tetrahedra (bits 0-4), distances needed (bit 7)
    __shared__ uint64_t s_pat[kBoxMax];  // usable triangles of each K-tetrahedron, 2 bits each
    const int lane = threadIdx.x;
    const long t = blockIdx.x;
    const Geom &g = a.g;
    const bool iface = t < a.nint;
    const long l = iface ? a.ilist[t] : a.nlist[t - a.nint];
    int i, j, k;
    node_of(g, l, i, j, k);
    const V3 P{g.x[l], g.y[l], g.z[l]};
    const double cmax = fmax(fabs(P.x), fmax(fabs(P.y), fabs(P.z)));
    int i0, i1, j0, j1, k0, k1;
    box(i, g.ni, i0, i1); box(j, g.nj, j0, j1); box(k, g.nk, k0, k1);
    const int bx = i1 - i0 + 1, by = j1 - j0 + 1, nbox = bx * by * (k1 - k0 + 1);
    // 1. interface nodes of the box, in box order (all flags loaded first)
    constexpr int kChunks = (kBoxMax + 63) / 64;
    int8_t f[kChunks];
    long ls[kChunks];
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
        const int e = c * 64 + lane;
        f[c] = 0;
        ls[c] = 0;
        if (e < nbox) {
            const int r = e / bx;
            ls[c] = (long)(i0 + e - r * bx) + g.sj * (j0 + r % by) + g.sk * (k0 + r / by);
            f[c] = a.nid[ls[c]];
        }
    }
    int mmv[kChunks];
#pragma unroll
    for (int c = 0; c < kChunks; ++c) mmv[c] = f[c] == 1 ? a.mid[ls[c]] : -1;
    int nm = 0;
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
        const unsigned long long bal = __ballot(mmv[c] >= 0);
        if (mmv[c] >= 0) s_m[nm + (int)__popcll(bal & ((1ull << lane) - 1ull))] = mmv[c];
        nm += (int)__popcll(bal);
    }
    __syncthreads();
    const double pc0 = iface ? 0.0 : a.pc[l];
    // a neighbour's updates with D >= 10 leave |phicorrected| <= 10 unchanged
    const bool can = iface || fabs(pc0) <= 10.0;
    // 2. which listed nodes need their distances
    double carry = iface ? 10.0 : fabs(pc0);
    for (int c0 = 0; c0 < nm; c0 += 64) {
        const int e = c0 + lane;
        long mm = -1;
        double lb = 1e300, ub = 1e300;   // no usable triangle: out of reach, no bound
        double R = -1.0;
        int cnt = 0;
        if (e < nm) {
            mm = s_m[e];
            const long lm = a.ilist[mm];
            R = a.mrad[mm];
            const double em = a.memin[mm];
            cnt = a.nkt[mm];
            const double dx = P.x - g.x[lm], dy = P.y - g.y[lm], dz = P.z - g.z[lm];
            if (R >= 0.0) {
                const double dP = sqrt(dx * dx + dy * dy + dz * dz);
                const double err = dist_err(dP + R, em, cmax);
                lb = err < 1e300 ? dP - R - err : -1e300;
                ub = (dP + R) * (1.0 + 1e-12) + 1e-300;
                const double *bx = a.mbox + mm * kMB;
                if (err < 1e300) {
                    // distance to the box, from the offsets along its axes
                    double q2 = 0.0;
                    for (int c = 0; c < 3; ++c) {
                        const double pr = dx * bx[kMF + 3 * c] + dy * bx[kMF + 3 * c + 1] + dz * bx[kMF + 3 * c + 2];
                        const double o = fmax(0.0, fmax(bx[kME + 2 * c] - pr, pr - bx[kME + 2 * c + 1]));
                        q2 += o * o;
                    }
                    lb = fmax(lb, sqrt(q2) - err);
                }
                // the reference's minimum includes every vertex distance
                double um = 1e300;
                for (int v = 0; v < 4; ++v) {
                    const double e0 = P.x - bx[kMV + 3 * v], e1 = P.y - bx[kMV + 3 * v + 1], e2 = P.z - bx[kMV + 3 * v + 2];
                    um = fmin(um, e0 * e0 + e1 * e1 + e2 * e2);
                }
                ub = fmin(ub, sqrt(um) * (1.0 + 1e-12) + 1e-300);
            }
        }
        // exclusive prefix minimum of the upper bounds over the list
        double incl = ub;
        for (int d = 1; d < 64; d <<= 1) {
            const double o = __shfl_up(incl, d);
            if (lane >= d) incl = fmin(incl, o);
        }
        double excl = __shfl_up(incl, 1);
        if (lane == 0) excl = 1e300;
        bool go = !(can && lb > fmin(carry, excl) + kSlack);
        uint64_t pat = 0;
        if (go && e < nm) {
            // the sphere cannot tell: bound every triangle
            double lo = 1e300, hi = 1e300;
            for (int q = 0; q < cnt; ++q) {
                const long rr = a.mrec[mm * 24 + q];
                const int nn = a.nuse[rr];
                pat |= (uint64_t)nn << (2 * q);
                for (int u = 0; u < nn; ++u) {
                    const double *T = a.trec + (rr * 2 + u) * kTS;
                    lo = fmin(lo, tri_lower(T, P, cmax));
                    hi = fmin(hi, tri_upper(T, P));
                }
            }
            if (R >= 0.0) {
                lb = lo;
                ub = fmin(ub, hi);
            }
        }
        incl = ub;
        for (int d = 1; d < 64; d <<= 1) {
            const double o = __shfl_up(incl, d);
            if (lane >= d) incl = fmin(incl, o);
        }
        excl = __shfl_up(incl, 1);
        if (lane == 0) excl = 1e300;
        go = !(can && lb > fmin(carry, excl) + kSlack);
        if (e < nm) {
            s_cnt[e] = (uint8_t)(cnt | (go ? 0x80 : 0));
            s_pat[e] = pat;
        }
        carry = fmin(carry, __shfl(incl, 63));
    }
    __syncthreads();
    // 3. the reference's updates, in order
    const double sgn0 = copysign(1.0, g.phi0[l]);
    double D = 10.0, SD = 99.9, pc = pc0;
    int e = 0;
    while (e < nm) {
        const int ce = s_cnt[e];
        if (!(ce & 0x80)) {
            // D stays; SD takes its per-tetrahedron update cnt times
            if (iface)
                for (int c = 0; c < (ce & 0x7f); ++c) {
                    const double s2 = sd_update(SD, D, sgn0);
                    if (s2 == SD) break;
                    SD = s2;
                }
            ++e;
            continue;
        }
        // a group of listed nodes whose usable triangles fit the 64 lanes;
        // this lane's triangle
        int gend = e, slots = 0;
        long myrec = -1;
        int myu = 0;
        while (gend < nm) {
            const int cg = s_cnt[gend];
            if (cg & 0x80) {
                const uint64_t pat = s_pat[gend];
                const int need = (int)__popcll(pat & 0x5555555555555555ull) + 2 * (int)__popcll(pat & 0xaaaaaaaaaaaaaaaaull);
                if (slots + need > 64) break;
                if (lane >= slots && lane < slots + need) {
                    int idx = lane - slots, q = 0;
                    for (;; ++q) {
                        const int nn = (int)((pat >> (2 * q)) & 3);
                        if (idx < nn) break;
                        idx -= nn;
                    }
                    myu = idx;
                    myrec = a.mrec[(long)s_m[gend] * 24 + q];
                }
                slots += need;
            }
            ++gend;
        }
        double myDa = 0.0;
        if (myrec >= 0) myDa = tri_distance(a.trec + (myrec * 2 + myu) * kTS, P);
        int tix = 0;
        for (int f2 = e; f2 < gend; ++f2) {
            const int cg = s_cnt[f2];
            const int cnt = cg & 0x7f;
            if (!(cg & 0x80)) {
                if (iface)
                    for (int c = 0; c < cnt; ++c) {
                        const double s2 = sd_update(SD, D, sgn0);
                        if (s2 == SD) break;
                        SD = s2;
                    }
                continue;
            }
            const uint64_t pat = s_pat[f2];
            if (!iface) D = 10.0;
            for (int q = 0; q < cnt; ++q) {
                const int nn = (int)((pat >> (2 * q)) & 3);
                for (int u = 0; u < nn; ++u) {
                    const double Da = lane_value(myDa, tix++);
                    const double ex = (copysign(1.0, D - Da) + 1.0) / 2.0;
                    D = D + ex * (Da - D);
                }
                // end of a K-tetrahedron (GetIsosurfaceAndDistances.F90:221-229,
                // NeighboursPhiCorrection.F90:203-215)
                if (iface) SD = sd_update(SD, D, sgn0);
                else pc = pc_update(pc, D);
            }
        }
        e = gend;
    }
    // NeighboursPhiCorrection.F90:89-130 with C = 0: phi = SD + 0
    if (lane == 0) a.pc[l] = iface ? SD + 0.0 : pc;
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

// the isosurface store: one per (device, stream), grown on demand, kept until exit
int store_take(size_t bytes, hipStream_t s, void **p)
{
    return cached_buffer(kBufStore, bytes, s, p);
}

unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

size_t geometric_work(const Blk &b)
{
    const size_t nc = (size_t)(b.e.ni - 1) * (b.e.nj - 1) * (b.e.nk - 1);
    const size_t n = (size_t)b.n;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    return al(n) * 3 + al(nc) + al(nc) + al(nc * 4) + al(n * 4) * 3 + 256 + 2048;
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
    uint8_t *nmark = (uint8_t *)c.take(n), *dil = (uint8_t *)c.take(n);
    uint8_t *ctets = (uint8_t *)c.take(ncell), *cpn = (uint8_t *)c.take(ncell);
    int32_t *cslot = (int32_t *)c.take(ncell * 4);
    int32_t *mid = (int32_t *)c.take(n * 4), *ilist = (int32_t *)c.take(n * 4), *nlist = (int32_t *)c.take(n * 4);
    int32_t *count = (int32_t *)c.take(256);
    if (!nid || !nmark || !dil || !ctets || !cpn || !cslot || !mid || !ilist || !nlist || !count)
        return fail(LSM_ERR_WORKSPACE, "workspace too small");
    if (n > 0x7fffffffL) return fail(LSM_ERR_ARG, "block of %ld nodes: at most 2^31 - 1", n);
    Geom g{b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, (long)(b.e.ni - 1), (long)(b.e.ni - 1) * (b.e.nj - 1),
           b.x, b.y, b.z, phi0, b.bl};
    int rc;
    // counts: cells with K-tetrahedra, interface nodes, neighbours
    geo_zero<<<1, 64, 0, s>>>(count);
    if ((rc = check_hip(hipGetLastError(), "geo_zero"))) return rc;
    geo_cells<<<nblk(ncell), 256, 0, s>>>(g, ncell, ctets, cpn, cslot, count);
    if ((rc = check_hip(hipGetLastError(), "geo_cells"))) return rc;
    geo_nodes<<<nblk(n), 256, 0, s>>>(g, n, cpn, nid, mid, ilist, count + 1);
    if ((rc = check_hip(hipGetLastError(), "geo_nodes"))) return rc;
    // interface flags dilated along i (into nmark, as scratch), then j (into dil)
    geo_dilate<<<nblk(n), 256, 0, s>>>(g, n, (const uint8_t *)nid, 1, nmark, 0);
    geo_dilate<<<nblk(n), 256, 0, s>>>(g, n, nmark, 0, dil, 1);
    geo_mark<<<nblk(n), 256, 0, s>>>(g, n, nid, dil, nmark, nlist, count + 2);
    if ((rc = check_hip(hipGetLastError(), "geo_mark"))) return rc;
    // the counts size the isosurface store and the distance launch
    int32_t cnt[3] = {0, 0, 0};
    if ((rc = check_hip(hipMemcpyAsync(cnt, count, sizeof(cnt), hipMemcpyDeviceToHost, s), "D2H")) ||
        (rc = check_hip(hipStreamSynchronize(s), "hipStreamSynchronize")))
        return rc;
    const long nslots = cnt[0], nint = cnt[1], nnb = cnt[2];
    if (nint > 0) {
        const size_t nrec = (size_t)nslots;   // K-tetrahedra
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t o_trec = al(nrec);
        const size_t o_mrec = o_trec + al(nrec * 2 * kTS * sizeof(double));
        const size_t o_nkt = o_mrec + al((size_t)nint * 24 * sizeof(int32_t));
        const size_t o_rad = o_nkt + al((size_t)nint);
        const size_t o_emin = o_rad + al((size_t)nint * sizeof(double));
        const size_t o_box = o_emin + al((size_t)nint * sizeof(double));
        const size_t total = o_box + al((size_t)nint * kMB * sizeof(double));
        void *store = nullptr;
        if ((rc = store_take(total, s, &store))) return rc;
        char *sb = (char *)store;
        int8_t *nuse = (int8_t *)sb;
        double *trec = (double *)(sb + o_trec);
        int32_t *mrec = (int32_t *)(sb + o_mrec);
        int8_t *nkt = (int8_t *)(sb + o_nkt);
        double *mrad = (double *)(sb + o_rad), *memin = (double *)(sb + o_emin), *mbox = (double *)(sb + o_box);
        geo_iso<<<nblk(ncell), 256, 0, s>>>(g, ncell, ctets, cslot, nuse, trec);
        if ((rc = check_hip(hipGetLastError(), "geo_iso"))) return rc;
        geo_iface<<<nblk(nint), 256, 0, s>>>(g, nint, ilist, ctets, cslot, nuse, trec, mrec, nkt, mrad, memin, mbox);
        if ((rc = check_hip(hipGetLastError(), "geo_iface"))) return rc;
        DistArgs a{g, nint, nnb, ilist, nlist, mid, nid, mrec, nkt, mrad, memin, mbox, nuse, trec, pc};
        {
            KTimer kt(LSM_KT_GEOMETRIC, s);
            geo_distance<<<(unsigned)(nint + nnb), 64, 0, s>>>(a);
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
