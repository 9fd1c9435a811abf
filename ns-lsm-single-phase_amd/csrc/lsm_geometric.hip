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

__global__ __launch_bounds__(256) void geo_iso(Geom g, long ncell, const uint8_t *__restrict__ ctets,
                                               const int32_t *__restrict__ cslot, int8_t *__restrict__ nuse,
                                               double *__restrict__ tris)
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
        nuse[slot + t] = (int8_t)nu;
        for (int q = 0; q < 9 * nu; ++q) tris[(slot + t) * 18 + q] = tri[q];
    }
}

__device__ __forceinline__ void node_of(const Geom &g, long l, int &i, int &j, int &k)
{
    i = (int)(l % g.ni);
    const long r = l / g.ni;
    j = (int)(r % g.nj);
    k = (int)(r / g.nj);
}

// a node is an interface node when it is a vertex of a K-tetrahedron of one
// of its cells (GetTetrahedraList.F90:262-269)
__global__ __launch_bounds__(256) void geo_nodes(Geom g, long n, const uint8_t *__restrict__ cpn,
                                                 int8_t *__restrict__ nid)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    int i, j, k;
    node_of(g, l, i, j, k);
    int8_t f = 0;
    for (int dk = -1; dk <= 0; ++dk)
        for (int dj = -1; dj <= 0; ++dj)
            for (int di = -1; di <= 0; ++di) {
                const int ci = i + di, cj = j + dj, ck = k + dk;
                if (ci < 0 || cj < 0 || ck < 0 || ci > g.ni - 2 || cj > g.nj - 2 || ck > g.nk - 2) continue;
                const int pn = cpn[(long)ci + g.csj * cj + g.csk * ck];
                // vertex index of this node in that cell
                const int ox = -di, oy = -dj, oz = -dk;
                const int vq = oz * 4 + (oy ? (ox ? 2 : 3) : ox);
                if (pn & (1 << vq)) f = 1;
            }
    nid[l] = f;
}

__device__ __forceinline__ void box(int c, int n, int &lo, int &hi)
{
    lo = c - 4 < 0 ? 0 : c - 4;
    hi = c + 4 > n - 1 ? n - 1 : c + 4;
}

// neighbours: the nodes the search boxes of GetIsosurfaceAndDistances.F90:153-240
// visit without being interface nodes (InterfaceNodesID = -1)
__global__ __launch_bounds__(256) void geo_mark(Geom g, long n, const int8_t *__restrict__ nid,
                                                uint8_t *__restrict__ nmark)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    int i, j, k;
    node_of(g, l, i, j, k);
    uint8_t m = 0;
    if (nid[l] == 0) {
        int i0, i1, j0, j1, k0, k1;
        box(i, g.ni, i0, i1); box(j, g.nj, j0, j1); box(k, g.nk, k0, k1);
        for (int ks = k0; ks <= k1 && !m; ++ks)
            for (int js = j0; js <= j1 && !m; ++js)
                for (int is = i0; is <= i1; ++is)
                    if (nid[(long)is + g.sj * js + g.sk * ks] == 1) { m = 1; break; }
    }
    nmark[l] = m;
}

struct DistArgs {
    Geom g;
    long n;
    const int8_t *nid;
    const uint8_t *nmark, *ctets;
    const int32_t *cslot;
    const int8_t *nuse;
    const double *tris;
    double *pc;
};

// the running minimum of GetIsosurfaceAndDistances.F90:201-233 over the
// isosurface of one K-tetrahedron
__device__ __forceinline__ double ktet_min(const DistArgs &a, long rec, const V3 &P, double D)
{
    const int nu = a.nuse[rec];
    for (int t = 0; t < nu; ++t) {
        const double Da = dist_point_triangle(a.tris + rec * 18 + 9 * t, P);
        const double ex = (copysign(1.0, D - Da) + 1.0) / 2.0;
        D = D + ex * (Da - D);
    }
    return D;
}

__global__ __launch_bounds__(256) void geo_distance(DistArgs a)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= a.n) return;
    const Geom &g = a.g;
    const bool iface = a.nid[l] == 1;
    if (!iface && !a.nmark[l]) return;
    int i, j, k;
    node_of(g, l, i, j, k);
    const V3 P{g.x[l], g.y[l], g.z[l]};
    int i0, i1, j0, j1, k0, k1;
    box(i, g.ni, i0, i1); box(j, g.nj, j0, j1); box(k, g.nk, k0, k1);
    const double sgn0 = copysign(1.0, g.phi0[l]);
    double D = 10.0, SD = 99.9;
    double pc = iface ? 0.0 : a.pc[l];
    for (int ks = k0; ks <= k1; ++ks)
        for (int js = j0; js <= j1; ++js)
            for (int is = i0; is <= i1; ++is) {
                if (a.nid[(long)is + g.sj * js + g.sk * ks] != 1) continue;
                if (!iface) D = 10.0;
                // the K-tetrahedra of interface node (is, js, ks) in the order
                // GetTetrahedraList.F90:346-418 appends them
                for (int dk = -1; dk <= 0; ++dk)
                    for (int dj = -1; dj <= 0; ++dj)
                        for (int di = -1; di <= 0; ++di) {
                            const int ci = is + di, cj = js + dj, ck = ks + dk;
                            if (ci < 0 || cj < 0 || ck < 0 || ci > g.ni - 2 || cj > g.nj - 2 || ck > g.nk - 2)
                                continue;
                            const long c = (long)ci + g.csj * cj + g.csk * ck;
                            const int tm = a.ctets[c];
                            if (!tm) continue;
                            const int ox = -di, oy = -dj, oz = -dk;
                            const int vq = oz * 4 + (oy ? (ox ? 2 : 3) : ox);
                            const long slot = (long)a.cslot[c] * 6;
                            for (int t = 0; t < 6; ++t) {
                                if (!((tm >> t) & 1) || !((kTetHas[t] >> vq) & 1)) continue;
                                D = ktet_min(a, slot + t, P, D);
                                if (iface) {
                                    const double ex2 = (copysign(1.0, fabs(SD) - D) + 1.0) / 2.0;
                                    SD = fabs(SD) + ex2 * (D - fabs(SD));
                                    SD = sgn0 * SD;
                                } else {
                                    const double ex2 = (copysign(1.0, fabs(pc) - D) + 1.0) / 2.0;
                                    const double aux = fabs(pc) + ex2 * (D - fabs(pc));
                                    pc = copysign(1.0, pc) * aux;
                                }
                            }
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

__global__ void geo_zero(int32_t *c) { *c = 0; }

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
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    return al(b.n) + al(b.n) + al(nc) + al(nc) + al(nc * 4) + 256 + 1024;
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
    int32_t *count = (int32_t *)c.take(256);
    if (!nid || !nmark || !ctets || !cpn || !cslot || !count) return fail(LSM_ERR_WORKSPACE, "workspace too small");
    Geom g{b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, (long)(b.e.ni - 1), (long)(b.e.ni - 1) * (b.e.nj - 1),
           b.x, b.y, b.z, phi0};
    int rc;
    geo_zero<<<1, 1, 0, s>>>(count);
    if ((rc = check_hip(hipGetLastError(), "geo_zero"))) return rc;
    geo_cells<<<nblk(ncell), 256, 0, s>>>(g, ncell, ctets, cpn, cslot, count);
    if ((rc = check_hip(hipGetLastError(), "geo_cells"))) return rc;
    // the number of cells with K-tetrahedra sizes the isosurface store
    int32_t nslots = 0;
    if ((rc = check_hip(hipMemcpyAsync(&nslots, count, sizeof(int32_t), hipMemcpyDeviceToHost, s), "D2H")) ||
        (rc = check_hip(hipStreamSynchronize(s), "hipStreamSynchronize")))
        return rc;
    const size_t nrec = (size_t)nslots * 6;
    const size_t tri_off = (nrec + 255) & ~(size_t)255;
    void *store = nullptr;
    if ((rc = store_take(tri_off + nrec * 18 * sizeof(double) + 256, &store))) return rc;
    int8_t *nuse = (int8_t *)store;
    double *tris = (double *)((char *)store + tri_off);
    geo_iso<<<nblk(ncell), 256, 0, s>>>(g, ncell, ctets, cslot, nuse, tris);
    if ((rc = check_hip(hipGetLastError(), "geo_iso"))) return rc;
    geo_nodes<<<nblk(n), 256, 0, s>>>(g, n, cpn, nid);
    if ((rc = check_hip(hipGetLastError(), "geo_nodes"))) return rc;
    geo_mark<<<nblk(n), 256, 0, s>>>(g, n, nid, nmark);
    if ((rc = check_hip(hipGetLastError(), "geo_mark"))) return rc;
    DistArgs a{g, n, nid, nmark, ctets, cslot, nuse, tris, pc};
    {
        KTimer kt(LSM_KT_GEOMETRIC, s);
        geo_distance<<<nblk(n), 256, 0, s>>>(a);
    }
    if ((rc = check_hip(hipGetLastError(), "geo_distance"))) return rc;
    if (ids) {
        geo_ids<<<nblk(n), 256, 0, s>>>(n, nid, nmark, ids);
        if ((rc = check_hip(hipGetLastError(), "geo_ids"))) return rc;
    }
    return LSM_OK;
}

}  // namespace lsm
