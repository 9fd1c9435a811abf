// Water volume of the phi > 0 region (TotalWaterVolume + TetrahedronMethods)
// and the AdvectionNodes interface mask.
//
// One thread per grid cell: a fully wet cell adds its Davies-Salmond
// hexahedron volume, a cut cell is split into six tetrahedra and the cut ones
// go through marching tetrahedra.  Per-workgroup partial sums (wavefront
// shuffles, then LDS) are written to scratch and summed by a single
// workgroup in a fixed order, so the result does not depend on scheduling.
// The AdvectionNodes mask is built by two stencil passes of its own.
#include "lsm_internal.hpp"

namespace lsm {

namespace {

// M44Det, TetrahedronMethods.F90:1698-1720 (A row-major)
__device__ double m44det(const double A[4][4])
{
#define a(r, c) A[(r) - 1][(c) - 1]
    return ((a(1,1)*((a(2,2)*(a(3,3)*a(4,4)-a(3,4)*a(4,3))+a(2,3)*(a(3,4)*a(4,2)-a(3,2)*a(4,4)))+a(2,4)*(a(3,2)*a(4,3)-
            a(3,3)*a(4,2)))-a(1,2)*((a(2,1)*(a(3,3)*a(4,4)-a(3,4)*a(4,3))+a(2,3)*(a(3,4)*a(4,1)-a(3,1)*a(4,4)))+
            a(2,4)*(a(3,1)*a(4,3)-a(3,3)*a(4,1))))+a(1,3)*((a(2,1)*(a(3,2)*a(4,4)-a(3,4)*a(4,2))+a(2,2)*(a(3,4)*a(4,1)-
            a(3,1)*a(4,4)))+a(2,4)*(a(3,1)*a(4,2)-a(3,2)*a(4,1))))-a(1,4)*((a(2,1)*(a(3,2)*a(4,3)-a(3,3)*a(4,2))+
            a(2,2)*(a(3,3)*a(4,1)-a(3,1)*a(4,3)))+a(2,3)*(a(3,1)*a(4,2)-a(3,2)*a(4,1)));
#undef a
}

// GetTetrahedronVolume, TetrahedronMethods.F90:666-688
__device__ double tet_volume(const double v[4][3])
{
    double A[4][4];
    for (int r = 0; r < 4; ++r) { A[r][0] = v[r][0]; A[r][1] = v[r][1]; A[r][2] = v[r][2]; A[r][3] = 1.0; }
    return (1.0 / 6.0) * fabs(m44det(A));
}

__device__ __forceinline__ void cross3(const double *u, const double *v, double *o)
{
    o[0] = u[1] * v[2] - u[2] * v[1];
    o[1] = u[2] * v[0] - u[0] * v[2];
    o[2] = u[0] * v[1] - u[1] * v[0];
}

__device__ __forceinline__ double dot3(const double *u, const double *v) { return (u[0] * v[0] + u[1] * v[1]) + u[2] * v[2]; }

// HexahedronVolume (Davies & Salmond 1985), TetrahedronMethods.F90:1724-1793
__device__ double hex_volume(const double V[8][3])
{
    const double *A = V[0], *B = V[1], *C = V[2], *D = V[3], *E = V[4], *F = V[5], *G = V[6], *H = V[7];
    double AC[3], AF[3], AH[3], BE[3], CH[3], DB[3], ED[3], FC[3];
    double GA[3], GB[3], GC[3], GD[3], GE[3], GF[3], GH[3], HF[3];
    for (int c = 0; c < 3; ++c) {
        AC[c] = C[c] - A[c]; AF[c] = F[c] - A[c]; AH[c] = H[c] - A[c];
        BE[c] = E[c] - B[c]; CH[c] = H[c] - C[c]; DB[c] = B[c] - D[c];
        ED[c] = D[c] - E[c]; FC[c] = C[c] - F[c];
        GA[c] = A[c] - G[c]; GB[c] = B[c] - G[c]; GC[c] = C[c] - G[c]; GD[c] = D[c] - G[c];
        GE[c] = E[c] - G[c]; GF[c] = F[c] - G[c]; GH[c] = H[c] - G[c]; HF[c] = F[c] - H[c];
    }
    double t[3], s[3], acc;
    for (int c = 0; c < 3; ++c) s[c] = GA[c] + GB[c];
    cross3(DB, AC, t); acc = dot3(s, t);
    for (int c = 0; c < 3; ++c) s[c] = GA[c] + GE[c];
    cross3(BE, AF, t); acc = acc + dot3(s, t);
    for (int c = 0; c < 3; ++c) s[c] = GA[c] + GD[c];
    cross3(ED, AH, t); acc = acc + dot3(s, t);
    cross3(HF, GE, t); acc = acc + dot3(GF, t);
    cross3(CH, GD, t); acc = acc + dot3(GH, t);
    cross3(FC, GB, t); acc = acc + dot3(GC, t);
    return -((1.0 / 12.0) * acc);
}

__constant__ int kTetEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
__constant__ int kTetEdgeFlags[16] = {0x00, 0x0d, 0x13, 0x1e, 0x26, 0x2b, 0x35, 0x38,
                                      0x38, 0x35, 0x2b, 0x26, 0x1e, 0x13, 0x0d, 0x00};
__constant__ int kCellTets[6][4] = {{0, 5, 1, 6}, {0, 1, 2, 6}, {0, 2, 3, 6},
                                    {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 4, 5, 6}};

__device__ __forceinline__ void cp3(double *d, const double *s) { d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; }

// GetWaterVolume, TetrahedronMethods.F90:710-954
__device__ double water_volume_cut(const double v[4][3], const double E[6][3], int flag)
{
    const int cs = flag + 1;
    const double T = tet_volume(v);
    double t1[4][3], t2[4][3], t3[4][3];
    switch (cs) {
    case 1: return T;
    case 16: return 0.0;
    case 2: case 15:
        cp3(t1[0], v[0]); cp3(t1[1], E[0]); cp3(t1[2], E[2]); cp3(t1[3], E[3]);
        return cs == 2 ? T - tet_volume(t1) : tet_volume(t1);
    case 3: case 14:
        cp3(t1[0], v[1]); cp3(t1[1], E[0]); cp3(t1[2], E[1]); cp3(t1[3], E[4]);
        return cs == 3 ? T - tet_volume(t1) : tet_volume(t1);
    case 5: case 12:
        cp3(t1[0], v[2]); cp3(t1[1], E[1]); cp3(t1[2], E[2]); cp3(t1[3], E[5]);
        return cs == 5 ? T - tet_volume(t1) : tet_volume(t1);
    case 8: case 9:
        cp3(t1[0], v[3]); cp3(t1[1], E[3]); cp3(t1[2], E[4]); cp3(t1[3], E[5]);
        return cs == 8 ? tet_volume(t1) : T - tet_volume(t1);
    case 4: case 13: {
        cp3(t1[0], v[0]); cp3(t1[1], E[2]); cp3(t1[2], E[3]); cp3(t1[3], E[4]);
        cp3(t2[0], v[0]); cp3(t2[1], E[1]); cp3(t2[2], E[2]); cp3(t2[3], E[4]);
        cp3(t3[0], v[0]); cp3(t3[1], v[1]); cp3(t3[2], E[1]); cp3(t3[3], E[4]);
        const double aux = (tet_volume(t1) + tet_volume(t2)) + tet_volume(t3);
        return cs == 4 ? T - aux : aux;
    }
    case 6: case 11: {
        cp3(t1[0], v[0]); cp3(t1[1], E[0]); cp3(t1[2], E[1]); cp3(t1[3], E[5]);
        cp3(t2[0], v[0]); cp3(t2[1], E[0]); cp3(t2[2], E[3]); cp3(t2[3], E[5]);
        cp3(t3[0], v[0]); cp3(t3[1], v[2]); cp3(t3[2], E[1]); cp3(t3[3], E[5]);
        const double aux = (tet_volume(t1) + tet_volume(t2)) + tet_volume(t3);
        return cs == 6 ? T - aux : aux;
    }
    case 7: case 10: {
        cp3(t1[0], v[1]); cp3(t1[1], E[0]); cp3(t1[2], E[2]); cp3(t1[3], E[5]);
        cp3(t2[0], v[1]); cp3(t2[1], E[0]); cp3(t2[2], E[4]); cp3(t2[3], E[5]);
        cp3(t3[0], v[1]); cp3(t3[1], v[2]); cp3(t3[2], E[2]); cp3(t3[3], E[5]);
        const double aux = (tet_volume(t1) + tet_volume(t2)) + tet_volume(t3);
        return cs == 7 ? T - aux : aux;
    }
    }
    return 0.0;
}

// water volume returned by MarchingTetrahedron, TetrahedronMethods.F90:26-664
__device__ double marching_tet_water(const double v[4][3], const double ph[4])
{
    int sg[4], npos = 0, nneg = 0, nzero = 0;
    for (int q = 0; q < 4; ++q) {
        sg[q] = 0;
        if (ph[q] > kTiny) sg[q] = 1;
        if (ph[q] < -kTiny) sg[q] = -1;
        npos += sg[q] == 1; nneg += sg[q] == -1; nzero += sg[q] == 0;
    }
    if (npos == 4) return tet_volume(v);
    if (nneg == 4) return 0.0;
    if (nzero > 0 && (nneg == 0 || npos == 0)) {
        // isosurface through vertices, edges or a face: full or empty
        int ref = -1;
        if (nzero == 1) ref = (sg[0] == 0) ? 1 : 0;
        else if (nzero == 2) {
            if (sg[0] == 0 && sg[1] == 0) ref = 2;
            else if (sg[1] == 0 && sg[2] == 0) ref = 0;
            else if (sg[2] == 0 && sg[0] == 0) ref = 1;
            else if (sg[0] == 0 && sg[3] == 0) ref = 1;
            else ref = 0;
        } else if (nzero == 3) {
            for (int q = 0; q < 4; ++q) if (sg[q] != 0) ref = q;
        } else {
            return tet_volume(v);   // four zero vertices: FlagIndex stays 0
        }
        return sg[ref] == 1 ? tet_volume(v) : 0.0;
    }
    int flag = 0;
    for (int q = 0; q < 4; ++q) if (sg[q] == -1) flag |= 1 << q;
    if (flag == 0) return tet_volume(v);
    if (flag == 15) return 0.0;
    const int ef = kTetEdgeFlags[flag];
    double E[6][3];
    for (int e = 0; e < 6; ++e) {
        E[e][0] = E[e][1] = E[e][2] = -9.9;
        if (!(ef & (1 << e))) continue;
        const int s0 = kTetEdge[e][0], s1 = kTetEdge[e][1];
        const double den = ph[s1] - ph[s0];
        double s = 0.0;
        if (!(fabs(den) < kEpsSims * kEpsSims * kEpsSims)) s = ph[s1] / den;
        for (int c = 0; c < 3; ++c) E[e][c] = v[s1][c] + s * (v[s0][c] - v[s1][c]);
    }
    return water_volume_cut(v, E, flag);
}

// CellAnalysis volumes, TetrahedronMethods.F90:1823-2121 (TriangulationCase 1)
__device__ __forceinline__ void cell_analysis(const double ph[8], const double V[8][3], double &pc_vol,
                                           double &st_vol)
{
    pc_vol = 0.0; st_vol = 0.0;
    bool pcell = false;
    for (int q = 0; q < 8; ++q) if (fabs(ph[q]) < kEpsSims) pcell = true;
    if (!pcell) {
        const double sgn = sign1(ph[0]);
        for (int q = 1; q < 8; ++q) if (sgn * ph[q] < 0.0) { pcell = true; break; }
    }
    if (!pcell) return;
    for (int t = 0; t < 6; ++t) {
        bool pct = false, second = true;
        double pa = ph[kCellTets[t][0]];
        if (pa <= kEpsSims) second = false;
        if (fabs(pa) < kEpsSims) pct = true;
        const double sgn = sign1(pa);
        for (int q = 1; q < 4; ++q) {
            pa = ph[kCellTets[t][q]];
            if (sgn * pa < 0.0 || fabs(pa) < kEpsSims) {
                if (pa <= kEpsSims) second = false;
                pct = true;
            }
        }
        double tv[4][3], tp[4];
        for (int q = 0; q < 4; ++q) {
            const int vq = kCellTets[t][q];
            cp3(tv[q], V[vq]);
            tp[q] = ph[vq];
        }
        if (pct) pc_vol = pc_vol + marching_tet_water(tv, tp);
        if (second && !pct) st_vol = st_vol + tet_volume(tv);
    }
}

// Cells are processed in workgroups of 64 (i) x 4 (j) cells and a chunk of
// VKC k layers.  Pass 1 (volume_wet) sums the hexahedron volumes of fully
// wet cells and flags the workgroups that contain cut cells; pass 2
// (volume_cut) lists the cut cells of a flagged workgroup in LDS (in a fixed
// order) and runs their marching tetrahedra on consecutive threads, so the
// heavy path neither costs registers in pass 1 nor occupies wavefronts whose
// cells are not cut.
constexpr int VTX = 64, VTY = 4, VKC = 8;

struct VolArgs {
    const double *phi, *x, *y, *z;
    double *partial;    // [nblocks][3]: K-tetrahedra, second tetrahedra, bulk cells
    int32_t *cutflag;   // [nblocks]
    int ni, nj, nk;
    long sj, sk;
    int tiles_i, tiles_j;
    Blank bl;           // cells whose lower corner lies in a box are not counted
};

struct CellTile {
    int i, j, k0, k1;   // this thread's cell column (i, j) and the workgroup's layers [k0, k1)
    bool ok;
};

__device__ __forceinline__ CellTile tile_of(const VolArgs &a)
{
    const int b = blockIdx.x;
    const int ti = b % a.tiles_i, t2 = b / a.tiles_i;
    CellTile c;
    c.i = ti * VTX + threadIdx.x;
    c.j = (t2 % a.tiles_j) * VTY + threadIdx.y;
    c.k0 = (t2 / a.tiles_j) * VKC;
    c.k1 = min(c.k0 + VKC, a.nk - 1);
    c.ok = c.i < a.ni - 1 && c.j < a.nj - 1;
    return c;
}

// the eight corners of cell (i,j,k), in CellAnalysis' vertex order
struct Corners {
    long off[8];
    __device__ Corners(long sj, long sk)
    {
        off[0] = 0; off[1] = 1; off[2] = 1 + sj; off[3] = sj;
        off[4] = sk; off[5] = 1 + sk; off[6] = 1 + sj + sk; off[7] = sj + sk;
    }
};

// block sum of per-thread values (wavefront shuffles, then LDS; fixed order)
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double *out)
{
    for (int o = 32; o > 0; o >>= 1)
        for (int q = 0; q < N; ++q) v[q] += __shfl_xor(v[q], o, 64);
    __shared__ double sm[VTY][N];
    const int lane = threadIdx.x, wv = threadIdx.y;
    if (lane == 0)
        for (int q = 0; q < N; ++q) sm[wv][q] = v[q];
    __syncthreads();
    if (wv == 0 && lane < N) out[lane] = ((sm[0][lane] + sm[1][lane]) + sm[2][lane]) + sm[3][lane];
}

// TotalWaterVolume.F90:136-294, fully wet cells (HexahedronVolume) and the
// classification of the rest; a cell whose lower corner node lies in a
// blanking box is skipped (TotalWaterVolume.F90:139-160)
__global__ __launch_bounds__(VTX * VTY) void volume_wet(VolArgs a)
{
    const CellTile c = tile_of(a);
    const Corners cn(a.sj, a.sk);
    double v[3] = {0.0, 0.0, 0.0};
    bool cut = false;
    if (c.ok) {
        for (int k = c.k0; k < c.k1; ++k) {
            if (a.bl.n && a.bl.inside(c.i, c.j, k)) continue;
            const long l = (long)c.i + a.sj * c.j + a.sk * k;
            bool allp = true, alln = true;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const double ph = a.phi[l + cn.off[q]];
                if (!(ph > kEpsSims)) allp = false;
                if (!(ph < -kEpsSims)) alln = false;
            }
            if (allp) {
                double V[8][3];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    V[q][0] = a.x[l + cn.off[q]]; V[q][1] = a.y[l + cn.off[q]]; V[q][2] = a.z[l + cn.off[q]];
                }
                v[2] += hex_volume(V);
            }
            cut = cut || (!allp && !alln);
        }
    }
    __shared__ int anycut;
    if (threadIdx.x == 0 && threadIdx.y == 0) anycut = 0;
    __syncthreads();
    if (cut) anycut = 1;
    block_sum<3>(v, a.partial + (long)blockIdx.x * 3);   // contains a barrier
    if (threadIdx.x == 0 && threadIdx.y == 0) a.cutflag[blockIdx.x] = anycut;
}

// CellAnalysis of the cut cells (TetrahedronMethods.F90:1823-2121) in the
// workgroups pass 1 flagged; their K- and second-tetrahedra sums replace the
// zeros pass 1 wrote
__global__ __launch_bounds__(VTX * VTY) void volume_cut(VolArgs a)
{
    if (!a.cutflag[blockIdx.x]) return;
    const CellTile c = tile_of(a);
    const Corners cn(a.sj, a.sk);
    const int tid = threadIdx.y * VTX + threadIdx.x;
    // list the cut cells: layer by layer, wavefront by wavefront, lane order
    __shared__ int list[VKC * VTX * VTY];
    __shared__ int wcount[VTY];
    int n = 0;
    for (int k = c.k0; k < c.k1; ++k) {
        bool cut = false;
        if (c.ok && !(a.bl.n && a.bl.inside(c.i, c.j, k))) {
            const long l = (long)c.i + a.sj * c.j + a.sk * k;
            bool allp = true, alln = true;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const double ph = a.phi[l + cn.off[q]];
                if (!(ph > kEpsSims)) allp = false;
                if (!(ph < -kEpsSims)) alln = false;
            }
            cut = !allp && !alln;
        }
        const unsigned long long bal = __ballot(cut);
        if (threadIdx.x == 0) wcount[threadIdx.y] = __popcll(bal);
        __syncthreads();
        int base = n;
        for (int w = 0; w < (int)threadIdx.y; ++w) base += wcount[w];
        int tot = 0;
        for (int w = 0; w < VTY; ++w) tot += wcount[w];
        if (cut) {
            const unsigned long long below = threadIdx.x ? (bal & ((1ull << threadIdx.x) - 1ull)) : 0ull;
            list[base + __popcll(below)] = (k - c.k0) * (VTX * VTY) + tid;
        }
        n += tot;
        __syncthreads();
    }
    double v[2] = {0.0, 0.0};
    const int ti = blockIdx.x % a.tiles_i, tj = (blockIdx.x / a.tiles_i) % a.tiles_j;
    for (int e = tid; e < n; e += VTX * VTY) {
        const int code = list[e];
        const int kk = c.k0 + code / (VTX * VTY), t = code % (VTX * VTY);
        const int ii = ti * VTX + t % VTX, jj = tj * VTY + t / VTX;
        const long l = (long)ii + a.sj * jj + a.sk * kk;
        double ph[8], V[8][3];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            ph[q] = a.phi[l + cn.off[q]];
            V[q][0] = a.x[l + cn.off[q]]; V[q][1] = a.y[l + cn.off[q]]; V[q][2] = a.z[l + cn.off[q]];
        }
        double vk, vs;
        cell_analysis(ph, V, vk, vs);
        v[0] += vk; v[1] += vs;
    }
    block_sum<2>(v, a.partial + (long)blockIdx.x * 3);
}

// AdvectionNodes (TotalWaterVolume.F90:187-233): a cell's lower corner node
// n is next to the interface when |phi(n)| < eps_sims or a node of its 3x3x3
// neighbourhood (clipped to the block) has the opposite sign or |phi| <
// eps_sims; the mask is 1 on the neighbourhood of every such node.  Pass 1
// writes the per-node flag (min / max / min |.| over the neighbourhood), pass
// 2 the dilation, so every mask entry is written exactly once.  Nodes in a
// blanking box are not flagged and nodes strictly inside one are left out of
// the neighbourhood search (TotalWaterVolume.F90:139-160, 190-207).
struct MaskArgs {
    const double *phi;
    uint8_t *chg;
    int32_t *mask;
    int ni, nj, nk;
    long sj, sk;
    int tiles_i, tiles_j;
    Blank bl;           // nodes in a box are not interface nodes; nodes strictly inside are not searched
};

__global__ __launch_bounds__(VTX * VTY) void mask_change(MaskArgs a)
{
    const int b = blockIdx.x;
    const int ti = b % a.tiles_i, t2 = b / a.tiles_i;
    const int i = ti * VTX + threadIdx.x, j = (t2 % a.tiles_j) * VTY + threadIdx.y, k = t2 / a.tiles_j;
    if (i >= a.ni || j >= a.nj) return;
    const long l = (long)i + a.sj * j + a.sk * k;
    uint8_t flag = 0;
    if (i <= a.ni - 2 && j <= a.nj - 2 && k <= a.nk - 2 && !(a.bl.n && a.bl.inside(i, j, k))) {
        const double p = a.phi[l];
        double mn = p, mx = p, ma = fabs(p);
        const int km = max(k - 1, 0), kp = min(k + 1, a.nk - 1);
        const int jm = max(j - 1, 0), jp = min(j + 1, a.nj - 1);
        const int im = max(i - 1, 0), ip = min(i + 1, a.ni - 1);
        for (int kk = km; kk <= kp; ++kk)
            for (int jj = jm; jj <= jp; ++jj)
                for (int ii = im; ii <= ip; ++ii) {
                    if (a.bl.n && a.bl.interior(ii, jj, kk)) continue;
                    const double q = a.phi[(long)ii + a.sj * jj + a.sk * kk];
                    mn = fmin(mn, q); mx = fmax(mx, q); ma = fmin(ma, fabs(q));
                }
        // sign(p) * q < 0 for some q  <=>  min q < 0 (p >= +0) or max q > 0 (p <= -0)
        const bool opp = (sign1(p) > 0.0) ? (mn < 0.0) : (mx > 0.0);
        flag = (fabs(p) < kEpsSims || opp || ma < kEpsSims) ? 1 : 0;
    }
    a.chg[l] = flag;
}

__global__ __launch_bounds__(VTX * VTY) void mask_dilate(MaskArgs a)
{
    const int b = blockIdx.x;
    const int ti = b % a.tiles_i, t2 = b / a.tiles_i;
    const int i = ti * VTX + threadIdx.x, j = (t2 % a.tiles_j) * VTY + threadIdx.y, k = t2 / a.tiles_j;
    if (i >= a.ni || j >= a.nj) return;
    const int km = max(k - 1, 0), kp = min(k + 1, a.nk - 1);
    const int jm = max(j - 1, 0), jp = min(j + 1, a.nj - 1);
    const int im = max(i - 1, 0), ip = min(i + 1, a.ni - 1);
    int m = 0;
    for (int kk = km; kk <= kp; ++kk)
        for (int jj = jm; jj <= jp; ++jj)
            for (int ii = im; ii <= ip; ++ii) m |= a.chg[(long)ii + a.sj * jj + a.sk * kk];
    a.mask[(long)i + a.sj * j + a.sk * k] = m;
}

// sum of the per-workgroup partials in a fixed order (1024 threads)
__global__ __launch_bounds__(1024) void volume_final(const double *partial, long nblocks, double *vol)
{
    __shared__ double sm[1024][3];
    double s[3] = {0.0, 0.0, 0.0};
    for (long b = threadIdx.x; b < nblocks; b += 1024)
        for (int q = 0; q < 3; ++q) s[q] += partial[b * 3 + q];
    for (int q = 0; q < 3; ++q) sm[threadIdx.x][q] = s[q];
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 3; ++q) sm[threadIdx.x][q] += sm[threadIdx.x + w][q];
        __syncthreads();
    }
    if (threadIdx.x < 3) vol[threadIdx.x] = sm[0][threadIdx.x];
}

long vol_blocks(const Blk &b)
{
    return (long)((b.e.ni - 1 + VTX - 1) / VTX) * ((b.e.nj - 1 + VTY - 1) / VTY) *
           ((b.e.nk - 1 + VKC - 1) / VKC);
}

}  // namespace

// scratch of one evaluation: per-workgroup partials and cut flags, and the
// per-node change flags of the mask when the mask is asked for
static size_t volume_scratch_bytes(const Blk &b, bool with_mask)
{
    const size_t nb = (size_t)vol_blocks(b);
    return ((nb * 3 * sizeof(double) + 255) & ~(size_t)255) + ((nb * sizeof(int32_t) + 255) & ~(size_t)255) +
           (with_mask ? (((size_t)b.n + 255) & ~(size_t)255) : 0) + 1024;
}

size_t water_volume_work(const Blk &b)
{
    return volume_scratch_bytes(b, true);
}

int launch_water_volume(const Blk &b, const double *phi, double *vol, int32_t *mask, Carver &w,
                        hipStream_t s, bool timed)
{
    const long nblk = vol_blocks(b);
    double *partial = (double *)w.take((size_t)nblk * 3 * sizeof(double));
    int32_t *cutflag = (int32_t *)w.take((size_t)nblk * sizeof(int32_t));
    uint8_t *chg = mask ? (uint8_t *)w.take((size_t)b.n) : nullptr;
    if (!partial || !cutflag || (mask && !chg)) return fail(LSM_ERR_WORKSPACE, "volume workspace too small");
    int rc;
    const dim3 blk(VTX, VTY, 1);
    // vol == NULL: the AdvectionNodes mask alone
    if (vol) {
        VolArgs a{phi, b.x, b.y, b.z, partial, cutflag, b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk,
                  (b.e.ni - 1 + VTX - 1) / VTX, (b.e.nj - 1 + VTY - 1) / VTY, b.bl};
        if (timed) {
            KTimer kt(LSM_KT_VOLUME, s);
            volume_wet<<<(unsigned)nblk, blk, 0, s>>>(a);
            volume_cut<<<(unsigned)nblk, blk, 0, s>>>(a);
        } else {
            volume_wet<<<(unsigned)nblk, blk, 0, s>>>(a);
            volume_cut<<<(unsigned)nblk, blk, 0, s>>>(a);
        }
        if ((rc = check_hip(hipGetLastError(), "volume kernels"))) return rc;
        volume_final<<<1, 1024, 0, s>>>(partial, nblk, vol);
        if ((rc = check_hip(hipGetLastError(), "volume_final"))) return rc;
    }
    if (mask) {
        MaskArgs m{phi, chg, mask, b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk, (b.e.ni + VTX - 1) / VTX,
                   (b.e.nj + VTY - 1) / VTY, b.bl};
        const unsigned mb = (unsigned)(m.tiles_i * m.tiles_j * b.e.nk);
        mask_change<<<mb, blk, 0, s>>>(m);
        mask_dilate<<<mb, blk, 0, s>>>(m);
        if ((rc = check_hip(hipGetLastError(), "mask kernels"))) return rc;
    }
    return LSM_OK;
}

}  // namespace lsm

// ---------------------------------------------------------------------------
// allocation-free water volume: the scratch of the evaluation lives in a
// fixed device array of the library, so the call allocates nothing
// ---------------------------------------------------------------------------
#include <mutex>

namespace {

constexpr size_t kVolumeScratchBytes = (size_t)32 << 20;   // 32 MiB per device

// The array's name is a host-side stand-in in host code: its device address
// is taken once per device with hipGetSymbolAddress and only that address is
// ever passed to kernels.
static __device__ __align__(256) unsigned char scratch[kVolumeScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_dev[64];

int scratch_address(void **out)
{
    int dev = 0;
    int rc = lsm::check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= 64) return lsm::fail(LSM_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        rc = lsm::check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        if (!p) return lsm::fail(LSM_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return LSM_OK;
}

}  // namespace

extern "C" int lsm_water_volume_noalloc(const LsmBlock *blk, const double *phi, double *vol_out,
                                        int32_t *adv_mask, void *stream)
{
    using namespace lsm;
    Blk b;
    int rc = make_blk(blk, &b, false, true, kBlankUse);
    if (rc) return rc;
    if (!phi || !vol_out) return fail(LSM_ERR_ARG, "NULL array");
    const size_t need = volume_scratch_bytes(b, adv_mask != nullptr);
    if (need > kVolumeScratchBytes)
        return fail(LSM_ERR_WORKSPACE,
                    "lsm_water_volume_noalloc: a %dx%dx%d block needs %zu bytes of scratch%s, the "
                    "library's fixed scratch holds %zu (use lsm_water_volume with a workspace)",
                    b.e.ni, b.e.nj, b.e.nk, need, adv_mask ? " with the mask" : "", kVolumeScratchBytes);
    void *p = nullptr;
    if ((rc = scratch_address(&p))) return rc;
    Carver c{(char *)p, kVolumeScratchBytes, 0};
    return launch_water_volume(b, phi, vol_out, adv_mask, c, (hipStream_t)stream, false);
}
