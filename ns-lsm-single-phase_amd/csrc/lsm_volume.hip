// Water volume of the phi > 0 region (TotalWaterVolume + TetrahedronMethods)
// and the AdvectionNodes interface mask.
//
// One thread per grid cell: a fully wet cell adds its Davies-Salmond
// hexahedron volume, a cut cell is split into six tetrahedra and the cut ones
// go through marching tetrahedra.  Per-workgroup partial sums (wavefront
// shuffles, then LDS) are written to scratch and summed by a single
// workgroup in a fixed order, so the result does not depend on scheduling.
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
__device__ __noinline__ void cell_analysis(const double ph[8], const double V[8][3], double &pc_vol,
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

struct VolArgs {
    const double *phi, *x, *y, *z;
    int32_t *mask;
    double *partial;   // [gridDim.x][3]
    int ni, nj, nk;
    long sj, sk;
    long ncells;
};

// TotalWaterVolume.F90:136-294
__global__ __launch_bounds__(256) void volume_kernel(VolArgs a)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double vk = 0.0, vs = 0.0, vb = 0.0;
    if (c < a.ncells) {
        const int ci = a.ni - 1, cj = a.nj - 1;
        const int i = (int)(c % ci);
        const long r = c / ci;
        const int j = (int)(r % cj);
        const int k = (int)(r / cj);
        const long l = (long)i + a.sj * j + a.sk * k;
        const long off[8] = {0, 1, 1 + a.sj, a.sj, a.sk, 1 + a.sk, 1 + a.sj + a.sk, a.sj + a.sk};
        double ph[8];
        bool allp = true, alln = true;
        for (int q = 0; q < 8; ++q) {
            ph[q] = a.phi[l + off[q]];
            if (!(ph[q] > kEpsSims)) allp = false;
            if (!(ph[q] < -kEpsSims)) alln = false;
        }
        if (a.mask) {
            // AdvectionNodes: 3x3x3 neighbourhood of a node next to a sign change
            const int km = max(k - 1, 0), kp = min(k + 1, a.nk - 1);
            const int jm = max(j - 1, 0), jp = min(j + 1, a.nj - 1);
            const int im = max(i - 1, 0), ip = min(i + 1, a.ni - 1);
            const double p = ph[0];
            const double sgn = sign1(p);
            bool change = fabs(p) < kEpsSims;
            for (int kk = km; kk <= kp && !change; ++kk)
                for (int jj = jm; jj <= jp && !change; ++jj)
                    for (int ii = im; ii <= ip; ++ii) {
                        const double q = a.phi[(long)ii + a.sj * jj + a.sk * kk];
                        if (sgn * q < 0.0 || fabs(q) < kEpsSims) { change = true; break; }
                    }
            if (change)
                for (int kk = km; kk <= kp; ++kk)
                    for (int jj = jm; jj <= jp; ++jj)
                        for (int ii = im; ii <= ip; ++ii) a.mask[(long)ii + a.sj * jj + a.sk * kk] = 1;
        }
        if (!alln) {
            double V[8][3];
            for (int q = 0; q < 8; ++q) {
                V[q][0] = a.x[l + off[q]]; V[q][1] = a.y[l + off[q]]; V[q][2] = a.z[l + off[q]];
            }
            if (allp) vb = hex_volume(V);
            else cell_analysis(ph, V, vk, vs);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        vk += __shfl_xor(vk, o, 64);
        vs += __shfl_xor(vs, o, 64);
        vb += __shfl_xor(vb, o, 64);
    }
    __shared__ double sm[4][3];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sm[wv][0] = vk; sm[wv][1] = vs; sm[wv][2] = vb; }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int q = threadIdx.x;
        a.partial[(long)blockIdx.x * 3 + q] = ((sm[0][q] + sm[1][q]) + sm[2][q]) + sm[3][q];
    }
}

__global__ __launch_bounds__(256) void volume_final(const double *partial, long nblocks, double *vol)
{
    __shared__ double sm[256][3];
    double s[3] = {0.0, 0.0, 0.0};
    for (long b = threadIdx.x; b < nblocks; b += 256)
        for (int q = 0; q < 3; ++q) s[q] += partial[b * 3 + q];
    for (int q = 0; q < 3; ++q) sm[threadIdx.x][q] = s[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 3; ++q) sm[threadIdx.x][q] += sm[threadIdx.x + w][q];
        __syncthreads();
    }
    if (threadIdx.x < 3) vol[threadIdx.x] = sm[0][threadIdx.x];
}

__global__ __launch_bounds__(256) void fill_i32(int32_t *p, long n, int32_t v)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) p[l] = v;
}

long vol_blocks(const Blk &b)
{
    const long nc = (long)(b.e.ni - 1) * (b.e.nj - 1) * (b.e.nk - 1);
    return (nc + 255) / 256;
}

}  // namespace

size_t water_volume_work(const Blk &b) { return (size_t)vol_blocks(b) * 3 * sizeof(double) + 1024; }

int launch_water_volume(const Blk &b, const double *phi, double *vol, int32_t *mask, Carver &w,
                        hipStream_t s)
{
    const long nblk = vol_blocks(b);
    double *partial = (double *)w.take((size_t)nblk * 3 * sizeof(double));
    if (!partial) return fail(LSM_ERR_WORKSPACE, "volume workspace too small");
    int rc;
    if (mask) {
        fill_i32<<<(unsigned)((b.n + 255) / 256), 256, 0, s>>>(mask, b.n, 0);
        if ((rc = check_hip(hipGetLastError(), "fill_i32"))) return rc;
    }
    VolArgs a{phi, b.x, b.y, b.z, mask, partial, b.e.ni, b.e.nj, b.e.nk, b.e.sj, b.e.sk,
              (long)(b.e.ni - 1) * (b.e.nj - 1) * (b.e.nk - 1)};
    {
        KTimer kt(LSM_KT_VOLUME, s);
        volume_kernel<<<(unsigned)nblk, 256, 0, s>>>(a);
    }
    if ((rc = check_hip(hipGetLastError(), "volume_kernel"))) return rc;
    volume_final<<<1, 256, 0, s>>>(partial, nblk, vol);
    return check_hip(hipGetLastError(), "volume_final");
}

}  // namespace lsm
