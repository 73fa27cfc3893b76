// MPFA-O transmissibilities and CSR assembly on gfx950 (Cartesian grids).
//
// Reference: O_MPFA.calculateTransmissibilities and the interaction region
// (rs/src/no/uib/cipr/rs/geometry/flux/O_MPFA.java:54-457), MPFA helpers
// (MPFA.java:43-54, 109-120, 314-354), Connection checks
// (geometry/Connection.java:61-107), the structured mesh numbering and
// geometry (meshgen/structured/CartesianTopology{,2D,3D}.java, HexCell.java,
// QuadCell.java, BilinearMap.java, TrilinearMap.java), the sparsity pattern
// (numerics/LinearSolver.java:111-168) and the flux assembly
// (numerics/PotentialCalculator.java:319-333).
//
// Kernels:
//   k_corner   one lane per corner point: region, basis solves, local
//              A,B,C,D, T = C A^-1 B - D (LAPACK 3.0 dgetf2/dgetrs and
//              reference-BLAS dgemm operation order), staged per
//              (connection, corner slot).
//   k_merge    one lane per connection: storeTransmissibilities replayed in
//              corner-point order (the order rs visits points), then
//              Connection.check.
//   k_pattern  one lane per row: {r} U MD elements of incident connections,
//              sorted (count pass + fill pass).
//   k_assemble one lane per row: A[r,k] accumulated over incident
//              connections in ascending connection order, MD order inside,
//              which is the order of PotentialCalculator.discretizeFlux.
#include <algorithm>

#include "rsgpu_impl.hpp"

namespace {

struct v3 {
    double x, y, z;
};
__device__ __forceinline__ v3 mk(double x, double y, double z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 vmul(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double vnorm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ v3 vcross(v3 a, v3 v)
{
    return mk(a.y * v.z - a.z * v.y, a.z * v.x - a.x * v.z, a.x * v.y - a.y * v.x);
}

struct Grid {
    int dim, nx, ny, nz, nf;
    int ncells, nconn, npoints;
    const double *xs, *ys, *zs;
};

__device__ __forceinline__ v3 point_at(const Grid &g, int i, int j, int k)
{
    return mk(g.xs[i], g.ys[j], g.dim == 3 ? g.zs[k] : 0.0);
}
__device__ __forceinline__ int elem_index(const Grid &g, int i, int j, int k)
{
    return i + j * g.nx + k * g.nx * g.ny;
}
__device__ __forceinline__ bool elem_valid(const Grid &g, int i, int j, int k)
{
    return i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
}

// forward-face connection numbering (z, then y, then x faces in 3-D)
__device__ int conn_index(const Grid &g, int i, int j, int k, int f)
{
    const int ni = g.nx, nj = g.ny, nk = g.nz;
    if (g.dim == 3) {
        if (f == 1)
            return i + j * ni + k * ni * nj;
        if (f == 3)
            return i + j * ni + k * ni * (nj - 1) + ni * nj * (nk - 1);
        return i + j * (ni - 1) + k * (ni - 1) * nj + ni * nj * (nk - 1) +
               ni * (nj - 1) * nk;
    }
    if (f == 1)
        return i + j * ni;
    return i + j * (ni - 1) + ni * (nj - 1);
}

__device__ void face_offset(const Grid &g, int f, int &a, int &b, int &c)
{
    a = b = c = 0;
    if (g.dim == 3) {
        switch (f) {
        case 0: c = -1; break;
        case 1: c = 1; break;
        case 2: b = -1; break;
        case 3: b = 1; break;
        case 4: a = -1; break;
        default: a = 1; break;
        }
    } else {
        switch (f) {
        case 0: b = -1; break;
        case 1: b = 1; break;
        case 2: a = -1; break;
        default: a = 1; break;
        }
    }
}

// connection of a face (-1 boundary) and whether it is the here side
__device__ int face_connection(const Grid &g, int i, int j, int k, int f, int &is_here)
{
    int a, b, c;
    face_offset(g, f, a, b, c);
    if (!elem_valid(g, i + a, j + b, k + c))
        return -1;
    if (f & 1) {
        is_here = 1;
        return conn_index(g, i, j, k, f);
    }
    is_here = 0;
    return conn_index(g, i + a, j + b, k + c, f + 1);
}

__device__ v3 bilinear_center(v3 p0, v3 p1, v3 p2, v3 p3)
{
    const double a = .5, b = .5;
    const double k0 = (1 - a) * (1 - b), k1 = a * (1 - b), k2 = (1 - a) * b, k3 = a * b;
    return mk(k0 * p0.x + k1 * p1.x + k2 * p2.x + k3 * p3.x,
              k0 * p0.y + k1 * p1.y + k2 * p2.y + k3 * p3.y,
              k0 * p0.z + k1 * p1.z + k2 * p2.z + k3 * p3.z);
}

__device__ void bilinear_area_normal(v3 x1, v3 x2, v3 x3, v3 x4, double &area, v3 &normal)
{
    v3 a = vcross(vsub(x2, x1), vsub(x3, x1));
    v3 b = vcross(vsub(x2, x1), vsub(x4, x2));
    v3 c = vcross(vsub(x4, x3), vsub(x3, x1));
    v3 d = vcross(vsub(x4, x3), vsub(x4, x2));
    v3 n = vmul(vadd(a, vadd(b, vadd(c, d))), .25);
    double length = vnorm(n);
    area = length;
    normal = (length < 1.0e-9) ? mk(0, 0, 0) : vmul(n, 1.0 / length);
}

__device__ void elem_points3(const Grid &g, int i, int j, int k, v3 cp[8])
{
    cp[0] = point_at(g, i, j, k);
    cp[1] = point_at(g, i + 1, j, k);
    cp[2] = point_at(g, i + 1, j + 1, k);
    cp[3] = point_at(g, i, j + 1, k);
    cp[4] = point_at(g, i, j, k + 1);
    cp[5] = point_at(g, i + 1, j, k + 1);
    cp[6] = point_at(g, i + 1, j + 1, k + 1);
    cp[7] = point_at(g, i, j + 1, k + 1);
}

__device__ v3 elem_center(const Grid &g, int i, int j, int k)
{
    if (g.dim == 3) {
        v3 cp[8];
        elem_points3(g, i, j, k, cp);
        const v3 x1 = cp[0], x2 = cp[1], x3 = cp[3], x4 = cp[2];
        const v3 x5 = cp[4], x6 = cp[5], x7 = cp[7], x8 = cp[6];
        const double a = .5, b = .5, c = .5;
#define TRI(C)                                                                  \
    (c * (b * (a * x4.C + (1 - a) * x3.C) + (1 - b) * (a * x2.C + (1 - a) * x1.C)) + \
     (1 - c) * (b * (a * x8.C + (1 - a) * x7.C) + (1 - b) * (a * x6.C + (1 - a) * x5.C)))
        return mk(TRI(x), TRI(y), TRI(z));
#undef TRI
    }
    v3 p0 = point_at(g, i, j, 0), p1 = point_at(g, i + 1, j, 0);
    v3 p2 = point_at(g, i + 1, j + 1, 0), p3 = point_at(g, i, j + 1, 0);
    return bilinear_center(p0, p1, p3, p2);
}

__device__ void face_geom(const Grid &g, int i, int j, int k, int f, v3 &center,
                          v3 &normal, double &area)
{
    if (g.dim == 3) {
        v3 cp[8];
        elem_points3(g, i, j, k, cp);
        const v3 p0 = cp[0], p1 = cp[1], p2 = cp[3], p3 = cp[2];
        const v3 p4 = cp[4], p5 = cp[5], p6 = cp[7], p7 = cp[6];
        v3 q0, q1, q2, q3;
        double scale = 1.0;
        switch (f) {
        case 0: q0 = p0; q1 = p1; q2 = p2; q3 = p3; break;
        case 1: q0 = p4; q1 = p5; q2 = p6; q3 = p7; scale = -1.0; break;
        case 2: q0 = p0; q1 = p1; q2 = p4; q3 = p5; scale = -1.0; break;
        case 3: q0 = p2; q1 = p3; q2 = p6; q3 = p7; break;
        case 4: q0 = p0; q1 = p2; q2 = p4; q3 = p6; break;
        default: q0 = p1; q1 = p3; q2 = p5; q3 = p7; scale = -1.0; break;
        }
        v3 nrm;
        bilinear_area_normal(q0, q1, q2, q3, area, nrm);
        normal = vmul(nrm, scale);
        center = bilinear_center(q0, q1, q2, q3);
        return;
    }
    v3 p0 = point_at(g, i, j, 0), p1 = point_at(g, i + 1, j, 0);
    v3 p2 = point_at(g, i + 1, j + 1, 0), p3 = point_at(g, i, j + 1, 0);
    v3 a, b;
    switch (f) {
    case 0: a = p0; b = p1; break;
    case 1: a = p2; b = p3; break;
    case 2: a = p3; b = p0; break;
    default: a = p1; b = p2; break;
    }
    v3 n = mk(b.y - a.y, -(b.x - a.x), 0);
    n = vmul(n, 1 / vnorm(n));
    normal = n;
    area = vnorm(vsub(b, a));
    center = vmul(vadd(a, b), 0.5);
}

// MPFA.calculateNormalArea (MPFA.java:314-354)
__device__ v3 normal_area(const Grid &g, int i, int j, int k, int f, int pi, int pj, int pk)
{
    v3 c, n;
    double area;
    face_geom(g, i, j, k, f, c, n, area);
    if (g.dim == 2)
        return vmul(n, -area / 2);
    // interface corner order (CartesianTopology3D.getInterfacePoints)
    const int ci[8][3] = {{i, j, k},         {i + 1, j, k},         {i + 1, j + 1, k},
                          {i, j + 1, k},     {i, j, k + 1},         {i + 1, j, k + 1},
                          {i + 1, j + 1, k + 1}, {i, j + 1, k + 1}};
    const int ord[6][4] = {{0, 1, 2, 3}, {7, 6, 5, 4}, {4, 5, 1, 0},
                           {6, 7, 3, 2}, {7, 4, 0, 3}, {5, 6, 2, 1}};
    int s = 0;
    for (; s < 4; ++s) {
        const int *q = ci[ord[f][s]];
        if (q[0] == pi && q[1] == pj && q[2] == pk)
            break;
    }
    const int i2 = s == 0 ? 3 : (s - 1) % 4, i3 = (s + 1) % 4, i4 = (s + 2) % 4;
    const int *a1 = ci[ord[f][s]], *a2 = ci[ord[f][i2]], *a3 = ci[ord[f][i3]],
              *a4 = ci[ord[f][i4]];
    v3 x1 = point_at(g, a1[0], a1[1], a1[2]), x2 = point_at(g, a2[0], a2[1], a2[2]);
    v3 x3 = point_at(g, a3[0], a3[1], a3[2]), x4 = point_at(g, a4[0], a4[1], a4[2]);
    v3 v21 = vsub(x2, x1), v31 = vsub(x3, x1), v42 = vsub(x4, x2), v43 = vsub(x4, x3);
    v3 n1 = vmul(vcross(v21, v31), 9);
    v3 n2 = vmul(vcross(v21, v42), 3);
    v3 n3 = vmul(vcross(v43, v31), 3);
    v3 n4 = vcross(v43, v42);
    return vmul(vadd(vadd(vadd(n1, n2), n3), n4), 1. / 64.);
}

__device__ __forceinline__ v3 tensor_mult(const double *K, v3 v)
{
    const double xx = K[0], yy = K[1], zz = K[2], xy = K[3], xz = K[4], yz = K[5];
    return mk(xx * v.x + xy * v.y + xz * v.z, xy * v.x + yy * v.y + yz * v.z,
              xz * v.x + yz * v.y + zz * v.z);
}

// LAPACK 3.0 dgesv (dgetf2 + dgetrs), column-major, n <= 12
__device__ int dgesv(int n, int nrhs, double *a, int lda, double *b, int ldb)
{
    int ipiv[12];
    int info = 0;
    for (int j = 0; j < n; ++j) {
        int jp = j;
        double amax = fabs(a[j + j * lda]);
        for (int i = j + 1; i < n; ++i) {
            const double v = fabs(a[i + j * lda]);
            if (v > amax) {
                amax = v;
                jp = i;
            }
        }
        ipiv[j] = jp;
        if (a[jp + j * lda] != 0) {
            if (jp != j)
                for (int c = 0; c < n; ++c) {
                    const double t = a[j + c * lda];
                    a[j + c * lda] = a[jp + c * lda];
                    a[jp + c * lda] = t;
                }
            if (j < n - 1) {
                const double r = 1.0 / a[j + j * lda];
                for (int i = j + 1; i < n; ++i)
                    a[i + j * lda] *= r;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        if (j < n - 1)
            for (int c = j + 1; c < n; ++c) {
                const double yj = a[j + c * lda];
                if (yj != 0) {
                    const double temp = -1.0 * yj;
                    for (int i = j + 1; i < n; ++i)
                        a[i + c * lda] = a[i + c * lda] + a[i + j * lda] * temp;
                }
            }
    }
    if (info)
        return 1;
    for (int j = 0; j < n; ++j)
        if (ipiv[j] != j)
            for (int c = 0; c < nrhs; ++c) {
                const double t = b[j + c * ldb];
                b[j + c * ldb] = b[ipiv[j] + c * ldb];
                b[ipiv[j] + c * ldb] = t;
            }
    for (int c = 0; c < nrhs; ++c)
        for (int k = 0; k < n; ++k)
            if (b[k + c * ldb] != 0)
                for (int i = k + 1; i < n; ++i)
                    b[i + c * ldb] = b[i + c * ldb] - b[k + c * ldb] * a[i + k * lda];
    for (int c = 0; c < nrhs; ++c)
        for (int k = n - 1; k >= 0; --k)
            if (b[k + c * ldb] != 0) {
                b[k + c * ldb] = b[k + c * ldb] / a[k + k * lda];
                for (int i = 0; i < k; ++i)
                    b[i + c * ldb] = b[i + c * ldb] - b[k + c * ldb] * a[i + k * lda];
            }
    return 0;
}

// dgesv keeping the pivots (the adjoint solves with A')
__device__ int dgesv_piv(int n, int nrhs, double *a, int lda, double *b, int ldb, int *ipiv)
{
    int info = 0;
    for (int j = 0; j < n; ++j) {
        int jp = j;
        double amax = fabs(a[j + j * lda]);
        for (int i = j + 1; i < n; ++i) {
            const double v = fabs(a[i + j * lda]);
            if (v > amax) {
                amax = v;
                jp = i;
            }
        }
        ipiv[j] = jp;
        if (a[jp + j * lda] != 0) {
            if (jp != j)
                for (int c = 0; c < n; ++c) {
                    const double t = a[j + c * lda];
                    a[j + c * lda] = a[jp + c * lda];
                    a[jp + c * lda] = t;
                }
            if (j < n - 1) {
                const double r = 1.0 / a[j + j * lda];
                for (int i = j + 1; i < n; ++i)
                    a[i + j * lda] *= r;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        if (j < n - 1)
            for (int c = j + 1; c < n; ++c) {
                const double yj = a[j + c * lda];
                if (yj != 0) {
                    const double temp = -1.0 * yj;
                    for (int i = j + 1; i < n; ++i)
                        a[i + c * lda] = a[i + c * lda] + a[i + j * lda] * temp;
                }
            }
    }
    if (info)
        return 1;
    for (int j = 0; j < n; ++j)
        if (ipiv[j] != j)
            for (int c = 0; c < nrhs; ++c) {
                const double t = b[j + c * ldb];
                b[j + c * ldb] = b[ipiv[j] + c * ldb];
                b[ipiv[j] + c * ldb] = t;
            }
    for (int c = 0; c < nrhs; ++c)
        for (int k = 0; k < n; ++k)
            if (b[k + c * ldb] != 0)
                for (int i = k + 1; i < n; ++i)
                    b[i + c * ldb] = b[i + c * ldb] - b[k + c * ldb] * a[i + k * lda];
    for (int c = 0; c < nrhs; ++c)
        for (int k = n - 1; k >= 0; --k)
            if (b[k + c * ldb] != 0) {
                b[k + c * ldb] = b[k + c * ldb] / a[k + k * lda];
                for (int i = 0; i < k; ++i)
                    b[i + c * ldb] = b[i + c * ldb] - b[k + c * ldb] * a[i + k * lda];
            }
    return 0;
}

struct RIntf {
    signed char ei, ej, ek, f;   // owning element offsets relative to corner - 1
    signed char elocal, boundary, is_here, lidx;
    int conn;
};

// Stage layout per connection: 4 corner slots (2 in 2-D) x 8 elements.
#define MPFA_SLOTS 4
#define MPFA_NE 8

__global__ void __launch_bounds__(64)
    k_corner(Grid g, const double *__restrict__ K6, double continuity,
             double *__restrict__ st_t, int *__restrict__ st_e,
             signed char *__restrict__ st_n, int *__restrict__ err)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.npoints)
        return;
    const int npx = g.nx + 1, npy = g.ny + 1;
    const int pi = p % npx, pj = (p / npx) % npy, pk = g.dim == 3 ? p / (npx * npy) : 0;
    // ---- region: elements i-major, interfaces in orientation order
    int ne = 0, nif = 0;
    int eidx[8], ei_[8], ej_[8], ek_[8];
    RIntf I[24];
    int eintf[8][3], neintf[8];
    const int kmin = g.dim == 3 ? pk - 1 : 0, kmax = g.dim == 3 ? pk : 0;
    for (int i = pi - 1; i <= pi; ++i)
        for (int j = pj - 1; j <= pj; ++j)
            for (int k = kmin; k <= kmax; ++k) {
                if (!elem_valid(g, i, j, k))
                    continue;
                const int le = ne++;
                eidx[le] = elem_index(g, i, j, k);
                ei_[le] = i;
                ej_[le] = j;
                ek_[le] = k;
                neintf[le] = 0;
                for (int f = 0; f < g.nf; ++f) {
                    bool on;
                    if (g.dim == 3) {
                        switch (f) {
                        case 0: on = (pk == k); break;
                        case 1: on = (pk == k + 1); break;
                        case 2: on = (pj == j); break;
                        case 3: on = (pj == j + 1); break;
                        case 4: on = (pi == i); break;
                        default: on = (pi == i + 1); break;
                        }
                    } else {
                        switch (f) {
                        case 0: on = (pj == j); break;
                        case 1: on = (pj == j + 1); break;
                        case 2: on = (pi == i); break;
                        default: on = (pi == i + 1); break;
                        }
                    }
                    if (!on)
                        continue;
                    RIntf &R = I[nif];
                    R.ei = (signed char)(i - pi + 1);
                    R.ej = (signed char)(j - pj + 1);
                    R.ek = (signed char)(k - pk + 1);
                    R.f = (signed char)f;
                    R.elocal = (signed char)le;
                    int ih = 0;
                    R.conn = face_connection(g, i, j, k, f, ih);
                    R.boundary = R.conn < 0;
                    R.is_here = R.boundary ? 1 : (signed char)ih;
                    eintf[le][neintf[le]++] = nif;
                    nif++;
                }
            }
    int nc = 0;
    for (int a = 0; a < nif; ++a) {
        int found = -1;
        if (!I[a].boundary)
            for (int b = 0; b < a; ++b)
                if (!I[b].boundary && I[b].conn == I[a].conn) {
                    found = I[b].lidx;
                    break;
                }
        I[a].lidx = (signed char)(found >= 0 ? found : nc++);
    }
    // ---- local matrices (column-major, ld = nc)
    double A[144], B[96], C[144], D[96];
    for (int q = 0; q < nc * nc; ++q) {
        A[q] = 0;
        C[q] = 0;
    }
    for (int q = 0; q < nc * ne; ++q) {
        B[q] = 0;
        D[q] = 0;
    }
    const v3 corner = point_at(g, pi, pj, pk);
    for (int le = 0; le < ne; ++le) {
        const int nI = neintf[le], nb = nI + 1;
        double Am[16], Psi[16];
        const v3 elp = elem_center(g, ei_[le], ej_[le], ek_[le]);
        for (int q = 0; q < 16; ++q) {
            Am[q] = 0;
            Psi[q] = 0;
        }
        Am[0] = 1;
        Am[0 + 1 * nb] = elp.x;
        Am[0 + 2 * nb] = elp.y;
        if (g.dim == 3)
            Am[0 + 3 * nb] = elp.z;
        for (int q = 0; q < nI; ++q) {
            const RIntf &R = I[eintf[le][q]];
            v3 cc, nn;
            double ar;
            face_geom(g, pi - 1 + R.ei, pj - 1 + R.ej, pk - 1 + R.ek, R.f, cc, nn, ar);
            v3 cp = cc;
            if (continuity != 1)
                cp = mk((cc.x - corner.x) * continuity + corner.x,
                        (cc.y - corner.y) * continuity + corner.y,
                        (cc.z - corner.z) * continuity + corner.z);
            Am[q + 1] = 1;
            Am[q + 1 + 1 * nb] = cp.x;
            Am[q + 1 + 2 * nb] = cp.y;
            if (g.dim == 3)
                Am[q + 1 + 3 * nb] = cp.z;
        }
        for (int q = 0; q < nb; ++q)
            Psi[q + q * nb] = 1;
        if (dgesv(nb, nb, Am, nb, Psi, nb)) {
            atomicMax(err, 1);
            return;
        }
        v3 grad[4];
        for (int q = 0; q < nb; ++q)
            grad[q] = mk(Psi[1 + q * nb], Psi[2 + q * nb], g.dim == 3 ? Psi[3 + q * nb] : 0);
        v3 Kn[3];
        for (int q = 0; q < nI; ++q) {
            const RIntf &R = I[eintf[le][q]];
            Kn[q] = tensor_mult(K6 + 6 * (size_t)eidx[le],
                                normal_area(g, pi - 1 + R.ei, pj - 1 + R.ej, pk - 1 + R.ek,
                                            R.f, pi, pj, pk));
        }
        for (int q = 0; q < nI; ++q) {
            const RIntf &R = I[eintf[le][q]];
            double f = vdot(Kn[q], grad[0]);
            const int is = R.lidx, ei = R.elocal;
            B[is + ei * nc] += -f;
            if (R.is_here)
                D[is + ei * nc] += -f;
            for (int jj = 1; jj < nI + 1; ++jj) {
                f = vdot(Kn[q], grad[jj]);
                const int js = I[eintf[le][jj - 1]].lidx;
                A[is + js * nc] += f;
                if (R.is_here)
                    C[is + js * nc] += f;
            }
        }
    }
    // AiB = A \ B (in B), T = C*AiB + (-D) accumulated into D
    if (dgesv(nc, ne, A, nc, B, nc)) {
        atomicMax(err, 2);
        return;
    }
    for (int q = 0; q < nc * ne; ++q)
        D[q] = D[q] * -1;
    for (int j = 0; j < ne; ++j)
        for (int l = 0; l < nc; ++l) {
            const double bl = B[l + j * nc];
            if (bl != 0) {
                const double temp = 1 * bl;
                for (int i = 0; i < nc; ++i)
                    D[i + j * nc] = D[i + j * nc] + temp * C[i + l * nc];
            }
        }
    // stage T rows of the connections owned here (here interfaces)
    for (int a = 0; a < nif; ++a) {
        const RIntf &R = I[a];
        if (R.boundary || !R.is_here)
            continue;
        const int c = R.conn;
        // slot: this corner's rank among the face's corners in point order.
        // The face's corners differ from p in the two tangential directions;
        // rank = bit pattern of (corner > lowest corner) per tangential axis.
        int slot;
        const int fi = pi - 1 + R.ei, fj = pj - 1 + R.ej, fk = pk - 1 + R.ek;
        if (g.dim == 3) {
            const int ax = R.f >> 1;   // 0: z-face, 1: y-face, 2: x-face
            int da, db;                // tangential offsets (low axis first)
            if (ax == 0) {             // tangential x, y -> point order x fastest
                da = pi - fi;
                db = pj - fj;
            } else if (ax == 1) {      // tangential x, z
                da = pi - fi;
                db = pk - fk;
            } else {                   // tangential y, z
                da = pj - fj;
                db = pk - fk;
            }
            slot = da + 2 * db;
        } else {
            slot = (R.f >> 1) == 0 ? (pi - fi) : (pj - fj);
        }
        double *tt = st_t + ((size_t)c * MPFA_SLOTS + slot) * MPFA_NE;
        int *te = st_e + ((size_t)c * MPFA_SLOTS + slot) * MPFA_NE;
        for (int le = 0; le < ne; ++le) {
            tt[le] = D[R.lidx + le * nc];
            te[le] = eidx[le];
        }
        st_n[(size_t)c * MPFA_SLOTS + slot] = (signed char)ne;
    }
}

__global__ void __launch_bounds__(64)
    k_corner_vjp(Grid g, const double *__restrict__ K6, double continuity,
                 const int *__restrict__ md_ptr, const int *__restrict__ md_elem,
                 const double *__restrict__ tbar, double *__restrict__ st_m,
                 int *__restrict__ err)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.npoints)
        return;
    const int npx = g.nx + 1, npy = g.ny + 1;
    const int pi = p % npx, pj = (p / npx) % npy, pk = g.dim == 3 ? p / (npx * npy) : 0;
    // ---- region: elements i-major, interfaces in orientation order
    int ne = 0, nif = 0;
    int eidx[8], ei_[8], ej_[8], ek_[8];
    RIntf I[24];
    int eintf[8][3], neintf[8];
    const int kmin = g.dim == 3 ? pk - 1 : 0, kmax = g.dim == 3 ? pk : 0;
    for (int i = pi - 1; i <= pi; ++i)
        for (int j = pj - 1; j <= pj; ++j)
            for (int k = kmin; k <= kmax; ++k) {
                if (!elem_valid(g, i, j, k))
                    continue;
                const int le = ne++;
                eidx[le] = elem_index(g, i, j, k);
                ei_[le] = i;
                ej_[le] = j;
                ek_[le] = k;
                neintf[le] = 0;
                for (int f = 0; f < g.nf; ++f) {
                    bool on;
                    if (g.dim == 3) {
                        switch (f) {
                        case 0: on = (pk == k); break;
                        case 1: on = (pk == k + 1); break;
                        case 2: on = (pj == j); break;
                        case 3: on = (pj == j + 1); break;
                        case 4: on = (pi == i); break;
                        default: on = (pi == i + 1); break;
                        }
                    } else {
                        switch (f) {
                        case 0: on = (pj == j); break;
                        case 1: on = (pj == j + 1); break;
                        case 2: on = (pi == i); break;
                        default: on = (pi == i + 1); break;
                        }
                    }
                    if (!on)
                        continue;
                    RIntf &R = I[nif];
                    R.ei = (signed char)(i - pi + 1);
                    R.ej = (signed char)(j - pj + 1);
                    R.ek = (signed char)(k - pk + 1);
                    R.f = (signed char)f;
                    R.elocal = (signed char)le;
                    int ih = 0;
                    R.conn = face_connection(g, i, j, k, f, ih);
                    R.boundary = R.conn < 0;
                    R.is_here = R.boundary ? 1 : (signed char)ih;
                    eintf[le][neintf[le]++] = nif;
                    nif++;
                }
            }
    int nc = 0;
    for (int a = 0; a < nif; ++a) {
        int found = -1;
        if (!I[a].boundary)
            for (int b = 0; b < a; ++b)
                if (!I[b].boundary && I[b].conn == I[a].conn) {
                    found = I[b].lidx;
                    break;
                }
        I[a].lidx = (signed char)(found >= 0 ? found : nc++);
    }
    // ---- local matrices (column-major, ld = nc)
    double A[144], B[96], C[144], D[96];
    for (int q = 0; q < nc * nc; ++q) {
        A[q] = 0;
        C[q] = 0;
    }
    for (int q = 0; q < nc * ne; ++q) {
        B[q] = 0;
        D[q] = 0;
    }
    const v3 corner = point_at(g, pi, pj, pk);
    double fst[8][3][4];   // each element's contributions f(q, jj), jj = 0 for B/D
    for (int le = 0; le < ne; ++le) {
        const int nI = neintf[le], nb = nI + 1;
        double Am[16], Psi[16];
        const v3 elp = elem_center(g, ei_[le], ej_[le], ek_[le]);
        for (int q = 0; q < 16; ++q) {
            Am[q] = 0;
            Psi[q] = 0;
        }
        Am[0] = 1;
        Am[0 + 1 * nb] = elp.x;
        Am[0 + 2 * nb] = elp.y;
        if (g.dim == 3)
            Am[0 + 3 * nb] = elp.z;
        for (int q = 0; q < nI; ++q) {
            const RIntf &R = I[eintf[le][q]];
            v3 cc, nn;
            double ar;
            face_geom(g, pi - 1 + R.ei, pj - 1 + R.ej, pk - 1 + R.ek, R.f, cc, nn, ar);
            v3 cp = cc;
            if (continuity != 1)
                cp = mk((cc.x - corner.x) * continuity + corner.x,
                        (cc.y - corner.y) * continuity + corner.y,
                        (cc.z - corner.z) * continuity + corner.z);
            Am[q + 1] = 1;
            Am[q + 1 + 1 * nb] = cp.x;
            Am[q + 1 + 2 * nb] = cp.y;
            if (g.dim == 3)
                Am[q + 1 + 3 * nb] = cp.z;
        }
        for (int q = 0; q < nb; ++q)
            Psi[q + q * nb] = 1;
        if (dgesv(nb, nb, Am, nb, Psi, nb)) {
            atomicMax(err, 1);
            return;
        }
        v3 grad[4];
        for (int q = 0; q < nb; ++q)
            grad[q] = mk(Psi[1 + q * nb], Psi[2 + q * nb], g.dim == 3 ? Psi[3 + q * nb] : 0);
        v3 Kn[3];
        for (int q = 0; q < nI; ++q) {
            const RIntf &R = I[eintf[le][q]];
            Kn[q] = tensor_mult(K6 + 6 * (size_t)eidx[le],
                                normal_area(g, pi - 1 + R.ei, pj - 1 + R.ej, pk - 1 + R.ek,
                                            R.f, pi, pj, pk));
        }
        for (int q = 0; q < nI; ++q) {
            const RIntf &R = I[eintf[le][q]];
            double f = vdot(Kn[q], grad[0]);
            const int is = R.lidx, ei = R.elocal;
            B[is + ei * nc] += -f;
            if (R.is_here)
                D[is + ei * nc] += -f;
            fst[le][q][0] = f;
            for (int jj = 1; jj < nI + 1; ++jj) {
                f = vdot(Kn[q], grad[jj]);
                const int js = I[eintf[le][jj - 1]].lidx;
                A[is + js * nc] += f;
                if (R.is_here)
                    C[is + js * nc] += f;
                fst[le][q][jj] = f;
            }
        }
    }
    // Tbar: this corner's rows of the connection lists' gradients
    double Tb[96];
    for (int q = 0; q < nc * ne; ++q)
        Tb[q] = 0;
    for (int a = 0; a < nif; ++a) {
        const RIntf &R = I[a];
        if (R.boundary || !R.is_here)
            continue;
        const int c = R.conn, i = R.lidx;
        for (int le = 0; le < ne; ++le)
            for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q)
                if (md_elem[q] == eidx[le]) {
                    Tb[i + le * nc] = tbar[q];
                    break;
                }
    }
    // LU of A with pivots and AiB = A^-1 B (dgetf2 + dgetrs, as dgesv)
    int piv[12];
    if (dgesv_piv(nc, ne, A, nc, B, nc, piv)) {
        atomicMax(err, 2);
        return;
    }
    // W = C' Tbar, then Bbar = A^-T W (in W)
    double W[96];
    for (int j = 0; j < ne; ++j)
        for (int l = 0; l < nc; ++l) {
            double sacc = 0;
            for (int i = 0; i < nc; ++i)
                sacc += C[i + l * nc] * Tb[i + j * nc];
            W[l + j * nc] = sacc;
        }
    for (int j = 0; j < ne; ++j) {
        double *b = W + j * nc;
        for (int i = 0; i < nc; ++i) {
            double t = b[i];
            for (int k = 0; k < i; ++k)
                t -= A[k + i * nc] * b[k];
            b[i] = t / A[i + i * nc];
        }
        for (int i = nc - 1; i >= 0; --i) {
            double t = b[i];
            for (int k = i + 1; k < nc; ++k)
                t -= A[k + i * nc] * b[k];
            b[i] = t;
        }
        for (int i = nc - 1; i >= 0; --i)
            if (piv[i] != i) {
                const double t = b[i];
                b[i] = b[piv[i]];
                b[piv[i]] = t;
            }
    }
    // element terms in assembly order; Abar(i,l) = -sum_j Bbar(i,j) AiB(l,j),
    // Cbar(i,l) = sum_j Tbar(i,j) AiB(l,j)
    for (int le = 0; le < ne; ++le) {
        double gsum = 0;
        const int nI = neintf[le];
        for (int q = 0; q < nI; ++q) {
            const RIntf &R = I[eintf[le][q]];
            const int is = R.lidx;
            const double bq = -fst[le][q][0];
            gsum += W[is + le * nc] * bq;
            if (R.is_here)
                gsum += (-Tb[is + le * nc]) * bq;
            for (int jj = 1; jj < nI + 1; ++jj) {
                const int js = I[eintf[le][jj - 1]].lidx;
                double sa = 0, sc = 0;
                for (int j = 0; j < ne; ++j) {
                    sa += W[is + j * nc] * B[js + j * nc];
                    sc += Tb[is + j * nc] * B[js + j * nc];
                }
                gsum += (-sa) * fst[le][q][jj];
                if (R.is_here)
                    gsum += sc * fst[le][q][jj];
            }
        }
        st_m[(size_t)p * 8 + le] = gsum;
    }
}

// mbar of a cell: its corners' terms in ascending point order
__global__ void k_vjp_gather(Grid g, const double *__restrict__ st_m, double *__restrict__ mbar)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.ncells)
        return;
    const int i = e % g.nx, j = (e / g.nx) % g.ny, k = g.dim == 3 ? e / (g.nx * g.ny) : 0;
    const int npx = g.nx + 1, npy = g.ny + 1;
    double acc = 0;
    const int kt = g.dim == 3 ? 1 : 0;
    for (int pk = k; pk <= k + kt; ++pk)
        for (int pj = j; pj <= j + 1; ++pj)
            for (int pi = i; pi <= i + 1; ++pi) {
                // local index of e among the corner's elements (i-major)
                int le = 0, me = -1;
                const int kmin = g.dim == 3 ? pk - 1 : 0, kmax = g.dim == 3 ? pk : 0;
                for (int a = pi - 1; a <= pi && me < 0; ++a)
                    for (int b = pj - 1; b <= pj && me < 0; ++b)
                        for (int c = kmin; c <= kmax; ++c) {
                            if (!elem_valid(g, a, b, c))
                                continue;
                            if (a == i && b == j && c == k) {
                                me = le;
                                break;
                            }
                            ++le;
                        }
                const int p = pi + pj * npx + (g.dim == 3 ? pk * npx * npy : 0);
                acc += st_m[(size_t)p * 8 + me];
            }
    mbar[e] = acc;
}

// storeTransmissibilities replayed per connection over its corners in point
// order, then Connection.check
__global__ void k_merge(int nconn, int nslots, int cap, const double *__restrict__ st_t,
                        const int *__restrict__ st_e, const signed char *__restrict__ st_n,
                        const int *__restrict__ here, const int *__restrict__ there,
                        int *__restrict__ md_cnt, int *__restrict__ md_elem,
                        double *__restrict__ md_t, int *__restrict__ fails)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nconn)
        return;
    int cnt = 0;
    int *el = md_elem + (size_t)c * cap;
    double *tv = md_t + (size_t)c * cap;
    for (int s = 0; s < nslots; ++s) {
        const int ne = st_n[(size_t)c * MPFA_SLOTS + s];
        const double *tt = st_t + ((size_t)c * MPFA_SLOTS + s) * MPFA_NE;
        const int *te = st_e + ((size_t)c * MPFA_SLOTS + s) * MPFA_NE;
        for (int le = 0; le < ne; ++le) {
            const double tk = tt[le];
            const int e = te[le];
            bool found = false;
            for (int q = 0; q < cnt; ++q)
                if (el[q] == e) {
                    found = true;
                    tv[q] += tk;
                    break;
                }
            if (!found && tk != 0 && cnt < cap) {
                el[cnt] = e;
                tv[cnt] = tk;
                cnt++;
            }
        }
    }
    // Connection.setDarcyTransmissibilities: multiplier 1, check
    double sum = 0;
    bool hh = false, th = false;
    for (int q = 0; q < cnt; ++q) {
        tv[q] = tv[q] * 1.0;
        sum += tv[q];
        hh |= el[q] == here[c];
        th |= el[q] == there[c];
    }
    md_cnt[c] = cnt;
    if (fabs(sum) > 1.0e-6 || !hh || !th)
        atomicAdd(fails, 1);
}

__global__ void k_connections(Grid g, int *__restrict__ here, int *__restrict__ there)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.ncells)
        return;
    const int i = e % g.nx, j = (e / g.nx) % g.ny, k = e / (g.nx * g.ny);
    for (int f = 1; f < g.nf; f += 2) {
        int ih;
        const int c = face_connection(g, i, j, k, f, ih);
        if (c < 0)
            continue;
        int a, b, cc;
        face_offset(g, f, a, b, cc);
        here[c] = e;
        there[c] = elem_index(g, i + a, j + b, k + cc);
    }
}

// incident connections of cell e in ascending connection order
__device__ int incident(const Grid &g, int e, int *cs)
{
    const int i = e % g.nx, j = (e / g.nx) % g.ny, k = e / (g.nx * g.ny);
    int m = 0;
    for (int f = 0; f < g.nf; ++f) {
        int ih;
        const int c = face_connection(g, i, j, k, f, ih);
        if (c >= 0)
            cs[m++] = c;
    }
    for (int a = 1; a < m; ++a) {   // insertion sort (<= 6)
        const int v = cs[a];
        int b = a - 1;
        while (b >= 0 && cs[b] > v) {
            cs[b + 1] = cs[b];
            --b;
        }
        cs[b + 1] = v;
    }
    return m;
}

#define ROWCAP 64

__global__ void k_pattern(Grid g, int with_diag, const int *__restrict__ md_ptr,
                          const int *__restrict__ md_elem, int *__restrict__ rowlen,
                          const int *__restrict__ rowptr, int *__restrict__ colind,
                          int *__restrict__ err)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= g.ncells)
        return;
    int row[ROWCAP];
    int len = 0;
    auto ins = [&](int v) {
        for (int q = 0; q < len; ++q)
            if (row[q] == v)
                return;
        if (len >= ROWCAP) {
            atomicMax(err, 3);
            return;
        }
        int q = len;
        while (q > 0 && row[q - 1] > v) {
            row[q] = row[q - 1];
            --q;
        }
        row[q] = v;
        len++;
    };
    if (with_diag)
        ins(r);
    int cs[6];
    const int m = incident(g, r, cs);
    for (int a = 0; a < m; ++a)
        for (int q = md_ptr[cs[a]]; q < md_ptr[cs[a] + 1]; ++q)
            ins(md_elem[q]);
    if (!colind) {
        rowlen[r] = len;
        return;
    }
    for (int q = 0; q < len; ++q)
        colind[rowptr[r] + q] = row[q];
}

__global__ void k_assemble(Grid g, const int *__restrict__ md_ptr,
                           const int *__restrict__ md_elem, const double *__restrict__ md_t,
                           const int *__restrict__ here, const int *__restrict__ rowptr,
                           const int *__restrict__ colind, double *__restrict__ val)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= g.ncells)
        return;
    const int r0 = rowptr[r], r1 = rowptr[r + 1];
    for (int q = r0; q < r1; ++q)
        val[q] = 0.0;
    int cs[6];
    const int m = incident(g, r, cs);
    for (int a = 0; a < m; ++a) {
        const int c = cs[a];
        const bool h = here[c] == r;
        for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q) {
            const int k = md_elem[q];
            int lo = r0, hi = r1 - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (colind[mid] < k)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            val[lo] += h ? md_t[q] : -md_t[q];
        }
    }
}

__global__ void k_compact(int nconn, int cap, const int *__restrict__ md_ptr,
                          const int *__restrict__ src_e, const double *__restrict__ src_t,
                          int *__restrict__ dst_e, double *__restrict__ dst_t)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nconn)
        return;
    const int n = md_ptr[c + 1] - md_ptr[c];
    for (int q = 0; q < n; ++q) {
        dst_e[md_ptr[c] + q] = src_e[(size_t)c * cap + q];
        dst_t[md_ptr[c] + q] = src_t[(size_t)c * cap + q];
    }
}

struct MpfaResult {
    int nconn = 0, ncells = 0;
    std::vector<int> md_ptr;
    DevBuf<int> d_md_ptr, d_md_elem, d_here, d_there;
    DevBuf<double> d_md_t;
    DevBuf<double> xs, ys, zs;
    Grid g{};
    int fails = 0;
};

void mpfa_run(rsgpu_ctx *ctx, int dim, int nx, int ny, int nz, double dx, double dy,
              double dz, const double *K6_host, double continuity, MpfaResult &R)
{
    if (dim != 2 && dim != 3)
        throw RsError{RSGPU_BAD_ARGUMENT, "dim must be 2 or 3"};
    if (nx < 1 || ny < 1 || (dim == 3 && nz < 1))
        throw RsError{RSGPU_BAD_ARGUMENT, "grid sizes must be positive"};
    if (!(continuity > 0 && continuity <= 1))
        throw RsError{RSGPU_BAD_ARGUMENT, "continuity must be in (0,1]"};
    hipStream_t s = ctx->stream;
    Grid &g = R.g;
    g.dim = dim;
    g.nx = nx;
    g.ny = ny;
    g.nz = dim == 3 ? nz : 1;
    g.nf = dim == 3 ? 6 : 4;
    g.ncells = nx * ny * g.nz;
    g.nconn = dim == 3 ? 3 * nx * ny * nz - nx * ny - nx * nz - ny * nz : 2 * nx * ny - nx - ny;
    g.npoints = dim == 3 ? (nx + 1) * (ny + 1) * (nz + 1) : (nx + 1) * (ny + 1);
    // ArrayData.getExpandedPoints: cumulative sums from 0
    std::vector<double> xs(nx + 1), ys(ny + 1), zs(g.nz + 1);
    double o = 0;
    xs[0] = 0;
    for (int i = 0; i < nx; ++i) xs[i + 1] = (o += dx);
    o = 0;
    ys[0] = 0;
    for (int i = 0; i < ny; ++i) ys[i + 1] = (o += dy);
    o = 0;
    zs[0] = 0;
    if (dim == 3)
        for (int i = 0; i < nz; ++i) zs[i + 1] = (o += dz);
    R.xs.upload(xs, s);
    R.ys.upload(ys, s);
    R.zs.upload(zs, s);
    g.xs = R.xs.p;
    g.ys = R.ys.p;
    g.zs = R.zs.p;
    R.nconn = g.nconn;
    R.ncells = g.ncells;
    const int cap = dim == 3 ? 18 : 6;
    const int nslots = dim == 3 ? 4 : 2;
    DevBuf<double> K, st_t, md_t_cap;
    DevBuf<int> st_e, md_cnt, md_e_cap, err;
    DevBuf<signed char> st_n;
    K.upload(K6_host, (size_t)g.ncells * 6, s);
    const size_t nc = std::max(g.nconn, 1);
    st_t.alloc(nc * MPFA_SLOTS * MPFA_NE);
    st_e.alloc(nc * MPFA_SLOTS * MPFA_NE);
    st_n.alloc(nc * MPFA_SLOTS);
    st_n.zero(s);
    md_cnt.alloc(nc);
    md_e_cap.alloc(nc * cap);
    md_t_cap.alloc(nc * cap);
    err.alloc(2);
    err.zero(s);
    R.d_here.alloc(nc);
    R.d_there.alloc(nc);
    hipLaunchKernelGGL(k_connections, dim3((g.ncells + 255) / 256), dim3(256), 0, s, g,
                       R.d_here.p, R.d_there.p);
    hipLaunchKernelGGL(k_corner, dim3((g.npoints + 63) / 64), dim3(64), 0, s, g, K.p,
                       continuity, st_t.p, st_e.p, st_n.p, err.p);
    RS_CHECK_LAUNCH();
    if (g.nconn > 0)
        hipLaunchKernelGGL(k_merge, dim3((g.nconn + 255) / 256), dim3(256), 0, s,
                           g.nconn, nslots, cap, st_t.p, st_e.p, st_n.p, R.d_here.p,
                           R.d_there.p, md_cnt.p, md_e_cap.p, md_t_cap.p, err.p + 1);
    RS_CHECK_LAUNCH();
    std::vector<int> cnt(g.nconn);
    int herr[2];
    RS_HIP(hipMemcpyAsync(herr, err.p, sizeof(herr), hipMemcpyDeviceToHost, s));
    if (g.nconn)
        RS_HIP(hipMemcpyAsync(cnt.data(), md_cnt.p, sizeof(int) * g.nconn,
                              hipMemcpyDeviceToHost, s));
    RS_HIP(hipStreamSynchronize(s));
    if (herr[0])
        throw RsError{RSGPU_SINGULAR, herr[0] == 1 ? "Severe grid distortion"
                                                   : "Conductivity tensor does not appear "
                                                     "to be symmetrical, positive definite"};
    R.fails = herr[1];
    R.md_ptr.assign(g.nconn + 1, 0);
    for (int c = 0; c < g.nconn; ++c)
        R.md_ptr[c + 1] = R.md_ptr[c] + cnt[c];
    const int tot = R.md_ptr[g.nconn];
    R.d_md_ptr.upload(R.md_ptr, s);
    R.d_md_elem.alloc(std::max(tot, 1));
    R.d_md_t.alloc(std::max(tot, 1));
    if (g.nconn)
        hipLaunchKernelGGL(k_compact, dim3((g.nconn + 255) / 256), dim3(256), 0, s,
                           g.nconn, cap, R.d_md_ptr.p, md_e_cap.p, md_t_cap.p,
                           R.d_md_elem.p, R.d_md_t.p);
    RS_CHECK_LAUNCH();
    RS_HIP(hipStreamSynchronize(s));
}

} // namespace

static int fail(rsgpu_ctx *ctx, const RsError &e)
{
    if (ctx)
        ctx->err = e.msg;
    return e.code;
}

extern "C" int rsgpu_mpfa_o(rsgpu_ctx *ctx, int dim, int nx, int ny, int nz, double dx,
                            double dy, double dz, const double *K6, double continuity,
                            int32_t *md_ptr, int32_t *md_elem, double *md_t,
                            int *check_fail)
{
    if (!ctx || !K6 || !md_ptr || !md_elem || !md_t)
        return RSGPU_BAD_ARGUMENT;
    try {
        RS_HIP(hipSetDevice(ctx->device));
        MpfaResult R;
        mpfa_run(ctx, dim, nx, ny, nz, dx, dy, dz, K6, continuity, R);
        std::copy(R.md_ptr.begin(), R.md_ptr.end(), md_ptr);
        const int tot = R.md_ptr[R.nconn];
        if (tot) {
            RS_HIP(hipMemcpy(md_elem, R.d_md_elem.p, sizeof(int) * tot,
                             hipMemcpyDeviceToHost));
            RS_HIP(hipMemcpy(md_t, R.d_md_t.p, sizeof(double) * tot,
                             hipMemcpyDeviceToHost));
        }
        if (check_fail)
            *check_fail = R.fails;
        return RSGPU_OK;
    } catch (const RsError &e) {
        return fail(ctx, e);
    }
}

rsgpu_flux *flux_from_device(rsgpu_ctx *ctx, int nconn, int ncells, const std::vector<int> &mp,
                             DevBuf<int> &md_elem, DevBuf<double> &md_t,
                             const std::vector<int> &here, const std::vector<int> &there);

extern "C" int rsgpu_mpfa_o_flux(rsgpu_ctx *ctx, int dim, int nx, int ny, int nz, double dx,
                                 double dy, double dz, const double *K6, double continuity,
                                 rsgpu_flux **F_out, int *check_fail)
{
    if (!ctx || !K6 || !F_out)
        return RSGPU_BAD_ARGUMENT;
    try {
        RS_HIP(hipSetDevice(ctx->device));
        MpfaResult R;
        mpfa_run(ctx, dim, nx, ny, nz, dx, dy, dz, K6, continuity, R);
        std::vector<int> here(R.nconn), there(R.nconn);
        if (R.nconn) {
            RS_HIP(hipMemcpy(here.data(), R.d_here.p, sizeof(int) * R.nconn,
                             hipMemcpyDeviceToHost));
            RS_HIP(hipMemcpy(there.data(), R.d_there.p, sizeof(int) * R.nconn,
                             hipMemcpyDeviceToHost));
        }
        *F_out = flux_from_device(ctx, R.nconn, R.ncells, R.md_ptr, R.d_md_elem, R.d_md_t, here,
                                  there);
        if (check_fail)
            *check_fail = R.fails;
        return RSGPU_OK;
    } catch (const RsError &e) {
        return fail(ctx, e);
    }
}

// adjoint of the transmissibility map (no reference counterpart)
int flux_md_lists(const rsgpu_flux *F, const int **md_ptr, const int **md_elem, int *nconn,
                  int *ncells);

extern "C" int rsgpu_mpfa_o_vjp(rsgpu_ctx *ctx, int dim, int nx, int ny, int nz, double dx,
                                double dy, double dz, const double *K6, double continuity,
                                const rsgpu_flux *F, const double *tbar_dev, double *mbar_dev)
{
    if (!ctx || !K6 || !F || !tbar_dev || !mbar_dev)
        return RSGPU_BAD_ARGUMENT;
    try {
        RS_HIP(hipSetDevice(ctx->device));
        if (dim != 2 && dim != 3)
            throw RsError{RSGPU_BAD_ARGUMENT, "dim must be 2 or 3"};
        if (nx < 1 || ny < 1 || (dim == 3 && nz < 1))
            throw RsError{RSGPU_BAD_ARGUMENT, "grid sizes must be positive"};
        if (!(continuity > 0 && continuity <= 1))
            throw RsError{RSGPU_BAD_ARGUMENT, "continuity must be in (0,1]"};
        hipStream_t s = ctx->stream;
        Grid g{};
        g.dim = dim;
        g.nx = nx;
        g.ny = ny;
        g.nz = dim == 3 ? nz : 1;
        g.nf = dim == 3 ? 6 : 4;
        g.ncells = nx * ny * g.nz;
        g.nconn = dim == 3 ? 3 * nx * ny * nz - nx * ny - nx * nz - ny * nz
                           : 2 * nx * ny - nx - ny;
        g.npoints = dim == 3 ? (nx + 1) * (ny + 1) * (nz + 1) : (nx + 1) * (ny + 1);
        const int *mdp, *mde;
        int fnc, fcells;
        flux_md_lists(F, &mdp, &mde, &fnc, &fcells);
        if (fnc != g.nconn || fcells != g.ncells)
            throw RsError{RSGPU_BAD_ARGUMENT, "flux operator does not match the grid"};
        std::vector<double> xs(nx + 1), ys(ny + 1), zs(g.nz + 1);
        double o = 0;
        xs[0] = 0;
        for (int i = 0; i < nx; ++i) xs[i + 1] = (o += dx);
        o = 0;
        ys[0] = 0;
        for (int i = 0; i < ny; ++i) ys[i + 1] = (o += dy);
        o = 0;
        zs[0] = 0;
        if (dim == 3)
            for (int i = 0; i < nz; ++i) zs[i + 1] = (o += dz);
        DevBuf<double> dxs, dys, dzs, K, st_m;
        DevBuf<int> err;
        dxs.upload(xs, s);
        dys.upload(ys, s);
        dzs.upload(zs, s);
        g.xs = dxs.p;
        g.ys = dys.p;
        g.zs = dzs.p;
        K.upload(K6, (size_t)g.ncells * 6, s);
        st_m.alloc((size_t)g.npoints * 8);
        err.alloc(1);
        err.zero(s);
        hipLaunchKernelGGL(k_corner_vjp, dim3((g.npoints + 63) / 64), dim3(64), 0, s, g, K.p,
                           continuity, mdp, mde, tbar_dev, st_m.p, err.p);
        hipLaunchKernelGGL(k_vjp_gather, dim3((g.ncells + 255) / 256), dim3(256), 0, s, g,
                           st_m.p, mbar_dev);
        RS_CHECK_LAUNCH();
        int herr = 0;
        RS_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
        RS_HIP(hipStreamSynchronize(s));
        if (herr)
            throw RsError{RSGPU_SINGULAR, "singular MPFA-O local system"};
        return RSGPU_OK;
    } catch (const RsError &e) {
        return fail(ctx, e);
    }
}

extern "C" int rsgpu_mpfa_o_assemble(rsgpu_ctx *ctx, int dim, int nx, int ny, int nz,
                                     double dx, double dy, double dz, const double *K6,
                                     double continuity, rsgpu_csr **A_out,
                                     int *check_fail)
{
    if (!ctx || !K6 || !A_out)
        return RSGPU_BAD_ARGUMENT;
    try {
        RS_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        MpfaResult R;
        mpfa_run(ctx, dim, nx, ny, nz, dx, dy, dz, K6, continuity, R);
        const int n = R.ncells;
        DevBuf<int> rowlen, err;
        rowlen.alloc(n);
        err.alloc(1);
        err.zero(s);
        const dim3 gr((n + 255) / 256), bl(256);
        hipLaunchKernelGGL(k_pattern, gr, bl, 0, s, R.g, 1, R.d_md_ptr.p, R.d_md_elem.p,
                           rowlen.p, nullptr, nullptr, err.p);
        RS_CHECK_LAUNCH();
        std::vector<int> len(n), rp(n + 1, 0);
        RS_HIP(hipMemcpyAsync(len.data(), rowlen.p, sizeof(int) * n, hipMemcpyDeviceToHost,
                              s));
        int herr = 0;
        RS_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
        RS_HIP(hipStreamSynchronize(s));
        if (herr)
            throw RsError{RSGPU_BAD_ARGUMENT, "row pattern exceeds 64 entries"};
        for (int r = 0; r < n; ++r)
            rp[r + 1] = rp[r] + len[r];
        const int nnz = rp[n];
        rsgpu_csr *A = new rsgpu_csr();
        A->ctx = ctx;
        A->n = A->ncols = n;
        A->nnz = nnz;
        A->h_rowptr = rp;
        A->rowptr.upload(rp, s);
        A->colind.alloc(std::max(nnz, 1));
        hipLaunchKernelGGL(k_pattern, gr, bl, 0, s, R.g, 1, R.d_md_ptr.p, R.d_md_elem.p,
                           rowlen.p, A->rowptr.p, A->colind.p, err.p);
        A->val.alloc(std::max(nnz, 1));
        hipLaunchKernelGGL(k_assemble, gr, bl, 0, s, R.g, R.d_md_ptr.p, R.d_md_elem.p,
                           R.d_md_t.p, R.d_here.p, A->rowptr.p, A->colind.p, A->val.p);
        RS_CHECK_LAUNCH();
        A->h_colind.resize(nnz);
        if (nnz)
            RS_HIP(hipMemcpyAsync(A->h_colind.data(), A->colind.p, sizeof(int) * nnz,
                                  hipMemcpyDeviceToHost, s));
        spmv_build_plans(A->h_rowptr, n, A, s);
        A->pattern_id = next_pattern_id();
        RS_HIP(hipStreamSynchronize(s));
        if (check_fail)
            *check_fail = R.fails;
        *A_out = A;
        return RSGPU_OK;
    } catch (const RsError &e) {
        return fail(ctx, e);
    }
}
