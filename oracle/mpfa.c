/*
 * Oracle (test infrastructure only): scalar restatement of the rs Cartesian
 * mesh (topology numbering, hexahedral/quadrilateral geometry), the MPFA
 * O-method transmissibility calculation and the CSR pattern/flux assembly
 * that consumes it.  Compile with -O2 -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- Dense LU: LAPACK 3.0 dgesv = dgetf2 + dgetrs (reference BLAS) ------ */
/* Column-major, 0-based.  Returns 0 or OR_SINGULAR (info > 0). */
int or_dgesv(int n, int nrhs, double *a, int lda, int *ipiv, double *b,
             int ldb)
{
    int info = 0;
    for (int j = 0; j < n; ++j) {
        /* idamax: first index of max |a(i,j)|, i >= j */
        int jp = j;
        double amax = fabs(a[j + j * lda]);
        for (int i = j + 1; i < n; ++i) {
            double v = fabs(a[i + j * lda]);
            if (v > amax) {
                amax = v;
                jp = i;
            }
        }
        ipiv[j] = jp;
        if (a[jp + j * lda] != 0) {
            if (jp != j)
                for (int c = 0; c < n; ++c) {
                    double t = a[j + c * lda];
                    a[j + c * lda] = a[jp + c * lda];
                    a[jp + c * lda] = t;
                }
            if (j < n - 1) {
                double r = 1.0 / a[j + j * lda];
                for (int i = j + 1; i < n; ++i)
                    a[i + j * lda] *= r;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        if (j < n - 1) {
            /* dger: A22 += x * (-y)' column by column */
            for (int c = j + 1; c < n; ++c) {
                double yj = a[j + c * lda];
                if (yj != 0) {
                    double temp = -1.0 * yj;
                    for (int i = j + 1; i < n; ++i)
                        a[i + c * lda] = a[i + c * lda] + a[i + j * lda] * temp;
                }
            }
        }
    }
    if (info)
        return OR_SINGULAR;
    /* dlaswp on B */
    for (int j = 0; j < n; ++j)
        if (ipiv[j] != j)
            for (int c = 0; c < nrhs; ++c) {
                double t = b[j + c * ldb];
                b[j + c * ldb] = b[ipiv[j] + c * ldb];
                b[ipiv[j] + c * ldb] = t;
            }
    /* dtrsm Left Lower NoTrans Unit */
    for (int c = 0; c < nrhs; ++c)
        for (int k = 0; k < n; ++k)
            if (b[k + c * ldb] != 0)
                for (int i = k + 1; i < n; ++i)
                    b[i + c * ldb] = b[i + c * ldb] - b[k + c * ldb] * a[i + k * lda];
    /* dtrsm Left Upper NoTrans NonUnit */
    for (int c = 0; c < nrhs; ++c)
        for (int k = n - 1; k >= 0; --k)
            if (b[k + c * ldb] != 0) {
                b[k + c * ldb] = b[k + c * ldb] / a[k + k * lda];
                for (int i = 0; i < k; ++i)
                    b[i + c * ldb] = b[i + c * ldb] - b[k + c * ldb] * a[i + k * lda];
            }
    return OR_OK;
}

/* ---- Vector helpers (rs geometry/Vector3D.java) ------------------------- */
typedef struct {
    double x, y, z;
} v3;

static v3 mk(double x, double y, double z)
{
    v3 r = {x, y, z};
    return r;
}
static v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 vmul(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static double vnorm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static v3 vcross(v3 a, v3 v)
{
    return mk(a.y * v.z - a.z * v.y, a.z * v.x - a.x * v.z,
              a.x * v.y - a.y * v.x);
}

/* ---- Cartesian mesh ------------------------------------------------------ */
/* Element (i,j,k) -> i + j*nx + k*nx*ny (NaturalIJKOrdering.java:99-104).
 * Interfaces: element*nf + local orientation index; 3D orientations
 * TOP(0,k-side) BOTTOM(1,k+1) FRONT(2,j) BACK(3,j+1) LEFT(4,i) RIGHT(5,i+1)
 * (Orientation.java), 2D FRONT(0) BACK(1) LEFT(2) RIGHT(3)
 * (CartesianTopology2D.java). */
typedef struct {
    int dim, nx, ny, nz, nf;
    int ncells, nconn, npoints;
    double *xs, *ys, *zs; /* point coordinate lines */
    double dx, dy, dz;
} mesh_t;

static int mesh_init(mesh_t *m, int dim, int nx, int ny, int nz, double dx,
                     double dy, double dz)
{
    memset(m, 0, sizeof(*m));
    if (dim != 2 && dim != 3)
        return OR_BAD_ARGUMENT;
    if (nx < 1 || ny < 1 || (dim == 3 && nz < 1))
        return OR_BAD_ARGUMENT;
    m->dim = dim;
    m->nx = nx;
    m->ny = ny;
    m->nz = dim == 3 ? nz : 1;
    m->nf = dim == 3 ? 6 : 4;
    m->ncells = nx * ny * m->nz;
    if (dim == 3) {
        m->nconn = 3 * nx * ny * nz - nx * ny - nx * nz - ny * nz;
        m->npoints = (nx + 1) * (ny + 1) * (nz + 1);
    } else {
        m->nconn = 2 * nx * ny - nx - ny;
        m->npoints = (nx + 1) * (ny + 1);
    }
    m->dx = dx;
    m->dy = dy;
    m->dz = dz;
    /* ArrayData.getExpandedPoints: cumulative sums from the origin 0 */
    m->xs = (double *)malloc(sizeof(double) * (nx + 1));
    m->ys = (double *)malloc(sizeof(double) * (ny + 1));
    m->zs = (double *)malloc(sizeof(double) * (m->nz + 1));
    double o = 0;
    m->xs[0] = 0;
    for (int i = 0; i < nx; ++i) {
        o += dx;
        m->xs[i + 1] = o;
    }
    o = 0;
    m->ys[0] = 0;
    for (int i = 0; i < ny; ++i) {
        o += dy;
        m->ys[i + 1] = o;
    }
    o = 0;
    m->zs[0] = 0;
    if (dim == 3)
        for (int i = 0; i < nz; ++i) {
            o += dz;
            m->zs[i + 1] = o;
        }
    return OR_OK;
}

static void mesh_free(mesh_t *m)
{
    free(m->xs);
    free(m->ys);
    free(m->zs);
}

static v3 point_at(const mesh_t *m, int i, int j, int k)
{
    return mk(m->xs[i], m->ys[j], m->dim == 3 ? m->zs[k] : 0.0);
}

static int elem_index(const mesh_t *m, int i, int j, int k)
{
    return i + j * m->nx + k * m->nx * m->ny;
}

static int elem_valid(const mesh_t *m, int i, int j, int k)
{
    return i >= 0 && i < m->nx && j >= 0 && j < m->ny && k >= 0 &&
           k < m->nz;
}

/* Connection numbering (CartesianTopology3D.java:104-132,
 * CartesianTopology2D.java:103-121), for the forward orientations. */
static int conn_index(const mesh_t *m, int i, int j, int k, int orient)
{
    int ni = m->nx, nj = m->ny, nk = m->nz;
    if (m->dim == 3) {
        switch (orient) {
        case 1: /* BOTTOM */
            return i + j * ni + k * ni * nj;
        case 3: /* BACK */
            return i + j * ni + k * ni * (nj - 1) + ni * nj * (nk - 1);
        case 5: /* RIGHT */
            return i + j * (ni - 1) + k * (ni - 1) * nj + ni * nj * (nk - 1) +
                   ni * (nj - 1) * nk;
        }
    } else {
        switch (orient) {
        case 1: /* BACK */
            return i + j * ni;
        case 3: /* RIGHT */
            return i + j * (ni - 1) + ni * (nj - 1);
        }
    }
    return -1;
}

/* Neighbour element across a local face; -1 if boundary. */
static int face_neighbour(const mesh_t *m, int i, int j, int k, int f,
                          int *ni_, int *nj_, int *nk_)
{
    int a = i, b = j, c = k;
    if (m->dim == 3) {
        switch (f) {
        case 0: c = k - 1; break;
        case 1: c = k + 1; break;
        case 2: b = j - 1; break;
        case 3: b = j + 1; break;
        case 4: a = i - 1; break;
        case 5: a = i + 1; break;
        }
    } else {
        switch (f) {
        case 0: b = j - 1; break;
        case 1: b = j + 1; break;
        case 2: a = i - 1; break;
        case 3: a = i + 1; break;
        }
    }
    if (!elem_valid(m, a, b, c))
        return -1;
    if (ni_) {
        *ni_ = a;
        *nj_ = b;
        *nk_ = c;
    }
    return elem_index(m, a, b, c);
}

/* Connection of a face and whether it is the 'here' side.  Here = the
 * forward face of the lower-index element (CartesianTopology.java:68-79). */
static int face_connection(const mesh_t *m, int i, int j, int k, int f,
                           int *is_here)
{
    int a, b, c;
    if (face_neighbour(m, i, j, k, f, &a, &b, &c) < 0)
        return -1;
    if (f & 1) { /* forward orientation: BOTTOM/BACK/RIGHT (3D), BACK/RIGHT */
        *is_here = 1;
        return conn_index(m, i, j, k, f);
    }
    *is_here = 0;
    return conn_index(m, a, b, c, f + 1);
}

int or_mesh_counts(int dim, int nx, int ny, int nz, int *ncells, int *nconn,
                   int *npoints)
{
    mesh_t m;
    int st = mesh_init(&m, dim, nx, ny, nz, 1, 1, -1);
    if (st)
        return st;
    *ncells = m.ncells;
    *nconn = m.nconn;
    *npoints = m.npoints;
    mesh_free(&m);
    return OR_OK;
}

int or_mesh_connections(int dim, int nx, int ny, int nz, int *here,
                        int *there)
{
    mesh_t m;
    int st = mesh_init(&m, dim, nx, ny, nz, 1, 1, -1);
    if (st)
        return st;
    for (int k = 0; k < m.nz; ++k)
        for (int j = 0; j < m.ny; ++j)
            for (int i = 0; i < m.nx; ++i) {
                int e = elem_index(&m, i, j, k);
                for (int f = 1; f < m.nf; f += 2) {
                    int is_here;
                    int c = face_connection(&m, i, j, k, f, &is_here);
                    if (c < 0)
                        continue;
                    here[c] = e;
                    there[c] = face_neighbour(&m, i, j, k, f, 0, 0, 0);
                }
            }
    mesh_free(&m);
    return OR_OK;
}

/* ---- Geometry ------------------------------------------------------------ */
/* Element corner points cp[0..7] (CartesianTopology3D.getElementPoints) */
static void elem_points3(const mesh_t *m, int i, int j, int k, v3 cp[8])
{
    cp[0] = point_at(m, i, j, k);
    cp[1] = point_at(m, i + 1, j, k);
    cp[2] = point_at(m, i + 1, j + 1, k);
    cp[3] = point_at(m, i, j + 1, k);
    cp[4] = point_at(m, i, j, k + 1);
    cp[5] = point_at(m, i + 1, j, k + 1);
    cp[6] = point_at(m, i + 1, j + 1, k + 1);
    cp[7] = point_at(m, i, j + 1, k + 1);
}

/* BilinearMap.getPoint(.5,.5) */
static v3 bilinear_center(v3 p0, v3 p1, v3 p2, v3 p3)
{
    double a = .5, b = .5;
    double k0 = (1 - a) * (1 - b);
    double k1 = a * (1 - b);
    double k2 = (1 - a) * b;
    double k3 = a * b;
    return mk(k0 * p0.x + k1 * p1.x + k2 * p2.x + k3 * p3.x,
              k0 * p0.y + k1 * p1.y + k2 * p2.y + k3 * p3.y,
              k0 * p0.z + k1 * p1.z + k2 * p2.z + k3 * p3.z);
}

/* BilinearMap.computeAreaNormal */
static void bilinear_area_normal(v3 x1, v3 x2, v3 x3, v3 x4, double *area,
                                 v3 *normal)
{
    v3 a = vcross(vsub(x2, x1), vsub(x3, x1));
    v3 b = vcross(vsub(x2, x1), vsub(x4, x2));
    v3 c = vcross(vsub(x4, x3), vsub(x3, x1));
    v3 d = vcross(vsub(x4, x3), vsub(x4, x2));
    v3 n = vmul(vadd(a, vadd(b, vadd(c, d))), .25);
    double length = vnorm(n);
    *area = length;
    if (length < 1.0e-9)
        *normal = mk(0, 0, 0);
    else
        *normal = vmul(n, 1.0 / length);
}

/* TrilinearMap.getPoint(.5,.5,.5) with HexCell's reordering */
static v3 hex_center(const v3 cp[8])
{
    v3 x1 = cp[0], x2 = cp[1], x3 = cp[3], x4 = cp[2];
    v3 x5 = cp[4], x6 = cp[5], x7 = cp[7], x8 = cp[6];
    double a = .5, b = .5, c = .5;
#define TRI(C)                                                               \
    (c * (b * (a * x4.C + (1 - a) * x3.C) + (1 - b) * (a * x2.C + (1 - a) * x1.C)) + \
     (1 - c) * (b * (a * x8.C + (1 - a) * x7.C) + (1 - b) * (a * x6.C + (1 - a) * x5.C)))
    return mk(TRI(x), TRI(y), TRI(z));
#undef TRI
}

static v3 elem_center(const mesh_t *m, int i, int j, int k)
{
    if (m->dim == 3) {
        v3 cp[8];
        elem_points3(m, i, j, k, cp);
        return hex_center(cp);
    }
    /* QuadCell: BilinearMap(p0, p1, p3, p2) */
    v3 p0 = point_at(m, i, j, 0), p1 = point_at(m, i + 1, j, 0);
    v3 p2 = point_at(m, i + 1, j + 1, 0), p3 = point_at(m, i, j + 1, 0);
    return bilinear_center(p0, p1, p3, p2);
}

/* Interface geometry: center, outward unit normal, area
 * (HexCell.QuadFace / QuadCell.BiFace). */
static void face_geom(const mesh_t *m, int i, int j, int k, int f, v3 *center,
                      v3 *normal, double *area)
{
    if (m->dim == 3) {
        v3 cp[8];
        elem_points3(m, i, j, k, cp);
        v3 p0 = cp[0], p1 = cp[1], p2 = cp[3], p3 = cp[2];
        v3 p4 = cp[4], p5 = cp[5], p6 = cp[7], p7 = cp[6];
        v3 q[4];
        double scale = 1.0;
        switch (f) {
        case 0: q[0] = p0; q[1] = p1; q[2] = p2; q[3] = p3; break;
        case 1: q[0] = p4; q[1] = p5; q[2] = p6; q[3] = p7; scale = -1.0; break;
        case 2: q[0] = p0; q[1] = p1; q[2] = p4; q[3] = p5; scale = -1.0; break;
        case 3: q[0] = p2; q[1] = p3; q[2] = p6; q[3] = p7; break;
        case 4: q[0] = p0; q[1] = p2; q[2] = p4; q[3] = p6; break;
        default: q[0] = p1; q[1] = p3; q[2] = p5; q[3] = p7; scale = -1.0; break;
        }
        v3 nrm;
        bilinear_area_normal(q[0], q[1], q[2], q[3], area, &nrm);
        *normal = vmul(nrm, scale);
        *center = bilinear_center(q[0], q[1], q[2], q[3]);
        return;
    }
    v3 p0 = point_at(m, i, j, 0), p1 = point_at(m, i + 1, j, 0);
    v3 p2 = point_at(m, i + 1, j + 1, 0), p3 = point_at(m, i, j + 1, 0);
    v3 a, b;
    switch (f) {
    case 0: a = p0; b = p1; break; /* front */
    case 1: a = p2; b = p3; break; /* back */
    case 2: a = p3; b = p0; break; /* left */
    default: a = p1; b = p2; break; /* right */
    }
    v3 n = mk(b.y - a.y, -(b.x - a.x), 0);
    n = vmul(n, 1 / vnorm(n));
    *normal = n;
    *area = vnorm(vsub(b, a));
    *center = vmul(vadd(a, b), 0.5);
}

/* Interface corner-point order (getInterfacePoints), as (i,j,k) triples. */
static int face_points(const mesh_t *m, int i, int j, int k, int f,
                       int pts[4][3])
{
    if (m->dim == 3) {
        int cp[8][3] = {{i, j, k},         {i + 1, j, k},
                        {i + 1, j + 1, k}, {i, j + 1, k},
                        {i, j, k + 1},     {i + 1, j, k + 1},
                        {i + 1, j + 1, k + 1}, {i, j + 1, k + 1}};
        static const int ord[6][4] = {{0, 1, 2, 3}, {7, 6, 5, 4},
                                      {4, 5, 1, 0}, {6, 7, 3, 2},
                                      {7, 4, 0, 3}, {5, 6, 2, 1}};
        for (int q = 0; q < 4; ++q)
            memcpy(pts[q], cp[ord[f][q]], sizeof(int) * 3);
        return 4;
    }
    int cp[4][3] = {{i, j, 0}, {i + 1, j, 0}, {i + 1, j + 1, 0}, {i, j + 1, 0}};
    static const int ord2[4][2] = {{1, 0}, {3, 2}, {0, 3}, {2, 1}};
    for (int q = 0; q < 2; ++q)
        memcpy(pts[q], cp[ord2[f][q]], sizeof(int) * 3);
    return 2;
}

/* MPFA.calculateNormalArea (MPFA.java:314-354) */
static v3 normal_area(const mesh_t *m, int i, int j, int k, int f, int pi,
                      int pj, int pk)
{
    v3 c, n;
    double area;
    face_geom(m, i, j, k, f, &c, &n, &area);
    if (m->dim == 2)
        return vmul(n, -area / 2);
    int pts[4][3];
    face_points(m, i, j, k, f, pts);
    int s = 0;
    for (; s < 4; ++s)
        if (pts[s][0] == pi && pts[s][1] == pj && pts[s][2] == pk)
            break;
    int i2 = s == 0 ? 3 : (s - 1) % 4, i3 = (s + 1) % 4, i4 = (s + 2) % 4;
    v3 x1 = point_at(m, pts[s][0], pts[s][1], pts[s][2]);
    v3 x2 = point_at(m, pts[i2][0], pts[i2][1], pts[i2][2]);
    v3 x3 = point_at(m, pts[i3][0], pts[i3][1], pts[i3][2]);
    v3 x4 = point_at(m, pts[i4][0], pts[i4][1], pts[i4][2]);
    v3 v21 = vsub(x2, x1), v31 = vsub(x3, x1), v42 = vsub(x4, x2),
       v43 = vsub(x4, x3);
    v3 n1 = vmul(vcross(v21, v31), 9);
    v3 n2 = vmul(vcross(v21, v42), 3);
    v3 n3 = vmul(vcross(v43, v31), 3);
    v3 n4 = vcross(v43, v42);
    return vmul(vadd(vadd(vadd(n1, n2), n3), n4), 1. / 64.);
}

/* Tensor3D.mult */
static v3 tensor_mult(const double *K, v3 v)
{
    double xx = K[0], yy = K[1], zz = K[2], xy = K[3], xz = K[4], yz = K[5];
    return mk(xx * v.x + xy * v.y + xz * v.z, xy * v.x + yy * v.y + yz * v.z,
              xz * v.x + yz * v.y + zz * v.z);
}

/* ---- O-method interaction region ---------------------------------------- */
typedef struct {
    int e;            /* element index */
    int ei, ej, ek;
} rel_t;
typedef struct {
    int ei, ej, ek, f; /* owning element and local orientation */
    int elocal;        /* local index of owning element */
    int boundary;
    int conn, is_here;
    int lidx;          /* local connection index */
} rintf_t;
typedef struct {
    int ne, nif, nc;
    rel_t el[8];
    rintf_t intf[24];
    int eintf[8][3];   /* per element: region-interface ids in common order */
    int neintf[8];
} region_t;

/* Point (pi,pj,pk): elements(p) and interfaces(p) in the order the
 * LinkedHashSets of Topology.java receive them: element iteration order of
 * NaturalIJKOrdering (i outer, j, k inner), orientations in Orientation
 * order (CartesianTopology.java:44-80, Topology.java:134-164). */
static void build_region(const mesh_t *m, int pi, int pj, int pk,
                         region_t *R)
{
    memset(R, 0, sizeof(*R));
    int kmin = m->dim == 3 ? pk - 1 : 0, kmax = m->dim == 3 ? pk : 0;
    for (int i = pi - 1; i <= pi; ++i)
        for (int j = pj - 1; j <= pj; ++j)
            for (int k = kmin; k <= kmax; ++k) {
                if (!elem_valid(m, i, j, k))
                    continue;
                int le = R->ne++;
                R->el[le].e = elem_index(m, i, j, k);
                R->el[le].ei = i;
                R->el[le].ej = j;
                R->el[le].ek = k;
                for (int f = 0; f < m->nf; ++f) {
                    /* does face f of (i,j,k) contain the point? */
                    int on;
                    if (m->dim == 3) {
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
                    rintf_t *I = &R->intf[R->nif];
                    I->ei = i;
                    I->ej = j;
                    I->ek = k;
                    I->f = f;
                    I->elocal = le;
                    I->conn = face_connection(m, i, j, k, f, &I->is_here);
                    I->boundary = I->conn < 0;
                    if (I->boundary)
                        I->is_here = 1; /* MPFA.isHere: boundary -> true */
                    R->eintf[le][R->neintf[le]++] = R->nif;
                    R->nif++;
                }
            }
    /* createLocalInterfaceMap (O_MPFA.java:139-152) */
    R->nc = 0;
    for (int a = 0; a < R->nif; ++a) {
        rintf_t *I = &R->intf[a];
        int found = -1;
        if (!I->boundary)
            for (int b = 0; b < a; ++b)
                if (!R->intf[b].boundary && R->intf[b].conn == I->conn) {
                    found = R->intf[b].lidx;
                    break;
                }
        I->lidx = found >= 0 ? found : R->nc++;
    }
}

static v3 intf_center(const mesh_t *m, const rintf_t *I)
{
    v3 c, n;
    double a;
    face_geom(m, I->ei, I->ej, I->ek, I->f, &c, &n, &a);
    return c;
}

/* O_MPFA.getInterfaceContinuity */
static v3 continuity_point(const mesh_t *m, const rintf_t *I, v3 corner,
                           double continuity)
{
    v3 center = intf_center(m, I);
    if (continuity == 1)
        return center;
    return mk((center.x - corner.x) * continuity + corner.x,
              (center.y - corner.y) * continuity + corner.y,
              (center.z - corner.z) * continuity + corner.z);
}

/* Local matrices of one corner (assembleElementContribution,
 * O_MPFA.java:422-457).  When Ae..De are given, each element's own
 * contribution is also kept (Ae + le*144 etc.): every entry is linear in
 * that element's tensor, which the adjoint uses. */
static int corner_local(const mesh_t *m, const double *K6, double continuity,
                        int pi, int pj, int pk, region_t *R, double *A,
                        double *B, double *C, double *D, double *Ae,
                        double *Be, double *Ce, double *De)
{
    build_region(m, pi, pj, pk, R);
    const int nc = R->nc, ne = R->ne;
    memset(A, 0, sizeof(double) * nc * nc);
    memset(C, 0, sizeof(double) * nc * nc);
    memset(B, 0, sizeof(double) * nc * ne);
    memset(D, 0, sizeof(double) * nc * ne);
    if (Ae) {
        memset(Ae, 0, sizeof(double) * 8 * 144);
        memset(Ce, 0, sizeof(double) * 8 * 144);
        memset(Be, 0, sizeof(double) * 8 * 96);
        memset(De, 0, sizeof(double) * 8 * 96);
    }
    v3 corner = point_at(m, pi, pj, pk);
    for (int le = 0; le < ne; ++le) {
        const rel_t *E = &R->el[le];
        int nI = R->neintf[le];
        int nb = nI + 1;
        /* calculateBasisFunctions2D/3D: psi_i(x_j) = delta_ij */
        double Am[16], Psi[16];
        int piv[4];
        v3 elp = elem_center(m, E->ei, E->ej, E->ek);
        memset(Am, 0, sizeof(Am));
        memset(Psi, 0, sizeof(Psi));
        Am[0 + 0 * nb] = 1;
        Am[0 + 1 * nb] = elp.x;
        Am[0 + 2 * nb] = elp.y;
        if (m->dim == 3)
            Am[0 + 3 * nb] = elp.z;
        for (int q = 0; q < nI; ++q) {
            v3 cp = continuity_point(m, &R->intf[R->eintf[le][q]], corner,
                                     continuity);
            Am[q + 1 + 0 * nb] = 1;
            Am[q + 1 + 1 * nb] = cp.x;
            Am[q + 1 + 2 * nb] = cp.y;
            if (m->dim == 3)
                Am[q + 1 + 3 * nb] = cp.z;
        }
        for (int q = 0; q < nb; ++q)
            Psi[q + q * nb] = 1;
        if (or_dgesv(nb, nb, Am, nb, piv, Psi, nb))
            return OR_SINGULAR;
        v3 grad[4];
        for (int q = 0; q < nb; ++q)
            grad[q] = mk(Psi[1 + q * nb], Psi[2 + q * nb],
                         m->dim == 3 ? Psi[3 + q * nb] : 0);
        /* calculateLocalConductivityDirections */
        v3 Kn[3];
        for (int q = 0; q < nI; ++q) {
            const rintf_t *I = &R->intf[R->eintf[le][q]];
            Kn[q] = tensor_mult(K6 + 6 * (size_t)E->e,
                                normal_area(m, I->ei, I->ej, I->ek, I->f, pi,
                                            pj, pk));
        }
        /* assembleElementContribution (O_MPFA.java:422-457) */
        for (int q = 0; q < nI; ++q) {
            const rintf_t *I = &R->intf[R->eintf[le][q]];
            double f = vdot(Kn[q], grad[0]);
            int is = I->lidx;
            int ei = I->elocal;
            B[is + ei * nc] += -f;
            if (I->is_here)
                D[is + ei * nc] += -f;
            if (Ae) {
                Be[le * 96 + is + ei * nc] += -f;
                if (I->is_here)
                    De[le * 96 + is + ei * nc] += -f;
            }
            for (int jj = 1; jj < nI + 1; ++jj) {
                f = vdot(Kn[q], grad[jj]);
                int js = R->intf[R->eintf[le][jj - 1]].lidx;
                A[is + js * nc] += f;
                if (I->is_here)
                    C[is + js * nc] += f;
                if (Ae) {
                    Ae[le * 144 + is + js * nc] += f;
                    if (I->is_here)
                        Ce[le * 144 + is + js * nc] += f;
                }
            }
        }
    }
    return OR_OK;
}

/* calculateT (O_MPFA.java:193-239) for one corner; matrices column-major
 * with leading dimension nc (A, C: nc x nc; B, D, T: nc x ne). */
static int corner_T(const mesh_t *m, const double *K6, double continuity,
                    int pi, int pj, int pk, region_t *R, double *A, double *B,
                    double *C, double *D, double *T)
{
    int st = corner_local(m, K6, continuity, pi, pj, pk, R, A, B, C, D, NULL,
                          NULL, NULL, NULL);
    if (st)
        return st;
    const int nc = R->nc, ne = R->ne;
    /* MPFA.calculateTmatrix: AiB = A\B (dgesv on copies), T = C*AiB - D via
     * dgemm(alpha=1, beta=1) on -D */
    double Acp[144], AiB[96];
    int piv[12];
    memcpy(Acp, A, sizeof(double) * nc * nc);
    memcpy(AiB, B, sizeof(double) * nc * ne);
    if (or_dgesv(nc, ne, Acp, nc, piv, AiB, nc))
        return OR_SINGULAR;
    for (int q = 0; q < nc * ne; ++q)
        T[q] = D[q] * -1;
    for (int j = 0; j < ne; ++j)
        for (int l = 0; l < nc; ++l) {
            double bl = AiB[l + j * nc];
            if (bl != 0) {
                double temp = 1 * bl;
                for (int i = 0; i < nc; ++i)
                    T[i + j * nc] = T[i + j * nc] + temp * C[i + l * nc];
            }
        }
    return OR_OK;
}

static void point_ijk(const mesh_t *m, int p, int *pi, int *pj, int *pk)
{
    int npx = m->nx + 1, npy = m->ny + 1;
    *pi = p % npx;
    *pj = (p / npx) % npy;
    *pk = m->dim == 3 ? p / (npx * npy) : 0;
}

int or_mpfa_corner(int dim, int nx, int ny, int nz, double dx, double dy,
                   double dz, const double *K6, double continuity, int point,
                   int *nc, int *ne, double *A, double *B, double *C,
                   double *D, double *T, int *elems)
{
    mesh_t m;
    int st = mesh_init(&m, dim, nx, ny, nz, dx, dy, dz);
    if (st)
        return st;
    region_t R;
    int pi, pj, pk;
    point_ijk(&m, point, &pi, &pj, &pk);
    st = corner_T(&m, K6, continuity, pi, pj, pk, &R, A, B, C, D, T);
    *nc = R.nc;
    *ne = R.ne;
    for (int q = 0; q < R.ne; ++q)
        elems[q] = R.el[q].e;
    mesh_free(&m);
    return st;
}

/* O_MPFA.calculateTransmissibilities (O_MPFA.java:54-112) + Connection
 * checks (Connection.java:61-107). */
int or_mpfa_o(int dim, int nx, int ny, int nz, double dx, double dy,
              double dz, const double *K6, double continuity, int *md_ptr,
              int *md_elem, double *md_t, int *check_fail)
{
    mesh_t m;
    int st = mesh_init(&m, dim, nx, ny, nz, dx, dy, dz);
    if (st)
        return st;
    if (!(continuity > 0 && continuity <= 1)) {
        mesh_free(&m);
        return OR_BAD_ARGUMENT;
    }
    const int cap = dim == 3 ? 18 : 6;
    int *cnt = (int *)calloc(m.nconn, sizeof(int));
    double A[144], B[96], C[144], D[96], T[96];
    region_t R;
    for (int p = 0; p < m.npoints && !st; ++p) {
        int pi, pj, pk;
        point_ijk(&m, p, &pi, &pj, &pk);
        st = corner_T(&m, K6, continuity, pi, pj, pk, &R, A, B, C, D, T);
        if (st)
            break;
        const int nc = R.nc;
        /* storeTransmissibilities for each connection touching p */
        for (int a = 0; a < R.nif; ++a) {
            const rintf_t *I = &R.intf[a];
            if (I->boundary || !I->is_here)
                continue;
            int c = I->conn;
            int i = I->lidx; /* index(hereInterface(c)) */
            int *el = md_elem + (size_t)c * cap;
            double *tv = md_t + (size_t)c * cap;
            for (int le = 0; le < R.ne; ++le) {
                double tk = T[i + le * nc];
                int e = R.el[le].e, found = 0;
                for (int q = 0; q < cnt[c]; ++q)
                    if (el[q] == e) {
                        found = 1;
                        tv[q] += tk;
                        break;
                    }
                if (!found && tk != 0) {
                    el[cnt[c]] = e;
                    tv[cnt[c]] = tk;
                    cnt[c]++;
                }
            }
        }
    }
    if (st) {
        free(cnt);
        mesh_free(&m);
        return st;
    }
    /* compact into CSR-like lists and run Connection.check */
    int *here = (int *)malloc(sizeof(int) * m.nconn);
    int *there = (int *)malloc(sizeof(int) * m.nconn);
    or_mesh_connections(dim, nx, ny, nz, here, there);
    int pos = 0, fails = 0;
    md_ptr[0] = 0;
    for (int c = 0; c < m.nconn; ++c) {
        double sum = 0;
        int hh = 0, tt = 0;
        for (int q = 0; q < cnt[c]; ++q) {
            int e = md_elem[(size_t)c * cap + q];
            double t = md_t[(size_t)c * cap + q] * 1.0; /* multiplier 1 */
            md_elem[pos] = e;
            md_t[pos] = t;
            sum += t;
            hh |= e == here[c];
            tt |= e == there[c];
            pos++;
        }
        md_ptr[c + 1] = pos;
        if (fabs(sum) > 1.0e-6 || !hh || !tt)
            fails++;
    }
    if (check_fail)
        *check_fail = fails;
    free(here);
    free(there);
    free(cnt);
    mesh_free(&m);
    return OR_OK;
}

int or_element_centers(int dim, int nx, int ny, int nz, double dx, double dy,
                       double dz, double *centers)
{
    mesh_t m;
    int st = mesh_init(&m, dim, nx, ny, nz, dx, dy, dz);
    if (st)
        return st;
    for (int k = 0; k < m.nz; ++k)
        for (int j = 0; j < m.ny; ++j)
            for (int i = 0; i < m.nx; ++i) {
                v3 c = elem_center(&m, i, j, k);
                int e = elem_index(&m, i, j, k);
                centers[3 * e] = c.x;
                centers[3 * e + 1] = c.y;
                centers[3 * e + 2] = c.z;
            }
    mesh_free(&m);
    return OR_OK;
}

/* Per connection: here-interface unit normal * area (x,y,z,area) + center */
int or_interface_geometry(int dim, int nx, int ny, int nz, double dx,
                          double dy, double dz, double *normal_area,
                          double *center)
{
    mesh_t m;
    int st = mesh_init(&m, dim, nx, ny, nz, dx, dy, dz);
    if (st)
        return st;
    for (int k = 0; k < m.nz; ++k)
        for (int j = 0; j < m.ny; ++j)
            for (int i = 0; i < m.nx; ++i)
                for (int f = 1; f < m.nf; f += 2) {
                    int is_here;
                    int c = face_connection(&m, i, j, k, f, &is_here);
                    if (c < 0)
                        continue;
                    v3 cc, n;
                    double a;
                    face_geom(&m, i, j, k, f, &cc, &n, &a);
                    normal_area[4 * c] = n.x;
                    normal_area[4 * c + 1] = n.y;
                    normal_area[4 * c + 2] = n.z;
                    normal_area[4 * c + 3] = a;
                    center[3 * c] = cc.x;
                    center[3 * c + 1] = cc.y;
                    center[3 * c + 2] = cc.z;
                }
    mesh_free(&m);
    return OR_OK;
}

/* ---- Pattern and assembly ------------------------------------------------ */
/* LinearSolver.determineSparsityPattern (LinearSolver.java:111-168):
 * row r = {r} (with_diag) U {element(t) : t in MD(c), r = here/there(c)},
 * sorted (integerSetToSortedArray). */
static void row_insert(int *row, int *len, int v)
{
    for (int q = 0; q < *len; ++q)
        if (row[q] == v)
            return;
    int q = *len;
    while (q > 0 && row[q - 1] > v) {
        row[q] = row[q - 1];
        --q;
    }
    row[q] = v;
    (*len)++;
}

#define PAT_CAP 64

static int build_rows(int n, int nconn, const int *here, const int *there,
                      const int *md_ptr, const int *md_elem, int with_diag,
                      int *rows, int *lens)
{
    for (int r = 0; r < n; ++r) {
        lens[r] = 0;
        if (with_diag)
            row_insert(rows + (size_t)r * PAT_CAP, &lens[r], r);
    }
    for (int c = 0; c < nconn; ++c) {
        int rh = here[c], rt = there[c];
        for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q) {
            int col = md_elem[q];
            if (rh >= 0 && rh < n) {
                if (lens[rh] >= PAT_CAP)
                    return OR_BAD_ARGUMENT;
                row_insert(rows + (size_t)rh * PAT_CAP, &lens[rh], col);
            }
            if (rt >= 0 && rt < n) {
                if (lens[rt] >= PAT_CAP)
                    return OR_BAD_ARGUMENT;
                row_insert(rows + (size_t)rt * PAT_CAP, &lens[rt], col);
            }
        }
    }
    return OR_OK;
}

int or_pattern_count(int n, int nconn, const int *here, const int *there,
                     const int *md_ptr, const int *md_elem, int with_diag,
                     int *rowlen)
{
    int *rows = (int *)malloc(sizeof(int) * (size_t)n * PAT_CAP);
    int st = build_rows(n, nconn, here, there, md_ptr, md_elem, with_diag,
                        rows, rowlen);
    free(rows);
    return st;
}

int or_pattern(int n, int nconn, const int *here, const int *there,
               const int *md_ptr, const int *md_elem, int with_diag,
               int *rowptr, int *colind)
{
    int *rows = (int *)malloc(sizeof(int) * (size_t)n * PAT_CAP);
    int *lens = (int *)malloc(sizeof(int) * n);
    int st = build_rows(n, nconn, here, there, md_ptr, md_elem, with_diag,
                        rows, lens);
    if (!st) {
        rowptr[0] = 0;
        for (int r = 0; r < n; ++r) {
            rowptr[r + 1] = rowptr[r] + lens[r];
            memcpy(colind + rowptr[r], rows + (size_t)r * PAT_CAP,
                   sizeof(int) * lens[r]);
        }
    }
    free(rows);
    free(lens);
    return st;
}

/* PotentialCalculator.discretizeFlux (PotentialCalculator.java:319-333) */
int or_assemble_flux(int n, int nconn, const int *here, const int *there,
                     const int *md_ptr, const int *md_elem, const double *md_t,
                     const int *rowptr, const int *colind, double *data)
{
    (void)n;
    for (int c = 0; c < nconn; ++c) {
        int i = here[c], j = there[c];
        for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q) {
            double tk = md_t[q];
            int k = md_elem[q];
            if (or_csr_add(rowptr, colind, data, i, k, tk))
                return OR_NOT_IN_PATTERN;
            if (or_csr_add(rowptr, colind, data, j, k, -tk))
                return OR_NOT_IN_PATTERN;
        }
    }
    return OR_OK;
}

/* PotentialCalculator.calculateFlux (PotentialCalculator.java:265-280) */
void or_connection_flux(int nconn, const int *md_ptr, const int *md_elem,
                        const double *md_t, const double *p, double *flux)
{
    for (int c = 0; c < nconn; ++c) {
        double f = 0;
        for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q)
            f += md_t[q] * p[md_elem[q]];
        flux[c] = f;
    }
}

/* Net outflow of the listed cells: sum over each cell's connections in
 * ascending connection order of +flux (here side) or -flux (there side). */
void or_cell_rates(int nconn, const int *here, const int *there,
                   const double *flux, int nsel, const int *cells,
                   double *rates)
{
    for (int t = 0; t < nsel; ++t) {
        double q = 0;
        for (int c = 0; c < nconn; ++c) {
            if (here[c] == cells[t])
                q += flux[c];
            if (there[c] == cells[t])
                q += -flux[c];
        }
        rates[t] = q;
    }
}

/* ---- Adjoint of the transmissibility map ------------------------------- */
/* Given tbar (one value per MD entry, the gradient of an objective with
 * respect to the stored transmissibilities), returns for every cell e
 * mbar[e] = d objective / d m_e where the cell's tensor is K_e = exp(m_e)
 * K0_e.  Per corner T = C A^-1 B - D (MPFA.calculateTmatrix) and every
 * entry of A, B, C, D is linear in one element's tensor, so with
 * Tbar the corner's share of tbar:
 *   Bbar = A^-T C' Tbar,  Abar = -Bbar (A^-1 B)',  Cbar = Tbar (A^-1 B)',
 *   Dbar = -Tbar,  mbar_e += <Abar,A_e> + <Bbar,B_e> + <Cbar,C_e> + <Dbar,D_e>,
 * the inner products taken over element e's contributions in assembly
 * order (each contribution of an element lands on its own entry).  This is
 * not a function of the reference (rs has no gradients). */
int or_mpfa_o_vjp(int dim, int nx, int ny, int nz, double dx, double dy,
                  double dz, const double *K6, double continuity,
                  const int *md_ptr, const int *md_elem, const double *tbar,
                  double *mbar)
{
    mesh_t m;
    int st = mesh_init(&m, dim, nx, ny, nz, dx, dy, dz);
    if (st)
        return st;
    for (int e = 0; e < m.ncells; ++e)
        mbar[e] = 0;
    double A[144], B[96], C[144], D[96];
    double *Ae = (double *)malloc(sizeof(double) * 8 * 144);
    double *Ce = (double *)malloc(sizeof(double) * 8 * 144);
    double *Be = (double *)malloc(sizeof(double) * 8 * 96);
    double *De = (double *)malloc(sizeof(double) * 8 * 96);
    region_t R;
    for (int p = 0; p < m.npoints && !st; ++p) {
        int pi, pj, pk;
        point_ijk(&m, p, &pi, &pj, &pk);
        st = corner_local(&m, K6, continuity, pi, pj, pk, &R, A, B, C, D, Ae, Be,
                          Ce, De);
        if (st)
            break;
        const int nc = R.nc, ne = R.ne;
        double Tb[96], AiB[96], W[96], Bb[96], Ab[144], Cb[144], LU[144];
        int piv[12];
        memset(Tb, 0, sizeof(Tb));
        for (int a = 0; a < R.nif; ++a) {
            const rintf_t *I = &R.intf[a];
            if (I->boundary || !I->is_here)
                continue;
            const int c = I->conn, i = I->lidx;
            for (int le = 0; le < ne; ++le)
                for (int q = md_ptr[c]; q < md_ptr[c + 1]; ++q)
                    if (md_elem[q] == R.el[le].e) {
                        Tb[i + le * nc] = tbar[q];
                        break;
                    }
        }
        /* LU of A (dgetf2) and AiB = A^-1 B */
        memcpy(LU, A, sizeof(double) * nc * nc);
        memcpy(AiB, B, sizeof(double) * nc * ne);
        if (or_dgesv(nc, ne, LU, nc, piv, AiB, nc)) {
            st = OR_SINGULAR;
            break;
        }
        /* W = C' Tbar */
        for (int j = 0; j < ne; ++j)
            for (int l = 0; l < nc; ++l) {
                double s = 0;
                for (int i = 0; i < nc; ++i)
                    s += C[i + l * nc] * Tb[i + j * nc];
                W[l + j * nc] = s;
            }
        /* Bbar = A^-T W: U' y = W, L' z = y, then undo the row swaps */
        memcpy(Bb, W, sizeof(double) * nc * ne);
        for (int j = 0; j < ne; ++j) {
            double *b = Bb + j * nc;
            for (int i = 0; i < nc; ++i) {
                double t = b[i];
                for (int k = 0; k < i; ++k)
                    t -= LU[k + i * nc] * b[k];
                b[i] = t / LU[i + i * nc];
            }
            for (int i = nc - 1; i >= 0; --i) {
                double t = b[i];
                for (int k = i + 1; k < nc; ++k)
                    t -= LU[k + i * nc] * b[k];
                b[i] = t;
            }
            for (int i = nc - 1; i >= 0; --i)
                if (piv[i] != i) {
                    double t = b[i];
                    b[i] = b[piv[i]];
                    b[piv[i]] = t;
                }
        }
        /* Abar = -Bbar AiB', Cbar = Tbar AiB' */
        for (int l = 0; l < nc; ++l)
            for (int i = 0; i < nc; ++i) {
                double sa = 0, sc = 0;
                for (int j = 0; j < ne; ++j) {
                    sa += Bb[i + j * nc] * AiB[l + j * nc];
                    sc += Tb[i + j * nc] * AiB[l + j * nc];
                }
                Ab[i + l * nc] = -sa;
                Cb[i + l * nc] = sc;
            }
        /* element terms in assembly order (the contribution walk of
         * corner_local: interface q of the element, then its basis jj) */
        for (int le = 0; le < ne; ++le) {
            double g = 0;
            const int nI = R.neintf[le];
            for (int q = 0; q < nI; ++q) {
                const rintf_t *I = &R.intf[R.eintf[le][q]];
                const int is = I->lidx;
                const double bq = Be[le * 96 + is + le * nc];
                g += Bb[is + le * nc] * bq;
                if (I->is_here)
                    g += (-Tb[is + le * nc]) * De[le * 96 + is + le * nc];
                for (int jj = 1; jj < nI + 1; ++jj) {
                    const int js = R.intf[R.eintf[le][jj - 1]].lidx;
                    g += Ab[is + js * nc] * Ae[le * 144 + is + js * nc];
                    if (I->is_here)
                        g += Cb[is + js * nc] * Ce[le * 144 + is + js * nc];
                }
            }
            mbar[R.el[le].e] += g;
        }
    }
    free(Ae);
    free(Ce);
    free(Be);
    free(De);
    mesh_free(&m);
    return st;
}
