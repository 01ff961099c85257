/*
 * psf_oracle.c -- CPU restatement of the reference's reduced-space fluid step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker: it is compiled
 * into oracle/_build/libpsf_oracle.so and loaded by tests/, by the smoke check
 * and by the CPU-baseline leg of bench.py.  The product library never links,
 * loads or calls it.
 *
 * Every function restates the reference algorithm literally (same loop order,
 * same float/double mixing) and cites the reference file:line it follows.
 * Paths are relative to the reference repository root.
 *
 *   DEC mesh construction    DEC/decmesh3d.cpp:12-106, 1026-1403, 1668-1853
 *   Hodge stars              DEC/dec.cpp:5-197
 *   derivative1/2            DEC/dec.cpp:351-455
 *   gravity field            Solver/SolverImpl/psfsolver.cpp:344-358
 *   buildAdvection           Solver/SolverImpl/psfsolver.cpp:431-551
 *   integrate (step, u=Phi*omega, particles)
 *                            Solver/SolverImpl/psfsolver.cpp:21-234
 *
 * Deliberate, documented departures (none changes a result the reference
 * defines):
 *   - the particle respawn branch draws from a counter-based hash instead of
 *     libc rand() (psfsolver.cpp:219-222 calls rand() inside an OpenMP loop,
 *     a data race with no defined result); the draw -> position formula is
 *     the reference's;
 *   - the file is built with -ffp-contract=off so that results do not depend
 *     on whether the compiler fuses a*b+c (the reference's -march=native build
 *     leaves that to the host CPU).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_EXPORT __attribute__((visibility("default")))

/* GridState (DEC/gridenums.h) */
enum { OUTSIDE = 1, UNINITIALIZED = 2, INSIDE = 3 };
/* DECMesh3D::FaceDirection / EdgeDirection (DEC/decmesh3d.h:18-36) */
enum { FRONT, BACK, TOP, BOTTOM, LEFT, RIGHT };
enum { POSZ, NEGZ, POSY, NEGY, POSX, NEGX };

typedef struct { unsigned id; double center[3]; int inside; unsigned numDual; int e[6]; } o_point;
typedef struct { int id; int inside; double center[3]; unsigned dualCount; int v[2]; int f[4]; } o_edge;
typedef struct { int id; int inside; double center[3]; double normal[3]; double signum; int e[4]; } o_face;
typedef struct { int id; int inside; double center[3]; int f[6]; } o_voxel;

typedef struct ora_mesh {
    int dx, dy, dz;             /* glm::ivec3 dims */
    float resolution;           /* DECMesh3D::resolution (float) */
    float min[3];
    unsigned nXF, nYF, nZF, nXE, nYE, nZE;
    unsigned nV, nF, nE, nP, nSign;
    o_voxel* vox;
    o_face* faces;
    o_edge* edges;
    o_point* pts;
    uint32_t* sign;
} ora_mesh;

/* ---- index closed forms: decmesh3d.cpp:12-60 ---------------------------- */
ORA_EXPORT unsigned ora_voxel_index(const ora_mesh* m, unsigned x, unsigned y, unsigned z)
{ return z * (unsigned)(m->dx * m->dy) + y * (unsigned)m->dx + x; }
ORA_EXPORT unsigned ora_zface_index(const ora_mesh* m, unsigned x, unsigned y, unsigned z)
{ return z * (unsigned)(m->dx * m->dy) + y * (unsigned)m->dx + x; }
ORA_EXPORT unsigned ora_yface_index(const ora_mesh* m, unsigned x, unsigned y, unsigned z)
{ return z * (unsigned)((m->dy + 1) * m->dx) + y * (unsigned)m->dx + x + m->nZF; }
ORA_EXPORT unsigned ora_xface_index(const ora_mesh* m, unsigned x, unsigned y, unsigned z)
{ return z * (unsigned)(m->dy * (m->dx + 1)) + y * (unsigned)(m->dx + 1) + x + m->nZF + m->nYF; }
static unsigned zedge_index(const ora_mesh* m, unsigned x, unsigned y, unsigned z)
{ return z * (unsigned)(m->dx + 1) * (unsigned)(m->dy + 1) + y * (unsigned)(m->dx + 1) + x; }
static unsigned yedge_index(const ora_mesh* m, unsigned x, unsigned y, unsigned z)
{ return z * (unsigned)(m->dx + 1) * (unsigned)m->dy + y * (unsigned)(m->dx + 1) + x + m->nZE; }
static unsigned xedge_index(const ora_mesh* m, unsigned x, unsigned y, unsigned z)
{ return z * (unsigned)m->dx * (unsigned)(m->dy + 1) + y * (unsigned)m->dx + x + m->nZE + m->nYE; }
static unsigned point_index(const ora_mesh* m, unsigned x, unsigned y, unsigned z)
{ return z * (unsigned)((m->dy + 1) * (m->dx + 1)) + y * (unsigned)(m->dx + 1) + x; }
static unsigned sid2idx(int id) { return (unsigned)(labs(id) - 1); }
static int idx2sid(unsigned idx, int s) { return s * (int)(idx + 1); }

/* getFaceSignum(vid, fidx): decmesh3d.cpp:1848-1853 */
ORA_EXPORT int ora_face_signum(const ora_mesh* m, unsigned vid, unsigned fidx)
{
    unsigned idx = vid / 5, off = vid % 5;
    return (m->sign[idx] >> (off * 6 + fidx)) & 1 ? 1 : -1;
}

static void vset(double* d, double a, double b, double c) { d[0] = a; d[1] = b; d[2] = c; }

/* dvec3(min) + dvec3(x*res, y*res, z*res): the products are float (unsigned*float). */
static void grid_pos(const ora_mesh* m, unsigned x, unsigned y, unsigned z, double* out)
{
    float fx = (float)x * m->resolution, fy = (float)y * m->resolution, fz = (float)z * m->resolution;
    out[0] = (double)m->min[0] + (double)fx;
    out[1] = (double)m->min[1] + (double)fy;
    out[2] = (double)m->min[2] + (double)fz;
}

/* addPoint: decmesh3d.cpp:1026-1040 */
static int add_point(ora_mesh* m, int sid, int inside, unsigned x, unsigned y, unsigned z)
{
    unsigned vid = sid2idx(sid);
    o_point* p = &m->pts[vid];
    if (p->inside == UNINITIALIZED) {
        grid_pos(m, x, y, z, p->center);
        p->id = vid;
        p->inside = inside;
        return 1;
    }
    return -1;
}

/* endpoint bookkeeping shared by every branch of addEdge (decmesh3d.cpp:1055-1157) */
static void edge_endpoints(ora_mesh* m, unsigned eid, int inside,
                           unsigned ax, unsigned ay, unsigned az,   /* v1 */
                           unsigned bx, unsigned by, unsigned bz,   /* v2 */
                           unsigned hx, unsigned hy, unsigned hz,   /* the "+1" point */
                           unsigned lx, unsigned ly, unsigned lz)   /* the base point */
{
    o_edge* e = &m->edges[eid];
    e->v[0] = idx2sid(point_index(m, ax, ay, az), 1);
    e->v[1] = idx2sid(point_index(m, bx, by, bz), 1);
    add_point(m, e->v[0], inside, ax, ay, az);
    add_point(m, e->v[1], inside, bx, by, bz);
    o_point* ph = &m->pts[point_index(m, hx, hy, hz)];
    o_point* pl = &m->pts[point_index(m, lx, ly, lz)];
    ph->e[ph->numDual] = idx2sid(eid, 1);
    pl->e[pl->numDual] = idx2sid(eid, 1);
    ph->numDual++;
    pl->numDual++;
}

/* addEdge: decmesh3d.cpp:1043-1199 */
static int add_edge(ora_mesh* m, int sid, int inside, int dir, unsigned x, unsigned y, unsigned z,
                    const double* offset)
{
    unsigned eid = sid2idx(sid);
    o_edge* e = &m->edges[eid];
    if (e->inside == UNINITIALIZED) {
        e->id = (int)eid;
        e->inside = inside;
        memcpy(e->center, offset, sizeof(double) * 3);
        switch (dir) {
        case NEGZ: edge_endpoints(m, eid, inside, x, y, z + 1, x, y, z, x, y, z + 1, x, y, z); break;
        case POSZ: edge_endpoints(m, eid, inside, x, y, z, x, y, z + 1, x, y, z + 1, x, y, z); break;
        case NEGY: edge_endpoints(m, eid, inside, x, y + 1, z, x, y, z, x, y + 1, z, x, y, z); break;
        case POSY: edge_endpoints(m, eid, inside, x, y, z, x, y + 1, z, x, y + 1, z, x, y, z); break;
        case NEGX: edge_endpoints(m, eid, inside, x + 1, y, z, x, y, z, x + 1, y, z, x, y, z); break;
        case POSX: edge_endpoints(m, eid, inside, x, y, z, x + 1, y, z, x + 1, y, z, x, y, z); break;
        }
        return 1;
    }
    int a, b;
    switch (dir) {
    case NEGZ: a = idx2sid(point_index(m, x, y, z + 1), 1); b = idx2sid(point_index(m, x, y, z), 1); break;
    case POSZ: a = idx2sid(point_index(m, x, y, z), 1); b = idx2sid(point_index(m, x, y, z + 1), 1); break;
    case NEGY: a = idx2sid(point_index(m, x, y + 1, z), 1); b = idx2sid(point_index(m, x, y, z), 1); break;
    case POSY: a = idx2sid(point_index(m, x, y, z), 1); b = idx2sid(point_index(m, x, y + 1, z), 1); break;
    case NEGX: a = idx2sid(point_index(m, x + 1, y, z), 1); b = idx2sid(point_index(m, x, y, z), 1); break;
    default:   a = idx2sid(point_index(m, x, y, z), 1); b = idx2sid(point_index(m, x + 1, y, z), 1); break;
    }
    return (e->v[0] == a && e->v[1] == b) ? 1 : -1;
}

/* one edge slot of addFace: register the face on the edge, then addEdge */
typedef struct { unsigned idx; int dir; unsigned x, y, z; double off[3]; } edge_spec;

/* addFace: decmesh3d.cpp:1202-1366 */
static int add_face(ora_mesh* m, int sid, int inside, int dir, unsigned x, unsigned y, unsigned z)
{
    unsigned fid = sid2idx(sid);
    o_face* f = &m->faces[fid];
    if (f->inside != UNINITIALIZED) return -1;
    f->id = (int)fid;
    f->inside = inside;
    double base[3];
    grid_pos(m, x, y, z, base);
    const double r = (double)m->resolution;
    edge_spec s[4];
    /* per direction: signum, normal, centre and the four edges with the
       direction they are created in (decmesh3d.cpp:1209-1358) */
    switch (dir) {
    case FRONT:
        f->signum = -1; vset(f->normal, 0, 0, -1);
        vset(f->center, base[0] + 0.5 * r, base[1] + 0.5 * r, base[2] + 0.5 * 0.0);
        s[0] = (edge_spec){xedge_index(m, x, y + 1, z), POSX, x, y + 1, z, {0, 0.5 * r, 0}};
        s[1] = (edge_spec){yedge_index(m, x + 1, y, z), NEGY, x + 1, y, z, {0.5 * r, 0, 0}};
        s[2] = (edge_spec){xedge_index(m, x, y, z), NEGX, x, y, z, {0, 0.5 * -r, 0}};
        s[3] = (edge_spec){yedge_index(m, x, y, z), POSY, x, y, z, {0.5 * -r, 0, 0}};
        break;
    case BACK:
        f->signum = 1; vset(f->normal, 0, 0, 1);
        vset(f->center, base[0] + 0.5 * r, base[1] + 0.5 * r, base[2] + 0.5 * 0.0);
        s[0] = (edge_spec){xedge_index(m, x, y + 1, z), NEGX, x, y + 1, z, {0, 0.5 * r, 0}};
        s[1] = (edge_spec){yedge_index(m, x, y, z), NEGY, x, y, z, {0.5 * -r, 0, 0}};
        s[2] = (edge_spec){xedge_index(m, x, y, z), POSX, x, y, z, {0, 0.5 * -r, 0}};
        s[3] = (edge_spec){yedge_index(m, x + 1, y, z), POSY, x + 1, y, z, {0.5 * r, 0, 0}};
        break;
    case TOP:
        f->signum = 1; vset(f->normal, 0, -1, 0);
        vset(f->center, base[0] + 0.5 * r, base[1] + 0.5 * 0.0, base[2] + 0.5 * r);
        s[0] = (edge_spec){xedge_index(m, x, y, z + 1), POSX, x, y, z + 1, {0, 0, 0.5 * r}};
        s[1] = (edge_spec){zedge_index(m, x + 1, y, z), NEGZ, x + 1, y, z, {0.5 * r, 0, 0}};
        s[2] = (edge_spec){xedge_index(m, x, y, z), NEGX, x, y, z, {0, 0, 0.5 * -r}};
        s[3] = (edge_spec){zedge_index(m, x, y, z), POSZ, x, y, z, {0.5 * -r, 0, 0}};
        break;
    case BOTTOM:
        f->signum = -1; vset(f->normal, 0, 1, 0);
        vset(f->center, base[0] + 0.5 * r, base[1] + 0.5 * 0.0, base[2] + 0.5 * r);
        s[0] = (edge_spec){xedge_index(m, x, y, z + 1), NEGX, x, y, z + 1, {0, 0, 0.5 * r}};
        s[1] = (edge_spec){zedge_index(m, x, y, z), NEGZ, x, y, z, {0.5 * -r, 0, 0}};
        s[2] = (edge_spec){xedge_index(m, x, y, z), POSX, x, y, z, {0, 0, 0.5 * -r}};
        s[3] = (edge_spec){zedge_index(m, x + 1, y, z), POSZ, x + 1, y, z, {0.5 * r, 0, 0}};
        break;
    case LEFT:
        f->signum = -1; vset(f->normal, 1, 0, 0);
        vset(f->center, base[0] + 0.5 * 0.0, base[1] + 0.5 * r, base[2] + 0.5 * r);
        s[0] = (edge_spec){zedge_index(m, x, y + 1, z), NEGZ, x, y + 1, z, {0, 0.5 * r, 0}};
        s[1] = (edge_spec){yedge_index(m, x, y, z), NEGY, x, y, z, {0, 0, 0.5 * -r}};
        s[2] = (edge_spec){zedge_index(m, x, y, z), POSZ, x, y, z, {0, 0.5 * -r, 0}};
        s[3] = (edge_spec){yedge_index(m, x, y, z + 1), POSY, x, y, z + 1, {0, 0, 0.5 * r}};
        break;
    default: /* RIGHT */
        f->signum = 1; vset(f->normal, -1, 0, 0);
        vset(f->center, base[0] + 0.5 * 0.0, base[1] + 0.5 * r, base[2] + 0.5 * r);
        s[0] = (edge_spec){zedge_index(m, x, y + 1, z), POSZ, x, y + 1, z, {0, 0.5 * r, 0}};
        s[1] = (edge_spec){yedge_index(m, x, y, z + 1), NEGY, x, y, z + 1, {0, 0, 0.5 * r}};
        s[2] = (edge_spec){zedge_index(m, x, y, z), NEGZ, x, y, z, {0, 0.5 * -r, 0}};
        s[3] = (edge_spec){yedge_index(m, x, y, z), POSY, x, y, z, {0, 0, 0.5 * -r}};
        break;
    }
    for (int k = 0; k < 4; ++k) f->e[k] = idx2sid(s[k].idx, 1);
    for (int k = 0; k < 4; ++k) {
        o_edge* e = &m->edges[s[k].idx];
        e->f[e->dualCount] = idx2sid(fid, 1);
    }
    for (int k = 0; k < 4; ++k) m->edges[s[k].idx].dualCount++;
    for (int k = 0; k < 4; ++k) {
        double off[3] = {f->center[0] + s[k].off[0], f->center[1] + s[k].off[1], f->center[2] + s[k].off[2]};
        f->e[k] *= add_edge(m, f->e[k], inside, s[k].dir, s[k].x, s[k].y, s[k].z, off);
    }
    return 1;
}

/* addVoxel: decmesh3d.cpp:1369-1403 */
static void add_voxel(ora_mesh* m, unsigned vid, int inside, unsigned x, unsigned y, unsigned z)
{
    o_voxel* v = &m->vox[vid];
    if (v->inside != UNINITIALIZED) return;
    v->id = (int)vid;
    v->inside = inside;
    double base[3];
    grid_pos(m, x, y, z, base);
    const double r = (double)m->resolution;
    vset(v->center, base[0] + 0.5 * r, base[1] + 0.5 * r, base[2] + 0.5 * r);
    v->f[0] = idx2sid(ora_zface_index(m, x, y, z), 1);
    v->f[1] = idx2sid(ora_zface_index(m, x, y, z + 1), 1);
    v->f[2] = idx2sid(ora_yface_index(m, x, y, z), 1);
    v->f[3] = idx2sid(ora_yface_index(m, x, y + 1, z), 1);
    v->f[4] = idx2sid(ora_xface_index(m, x, y, z), 1);
    v->f[5] = idx2sid(ora_xface_index(m, x + 1, y, z), 1);
    v->f[0] *= add_face(m, v->f[0], inside, FRONT, x, y, z);
    v->f[1] *= add_face(m, v->f[1], inside, BACK, x, y, z + 1);
    v->f[2] *= add_face(m, v->f[2], inside, BOTTOM, x, y, z);
    v->f[3] *= add_face(m, v->f[3], inside, TOP, x, y + 1, z);
    v->f[4] *= add_face(m, v->f[4], inside, LEFT, x, y, z);
    v->f[5] *= add_face(m, v->f[5], inside, RIGHT, x + 1, y, z);
    unsigned idx = vid / 5, off = vid % 5;
    uint32_t orField = 0;
    for (int j = 0; j < 6; ++j) orField |= (uint32_t)(v->f[j] < 0 ? 0 : 1) << (off * 6 + j);
    m->sign[idx] |= orField;
}

/* DECMesh3D(resolution, dims, voxelSize, min): decmesh3d.cpp:62-106 */
ORA_EXPORT ora_mesh* ora_mesh_create(int dx, int dy, int dz, float resolution,
                                     float minx, float miny, float minz)
{
    if (dx <= 0 || dy <= 0 || dz <= 0) return NULL;
    ora_mesh* m = (ora_mesh*)calloc(1, sizeof(ora_mesh));
    m->dx = dx; m->dy = dy; m->dz = dz;
    m->resolution = resolution;
    m->min[0] = minx; m->min[1] = miny; m->min[2] = minz;
    m->nXF = (unsigned)((dx + 1) * (dy * dz));
    m->nYF = (unsigned)((dy + 1) * (dx * dz));
    m->nZF = (unsigned)((dz + 1) * (dx * dy));
    m->nXE = (unsigned)(dx * (dy + 1) * (dz + 1));
    m->nYE = (unsigned)((dx + 1) * dy * (dz + 1));
    m->nZE = (unsigned)((dx + 1) * (dy + 1) * dz);
    m->nV = (unsigned)(dx * dy * dz);
    m->nF = m->nXF + m->nYF + m->nZF;
    m->nE = m->nXE + m->nYE + m->nZE;
    m->nP = (unsigned)((dx + 1) * (dy + 1) * (dz + 1));
    m->nSign = (unsigned)ceil(m->nV / 5.0);
    m->vox = (o_voxel*)calloc(m->nV, sizeof(o_voxel));
    m->faces = (o_face*)calloc(m->nF, sizeof(o_face));
    m->edges = (o_edge*)calloc(m->nE, sizeof(o_edge));
    m->pts = (o_point*)calloc(m->nP, sizeof(o_point));
    m->sign = (uint32_t*)calloc(m->nSign, sizeof(uint32_t));
    for (unsigned i = 0; i < m->nV; ++i) m->vox[i].inside = UNINITIALIZED;
    for (unsigned i = 0; i < m->nF; ++i) m->faces[i].inside = UNINITIALIZED;
    for (unsigned i = 0; i < m->nE; ++i) m->edges[i].inside = UNINITIALIZED;
    for (unsigned i = 0; i < m->nP; ++i) m->pts[i].inside = UNINITIALIZED;
    for (unsigned z = 0; z < (unsigned)dz; ++z)
        for (unsigned y = 0; y < (unsigned)dy; ++y)
            for (unsigned x = 0; x < (unsigned)dx; ++x)
                add_voxel(m, z * (unsigned)(dx * dy) + y * (unsigned)dx + x, OUTSIDE, x, y, z);
    return m;
}

ORA_EXPORT void ora_mesh_destroy(ora_mesh* m)
{
    if (!m) return;
    free(m->vox); free(m->faces); free(m->edges); free(m->pts); free(m->sign); free(m);
}

/* setVoxelInside -> setFaceInside -> setEdgeInside -> setPointInside:
   decmesh3d.cpp:1668-1704 */
ORA_EXPORT void ora_mesh_set_voxel_inside(ora_mesh* m, unsigned vid)
{
    o_voxel* v = &m->vox[vid];
    v->inside = INSIDE;
    for (int j = 0; j < 6; ++j) {
        o_face* f = &m->faces[sid2idx(v->f[j])];
        f->inside = INSIDE;
        for (int k = 0; k < 4; ++k) {
            o_edge* e = &m->edges[sid2idx(f->e[k])];
            e->inside = INSIDE;
            m->pts[sid2idx(e->v[0])].inside = INSIDE;
            m->pts[sid2idx(e->v[1])].inside = INSIDE;
        }
    }
}

/* ---- table export (for tests) ------------------------------------------- */
ORA_EXPORT void ora_mesh_counts(const ora_mesh* m, unsigned* out)
{
    out[0] = m->nV; out[1] = m->nF; out[2] = m->nE; out[3] = m->nP; out[4] = m->nSign;
    out[5] = m->nZF; out[6] = m->nYF; out[7] = m->nXF; out[8] = m->nZE; out[9] = m->nYE; out[10] = m->nXE;
}
ORA_EXPORT void ora_mesh_sign_words(const ora_mesh* m, uint32_t* out)
{ memcpy(out, m->sign, sizeof(uint32_t) * m->nSign); }
ORA_EXPORT void ora_mesh_voxels(const ora_mesh* m, int* f6, int* inside, double* center)
{
    for (unsigned i = 0; i < m->nV; ++i) {
        memcpy(f6 + 6 * i, m->vox[i].f, sizeof(int) * 6);
        inside[i] = m->vox[i].inside;
        memcpy(center + 3 * i, m->vox[i].center, sizeof(double) * 3);
    }
}
ORA_EXPORT void ora_mesh_faces(const ora_mesh* m, int* e4, double* normal, double* center, int* inside)
{
    for (unsigned i = 0; i < m->nF; ++i) {
        memcpy(e4 + 4 * i, m->faces[i].e, sizeof(int) * 4);
        memcpy(normal + 3 * i, m->faces[i].normal, sizeof(double) * 3);
        memcpy(center + 3 * i, m->faces[i].center, sizeof(double) * 3);
        inside[i] = m->faces[i].inside;
    }
}
ORA_EXPORT void ora_mesh_edges(const ora_mesh* m, int* v2, int* f4, unsigned* dualCount, double* center, int* inside)
{
    for (unsigned i = 0; i < m->nE; ++i) {
        memcpy(v2 + 2 * i, m->edges[i].v, sizeof(int) * 2);
        memcpy(f4 + 4 * i, m->edges[i].f, sizeof(int) * 4);
        dualCount[i] = m->edges[i].dualCount;
        memcpy(center + 3 * i, m->edges[i].center, sizeof(double) * 3);
        inside[i] = m->edges[i].inside;
    }
}
ORA_EXPORT void ora_mesh_points(const ora_mesh* m, int* e6, unsigned* numDual, double* center, int* inside)
{
    for (unsigned i = 0; i < m->nP; ++i) {
        memcpy(e6 + 6 * i, m->pts[i].e, sizeof(int) * 6);
        numDual[i] = m->pts[i].numDual;
        memcpy(center + 3 * i, m->pts[i].center, sizeof(double) * 3);
        inside[i] = m->pts[i].inside;
    }
}

/* ---- operators ----------------------------------------------------------- */
static double len3(const double* a, const double* b)
{
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}
static void cross3(const double* a, const double* b, double* o)
{
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
}
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static const o_face* get_face(const ora_mesh* m, int id)
{
    long i = labs(id) - 1;
    return &m->faces[i < 0 ? 0 : i];   /* getFace: std::max(labs(id)-1,0l), decmesh3d.cpp:1817-1820 */
}

/* hodge2(mesh, dual=false): F x F diagonal, decmesh3d voxel loop, dec.cpp:143-193 */
ORA_EXPORT void ora_hodge2_primal(const ora_mesh* m, double* diag)
{
    memset(diag, 0, sizeof(double) * m->nF);
    for (unsigned v = 0; v < m->nV; ++v) {
        const o_voxel* vx = &m->vox[v];
        for (int i = 0; i < 6; ++i) {
            const o_face* f = get_face(m, vx->f[i]);
            const o_edge* e[4];
            for (int k = 0; k < 4; ++k) e[k] = &m->edges[sid2idx(f->e[k])];
            double area = 0.0;
            for (int k = 0; k < 4; ++k)
                area += len3(f->center, e[k]->center) * len3(f->center, e[(k + 1) % 4]->center);
            diag[sid2idx(vx->f[i])] = len3(vx->center, f->center) / area;
        }
    }
}

/* hodge2(mesh, dual=true): E x E diagonal over inside edges, dec.cpp:10-138 */
ORA_EXPORT void ora_hodge2_dual(const ora_mesh* m, double* diag)
{
    memset(diag, 0, sizeof(double) * m->nE);
    for (unsigned p = 0; p < m->nP; ++p) {
        const o_point* pt = &m->pts[p];
        for (int i = 0; i < 6; ++i) {
            if (pt->e[i] == 0) continue;
            const o_edge* f = &m->edges[sid2idx(pt->e[i])];
            if (f->inside != INSIDE) continue;
            const double* v1 = m->pts[sid2idx(f->v[0])].center;
            const double* v2 = m->pts[sid2idx(f->v[1])].center;
            double n1[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
            double ln = sqrt(dot3(n1, n1));
            n1[0] /= ln; n1[1] /= ln; n1[2] /= ln;
            double areaPrim = 0.0;
            for (int j = 0; j < 4; ++j) {
                const o_face* e1 = get_face(m, f->f[j]);
                for (int k = 0; k < 4; ++k) {
                    const o_face* e2 = get_face(m, f->f[k]);
                    if (e2->id == e1->id) continue;
                    double side1 = len3(f->center, e1->center);
                    double side2 = len3(f->center, e2->center);
                    double cr[3];
                    cross3(e1->normal, e2->normal, cr);
                    areaPrim += side1 * side2 * fabs(dot3(n1, cr)) * (f->f[j] != 0) * (f->f[k] != 0);
                }
            }
            areaPrim /= 2;
            if (areaPrim != 0.0)
                diag[sid2idx(pt->e[i])] = 1.0 / (len3(pt->center, f->center) / areaPrim);
        }
    }
}

/* derivative1(mesh, dual=true): E x F triplets (edge, face, sign), dec.cpp:381-455.
   rows/cols/vals must hold 4*F entries; entry 4*f+k is edge k of face f. */
ORA_EXPORT void ora_derivative1(const ora_mesh* m, int* rows, int* cols, double* vals)
{
    for (unsigned i = 0; i < m->nF; ++i) {
        const o_face* f = &m->faces[i];
        for (int k = 0; k < 4; ++k) {
            rows[4 * i + k] = (int)sid2idx(f->e[k]);
            cols[4 * i + k] = f->id;
            vals[4 * i + k] = f->e[k] >= 0 ? 1.0 : -1.0;
        }
    }
}

/* derivative2(mesh): V x F signed incidence over inside voxels, dec.cpp:351-378.
   Writes 6*V triplets; entries of voxels that are not inside are the
   default-constructed Eigen::Triplet (0,0,0), exactly as the reference's
   pre-sized triplet list leaves them. */
ORA_EXPORT void ora_derivative2(const ora_mesh* m, int* vox_rows, int* face_cols, double* vals)
{
    for (unsigned i = 0; i < m->nV; ++i) {
        const o_voxel* v = &m->vox[i];
        for (int k = 0; k < 6; ++k) {
            if (v->inside == INSIDE) {
                vox_rows[6 * i + k] = v->id;
                face_cols[6 * i + k] = (int)sid2idx(v->f[k]);
                vals[6 * i + k] = v->f[k] > 0 ? 1.0 : -1.0;
            } else {
                vox_rows[6 * i + k] = 0; face_cols[6 * i + k] = 0; vals[6 * i + k] = 0.0;
            }
        }
    }
}

/* gravity face field before projection: psfsolver.cpp:344-357 */
ORA_EXPORT void ora_gravity_field(const ora_mesh* m, double* f)
{
    memset(f, 0, sizeof(double) * m->nF);
    const double ey[3] = {0.0, 1.0, 0.0};
    const double eps = 2.220446049250313e-16;
    for (unsigned i = 0; i < m->nF; ++i) {
        const o_face* fc = &m->faces[i];
        if (fc->inside != INSIDE) continue;
        if (fabs(dot3(ey, fc->normal)) > eps) {
            if (dot3(fc->normal, ey) != 0.0)
                f[fc->id] = -(fc->normal[1] > 0 ? -1 : 1) * 0.00981;
        }
    }
}

/* buildAdvection: psfsolver.cpp:431-551.
 * Phi: F x K column-major (Eigen MatrixXd), V = curl*Phi: E x K column-major,
 * lambda: K.  Output C: K matrices K x K, each column-major, matrix i at
 * C + i*K*K (the layout of std::vector<Eigen::MatrixXd> advection).
 * W_i is built one i at a time; its entries accumulate in the reference's
 * voxel order, so each W_i is the reference's wedges[i] to the bit. */
ORA_EXPORT int ora_build_advection(const ora_mesh* m, int K, const double* Phi, const double* lambda,
                                   const double* V, double* C, int nthreads)
{
    const unsigned E = m->nE, F = m->nF;
    const float hr = m->resolution / 2;
    const double scale = (double)(hr * hr);
    const double one3[3] = {1.0, 1.0, 1.0};
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    int fail = 0;
#pragma omp parallel
    {
        double* W = (double*)malloc(sizeof(double) * (size_t)E * K);
        if (!W) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(dynamic, 1)
        for (int i = 0; i < K; ++i) {
            if (!W) continue;
            memset(W, 0, sizeof(double) * (size_t)E * K);
            for (unsigned vi = 0; vi < m->nV; ++vi) {
                const o_voxel* fit = &m->vox[vi];
                if (fit->inside != INSIDE) continue;
                const o_face* f[6];
                for (int k = 0; k < 6; ++k) f[k] = get_face(m, fit->f[k]);
                const unsigned ie[12] = {
                    sid2idx(f[0]->e[0]), sid2idx(f[0]->e[1]), sid2idx(f[0]->e[2]), sid2idx(f[0]->e[3]),
                    sid2idx(f[1]->e[0]), sid2idx(f[1]->e[1]), sid2idx(f[1]->e[2]), sid2idx(f[1]->e[3]),
                    sid2idx(f[4]->e[0]), sid2idx(f[4]->e[2]), sid2idx(f[5]->e[0]), sid2idx(f[5]->e[2])};
                double c[6][6];
                for (int a = 0; a < 6; ++a)
                    for (int b = 0; b < 6; ++b) {
                        double cr[3];
                        cross3(f[a]->normal, f[b]->normal, cr);
                        c[a][b] = dot3(cr, one3);
                    }
                size_t fr[6];
                for (int k = 0; k < 6; ++k) fr[k] = (size_t)(labs(fit->f[k]) - 1);
                const double va1 = Phi[fr[0] + (size_t)i * F], va2 = Phi[fr[1] + (size_t)i * F];
                const double va3 = Phi[fr[2] + (size_t)i * F], va4 = Phi[fr[3] + (size_t)i * F];
                const double va5 = Phi[fr[4] + (size_t)i * F], va6 = Phi[fr[5] + (size_t)i * F];
                for (int j = 0; j < K; ++j) {
                    const double vb1 = Phi[fr[0] + (size_t)j * F], vb2 = Phi[fr[1] + (size_t)j * F];
                    const double vb3 = Phi[fr[2] + (size_t)j * F], vb4 = Phi[fr[3] + (size_t)j * F];
                    const double vb5 = Phi[fr[4] + (size_t)j * F], vb6 = Phi[fr[5] + (size_t)j * F];
                    double* Wj = W + (size_t)j * E;
                    Wj[ie[0]] += scale * (va1 * vb4 - vb1 * va4) * c[0][3];
                    Wj[ie[1]] += scale * (va6 * vb1 - vb6 * va1) * c[5][0];
                    Wj[ie[2]] += scale * (va3 * vb1 - vb3 * va1) * c[2][0];
                    Wj[ie[3]] += scale * (va1 * vb5 - vb1 * va5) * c[0][4];
                    Wj[ie[4]] += scale * (va4 * vb2 - vb4 * va2) * c[3][1];
                    Wj[ie[5]] += scale * (va5 * vb2 - vb5 * va2) * c[4][1];
                    Wj[ie[6]] += scale * (va2 * vb3 - vb2 * va3) * c[1][2];
                    Wj[ie[7]] += scale * (va2 * vb6 - vb2 * va6) * c[1][5];
                    Wj[ie[8]] += scale * (va5 * vb4 - vb5 * va4) * c[4][3];
                    Wj[ie[9]] += scale * (va3 * vb5 - vb3 * va5) * c[2][4];
                    Wj[ie[10]] += scale * (va4 * vb6 - vb4 * va6) * c[3][5];
                    Wj[ie[11]] += scale * (va6 * vb3 - vb6 * va3) * c[5][2];
                }
            }
            /* advection[i].col(j) = (1/lambda_i) * wedges[i].col(j); then
               advection[i] = V^T * advection[i]  (psfsolver.cpp:539-550) */
            const double inv = 1.0 / lambda[i];
            for (size_t t = 0; t < (size_t)E * K; ++t) W[t] = inv * W[t];
            double* Ci = C + (size_t)i * K * K;
            for (int j = 0; j < K; ++j)
                for (int k = 0; k < K; ++k) {
                    const double* Vk = V + (size_t)k * E;
                    const double* Wj = W + (size_t)j * E;
                    double s = 0.0;
                    for (unsigned e = 0; e < E; ++e) s += Vk[e] * Wj[e];
                    Ci[(size_t)j * K + k] = s;
                }
        }
        free(W);
    }
    return fail ? -1 : 0;
}

/* ---- the step (psfsolver.cpp:34-60) ------------------------------------- */

/* One reduced step, in place on omega (K), with C as K column-major K x K
   matrices.  y_k = (omega^T C_k) omega  (psfsolver.cpp:39-44). */
ORA_EXPORT void ora_reduced_step(int K, double* omega, const double* C, const double* lambda,
                                 const double* g, double dt, double nu, int gravity_active,
                                 double* y_out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    double e1 = 0.0;
    for (int k = 0; k < K; ++k) e1 += omega[k] * omega[k];
    double* vel = (double*)malloc(sizeof(double) * K);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; ++k) {
        const double* Ck = C + (size_t)k * K * K;
        double yk = 0.0;
        for (int j = 0; j < K; ++j) {
            double t = 0.0;                                   /* (omega^T C_k)_j */
            for (int a = 0; a < K; ++a) t += omega[a] * Ck[(size_t)j * K + a];
            yk += t * omega[j];
        }
        vel[k] = yk;
    }
    for (int k = 0; k < K; ++k) omega[k] += dt * vel[k];
    double e2 = 0.0;
    for (int k = 0; k < K; ++k) e2 += omega[k] * omega[k];
    const double fac = sqrt(e1 / e2);
    for (int k = 0; k < K; ++k) omega[k] *= fac;
    for (int k = 0; k < K; ++k) omega[k] *= exp(nu * lambda[k] * dt);
    if (gravity_active)
        for (int k = 0; k < K; ++k) omega[k] += dt * g[k];
    if (y_out) memcpy(y_out, vel, sizeof(double) * K);
    free(vel);
}

/* velocityField = (velBasisField * basisCoeff).cast<float>()  (psfsolver.cpp:60).
   Phi is F x K column-major. */
ORA_EXPORT void ora_reconstruct(int F, int K, const double* Phi, const double* omega, float* u, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(static)
    for (int f = 0; f < F; ++f) {
        double s = 0.0;
        for (int k = 0; k < K; ++k) s += Phi[(size_t)k * F + f] * omega[k];
        u[f] = (float)s;
    }
}

/* ---- particles (psfsolver.cpp:70-232) ------------------------------------ */

/* counter-based draws in [0,1024), standing in for rand()%1024: one
 * SplitMix64 output (Steele, Lea and Flood 2014: a Weyl increment, then the
 * variant-13 finaliser) of a key mixing (seed, step, particle); draw k
 * (0 <= k < 6) is bits 10k .. 10k+9 of it.  The particle step uses k = 0, 1, 2. */
ORA_EXPORT uint32_t ora_respawn_draw(uint32_t seed, uint32_t step, uint32_t particle, uint32_t k)
{
    uint64_t z = ((uint64_t)seed << 32) ^ ((uint64_t)step * 0x9E3779B97F4A7C15ull)
               ^ ((uint64_t)particle * 0xD1B54A32D192ED03ull);
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return k < 6 ? (uint32_t)((z >> (10 * k)) & 1023u) : 0u;
}

static float fmix(float x, float y, float a) { return x * (1.0f - a) + y * a; }   /* glm::mix */
static int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static float fclampf(float v, float lo, float hi)
{
    float t = (v < lo) ? lo : v;        /* glm::max: (x < y) ? y : x */
    return (hi < t) ? hi : t;           /* glm::min: (y < x) ? y : x */
}

typedef struct {
    float aabb_min[3], aabb_max[3];
    double resolution;      /* AbstractSolver::resolution (double) */
    double timestep;        /* AbstractSolver::timeStep */
    float lifetime;         /* AbstractSolver::lifeTime */
    uint32_t seed, step;
} ora_particle_params;

/* One pass of the particle loop over N particles (x,y,z,w float4 records).
   cells_out (optional, 12 ints per particle) receives the clamped
   x/y/z Ofs, Ofs-, Ofs+ of particles that advected this pass, -1 otherwise. */
ORA_EXPORT void ora_particles(const ora_mesh* m, const ora_particle_params* pp, const float* u,
                              float* parts, int N, int* cells_out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const int dx = m->dx, dy = m->dy, dz = m->dz;
    const double res = pp->resolution;
    const float ext[3] = {pp->aabb_max[0] - pp->aabb_min[0], pp->aabb_max[1] - pp->aabb_min[1],
                          pp->aabb_max[2] - pp->aabb_min[2]};
    const float hf = (float)res;
    const float inv = (float)(1.0 / res);
    const float dtf = (float)pp->timestep;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        float* p = parts + 4 * (size_t)n;
        int* co = cells_out ? cells_out + 12 * (size_t)n : NULL;
        if (p[3] > 0.0f) {
            /* (aabb.min + aabb.getExtent() + p) is float; "/resolution" and
               "-resolution/2" promote to double (psfsolver.cpp:80-89) */
            const float sx = (pp->aabb_min[0] + ext[0]) + p[0];
            const float sy = (pp->aabb_min[1] + ext[1]) + p[1];
            const float sz = (pp->aabb_min[2] + ext[2]) + p[2];
            int yOfs = (int)((double)sy / res);
            int xOfs = (int)((double)sx / res);
            int zOfs = (int)((double)sz / res);
            int zM = (int)(((double)sz - res / 2) / res);
            int zP = (int)(((double)sz + res / 2) / res);
            int yM = (int)(((double)sy - res / 2) / res);
            int yP = (int)(((double)sy + res / 2) / res);
            int xM = (int)(((double)sx - res / 2) / res);
            int xP = (int)(((double)sx + res / 2) / res);
            xOfs = iclamp(xOfs, 0, dx - 1); yOfs = iclamp(yOfs, 0, dy - 1); zOfs = iclamp(zOfs, 0, dz - 1);
            xM = iclamp(xM, 0, dx - 1); xP = iclamp(xP, 0, dx - 1);
            yM = iclamp(yM, 0, dy - 1); yP = iclamp(yP, 0, dy - 1);
            zM = iclamp(zM, 0, dz - 1); zP = iclamp(zP, 0, dz - 1);
            if (co) { co[0] = xOfs; co[1] = yOfs; co[2] = zOfs; co[3] = xM; co[4] = xP;
                      co[5] = yM; co[6] = yP; co[7] = zM; co[8] = zP; co[9] = co[10] = co[11] = 0; }

            unsigned v1x = ora_voxel_index(m, xOfs, yM, zM), v2x = ora_voxel_index(m, xOfs, yP, zM);
            unsigned v3x = ora_voxel_index(m, xOfs, yM, zP), v4x = ora_voxel_index(m, xOfs, yP, zP);
            unsigned v1y = ora_voxel_index(m, xM, yOfs, zM), v2y = ora_voxel_index(m, xP, yOfs, zM);
            unsigned v3y = ora_voxel_index(m, xM, yOfs, zP), v4y = ora_voxel_index(m, xP, yOfs, zP);
            unsigned v1z = ora_voxel_index(m, xM, yM, zOfs), v2z = ora_voxel_index(m, xP, yM, zOfs);
            unsigned v3z = ora_voxel_index(m, xM, yP, zOfs), v4z = ora_voxel_index(m, xP, yP, zOfs);

            float cf1x[3], cf1y[3], cf1z[3];
            cf1x[0] = (pp->aabb_min[0] + hf * (float)xOfs) + 0.5f * 0.0f;
            cf1x[1] = (pp->aabb_min[1] + hf * (float)yM) + 0.5f * hf;
            cf1x[2] = (pp->aabb_min[2] + hf * (float)zM) + 0.5f * hf;
            cf1y[0] = (pp->aabb_min[0] + hf * (float)xM) + 0.5f * hf;
            cf1y[1] = (pp->aabb_min[1] + hf * (float)yOfs) + 0.5f * 0.0f;
            cf1y[2] = (pp->aabb_min[2] + hf * (float)zM) + 0.5f * hf;
            cf1z[0] = (pp->aabb_min[0] + hf * (float)xM) + 0.5f * hf;
            cf1z[1] = (pp->aabb_min[1] + hf * (float)yM) + 0.5f * hf;
            cf1z[2] = (pp->aabb_min[2] + hf * (float)zOfs) + 0.5f * 0.0f;

            unsigned v1xf5 = ora_xface_index(m, xOfs, yM, zM), v1xf6 = ora_xface_index(m, xOfs + 1, yM, zM);
            unsigned v2xf5 = ora_xface_index(m, xOfs, yP, zM), v2xf6 = ora_xface_index(m, xOfs + 1, yP, zM);
            unsigned v3xf5 = ora_xface_index(m, xOfs, yM, zP), v3xf6 = ora_xface_index(m, xOfs + 1, yM, zP);
            unsigned v4xf5 = ora_xface_index(m, xOfs, yP, zP), v4xf6 = ora_xface_index(m, xOfs + 1, yP, zP);
            unsigned v1yf3 = ora_yface_index(m, xM, yOfs, zM), v1yf4 = ora_yface_index(m, xM, yOfs + 1, zM);
            unsigned v2yf3 = ora_yface_index(m, xP, yOfs, zM), v2yf4 = ora_yface_index(m, xP, yOfs + 1, zM);
            unsigned v3yf3 = ora_yface_index(m, xM, yOfs, zP), v3yf4 = ora_yface_index(m, xM, yOfs + 1, zP);
            unsigned v4yf3 = ora_yface_index(m, xP, yOfs, zP), v4yf4 = ora_yface_index(m, xP, yOfs + 1, zP);
            unsigned v1zf1 = ora_zface_index(m, xM, yM, zOfs), v1zf2 = ora_zface_index(m, xM, yM, zOfs + 1);
            unsigned v2zf1 = ora_zface_index(m, xP, yM, zOfs), v2zf2 = ora_zface_index(m, xP, yM, zOfs + 1);
            unsigned v3zf1 = ora_zface_index(m, xM, yP, zOfs), v3zf2 = ora_zface_index(m, xM, yP, zOfs + 1);
            unsigned v4zf1 = ora_zface_index(m, xP, yP, zOfs), v4zf2 = ora_zface_index(m, xP, yP, zOfs + 1);

#define SG(v, j) ((float)ora_face_signum(m, (v), (j)))
            float vel1x = -SG(v1x, 4) * u[v1xf5], vel3x = SG(v1x, 5) * u[v1xf6];
            float vel2x = -SG(v2x, 4) * u[v2xf5], vel4x = SG(v2x, 5) * u[v2xf6];
            float vel5x = -SG(v3x, 4) * u[v3xf5], vel7x = SG(v3x, 5) * u[v3xf6];
            float vel6x = -SG(v4x, 4) * u[v4xf5], vel8x = SG(v4x, 5) * u[v4xf6];
            float vel1y = -SG(v1y, 2) * u[v1yf3], vel2y = SG(v1y, 3) * u[v1yf4];
            float vel3y = -SG(v2y, 2) * u[v2yf3], vel4y = SG(v2y, 3) * u[v2yf4];
            float vel5y = -SG(v3y, 2) * u[v3yf3], vel6y = SG(v3y, 3) * u[v3yf4];
            float vel7y = -SG(v4y, 2) * u[v4yf3], vel8y = SG(v4y, 3) * u[v4yf4];
            float vel1z = SG(v1z, 0) * u[v1zf1], vel2z = SG(v3z, 0) * u[v3zf1];
            float vel3z = SG(v2z, 0) * u[v2zf1], vel4z = SG(v4z, 0) * u[v4zf1];
            float vel5z = -SG(v1z, 1) * u[v1zf2], vel6z = -SG(v3z, 1) * u[v3zf2];
            float vel7z = -SG(v2z, 1) * u[v2zf2], vel8z = -SG(v4z, 1) * u[v4zf2];
#undef SG
            float nx[3], ny[3], nz[3], vel[3];
            for (int c = 0; c < 3; ++c) {
                nx[c] = inv * (p[c] - cf1x[c]);
                ny[c] = inv * (p[c] - cf1y[c]);
                nz[c] = inv * (p[c] - cf1z[c]);
            }
            {   /* x: mix over y, then x, then z */
                float a0 = fmix(vel1x, vel2x, nx[1]), a1 = fmix(vel3x, vel4x, nx[1]);
                float a2 = fmix(vel5x, vel6x, nx[1]), a3 = fmix(vel7x, vel8x, nx[1]);
                float b0 = fmix(a0, a1, nx[0]), b1 = fmix(a2, a3, nx[0]);
                vel[0] = fmix(b0, b1, nx[2]);
            }
            {
                float a0 = fmix(vel1y, vel2y, ny[1]), a1 = fmix(vel3y, vel4y, ny[1]);
                float a2 = fmix(vel5y, vel6y, ny[1]), a3 = fmix(vel7y, vel8y, ny[1]);
                float b0 = fmix(a0, a1, ny[0]), b1 = fmix(a2, a3, ny[0]);
                vel[1] = fmix(b0, b1, ny[2]);
            }
            {
                float a0 = fmix(vel1z, vel2z, nz[1]), a1 = fmix(vel3z, vel4z, nz[1]);
                float a2 = fmix(vel5z, vel6z, nz[1]), a3 = fmix(vel7z, vel8z, nz[1]);
                float b0 = fmix(a0, a1, nz[0]), b1 = fmix(a2, a3, nz[0]);
                vel[2] = fmix(b0, b1, nz[2]);
            }
            for (int c = 0; c < 3; ++c)
                p[c] = fclampf(p[c] + dtf * vel[c], pp->aabb_min[c], pp->aabb_max[c]);
            p[3] = p[3] - 1.0f;
        } else {
            /* respawn: psfsolver.cpp:219-222 */
            if (co) for (int c = 0; c < 12; ++c) co[c] = -1;
            const float center_z = pp->aabb_min[2] + 0.5f * (pp->aabb_max[2] - pp->aabb_min[2]);
            uint32_t r0 = ora_respawn_draw(pp->seed, pp->step, (uint32_t)n, 0);
            uint32_t r1 = ora_respawn_draw(pp->seed, pp->step, (uint32_t)n, 1);
            uint32_t r2 = ora_respawn_draw(pp->seed, pp->step, (uint32_t)n, 2);
            p[0] = (float)((r0 / 1024.0 - 0.5) * 0.25);
            p[1] = pp->aabb_min[1] + 0.2f;
            p[2] = (float)(-(double)center_z + (r1 / 1024.0 - 0.5) * 0.25);
            p[3] = (float)(pp->lifetime * (r2 / 1024.0));
        }
    }
}
