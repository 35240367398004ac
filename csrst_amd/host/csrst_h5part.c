/*
 * csrst_h5part -- converts the driver's restart records (host/checkpoint.hpp)
 * to and from an H5Part file in the layout the reference's H5hut calls give
 * (SaveH5Part.cpp:79-353, LoadH5Part.cpp:60-321):
 *
 *   /                       file attributes: units (tUnit ... TIMEUnit) and
 *                           every command-line argument (string, float64,
 *                           int64; switches as int64 0/1)
 *   /Step#k                 one group per saved step, attribute TIME [ns]
 *   /Step#k/x y z oldx oldy px py pz q   float64[n] (z = 1.00000001, pz = 0)
 *   /Step#k/id              int64[n]
 *   /Step#k/Block/Efield    datasets 0 1 2: float64[1][Ny][Nx] (Ex, Ey, 0),
 *                           attributes __Spacing__ (dx, dy, 1) and
 *                           __Origin__ (ox, oy, 0)
 *   /Step#k/Block/Hfield    (Hzx, Hzy, Hzx + Hzy), origin shifted by half a cell
 *   /Step#k/Block/Jfield    (Jx, Jy, 0)
 *
 *   csrst_h5part write RECORDS OUT.h5part   append every record as a step
 *                                            (a new file gets all attributes;
 *                                            an existing one the reduced set
 *                                            of SaveH5Part.cpp:355-428)
 *   csrst_h5part read IN.h5part RECORDS     the last step and the file
 *                                            attributes as one record
 *
 * Plain C over the HDF5 C library, so that no C++ runtime other than the
 * system's is involved; the GPU driver never links HDF5.
 */
#include <hdf5.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    char *name;
    char kind;  /* 's', 'd', 'i' */
    char *s;
    double d;
    int64_t i;
} attr_t;

typedef struct {
    int64_t Nx, Ny, n, step;
    double time, hx, hy, ox, oy;
    double *E, *H, *J;                                /* 2 * Nx * Ny each */
    double *x, *y, *oldx, *oldy, *px, *py, *q;         /* n each */
    int64_t *id;
    int64_t nattr;
    attr_t *attrs;
} record_t;

static void die(const char *msg, const char *what)
{
    fprintf(stderr, "csrst_h5part: %s%s%s\n", msg, what ? ": " : "", what ? what : "");
    exit(1);
}

static void *xmalloc(size_t b)
{
    void *p = malloc(b ? b : 1);
    if (!p) die("out of memory", NULL);
    return p;
}

/* ---- restart records -------------------------------------------------------- */

static void rd(FILE *f, void *p, size_t b)
{
    if (b && fread(p, 1, b, f) != b) die("truncated restart file", NULL);
}

static char *rd_str(FILE *f)
{
    int32_t l;
    rd(f, &l, 4);
    if (l < 0 || l > (1 << 20)) die("bad string length", NULL);
    char *s = xmalloc((size_t)l + 1);
    rd(f, s, (size_t)l);
    s[l] = 0;
    return s;
}

/* reads the next record; returns 0 at the end of the file */
static int read_record(FILE *f, record_t *r)
{
    char tag[4];
    if (fread(tag, 1, 4, f) != 4) return 0;
    if (memcmp(tag, "STEP", 4)) die("bad record tag", NULL);
    rd(f, &r->Nx, 8); rd(f, &r->Ny, 8); rd(f, &r->n, 8); rd(f, &r->step, 8);
    if (r->Nx <= 0 || r->Ny <= 0 || r->n < 0 || r->Nx * r->Ny > ((int64_t)1 << 34) || r->n > ((int64_t)1 << 34))
        die("bad record sizes", NULL);
    rd(f, &r->time, 8); rd(f, &r->hx, 8); rd(f, &r->hy, 8); rd(f, &r->ox, 8); rd(f, &r->oy, 8);
    const size_t cells = (size_t)(r->Nx * r->Ny), n = (size_t)r->n;
    double **fl[3] = {&r->E, &r->H, &r->J};
    for (int k = 0; k < 3; ++k) {
        *fl[k] = xmalloc(2 * cells * 8);
        rd(f, *fl[k], 2 * cells * 8);
    }
    double **pa[7] = {&r->x, &r->y, &r->oldx, &r->oldy, &r->px, &r->py, &r->q};
    for (int k = 0; k < 7; ++k) {
        *pa[k] = xmalloc(n * 8);
        rd(f, *pa[k], n * 8);
    }
    r->id = xmalloc(n * 8);
    rd(f, r->id, n * 8);
    rd(f, &r->nattr, 8);
    if (r->nattr < 0 || r->nattr > 4096) die("bad attribute count", NULL);
    r->attrs = xmalloc((size_t)r->nattr * sizeof(attr_t));
    for (int64_t k = 0; k < r->nattr; ++k) {
        attr_t *a = &r->attrs[k];
        memset(a, 0, sizeof *a);
        a->name = rd_str(f);
        rd(f, &a->kind, 1);
        if (a->kind == 's') a->s = rd_str(f);
        else if (a->kind == 'd') rd(f, &a->d, 8);
        else if (a->kind == 'i') rd(f, &a->i, 8);
        else die("bad attribute kind", a->name);
    }
    return 1;
}

static void wr(FILE *f, const void *p, size_t b)
{
    if (b && fwrite(p, 1, b, f) != b) die("write failed", NULL);
}

static void wr_str(FILE *f, const char *s)
{
    const int32_t l = (int32_t)strlen(s);
    wr(f, &l, 4);
    wr(f, s, (size_t)l);
}

static void write_record(FILE *f, const record_t *r)
{
    wr(f, "STEP", 4);
    wr(f, &r->Nx, 8); wr(f, &r->Ny, 8); wr(f, &r->n, 8); wr(f, &r->step, 8);
    wr(f, &r->time, 8); wr(f, &r->hx, 8); wr(f, &r->hy, 8); wr(f, &r->ox, 8); wr(f, &r->oy, 8);
    const size_t cells = (size_t)(r->Nx * r->Ny), n = (size_t)r->n;
    const double *fl[3] = {r->E, r->H, r->J};
    for (int k = 0; k < 3; ++k) wr(f, fl[k], 2 * cells * 8);
    const double *pa[7] = {r->x, r->y, r->oldx, r->oldy, r->px, r->py, r->q};
    for (int k = 0; k < 7; ++k) wr(f, pa[k], n * 8);
    wr(f, r->id, n * 8);
    wr(f, &r->nattr, 8);
    for (int64_t k = 0; k < r->nattr; ++k) {
        const attr_t *a = &r->attrs[k];
        wr_str(f, a->name);
        wr(f, &a->kind, 1);
        if (a->kind == 's') wr_str(f, a->s);
        else if (a->kind == 'd') wr(f, &a->d, 8);
        else wr(f, &a->i, 8);
    }
}

/* ---- HDF5 helpers -------------------------------------------------------------- */

#define H5CHK(x)                                                   \
    do {                                                           \
        if ((x) < 0) die("HDF5 call failed", #x);                  \
    } while (0)

static void attr_write(hid_t obj, const char *name, hid_t type, const void *v, hsize_t count)
{
    if (H5Aexists(obj, name) > 0) H5CHK(H5Adelete(obj, name));
    hid_t sp = H5Screate_simple(1, &count, NULL);
    hid_t a = H5Acreate2(obj, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5CHK(a);
    H5CHK(H5Awrite(a, type, v));
    H5Aclose(a);
    H5Sclose(sp);
}

static void attr_string(hid_t obj, const char *name, const char *v)
{
    if (H5Aexists(obj, name) > 0) H5CHK(H5Adelete(obj, name));
    hid_t t = H5Tcopy(H5T_C_S1);
    H5CHK(H5Tset_size(t, strlen(v) + 1));
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(obj, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5CHK(a);
    H5CHK(H5Awrite(a, t, v));
    H5Aclose(a);
    H5Sclose(sp);
    H5Tclose(t);
}

static void dset_write(hid_t g, const char *name, hid_t type, int rank, const hsize_t *dims, const void *v)
{
    hid_t sp = H5Screate_simple(rank, dims, NULL);
    hid_t d = H5Dcreate2(g, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5CHK(d);
    H5CHK(H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v));
    H5Dclose(d);
    H5Sclose(sp);
}

static int is_step_name(const char *s, long *k)
{
    if (strncmp(s, "Step#", 5) || !s[5]) return 0;
    char *end;
    const long v = strtol(s + 5, &end, 10);
    if (*end || v < 0) return 0;
    *k = v;
    return 1;
}

/* number of steps and the largest step index (H5GetNumSteps) */
static long count_steps(hid_t file, long *last)
{
    H5G_info_t info;
    H5CHK(H5Gget_info(file, &info));
    long count = 0;
    *last = -1;
    for (hsize_t k = 0; k < info.nlinks; ++k) {
        char name[256];
        if (H5Lget_name_by_idx(file, ".", H5_INDEX_NAME, H5_ITER_INC, k, name, sizeof name, H5P_DEFAULT) < 0)
            continue;
        long s;
        if (is_step_name(name, &s)) {
            ++count;
            if (s > *last) *last = s;
        }
    }
    return count;
}

static const attr_t *find_attr(const record_t *r, const char *name)
{
    for (int64_t k = 0; k < r->nattr; ++k)
        if (!strcmp(r->attrs[k].name, name)) return &r->attrs[k];
    return NULL;
}

static void write_cli_attr(hid_t file, const attr_t *a)
{
    if (a->kind == 's') attr_string(file, a->name, a->s);
    else if (a->kind == 'd') attr_write(file, a->name, H5T_NATIVE_DOUBLE, &a->d, 1);
    else attr_write(file, a->name, H5T_NATIVE_INT64, &a->i, 1);
}

/* SaveH5Part::write for one record, as step group number k */
static void write_step(hid_t file, long k, const record_t *r)
{
    char name[64];
    snprintf(name, sizeof name, "Step#%ld", k);
    hid_t g = H5Gcreate2(file, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5CHK(g);
    const double tns = r->time * 1e9;
    attr_write(g, "TIME", H5T_NATIVE_DOUBLE, &tns, 1);
    /* at least one particle (SaveH5Part.cpp:93, 117, 173): 0 and id -1 */
    const size_t n = (size_t)r->n, nloc = n ? n : 1;
    double *buf = xmalloc(nloc * 8);
    const hsize_t dn = nloc;
    const double *src[9] = {r->x, r->y, NULL, r->oldx, r->oldy, r->px, r->py, NULL, r->q};
    const char *names[9] = {"x", "y", "z", "oldx", "oldy", "px", "py", "pz", "q"};
    for (int c = 0; c < 9; ++c) {
        buf[0] = 0.0;
        for (size_t p = 0; p < n; ++p) buf[p] = src[c] ? src[c][p] : (c == 2 ? 1.00000001 : 0.0);
        dset_write(g, names[c], H5T_NATIVE_DOUBLE, 1, &dn, buf);
    }
    int64_t *ibuf = (int64_t *)buf;
    ibuf[0] = -1;
    for (size_t p = 0; p < n; ++p) ibuf[p] = r->id[p];
    dset_write(g, "id", H5T_NATIVE_INT64, 1, &dn, ibuf);
    free(buf);

    hid_t b = H5Gcreate2(g, "Block", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5CHK(b);
    const size_t cells = (size_t)(r->Nx * r->Ny);
    const hsize_t dims[3] = {1, (hsize_t)r->Ny, (hsize_t)r->Nx};
    double *z = xmalloc(cells * 8);
    const char *fields[3] = {"Efield", "Hfield", "Jfield"};
    const double *planes[3] = {r->E, r->H, r->J};
    for (int f = 0; f < 3; ++f) {
        hid_t fg = H5Gcreate2(b, fields[f], H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        H5CHK(fg);
        dset_write(fg, "0", H5T_NATIVE_DOUBLE, 3, dims, planes[f]);
        dset_write(fg, "1", H5T_NATIVE_DOUBLE, 3, dims, planes[f] + cells);
        for (size_t c = 0; c < cells; ++c) z[c] = f == 1 ? planes[f][c] + planes[f][cells + c] : 0.0;
        dset_write(fg, "2", H5T_NATIVE_DOUBLE, 3, dims, z);
        const double sp[3] = {r->hx, r->hy, 1.0};
        const double half = f == 1 ? 0.5 : 0.0;
        const double org[3] = {r->ox + half * r->hx, r->oy + half * r->hy, 0.0};
        attr_write(fg, "__Spacing__", H5T_NATIVE_DOUBLE, sp, 3);
        attr_write(fg, "__Origin__", H5T_NATIVE_DOUBLE, org, 3);
        H5Gclose(fg);
    }
    free(z);
    H5Gclose(b);
    H5Gclose(g);
}

static int do_write(const char *in, const char *out)
{
    FILE *f = fopen(in, "rb");
    if (!f) die("cannot open", in);
    char m[8];
    if (fread(m, 1, 8, f) != 8 || memcmp(m, "CSRSTRS1", 8)) die("not a restart file", in);
    record_t r;
    if (!read_record(f, &r)) die("no records", in);
    FILE *probe = fopen(out, "rb");
    hid_t file;
    long next = 0;
    if (probe) {
        fclose(probe);
        file = H5Fopen(out, H5F_ACC_RDWR, H5P_DEFAULT);
        H5CHK(file);
        long last;
        next = count_steps(file, &last);
        if (last + 1 > next) next = last + 1;
        /* SaveH5Part.cpp:355-428: the reduced set, int-valued ones only */
        const char *reduced[5] = {"startstep", "numsteps", "savemodulus", "order", "debug"};
        for (int k = 0; k < 5; ++k) {
            const attr_t *a = find_attr(&r, reduced[k]);
            if (a && a->kind == 'i') write_cli_attr(file, a);
        }
    } else {
        file = H5Fcreate(out, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
        H5CHK(file);
        /* SaveH5Part.cpp:274-286 units, then every argument */
        const char *units[11][2] = {{"tUnit", "s"},      {"xUnit", "m"},           {"yUnit", "m"},
                                    {"zUnit", "m"},      {"pxUnit", "#beta#gamma"}, {"pyUnit", "#beta#gamma"},
                                    {"pzUnit", "#beta#gamma"}, {"qUnit", "Cb"},    {"idUnit", "1"},
                                    {"SPOSUnit", "m"},   {"TIMEUnit", "ns"}};
        for (int k = 0; k < 11; ++k) attr_string(file, units[k][0], units[k][1]);
        for (int64_t k = 0; k < r.nattr; ++k) write_cli_attr(file, &r.attrs[k]);
    }
    do {
        write_step(file, next++, &r);
    } while (read_record(f, &r));
    fclose(f);
    H5CHK(H5Fclose(file));
    return 0;
}

static double *read_dset(hid_t g, const char *name, hid_t type, size_t expect, size_t *count)
{
    hid_t d = H5Dopen2(g, name, H5P_DEFAULT);
    if (d < 0) die("missing dataset", name);
    hid_t sp = H5Dget_space(d);
    const hssize_t np = H5Sget_simple_extent_npoints(sp);
    if (np < 0 || (expect && (size_t)np != expect)) die("dataset size", name);
    double *v = xmalloc((size_t)np * 8);
    H5CHK(H5Dread(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v));
    if (count) *count = (size_t)np;
    H5Sclose(sp);
    H5Dclose(d);
    return v;
}

typedef struct {
    record_t *r;
} iter_ctx;

static herr_t collect_attr(hid_t loc, const char *name, const H5A_info_t *info, void *op)
{
    (void)info;
    record_t *r = ((iter_ctx *)op)->r;
    hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
    if (a < 0) return 0;
    hid_t t = H5Aget_type(a), sp = H5Aget_space(a);
    const hssize_t np = H5Sget_simple_extent_npoints(sp);
    attr_t v;
    memset(&v, 0, sizeof v);
    int keep = 1;
    switch (H5Tget_class(t)) {
    case H5T_STRING: {
        const size_t sz = H5Tget_size(t);
        v.kind = 's';
        v.s = xmalloc(sz + 1);
        hid_t mt = H5Tcopy(H5T_C_S1);
        H5Tset_size(mt, sz);
        if (np != 1 || H5Aread(a, mt, v.s) < 0) keep = 0;
        v.s[sz] = 0;
        H5Tclose(mt);
        break;
    }
    case H5T_FLOAT:
        v.kind = 'd';
        if (np != 1 || H5Aread(a, H5T_NATIVE_DOUBLE, &v.d) < 0) keep = 0;
        break;
    case H5T_INTEGER:
        v.kind = 'i';
        if (np != 1 || H5Aread(a, H5T_NATIVE_INT64, &v.i) < 0) keep = 0;
        break;
    default:
        keep = 0;
    }
    if (keep) {
        v.name = strdup(name);
        r->attrs = realloc(r->attrs, (size_t)(r->nattr + 1) * sizeof(attr_t));
        if (!r->attrs || !v.name) die("out of memory", NULL);
        r->attrs[r->nattr++] = v;
    }
    H5Sclose(sp);
    H5Tclose(t);
    H5Aclose(a);
    return 0;
}

/* LoadH5Part: the last step, the fields, the bunch and the file attributes */
static int do_read(const char *in, const char *out)
{
    hid_t file = H5Fopen(in, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) die("cannot open", in);
    long last;
    if (count_steps(file, &last) <= 0) die("no steps in", in);
    char name[64];
    snprintf(name, sizeof name, "Step#%ld", last);
    hid_t g = H5Gopen2(file, name, H5P_DEFAULT);
    H5CHK(g);
    record_t r;
    memset(&r, 0, sizeof r);
    r.step = last;
    hid_t ta = H5Aopen(g, "TIME", H5P_DEFAULT);
    if (ta < 0) die("missing TIME", name);
    H5CHK(H5Aread(ta, H5T_NATIVE_DOUBLE, &r.time));
    H5Aclose(ta);
    r.time *= 1e-9;   /* LoadH5Part.cpp:181 */
    size_t n;
    r.x = read_dset(g, "x", H5T_NATIVE_DOUBLE, 0, &n);
    r.n = (int64_t)n;
    r.y = read_dset(g, "y", H5T_NATIVE_DOUBLE, n, NULL);
    r.oldx = read_dset(g, "oldx", H5T_NATIVE_DOUBLE, n, NULL);
    r.oldy = read_dset(g, "oldy", H5T_NATIVE_DOUBLE, n, NULL);
    r.px = read_dset(g, "px", H5T_NATIVE_DOUBLE, n, NULL);
    r.py = read_dset(g, "py", H5T_NATIVE_DOUBLE, n, NULL);
    r.q = read_dset(g, "q", H5T_NATIVE_DOUBLE, n, NULL);
    if (H5Lexists(g, "id", H5P_DEFAULT) > 0) {
        r.id = (int64_t *)read_dset(g, "id", H5T_NATIVE_INT64, n, NULL);
    } else {
        r.id = xmalloc(n * 8);
        for (size_t p = 0; p < n; ++p) r.id[p] = (int64_t)p;
    }
    const char *fields[3] = {"Block/Efield", "Block/Hfield", "Block/Jfield"};
    double **planes[3] = {&r.E, &r.H, &r.J};
    for (int f = 0; f < 3; ++f) {
        hid_t fg = H5Gopen2(g, fields[f], H5P_DEFAULT);
        if (fg < 0) die("missing field", fields[f]);
        hid_t d = H5Dopen2(fg, "0", H5P_DEFAULT);
        if (d < 0) die("missing component 0 of", fields[f]);
        hid_t sp = H5Dget_space(d);
        hsize_t dims[3] = {0, 0, 0};
        if (H5Sget_simple_extent_ndims(sp) != 3) die("field rank", fields[f]);
        H5Sget_simple_extent_dims(sp, dims, NULL);
        H5Sclose(sp);
        H5Dclose(d);
        if (dims[0] != 1) die("field with more than one z layer", fields[f]);
        if (f == 0) {
            r.Nx = (int64_t)dims[2];
            r.Ny = (int64_t)dims[1];
            double v[3];
            hid_t a = H5Aopen(fg, "__Spacing__", H5P_DEFAULT);
            if (a < 0) die("missing __Spacing__", fields[f]);
            H5CHK(H5Aread(a, H5T_NATIVE_DOUBLE, v));
            H5Aclose(a);
            r.hx = v[0]; r.hy = v[1];
            a = H5Aopen(fg, "__Origin__", H5P_DEFAULT);
            if (a < 0) die("missing __Origin__", fields[f]);
            H5CHK(H5Aread(a, H5T_NATIVE_DOUBLE, v));
            H5Aclose(a);
            r.ox = v[0]; r.oy = v[1];
        } else if ((int64_t)dims[2] != r.Nx || (int64_t)dims[1] != r.Ny) {
            die("field sizes differ", fields[f]);
        }
        const size_t cells = (size_t)(r.Nx * r.Ny);
        *planes[f] = xmalloc(2 * cells * 8);
        double *c0 = read_dset(fg, "0", H5T_NATIVE_DOUBLE, cells, NULL);
        double *c1 = read_dset(fg, "1", H5T_NATIVE_DOUBLE, cells, NULL);
        memcpy(*planes[f], c0, cells * 8);
        memcpy(*planes[f] + cells, c1, cells * 8);
        free(c0);
        free(c1);
        H5Gclose(fg);
    }
    H5Gclose(g);
    iter_ctx ctx = {&r};
    hsize_t idx = 0;
    H5CHK(H5Aiterate2(file, H5_INDEX_NAME, H5_ITER_INC, &idx, collect_attr, &ctx));
    H5Fclose(file);
    FILE *f = fopen(out, "wb");
    if (!f) die("cannot write", out);
    wr(f, "CSRSTRS1", 8);
    write_record(f, &r);
    if (fclose(f)) die("write failed", out);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 4 && !strcmp(argv[1], "write")) return do_write(argv[2], argv[3]);
    if (argc == 4 && !strcmp(argv[1], "read")) return do_read(argv[2], argv[3]);
    fprintf(stderr, "usage: %s write RECORDS OUT.h5part | read IN.h5part RECORDS\n", argv[0]);
    return 2;
}
