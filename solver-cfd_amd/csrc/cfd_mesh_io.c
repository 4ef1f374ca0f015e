/* cfd_mesh_io.c — the producers and consumers either side of the pseudo-time iteration:
 * the reference mesh and porous-block flags, the nodal velocity average, and the state
 * file formats (final-state table and restart files).
 *
 * Mesh and flags follow OpenMp_Multicore/nonsymetric_mesh.f90 statement by statement (the
 * loops update the node arrays in place, so their order is part of the result). The file
 * formats are those of output.f90 (writes) and initial.f90 (restart reads, whose loop order
 * differs from the write order; both orders are offered).
 */
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cfd_internal.h"

int cfd__fail(int code, const char *fmt, ...);

/* nonsymetric_mesh.f90:7-44. dx_c = Lhori/(imax-1) (comum.f90:45); the y split index is
 * int(y_down/dx_c)+1, and the lower branch mirrors about -Y(jmax) as the reference does
 * (which is why its Y(1) is not -y_down when the domain is not symmetric). */
int cfd_mesh_reference(int32_t imax, int32_t jmax, double lhori, double y_down, double px, double py, double q,
                       double *x, double *y)
{
    if (imax < 3 || jmax < 3 || !x || !y) return cfd__fail(CFD_ERR_ARG, "cfd_mesh_reference: bad arguments");
    if (!(lhori > 0) || !(q > 0)) return cfd__fail(CFD_ERR_ARG, "cfd_mesh_reference: Lhori and Q must be positive");
    const double dxc = lhori / (double)(imax - 1);
    const int js = (int)(y_down / dxc) + 1;
    if (js < 1 || js > jmax) return cfd__fail(CFD_ERR_ARG, "cfd_mesh_reference: y split %d outside 1..jmax", js);
    for (int i = 1; i <= imax; ++i) x[i] = ((double)i - 1.0) * dxc - lhori / 2.0;
    for (int j = 1; j <= jmax; ++j) y[j] = ((double)j - 1.0) * dxc - y_down;
    const int c = imax / 2 + 1;
    const double tq = tanh(q);
    for (int i = c; i <= imax; ++i) {
        const double e = (x[i] - x[imax]) / (x[c] - x[imax]);
        const double s = px * e + (1.0 - px) * (1.0 - (tanh(q * (1.0 - e))) / tq);
        x[i] = x[imax] - s * (x[imax] - x[c]);
    }
    for (int i = c; i >= 1; --i) {
        const double e = (x[i] + x[imax]) / (x[c] + x[imax]);
        const double s = px * e + (1.0 - px) * (1.0 - (tanh(q * (1.0 - e))) / tq);
        x[i] = -x[imax] - s * (-x[imax] - x[c]);
    }
    for (int j = js; j <= jmax; ++j) {
        const double e = (y[j] - y[jmax]) / (y[js] - y[jmax]);
        const double s = py * e + (1.0 - py) * (1.0 - (tanh(q * (1.0 - e))) / tq);
        y[j] = y[jmax] - s * (y[jmax] - y[js]);
    }
    for (int j = js; j >= 1; --j) {
        const double e = (y[j] + y[jmax]) / (y[js] + y[jmax]);
        const double s = py * e + (1.0 - py) * (1.0 - (tanh(q * (1.0 - e))) / tq);
        y[j] = -y[jmax] - s * (-y[jmax] - y[js]);
    }
    return CFD_OK;
}

/* nonsymetric_mesh.f90:165-218 with the block x <= x_edge, y >= y_edge (the reference uses
 * -4 and 0.5). Flags: 0 fluid, 1 boundary, 2 interior, 3 fluid next to a boundary cell;
 * every non-fluid cell is porous (eps = porosity, L = 1). The updates are in place and in
 * the reference's loop order. The third loop reads flag(0, j) and flag(i, 0): with the
 * reference's column-major storage flag(0, j) is flag(imax, j-1); reads before the array
 * start are taken as fluid. */
int cfd_porous_flags_reference(int32_t imax, int32_t jmax, const double *x, const double *y, double x_edge,
                               double y_edge, double porosity, double *eps, double *lp, int32_t *flags)
{
    enum { F = 0, B = 1, I = 2, BS = 3 };
    if (imax < 3 || jmax < 3 || !x || !y || !eps || !lp) return cfd__fail(CFD_ERR_ARG, "cfd_porous_flags_reference: bad arguments");
    const int ld = jmax + 1;
    int *fl = (int *)calloc((size_t)(imax + 2) * ld, sizeof(int));
    if (!fl) return cfd__fail(CFD_ERR_ARG, "out of host memory");
#define FL(i, j) fl[(size_t)(i) * ld + (j)]
    for (int j = 1; j <= jmax; ++j)
        for (int i = 1; i <= imax; ++i) FL(i, j) = (x[i] <= x_edge && y[j] >= y_edge) ? I : F;
    for (int j = 2; j <= jmax - 1; ++j)
        for (int i = 2; i <= imax - 1; ++i)
            if ((FL(i, j) == I && FL(i - 1, j) == F) || (FL(i, j) == I && FL(i, j - 1) == F)
                || (FL(i, j) == I && FL(i + 1, j) == F) || (FL(i, j) == I && FL(i, j + 1) == F))
                FL(i, j) = B;
    for (int j = 2; j <= jmax - 1; ++j)
        if ((FL(1, j) == I && FL(1, j - 1) == F) || (FL(1, j) == I && FL(1, j + 1) == F)) FL(1, j) = B;
    for (int j = 1; j <= jmax - 1; ++j)
        for (int i = 1; i <= imax - 1; ++i) {
            const int w = i > 1 ? FL(i - 1, j) : (j > 1 ? FL(imax, j - 1) : F);
            const int s = j > 1 ? FL(i, j - 1) : F;
            if ((FL(i, j) == F && w == B) || (FL(i, j) == F && s == B)
                || (FL(i, j) == F && FL(i + 1, j) == B) || (FL(i, j) == F && FL(i, j + 1) == B))
                FL(i, j) = BS;
        }
    for (int j = 1; j <= jmax; ++j)
        for (int i = 1; i <= imax; ++i) {
            const size_t k = (size_t)i * ld + j;
            const int porous = FL(i, j) != F;
            eps[k] = porous ? porosity : 1.0;
            lp[k] = porous ? 1.0 : 0.0;
            if (flags) flags[k] = FL(i, j);
        }
#undef FL
    free(fl);
    return CFD_OK;
}

/* comp_mean.f90:1-25 on the host: u(i,j) = (um(i+1,j) + um(i,j)) / 2, v(i,j) = (vm(i,j+1) +
 * vm(i,j)) / 2, C layouts in, p layout out. */
int cfd_nodal_velocity(int32_t imax, int32_t jmax, const double *um, const double *vm, double *u, double *v)
{
    if (imax < 1 || jmax < 1 || !um || !vm || !u || !v) return cfd__fail(CFD_ERR_ARG, "cfd_nodal_velocity: bad arguments");
    const size_t lu = (size_t)jmax + 1, lv = (size_t)jmax + 2;
    for (int i = 1; i <= imax; ++i)
        for (int j = 1; j <= jmax; ++j) {
            u[i * lu + j] = (um[(i + 1) * lu + j] + um[i * lu + j]) * 0.5;
            v[i * lu + j] = (vm[i * lv + j + 1] + vm[i * lv + j]) * 0.5;
        }
    return CFD_OK;
}

/* ---- files ------------------------------------------------------------------------- */
static int put(FILE *f, double v) { return fprintf(f, " %.17E", v) < 0; }

/* output.f90:12-18: one record per node, i outer, j inner: X(i) Y(j) u v p T C. */
int cfd_write_state_table(const char *path, int32_t imax, int32_t jmax, const double *x, const double *y,
                          const double *u, const double *v, const double *p, const double *T, const double *C)
{
    if (!path || !x || !y || !u || !v || !p || !T || !C) return cfd__fail(CFD_ERR_ARG, "cfd_write_state_table: NULL argument");
    FILE *f = fopen(path, "w");
    if (!f) return cfd__fail(CFD_ERR_ARG, "cannot open %s: %s", path, strerror(errno));
    const size_t ld = (size_t)jmax + 1;
    int bad = 0;
    for (int i = 1; i <= imax && !bad; ++i)
        for (int j = 1; j <= jmax && !bad; ++j) {
            const size_t k = i * ld + j;
            bad = put(f, x[i]) || put(f, y[j]) || put(f, u[k]) || put(f, v[k]) || put(f, p[k]) || put(f, T[k])
                  || put(f, C[k]) || fputc('\n', f) == EOF;
        }
    if (fclose(f) != 0) bad = 1;
    return bad ? cfd__fail(CFD_ERR_ARG, "write error on %s", path) : CFD_OK;
}

/* All numbers of a list-directed text file (records may wrap over lines; Fortran 'D'
 * exponents are accepted). Returns a malloc'd array and its length. */
static int read_numbers(const char *path, double **out, size_t *n)
{
    FILE *f = fopen(path, "r");
    if (!f) return cfd__fail(CFD_ERR_ARG, "cannot open %s: %s", path, strerror(errno));
    size_t cap = 1 << 16, len = 0;
    double *a = (double *)malloc(cap * sizeof(double));
    char tok[128];
    int c, r = CFD_OK;
    size_t tl = 0;
    while (a && r == CFD_OK) {
        c = fgetc(f);
        if (c == EOF || isspace(c) || c == ',') {
            if (tl) {
                tok[tl] = 0;
                for (size_t q = 0; q < tl; ++q) if (tok[q] == 'd' || tok[q] == 'D') tok[q] = 'E';
                char *end = NULL;
                const double v = strtod(tok, &end);
                if (!end || *end) { r = cfd__fail(CFD_ERR_ARG, "%s: not a number: '%s'", path, tok); break; }
                if (len == cap) {
                    cap *= 2;
                    double *b = (double *)realloc(a, cap * sizeof(double));
                    if (!b) { free(a); a = NULL; break; }
                    a = b;
                }
                a[len++] = v;
                tl = 0;
            }
            if (c == EOF) break;
        } else if (tl + 1 < sizeof tok) {
            tok[tl++] = (char)c;
        } else {
            r = cfd__fail(CFD_ERR_ARG, "%s: token too long", path);
        }
    }
    fclose(f);
    if (!a) return cfd__fail(CFD_ERR_ARG, "out of host memory");
    if (r != CFD_OK) { free(a); return r; }
    *out = a;
    *n = len;
    return CFD_OK;
}

int cfd_read_state_table(const char *path, int32_t imax, int32_t jmax, double *x, double *y, double *u, double *v,
                         double *p, double *T, double *C)
{
    if (!path) return cfd__fail(CFD_ERR_ARG, "cfd_read_state_table: NULL path");
    double *a = NULL;
    size_t n = 0;
    int r = read_numbers(path, &a, &n);
    if (r != CFD_OK) return r;
    const size_t want = (size_t)7 * imax * jmax;
    if (n != want) { free(a); return cfd__fail(CFD_ERR_ARG, "%s: %zu numbers, expected %zu (7 x %d x %d)", path, n, want, imax, jmax); }
    const size_t ld = (size_t)jmax + 1;
    size_t q = 0;
    for (int i = 1; i <= imax; ++i)
        for (int j = 1; j <= jmax; ++j, q += 7) {
            const size_t k = i * ld + j;
            if (x && j == 1) x[i] = a[q];
            if (y && i == 1) y[j] = a[q + 1];
            if (u) u[k] = a[q + 2];
            if (v) v[k] = a[q + 3];
            if (p) p[k] = a[q + 4];
            if (T) T[k] = a[q + 5];
            if (C) C[k] = a[q + 6];
        }
    free(a);
    return CFD_OK;
}

static int join(char *buf, size_t n, const char *dir, const char *name)
{
    const int w = snprintf(buf, n, "%s/%s", dir, name);
    return (w < 0 || (size_t)w >= n) ? cfd__fail(CFD_ERR_ARG, "path too long") : CFD_OK;
}

/* output.f90:20-43: restartU.dat (um, i = 1..imax+1 outer, j = 1..jmax inner), restartV.dat
 * (vm, i = 1..imax, j = 1..jmax+1), restartPTC.dat (p T C per line, i outer). */
int cfd_write_restart(const char *dir, int32_t imax, int32_t jmax, const double *um, const double *vm,
                      const double *p, const double *T, const double *C)
{
    if (!dir || !um || !vm || !p || !T || !C) return cfd__fail(CFD_ERR_ARG, "cfd_write_restart: NULL argument");
    char path[4096];
    const size_t lu = (size_t)jmax + 1, lv = (size_t)jmax + 2;
    int bad = 0;
    if (join(path, sizeof path, dir, "restartU.dat") != CFD_OK) return CFD_ERR_ARG;
    FILE *f = fopen(path, "w");
    if (!f) return cfd__fail(CFD_ERR_ARG, "cannot open %s: %s", path, strerror(errno));
    for (int i = 1; i <= imax + 1 && !bad; ++i)
        for (int j = 1; j <= jmax && !bad; ++j) bad = fprintf(f, " %.17E\n", um[i * lu + j]) < 0;
    if (fclose(f) != 0) bad = 1;
    if (bad || join(path, sizeof path, dir, "restartV.dat") != CFD_OK) return cfd__fail(CFD_ERR_ARG, "restart write failed");
    f = fopen(path, "w");
    if (!f) return cfd__fail(CFD_ERR_ARG, "cannot open %s: %s", path, strerror(errno));
    for (int i = 1; i <= imax && !bad; ++i)
        for (int j = 1; j <= jmax + 1 && !bad; ++j) bad = fprintf(f, " %.17E\n", vm[i * lv + j]) < 0;
    if (fclose(f) != 0) bad = 1;
    if (bad || join(path, sizeof path, dir, "restartPTC.dat") != CFD_OK) return cfd__fail(CFD_ERR_ARG, "restart write failed");
    f = fopen(path, "w");
    if (!f) return cfd__fail(CFD_ERR_ARG, "cannot open %s: %s", path, strerror(errno));
    for (int i = 1; i <= imax && !bad; ++i)
        for (int j = 1; j <= jmax && !bad; ++j) {
            const size_t k = i * lu + j;
            bad = fprintf(f, " %.17E %.17E %.17E\n", p[k], T[k], C[k]) < 0;
        }
    if (fclose(f) != 0) bad = 1;
    return bad ? cfd__fail(CFD_ERR_ARG, "restart write failed") : CFD_OK;
}

/* initial.f90:52-87 reads each file with j outer and i inner, which is not the order
 * output.f90 writes (CFD_RESTART_ORDER_REFERENCE reproduces that read; a file written by
 * the reference and read back this way comes in transposed within the flattened order).
 * CFD_RESTART_ORDER_WRITTEN reads i outer, the order of the writes, for exact round trips. */
int cfd_read_restart(const char *dir, int32_t imax, int32_t jmax, int32_t order, double *um, double *vm, double *p,
                     double *T, double *C)
{
    if (!dir || !um || !vm || !p || !T || !C) return cfd__fail(CFD_ERR_ARG, "cfd_read_restart: NULL argument");
    if (order != CFD_RESTART_ORDER_REFERENCE && order != CFD_RESTART_ORDER_WRITTEN)
        return cfd__fail(CFD_ERR_ARG, "unknown restart order %d", order);
    const int jfirst = order == CFD_RESTART_ORDER_REFERENCE;
    const size_t lu = (size_t)jmax + 1, lv = (size_t)jmax + 2;
    char path[4096];
    double *a = NULL;
    size_t n = 0, q = 0;
    /* um: (imax+1) x jmax */
    if (join(path, sizeof path, dir, "restartU.dat") != CFD_OK) return CFD_ERR_ARG;
    int r = read_numbers(path, &a, &n);
    if (r != CFD_OK) return r;
    if (n != (size_t)(imax + 1) * jmax) { free(a); return cfd__fail(CFD_ERR_ARG, "%s: %zu values, expected %d", path, n, (imax + 1) * jmax); }
    q = 0;
    if (jfirst) { for (int j = 1; j <= jmax; ++j) for (int i = 1; i <= imax + 1; ++i) um[i * lu + j] = a[q++]; }
    else { for (int i = 1; i <= imax + 1; ++i) for (int j = 1; j <= jmax; ++j) um[i * lu + j] = a[q++]; }
    free(a);
    /* vm: imax x (jmax+1) */
    if (join(path, sizeof path, dir, "restartV.dat") != CFD_OK) return CFD_ERR_ARG;
    r = read_numbers(path, &a, &n);
    if (r != CFD_OK) return r;
    if (n != (size_t)imax * (jmax + 1)) { free(a); return cfd__fail(CFD_ERR_ARG, "%s: %zu values, expected %d", path, n, imax * (jmax + 1)); }
    q = 0;
    if (jfirst) { for (int j = 1; j <= jmax + 1; ++j) for (int i = 1; i <= imax; ++i) vm[i * lv + j] = a[q++]; }
    else { for (int i = 1; i <= imax; ++i) for (int j = 1; j <= jmax + 1; ++j) vm[i * lv + j] = a[q++]; }
    free(a);
    /* p, T, C: imax x jmax records of three */
    if (join(path, sizeof path, dir, "restartPTC.dat") != CFD_OK) return CFD_ERR_ARG;
    r = read_numbers(path, &a, &n);
    if (r != CFD_OK) return r;
    if (n != (size_t)3 * imax * jmax) { free(a); return cfd__fail(CFD_ERR_ARG, "%s: %zu values, expected %d", path, n, 3 * imax * jmax); }
    q = 0;
    if (jfirst) {
        for (int j = 1; j <= jmax; ++j)
            for (int i = 1; i <= imax; ++i, q += 3) { p[i * lu + j] = a[q]; T[i * lu + j] = a[q + 1]; C[i * lu + j] = a[q + 2]; }
    } else {
        for (int i = 1; i <= imax; ++i)
            for (int j = 1; j <= jmax; ++j, q += 3) { p[i * lu + j] = a[q]; T[i * lu + j] = a[q + 1]; C[i * lu + j] = a[q + 2]; }
    }
    free(a);
    return CFD_OK;
}

/* ParaView file of transient.f90:55-103: structured grid of the nodes (x(i), y(j), 0), point
 * data |u| (nodal velocity magnitude), T, P, and the vector (u, v, 0); i outer, j inner,
 * fixed-point numbers with 4 decimals in 14 columns. The header counts use the reference's
 * widths (I3 for the dimensions, I6 for the point counts) and are written in full where they
 * do not fit (the reference's format would print asterisks there). */
int cfd_write_vtk(const char *path, const char *title, int32_t imax, int32_t jmax, const double *x, const double *y,
                  const double *u, const double *v, const double *p, const double *T)
{
    if (!path || !x || !y || !u || !v || !p || !T || imax < 1 || jmax < 1)
        return cfd__fail(CFD_ERR_ARG, "cfd_write_vtk: bad arguments");
    FILE *f = fopen(path, "w");
    if (!f) return cfd__fail(CFD_ERR_ARG, "cannot open %s: %s", path, strerror(errno));
    const size_t ld = (size_t)jmax + 1;
    const long n = (long)imax * jmax;
    int bad = fprintf(f, "# vtk DataFile Version 3.0\n%s\nASCII\nDATASET STRUCTURED_GRID\n",
                      title ? title : "Droplet Combustion") < 0;
    bad = bad || fprintf(f, "DIMENSIONS %3d %3d 1\nPOINTS %6ld float\n", (int)jmax, (int)imax, n) < 0;
    for (int i = 1; i <= imax && !bad; ++i)
        for (int j = 1; j <= jmax && !bad; ++j) bad = fprintf(f, "%14.4f%14.4f%14.4f\n", x[i], y[j], 0.0) < 0;
    bad = bad || fprintf(f, "POINT_DATA %6ld\nSCALARS VEL_MAGNITUDE float\nLOOKUP_TABLE default\n", n) < 0;
    for (int i = 1; i <= imax && !bad; ++i)
        for (int j = 1; j <= jmax && !bad; ++j) {
            const size_t k = i * ld + j;
            bad = fprintf(f, "%14.4f\n", sqrt(u[k] * u[k] + v[k] * v[k])) < 0;
        }
    const double *sc[2] = {T, p};
    const char *nm[2] = {"T", "P"};
    for (int s = 0; s < 2 && !bad; ++s) {
        bad = fprintf(f, "SCALARS %s float\nLOOKUP_TABLE default\n", nm[s]) < 0;
        for (int i = 1; i <= imax && !bad; ++i)
            for (int j = 1; j <= jmax && !bad; ++j) bad = fprintf(f, "%14.4f\n", sc[s][i * ld + j]) < 0;
    }
    bad = bad || fprintf(f, "\n\nVECTORS Vectors float\n") < 0;
    for (int i = 1; i <= imax && !bad; ++i)
        for (int j = 1; j <= jmax && !bad; ++j) {
            const size_t k = i * ld + j;
            bad = fprintf(f, "%14.4f%14.4f%14.4f\n", u[k], v[k], 0.0) < 0;
        }
    if (fclose(f) != 0) bad = 1;
    return bad ? cfd__fail(CFD_ERR_ARG, "write error on %s", path) : CFD_OK;
}
