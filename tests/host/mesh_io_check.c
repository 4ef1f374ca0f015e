/* Host-only exercise of the mesh, porous-flag and state-file routines of cfd_mesh_io.c,
 * built with AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py
 * (no GPU code is compiled or run). Every routine runs on a few grid shapes, including the
 * wrapped-record reader on a file written with records split over lines, and the round trips
 * must return what was written. Exit status 0 on success. */
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cfd.h"

int cfd__fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
    return code;
}

#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__);       \
            return 1;                                                            \
        }                                                                        \
    } while (0)

static double *arr(size_t n, double base)
{
    double *a = (double *)malloc(n * sizeof(double));
    for (size_t k = 0; k < n; ++k) a[k] = base + 1e-3 * (double)(k % 977) - 0.25 * (double)((k * 7) % 5);
    return a;
}

static int one_shape(int im, int jm, const char *dir)
{
    const size_t np = (size_t)(im + 1) * (jm + 1), nu = (size_t)(im + 2) * (jm + 1), nv = (size_t)(im + 1) * (jm + 2);
    double *x = (double *)calloc(im + 1, sizeof(double)), *y = (double *)calloc(jm + 1, sizeof(double));
    CHECK(cfd_mesh_reference(im, jm, 12.0, 5.0, 0.6, 1.6, 1.8, x, y) == 0);
    for (int i = 2; i <= im; ++i) CHECK(x[i] > x[i - 1]);
    double *eps = (double *)calloc(np, sizeof(double)), *lp = (double *)calloc(np, sizeof(double));
    int32_t *fl = (int32_t *)calloc(np, sizeof(int32_t));
    CHECK(cfd_porous_flags_reference(im, jm, x, y, x[im / 2], y[jm / 3], 0.6, eps, lp, fl) == 0);
    double *um = arr(nu, 0.5), *vm = arr(nv, -0.2), *p = arr(np, 1.0), *T = arr(np, 0.3), *C = arr(np, 0.7);
    double *u = (double *)calloc(np, sizeof(double)), *v = (double *)calloc(np, sizeof(double));
    CHECK(cfd_nodal_velocity(im, jm, um, vm, u, v) == 0);

    char path[512];
    snprintf(path, sizeof path, "%s/state.dat", dir);
    CHECK(cfd_write_state_table(path, im, jm, x, y, u, v, p, T, C) == 0);
    double *r[7];
    for (int q = 0; q < 7; ++q) r[q] = (double *)calloc(np > (size_t)(im + 1) ? np : (size_t)(im + 1), sizeof(double));
    CHECK(cfd_read_state_table(path, im, jm, r[0], r[1], r[2], r[3], r[4], r[5], r[6]) == 0);
    CHECK(cfd_read_state_table(path, im, jm, NULL, NULL, r[2], NULL, NULL, NULL, r[6]) == 0);

    /* the same records wrapped over lines (three values, then four) */
    FILE *f = fopen(path, "r");
    snprintf(path, sizeof path, "%s/state_wrapped.dat", dir);
    FILE *g = fopen(path, "w");
    CHECK(f && g);
    double val[7];
    while (fscanf(f, "%lf %lf %lf %lf %lf %lf %lf", &val[0], &val[1], &val[2], &val[3], &val[4], &val[5], &val[6]) == 7)
        fprintf(g, " %.17g %.17g %.17g\n %.17g %.17g %.17g %.17g\n", val[0], val[1], val[2], val[3], val[4], val[5], val[6]);
    fclose(f);
    fclose(g);
    CHECK(cfd_read_state_table(path, im, jm, r[0], r[1], r[2], r[3], r[4], r[5], r[6]) == 0);

    CHECK(cfd_write_restart(dir, im, jm, um, vm, p, T, C) == 0);
    double *um2 = (double *)calloc(nu, sizeof(double)), *vm2 = (double *)calloc(nv, sizeof(double));
    double *p2 = (double *)calloc(np, sizeof(double)), *T2 = (double *)calloc(np, sizeof(double));
    double *C2 = (double *)calloc(np, sizeof(double));
    CHECK(cfd_read_restart(dir, im, jm, CFD_RESTART_ORDER_WRITTEN, um2, vm2, p2, T2, C2) == 0);
    CHECK(cfd_read_restart(dir, im, jm, CFD_RESTART_ORDER_REFERENCE, um2, vm2, p2, T2, C2) == 0);

    snprintf(path, sizeof path, "%s/out.vtk", dir);
    CHECK(cfd_write_vtk(path, NULL, im, jm, x, y, u, v, p, T) == 0);
    CHECK(cfd_write_vtk(path, "title", im, jm, x, y, u, v, p, T) == 0);

    free(x); free(y); free(eps); free(lp); free(fl); free(um); free(vm); free(p); free(T); free(C);
    free(u); free(v); free(um2); free(vm2); free(p2); free(T2); free(C2);
    for (int q = 0; q < 7; ++q) free(r[q]);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const int shapes[][2] = {{51, 84}, {24, 20}, {5, 7}, {64, 40}};
    for (size_t s = 0; s < sizeof shapes / sizeof shapes[0]; ++s)
        if (one_shape(shapes[s][0], shapes[s][1], argv[1])) {
            fprintf(stderr, "shape %dx%d\n", shapes[s][0], shapes[s][1]);
            return 1;
        }
    /* a y split outside the grid is an error, not an out-of-bounds write */
    double x[65], y[4];
    if (cfd_mesh_reference(64, 3, 12.0, 5.0, 0.6, 1.6, 1.8, x, y) == 0) return 1;
    puts("mesh_io_check ok");
    return 0;
}
