/*
 * cfd_fortran.h — the reference's Fortran hot-path subroutines, callable from Fortran
 * unchanged: same names (Fortran external-procedure mangling, lower case with a trailing
 * underscore, as flang and gfortran emit for OpenMp_Multicore/main.f90's CALLs), same
 * argument order and meaning, every argument by reference, arrays on the HOST in
 * Fortran column-major order with the reference's bounds:
 *   u-like  um(1:imax+1, 1:jmax)       v-like  vm(1:imax, 1:jmax+1)
 *   p-like  p, T, C, Z (1:imax, 1:jmax)
 * The residual of solve_U / solve_V / solve_P is the last argument (equations.f90:244,
 * 546, 607), as in the reference; solve_Z / solve_C return none (equations.f90:669, 769).
 *
 * A driver such as OpenMp_Multicore/main.f90 links this library in place of equations.f90
 * and boundary.f90 and keeps its loop (main.f90:116-157) as written. The reference's
 * module state (comum.f90: imax, jmax, the mesh, epsilon1, liga_poros, Re, Da, Cf, dtau,
 * dt, beta, ...) is handed over once by cfd_fortran_setup() before the first call.
 *
 * Each call copies its array arguments to the GPU, runs the per-operator HIP path of
 * include/cfd_legacy.h (the reference's arithmetic, bit for bit) and copies back the
 * arrays the reference writes, including the boundary rewrites of bcUV inside
 * solve_U / solve_V (vm_tau / um_tau). Arrays the reference writes only in part (RU, the
 * n_tau outputs) are copied both ways, so the cells it leaves alone keep the caller's
 * values. The copies dominate: this entry level serves unchanged Fortran callers; a caller
 * that can change its loop keeps the state on the GPU with cfd_physical_step (cfd.h).
 * Errors set cfd_last_error(); a failed solve returns a NaN residual.
 */
#ifndef CFD_AMD_FORTRAN_H
#define CFD_AMD_FORTRAN_H
#include "cfd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces the module state of comum.f90:50-146 set by mesh() (nonsymetric_mesh.f90) and
 * main.f90:105-106: parameters, node coordinates x(1:imax), y(1:jmax) and the porosity
 * fields epsilon1(1:imax,1:jmax), liga_poros(1:imax,1:jmax) (NULL: 1 and 0). Creates (or
 * re-creates) the default context of cfd_legacy.h. Returns CFD_OK or an error code. */
int cfd_fortran_setup(const cfd_params *p, const double *x, const double *y, const double *epsilon1,
                      const double *liga_poros);
/* Releases the default context and the transfer buffers. */
void cfd_fortran_finalize(void);

/* equations.f90:244-301 */
void solve_u_(const double *um, const double *vm, const double *um_n, double *um_tau, double *vm_tau,
              double *um_n_tau, const double *p, double *residual_u);
/* equations.f90:546-603 */
void solve_v_(const double *um, const double *vm, const double *vm_n, double *um_tau, double *vm_tau,
              double *vm_n_tau, const double *p, const double *T, double *residual_v);
/* equations.f90:607-665 */
void solve_p_(const double *p, const double *um_n, const double *vm_n, double *pn, double *residual_p);
/* equations.f90:669-719 */
void solve_z_(const double *um_n, const double *vm_n, const double *Z, double *Z_n_tau, const double *Z_tau);
/* equations.f90:769-819 */
void solve_c_(const double *um_n, const double *vm_n, const double *C, double *C_n_tau, const double *C_tau);

/* equations.f90:2-116, 305-420, 723-765, 823-867 */
void resu_(const double *um, const double *vm, const double *p, double *RU);
void resv_(const double *um, const double *vm, const double *p, const double *T, double *RV);
void resz_(const double *um_n, const double *vm_n, const double *Z, double *RZ);
void resc_(const double *um_n, const double *vm_n, const double *C, double *RC);

/* boundary.f90:1-49, 52-73, 76-116, 119-159 */
void bcuv_(double *um, double *vm);
void bcp_(double *pn);
void bcz_(double *Zt);
void bcc_(double *Cn);

#ifdef __cplusplus
}
#endif
#endif
