/*
 * mdelast_setfl.h - native SETFL (LAMMPS eam/alloy) loader and node-derivative
 * builder of libmdelast_hip.so.
 *
 * Replaces the reading half of get_EAM_alloy_file (get_potential.py:58-147) and
 * get_derivatives (get_potential.py:49-55, SciPy UnivariateSpline(k=5, s=0)):
 * the file is parsed, converted to CGS and given first and second derivatives at
 * its nodes on the host, without Python or SciPy. The node arrays feed
 * mdl_table_create (mdelast.h) as (f, f', f'').
 *
 * Host-only: no call here touches a GPU. Status codes and mdl_last_error() as in
 * mdelast.h.
 */
#ifndef MDELAST_SETFL_H
#define MDELAST_SETFL_H

#include "mdelast.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdl_setfl mdl_setfl;

/* Node arrays of one SETFL file, mdl_setfl_nodes(which = ...) */
enum { MDL_SETFL_F = 0, MDL_SETFL_RHO = 1, MDL_SETFL_PHI = 2 };

/* Parse a SETFL file (get_potential.py:58-147): three comment lines, "Nel names",
 * "Nrho drho Nr dr rcut", then per element "Z mass a lattice", Nrho values of
 * F(rho) and Nr values of rho(r), then the Nel(Nel+1)/2 blocks of r*phi(r) in the
 * order 11, 21, 22, 31, ... Values are read as whitespace-separated tokens; a
 * block must end at a line end. CGS: F x eV, rho / A^3, phi = (r phi) / r x eV x A
 * with phi[0] = phi[1]; grids 0 .. d (N - 1) as numpy.linspace forms them.
 * with_derivatives = 0 skips the spline derivatives (the d1, d2 arrays are then 0).
 * MDL_ERR_ARG on a missing or malformed file, with the reason in mdl_last_error(). */
mdl_status mdl_setfl_read(const char *path, int32_t with_derivatives, mdl_setfl **out);
mdl_status mdl_setfl_destroy(mdl_setfl *s);

/* Grid of the file, CGS. rcut = r_grid[N_r - 1] and rhomax = rho_grid[N_rho - 1]
 * are what the reference hands to the driver (get_potential.py:94-95, :117, :134);
 * rcut_header is the header's own cut-off (in cm). Any pointer may be NULL. */
mdl_status mdl_setfl_grid(const mdl_setfl *s, int32_t *n_elements, int64_t *n_rho, double *drho,
                          int64_t *n_r, double *dr, double *rcut, double *rhomax,
                          double *rcut_header);

/* Element e (file order): symbol (NUL-terminated, truncated to name_len), atomic
 * number, atomic mass (g/mol), lattice constant (Angstrom), lattice name. */
mdl_status mdl_setfl_element(const mdl_setfl *s, int32_t e, char *name, int32_t name_len,
                             double *Z, double *mass, double *a_lattice, char *lattice,
                             int32_t lattice_len);

/* Node values and spline derivatives, CGS. which = MDL_SETFL_F (index = element,
 * N_rho values), MDL_SETFL_RHO (element, N_r values) or MDL_SETFL_PHI (pair slot
 * j(j+1)/2 + i, i <= j, N_r values). f, d1, d2 may be NULL. */
mdl_status mdl_setfl_nodes(const mdl_setfl *s, int32_t which, int32_t index, double *f, double *d1,
                           double *d2);

/* First and second derivatives at the nodes x[0..n-1] (strictly increasing, n >= 6)
 * of the degree-5 spline interpolating y there, with interior knots x[3..n-4] (the
 * spline FITPACK builds for UnivariateSpline(k=5, s=0), get_potential.py:49-55). */
mdl_status mdl_spline5_node_derivatives(int64_t n, const double *x, const double *y, double *d1,
                                        double *d2);

#ifdef __cplusplus
}
#endif

#endif /* MDELAST_SETFL_H */
