/*
 * cf_tables.h -- host-side builders of the column-independent wall-normal
 * tables of the general (order 2 or 4, optional scalar) path: ghost
 * closures, wall-reduced banded operators and their factors.  Plain C++,
 * no device code.
 *
 * Levels are device levels: cells 0..nz-1, faces 0..nz-2 interior with the
 * wall faces at -1 and nz-1.  dzl points at level 0 of a dz table valid on
 * levels -3..nz+2 (mirrored ghosts, ghost.c:31-54).
 */
#ifndef CF_TABLES_H
#define CF_TABLES_H

#include <vector>

/* The ghost levels of one wall as affine functions of the three unknowns
 * next to it: ghost m (m = 0 nearest the wall) = sum_t G[m][t] x_t + B[m] bc,
 * x_t the t-th unknown counted from the wall.  Faces: ghost 0 is the wall
 * face itself. */
struct Clos1 {
    double G[3][3];
    double B[3];
};

namespace cft {

enum { CELL = 0, FACE = 1 };
enum { DIRICHLET = 0, NEUMANN = 1 };

/* one banded row: c[t] multiplies the unknown at level (row level) - 3 + t */
struct Row7 {
    double c[7];
    double r;
};

/* ghost closure of one wall (top = 0 bottom, 1 top) for a cell or face
 * field of the given order; cells take the Dirichlet or Neumann rows of
 * diffuseu.c:617-802 / pressure.c:446-726, faces those of
 * diffusew.c:625-766.  Returns 0, or -1 if the ghost rows are singular. */
int closure(int fam, int order, int top, int bctype, const double *dzl, int nz, Clos1 &out);

/* interior rows of (I + nudt d2/dz2) * dvol scaled as solve_helmholtz1d
 * assembles them (diffuseu.c:192-266, diffusew.c:196-270) when with_dvol,
 * or of the Poisson operator dt * d2/dz2 * dxy dz (pressure.c:318-339,
 * :382-420) when not; dxy = 1 for the 1-D solves, dx dy for Poisson. */
void op_rows(int fam, int order, const double *dzl, int nz, double nudt, double dxy, bool with_dvol,
             std::vector<Row7> &rows);

/* substitute the ghost unknowns of every row by their closures (bc = 0),
 * leaving a banded system over the interior unknowns only */
void eliminate(int fam, int nz, const Clos1 &bot, const Clos1 &top, std::vector<Row7> &rows);

/* seven-band LU without pivoting in the recurrence form of heptdag.c:5-61:
 * tab = [a | al | be | iden | p | q | r], 7 n doubles.  Returns -1 on a
 * zero pivot. */
int hept_factor(const std::vector<Row7> &rows, std::vector<double> &tab);
void hept_solve(const std::vector<double> &tab, int n, const std::vector<double> &rhs, std::vector<double> &x);

/* the explicit 1-D operator q + nudt d2q/dz2 of helmholtz1d_operator_u/w
 * (diffuseu.c:107-170, diffusew.c:107-172) at the interior levels of a
 * column given with its ghosts: col[k + 3] = level k, k in [-3, n + 3) */
void helm_apply(int fam, int order, const double *dzl, int nz, double nudt, const double *col, double *out);

}   /* namespace cft */

#endif
