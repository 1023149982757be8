/* ast_ref.c -- test infrastructure only: a driver that evaluates libast
 * expressions with the reference's own libast (lib/libast.c, compiled from
 * where it lies) the way the reference's ASCII reader does
 * (io/read_ascii.c:233-340: ascii_read_sel / ascii_read_double set every
 * variable with its column's type -- int for %d, double for %lf -- and
 * evaluate; positions, weights and n(z) are AST_DTYPE_DOUBLE trees, the
 * selection an AST_DTYPE_BOOL tree, io/read_ascii.c:656-735).
 *
 * stdin:
 *   ncol                      number of value columns
 *   t_1 ... t_ncol            'd' (double) or 'i' (int) per column
 *   nexpr
 *   nexpr lines "D <expr>" or "B <expr>"  (double or bool tree)
 *   nrow
 *   nrow lines of ncol values
 * stdout: one line per expression, "BUILD_ERR" if the tree does not build,
 * else "OK"; then per row one line with each expression's value (%a for
 * doubles, 0/1 for bools, "E" where the evaluation fails). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdbool.h>

#include "libast.h"

int main(void) {
  int ncol, nexpr;
  if (scanf("%d", &ncol) != 1 || ncol < 1 || ncol > 64) return 1;
  char types[64];
  for (int c = 0; c < ncol; c++)
    if (scanf(" %c", &types[c]) != 1) return 1;
  if (scanf("%d", &nexpr) != 1 || nexpr < 1 || nexpr > 256) return 1;
  ast_t *ast[256];
  char kind[256];
  bool ok[256];
  char line[4096];
  if (!fgets(line, sizeof line, stdin)) return 1;   /* rest of the count line */
  for (int e = 0; e < nexpr; e++) {
    if (!fgets(line, sizeof line, stdin)) return 1;
    line[strcspn(line, "\n")] = 0;
    kind[e] = line[0];
    ast[e] = ast_init();
    if (!ast[e]) return 1;
    ok[e] = ast_build(ast[e], line + 2, kind[e] == 'B' ? AST_DTYPE_BOOL : AST_DTYPE_DOUBLE,
                      true) == 0;
    /* every variable must name a value column */
    for (long i = 0; ok[e] && i < ast[e]->nvar; i++)
      if (ast[e]->vidx[i] > ncol) ok[e] = false;
    puts(ok[e] ? "OK" : "BUILD_ERR");
  }
  long nrow;
  if (scanf("%ld", &nrow) != 1) return 1;
  double dv[64];
  int iv[64];
  for (long r = 0; r < nrow; r++) {
    for (int c = 0; c < ncol; c++) {
      if (types[c] == 'i') {
        if (scanf("%d", &iv[c]) != 1) return 1;
      } else if (scanf("%lf", &dv[c]) != 1) return 1;
    }
    for (int e = 0; e < nexpr; e++) {
      if (e) putchar(' ');
      if (!ok[e]) {
        putchar('-');
        continue;
      }
      bool fail = false;
      for (long i = 0; i < ast[e]->nvar && !fail; i++) {
        const long idx = ast[e]->vidx[i];
        const int c = (int)idx - 1;
        if (types[c] == 'i')
          fail = ast_set_var(ast[e], idx, &iv[c], 0, AST_DTYPE_INT) != 0;
        else
          fail = ast_set_var(ast[e], idx, &dv[c], 0, AST_DTYPE_DOUBLE) != 0;
      }
      if (kind[e] == 'B') {
        bool b = false;
        if (!fail) fail = ast_eval(ast[e], &b) != 0;
        if (fail) putchar('E');
        else putchar(b ? '1' : '0');
      } else {
        double d = 0;
        if (!fail) fail = ast_eval(ast[e], &d) != 0;
        if (fail) putchar('E');
        else printf("%a", d);
      }
      /* an evaluation error sticks to the tree: rebuild it for later rows */
      if (fail) {
        ast_destroy(ast[e]);
        ok[e] = false;
      }
    }
    putchar('\n');
  }
  return 0;
}
