/*
 * ref_gs_wrap.c -- a thin C entry point over the reference's vendored gslib
 * (3rd_party/gslib/src/gs.c, built serially with MPI undefined) so that
 * ctypes can call gs_setup + gs on an id list.  TEST INFRASTRUCTURE ONLY:
 * it is linked into oracle/_ref/libnrsref.so, which exists only in the build
 * container and pins the oracle's gather-scatter against gslib's own.
 */
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include "c99.h"
#include "name.h"
#include "fail.h"
#include "types.h"
#include "comm.h"
#include "mem.h"
#include "gs_defs.h"
#include "gs.h"

/* op: 0 add, 1 mul, 2 min, 3 max ; dom: 0 double, 1 float, 2 int */
int ref_gslib_gs(const long long *ids, unsigned n, void *v, int dom, int op)
{
  struct comm c;
  comm_init(&c, 0);
  slong *id = (slong *)malloc(sizeof(slong) * (n ? n : 1));
  for (unsigned i = 0; i < n; ++i) id[i] = (slong)ids[i];
  struct gs_data *g = gs_setup(id, n, &c, 0, gs_pairwise, 0);
  const gs_dom d = dom == 0 ? gs_double : (dom == 1 ? gs_float : gs_int);
  const gs_op o = op == 0 ? gs_add : (op == 1 ? gs_mul : (op == 2 ? gs_min : gs_max));
  gs(v, d, o, 0, g, 0);
  gs_free(g);
  comm_free(&c);
  free(id);
  return 0;
}
