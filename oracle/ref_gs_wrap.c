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

/* The same with a handle kept across calls, for timing the reference's
 * gather-scatter as a CPU baseline: setup once, apply many times. */
struct ref_gs_handle {
  struct comm c;
  struct gs_data *g;
};

void *ref_gslib_setup(const long long *ids, unsigned n)
{
  struct ref_gs_handle *h = (struct ref_gs_handle *)malloc(sizeof(*h));
  slong *id = (slong *)malloc(sizeof(slong) * (n ? n : 1));
  for (unsigned i = 0; i < n; ++i) id[i] = (slong)ids[i];
  comm_init(&h->c, 0);
  h->g = gs_setup(id, n, &h->c, 0, gs_pairwise, 0);
  free(id);
  return h;
}

int ref_gslib_apply(void *handle, void *v, int dom, int op)
{
  struct ref_gs_handle *h = (struct ref_gs_handle *)handle;
  const gs_dom d = dom == 0 ? gs_double : (dom == 1 ? gs_float : gs_int);
  const gs_op o = op == 0 ? gs_add : (op == 1 ? gs_mul : (op == 2 ? gs_min : gs_max));
  gs(v, d, o, 0, h->g, 0);
  return 0;
}

void ref_gslib_free(void *handle)
{
  struct ref_gs_handle *h = (struct ref_gs_handle *)handle;
  gs_free(h->g);
  comm_free(&h->c);
  free(h);
}
