/*
 * Test infrastructure (oracle side): plain-pointer wrappers around
 * reference functions whose arguments are reference structs, so tests can
 * call them through ctypes.  Compiled against the reference headers where
 * they lie and linked with oracle/_ref/libherc_ref.so.
 */
#include <stdlib.h>
#include <string.h>

#include "../config.h"
#include "../globals.h"
#include "hercules_index.h"
#include "hercules_node.h"
#include "hercules_node_split.h"

/* calculate_node_min_distance (hercules_node.c:879-924) on a node whose
   segmentation and sketches are given as arrays; the function does not use
   its index argument */
float ref_node_min_distance(const float *query, int nseg, const short *seg_end, const float *ind)
{
  struct hercules_node node;
  memset(&node, 0, sizeof(node));
  short *pts = malloc(sizeof(short) * nseg);
  struct segment_sketch *sk = calloc(nseg, sizeof(struct segment_sketch));
  float *vals = malloc(sizeof(float) * 4 * nseg);
  memcpy(pts, seg_end, sizeof(short) * nseg);
  memcpy(vals, ind, sizeof(float) * 4 * nseg);
  for (int i = 0; i < nseg; ++i) {
    sk[i].indicators = vals + 4 * i;
    sk[i].num_indicators = 4;
  }
  node.node_points = pts;
  node.num_node_points = (short)nseg;
  node.node_segment_sketches = sk;
  float r = calculate_node_min_distance(NULL, &node, (ts_type *)query);
  free(pts);
  free(sk);
  free(vals);
  return r;
}
