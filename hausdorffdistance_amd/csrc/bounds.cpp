// Bounds on the directed HD from the two MBRs alone (src/geoms/hd_bounds.h:
// 111-160, over Mbr::GetMinDist2 / GetMaxDist2 of src/geoms/mbr.h:193-266),
// as the reference's hybrid writes them into its stats ("HDLowerBound",
// "HDUpperBound", hausdorff_distance_hybrid.h:95-106). Host code, FP32.
//
// Each MBR is replaced by its faces, taken as boxes with one axis collapsed:
// 2D bottom, right, top, left; 3D y=lower, y=upper, x=lower, x=upper,
// z=lower, z=upper. The order matters: the upper bound's inner loop stops at
// the first face pair whose farthest distance does not exceed the running
// bound, keeping the smallest distance seen before it.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "hd_internal.h"

namespace hd {
namespace {

struct Box {
  float lo[3], hi[3];
};

int Faces(int nd, const float* lo, const float* hi, Box* faces) {
  static const int kAxis2[4] = {1, 0, 1, 0}, kUpper2[4] = {0, 1, 1, 0};
  static const int kAxis3[6] = {1, 1, 0, 0, 2, 2}, kUpper3[6] = {0, 1, 0, 1, 0, 1};
  const int nf = nd == 2 ? 4 : 6;
  for (int f = 0; f < nf; ++f) {
    const int axis = nd == 2 ? kAxis2[f] : kAxis3[f];
    const bool upper = (nd == 2 ? kUpper2[f] : kUpper3[f]) != 0;
    for (int d = 0; d < nd; ++d) {
      faces[f].lo[d] = lo[d];
      faces[f].hi[d] = hi[d];
    }
    faces[f].lo[axis] = faces[f].hi[axis] = upper ? hi[axis] : lo[axis];
  }
  return nf;
}

// Squared distance between the nearest points of two boxes.
float NearestDist2(int nd, const Box& a, const float* olo, const float* ohi) {
  float s = 0.0f;
  for (int d = 0; d < nd; ++d) {
    float gap = 0.0f;
    if (ohi[d] < a.lo[d])
      gap = a.lo[d] - ohi[d];
    else if (a.hi[d] < olo[d])
      gap = olo[d] - a.hi[d];
    s += gap * gap;
  }
  return s;
}

// Squared distance between the farthest points of two boxes.
float FarthestDist2(int nd, const Box& a, const Box& b) {
  float s = 0.0f;
  for (int d = 0; d < nd; ++d) {
    const float span = std::max(std::fabs(a.lo[d] - b.hi[d]), std::fabs(a.hi[d] - b.lo[d]));
    s += span * span;
  }
  return s;
}

}  // namespace

void HdBounds2(int nd, const float* lo_a, const float* hi_a, const float* lo_b,
               const float* hi_b, float* lb2, float* ub2) {
  Box fa[6], fb[6];
  const int nf = Faces(nd, lo_a, hi_a, fa);
  Faces(nd, lo_b, hi_b, fb);
  float lower = 0.0f;
  for (int i = 0; i < nf; ++i) lower = std::max(lower, NearestDist2(nd, fa[i], lo_b, hi_b));
  float upper = 0.0f;
  for (int i = 0; i < nf; ++i) {
    float best = FLT_MAX;
    for (int j = 0; j < nf; ++j) {
      const float d = FarthestDist2(nd, fa[i], fb[j]);
      if (d <= upper) break;
      best = std::min(best, d);
    }
    if (best != FLT_MAX) upper = std::max(upper, best);
  }
  *lb2 = lower;
  *ub2 = upper;
}

}  // namespace hd

using namespace hd;

extern "C" int hd_mbr_hd_bounds(const float* mbr_a, const float* mbr_b, int n_dims,
                                float* lower, float* upper) {
  ClearError();
  if (n_dims != 2 && n_dims != 3)
    return Fail(HD_EINVAL, "n_dims must be 2 or 3, got %d", n_dims);
  if (!mbr_a || !mbr_b || !lower || !upper) return Fail(HD_EINVAL, "null argument");
  float lo_a[3], hi_a[3], lo_b[3], hi_b[3], lb2, ub2;
  NonEmptyMbr(n_dims, mbr_a, lo_a, hi_a);
  NonEmptyMbr(n_dims, mbr_b, lo_b, hi_b);
  HdBounds2(n_dims, lo_a, hi_a, lo_b, hi_b, &lb2, &ub2);
  *lower = std::sqrt(lb2);
  *upper = std::sqrt(ub2);
  return HD_OK;
}
