// body_geom.hpp -- device geometry of rigid bodies shared by the flow-move
// and the field evaluators.
#pragma once

#include "common.hpp"

namespace vvhd {

// TBody::isPointInContour over the corners (TBody.cpp:238-281): index of the
// nearest segment when p is inside the body, else -1.
__device__ inline int point_invalid(const vvhd_body& B, const vvhd_seg* segs, const double* geo, double px, double py) {
    bool in = B.area <= 0;
    if (!in && (px < geo[1] || py < geo[2] || px > geo[3] || py > geo[4] ||
                (px - B.cofm[0]) * (px - B.cofm[0]) + (py - B.cofm[1]) * (py - B.cofm[1]) > geo[0]))
        return -1;
    const int f = B.first_seg, n = B.nseg;
    for (int i = 0, j = n - 1; i < n; j = i++) {
        const double vix = segs[f + i].cx, viy = segs[f + i].cy, vjx = segs[f + j].cx, vjy = segs[f + j].cy;
        if (((viy < vjy) && (viy < py) && (py <= vjy) && ((vjy - viy) * (px - vix) > (vjx - vix) * (py - viy))) ||
            ((viy > vjy) && (viy > py) && (py >= vjy) && ((vjy - viy) * (px - vix) < (vjx - vix) * (py - viy))))
            in = !in;
    }
    if (!in) return -1;
    int best = -1;
    double bd = DBL_MAX;
    for (int k = 0; k < n; k++) {
        const double dx = segs[f + k].x - px, dy = segs[f + k].y - py;
        const double d2 = dx * dx + dy * dy;
        if (d2 < bd) {
            best = f + k;
            bd = d2;
        }
    }
    return best;
}

}  // namespace vvhd
