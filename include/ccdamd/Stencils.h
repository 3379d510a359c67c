// Stencil keys of the swept-CCD interface (reference: src/Stencils.h).  Same fields, same canonical
// order, same lexicographic operator< -- so std::set<VertexFaceStencil> iterates in the reference's
// order.  `isnew` is carried but, as in the reference, is not part of the key.
#pragma once

#include <algorithm>
#include <utility>

struct VertexFaceStencil {
    int p;
    int q0, q1, q2;
    bool isnew;

    // Stencils.h:13-57: the face's vertices in ascending order.
    VertexFaceStencil(int p_, int q0_, int q1_, int q2_) : p(p_), isnew(true)
    {
        int a = q0_, b = q1_, c = q2_;
        if (a > b) std::swap(a, b);
        if (b > c) std::swap(b, c);
        if (a > b) std::swap(a, b);
        q0 = a;
        q1 = b;
        q2 = c;
    }

    bool operator<(const VertexFaceStencil &o) const
    {
        if (p != o.p) return p < o.p;
        if (q0 != o.q0) return q0 < o.q0;
        if (q1 != o.q1) return q1 < o.q1;
        return q2 < o.q2;
    }
};

struct EdgeEdgeStencil {
    int p0, p1;
    int q0, q1;
    bool isnew;

    // Stencils.h:86-117: each edge ascending, then the edge with the smaller first vertex first.
    EdgeEdgeStencil(int p0_, int p1_, int q0_, int q1_) : isnew(true)
    {
        p0 = std::min(p0_, p1_);
        p1 = std::max(p0_, p1_);
        q0 = std::min(q0_, q1_);
        q1 = std::max(q0_, q1_);
        if (p0 > q0) {
            std::swap(p0, q0);
            std::swap(p1, q1);
        }
    }

    bool operator<(const EdgeEdgeStencil &o) const
    {
        if (p0 != o.p0) return p0 < o.p0;
        if (p1 != o.p1) return p1 < o.p1;
        if (q0 != o.q0) return q0 < o.q0;
        return q1 < o.q1;
    }
};
