// Complex arithmetic of the plain-double path's series approximation, shared by the
// host (coefficient recurrence) and the device (per-pixel Horner seed).
//
// The engine uses std::complex<double> (mandel_perturbation.cpp:2319-2323,
// 4273-4287), whose rounding is library-specific; this build fixes explicit
// formulas: textbook multiplication and Smith's division (the form of libgcc's
// __divdc3 for finite operands).  Built without FP contraction on both sides.
#pragma once

#include "mx_fe.h"

#define MX_SA_N 30

struct MxC {
    double r, i;
};

MX_HD MxC mxc_add(MxC a, MxC b) { return MxC{a.r + b.r, a.i + b.i}; }

MX_HD MxC mxc_mul(MxC a, MxC b) { return MxC{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }

MX_HD MxC mxc_div(MxC a, MxC b) {
    MxC q;
    if (fabs(b.r) < fabs(b.i)) {
        const double ratio = b.r / b.i;
        const double denom = (b.r * ratio) + b.i;
        q.r = ((a.r * ratio) + a.i) / denom;
        q.i = ((a.i * ratio) - a.r) / denom;
    } else {
        const double ratio = b.i / b.r;
        const double denom = (b.i * ratio) + b.r;
        q.r = ((a.i * ratio) + a.r) / denom;
        q.i = (a.i - (a.r * ratio)) / denom;
    }
    return q;
}
