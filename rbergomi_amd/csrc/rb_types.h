// rb_types.h -- work tables shared by the host planner and the kernels.
//
// The reference walks a flat option index I = h*op_size + t and recomputes state when a
// "trigger" fires (rBergLib/ParamTot.h:32-50).  The planner instead deduplicates:
//   unit   = one distinct H: one Volterra convolution per path       (updateWtilde)
//   group  = one distinct (H, eta, T[, rho for plain MC]): one pass of the variance
//            process and its integrals per path                       (scaleWtilde/updateV/intVdt/intRootVdW)
//   option = one output I: one conditioned Black-Scholes payoff per path (BS_call_price)
#pragma once
#include <stdint.h>

namespace rb {

struct Unit {
    double H;
    double a;   // W1 weight of the first-cell term: rho_H / sqrt(2H) = 1/(H+1/2)
    double b;   // W1perp weight: sqrt(1 - rho_H^2) / sqrt(2H)
    double sc;  // sqrt(2H) * N^-H
};

struct Group {
    int unit;
    int pad;
    double xi;
    double eta;
    double TH;   // T^H
    double dt;   // T/N
    double sdt;  // sqrt(dt)
    double rho;  // plain MC only
    double crho; // sqrt(1 - rho^2), plain MC only
};

struct Option {
    int group;
    int pad;
    double rho;
    double K;
    double lnK;
    double T;
};

enum Method { METHOD_RT = 0, METHOD_MC = 1 };

}  // namespace rb
