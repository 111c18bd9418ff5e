// Hosek-Wilkie (2012) RGB sky-model state: the nine distribution coefficients
// and the radiance scale per channel for one (turbidity, ground albedo, solar
// elevation). Restates the state initialisation of the reference's vendored
// model, src/ArHosekSkyModel.cpp:147-289 and :306-339: each quantity is a
// quintic Bezier in elevation^(1/3) over six control points, blended linearly
// between the two albedo tables and the two integer turbidities that bracket
// the requested turbidity. Computed in double, as the reference does.
#include <cmath>

#include "drk_internal.h"

namespace drk {
namespace {

const double kCoeffs[] = {
#include "../data/hosek_rgb_coeffs.inc"
};
const double kRadiance[] = {
#include "../data/hosek_rgb_radiance.inc"
};
static_assert(sizeof(kCoeffs) / sizeof(double) == 3 * 2 * 10 * 6 * 9, "Hosek coefficient table size");
static_assert(sizeof(kRadiance) / sizeof(double) == 3 * 2 * 10 * 6, "Hosek radiance table size");

// Bernstein weights of degree 5 at x, in the order of the control points.
void bernstein5(double x, double w[6]) {
    const double y = 1.0 - x;
    w[0] = std::pow(y, 5.0);
    w[1] = 5.0 * std::pow(y, 4.0) * x;
    w[2] = 10.0 * std::pow(y, 3.0) * std::pow(x, 2.0);
    w[3] = 10.0 * std::pow(y, 2.0) * std::pow(x, 3.0);
    w[4] = 5.0 * y * std::pow(x, 4.0);
    w[5] = std::pow(x, 5.0);
}

// One blend term: table block for (albedo index, turbidity index), stride =
// number of values per control point (9 coefficients or 1 radiance).
double bezier(const double *block, int stride, int comp, const double w[6]) {
    double s = 0.0;
    for (int k = 0; k < 6; ++k) s += w[k] * block[k * stride + comp];
    return s;
}

}  // namespace

void hosek_state(double turbidity, double albedo, double elevation, double cfg[3][9], double rad[3]) {
    const int ti = (int)turbidity;
    const double tr = turbidity - (double)ti;
    const double x = std::pow(elevation / (M_PI / 2.0), 1.0 / 3.0);
    double w[6];
    bernstein5(x, w);
    for (int ch = 0; ch < 3; ++ch) {
        const double *cbase = kCoeffs + ch * (2 * 10 * 6 * 9);
        const double *rbase = kRadiance + ch * (2 * 10 * 6);
        // (albedo weight, turbidity weight, albedo index, turbidity index)
        struct Term { double wa, wt; int a, t; } terms[4] = {
            {1.0 - albedo, 1.0 - tr, 0, ti - 1},
            {albedo, 1.0 - tr, 1, ti - 1},
            {1.0 - albedo, tr, 0, ti},
            {albedo, tr, 1, ti},
        };
        const int nterms = (ti == 10) ? 2 : 4;
        for (int i = 0; i < 9; ++i) cfg[ch][i] = 0.0;
        rad[ch] = 0.0;
        for (int k = 0; k < nterms; ++k) {
            const Term &T = terms[k];
            const double *cb = cbase + (T.a * 10 + T.t) * 6 * 9;
            const double *rb = rbase + (T.a * 10 + T.t) * 6;
            for (int i = 0; i < 9; ++i) cfg[ch][i] += T.wa * T.wt * bezier(cb, 9, i, w);
            rad[ch] += T.wa * T.wt * bezier(rb, 1, 0, w);
        }
    }
}

}  // namespace drk
