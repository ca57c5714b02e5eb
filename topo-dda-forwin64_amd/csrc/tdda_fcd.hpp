// tdda_fcd.hpp -- the filtered coupled-dipole (FCD) interaction tensor of
// the reference (AProductCore::FCD_inter, AProductCore.cpp:625-695), for the
// host library and its device kernels.
//
// FCD needs the sine and cosine integrals Si and Ci.  The reference reads
// them from tables of x = 0.1, 0.2, ... (SiCi/SiCi.py writes them with
// scipy.special.sici) and interpolates linearly between the two nodes around
// x (SiCi.cpp:27-46); the tables themselves are not in its repository.
// Here the node values are computed on the fly and interpolated the same
// way, so the tensor follows the reference's, table error included.
//
// Si / Ci at a node: power series for x <= 4; above, the continued
// fraction of E1(ix) = -Ci(x) + i (Si(x) - pi/2) evaluated with the
// modified Lentz method (the published "cisi" scheme of Numerical Recipes,
// written here from its mathematics).  Absolute error about 1e-16.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace tdda {

constexpr double kSiCiStep = 0.1;  // node spacing of the reference's tables (SiCi.py:10)

__host__ __device__ inline void sici_exact(double x, double& si, double& ci)
{
    const double euler = 0.57721566490153286061;
    if (x <= 4.0) {
        const double x2 = x * x;
        double term = x, s = x;
        for (int k = 1; k < 60; k++) {
            term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
            const double add = term / (2.0 * k + 1.0);
            s += add;
            if (fabs(add) < 1e-17 * fabs(s)) break;
        }
        double t = 1.0, c = 0.0;
        for (int k = 1; k < 60; k++) {
            t *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
            const double add = t / (2.0 * k);
            c += add;
            if (fabs(add) < 1e-17 * fabs(c)) break;
        }
        si = s;
        ci = euler + log(x) + c;
        return;
    }
    // E1(z) = e^-z / (z + 1 - 1 / (z + 3 - 4 / (z + 5 - ...))), z = i x
    double br = 1.0, bi = x;                          // b = z + 1
    double cr = 1e300, ci2 = 0.0;                     // c = 1 / tiny
    double q = br * br + bi * bi;
    double dr = br / q, di = -bi / q;                 // d = 1 / b
    double hr = dr, hi = di;                          // h = d
    for (int i = 1; i < 1000; i++) {
        const double a = -(double)i * i;
        br += 2.0;
        // d = 1 / (a d + b)
        double tr = a * dr + br, ti = a * di + bi;
        q = tr * tr + ti * ti;
        dr = tr / q;
        di = -ti / q;
        // c = b + a / c
        q = cr * cr + ci2 * ci2;
        const double ir = cr / q, ii = -ci2 / q;
        cr = br + a * ir;
        ci2 = bi + a * ii;
        // del = c d ; h *= del
        const double er = cr * dr - ci2 * di, ei = cr * di + ci2 * dr;
        const double nhr = hr * er - hi * ei, nhi = hr * ei + hi * er;
        hr = nhr;
        hi = nhi;
        if (fabs(er - 1.0) + fabs(ei) < 1e-16) break;
    }
    // E1 = (cos x - i sin x) h
    const double cs = cos(x), sn = sin(x);
    const double e1r = cs * hr + sn * hi, e1i = cs * hi - sn * hr;
    si = 1.5707963267948966192 + e1i;
    ci = -e1r;
}

// The reference's table lookup (SiCi::get_Si / get_Ci, SiCi.cpp:27-46):
// linear interpolation between the nodes floor(x / 0.1) 0.1 and
// ceil(x / 0.1) 0.1 (on a node, the node's value, where the reference
// divides 0 by 0).  Needs x >= 0.1, the tables' first node.
__host__ __device__ inline void sici_table(double x, double& si, double& ci)
{
    const double a1 = floor(x / kSiCiStep) * kSiCiStep, a2 = ceil(x / kSiCiStep) * kSiCiStep;
    double s1, c1, s2, c2;
    if (a1 == a2) {  // on a node (the reference's 0 / 0)
        sici_exact(a1, si, ci);
        return;
    }
    sici_exact(a1, s1, c1);
    sici_exact(a2, s2, c2);
    si = s1 + ((s2 - s1) / (a2 - a1)) * (x - a1);
    ci = c1 + ((c2 - c1) / (a2 - a1)) * (x - a1);
}

// FCD tensor at offset (x, y, z) as [xx, xy, xz, yy, yz, zz] (re, im)
// pairs; zero at r = 0 (AProductCore.cpp:625-695, with Kf = pi / d).
__host__ __device__ inline void fcd_inter(double K, double d, double x, double y, double z, double* out12)
{
    const double xs = x * x, ys = y * y, zs = z * z;
    const double rn = sqrt(xs + ys + zs);
    if (rn == 0.0) {
        for (int l = 0; l < 12; l++) out12[l] = 0.0;
        return;
    }
    const double pi = 3.14159265358979323846;
    const double r2 = rn * rn, r3 = rn * rn * rn;
    const double Kf = pi / d;
    const double up = (Kf + K) * rn, dn = (Kf - K) * rn;
    double Sz, Cz, Sf, Cf;
    sici_table(up, Sz, Cz);
    sici_table(dn, Sf, Cf);
    const double Cz1 = cos(up) / rn, Cf1 = cos(dn) / rn;
    const double Sz1 = sin(up) / rn, Sf1 = sin(dn) / rn;
    const double Cz2 = (-up * sin(up) - cos(up)) / r2, Cf2 = (-dn * sin(dn) - cos(dn)) / r2;
    const double Sz2 = (up * cos(up) - sin(up)) / r2, Sf2 = (dn * cos(dn) - sin(dn)) / r2;
    const double kr = K * rn;
    const double sikr = sin(kr), cskr = cos(kr);
    const double sikfr = sin(Kf * rn), cskfr = cos(Kf * rn);
    // CC = pi i + Cf - Cz
    const double CCr = Cf - Cz, CCi = pi;
    const double C1C1 = Cf1 - Cz1, C2C2 = Cf2 - Cz2;
    const double SS = Sz + Sf, S1S1 = Sz1 + Sf1, S2S2 = Sz2 + Sf2;
    // gf = (sin kr CC + cos kr SS) / (pi r)
    const double gfr = (sikr * CCr + cskr * SS) / (pi * rn), gfi = (sikr * CCi) / (pi * rn);
    // const1 = (kr cos kr - sin kr) CC - (kr sin kr + cos kr) SS + r sin kr C1C1 + r cos kr S1S1
    const double a1 = kr * cskr - sikr;
    const double k1r = a1 * CCr - (kr * sikr + cskr) * SS + rn * sikr * C1C1 + rn * cskr * S1S1;
    const double k1i = a1 * CCi;
    // const2 = -K kr sin kr CC - K kr cos kr SS + 2 kr cos kr C1C1 - 2 kr sin kr S1S1
    //          + r sin kr C2C2 + r cos kr S2S2
    const double b1 = -K * kr * sikr;
    const double k2r = b1 * CCr - (K * kr * cskr) * SS + (2 * kr * cskr) * C1C1 - (2 * kr * sikr) * S1S1 +
                       rn * sikr * C2C2 + rn * cskr * S2S2;
    const double k2i = b1 * CCi;
    const double g1r = k1r / (pi * r2), g1i = k1i / (pi * r2);
    const double g2r = (rn * k2r - 2.0 * k1r) / (pi * r3), g2i = (rn * k2i - 2.0 * k1i) / (pi * r3);
    const double hr = (sikfr - Kf * rn * cskfr) / (2.0 * pi * pi * r3);
    const double t1r = K * K * gfr + g1r / rn + (4.0 * pi / 3.0) * hr, t1i = K * K * gfi + g1i / rn;
    const double t2r = (g2r - g1r / rn) / r2, t2i = (g2i - g1i / rn) / r2;
    const double w[6] = {xs, x * y, z * x, ys, y * z, zs};
    const bool diag[6] = {true, false, false, true, false, true};
    for (int l = 0; l < 6; l++) {
        out12[2 * l] = -((diag[l] ? t1r : 0.0) + w[l] * t2r);
        out12[2 * l + 1] = -((diag[l] ? t1i : 0.0) + w[l] * t2i);
    }
}

}  // namespace tdda
