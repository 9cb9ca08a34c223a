// Exchange-correlation functionals on the device, unpolarised, FP64: Becke 88 and PBE exchange, LYP and PBE correlation
// (the latter on the "modified" PW92 parameters), Slater exchange and VWN correlation (both fits), and B3LYP's XC part
// (lda_x.c:121-140, lda_c_vwn.c:42-63,103-160,178-236, work_lda.i:44-86, hyb_gga_xc_b3lyp.c:50-56, mix_func.c:28-107).  The thresholds and variable choices are those of libxc 1.2.0 as vendored
// by pDynamo3 (pMolecule/QCModel/__extensions__/cxc/work_gga_x.i:37-183, work_gga_c.i:25-245, gga_x_b88.c:48,109-148,
// gga_x_pbe.c:48,64,128-162, gga_c_lyp.c:40,65-223, gga_c_pbe.c:40,75,80-264, lda_c_pw.c:60-236, util.h:67-84).  The
// cube roots are pow(x, 1/3), as in that build (cxc/xc_config.h falls back to pow because config.h is not included).
#pragma once

#include <hip/hip_runtime.h>

namespace xq {

constexpr double kXMinDens = 5.0e-13;
constexpr double kXMinGrad = 5.0e-13;
constexpr double kXFactorC = 0.9305257363491000250020102180716672510262;   // 3/8 (3/pi)^(1/3) 4^(2/3)
constexpr double kX2S = 0.1282782438530421943003109254455883701296;        // 1/(2 (6 pi^2)^(1/3))
constexpr double kCbrt2 = 1.259921049894873164767210607278228350570;

__device__ __forceinline__ double xq_cbrt(double x) { return pow(x, 1.0 / 3.0); }

// Exchange enhancement factors F(x) and dF/dx.
__device__ __forceinline__ void b88_enhancement(double x, double &F, double &dF) {
    const double beta = 0.0042, gamma = 6.0;
    const double ash = asinh(x);
    const double f1 = beta / kXFactorC * x * x, f2 = 1.0 + gamma * beta * x * ash;
    const double df1 = 2.0 * beta / kXFactorC * x, df2 = gamma * beta * (ash + x / sqrt(1.0 + x * x));
    F = 1.0 + f1 / f2;
    dF = (df1 * f2 - f1 * df2) / (f2 * f2);
}

__device__ __forceinline__ void pbe_x_enhancement(double x, double &F, double &dF) {
    const double kappa = 0.8040, mu = 0.2195149727645171;
    const double s = kX2S * x, f0 = kappa + mu * s * s;
    F = 1.0 + kappa * (1.0 - kappa / f0);
    dF = kX2S * kappa * kappa * (2.0 * mu * s) / (f0 * f0);
}

// Exchange: eps (per particle), vrho, vsigma.
__device__ __forceinline__ void gga_exchange(int which, double rho, double sigma, double &eps, double &vr, double &vs) {
    eps = 0.0; vr = 0.0; vs = 0.0;
    if (rho < kXMinDens) return;
    const double ds = 0.5 * rho;
    double gdm = sqrt(sigma) * 0.5;
    if (gdm < kXMinGrad) gdm = kXMinGrad;
    const double rl = pow(ds, 4.0 / 3.0), x = gdm / rl;
    double F, dF;
    if (which == 0) b88_enhancement(x, F, dF);
    else pbe_x_enhancement(x, F, dF);
    eps = (2.0 * (-kXFactorC) * rl * F) / rho;
    vr = (-kXFactorC) * (rl / ds) * ((4.0 / 3.0) * F - (4.0 / 3.0) * dF * x);
    if (gdm > kXMinGrad) vs = 2.0 * (-kXFactorC) * rl * (dF * x / (2.0 * sigma));
}

// LYP at zeta = 0 in the variables rs, xt and xs = 2^(1/3) xt (Miehlich et al. 1989 form).
__device__ __forceinline__ void lyp(double rs, double xt, double xs, double &f, double &frs, double &fxt, double &fxs) {
    const double A = 0.04918, B = 0.132;
    const double crs = xq_cbrt(4.0 * M_PI / 3.0), Cf = 3.0 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0) / 10.0;
    const double c = 0.2533 * crs, d = 0.349 * crs;
    const double xt2 = xt * xt, xs2 = xs * xs;
    const double opd = 1.0 / (1.0 + d * rs), om = B * exp(-c * rs) * opd, de = (c + d * opd) * rs;
    const double a6 = 1.0 / pow(2.0, 8.0 / 3.0), a4 = a6 / 4.0, a5 = a4 / 18.0;
    const double t1 = -1.0 / (1.0 + d * rs);
    const double t2 = -xt2 * ((47.0 - 7.0 * de) / 72.0 - 2.0 / 3.0);
    const double t3 = -Cf;
    const double t4 = a4 * (2.5 - de / 18.0) * (2.0 * xs2);
    const double t5 = a5 * (de - 11.0) * (2.0 * xs2);
    const double t6 = -a6 * (2.0 / 3.0 * (2.0 * xs2) - xs2 / 4.0 - xs2 / 4.0);
    const double tt = t2 + t3 + t4 + t5 + t6;
    f = A * (t1 + om * tt);
    const double dom = -om * (c + d * opd), dde = c + d * opd * opd;
    const double dt1 = -d * t1 / (1.0 + d * rs);
    const double dt2 = xt2 * dde * 7.0 / 72.0;
    const double dt4 = -a4 * dde / 18.0 * (2.0 * xs2);
    const double dt5 = a5 * dde * (2.0 * xs2);
    frs = A * (dt1 + dom * tt + om * (dt2 + dt4 + dt5));
    fxt = -2.0 * A * om * xt * ((47.0 - 7.0 * de) / 72.0 - 2.0 / 3.0);
    fxs = A * om * 2.0 * xs * (a4 * (2.5 - de / 18.0) + a5 * (de - 11.0) - a6 * (2.0 / 3.0 - 0.25));
}

// PW92 correlation of the unpolarised gas with the PBE ("modified") constants: e_c(rs) and de_c/drs.
__device__ __forceinline__ void pw92_mod(double rs, double &ec, double &dec) {
    const double a = 0.0310907, alpha = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
    const double srs = sqrt(rs);
    const double q0 = -2.0 * a * (1.0 + alpha * rs);
    const double q1 = 2.0 * a * (b1 * srs + b2 * rs + b3 * srs * rs + b4 * rs * rs);
    const double q2 = log(1.0 + 1.0 / q1);
    ec = q0 * q2;
    const double dq0 = -2.0 * a * alpha;
    const double dq1 = a * (b1 / srs + 2.0 * b2 + 3.0 * b3 * srs + 4.0 * b4 * rs);
    dec = dq0 * q2 - q0 * dq1 / (q1 * (1.0 + q1));
}

// PBE correlation at zeta = 0 (phi = 1), t = xt / (4 2^(1/3) sqrt(rs)).
__device__ __forceinline__ void pbe_c(double rs, double xt, double &f, double &frs, double &fxt) {
    const double beta = 0.06672455060314922;
    const double gam = (1.0 - log(2.0)) / (M_PI * M_PI);
    double ec, dec;
    pw92_mod(rs, ec, dec);
    const double tconv = 4.0 * kCbrt2, srs = sqrt(rs), t = xt / (tconv * srs);
    const double e1 = ec / gam, e2 = exp(-e1), e3 = e2 - 1.0;
    const double A = beta / (gam * e3);
    const double dAdec = (A * e2 / e3) / gam;
    const double t2 = t * t, g1 = t2 + A * t2 * t2, g3 = 1.0 + A * g1, g2 = beta * g1 / (gam * g3);
    const double H = gam * log(1.0 + g2);
    const double dHdt = gam * (beta / (gam * g3 * g3) * (t * (2.0 + 4.0 * A * t2))) / (1.0 + g2);
    const double dHdA = gam * (beta / (gam * g3 * g3) * (t2 * t2 - g1 * g1)) / (1.0 + g2);
    f = ec + H;
    const double dtdrs = -xt / (2.0 * tconv * rs * srs);
    frs = (1.0 + dHdA * dAdec) * dec + dHdt * dtdrs;
    fxt = dHdt * (t / xt);
}

// Correlation: eps (per particle), vrho, vsigma.
__device__ __forceinline__ void gga_correlation(int which, double rho, double sigma, double &eps, double &vr, double &vs) {
    eps = 0.0; vr = 0.0; vs = 0.0;
    const double dens = rho > 0.0 ? rho : 0.0;
    if (dens < kXMinDens) return;
    const double rs = xq_cbrt(3.0 / (4.0 * M_PI * dens));
    const double st = sigma > kXMinGrad * kXMinGrad ? sigma : kXMinGrad * kXMinGrad;
    const double xt = sqrt(st) / pow(dens, 4.0 / 3.0), xs = xq_cbrt(2.0) * xt;
    double f, frs, fxt, fxs = 0.0;
    if (which == 0) lyp(rs, xt, xs, f, frs, fxt, fxs);
    else pbe_c(rs, xt, f, frs, fxt);
    const double drs = -rs / (3.0 * dens), dxt = -4.0 * xt / (3.0 * dens), dxtds = xt / (2.0 * st);
    eps = f;
    vr = f + dens * (frs * drs + fxt * dxt) + 2.0 * dens * fxs * (kCbrt2 * dxt);
    vs = dens * fxt * dxtds + 2.0 * dens * fxs * (kCbrt2 * dxtds);
}

// LDA pieces: rs = (3/(4 pi))^(1/3) rho^(-1/3), vrho = eps + rho (d eps/d rs)(-rs/(3 rho)); zero below the density threshold.
constexpr double kRsFactor = 0.6203504908994000866;              // (3/(4 pi))^(1/3)
constexpr double kSlaterAx = -0.458165293283142893475554485052;  // -(3/4) (3/(2 pi))^(2/3)

__device__ __forceinline__ void lda_finish(double dens, double rs, double e, double dedrs, double &eps, double &vr) {
    eps = e;
    vr = e + dens * dedrs * (-rs / (3.0 * dens));
}

__device__ __forceinline__ void slater_exchange(double rho, double &eps, double &vr, double &vs) {
    eps = 0.0; vr = 0.0; vs = 0.0;
    const double dens = rho > 0.0 ? rho : 0.0;
    if (dens < kXMinDens) return;
    const double rs = kRsFactor * pow(dens, -1.0 / 3.0);
    lda_finish(dens, rs, kSlaterAx / rs, -kSlaterAx / (rs * rs), eps, vr);
}

// Vosko-Wilk-Nusair correlation at zeta = 0 in x = sqrt(rs): eps = A [ln(x^2/X) + (2b/Q - f2 f3) atan(Q/(2x + b))
// - f2 ln((x - x0)^2/X)], X = x^2 + b x + c, Q = sqrt(4c - b^2), f2 = b x0 / X(x0), f3 = 2 (2 x0 + b)/Q; rpa selects the
// fit to the RPA energies (functional 8) over the default fit (functional 7).
__device__ __forceinline__ void vwn_correlation(int rpa, double rho, double &eps, double &vr, double &vs) {
    eps = 0.0; vr = 0.0; vs = 0.0;
    const double dens = rho > 0.0 ? rho : 0.0;
    if (dens < kXMinDens) return;
    const double A = 0.0310907;
    const double b = rpa ? 13.0720 : 3.72744, c = rpa ? 42.7198 : 12.9352, x0 = rpa ? -0.409286 : -0.10498;
    const double Q = sqrt(4.0 * c - b * b);
    const double f1 = 2.0 * b / Q, f2 = b * x0 / (x0 * x0 + b * x0 + c), f3 = 2.0 * (2.0 * x0 + b) / Q;
    const double rs = kRsFactor * pow(dens, -1.0 / 3.0), x = sqrt(rs);
    const double X = x * x + b * x + c, at = atan(Q / (2.0 * x + b)), xx0 = x - x0;
    const double t1 = 2.0 * x + b, t2 = 2.0 * c + b * x, t3 = t1 * t1 + Q * Q;
    const double dedx = A * (-2.0 * f2 / xx0 + (f2 * t1 + t2 / x) / X - 2.0 * Q * (f1 - f2 * f3) / t3);
    lda_finish(dens, rs, A * (log(x * x / X) + (f1 - f2 * f3) * at - f2 * log(xx0 * xx0 / X)), dedx / (2.0 * x), eps, vr);
}

// B3LYP's XC part: 0.08 Slater + 0.72 B88 + 0.19 VWN (RPA fit) + 0.81 LYP, summed in that order (the exact exchange,
// 0.20, is the caller's).
__device__ __forceinline__ void b3lyp(double rho, double sigma, double &eps, double &vr, double &vs) {
    const double c0 = 1.0 - 0.20 - 0.72, c1 = 0.72, c2 = 1.0 - 0.81, c3 = 0.81;
    double e, a, b;
    eps = 0.0; vr = 0.0; vs = 0.0;
    slater_exchange(rho, e, a, b);          eps += c0 * e; vr += c0 * a;
    gga_exchange(0, rho, sigma, e, a, b);    eps += c1 * e; vr += c1 * a; vs += c1 * b;
    vwn_correlation(1, rho, e, a, b);        eps += c2 * e; vr += c2 * a;
    gga_correlation(0, rho, sigma, e, a, b); eps += c3 * e; vr += c3 * a; vs += c3 * b;
}

// Sum over up to four functionals given by libxc ID (1 Slater, 7 VWN, 8 VWN-RPA, 101 PBE-x, 106 B88, 130 PBE-c, 131 LYP,
// 402 B3LYP).
struct FunctionalSet {
    int n;
    int id[4];
};

__device__ __forceinline__ void evaluate_functional(const FunctionalSet &fs, double rho, double sigma, double &eps,
                                                    double &vr, double &vs) {
    eps = 0.0; vr = 0.0; vs = 0.0;
    for (int k = 0; k < fs.n; k++) {
        double e, a, b;
        switch (fs.id[k]) {
            case 106: gga_exchange(0, rho, sigma, e, a, b); break;
            case 101: gga_exchange(1, rho, sigma, e, a, b); break;
            case 131: gga_correlation(0, rho, sigma, e, a, b); break;
            case 130: gga_correlation(1, rho, sigma, e, a, b); break;
            case 1: slater_exchange(rho, e, a, b); break;
            case 7: vwn_correlation(0, rho, e, a, b); break;
            case 8: vwn_correlation(1, rho, e, a, b); break;
            default: b3lyp(rho, sigma, e, a, b); break;
        }
        eps += e; vr += a; vs += b;
    }
}

}  // namespace xq
