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

// ------------------------------------------------------------------------------------------------------------------------
// Spin-polarised forms (libxc 1.2.0 with nspin = 2: work_gga_x.i:78-158, work_gga_c.i:55-131, gga_c_lyp.c:65-158,
// gga_c_pbe.c:80-218, lda_c_pw.c:62-216, work_lda.i:44-86, lda_x.c:121-149, lda_c_vwn.c:42-63,103-236,
// hyb_gga_xc_b3lyp.c:50-56): rho = (rho_a, rho_b), sigma = (s_aa, s_ab, s_bb) -> eps, vrho (a, b), vsigma (aa, ab, bb).
// ------------------------------------------------------------------------------------------------------------------------
struct PolOut {
    double e, va, vb, vsaa, vsab, vsbb;
};

constexpr double kMinZeta = 5.0e-13;
constexpr double kFZetaDen = 0.519842099789746380;   // 2^(4/3) - 2

__device__ __forceinline__ double f_zeta(double z) {
    return (pow(1.0 + z, 4.0 / 3.0) + pow(1.0 - z, 4.0 / 3.0) - 2.0) / kFZetaDen;
}
__device__ __forceinline__ double df_zeta(double z) { return (xq_cbrt(1.0 + z) - xq_cbrt(1.0 - z)) * (4.0 / 3.0) / kFZetaDen; }

__device__ __forceinline__ void dens_zeta(double ra, double rb, double &dens, double &zeta) {
    dens = ra + rb > 0.0 ? ra + rb : 0.0;
    zeta = dens > kXMinDens ? (ra - rb) / dens : 0.0;
    zeta = zeta > 1.0 ? 1.0 : (zeta < -1.0 ? -1.0 : zeta);
}

__device__ __forceinline__ void lda_finish_pol(double dens, double rs, double zeta, double e, double dedrs, double dedz,
                                               PolOut &o) {
    const double v = e + dens * dedrs * (-rs / (3.0 * dens));
    o.e = e;
    o.vb = v - (zeta + 1.0) * dedz;
    o.va = v - (zeta - 1.0) * dedz;
}

__device__ __forceinline__ void slater_pol(double ra, double rb, PolOut &o) {
    double dens, zeta;
    dens_zeta(ra, rb, dens, zeta);
    if (dens < kXMinDens) return;
    const double rs = kRsFactor * pow(dens, -1.0 / 3.0);
    const double fz = 0.5 * (pow(1.0 + zeta, 4.0 / 3.0) + pow(1.0 - zeta, 4.0 / 3.0));
    const double dfz = 2.0 / 3.0 * (xq_cbrt(1.0 + zeta) - xq_cbrt(1.0 - zeta));
    lda_finish_pol(dens, rs, zeta, kSlaterAx / rs * fz, -kSlaterAx / (rs * rs) * fz, kSlaterAx / rs * dfz, o);
}

// One VWN fit k (0 paramagnetic, 1 ferromagnetic, 2 spin stiffness) at x = sqrt(rs).
__device__ __forceinline__ void vwn_fit(int rpa, int k, double x, double &e, double &dedrs) {
    const double A = k == 0 ? 0.0310907 : (k == 1 ? 0.01554535 : -1.0 / (6.0 * M_PI * M_PI));
    double b, c, x0;
    if (rpa) {
        b = k == 0 ? 13.0720 : (k == 1 ? 20.1231 : 1.06835);
        c = k == 0 ? 42.7198 : (k == 1 ? 101.578 : 11.4813);
        x0 = k == 0 ? -0.409286 : (k == 1 ? -0.743294 : -0.228344);
    } else {
        b = k == 0 ? 3.72744 : (k == 1 ? 7.06042 : 1.13107);
        c = k == 0 ? 12.9352 : (k == 1 ? 18.0578 : 13.0045);
        x0 = k == 0 ? -0.10498 : (k == 1 ? -0.32500 : -0.0047584);
    }
    const double Q = sqrt(4.0 * c - b * b);
    const double f1 = 2.0 * b / Q, f2 = b * x0 / (x0 * x0 + b * x0 + c), f3 = 2.0 * (2.0 * x0 + b) / Q;
    const double X = x * x + b * x + c, at = atan(Q / (2.0 * x + b)), xx0 = x - x0;
    const double t1 = 2.0 * x + b, t2 = 2.0 * c + b * x, t3 = t1 * t1 + Q * Q;
    e = A * (log(x * x / X) + (f1 - f2 * f3) * at - f2 * log(xx0 * xx0 / X));
    dedrs = A * (-2.0 * f2 / xx0 + (f2 * t1 + t2 / x) / X - 2.0 * Q * (f1 - f2 * f3) / t3) / (2.0 * x);
}

// hl = 1: e_P + (e_F - e_P) f(zeta), the interpolation B3LYP sets for its VWN part; hl = 0: VWN's own.
__device__ __forceinline__ void vwn_pol(int rpa, int hl, double ra, double rb, PolOut &o) {
    double dens, zeta;
    dens_zeta(ra, rb, dens, zeta);
    if (dens < kXMinDens) return;
    const double rs = kRsFactor * pow(dens, -1.0 / 3.0), x = sqrt(rs);
    double e1, e2, e3, v1, v2, v3;
    vwn_fit(rpa, 0, x, e1, v1);
    vwn_fit(rpa, 1, x, e2, v2);
    vwn_fit(rpa, 2, x, e3, v3);
    const double fpp = 4.0 / (9.0 * (xq_cbrt(2.0) - 1.0));
    const double fz = f_zeta(zeta), dfz = df_zeta(zeta), z3 = pow(zeta, 3), z4 = z3 * zeta;
    double t1, t2, dt1, dt2;
    if (hl) {
        t1 = 0.0; t2 = fz; dt1 = 0.0; dt2 = dfz;
    } else {
        t1 = (fz / fpp) * (1.0 - z4); t2 = fz * z4;
        dt1 = (dfz * (1.0 - z4) - 4.0 * fz * z3) / fpp; dt2 = dfz * z4 + 4.0 * fz * z3;
    }
    lda_finish_pol(dens, rs, zeta, e1 + e3 * t1 + (e2 - e1) * t2, v1 + v3 * t1 + (v2 - v1) * t2,
                   e3 * dt1 + (e2 - e1) * dt2, o);
}

__device__ __forceinline__ void gga_exchange_pol(int which, double ra, double rb, double saa, double sbb, PolOut &o) {
    const double dens = ra + rb;
    if (dens < kXMinDens) return;
    double zk = 0.0;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double r = s == 0 ? ra : rb, sg = s == 0 ? saa : sbb;
        if (r < kXMinDens) continue;
        double gdm = sqrt(sg);
        if (gdm < kXMinGrad) gdm = kXMinGrad;
        const double rl = pow(r, 4.0 / 3.0), x = gdm / pow(r, 4.0 / 3.0);
        double F, dF;
        if (which == 0) b88_enhancement(x, F, dF);
        else pbe_x_enhancement(x, F, dF);
        zk += (-kXFactorC) * rl * F;
        const double v = (-kXFactorC) * (rl / r) * ((4.0 / 3.0) * F - (4.0 / 3.0) * dF * x);
        const double vs = gdm > kXMinGrad ? (-kXFactorC) * rl * (dF * x / (2.0 * sg)) : 0.0;
        if (s == 0) { o.va += v; o.vsaa = vs; }
        else { o.vb += v; o.vsbb = vs; }
    }
    o.e = zk / dens;
}

// LYP at general zeta: f and its derivatives in rs, zeta, xt, xs_a, xs_b.
__device__ __forceinline__ void lyp_z(double rs, double zeta, double xt, double xsa, double xsb, double &f, double &frs,
                                      double &fz, double &fxt, double &fxa, double &fxb) {
    const double A = 0.04918, B = 0.132;
    const double crs = xq_cbrt(4.0 * M_PI / 3.0), Cf = 3.0 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0) / 10.0;
    const double c = 0.2533 * crs, d = 0.349 * crs;
    const double xt2 = xt * xt, z2 = zeta * zeta, opz = 1.0 + zeta, omz = 1.0 - zeta;
    const double opz23 = pow(opz, 2.0 / 3.0), omz23 = pow(omz, 2.0 / 3.0);
    const double opz53 = opz * opz23, omz53 = omz * omz23, opz83 = opz * opz53, omz83 = omz * omz53;
    const double xa2 = xsa * xsa, xb2 = xsb * xsb;
    const double opd = 1.0 / (1.0 + d * rs), om = B * exp(-c * rs) * opd, de = (c + d * opd) * rs;
    const double a6 = 1.0 / pow(2.0, 8.0 / 3.0), a4 = a6 / 4.0, a5 = a4 / (9.0 * 2.0);
    const double S8 = xa2 * opz83 + xb2 * omz83, S11 = xa2 * opz * opz83 + xb2 * omz * omz83;
    const double t1 = -(1.0 - z2) / (1.0 + d * rs);
    const double t2 = -xt2 * ((1.0 - z2) * (47.0 - 7.0 * de) / (4.0 * 18.0) - 2.0 / 3.0);
    const double t3 = -Cf / 2.0 * (1.0 - z2) * (opz83 + omz83);
    const double t4 = a4 * (1.0 - z2) * (5.0 / 2.0 - de / 18.0) * S8;
    const double t5 = a5 * (1.0 - z2) * (de - 11.0) * S11;
    const double t6 = -a6 * (2.0 / 3.0 * S8 - opz * opz * xb2 * omz83 / 4.0 - omz * omz * xa2 * opz83 / 4.0);
    const double tt = t2 + t3 + t4 + t5 + t6;
    const double dom = -om * (c + d * opd), dde = c + d * opd * opd;
    f = A * (t1 + om * tt);
    frs = A * (-d * t1 / (1.0 + d * rs) + dom * tt +
               om * (xt2 * (1.0 - z2) * dde * 7.0 / (4.0 * 18.0) - a4 * (1.0 - z2) * dde / 18.0 * S8 +
                     a5 * (1.0 - z2) * dde * S11));
    fz = A * (2.0 * zeta / (1.0 + d * rs) +
              om * (xt2 * 2.0 * zeta * (47.0 - 7.0 * de) / (4.0 * 18.0) -
                    Cf / 2.0 * (-2.0 * zeta * (opz83 + omz83) + (1.0 - z2) * 8.0 / 3.0 * (opz53 - omz53)) +
                    a4 * (5.0 / 2.0 - de / 18.0) * (-2.0 * zeta * S8 + (1.0 - z2) * 8.0 / 3.0 * (xa2 * opz53 - xb2 * omz53)) +
                    a5 * (de - 11.0) * (-2.0 * zeta * S11 + (1.0 - z2) * 11 / 3.0 * (xa2 * opz83 - xb2 * omz83)) -
                    a6 * (16.0 / 9.0 * (xa2 * opz53 - xb2 * omz53) - 1.0 / 2.0 * (opz * xb2 * omz83 - omz * xa2 * opz83) +
                          2.0 / 3.0 * (opz * opz * xb2 * omz53 - omz * omz * xa2 * opz53))));
    fxt = -2.0 * A * om * xt * ((1.0 - z2) * (47.0 - 7.0 * de) / (4.0 * 18.0) - 2.0 / 3.0);
    fxa = A * om * 2.0 * xsa * (a4 * (1.0 - z2) * (5.0 / 2.0 - de / 18.0) * opz83 + a5 * (1.0 - z2) * (de - 11.0) * opz * opz83 -
                                a6 * (2.0 / 3.0 * opz83 - omz * omz * opz83 / 4.0));
    fxb = A * om * 2.0 * xsb * (a4 * (1.0 - z2) * (5.0 / 2.0 - de / 18.0) * omz83 + a5 * (1.0 - z2) * (de - 11.0) * omz * omz83 -
                                a6 * (2.0 / 3.0 * omz83 - opz * opz * omz83 / 4.0));
}

// PW92 "modified" fit k (0 paramagnetic, 1 ferromagnetic, 2 minus the spin stiffness).
__device__ __forceinline__ void pw92_fit(int k, double rs, double &e, double &dedrs) {
    const double a = k == 0 ? 0.0310907 : (k == 1 ? 0.01554535 : 0.0168869);
    const double al = k == 0 ? 0.21370 : (k == 1 ? 0.20548 : 0.11125);
    const double b1 = k == 0 ? 7.5957 : (k == 1 ? 14.1189 : 10.357), b2 = k == 0 ? 3.5876 : (k == 1 ? 6.1977 : 3.6231);
    const double b3 = k == 0 ? 1.6382 : (k == 1 ? 3.3662 : 0.88026), b4 = k == 0 ? 0.49294 : (k == 1 ? 0.62517 : 0.49671);
    const double srs = sqrt(rs);
    const double q0 = -2.0 * a * (1.0 + al * rs);
    const double q1 = 2.0 * a * (b1 * srs + b2 * rs + b3 * srs * rs + b4 * pow(rs, 2.0));
    const double q2 = log(1.0 + 1.0 / q1);
    const double dq1 = a * (b1 / srs + 2.0 * b2 + 3.0 * b3 * srs + 2.0 * 2.0 * b4 * pow(rs, 1.0));
    e = q0 * q2;
    dedrs = -2.0 * a * al * q2 - q0 * dq1 / (q1 * (1.0 + q1));
}

__device__ __forceinline__ void pbe_c_z(double rs, double zeta, double xt, double &f, double &frs, double &fz, double &fxt) {
    const double beta = 0.06672455060314922, gam = (1.0 - log(2.0)) / (M_PI * M_PI);
    const double fz20 = 1.709920934161365617563962776245;
    double ecp, vcp, ecf, vcf, al, dal;
    pw92_fit(0, rs, ecp, vcp);
    pw92_fit(1, rs, ecf, vcf);
    pw92_fit(2, rs, al, dal);
    al = -al;
    dal = -dal;
    const double fzz = f_zeta(zeta), dfz = df_zeta(zeta), z3 = zeta * zeta * zeta, z4 = zeta * z3;
    const double ec = ecp + z4 * fzz * (ecf - ecp - al / fz20) + fzz * al / fz20;
    const double decdrs = vcp + z4 * fzz * (vcf - vcp - dal / fz20) + fzz * dal / fz20;
    const double decdz = (4.0 * z3 * fzz + z4 * dfz) * (ecf - ecp - al / fz20) + dfz * al / fz20;
    const double tconv = 4.0 * kCbrt2, auxp = xq_cbrt(1.0 + zeta), auxm = xq_cbrt(1.0 - zeta);
    const double phi = 0.5 * (auxp * auxp + auxm * auxm), phi3 = pow(phi, 3), srs = sqrt(rs);
    const double t = xt / (tconv * phi * srs), t2 = t * t;
    const double e1 = ec / (gam * phi3), e2 = exp(-e1), e3 = e2 - 1.0;
    const double Aa = beta / (gam * e3), dx = Aa * e2 / e3;
    const double dAdec = dx / (gam * phi3), dAdphi = dx * (-3.0 * e1 / phi);
    const double g1 = t2 + Aa * t2 * t2, g3 = 1.0 + Aa * g1, g2 = beta * g1 / (gam * g3);
    const double H = gam * phi3 * log(1.0 + g2);
    const double dHdphi = 3.0 * H / phi;
    const double dHdt = gam * phi3 * (beta / (gam * g3 * g3) * t * (2.0 + 4.0 * Aa * t2)) / (1.0 + g2);
    const double dHdA = gam * phi3 * (beta / (gam * g3 * g3) * (t2 * t2 - g1 * g1)) / (1.0 + g2);
    const double dfdphi = dHdphi + dHdA * dAdphi, dfdec = 1.0 + dHdA * dAdec;
    double dphidz = 0.0;
    if (auxp > kMinZeta) dphidz += 1 / auxp;
    if (auxm > kMinZeta) dphidz -= 1 / auxm;
    dphidz *= 1.0 / 3.0;
    f = ec + H;
    frs = dfdec * decdrs + dHdt * (-xt / (2.0 * tconv * phi * rs * srs));
    fz = dfdec * decdz + (dfdphi + dHdt * (-t / phi)) * dphidz;
    fxt = dHdt * (t / xt);
}

__device__ __forceinline__ void gga_correlation_pol(int which, double ra, double rb, double saa, double sab, double sbb,
                                                    PolOut &o) {
    const double mg2 = kXMinGrad * kXMinGrad;
    double dens, zeta;
    dens_zeta(ra, rb, dens, zeta);
    if (dens < kXMinDens) return;
    const double rs = xq_cbrt(3.0 / (4.0 * M_PI * dens));
    const double da = ra > kXMinDens ? ra : kXMinDens, db = rb > kXMinDens ? rb : kXMinDens;
    double st = saa + 2.0 * sab + sbb;
    if (st < mg2) st = mg2;
    const double xt = sqrt(st) / pow(dens, 4.0 / 3.0);
    const double s0 = saa > mg2 ? saa : mg2, s2 = sbb > mg2 ? sbb : mg2;
    const double xsa = sqrt(s0) / pow(da, 4.0 / 3.0), xsb = sqrt(s2) / pow(db, 4.0 / 3.0);
    double f, frs, fz, fxt, fxa = 0.0, fxb = 0.0;
    if (which == 0) lyp_z(rs, zeta, xt, xsa, xsb, f, frs, fz, fxt, fxa, fxb);
    else pbe_c_z(rs, zeta, xt, f, frs, fz, fxt);
    const double drs = -rs / (3.0 * dens), dxt = -4.0 * xt / (3.0 * dens), dxtds = xt / (2.0 * st);
    const double v0 = f + dens * (frs * drs + fxt * dxt), vs0 = dens * fxt * dxtds;
    o.e = f;
    o.vb = v0 + fz * (-(zeta + 1.0)) + dens * fxb * (-4.0 / 3.0 * xsb / db);
    o.va = v0 + fz * (-(zeta - 1.0)) + dens * fxa * (-4.0 / 3.0 * xsa / da);
    o.vsbb = vs0 + dens * fxb * (xsb / (2.0 * s2));
    o.vsab = 2.0 * vs0;
    o.vsaa = vs0 + dens * fxa * (xsa / (2.0 * s0));
}

__device__ __forceinline__ void functional_pol(int id, double ra, double rb, double saa, double sab, double sbb, PolOut &o) {
    o = PolOut{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    switch (id) {
        case 106: gga_exchange_pol(0, ra, rb, saa, sbb, o); break;
        case 101: gga_exchange_pol(1, ra, rb, saa, sbb, o); break;
        case 131: gga_correlation_pol(0, ra, rb, saa, sab, sbb, o); break;
        case 130: gga_correlation_pol(1, ra, rb, saa, sab, sbb, o); break;
        case 1: slater_pol(ra, rb, o); break;
        case 7: vwn_pol(0, 0, ra, rb, o); break;
        case 8: vwn_pol(1, 0, ra, rb, o); break;
        default: {   // 402: B3LYP
            const double coef[4] = {1.0 - 0.20 - 0.72, 0.72, 1.0 - 0.81, 0.81};
            PolOut q;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                q = PolOut{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                if (k == 0) slater_pol(ra, rb, q);
                else if (k == 1) gga_exchange_pol(0, ra, rb, saa, sbb, q);
                else if (k == 2) vwn_pol(1, 1, ra, rb, q);
                else gga_correlation_pol(0, ra, rb, saa, sab, sbb, q);
                o.e += coef[k] * q.e; o.va += coef[k] * q.va; o.vb += coef[k] * q.vb;
                if (k == 1 || k == 3) { o.vsaa += coef[k] * q.vsaa; o.vsab += coef[k] * q.vsab; o.vsbb += coef[k] * q.vsbb; }
            }
        }
    }
}

// Sum over the set's functionals, polarised.
__device__ __forceinline__ void evaluate_functional_pol(const FunctionalSet &fs, double ra, double rb, double saa,
                                                        double sab, double sbb, PolOut &out) {
    out = PolOut{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < fs.n; k++) {
        PolOut q;
        functional_pol(fs.id[k], ra, rb, saa, sab, sbb, q);
        out.e += q.e; out.va += q.va; out.vb += q.vb; out.vsaa += q.vsaa; out.vsab += q.vsab; out.vsbb += q.vsbb;
    }
}

}  // namespace xq
