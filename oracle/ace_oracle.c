/*
 * ace_oracle.c -- FP64 CPU restatement of the reference hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by the tests, the smoke check and the bench's CPU-baseline leg.
 * The product path never links or calls it.  Built with -ffp-contract=off.
 * Citations are file:line in the ace source tree.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "ace_oracle.h"

#define OR_PI   3.141592653589793   /* math/const.hpp:9  */
#define OR_RAD2 1.414213562373095   /* math/const.hpp:11 */
#define OR_LOG2 0.693147180559945   /* math/const.hpp:15 */
#define OR_ZERO 1e-12               /* math/const.hpp:8  */

/* ------------------------------------------------------------------------------------------- */
/* scalar math                                                                                  */
/* ------------------------------------------------------------------------------------------- */

/* degree-15 Taylor series of cos about pi/2, special.cpp:21-27 */
double or_coscut(double x) {
    x -= 0.5 * OR_PI;
    const double x2 = x * x;
    return x * (-1.0 + x2 * (1.0 / 6.0 + x2 * (-1.0 / 120.0 + x2 * (1.0 / 5040.0 + x2 * (-1.0 / 362880.0
           + x2 * (1.0 / 39916800.0 + x2 * (-1.0 / 6227020800.0 + x2 * 1.0 / 1307674368000.0)))))));
}

/* degree-16 Taylor series of sin about pi/2, special.cpp:35-41 */
double or_sincut(double x) {
    x -= 0.5 * OR_PI;
    const double x2 = x * x;
    return 1.0 + x2 * (-1.0 / 2.0 + x2 * (1.0 / 24.0 + x2 * (-1.0 / 720.0 + x2 * (1.0 / 40320.0
           + x2 * (-1.0 / 3628800.0 + x2 * (1.0 / 479001600.0 + x2 * (-1.0 / 87178291200.0
           + x2 * 1.0 / 20922789888000.0)))))));
}

/* exp2 by exponent assembly and a [-0.5,0.5) Pade fraction, special.cpp:226-267 */
static double or_exp2(double x) {
    static const double q[2] = {2.33184211722314911771e2, 4.36821166879210612817e3};
    static const double p[3] = {2.30933477057345225087e-2, 2.02020656693165307700e1,
                                1.51390680115615096133e3};
    const double ipart = floor(x + 0.5);
    const double fpart = x - ipart;
    const uint64_t bits = (uint64_t)(uint32_t)((((int)ipart) + 1023) << 20) << 32;
    double epart;
    memcpy(&epart, &bits, sizeof(double));
    x = fpart * fpart;
    double px = p[0];
    px = px * x + p[1];
    double qx = x + q[0];
    px = px * x + p[2];
    qx = qx * x + q[1];
    px = px * fpart;
    x = 1.0 + 2.0 * (px / (qx - px));
    return epart * x;
}

/* special.cpp:269-276 */
double or_fmexp(double x) {
    const double l2e = 1.4426950408889634074;
    if (x < -1022.0 / l2e) return 0;
    if (x > 1023.0 / l2e) return INFINITY;
    return or_exp2(l2e * x);
}

/* special.cpp:139-148 */
double or_powint(double x, int n) {
    double yy, ww;
    if (n == 0) return 1.0;
    if (x == 0.0) return 0.0;
    int nn = (n > 0) ? n : -n;
    ww = x;
    for (yy = 1.0; nn != 0; nn >>= 1, ww *= ww)
        if (nn & 1) yy *= ww;
    return (n > 0) ? yy : 1.0 / yy;
}

static double or_sech(double x) {   /* special.hpp sech: 1/cosh */
    return 1.0 / cosh(x);
}

/* ------------------------------------------------------------------------------------------- */
/* cutoff, cutoff.cpp:62-117                                                                    */
/* ------------------------------------------------------------------------------------------- */

double or_cutf(int name, double rc, double dr) {
    const double rci = 1.0 / rc, pirci = OR_PI / rc;
    switch (name) {
        case ACE_CUT_STEP: return (dr < rc) ? 1.0 : 0.0;
        case ACE_CUT_COS: return (dr < rc) ? 0.5 * (or_coscut(dr * pirci) + 1.0) : 0.0;
        case ACE_CUT_TANH: {
            const double f = (dr > rc) ? 0.0 : tanh(1.0 - dr * rci);
            return f * f * f;
        }
        case ACE_CUT_POLY3:
            if (dr < rc) { dr *= rci; return (2.0 * dr - 3.0) * dr * dr + 1.0; }
            return 0.0;
        default: return NAN;
    }
}

double or_cutg(int name, double rc, double dr) {
    const double rci = 1.0 / rc, pirci = OR_PI / rc;
    switch (name) {
        case ACE_CUT_STEP: return 0.0;
        case ACE_CUT_COS: return (dr < rc) ? -0.5 * pirci * or_sincut(dr * pirci) : 0.0;
        case ACE_CUT_TANH:
            if (dr < rc) { const double f = tanh(1.0 - dr * rci); return -3.0 * f * f * (1.0 - f * f) * rci; }
            return 0.0;
        case ACE_CUT_POLY3:
            if (dr < rc) { dr *= rci; return 6.0 * dr * (dr - 1.0) * rci; }
            return 0.0;
        default: return NAN;
    }
}

/* ------------------------------------------------------------------------------------------- */
/* single radial function and derivative, basis_radial.cpp:175-257                              */
/* ------------------------------------------------------------------------------------------- */

double or_radial_f(int form, int cut, double rc, double eta, double rs, double dr) {
    const double cutf = or_cutf(cut, rc, dr);
    switch (form) {
        case ACE_RAD_GAUSSIAN: { const double arg = eta * (dr - rs); return or_fmexp(-arg * arg) * cutf; }
        case ACE_RAD_SECH: return or_sech(eta * (dr - rs)) * cutf;
        case ACE_RAD_LOGISTIC: { const double s = or_sech(eta * (dr - rs)); return s * s * cutf; }
        case ACE_RAD_TANH: return 0.5 * (tanh(-eta * (dr - rs)) + 1.0) * cutf;
        case ACE_RAD_LOGCOSH: return 0.5 * log1p(or_fmexp(-2.0 * eta * (dr - rs))) * cutf;
        case ACE_RAD_LOGCOSH2: {
            const double arg = -eta * (dr - rs);
            const double fexp = or_fmexp(2.0 * OR_RAD2 * arg);
            return (arg == 0.0) ? 0.5 * cutf / OR_RAD2 : arg * fexp / (fexp - 1.0) * cutf;
        }
        default: return NAN;
    }
}

double or_radial_d(int form, int cut, double rc, double eta, double rs, double dr) {
    const double cutf = or_cutf(cut, rc, dr);
    const double cutg = or_cutg(cut, rc, dr);
    switch (form) {
        case ACE_RAD_GAUSSIAN: {
            const double arg = eta * (dr - rs);
            return or_fmexp(-arg * arg) * (-2.0 * eta * arg * cutf + cutg);
        }
        case ACE_RAD_SECH: {
            const double expf = or_fmexp(-eta * (dr - rs));
            const double expf2 = expf * expf;
            const double den = 1.0 / (1.0 + expf2);
            return 2.0 * expf * den * (-1.0 * eta * (1.0 - expf2) * den * cutf + cutg);
        }
        case ACE_RAD_LOGISTIC: {
            const double s = or_sech(eta * (dr - rs));
            const double t = tanh(eta * (dr - rs));
            return s * s * (-2.0 * eta * t * cutf + cutg);
        }
        case ACE_RAD_TANH: {
            const double t = tanh(-eta * (dr - rs));
            return 0.5 * (-eta * (1.0 - t * t) * cutf + (1.0 + t) * cutg);
        }
        case ACE_RAD_LOGCOSH: {
            const double fexp = or_fmexp(-2.0 * eta * (dr - rs));
            return -eta * fexp / (1.0 + fexp) * cutf + 0.5 * log1p(fexp) * cutg;
        }
        case ACE_RAD_LOGCOSH2: {
            const double arg = -eta * (dr - rs);
            const double fexp = or_fmexp(2.0 * OR_RAD2 * arg);
            const double den = 1.0 / (fexp - 1.0);
            return (arg == 0.0) ? 0.5 * (-eta * cutg + cutf / OR_RAD2)
                                : fexp * (eta * (1.0 + 2.0 * OR_RAD2 * arg - fexp) * den * cutf + arg * cutg) * den;
        }
        default: return NAN;
    }
}

/* block evaluation, BasisR::symm basis_radial.cpp:263-309 */
static void radial_symm(int form, int cut, double rc, int n, const double* rs, const double* eta,
                        double dr, double* out) {
    const double cutf = or_cutf(cut, rc, dr);
    for (int i = 0; i < n; ++i) {
        switch (form) {
            case ACE_RAD_GAUSSIAN: { const double arg = eta[i] * (dr - rs[i]); out[i] = or_fmexp(-arg * arg) * cutf; } break;
            case ACE_RAD_SECH: { const double e = or_fmexp(-eta[i] * (dr - rs[i])); out[i] = 2.0 * e / (1.0 + e * e) * cutf; } break;
            case ACE_RAD_LOGISTIC: { const double e = or_fmexp(-2.0 * eta[i] * (dr - rs[i])); out[i] = 4.0 * e / ((1.0 + e) * (1.0 + e)) * cutf; } break;
            case ACE_RAD_TANH: out[i] = 0.5 * (tanh(-eta[i] * (dr - rs[i])) + 1.0) * cutf; break;
            case ACE_RAD_LOGCOSH: out[i] = 0.5 * log1p(or_fmexp(-2.0 * eta[i] * (dr - rs[i]))) * cutf; break;
            case ACE_RAD_LOGCOSH2: {
                const double arg = -eta[i] * (dr - rs[i]);
                const double fexp = or_fmexp(2.0 * OR_RAD2 * arg);
                out[i] = (arg == 0.0) ? 0.5 * cutf / OR_RAD2 : arg * fexp / (fexp - 1.0) * cutf;
            } break;
            default: out[i] = NAN;
        }
    }
}

/* BasisR::force basis_radial.cpp:316-376: amp = -sum_i dEdG[i] dG_i/dr */
static double radial_force(int form, int cut, double rc, int n, const double* rs, const double* eta,
                           double dr, const double* dEdG) {
    double amp = 0;
    const double cutf = or_cutf(cut, rc, dr);
    const double cutg = or_cutg(cut, rc, dr);
    for (int i = 0; i < n; ++i) {
        switch (form) {
            case ACE_RAD_GAUSSIAN: {
                const double arg = eta[i] * (dr - rs[i]);
                amp -= dEdG[i] * or_fmexp(-arg * arg) * (-2.0 * eta[i] * arg * cutf + cutg);
            } break;
            case ACE_RAD_SECH: {
                const double expf = or_fmexp(-eta[i] * (dr - rs[i]));
                const double expf2 = expf * expf;
                const double den = 1.0 / (1.0 + expf2);
                amp -= dEdG[i] * 2.0 * expf * den * (-eta[i] * (1.0 - expf2) * den * cutf + cutg);
            } break;
            case ACE_RAD_LOGISTIC: {
                const double fexp2 = or_fmexp(-2.0 * eta[i] * (dr - rs[i]));
                const double den = 1.0 / (1.0 + fexp2);
                const double tanhf = (1.0 - fexp2) * den;
                const double sech2f = 4.0 * fexp2 * den * den;
                amp -= dEdG[i] * sech2f * (-2.0 * eta[i] * tanhf * cutf + cutg);
            } break;
            case ACE_RAD_TANH: {
                const double t = tanh(-eta[i] * (dr - rs[i]));
                amp -= dEdG[i] * 0.5 * (-eta[i] * (1.0 - t * t) * cutf + (1.0 + t) * cutg);
            } break;
            case ACE_RAD_LOGCOSH: {
                const double fexp = or_fmexp(-2.0 * eta[i] * (dr - rs[i]));
                amp -= dEdG[i] * (-eta[i] * fexp / (1.0 + fexp) * cutf + 0.5 * log1p(fexp) * cutg);
            } break;
            case ACE_RAD_LOGCOSH2: {
                const double arg = -eta[i] * (dr - rs[i]);
                const double fexp = or_fmexp(2.0 * OR_RAD2 * arg);
                const double den = 1.0 / (fexp - 1.0);
                const double val = (arg == 0.0) ? 0.5 * (-eta[i] * cutg + cutf / OR_RAD2)
                    : fexp * (eta[i] * (1.0 + 2.0 * OR_RAD2 * arg - fexp) * den * cutf + arg * cutg) * den;
                amp -= dEdG[i] * val;
            } break;
            default: amp = NAN;
        }
    }
    return amp;
}

/* ------------------------------------------------------------------------------------------- */
/* angular functions, basis_angular.cpp                                                         */
/* ------------------------------------------------------------------------------------------- */

/* single function value, basis_angular.cpp:212-249 */
double or_angular_f(int form, int cut, double rc, double eta, double zeta, int lambda,
                    double cosv, double rij, double rik) {
    const double c0 = or_cutf(cut, rc, rij), c1 = or_cutf(cut, rc, rik);
    const double hcos = 0.5 * cosv;
    const double cprod = c0 * c1;
    const double ieta2 = 1.0 / (eta * eta);
    const double r2s = rij * rij + rik * rik;
    const double pw = pow(fabs(0.5 + lambda * hcos), zeta);
    switch (form) {
        case ACE_ANG_GAUSS: return cprod * pw * or_fmexp(-ieta2 * r2s);
        case ACE_ANG_GAUSS2: { const double f = or_fmexp(-0.5 * OR_PI * ieta2 * r2s); return cprod * pw * 2.0 * f / (1.0 + f); }
        case ACE_ANG_SECH: { const double f = or_fmexp(-0.5 * OR_PI * ieta2 * r2s); return cprod * pw * 2.0 * f / (1.0 + f * f); }
        case ACE_ANG_STUDENT3: return cprod * pw * or_powint(1.0 / sqrt(1.0 + ieta2 * r2s), 4);
        case ACE_ANG_STUDENT4: return cprod * pw * or_powint(1.0 / sqrt(1.0 + ieta2 * r2s), 5);
        case ACE_ANG_STUDENT5: return cprod * pw * or_powint(1.0 / sqrt(1.0 + ieta2 * r2s), 6);
        default: return NAN;
    }
}

/* single function derivatives (d/dcos, d/drij, d/drik), basis_angular.cpp:255-364 */
void or_angular_d(int form, int cut, double rc, double eta, double zeta, int lambda,
                  double cosv, double rij, double rik, double* fphi, double* feta) {
    const double c[2] = {or_cutf(cut, rc, rij), or_cutf(cut, rc, rik)};
    const double g[2] = {or_cutg(cut, rc, rij), or_cutg(cut, rc, rik)};
    const double lambdaf = 0.5 * lambda;
    const double ieta2 = 1.0 / (eta * eta);
    const double r2s = rij * rij + rik * rik;
    const double dij = 2.0 * rij * c[0];
    const double dik = 2.0 * rik * c[1];
    const double cw = fabs(0.5 + lambdaf * cosv);
    double radf, rscale;   /* radial damping value and d(ln damping)/d(r2s) * 2 scale */
    *fphi = 0; feta[0] = 0; feta[1] = 0;
    switch (form) {
        case ACE_ANG_GAUSS: {
            const double gangle = pow(cw, zeta - 1.0) * or_fmexp(-ieta2 * r2s);
            const double fangle = cw * gangle;
            *fphi = zeta * lambdaf * gangle;
            feta[0] = fangle * (-dij * ieta2 + g[0]);
            feta[1] = fangle * (-dik * ieta2 + g[1]);
        } break;
        case ACE_ANG_GAUSS2: {
            const double fexp = or_fmexp(-0.5 * OR_PI * ieta2 * r2s);
            const double den = 1.0 / (1.0 + fexp);
            const double gangle = pow(cw, zeta - 1.0) * 2.0 * fexp * den;
            const double fangle = cw * gangle;
            *fphi = zeta * lambdaf * gangle;
            feta[0] = fangle * (-dij * ieta2 * 0.5 * OR_PI * den + g[0]);
            feta[1] = fangle * (-dik * ieta2 * 0.5 * OR_PI * den + g[1]);
        } break;
        case ACE_ANG_SECH: {
            const double fexp = or_fmexp(-0.5 * OR_PI * ieta2 * r2s);
            const double fexp2 = fexp * fexp;
            const double den = 1.0 / (1.0 + fexp2);
            const double ftanh = (1.0 - fexp2) * den;
            const double fsech = 2.0 * fexp * den;
            const double gangle = pow(cw, zeta - 1.0) * fsech;
            const double fangle = cw * gangle;
            *fphi = zeta * lambdaf * gangle;
            feta[0] = fangle * (-dij * 0.5 * OR_PI * ieta2 * ftanh + g[0]);
            feta[1] = fangle * (-dik * 0.5 * OR_PI * ieta2 * ftanh + g[1]);
        } break;
        case ACE_ANG_STUDENT3: case ACE_ANG_STUDENT4: case ACE_ANG_STUDENT5: {
            const int p = (form == ACE_ANG_STUDENT3) ? 4 : (form == ACE_ANG_STUDENT4) ? 5 : 6;
            const double den = 1.0 / sqrt(1.0 + ieta2 * r2s);
            const double gangle = pow(cw, zeta - 1.0) * or_powint(den, p);
            const double fangle = cw * gangle;
            *fphi = zeta * lambdaf * gangle;
            feta[0] = fangle * (-dij * ieta2 * p / 2.0 * den * den + g[0]);
            feta[1] = fangle * (-dik * ieta2 * p / 2.0 * den * den + g[1]);
        } break;
        default: *fphi = NAN; return;
    }
    (void)radf; (void)rscale;
    *fphi *= c[0] * c[1];
    feta[0] *= c[1];
    feta[1] *= c[0];
}

/* angular block with the BasisA::init run grouping of equal eta (basis_angular.cpp:178-205) */
typedef struct {
    int form, cut, n, nr;
    double rc;
    const double* eta;
    const double* zeta;
    const int* lambda;
    double* ieta2;
    double* lambdaf;
    int* rflag;
    double* ietar2;
    double* damp;
} or_ablk;

static void ablk_init(const ace_model_desc* m, int blk, or_ablk* b) {
    b->form = m->a_form[blk];
    b->cut = m->a_cut[blk];
    b->rc = m->a_rc[blk];
    b->n = m->a_n[blk];
    b->eta = m->a_eta + m->a_off[blk];
    b->zeta = m->a_zeta + m->a_off[blk];
    b->lambda = m->a_lambda + m->a_off[blk];
    const int n = b->n > 0 ? b->n : 1;
    b->ieta2 = (double*)malloc(sizeof(double) * n);
    b->lambdaf = (double*)malloc(sizeof(double) * n);
    b->rflag = (int*)malloc(sizeof(int) * n);
    b->ietar2 = (double*)malloc(sizeof(double) * n);
    b->damp = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < b->n; ++i) {
        b->ieta2[i] = 1.0 / (b->eta[i] * b->eta[i]);
        b->lambdaf[i] = 0.5 * b->lambda[i];
    }
    b->nr = 0;
    if (b->n > 0) {
        int flag = 0;
        double etar[ACE_MAX_TYPES * 64];
        int ne = 0;
        b->rflag[0] = flag;
        etar[ne++] = b->eta[0];
        for (int i = 1; i < b->n; ++i) {
            if (fabs(b->eta[i] - b->eta[i - 1]) < 1.0e-6) b->rflag[i] = flag;
            else { b->rflag[i] = ++flag; etar[ne++] = b->eta[i]; }
        }
        b->nr = flag + 1;
        for (int i = 0; i < b->nr; ++i) b->ietar2[i] = 1.0 / (etar[i] * etar[i]);
    }
}

static void ablk_free(or_ablk* b) {
    free(b->ieta2); free(b->lambdaf); free(b->rflag); free(b->ietar2); free(b->damp);
}

/* BasisA::symm, basis_angular.cpp:371-420 */
static void angular_symm(or_ablk* b, double drIJ, double drIK, double cosIJK, double* out) {
    const double cprod = or_cutf(b->cut, b->rc, drIJ) * or_cutf(b->cut, b->rc, drIK);
    const double r2s = drIJ * drIJ + drIK * drIK;
    switch (b->form) {
        case ACE_ANG_GAUSS:
            for (int j = 0; j < b->nr; ++j) b->damp[j] = cprod * or_fmexp(-b->ietar2[j] * r2s);
            for (int i = 0; i < b->n; ++i)
                out[i] = pow(fabs(0.5 + b->lambdaf[i] * cosIJK), b->zeta[i]) * b->damp[b->rflag[i]];
            break;
        case ACE_ANG_GAUSS2:
            for (int j = 0; j < b->nr; ++j) b->damp[j] = or_fmexp(-0.5 * OR_PI * b->ietar2[j] * r2s);
            for (int i = 0; i < b->n; ++i) {
                const double f = b->damp[b->rflag[i]];
                out[i] = cprod * pow(fabs(0.5 + b->lambdaf[i] * cosIJK), b->zeta[i]) * 2.0 * f / (1.0 + f);
            }
            break;
        case ACE_ANG_SECH:
            for (int j = 0; j < b->nr; ++j) b->damp[j] = or_fmexp(-0.5 * OR_PI * b->ietar2[j] * r2s);
            for (int i = 0; i < b->n; ++i) {
                const double f = b->damp[b->rflag[i]];
                out[i] = cprod * pow(fabs(0.5 + b->lambdaf[i] * cosIJK), b->zeta[i]) * 2.0 * f / (1.0 + f * f);
            }
            break;
        case ACE_ANG_STUDENT3: case ACE_ANG_STUDENT4: case ACE_ANG_STUDENT5: {
            const int p = (b->form == ACE_ANG_STUDENT3) ? 4 : (b->form == ACE_ANG_STUDENT4) ? 5 : 6;
            for (int j = 0; j < b->nr; ++j) b->damp[j] = cprod * or_powint(1.0 / sqrt(1.0 + b->ietar2[j] * r2s), p);
            for (int i = 0; i < b->n; ++i)
                out[i] = pow(fabs(0.5 + b->lambdaf[i] * cosIJK), b->zeta[i]) * b->damp[b->rflag[i]];
        } break;
        default:
            for (int i = 0; i < b->n; ++i) out[i] = NAN;
    }
}

/* BasisA::force, basis_angular.cpp:430-563: phi = -sum dEdG dG/dcos, eta = -sum dEdG dG/dr */
static void angular_force(or_ablk* b, double drIJ, double drIK, double cosIJK, double* phi,
                          double* eta, const double* dEdG) {
    const double c[2] = {or_cutf(b->cut, b->rc, drIJ), or_cutf(b->cut, b->rc, drIK)};
    const double g[2] = {or_cutg(b->cut, b->rc, drIJ), or_cutg(b->cut, b->rc, drIK)};
    *phi = 0; eta[0] = 0; eta[1] = 0;
    const double r2s = drIJ * drIJ + drIK * drIK;
    switch (b->form) {
        case ACE_ANG_GAUSS: {
            const double dij = 2.0 * drIJ * c[0], dik = 2.0 * drIK * c[1];
            for (int j = 0; j < b->nr; ++j) b->damp[j] = or_fmexp(-b->ietar2[j] * r2s);
            for (int i = 0; i < b->n; ++i) {
                const double cw = fabs(0.5 + b->lambdaf[i] * cosIJK);
                const double gangle = dEdG[i] * pow(cw, b->zeta[i] - 1.0) * b->damp[b->rflag[i]];
                const double fangle = cw * gangle;
                *phi -= b->zeta[i] * b->lambdaf[i] * gangle;
                eta[0] -= fangle * (-dij * b->ieta2[i] + g[0]);
                eta[1] -= fangle * (-dik * b->ieta2[i] + g[1]);
            }
        } break;
        case ACE_ANG_GAUSS2: {
            const double dij = OR_PI * drIJ * c[0], dik = OR_PI * drIK * c[1];
            for (int j = 0; j < b->nr; ++j) b->damp[j] = or_fmexp(-0.5 * OR_PI * b->ietar2[j] * r2s);
            for (int i = 0; i < b->n; ++i) {
                const double fexp = b->damp[b->rflag[i]];
                const double den = 1.0 / (1.0 + fexp);
                const double cw = fabs(0.5 + b->lambdaf[i] * cosIJK);
                const double gangle = dEdG[i] * pow(cw, b->zeta[i] - 1.0) * 2.0 * fexp * den;
                const double fangle = cw * gangle;
                *phi -= b->zeta[i] * b->lambdaf[i] * gangle;
                eta[0] -= fangle * (-dij * b->ieta2[i] * den + g[0]);
                eta[1] -= fangle * (-dik * b->ieta2[i] * den + g[1]);
            }
        } break;
        case ACE_ANG_SECH: {
            const double dij = OR_PI * drIJ * c[0], dik = OR_PI * drIK * c[1];
            for (int j = 0; j < b->nr; ++j) b->damp[j] = or_fmexp(-0.5 * OR_PI * b->ietar2[j] * r2s);
            for (int i = 0; i < b->n; ++i) {
                const double fexp = b->damp[b->rflag[i]];
                const double fexp2 = fexp * fexp;
                const double den = 1.0 / (1.0 + fexp2);
                const double ftanh = (1.0 - fexp2) * den;
                const double fsech = 2.0 * fexp * den;
                const double cw = fabs(0.5 + b->lambdaf[i] * cosIJK);
                const double gangle = pow(cw, b->zeta[i] - 1.0) * dEdG[i] * fsech;
                const double fangle = cw * gangle;
                *phi -= b->zeta[i] * b->lambdaf[i] * gangle;
                eta[0] -= fangle * (-dij * b->ieta2[i] * ftanh + g[0]);
                eta[1] -= fangle * (-dik * b->ieta2[i] * ftanh + g[1]);
            }
        } break;
        case ACE_ANG_STUDENT3: case ACE_ANG_STUDENT4: case ACE_ANG_STUDENT5: {
            const int p = (b->form == ACE_ANG_STUDENT3) ? 4 : (b->form == ACE_ANG_STUDENT4) ? 5 : 6;
            const double dij = p * drIJ * c[0], dik = p * drIK * c[1];
            for (int j = 0; j < b->nr; ++j) b->damp[j] = 1.0 / sqrt(1.0 + b->ietar2[j] * r2s);
            for (int i = 0; i < b->n; ++i) {
                const double den = b->damp[b->rflag[i]];
                const double cw = fabs(0.5 + b->lambdaf[i] * cosIJK);
                const double gangle = pow(cw, b->zeta[i] - 1.0) * dEdG[i] * or_powint(den, p);
                const double fangle = cw * gangle;
                *phi -= b->zeta[i] * b->lambdaf[i] * gangle;
                eta[0] -= fangle * (-dij * b->ieta2[i] * den * den + g[0]);
                eta[1] -= fangle * (-dik * b->ieta2[i] * den * den + g[1]);
            }
        } break;
        default: *phi = NAN; return;
    }
    *phi *= c[0] * c[1];
    eta[0] *= c[1];
    eta[1] *= c[0];
}

/* ------------------------------------------------------------------------------------------- */
/* model layout, NNH::init_input (nnh.cpp:183-201)                                              */
/* ------------------------------------------------------------------------------------------- */

static int nab_of(int nt) { return nt * (nt + 1) / 2; }
static int tri_index(int j, int k) { return (j > k) ? j * (j + 1) / 2 + k : k * (k + 1) / 2 + j; }

static int input_r(const ace_model_desc* m, int I) {
    int s = 0;
    for (int J = 0; J < m->ntypes; ++J) s += m->r_n[I * m->ntypes + J];
    return s;
}
static int offset_r(const ace_model_desc* m, int I, int J) {
    int s = 0;
    for (int j = 0; j < J; ++j) s += m->r_n[I * m->ntypes + j];
    return s;
}
static int offset_a(const ace_model_desc* m, int I, int b) {
    int s = 0;
    const int nab = nab_of(m->ntypes);
    for (int j = 0; j < b; ++j) s += m->a_n[I * nab + j];
    return s;
}

int or_input_width(const ace_model_desc* m, int I) {
    const int nab = nab_of(m->ntypes);
    int s = input_r(m, I);
    for (int b = 0; b < nab; ++b) s += m->a_n[I * nab + b];
    return s;
}

int or_param_count(const ace_model_desc* m, int I) {
    const int* w = m->nn_width + I * (ACE_MAX_LAYERS + 1);
    int s = 0;
    for (int l = 0; l < m->nn_nlayer[I]; ++l) s += w[l + 1] + w[l + 1] * w[l];
    return s;
}

/* ------------------------------------------------------------------------------------------- */
/* neighbour list, neighbor.cpp:46-87                                                           */
/* ------------------------------------------------------------------------------------------- */

long long or_neighbors(const double* R, int n, const double* pos, double rc, long long cap,
                       long long* offsets, int* jidx, int* img) {
    int cf[3] = {0, 0, 0};
    double rn = 0;
    for (int i = 0; i < 9; ++i) rn += R[i] * R[i];
    if (sqrt(rn) > OR_ZERO) {
        for (int d = 0; d < 3; ++d) {
            const double* col = R + 3 * d;
            cf[d] = (int)ceil(3.0 * rc / sqrt(col[0] * col[0] + col[1] * col[1] + col[2] * col[2]));
        }
    }
    const double rc2 = rc * rc;
    long long cnt = 0;
    for (int a = 0; a < n; ++a) {
        offsets[a] = cnt;
        for (int b = 0; b < n; ++b) {
            const double dr[3] = {pos[3 * a] - pos[3 * b], pos[3 * a + 1] - pos[3 * b + 1], pos[3 * a + 2] - pos[3 * b + 2]};
            for (int i = -cf[0]; i <= cf[0]; ++i)
                for (int j = -cf[1]; j <= cf[1]; ++j)
                    for (int k = -cf[2]; k <= cf[2]; ++k) {
                        const double fi = 1.0 * i, fj = 1.0 * j, fk = 1.0 * k;
                        double ds2 = 0;
                        for (int d = 0; d < 3; ++d) {
                            const double sh = R[d] * fi + R[d + 3] * fj + R[d + 6] * fk;
                            const double v = dr[d] - sh;
                            ds2 = (d == 0) ? v * v : ds2 + v * v;
                        }
                        if (b != a || ds2 > OR_ZERO) {
                            if (ds2 < rc2) {
                                if (cnt < cap) {
                                    jidx[cnt] = b;
                                    img[3 * cnt] = i; img[3 * cnt + 1] = j; img[3 * cnt + 2] = k;
                                }
                                ++cnt;
                            }
                        }
                    }
        }
    }
    offsets[n] = cnt;
    return cnt;
}

/* r_IJ = posn(i) - (posn(j) + R*img), nnp.cpp:363 */
static void disp(const double* R, const double* pos, int i, int j, const int* im, double* out) {
    for (int d = 0; d < 3; ++d) {
        const double sh = R[d] * im[0] + R[d + 3] * im[1] + R[d + 6] * im[2];
        out[d] = pos[3 * i + d] - (pos[3 * j + d] + sh);
    }
}
static double norm3(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* ------------------------------------------------------------------------------------------- */
/* NNP::compute_symm, nnp.cpp:350-392                                                           */
/* ------------------------------------------------------------------------------------------- */

void or_symm(const ace_model_desc* m, const double* R, int n, const double* pos, const int* type,
             const long long* offsets, const int* jidx, const int* img, int ldg, double* G) {
    const int nt = m->ntypes, nab = nab_of(nt);
    or_ablk* blk = (or_ablk*)malloc(sizeof(or_ablk) * nt * nab);
    for (int b = 0; b < nt * nab; ++b) ablk_init(m, b, &blk[b]);
    double buf[1024];
    for (int i = 0; i < n; ++i) {
        double* Gi = G + (long long)i * ldg;
        for (int g = 0; g < ldg; ++g) Gi[g] = 0;
        const int II = type[i];
        const int nR = input_r(m, II);
        for (long long j = offsets[i]; j < offsets[i + 1]; ++j) {
            const int jj = jidx[j], JJ = type[jj];
            double rIJ[3];
            disp(R, pos, i, jj, img + 3 * j, rIJ);
            const double drIJ = norm3(rIJ);
            const int rb = II * nt + JJ;
            radial_symm(m->r_form[rb], m->r_cut[rb], m->r_rc[rb], m->r_n[rb], m->r_rs + m->r_off[rb],
                        m->r_eta + m->r_off[rb], drIJ, buf);
            const int oR = offset_r(m, II, JJ);
            for (int nr = 0; nr < m->r_n[rb]; ++nr) Gi[oR + nr] += buf[nr];
            for (long long k = j + 1; k < offsets[i + 1]; ++k) {
                const int kk = jidx[k], KK = type[kk];
                double rIK[3];
                disp(R, pos, i, kk, img + 3 * k, rIK);
                const double drIK = norm3(rIK);
                const double cosIJK = dot3(rIJ, rIK) / (drIJ * drIK);
                const int b = tri_index(JJ, KK);
                or_ablk* ab = &blk[II * nab + b];
                angular_symm(ab, drIJ, drIK, cosIJK, buf);
                const int oA = nR + offset_a(m, II, b);
                for (int na = 0; na < ab->n; ++na) Gi[oA + na] += buf[na];
            }
        }
    }
    for (int b = 0; b < nt * nab; ++b) ablk_free(&blk[b]);
    free(blk);
}

/* ------------------------------------------------------------------------------------------- */
/* network: activations nn.cpp:149-240, 465-512, 671-715; ANN::fpbp nn.cpp:1402-1420             */
/* ------------------------------------------------------------------------------------------- */

/* AFFPBP2 semantics: a = f(z), d = df/dz, d2 = d2f/dz2 */
static void act_eval(int act, double c, double z, double* a, double* d, double* d2) {
    switch (act) {
        case ACE_ACT_LINEAR: *a = z; *d = 1.0; *d2 = 0.0; break;
        case ACE_ACT_SIGMOID:
            if (z >= 0) {
                const double e = exp(-c * z), f = 1.0 / (1.0 + e);
                *a = f; *d = c * e * f * f; *d2 = -c * c * e * (1.0 - e) * f * f * f;
            } else {
                const double e = exp(c * z), f = 1.0 / (1.0 + e);
                *a = e * f; *d = c * e * f * f; *d2 = -c * c * e * (e - 1.0) * f * f * f;
            }
            break;
        case ACE_ACT_TANH: {
            const double t = tanh(c * z);
            *a = t; *d = c * (1.0 - t * t); *d2 = c * c * (-2.0 * t * (1.0 - t * t));
        } break;
        case ACE_ACT_SWISH: {
            const double zz = c * z;
            if (zz > 0.0) {
                const double e = exp(-zz), den = 1.0 / (1.0 + e);
                *a = z * den; *d = (1.0 + e * (1.0 + zz)) * den * den;
                *d2 = c * e * (2.0 - zz + e * (2.0 + zz)) * den * den * den;
            } else {
                const double e = exp(zz), den = 1.0 / (1.0 + e);
                *a = z * e * den; *d = e * (1.0 + zz + e) * den * den;
                *d2 = c * e * (2.0 + zz + e * (2.0 - zz)) * den * den * den;
            }
        } break;
        case ACE_ACT_SOFTPLUS: {
            const double ci = 1.0 / c;
            if (z >= 0.0) {
                const double e = exp(-c * z), den = 1.0 / (e + 1.0);
                *a = z + ci * (log1p(e) - OR_LOG2); *d = den; *d2 = c * e * den * den;
            } else {
                const double e = exp(c * z), den = 1.0 / (e + 1.0);
                *a = ci * (log1p(e) - OR_LOG2); *d = e * den; *d2 = c * e * den * den;
            }
        } break;
        default: *a = NAN; *d = NAN; *d2 = NAN;
    }
}

#define OR_MAXW 256
typedef struct {
    int L;
    const int* w;                 /* widths, w[0] = inputs */
    const double* b[ACE_MAX_LAYERS];
    const double* W[ACE_MAX_LAYERS];   /* column-major w[l+1] x w[l] */
    const double* inpw;
    const double* inpb;
    double outw, outb, c;
    int act;
    /* work */
    double ins[OR_MAXW];
    double z[ACE_MAX_LAYERS][OR_MAXW], a[ACE_MAX_LAYERS][OR_MAXW];
    double d[ACE_MAX_LAYERS][OR_MAXW], d2[ACE_MAX_LAYERS][OR_MAXW];
    double out;
} or_net;

static void net_bind(const ace_model_desc* m, int I, const double* params, or_net* nn) {
    nn->L = m->nn_nlayer[I];
    nn->w = m->nn_width + I * (ACE_MAX_LAYERS + 1);
    const double* p = params ? params : m->params + m->par_off[I];
    int cnt = 0;
    for (int l = 0; l < nn->L; ++l) { nn->b[l] = p + cnt; cnt += nn->w[l + 1]; }
    for (int l = 0; l < nn->L; ++l) { nn->W[l] = p + cnt; cnt += nn->w[l + 1] * nn->w[l]; }
    nn->inpw = m->inpw + m->in_off[I];
    nn->inpb = m->inpb + m->in_off[I];
    nn->outw = m->outw[I];
    nn->outb = m->outb[I];
    nn->c = m->nn_c[I];
    nn->act = m->nn_act[I];
}

/* ANN::fpbp2 (forward with first and second activation derivatives), nn.cpp:1378-1420 */
static double net_fpbp(or_net* nn, const double* G) {
    const int K = nn->w[0];
    for (int i = 0; i < K; ++i) nn->ins[i] = nn->inpw[i] * (G[i] + nn->inpb[i]);
    for (int l = 0; l < nn->L; ++l) {
        const int no = nn->w[l + 1], ni = nn->w[l];
        const double* in = (l == 0) ? nn->ins : nn->a[l - 1];
        const int act = (l == nn->L - 1) ? ACE_ACT_LINEAR : nn->act;
        for (int o = 0; o < no; ++o) {
            double s = 0;
            for (int i = 0; i < ni; ++i) s += nn->W[l][o + i * no] * in[i];
            nn->z[l][o] = nn->b[l][o] + s;
            act_eval(act, nn->c, nn->z[l][o], &nn->a[l][o], &nn->d[l][o], &nn->d2[l][o]);
        }
    }
    nn->out = nn->outb + nn->a[nn->L - 1][0] * nn->outw;
    return nn->out;
}

/* DODZ::grad row 0: dE/dG (nn.cpp:1836-1862) */
static void net_dodi(const or_net* nn, double* dodi) {
    double cur[OR_MAXW], nxt[OR_MAXW];
    const int L = nn->L;
    cur[0] = nn->outw * nn->d[L - 1][0];
    for (int l = L - 1; l > 0; --l) {
        const int no = nn->w[l + 1], ni = nn->w[l];
        for (int i = 0; i < ni; ++i) {
            double s = 0;
            for (int o = 0; o < no; ++o) s += cur[o] * nn->W[l][o + i * no];
            nxt[i] = s * nn->d[l - 1][i];
        }
        memcpy(cur, nxt, sizeof(double) * ni);
    }
    const int no = nn->w[1], K = nn->w[0];
    for (int i = 0; i < K; ++i) {
        double s = 0;
        for (int o = 0; o < no; ++o) s += cur[o] * nn->W[0][o + i * no];
        dodi[i] = s * nn->inpw[i];
    }
}

/* Cost::grad with dcdo = scale (nn.cpp:1747-1808): grad += scale * dE/dparams */
static void net_cost_grad(const or_net* nn, double scale, double* grad) {
    double dz[ACE_MAX_LAYERS][OR_MAXW];
    const int L = nn->L;
    dz[L - 1][0] = scale * nn->outw * nn->d[L - 1][0];
    for (int l = L - 1; l > 0; --l) {
        const int no = nn->w[l + 1], ni = nn->w[l];
        for (int i = 0; i < ni; ++i) {
            double s = 0;
            for (int o = 0; o < no; ++o) s += nn->W[l][o + i * no] * dz[l][o];
            dz[l - 1][i] = s * nn->d[l - 1][i];
        }
    }
    int cnt = 0;
    for (int l = 0; l < L; ++l)
        for (int o = 0; o < nn->w[l + 1]; ++o) grad[cnt++] += dz[l][o];
    for (int l = 0; l < L; ++l) {
        const int no = nn->w[l + 1], ni = nn->w[l];
        const double* in = (l == 0) ? nn->ins : nn->a[l - 1];
        for (int i = 0; i < ni; ++i)
            for (int o = 0; o < no; ++o) grad[cnt++] += dz[l][o] * in[i];
    }
}

/*
 * Forward-over-reverse: grad += scale * d/dparams [ dE/dG . v ] (network factor of d(F)/dW,
 * the contraction of NN::D2ODZDI nn.cpp:2047-2140 with the input direction v).
 * Tangent: s' = inpw*v, z'_l = W_l a'_{l-1}, a'_l = f'(z_l) z'_l, Q = outw a'_last.
 */
static void net_fwd_over_rev(const or_net* nn, const double* v, double scale, double* grad) {
    const int L = nn->L, K = nn->w[0];
    double ins_t[OR_MAXW];
    double zt[ACE_MAX_LAYERS][OR_MAXW], at[ACE_MAX_LAYERS][OR_MAXW];
    for (int i = 0; i < K; ++i) ins_t[i] = nn->inpw[i] * v[i];
    for (int l = 0; l < L; ++l) {
        const int no = nn->w[l + 1], ni = nn->w[l];
        const double* in = (l == 0) ? ins_t : at[l - 1];
        for (int o = 0; o < no; ++o) {
            double s = 0;
            for (int i = 0; i < ni; ++i) s += nn->W[l][o + i * no] * in[i];
            zt[l][o] = s;
            at[l][o] = nn->d[l][o] * s;
        }
    }
    /* reverse sweep: primal adjoint ab, tangent adjoint atb */
    double ab[OR_MAXW], atb[OR_MAXW], zb[OR_MAXW], ztb[OR_MAXW];
    ab[0] = 0.0;
    atb[0] = scale * nn->outw;
    /* parameter offsets */
    int boff[ACE_MAX_LAYERS], woff[ACE_MAX_LAYERS], cnt = 0;
    for (int l = 0; l < L; ++l) { boff[l] = cnt; cnt += nn->w[l + 1]; }
    for (int l = 0; l < L; ++l) { woff[l] = cnt; cnt += nn->w[l + 1] * nn->w[l]; }
    for (int l = L - 1; l >= 0; --l) {
        const int no = nn->w[l + 1], ni = nn->w[l];
        for (int o = 0; o < no; ++o) {
            ztb[o] = nn->d[l][o] * atb[o];
            zb[o] = nn->d2[l][o] * zt[l][o] * atb[o] + nn->d[l][o] * ab[o];
        }
        const double* in = (l == 0) ? nn->ins : nn->a[l - 1];
        const double* in_t = (l == 0) ? ins_t : at[l - 1];
        for (int o = 0; o < no; ++o) grad[boff[l] + o] += zb[o];
        for (int i = 0; i < ni; ++i)
            for (int o = 0; o < no; ++o)
                grad[woff[l] + o + i * no] += zb[o] * in[i] + ztb[o] * in_t[i];
        if (l > 0) {
            for (int i = 0; i < ni; ++i) {
                double s = 0, st = 0;
                for (int o = 0; o < no; ++o) {
                    s += nn->W[l][o + i * no] * zb[o];
                    st += nn->W[l][o + i * no] * ztb[o];
                }
                ab[i] = s;
                atb[i] = st;
            }
        }
    }
}

double or_network(const ace_model_desc* m, int I, const double* G, double* dEdG) {
    or_net* nn = (or_net*)malloc(sizeof(or_net));
    net_bind(m, I, NULL, nn);
    const double out = net_fpbp(nn, G);
    if (dEdG) net_dodi(nn, dEdG);
    free(nn);
    return out;
}

double or_energy(const ace_model_desc* m, int n, const int* type, int ldg, const double* G) {
    or_net* nn = (or_net*)malloc(sizeof(or_net));
    double pe = 0;
    for (int i = 0; i < n; ++i) {
        net_bind(m, type[i], NULL, nn);
        pe += net_fpbp(nn, G + (long long)i * ldg) + m->E0[type[i]];
    }
    free(nn);
    return pe;
}

/* ------------------------------------------------------------------------------------------- */
/* NNP::compute_force, nnp.cpp:412-476                                                          */
/* ------------------------------------------------------------------------------------------- */

double or_force(const ace_model_desc* m, const double* R, int n, const double* pos, const int* type,
                const long long* offsets, const int* jidx, const int* img, int ldg,
                const double* G, double* F, double* virial) {
    const int nt = m->ntypes, nab = nab_of(nt);
    or_ablk* blk = (or_ablk*)malloc(sizeof(or_ablk) * nt * nab);
    for (int b = 0; b < nt * nab; ++b) ablk_init(m, b, &blk[b]);
    or_net* nn = (or_net*)malloc(sizeof(or_net));
    double dEdG[OR_MAXW];
    double pe = 0;
    double W[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3 * n; ++i) F[i] = 0;
    for (int i = 0; i < n; ++i) {
        const int II = type[i];
        net_bind(m, II, NULL, nn);
        pe += net_fpbp(nn, G + (long long)i * ldg);
        net_dodi(nn, dEdG);
        const int nR = input_r(m, II);
        for (long long j = offsets[i]; j < offsets[i + 1]; ++j) {
            const int jj = jidx[j], JJ = type[jj];
            double rIJ[3];
            disp(R, pos, i, jj, img + 3 * j, rIJ);
            const double drIJ = norm3(rIJ);
            const double dIJi = 1.0 / drIJ;
            const int rb = II * nt + JJ;
            const double amp = radial_force(m->r_form[rb], m->r_cut[rb], m->r_rc[rb], m->r_n[rb],
                                            m->r_rs + m->r_off[rb], m->r_eta + m->r_off[rb], drIJ,
                                            dEdG + offset_r(m, II, JJ)) * dIJi;
            for (int d = 0; d < 3; ++d) {
                F[3 * i + d] += amp * rIJ[d];
                F[3 * jj + d] -= amp * rIJ[d];
            }
            /* dE/dd = -amp * rIJ */
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) W[3 * a + c] += amp * rIJ[a] * rIJ[c];
            for (long long k = j + 1; k < offsets[i + 1]; ++k) {
                const int kk = jidx[k], KK = type[kk];
                double rIK[3];
                disp(R, pos, i, kk, img + 3 * k, rIK);
                const double drIK = norm3(rIK);
                const double dIKi = 1.0 / drIK;
                const int b = tri_index(JJ, KK);
                const int oA = nR + offset_a(m, II, b);
                double phi = 0, eta[2] = {0, 0};
                const double cosIJK = dot3(rIJ, rIK) * dIJi * dIKi;
                angular_force(&blk[II * nab + b], drIJ, drIK, cosIJK, &phi, eta, dEdG + oA);
                const double fi0 = (phi * (dIKi - cosIJK * dIJi) + eta[0]) * dIJi;
                const double fi1 = (phi * (dIJi - cosIJK * dIKi) + eta[1]) * dIKi;
                const double fj0 = (-phi * cosIJK * dIJi + eta[0]) * dIJi, fj1 = phi * dIJi * dIKi;
                const double fk0 = (-phi * cosIJK * dIKi + eta[1]) * dIKi, fk1 = phi * dIKi * dIJi;
                for (int d = 0; d < 3; ++d) {
                    F[3 * i + d] += fi0 * rIJ[d];
                    F[3 * i + d] += fi1 * rIK[d];
                    F[3 * jj + d] -= fj0 * rIJ[d] + fj1 * rIK[d];
                    F[3 * kk + d] -= fk0 * rIK[d] + fk1 * rIJ[d];
                }
                /* dE/dd_ij = -(fj0 rIJ + fj1 rIK), dE/dd_ik = -(fk0 rIK + fk1 rIJ) */
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c)
                        W[3 * a + c] += (fj0 * rIJ[a] + fj1 * rIK[a]) * rIJ[c] + (fk0 * rIK[a] + fk1 * rIJ[a]) * rIK[c];
            }
        }
    }
    if (virial) for (int q = 0; q < 9; ++q) virial[q] = W[q];
    for (int b = 0; b < nt * nab; ++b) ablk_free(&blk[b]);
    free(blk);
    free(nn);
    return pe;
}

/* ------------------------------------------------------------------------------------------- */
/* training, NNPTEFR::error_cost_o1 nnptefr.cpp:1292-1425 + force-matching term                 */
/* ------------------------------------------------------------------------------------------- */

static double norm_factor(int norm, double nAtoms) {   /* nnptefr.cpp:609-632 */
    switch (norm) {
        case ACE_NORM_IDENTITY: return 1.0;
        case ACE_NORM_LINEAR: return 1.0 / nAtoms;
        case ACE_NORM_SQRT: return 1.0 / sqrt(nAtoms);
        case ACE_NORM_CBRT: return 1.0 / cbrt(nAtoms);
        case ACE_NORM_LOG: return 1.0 / (log(nAtoms) + 1.0);
        default: return NAN;
    }
}

void or_train_struct(const ace_model_desc* m, const ace_loss_desc* L, const double* R, int n,
                     const double* pos, const int* type, const long long* offsets, const int* jidx,
                     const int* img, int ldg, const double* G, double E_ref, const double* F_ref,
                     double* grad, double* out_loss) {
    const int nt = m->ntypes;
    int goff[ACE_MAX_TYPES + 1];
    goff[0] = 0;
    for (int t = 0; t < nt; ++t) goff[t + 1] = goff[t] + or_param_count(m, t);
    or_net* nn = (or_net*)malloc(sizeof(or_net));
    /* energy: pe = sum(out) + sum(E0) */
    double pe = 0;
    for (int i = 0; i < n; ++i) {
        net_bind(m, type[i], NULL, nn);
        pe += net_fpbp(nn, G + (long long)i * ldg) + m->E0[type[i]];
    }
    const double nAtoms = n;
    const double dE = pe - E_ref;
    const double nf = norm_factor(L->norm, nAtoms);
    const double dU = dE * nf;
    double gfac = nf, le = 0;
    switch (L->loss) {
        case ACE_LOSS_MSE: le = 0.5 * dU * dU; gfac *= dU; break;
        case ACE_LOSS_MAE: le = fabs(dU); gfac *= (dU > 0) - (dU < 0); break;
        case ACE_LOSS_HUBER: {
            const double arg = dU / L->delta, sq = sqrt(1.0 + arg * arg);
            le = L->delta * L->delta * (sq - 1.0); gfac *= dU / sq;
        } break;
        case ACE_LOSS_ASINH: {
            const double arg = dU / L->delta, sq = sqrt(1.0 + arg * arg), lg = log(arg + sq);
            le = L->delta * L->delta * (1.0 - sq + arg * lg); gfac *= lg * L->delta;
        } break;
    }
    out_loss[0] = L->w_energy * le;
    out_loss[2] = (dE * dE) / (nAtoms * nAtoms);
    gfac *= L->w_energy;
    for (int i = 0; i < n; ++i) {
        net_bind(m, type[i], NULL, nn);
        net_fpbp(nn, G + (long long)i * ldg);
        net_cost_grad(nn, gfac, grad + goff[type[i]]);
    }
    out_loss[1] = 0;
    if (L->w_force != 0.0 && F_ref) {
        double* F = (double*)malloc(sizeof(double) * 3 * n);
        double* eps = (double*)malloc(sizeof(double) * 3 * n);
        or_force(m, R, n, pos, type, offsets, jidx, img, ldg, G, F, NULL);
        double lf = 0;
        for (int q = 0; q < 3 * n; ++q) {
            const double r = F[q] - F_ref[q];
            lf += 0.5 * r * r;
            eps[q] = L->w_force * r;          /* dL/dF */
        }
        out_loss[1] = L->w_force * lf;
        /* v_i = sum_a eps_a . dG_i/dr_a, then grad -= d/dW (dE_i/dG . v_i) */
        const int nab = nab_of(nt);
        double v[OR_MAXW];
        for (int i = 0; i < n; ++i) {
            const int II = type[i];
            const int K = or_input_width(m, II);
            const int nR = input_r(m, II);
            for (int g = 0; g < K; ++g) v[g] = 0;
            for (long long j = offsets[i]; j < offsets[i + 1]; ++j) {
                const int jj = jidx[j], JJ = type[jj];
                double rIJ[3], eij[3];
                disp(R, pos, i, jj, img + 3 * j, rIJ);
                const double drIJ = norm3(rIJ);
                for (int d = 0; d < 3; ++d) eij[d] = eps[3 * i + d] - eps[3 * jj + d];
                const double pIJ = dot3(rIJ, eij) / drIJ;     /* rhat_ij . (eps_i - eps_j) */
                const int rb = II * nt + JJ;
                const int oR = offset_r(m, II, JJ);
                for (int q = 0; q < m->r_n[rb]; ++q)
                    v[oR + q] += or_radial_d(m->r_form[rb], m->r_cut[rb], m->r_rc[rb],
                                             m->r_eta[m->r_off[rb] + q], m->r_rs[m->r_off[rb] + q], drIJ) * pIJ;
                for (long long k = j + 1; k < offsets[i + 1]; ++k) {
                    const int kk = jidx[k], KK = type[kk];
                    double rIK[3], eik[3];
                    disp(R, pos, i, kk, img + 3 * k, rIK);
                    const double drIK = norm3(rIK);
                    for (int d = 0; d < 3; ++d) eik[d] = eps[3 * i + d] - eps[3 * kk + d];
                    const double cosv = dot3(rIJ, rIK) / (drIJ * drIK);
                    const double pIK = dot3(rIK, eik) / drIK;
                    /* dcos/dd_ij . e_ij + dcos/dd_ik . e_ik */
                    const double dcos = (dot3(rIK, eij) / drIK - cosv * pIJ) / drIJ
                                      + (dot3(rIJ, eik) / drIJ - cosv * pIK) / drIK;
                    const int b = tri_index(JJ, KK);
                    const int ab = II * nab + b;
                    const int oA = nR + offset_a(m, II, b);
                    for (int q = 0; q < m->a_n[ab]; ++q) {
                        double fphi, feta[2];
                        const int o = m->a_off[ab] + q;
                        or_angular_d(m->a_form[ab], m->a_cut[ab], m->a_rc[ab], m->a_eta[o], m->a_zeta[o],
                                     m->a_lambda[o], cosv, drIJ, drIK, &fphi, feta);
                        v[oA + q] += feta[0] * pIJ + feta[1] * pIK + fphi * dcos;
                    }
                }
            }
            net_bind(m, II, NULL, nn);
            net_fpbp(nn, G + (long long)i * ldg);
            net_fwd_over_rev(nn, v, -1.0, grad + goff[II]);
        }
        free(F);
        free(eps);
    }
    free(nn);
}

/* ADAM::step, algo.cpp:280-296 */
void or_adam_step(int dim, double* p, const double* g, double gamma, double eps, double* m,
                  double* v, double* b) {
    const double beta1 = 0.9, beta2 = 0.999;
    const double b1ii = 1.0 / (1.0 - b[0]);
    const double b2ii = 1.0 / (1.0 - b[1]);
    for (int n = 0; n < dim; ++n) {
        m[n] *= beta1; m[n] += (1.0 - beta1) * g[n];
        v[n] *= beta2; v[n] += (1.0 - beta2) * g[n] * g[n];
        p[n] -= gamma * b1ii * m[n] / (sqrt(b2ii * v[n]) + eps);
    }
    b[0] *= beta1;
    b[1] *= beta2;
}
