/*
 * oracle_geom.h -- TEST INFRASTRUCTURE ONLY (the oracle).  Shared,
 * precision-independent parts of the CPU restatement: the D3Q19 tables, the FP32
 * inside tests, the motion law and the brute-force boundary-link build.
 *
 * Nothing in the shipped library includes or links this file; tests and the
 * benchmark's CPU-baseline leg use it as the checker.
 */
#ifndef ORACLE_GEOM_H
#define ORACLE_GEOM_H

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/lbm_hip.h"

/* D3Q19 in the reference order (src/Utilities.jl:950-952), 0-based. */
static const int OCX[19] = {1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0};
static const int OCY[19] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 0};
static const int OCZ[19] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 0};
/* bbOpp (src/Utilities.jl:960) minus one. */
static const int OOPP[19] = {1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 17, 16, 18};
/* weight class: 0 -> 1/18, 1 -> 1/36, 2 -> 1/3 (src/Utilities.jl:954-958). */
static const int OWCLS[19] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2};

/* D3Q27 in the reference order (src/Utilities.jl:962-976), 0-based: the D3Q19
 * moving directions, then the eight corners, then rest. */
static const int OCX27[27] = {1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0,
                              1, -1, 1, -1, 1, -1, -1, 1, 0};
static const int OCY27[27] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1,
                              1, -1, 1, -1, -1, 1, 1, -1, 0};
static const int OCZ27[27] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1,
                              1, -1, -1, 1, 1, -1, 1, -1, 0};
static const int OOPP27[27] = {1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 17, 16,
                               19, 18, 21, 20, 23, 22, 25, 24, 26};
/* D3Q19 rest direction in the 27-entry layout of the oracle's tables */
static const int OCX19[27] = {1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0};
static const int OCY19[27] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 0};
static const int OCZ19[27] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 0};
static const int OOPP19[27] = {1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 17, 16, 18};

/* Planted mistakes, for tests that show a closed-form pin fails when the
 * oracle is wrong (set by orc_plant, 0 = none).  Never set outside tests. */
enum {
    ORC_PLANT_NONE = 0,
    ORC_PLANT_SPIN_SIGN = 1,   /* wall velocity U + r x Omega instead of U + Omega x r */
    ORC_PLANT_ARM_SIGN = 2,    /* torque arm X - x_w instead of x_w - X */
    ORC_PLANT_ARM_MIRROR = 3,  /* torque arm to x - h c_d instead of x + h c_d */
    ORC_PLANT_WALL_HALF = 4,   /* moving-wall term w rho 3 (c.u_w), half of 2 w rho 3 (c.u_w) */
};
static int orc_planted = ORC_PLANT_NONE;

/* Per-body data prepared on the host once per pose (FP32 mask, FP64 kinematics). */
typedef struct {
    int shape;
    float c[3];
    float r2;          /* sphere, cylinder R*R */
    float ia2[3];      /* ellipsoid 1/(a*a), 1/(b*b), 1/(c*c) */
    float rot[9];
    float semi[3];
    float naca[5];     /* m, p, t, x_int, x_le_offset */
    float rh[9];       /* half-step rotation, FP32 */
    double ch[3], v[3], om[3];
} orc_body;

static void orc_prepare_body(const lbm_body* b, orc_body* o) {
    memset(o, 0, sizeof(*o));
    o->shape = b->shape;
    for (int k = 0; k < 3; ++k) {
        o->c[k] = b->center[k];
        o->ch[k] = b->center_half[k];
        o->v[k] = b->vel[k];
        o->om[k] = b->omega[k];
    }
    for (int k = 0; k < 9; ++k) {
        o->rot[k] = b->rot[k];
        o->rh[k] = (float)b->rot_half[k];
    }
    for (int k = 0; k < 3; ++k) o->semi[k] = b->semi[k];
    for (int k = 0; k < 5; ++k) o->naca[k] = b->naca[k];
    if (b->shape == LBM_SPHERE || b->shape == LBM_CYLINDER) {
        float R = b->semi[0];
        o->r2 = R * R;
    } else if (b->shape == LBM_ELLIPSOID) {
        for (int k = 0; k < 3; ++k) {
            float a = b->semi[k];
            o->ia2[k] = 1.0f / (a * a);
        }
    }
}

/* NACA 4-digit surfaces at chordwise position x from the leading edge, FP32,
 * in the reference's operation order: _naca_thickness_y
 * (src/Geometry.jl:325-330), _naca_camber_y_dydx (:332-345) and
 * _naca_upper_lower_y (:347-356).  The camber line stays in chord fractions
 * (not multiplied by the chord), as there.  xi^4 is evaluated in FP64 and
 * rounded once (Julia's ^(::Float32, ::Integer) does that), and
 * cos(atan(s)) = 1/sqrt(1 + s^2) exactly. */
static void orc_naca_lines(float x, float chord, float m, float p, float t, float x_int, float* yl, float* yu) {
    float xbar = x / chord;
    float xi = x_int * xbar;
    if (!(xi > 0.0f)) xi = 0.0f;
    float rt = sqrtf(xi);
    float x3 = xi * xi * xi;
    double xd = (double)xi;
    float x4 = (float)((xd * xd) * (xd * xd));
    float poly = 0.2969f * rt - 0.1260f * xi;
    poly = poly - 0.3516f * xi * xi;
    poly = poly + 0.2843f * x3;
    poly = poly - 0.1015f * x4;
    float yt = 5.0f * t * chord * poly;
    float yc = 0.0f, dyb = 0.0f;
    if (m == 0.0f || p == 0.0f) {
        yc = 0.0f;
        dyb = 0.0f;
    } else if (xbar < p) {
        yc = m / (p * p) * (2.0f * p * xbar - xbar * xbar);
        dyb = 2.0f * m / (p * p) * (p - xbar);
    } else {
        float den = 1.0f - p;
        float den2 = den * den;
        yc = m / den2 * ((1.0f - 2.0f * p) + 2.0f * p * xbar - xbar * xbar);
        dyb = 2.0f * m / den2 * (p - xbar);
    }
    float sl = dyb / chord;
    float cs = 1.0f / sqrtf(1.0f + sl * sl);
    *yu = yc + yt * cs;
    *yl = yc - yt * cs;
}

/* Minimum-image displacement along a periodic axis of length L (FP32). */
static float orc_min_image(float d, int periodic, int L) {
    if (periodic) {
        float h = 0.5f * (float)L;
        if (d > h) d -= (float)L;
        else if (d < -h) d += (float)L;
    }
    return d;
}

/* Inside test in FP32 with the reference's tolerance R^2 + 1e-6 R^2
 * (src/Geometry.jl:73; the airfoil src/Geometry.jl:75-82 with its 1e-6 chord
 * tolerance), evaluated left to right with no contraction.  The cylinder
 * ignores z (extrusion); the sphere adds the z term. */
static int orc_inside(const orc_body* b, float px, float py, float pz, const int per[3],
                      const int L[3]) {
    float dx = orc_min_image(px - b->c[0], per[0], L[0]);
    float dy = orc_min_image(py - b->c[1], per[1], L[1]);
    float dz = orc_min_image(pz - b->c[2], per[2], L[2]);
    if (b->shape == LBM_SPHERE) {
        float s = dx * dx + dy * dy;
        s = s + dz * dz;
        return s <= b->r2 + 1e-6f * b->r2;
    } else if (b->shape == LBM_CYLINDER) {
        float s = dx * dx + dy * dy;
        return s <= b->r2 + 1e-6f * b->r2;
    } else if (b->shape == LBM_NACA4) {
        /* _rot_to_body (src/Geometry.jl:321-323): rot = [[c,-s,0],[s,c,0],[0,0,1]] */
        float xb = b->rot[0] * dx + b->rot[3] * dy;
        float yb = b->rot[1] * dx + b->rot[4] * dy;
        float xle = xb + b->naca[4];
        float chord = b->semi[0];
        if (!(0.0f <= xle && xle <= chord)) return 0;
        float yl, yu;
        orc_naca_lines(xle, chord, b->naca[0], b->naca[1], b->naca[2], b->naca[3], &yl, &yu);
        float tol = 1e-6f * chord;
        return yl - tol <= yb && yb <= yu + tol;
    } else {
        float xb = b->rot[0] * dx + b->rot[3] * dy;
        xb = xb + b->rot[6] * dz;
        float yb = b->rot[1] * dx + b->rot[4] * dy;
        yb = yb + b->rot[7] * dz;
        float zb = b->rot[2] * dx + b->rot[5] * dy;
        zb = zb + b->rot[8] * dz;
        float q = xb * xb * b->ia2[0] + yb * yb * b->ia2[1];
        q = q + zb * zb * b->ia2[2];
        return q <= 1.0f + 1e-6f;
    }
}

/* Level set at p (FP32, half-step pose), negative inside: phi_cylinder_lat
 * (src/Geometry.jl:368-372) and phi_airfoil_lat (src/Geometry.jl:375-396) with
 * the half-step centre and rotation (src/Motion.jl:104-116); sphere |d| - R;
 * ellipsoid sqrt(sum (x_i/a_i)^2) - 1.  hypot(a, b) is sqrt(a^2 + b^2). */
static float orc_phi(const orc_body* b, float px, float py, float pz, const int per[3], const int L[3]) {
    float dx = orc_min_image(px - (float)b->ch[0], per[0], L[0]);
    float dy = orc_min_image(py - (float)b->ch[1], per[1], L[1]);
    float dz = orc_min_image(pz - (float)b->ch[2], per[2], L[2]);
    if (b->shape == LBM_SPHERE) {
        float s = dx * dx + dy * dy;
        s = s + dz * dz;
        return sqrtf(s) - b->semi[0];
    }
    if (b->shape == LBM_CYLINDER) {
        float s = dx * dx + dy * dy;
        return sqrtf(s) - b->semi[0];
    }
    if (b->shape == LBM_ELLIPSOID) {
        float e = 0.0f;
        for (int k = 0; k < 3; ++k) {
            float r = b->rh[k] * dx + b->rh[3 + k] * dy;
            r = r + b->rh[6 + k] * dz;
            float t = r / b->semi[k];
            e = e + t * t;
        }
        return sqrtf(e) - 1.0f;
    }
    float xb = b->rh[0] * dx + b->rh[3] * dy;
    float yb = b->rh[1] * dx + b->rh[4] * dy;
    float xle = xb + b->naca[4];
    float chord = b->semi[0];
    if (xle < 0.0f) return sqrtf(xle * xle + yb * yb);
    if (xle > chord) return sqrtf((xle - chord) * (xle - chord) + yb * yb);
    float yl, yu;
    orc_naca_lines(xle, chord, b->naca[0], b->naca[1], b->naca[2], b->naca[3], &yl, &yu);
    if (yb >= yl && yb <= yu) return -fminf(yu - yb, yb - yl);
    return yb > yu ? yb - yu : yl - yb;
}

/* Rodrigues rotation about unit axis k by angle th, composed as Rot * R0. */
static void orc_rotation_at(const double om[3], double t, const double R0[9], double out[9]) {
    double w = sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
    double Rr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (w > 0.0) {
        double k[3] = {om[0] / w, om[1] / w, om[2] / w};
        double th = w * t, s = sin(th), c1 = 1.0 - cos(th);
        double K[9] = {0, -k[2], k[1], k[2], 0, -k[0], -k[1], k[0], 0};
        double K2[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double a = 0.0;
                for (int m = 0; m < 3; ++m) a += K[i * 3 + m] * K[m * 3 + j];
                K2[i * 3 + j] = a;
            }
        for (int i = 0; i < 9; ++i) Rr[i] += s * K[i] + c1 * K2[i];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double a = 0.0;
            for (int m = 0; m < 3; ++m) a += Rr[i * 3 + m] * R0[m * 3 + j];
            out[i * 3 + j] = a;
        }
}

/* motion_at for step n (1-based): mask pose at t = n-1, kinematics at t = n-1/2
 * (src/Motion.jl:26-36 with dt = 1).  Centre c0 + v t + A sin(w t + phase),
 * velocity v + A w cos(w t + phase) (the reference's heave law,
 * src/MainCore.jl:72-80, is A = (0, -0.4 R, 0), phase 0). */
static void orc_body_from_law(const lbm_motion* m, int64_t n, lbm_body* b) {
    memset(b, 0, sizeof(*b));
    b->shape = m->shape;
    double t0 = (double)(n - 1), th = (double)(n - 1) + 0.5;
    double R[9], Rh[9];
    orc_rotation_at(m->omega, t0, m->rot0, R);
    orc_rotation_at(m->omega, th, m->rot0, Rh);
    double s0 = sin(m->freq * t0 + m->phase), sh = sin(m->freq * th + m->phase);
    double chh = cos(m->freq * th + m->phase);
    for (int k = 0; k < 3; ++k) {
        b->center[k] = (float)(m->center0[k] + m->vel[k] * t0 + m->amp[k] * s0);
        b->semi[k] = (float)m->semi[k];
        b->center_half[k] = m->center0[k] + m->vel[k] * th + m->amp[k] * sh;
        b->vel[k] = m->vel[k] + m->amp[k] * m->freq * chh;
        b->omega[k] = m->omega[k];
    }
    for (int k = 0; k < 9; ++k) {
        b->rot[k] = (float)R[k];
        b->rot_half[k] = Rh[k];
    }
    for (int k = 0; k < 5; ++k) b->naca[k] = (float)m->naca[k];
}

#endif /* ORACLE_GEOM_H */
